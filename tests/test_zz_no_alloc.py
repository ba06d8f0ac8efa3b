"""The allocation-free training step (f3d_nn_train_step_static: partial
gradients in the library's static device scratch) against f3d_nn_train_step
(partials in a buffer the context allocates), through the C-ABI: two calls
in a row, the second with a smaller batch than the first, each from the same
parameters and data on two contexts, must give bit-identical losses and
parameters (same kernels, same summation order); a batch beyond the static
scratch is refused with an error."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def test_train_step_static_matches_train_step():
    import flame3d_amd.abi as abi
    import flame3d_amd.mech as mm
    import flame3d_amd.surrogate as sur
    print("\n[no-alloc] f3d_nn_train_step_static vs f3d_nn_train_step", flush=True)
    m = mm.load("h2_lidryer")
    cfg, params, _ = sur.load()
    ctxs = []
    for _ in range(2):
        c = abi.Context(8, 8, 8, m, "fp32")
        c.nn_init(cfg)
        c.nn_set_params(params)
        ctxs.append(c)
    g = torch.Generator(device="cuda").manual_seed(5)
    for B in (8192, 1024):  # the scratch's full size, then a smaller batch
        X = torch.randn((B, m.Ns + 2), generator=g, device="cuda", dtype=torch.float32)
        Y = torch.randn((B, m.Ns), generator=g, device="cuda", dtype=torch.float32)
        la = torch.zeros(1, device="cuda", dtype=torch.float32)
        lb = torch.zeros(1, device="cuda", dtype=torch.float32)
        ctxs[0].nn_train_step(X, Y, 1e-3, la)
        ctxs[1].nn_train_step_static(X, Y, 1e-3, lb)
        torch.cuda.synchronize()
        pa, pb = np.asarray(ctxs[0].nn_get_params()), np.asarray(ctxs[1].nn_get_params())
        print(f"  B={B}: loss {float(la):.6e} vs {float(lb):.6e}, "
              f"params max |diff| {np.abs(pa - pb).max():.1e}", flush=True)
        assert float(la) == float(lb)
        assert np.array_equal(pa, pb)
        assert not np.array_equal(pa, params)  # the step moved the weights
    X = torch.zeros((8256, m.Ns + 2), device="cuda", dtype=torch.float32)
    Y = torch.zeros((8256, m.Ns), device="cuda", dtype=torch.float32)
    with pytest.raises(abi.F3DError, match="8192"):
        ctxs[1].nn_train_step_static(X, Y, 1e-3)
    for c in ctxs:
        c.close()
