"""GPU parity of the native time-step driver (f3d_time_step, solver.jl:240-378)
and of the NaN check (solver.jl:388) against the FP64 oracle driven through
the same sequence: chemistry over dt/2 + fills, three RK3 stages, chemistry
over dt/2 + fills.  Workload: the reacting H2/air mixing layer on a small
stretched block.  Bars: 1e-12 normwise relative (FP64); the FP32 + surrogate
steps against the FP64 Newton steps: bounds on the accumulated difference
(written in the test)."""
import os

import numpy as np
import pytest
import torch

from helpers import rel_err_per_var, to_np, to_torch

pytestmark = pytest.mark.gpu
SIZE = (16, 12, 10)


def _setup(prec="fp64", size=SIZE):
    import flame3d_amd.abi as abi
    import flame3d_amd.cases as cases
    import flame3d_amd.mech as mm
    m = mm.load("h2_lidryer")
    case = cases.reacting_mixing_layer(size, m, dev="cpu")
    ctx = abi.Context(*size, m, prec)
    ctx.set_metrics([x.cuda() for x in case.grid.metrics], case.grid.J.cuda())
    ctx.set_boundary(case.bc)
    return m, case, ctx


def _oracle_step(orc, case, U, Un, r, rn, Q, ws, M, J, sub):
    def half():
        for _ in range(sub):
            orc.chem_stiff(U, Q, r, case.dt / 2 / sub)
        orc.fill(case.bc, Q, U, r)
    half()
    for k in (1, 2, 3):
        orc.rk3_stage(k, U, Un, r, rn, Q, ws, M, J, case.dt, case.bc)
    half()


def test_time_step_parity_stiff():
    """Two time steps with eval_gpu_stiff chemistry (2 substeps per half
    step) against the oracle driven through the same sequence: 1e-12."""
    import flame3d_amd.abi as abi
    from oracle import Oracle
    print(f"\n[driver] time_step chem=stiff sub_step=2 size={SIZE}", flush=True)
    m, case, ctx = _setup()
    orc = Oracle(*SIZE, m)
    Q, r = to_np(case.Q), to_np(case.rhoi)
    M, J = [to_np(x) for x in case.grid.metrics], to_np(case.grid.J)
    U = orc.prim2c(Q)
    Un, rn, ws = orc.zeros(5), orc.zeros(m.Ns), orc.workspace()
    Qg, rg, Ug = to_torch(Q), to_torch(r), to_torch(U)
    Ung, rng_ = ctx.zeros(5), ctx.zeros(m.Ns)
    for _ in range(2):
        _oracle_step(orc, case, U, Un, r, rn, Q, ws, M, J, 2)
        assert ctx.time_step(Ug, Ung, rg, rng_, Qg, case.dt, abi.CHEM_STIFF, 2) == 0
    torch.cuda.synchronize()
    for name, got, ref in (("U", Ug, U), ("rhoi", rg, r), ("Q", Qg, Q)):
        errs = rel_err_per_var(to_np(got), ref)
        print(f"  {name} rel err per variable {['%.1e' % e for e in errs]}", flush=True)
        assert max(errs) <= 1e-12, (name, errs)
    ctx.close()


def test_time_step_newton_is_the_call_sequence():
    """f3d_time_step with Newton chemistry equals, bit for bit, the same
    sequence of library calls made from the host (the Newton step itself is
    checked against the oracle, iteration counts included, in
    test_gpu_chem.py; over several steps an iterative solve cannot be held to
    1e-12 against a second implementation: a cell whose update sits at the
    tolerance may stop one iteration apart)."""
    import flame3d_amd.abi as abi
    print(f"\n[driver] time_step chem=newton vs call sequence size={SIZE}", flush=True)
    m, case, ctx = _setup()
    Q, r = case.Q.cuda(), case.rhoi.cuda()
    U = ctx.zeros(5)
    ctx.prim2c(U, Q)
    Q2, r2, U2 = Q.clone(), r.clone(), U.clone()
    Un, rn = ctx.zeros(5), ctx.zeros(m.Ns)
    for _ in range(2):
        ctx.time_step(U, Un, r, rn, Q, case.dt, abi.CHEM_NEWTON)
        for part in ("chem", "rk", "chem"):
            if part == "chem":
                ctx.chem_newton(U2, Q2, r2, case.dt / 2, 50, 1e-10, 1e-6)
                ctx.fill(Q2, U2, r2)
            else:
                for k in (1, 2, 3):
                    ctx.rk3_stage(k, U2, Un, r2, rn, Q2, case.dt)
    torch.cuda.synchronize()
    assert torch.equal(U, U2) and torch.equal(r, r2) and torch.equal(Q, Q2)
    assert not ctx.nan_check(U)
    ctx.close()


def test_nan_check():
    print("\n[driver] nan_check", flush=True)
    m, case, ctx = _setup()
    U = ctx.zeros(5)
    ctx.prim2c(U, case.Q.cuda())
    assert ctx.nan_check(U) is False
    U[4, 0, 0, 0] = float("nan")  # a ghost corner counts too
    assert ctx.nan_check(U) is True
    U[4, 0, 0, 0] = 1.0
    U[2, 7, 6, 9] = float("nan")
    assert ctx.nan_check(U) is True
    ctx.close()


SURROGATE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                         "flame3d.jl_amd", "mechanisms", "h2_surrogate.npz")


def _run_steps(size, prec, nn, steps, perturb=0.0):
    """`steps` time steps of the reacting mixing layer from its initial
    state; returns the interior (T, Y) and the surrogate's fallback count."""
    import flame3d_amd.abi as abi
    import flame3d_amd.surrogate as sur
    m, case, ctx = _setup(prec, size)
    if nn:
        sur.attach(ctx)
    Q, r = case.Q.cuda(), case.rhoi.cuda()
    if perturb:
        g = torch.Generator(device="cuda").manual_seed(7)
        for n in (1, 2, 3):
            Q[n] *= 1 + perturb * torch.randn(Q[n].shape, generator=g, device="cuda",
                                              dtype=torch.float64)
    U = ctx.zeros(5)
    ctx.prim2c(U, Q)
    Un, rn = ctx.zeros(5), ctx.zeros(m.Ns)
    nfb = 0
    for _ in range(steps):
        nfb += ctx.time_step(U, Un, r, rn, Q, case.dt,
                             abi.CHEM_SURROGATE if nn else abi.CHEM_NEWTON)
    torch.cuda.synchronize()
    ctx.close()
    sl = (slice(4, -4),) * 3
    return Q[5][sl].clone(), r[(slice(None),) + sl] / Q[0][sl].unsqueeze(0), nfb


@pytest.mark.skipif(not os.path.exists(SURROGATE), reason="trained surrogate not present")
def test_time_step_fp32_surrogate_vs_fp64_newton():
    """The benchmark's two paths on a 48x40x32 block of the reacting mixing
    layer over three time steps (six chemistry half-steps, nine RK3 stages),
    against FP64 plain (FP64 stage + FP64 Newton, which matches the oracle to
    1e-12), at the accuracy bars Y <= 1e-3 and T <= 1 K everywhere, for the
    benchmarked path (FP32 stage + surrogate) and for each of its two parts:
    the surrogate with the FP64 stage, the FP32 stage with Newton."""
    size, steps = (48, 40, 32), 3
    print(f"\n[driver] accuracy vs FP64 plain, size={size}, {steps} steps", flush=True)
    T0, Y0, _ = _run_steps(size, "fp64", False, steps)
    cells = size[0] * size[1] * size[2]
    for name, prec, nn in (("fp64 stage + surrogate", "fp64", True),
                           ("fp32 stage + Newton", "fp32", False),
                           ("fp32 stage + surrogate", "fp32", True)):
        T, Y, nfb = _run_steps(size, prec, nn, steps)
        dT, dY = (T - T0).abs(), (Y - Y0).abs()
        eT, eY = float(dT.max()), float(dY.max())
        print(f"  {name}: T linf {eT:.3f} K rms {float(dT.pow(2).mean().sqrt()):.4f} K, "
              f"Y linf {eY:.2e}, fallback {nfb / (2 * steps * cells):.3f}", flush=True)
        assert eT <= 1.0 and eY <= 1e-3, (name, eT, eY)
