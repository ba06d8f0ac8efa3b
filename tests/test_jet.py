"""The hot supersonic air jet (5-species air at 5000 K, Ma 1.5, into 300 K
still air at 3596 Pa; WENO7-Z in the middle branch of the hybrid switch;
Newton air chemistry): the case itself on the CPU, then one time step of the
FP64 GPU path against the oracle driven through the same sequence, and the
FP32 stage against the FP64 one."""
import math

import numpy as np
import pytest
import torch

from helpers import rel_err_per_var, to_np, to_torch

SIZE = (24, 16, 16)


def _case(dev="cpu"):
    import flame3d_amd.cases as cases
    import flame3d_amd.mech as mm
    m = mm.load("air5")
    return m, cases.reacting_jet(SIZE, m, dev=dev)


def test_jet_case():
    import flame3d_amd.abi as abi
    import flame3d_amd.cases as cases
    m, case = _case()
    q = case.Q[:, 4:-4, 4:-4, 4:-4]
    T, u, p = q[5], q[1], q[4]
    assert abs(float(T.max()) - 5000.0) < 1.0 and abs(float(T.min()) - 300.0) < 1e-9
    assert torch.allclose(p, torch.full_like(p, 3596.0))
    # the jet core moves at Mach 1.5 of its own sound speed (2 % noise)
    Tj = torch.tensor([5000.0], dtype=torch.float64)
    Yj = torch.zeros((m.Ns, 1), dtype=torch.float64)
    Yj[m.index("O2")], Yj[m.index("N2")] = 0.233, 0.767
    W = 1.0 / (0.233 / m.mw[m.index("O2")] + 0.767 / m.mw[m.index("N2")])
    R = m.Ru / W
    cp = float(cases.cp_mix(m, Tj, Yj)[0])
    c = math.sqrt(cp / (cp - R) * R * 5000.0)
    core = T > 4999.0
    assert abs(float(u[core].mean()) / (1.5 * c) - 1) < 0.01
    assert float(u[T < 301.0].abs().max()) < 1e-4 * 1.5 * c  # the tanh edge's tail
    assert case.bc[0] == abi.BC_KEEP and case.dt > 0


def _oracle_step(orc, case, U, Un, r, rn, Q, ws, M, J):
    def half():
        orc.chem_newton(U, Q, r, case.dt / 2)
        orc.fill(case.bc, Q, U, r)
    half()
    for k in (1, 2, 3):
        orc.rk3_stage(k, U, Un, r, rn, Q, ws, M, J, case.dt, case.bc)
    half()


@pytest.mark.gpu
def test_jet_time_step_parity():
    import flame3d_amd.abi as abi
    from oracle import Oracle
    print(f"\n[jet] one time step, WENO7-Z + Newton air chemistry, size={SIZE}", flush=True)
    m, case = _case()
    orc = Oracle(*SIZE, m, scheme=dict(recon_kind=1))
    Q, r = to_np(case.Q), to_np(case.rhoi)
    M, J = [to_np(x) for x in case.grid.metrics], to_np(case.grid.J)
    U = np.asfortranarray(orc.prim2c(Q))
    Un, rn, ws = orc.zeros(5), orc.zeros(m.Ns), orc.workspace()
    _oracle_step(orc, case, U, Un, r, rn, Q, ws, M, J)
    res = {}
    for prec in ("fp64", "fp32"):
        ctx = abi.Context(*SIZE, m, prec, scheme=dict(recon_kind=abi.RECON_WENO7))
        ctx.set_metrics([x.cuda() for x in case.grid.metrics], case.grid.J.cuda())
        ctx.set_boundary(case.bc)
        Qg, rg = case.Q.cuda(), case.rhoi.cuda()
        Ug = ctx.zeros(5)
        ctx.prim2c(Ug, Qg)
        ctx.time_step(Ug, ctx.zeros(5), rg, ctx.zeros(m.Ns), Qg, case.dt, abi.CHEM_NEWTON)
        torch.cuda.synchronize()
        assert not ctx.nan_check(Ug)
        res[prec] = (to_np(Ug), to_np(rg), to_np(Qg))
        ctx.close()
    for name, got, ref in zip(("U", "rhoi", "Q"), res["fp64"], (U, r, Q)):
        errs = rel_err_per_var(got, ref)
        print(f"  fp64 {name} {['%.1e' % e for e in errs]}", flush=True)
        assert max(errs) <= 1e-12, (name, errs)
    # FP32 stage (FP64 chemistry) against the FP64 path after the step
    sl = (slice(4, -4),) * 3
    dT = np.max(np.abs(res["fp32"][2][sl + (5,)] - res["fp64"][2][sl + (5,)]))
    print(f"  fp32 vs fp64 T max diff {dT:.3e} K", flush=True)
    assert dT < 1.0
