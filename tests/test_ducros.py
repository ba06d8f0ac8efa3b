"""Ducros sensor theta = (div u)^2 / ((div u)^2 + |curl u|^2 + 1e-30) (the
switch of the shock-laden mixing-layer workload: UP7 below theta = 0.1,
WENO5-Z above; the reference itself switches on its pressure sensor,
schemes.jl:1-24).  Oracle closed forms on linear velocity fields (central
differences and constant metrics are exact there), then the GPU sensor and
a whole RK3 stage against the oracle."""
import numpy as np
import pytest
import torch

from helpers import rel_err, rel_err_per_var, to_np, to_torch

NG = 4


def _linear_state(grid, A):
    """Q with velocity u = A (x, y, z) on every padded point."""
    X = torch.stack([grid.x, grid.y, grid.z])
    Q = torch.zeros((7,) + tuple(grid.x.shape), dtype=torch.float64)
    Q[0] = 1.0
    for c in range(3):
        Q[1 + c] = sum(A[c][j] * X[j] for j in range(3))
    Q[4] = 1e5
    Q[5] = 300.0
    return Q


def _theta(A):
    A = np.asarray(A, dtype=np.float64)
    dv = np.trace(A)
    w = np.array([A[2, 1] - A[1, 2], A[0, 2] - A[2, 0], A[1, 0] - A[0, 1]])
    return dv * dv / (dv * dv + w @ w + 1e-30)


@pytest.mark.parametrize("A", [
    [[1.0, 0, 0], [0, 1.0, 0], [0, 0, 1.0]],          # pure dilatation: theta = 1
    [[0, -1.0, 0], [1.0, 0, 0], [0, 0, 0]],           # solid rotation: theta = 0
    [[0.3, -0.7, 0.2], [0.5, -0.1, 0.9], [-0.4, 0.6, 0.25]],  # general linear field
])
def test_ducros_oracle_linear_fields(A):
    import flame3d_amd.cases as cases
    import flame3d_amd.mech as mm
    from oracle import Oracle
    m = mm.load("air5")
    grid = cases.uniform_grid(12, 10, 8, 1.2, 0.7, 0.5, dev="cpu", periodic=False)
    Q = _linear_state(grid, A)
    orc = Oracle(12, 10, 8, m)
    phi = orc.ducros_sensor(to_np(Q), [to_np(x) for x in grid.metrics], to_np(grid.J))
    inner = phi[1:-1, 1:-1, 1:-1]  # the points the sensor covers
    assert np.max(np.abs(inner - _theta(A))) < 1e-12
    assert np.all(phi[0] == 0) and np.all(phi[-1] == 0)


@pytest.mark.gpu
def test_ducros_gpu_matches_oracle_and_stage():
    """Skewed grid, smooth random velocity: the sensor to 1e-12 (FP64) and
    1e-6 (FP32 path, stored in float); one RK3 stage of the non-reacting air
    mixing layer with the Ducros switch against the oracle to 1e-12."""
    import flame3d_amd.abi as abi
    import flame3d_amd.cases as cases
    import flame3d_amd.mech as mm
    from oracle import Oracle
    print("\n[ducros] sensor + stage parity", flush=True)
    m = mm.load("air5")
    N = (16, 12, 10)
    grid = cases.skewed_grid(*N)
    g = torch.Generator().manual_seed(7)
    Q = _linear_state(grid, [[0.2, 0.1, 0], [0, -0.3, 0.4], [0.1, 0, 0.1]])
    Q[1:4] += 0.05 * torch.randn((3,) + tuple(Q.shape[1:]), generator=g, dtype=torch.float64)
    orc = Oracle(*N, m)
    ref = orc.ducros_sensor(to_np(Q), [to_np(x) for x in grid.metrics], to_np(grid.J))
    for prec, tol in (("fp64", 1e-12), ("fp32", 1e-6)):
        ctx = abi.Context(*N, m, prec, scheme=abi.DUCROS_SCHEME)
        with pytest.raises(abi.F3DError):  # the Ducros sensor needs the metrics
            ctx.shock_sensor(ctx.rzeros(), Q.cuda())
        ctx.set_metrics([x.cuda() for x in grid.metrics], grid.J.cuda())
        phi = ctx.rzeros()
        ctx.shock_sensor(phi, Q.cuda())
        torch.cuda.synchronize()
        e = rel_err(to_np(phi.double()), ref)
        print(f"  {prec} sensor rel err {e:.2e}", flush=True)
        assert e <= tol
        ctx.close()
    # one stage of the non-reacting mixing layer with the Ducros switch
    case = cases.mixing_layer((16, 16, 12), m, dev="cpu")
    n3 = (16, 16, 12)
    sch = {"phi1": 0.1, "phi2": 2.0, "sensor_kind": 1}
    orc = Oracle(*n3, m, scheme=sch)
    Qo, ro = to_np(case.Q), to_np(case.rhoi)
    M, J = [to_np(x) for x in case.grid.metrics], to_np(case.grid.J)
    Uo = orc.prim2c(Qo)
    Uno, rno, ws = orc.zeros(5), orc.zeros(m.Ns), orc.workspace()
    orc.rk3_stage(1, Uo, Uno, ro, rno, Qo, ws, M, J, case.dt, case.bc)
    phio = ws[4]
    ctx = abi.Context(*n3, m, "fp64", scheme=abi.DUCROS_SCHEME)
    ctx.set_metrics([x.cuda() for x in case.grid.metrics], case.grid.J.cuda())
    ctx.set_boundary(case.bc)
    Qg, rg = case.Q.cuda(), case.rhoi.cuda()
    Ug = ctx.zeros(5)
    ctx.prim2c(Ug, Qg)
    ctx.rk3_stage(1, Ug, ctx.zeros(5), rg, ctx.zeros(m.Ns), Qg, case.dt)
    torch.cuda.synchronize()
    frac = float(np.mean(phio[1:-1, 1:-1, 1:-1] >= 0.1))
    print(f"  stage: theta >= 0.1 on {frac:.3f} of the points", flush=True)
    assert 0.0 < frac < 1.0  # both branches are exercised
    for name, got, refv in (("U", Ug, Uo), ("rhoi", rg, ro), ("Q", Qg, Qo)):
        errs = rel_err_per_var(to_np(got), refv)
        print(f"  {name} {['%.1e' % x for x in errs]}", flush=True)
        assert max(errs) <= 1e-12
    ctx.close()
