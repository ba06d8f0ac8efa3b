"""The fused FP32 stage (per-direction sweeps into one right-hand side,
F3D_STAGE_FUSED) against the kernel-granular FP32 stage (the reference's
launch sequence, solver.jl:22-56) and against the FP64 oracle.

Both float paths evaluate the same face formulas; they differ in how the
face-flux differences of the three directions are summed, so one stage from
the same state agrees to float rounding of the increments.  Sizes cover
more than one 256-cell x segment (the segment's last face is evaluated by a
separate round), x extents that are not a multiple of the 64-lane marches,
and several march segments along y and z."""
import numpy as np
import pytest
import torch

from helpers import rel_err_per_var, to_np

pytestmark = pytest.mark.gpu


def _stage_pair(mech_name, N, scheme=None, seed=3, KRK=(1,)):
    import flame3d_amd.abi as abi
    import flame3d_amd.cases as cases
    import flame3d_amd.mech as mm
    m = mm.load(mech_name)
    grid = cases.skewed_grid(*N)
    Q, rhoi = cases.random_state(grid, m, seed=seed, shock=True)
    dt = cases.cfl_dt(grid, Q, 0.3)
    kinds = (abi.BC_COPY, abi.BC_COPY, abi.BC_COPY, abi.BC_COPY, abi.BC_PERIODIC,
             abi.BC_PERIODIC)
    out = {}
    for kind in (abi.STAGE_FUSED, abi.STAGE_GRANULAR):
        sc = dict(scheme or {})
        sc["stage_kind"] = kind
        ctx = abi.Context(*N, m, "fp32", scheme=sc)
        ctx.set_metrics([x.cuda() for x in grid.metrics], grid.J.cuda())
        ctx.set_boundary(kinds)
        Qg, rg = Q.cuda().clone(), rhoi.cuda().clone()
        U = ctx.zeros(5)
        ctx.prim2c(U, Qg)
        U0, r0 = U.clone(), rg.clone()
        Un, rn = ctx.zeros(5), ctx.zeros(m.Ns)
        for k in KRK:
            ctx.rk3_stage(k, U, Un, rg, rn, Qg, dt)
        torch.cuda.synchronize()
        out[kind] = (U - U0, rg - r0, Qg.clone())
        ctx.close()
    return out[abi.STAGE_FUSED], out[abi.STAGE_GRANULAR]


def _cmp(a, b, what, tol):
    errs = rel_err_per_var(to_np(a), to_np(b))
    print(f"  {what}: max rel err {max(errs):.2e}", flush=True)
    assert max(errs) <= tol, (what, errs)


@pytest.mark.parametrize("mech", ["air5", "h2_lidryer"])
@pytest.mark.parametrize("N", [(20, 12, 10), (300, 9, 7), (70, 70, 6), (13, 6, 70)])
def test_fused_stage_matches_granular(mech, N):
    import flame3d_amd.abi as abi
    print(f"\n[sweeps] {mech} N={N}", flush=True)
    f, g = _stage_pair(mech, N)
    _cmp(f[0], g[0], "dU", 1e-4)
    _cmp(f[1], g[1], "drho_k", 1e-4)
    _cmp(f[2], g[2], "Q", 1e-5)


@pytest.mark.parametrize("scheme", ["weno7", "ducros"])
def test_fused_stage_scheme_options(scheme):
    import flame3d_amd.abi as abi
    sc = {"recon_kind": abi.RECON_WENO7} if scheme == "weno7" else dict(abi.DUCROS_SCHEME)
    print(f"\n[sweeps] scheme={scheme}", flush=True)
    f, g = _stage_pair("air5", (40, 18, 14), scheme=sc)
    _cmp(f[0], g[0], "dU", 1e-4)
    _cmp(f[1], g[1], "drho_k", 1e-4)


def test_fused_three_stages_vs_oracle():
    """Three RK3 stages of the fused float stage against the FP64 oracle on
    the reacting mixing layer's states (the benchmark's mechanism)."""
    import flame3d_amd.abi as abi
    import flame3d_amd.cases as cases
    import flame3d_amd.mech as mm
    from oracle import Oracle
    m = mm.load("h2_lidryer")
    N = (24, 20, 12)
    case = cases.reacting_mixing_layer(N, m, dev="cpu")
    print(f"\n[sweeps] reacting mixing layer N={N} three stages vs oracle", flush=True)

    def f(t):
        return np.array(t.numpy().T, order="F", copy=True)

    orc = Oracle(*N, m)
    Qo, ro = f(case.Q), f(case.rhoi)
    M, J = [f(x) for x in case.grid.metrics], f(case.grid.J)
    Uo = orc.prim2c(Qo)
    Uno, rno, ws = orc.zeros(5), orc.zeros(m.Ns), orc.workspace()
    ctx = abi.Context(*N, m, "fp32")
    ctx.set_metrics([x.cuda() for x in case.grid.metrics], case.grid.J.cuda())
    ctx.set_boundary(case.bc)
    Qg, rg = case.Q.cuda(), case.rhoi.cuda()
    Ug = ctx.zeros(5)
    ctx.prim2c(Ug, Qg)
    Ung, rng_ = ctx.zeros(5), ctx.zeros(m.Ns)
    for k in (1, 2, 3):
        orc.rk3_stage(k, Uo, Uno, ro, rno, Qo, ws, M, J, case.dt, case.bc)
        ctx.rk3_stage(k, Ug, Ung, rg, rng_, Qg, case.dt)
    torch.cuda.synchronize()
    # errors against the magnitude of each variable's group: density,
    # momentum vector, energy; the partial densities against the density.
    # (The mixing layer's w is a 1 % perturbation, so a per-component norm
    # would divide the float rounding of the pressure flux by |rho w|.)
    Un_, Uo_ = to_np(Ug), Uo
    it = (slice(4, -4),) * 3
    Ui = Uo_[it]
    scales = [np.abs(Ui[..., 0]).max()] + [np.sqrt((Ui[..., 1:4] ** 2).sum(-1)).max()] * 3 + \
        [np.abs(Ui[..., 4]).max()]
    errs = [np.abs(Un_[it][..., n] - Ui[..., n]).max() / scales[n] for n in range(5)]
    print(f"  U: max err / group scale {max(errs):.2e} {['%.1e' % e for e in errs]}", flush=True)
    assert max(errs) <= 1e-5, errs
    er = np.abs(to_np(rg)[it] - ro[it]).max() / scales[0]
    print(f"  rhoi: max err / rho scale {er:.2e}", flush=True)
    assert er <= 1e-5
    errs = rel_err_per_var(to_np(Qg), Qo)
    print(f"  Q: max rel err {max(errs):.2e}", flush=True)
    assert max(errs[:2] + errs[4:]) <= 1e-4, errs
    dT = np.abs(to_np(Qg)[4:-4, 4:-4, 4:-4, 5] - Qo[4:-4, 4:-4, 4:-4, 5]).max()
    print(f"  T max diff {dT:.2e} K", flush=True)
    assert dT < 0.05
    ctx.close()
