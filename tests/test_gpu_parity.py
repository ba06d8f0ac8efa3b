"""GPU parity of every stage kernel against the FP64 oracle.

Each kernel is driven through the C-ABI with the same inputs as the oracle
(inputs of downstream kernels are taken from the oracle, so a failure points
at one kernel), on a small skewed curvilinear block with a shocked random
state, for the 5-species air and the 9-species H2 mechanisms.

Bar: the FP64 path matches the oracle to 1e-12 normwise relative error per
output variable.  The FP32 path is checked to 2e-4 normwise relative error
(float arithmetic on the same inputs; the full-stage FP32 errors are reported
in the benchmark).
"""
import numpy as np
import pytest
import torch

from helpers import rel_err, rel_err_per_var, to_np, to_torch

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32 = 2e-4
SIZE = (20, 12, 10)


@pytest.fixture(scope="module", params=["air5", "h2_lidryer"])
def S(request):
    import flame3d_amd.abi as abi
    import flame3d_amd.cases as cases
    import flame3d_amd.mech as mm
    from oracle import Oracle

    m = mm.load(request.param)
    Nx, Ny, Nz = SIZE
    grid = cases.skewed_grid(Nx, Ny, Nz)
    Q, rhoi = cases.random_state(grid, m, seed=7, shock=True)
    print(f"\n[parity] mech={request.param} size={SIZE}", flush=True)
    st = dict(m=m, grid=grid, Qt=Q, rhoit=rhoi, orc=Oracle(Nx, Ny, Nz, m),
              c64=abi.Context(Nx, Ny, Nz, m, "fp64"), c32=abi.Context(Nx, Ny, Nz, m, "fp32"),
              Q=to_np(Q), rhoi=to_np(rhoi), M=[to_np(x) for x in grid.metrics], J=to_np(grid.J))
    st["Md"] = [x.cuda() for x in grid.metrics]
    st["Jd"] = grid.J.cuda()
    yield st
    st["c64"].close()
    st["c32"].close()


def _ctx(S, prec):
    return S["c64"] if prec == "fp64" else S["c32"]


def _tol(prec):
    return TOL64 if prec == "fp64" else TOL32


def _r(ctx, a):
    return to_torch(a, dtype=ctx.rdtype)


def _check(got, ref, prec, what):
    errs = rel_err_per_var(got, ref) if ref.ndim == 4 else [rel_err(got, ref)]
    print(f"  {what} [{prec}] max rel err {max(errs):.2e}", flush=True)
    assert max(errs) <= _tol(prec), (what, errs)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("d", [0, 1, 2])
def test_flux_split(S, prec, d):
    ctx, orc = _ctx(S, prec), S["orc"]
    M = S["M"][3 * d:3 * d + 3]
    Fp_r, Fm_r = orc.flux_split(S["Q"], *M)
    Fp, Fm = ctx.rzeros(5), ctx.rzeros(5)
    ctx.flux_split(S["Qt"].cuda(), Fp, Fm, *S["Md"][3 * d:3 * d + 3])
    _check(to_np(Fp), Fp_r, prec, f"fluxSplit d={d} F+")
    _check(to_np(Fm), Fm_r, prec, f"fluxSplit d={d} F-")


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_species_split(S, prec):
    ctx, orc = _ctx(S, prec), S["orc"]
    Fp_r, Fm_r = orc.species_split(S["rhoi"], S["Q"], *S["M"][3:6])
    Fp, Fm = ctx.rzeros(ctx.Ns), ctx.rzeros(ctx.Ns)
    ctx.species_split(S["rhoit"].cuda(), S["Qt"].cuda(), Fp, Fm, *S["Md"][3:6])
    _check(to_np(Fp), Fp_r, prec, "split F+")
    _check(to_np(Fm), Fm_r, prec, "split F-")


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_shock_sensor(S, prec):
    ctx, orc = _ctx(S, prec), S["orc"]
    ref = orc.shock_sensor(S["Q"])
    phi = ctx.rzeros()
    ctx.shock_sensor(phi, S["Qt"].cuda())
    _check(to_np(phi), ref, prec, "shockSensor")


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("d", [0, 1, 2])
def test_weno(S, prec, d):
    ctx, orc = _ctx(S, prec), S["orc"]
    M = S["M"][3 * d:3 * d + 3]
    phi = orc.shock_sensor(S["Q"])
    Fp, Fm = orc.flux_split(S["Q"], *M)
    ref = orc.weno(d, phi, Fp, Fm, 5)
    # branch coverage: sensor values on the faces of this sweep
    n1 = np.mean(phi < 1e-3)
    n2 = np.mean((phi >= 1e-3) & (phi < 0.1))
    print(f"  sensor fractions UP7 {n1:.2f} WENO {n2:.2f} NND {1 - n1 - n2:.2f}")
    assert n1 > 0.01 and n2 > 0.01 and 1 - n1 - n2 > 0.01
    F = ctx.face_zeros(d, 5)
    ctx.weno(d, F, _r(ctx, phi), _r(ctx, Fp), _r(ctx, Fm), 5)
    _check(to_np(F), ref, prec, f"WENO d={d}")
    # species sweep with NV = Ns
    Fpi, Fmi = orc.species_split(S["rhoi"], S["Q"], *M)
    refi = orc.weno(d, phi, Fpi, Fmi, ctx.Ns)
    Fi = ctx.face_zeros(d, ctx.Ns)
    ctx.weno(d, Fi, _r(ctx, phi), _r(ctx, Fpi), _r(ctx, Fmi), ctx.Ns)
    _check(to_np(Fi), refi, prec, f"WENO species d={d}")


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_mixture(S, prec):
    ctx, orc = _ctx(S, prec), S["orc"]
    Yr, lr, mr, Dr = orc.mixture(S["Q"], S["rhoi"])
    Yi, lam, mu, D = ctx.rzeros(ctx.Ns), ctx.rzeros(), ctx.rzeros(), ctx.rzeros(ctx.Ns)
    ctx.mixture(S["Qt"].cuda(), S["rhoit"].cuda(), Yi, lam, mu, D)
    _check(to_np(Yi), Yr, prec, "mixture Yi")
    _check(to_np(lam), lr, prec, "mixture lambda")
    _check(to_np(mu), mr, prec, "mixture mu")
    _check(to_np(D), Dr, prec, "mixture D")


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("d", [0, 1, 2])
def test_viscous(S, prec, d):
    ctx, orc = _ctx(S, prec), S["orc"]
    Yr, lr, mr, Dr = orc.mixture(S["Q"], S["rhoi"])
    Fd_r, Fh_r = orc.spec_viscous_flux(d, S["Q"], Yr, S["M"], S["J"], Dr)
    Fv_r = orc.viscous_flux(d, S["Q"], S["M"], S["J"], lr, mr, Fh_r)
    Fd, Fh = ctx.face_zeros(d, ctx.Ns), ctx.face_zeros(d, 3)
    ctx.spec_viscous_flux(d, Fd, S["Qt"].cuda(), _r(ctx, Yr), S["Md"], S["Jd"], _r(ctx, Dr), Fh)
    _check(to_np(Fd), Fd_r, prec, f"specViscousFlux d={d} Fd")
    _check(to_np(Fh), Fh_r, prec, f"specViscousFlux d={d} Fh")
    Fv = ctx.face_zeros(d, 4)
    ctx.viscous_flux(d, Fv, S["Qt"].cuda(), S["Md"], S["Jd"], _r(ctx, lr), _r(ctx, mr),
                     _r(ctx, Fh_r))
    _check(to_np(Fv), Fv_r, prec, f"viscousFlux d={d}")


def _faces(S, ctx):
    orc = S["orc"]
    phi = orc.shock_sensor(S["Q"])
    Yr, lr, mr, Dr = orc.mixture(S["Q"], S["rhoi"])
    F, Fi, Fv, Fd = [], [], [], []
    for d in range(3):
        M = S["M"][3 * d:3 * d + 3]
        Fp, Fm = orc.flux_split(S["Q"], *M)
        F.append(orc.weno(d, phi, Fp, Fm, 5))
        Fpi, Fmi = orc.species_split(S["rhoi"], S["Q"], *M)
        Fi.append(orc.weno(d, phi, Fpi, Fmi, ctx.Ns))
        fd, fh = orc.spec_viscous_flux(d, S["Q"], Yr, S["M"], S["J"], Dr)
        Fd.append(fd)
        Fv.append(orc.viscous_flux(d, S["Q"], S["M"], S["J"], lr, mr, fh))
    return F, Fi, Fv, Fd


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_div(S, prec):
    ctx, orc = _ctx(S, prec), S["orc"]
    F, Fi, Fv, Fd = _faces(S, ctx)
    dt = 1e-6  # increments of a few % of U, so the comparison is not roundoff-bound
    U0 = orc.prim2c(S["Q"])
    U_r = U0.copy(order="F")
    orc.div(U_r, *F, *Fv, dt, S["J"])
    U = to_torch(U0)
    ctx.div(U, *[_r(ctx, f) for f in F], *[_r(ctx, f) for f in Fv], dt, S["Jd"])
    # compare the increments (U itself is dominated by the unchanged part)
    _check(to_np(U) - U0, U_r - U0, prec, "div dU")
    r_r = S["rhoi"].copy(order="F")
    orc.div_specs(r_r, *Fi, *Fd, dt, S["J"])
    r = S["rhoit"].cuda().clone()
    ctx.div_specs(r, *[_r(ctx, f) for f in Fi], *[_r(ctx, f) for f in Fd], dt, S["Jd"])
    _check(to_np(r) - S["rhoi"], r_r - S["rhoi"], prec, "divSpecs drho")


def test_c2prim_prim2c_lincomb_gety(S):
    ctx, orc = S["c64"], S["orc"]
    U_r = orc.prim2c(S["Q"])
    U = ctx.zeros(5)
    ctx.prim2c(U, S["Qt"].cuda())
    _check(to_np(U), U_r, "fp64", "prim2c")
    # perturb U and rho_i, then c2Prim (GetT Newton in FP64)
    rng = np.random.default_rng(3)
    U_p = U_r * (1 + 1e-3 * rng.standard_normal(U_r.shape))
    U_p = np.asfortranarray(U_p)
    r_p = np.asfortranarray(S["rhoi"] * (1 + 1e-3 * rng.standard_normal(S["rhoi"].shape)))
    Qo, Uo, ro = S["Q"].copy(order="F"), U_p.copy(order="F"), r_p.copy(order="F")
    orc.c2prim(Uo, Qo, ro)
    Qg, Ug, rg = S["Qt"].cuda().clone(), to_torch(U_p), to_torch(r_p)
    ctx.c2prim(Ug, Qg, rg)
    _check(to_np(Qg), Qo, "fp64", "c2Prim Q")
    _check(to_np(rg), ro, "fp64", "c2Prim rhoi")
    # linComb
    Ul = U_p.copy(order="F")
    orc.lincomb(Ul, U_r, 5, 2.0 / 3.0, 1.0 / 3.0)
    Ulg = to_torch(U_p)
    ctx.lincomb(Ulg, to_torch(U_r), 5, 2.0 / 3.0, 1.0 / 3.0)
    _check(to_np(Ulg), Ul, "fp64", "linComb")
    # getY
    Yr = orc.get_y(S["rhoi"], S["Q"])
    Y = ctx.zeros(ctx.Ns)
    ctx.get_y(Y, S["rhoit"].cuda(), S["Qt"].cuda())
    _check(to_np(Y), Yr, "fp64", "getY")


@pytest.mark.parametrize("kinds", [(1, 1, 1, 1, 1, 1), (3, 2, 2, 2, 1, 1), (2, 0, 0, 3, 2, 2)])
def test_fill(S, kinds):
    import flame3d_amd.abi as abi
    ctx, orc = S["c64"], S["orc"]
    m = S["m"]
    inQ = [1.2, 300.0, 1.0, -2.0, 1.0e5, 310.0, 2.0e5]
    inr = list(np.linspace(0.1, 0.2, m.Ns))
    U0 = orc.prim2c(S["Q"])
    Qo, Uo, ro = S["Q"].copy(order="F"), U0.copy(order="F"), S["rhoi"].copy(order="F")
    orc.fill(kinds, Qo, Uo, ro, inQ, inr)
    ctx.set_boundary(kinds, inQ, inr)
    Qg, Ug, rg = S["Qt"].cuda().clone(), to_torch(U0), S["rhoit"].cuda().clone()
    ctx.fill(Qg, Ug, rg)
    assert np.array_equal(to_np(Qg), Qo)
    assert np.array_equal(to_np(Ug), Uo)
    assert np.array_equal(to_np(rg), ro)
    ctx.set_boundary((abi.BC_KEEP,) * 6)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_rk3_step(S, prec):
    """Three stages of solver.jl:298-326 through f3d_rk3_stage vs the oracle."""
    import flame3d_amd.cases as cases
    ctx, orc = _ctx(S, prec), S["orc"]
    m = S["m"]
    kinds = (2, 2, 2, 2, 1, 1)
    dt = cases.cfl_dt(S["grid"], S["Qt"], 0.3)
    Q, r = S["Q"].copy(order="F"), S["rhoi"].copy(order="F")
    U = orc.prim2c(Q)
    Un, rn = orc.zeros(5), orc.zeros(m.Ns)
    ws = orc.workspace()
    Qg, rg = S["Qt"].cuda().clone(), S["rhoit"].cuda().clone()
    Ug = to_torch(U)
    Ung, rng_ = ctx.zeros(5), ctx.zeros(m.Ns)
    ctx.set_metrics(S["Md"], S["Jd"])
    ctx.set_boundary(kinds)
    for KRK in (1, 2, 3):
        orc.rk3_stage(KRK, U, Un, r, rn, Q, ws, S["M"], S["J"], dt, kinds)
        ctx.rk3_stage(KRK, Ug, Ung, rg, rng_, Qg, dt)
    torch.cuda.synchronize()
    tol = TOL64 if prec == "fp64" else 1e-4
    for name, got, ref in (("U", Ug, U), ("rhoi", rg, r), ("Q", Qg, Q)):
        errs = rel_err_per_var(to_np(got), ref)
        print(f"  rk3 step {name} [{prec}] max rel err {max(errs):.2e}", flush=True)
        assert max(errs) <= tol, (name, errs)
    ctx.set_boundary((0,) * 6)


@pytest.mark.parametrize("d", [0, 1, 2])
def test_weno_fp32_flat_stencil_next_to_jump(S, d):
    """Regression: four equal stencil values next to a jump make one WENO5-Z
    smoothness indicator exactly 0, and (tau / (1e-26 + 0))^2 overflows float.
    The FP32 weights must stay finite and match the FP64 oracle."""
    ctx, orc = S["c32"], S["orc"]
    Fp, Fm = orc.zeros(5), orc.zeros(5)
    ax = (slice(None),) * d
    for n in range(5):
        Fp[..., n] = 0.0646 + 0.01 * n
        Fm[..., n] = 0.0411 + 0.01 * n
    jump = [slice(None)] * 3
    jump[d] = slice(9, None)
    Fp[tuple(jump) + (slice(None),)] += 0.02
    Fm[tuple(jump) + (slice(None),)] -= 0.015
    phi = orc.zeros() + 0.0046  # WENO branch everywhere
    ref = orc.weno(d, phi, Fp, Fm, 5)
    F = ctx.face_zeros(d, 5)
    ctx.weno(d, F, _r(ctx, phi), _r(ctx, Fp), _r(ctx, Fm), 5)
    got = to_np(F)
    assert np.isfinite(got).all()
    _check(got, ref, "fp32", f"WENO flat-next-to-jump d={d}")
