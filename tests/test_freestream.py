"""Free-stream preservation on the reference mesh generator's grids
(parse_mesh.jl:24-50 coordinates, :132-221 metrics).

A uniform state has no viscous or diffusive fluxes, its sensor is zero (the
UP7 branch everywhere) and, when it is supersonic along every grid
direction, F- = 0 and F+ is the physical flux through the cofactor metrics.
One stage then changes U by exactly the metric-identity residual
  dU_n = J dt sum_d sum_x f_{n,x} (D_d M_{d,x}),
D_d the UP7 face-flux difference along direction d applied to the metric
array M_{d,x} = d xi_d / d x (J^-1-weighted cofactors), f_{n,x} the uniform
physical flux of variable n along x.  On the generator's own grid (x, y, z
each a function of one index) the cofactors are constant along their
direction, so the residual and the stage's change are exactly zero; on the
same grid with a smooth wave added to the nodes the residual is non-zero and
the oracle's stage equals it.  GPU: the FP64 stage against the oracle, the
fused float stage against the closed form."""
import numpy as np
import pytest
import torch

from helpers import to_np, to_torch

NG = 4
UP7 = (0.0, 17 / 600, -23 / 120, 22 / 30, 0.5, -0.075, 0.005)


def grid(N, wavy):
    import flame3d_amd.cases as cases
    from oracle import mesh
    x, y, z = [c.numpy() for c in cases.parse_mesh_coords(*N, n_uniform=N[0] - 6)]
    if wavy:
        # node displacements of a fraction of the local spacing (the y and z
        # spacings vary by two orders of magnitude across the clustering)
        K, J, I = np.meshgrid(*[np.arange(s, dtype=float) for s in x.shape], indexing="ij")
        it = (slice(NG, -NG),) * 3
        # tapered to zero at the boundary nodes, so the reflected ghosts stay
        # an admissible (non-folding) continuation
        hx, hy, hz = np.gradient(x, axis=2), np.gradient(y, axis=1), np.gradient(z, axis=0)
        w = 1.0
        for A, n in ((I, N[0]), (J, N[1]), (K, N[2])):
            w = w * np.sin(np.pi * np.clip((A - NG) / (n - 1), 0, 1)) ** 2
        x[it] += (0.2 * w * hx * np.sin(2 * np.pi * J / N[1]) * np.cos(2 * np.pi * K / N[2]))[it]
        y[it] += (0.15 * w * hy * np.sin(2 * np.pi * I / N[0] + 0.5 * K))[it]
        z[it] += (0.15 * w * hz * np.cos(2 * np.pi * I / N[0]) * np.sin(0.7 * J))[it]
    mesh.ghosts(x, y, z)
    M, Jm = mesh.metrics(x, y, z)
    f = lambda a: np.asfortranarray(a.T)  # [k, j, i] -> oracle [i, j, k]
    return [f(m) for m in M], f(Jm)


def uniform_state(orc, mech, V=5000.0):
    import flame3d_amd.cases as cases
    sh = orc.shape[::-1]
    one = torch.ones(sh, dtype=torch.float64)
    Y = torch.zeros((mech.Ns,) + sh, dtype=torch.float64)
    Y[mech.index("O2")], Y[mech.index("N2")] = 0.233, 0.767
    Q, r = cases.state_from_TPY(mech, 300.0 * one, 1e5 * one, Y, V * one, V * one, V * one)
    return to_np(Q), to_np(r)


def physical_flux(Q, rhoi):
    """f[n][x] of the uniform state (rho u_x, rho u_k u_x + p d_kx, (E + p) u_x)
    and rho_k u_x."""
    q = Q[0, 0, 0]
    rho, u, p, e = q[0], q[1:4], q[4], q[6]
    E = e + 0.5 * rho * (u @ u)
    f = [rho * u] + [rho * u[k] * u + p * np.eye(3)[k] for k in range(3)] + [(E + p) * u]
    fs = [rhoi[0, 0, 0, n] * u for n in range(rhoi.shape[-1])]
    return np.array(f), np.array(fs)


def closed_form(M, J, f, dt, N):
    """J dt sum_d D_d (sum_x f_x M_{d,x}) over the interior, per variable."""
    out = np.zeros(tuple(N) + (f.shape[0],))
    for d in range(3):
        for n in range(f.shape[0]):
            G = sum(f[n, x] * M[3 * d + x] for x in range(3))
            G = np.moveaxis(G, d, 0)
            nd = N[d]
            face = sum(UP7[m] * G[m:m + nd + 1] for m in range(7))  # faces 0..nd
            diff = np.moveaxis(face[:-1] - face[1:], 0, d)
            sl = [slice(NG, NG + N[a]) for a in range(3)]
            sl[d] = slice(None)
            out[..., n] += diff[tuple(sl)]
    return out * (J[NG:-NG, NG:-NG, NG:-NG] * dt)[..., None]


def supersonic_everywhere(M, Q):
    q = Q[0, 0, 0]
    c = np.sqrt((q[4] / q[6] + 1) * q[4] / q[0])
    for d in range(3):
        un = sum(M[3 * d + x] * q[1 + x] for x in range(3))
        mag = np.sqrt(sum(M[3 * d + x] ** 2 for x in range(3)))
        if not np.all(un - c * mag > 0):
            return False
    return True


def oracle_stage(N, wavy, dt=1e-9):
    import flame3d_amd.mech as mm
    from oracle import Oracle
    m = mm.load("air5")
    orc = Oracle(*N, m)
    M, J = grid(N, wavy)
    Q, r = uniform_state(orc, m)
    assert supersonic_everywhere(M, Q)
    U = orc.prim2c(Q)
    U0, r0 = U.copy(order="F"), r.copy(order="F")
    Qs, rs = Q.copy(order="F"), r.copy(order="F")
    orc.rk3_stage(1, U, orc.zeros(5), rs, orc.zeros(m.Ns), Qs, orc.workspace(), M, J, dt, (0,) * 6)
    it = (slice(NG, -NG),) * 3
    f, fs = physical_flux(Q, r)
    ref = closed_form(M, J, f, dt, N)
    refs = closed_form(M, J, fs, dt, N)
    return dict(m=m, M=M, J=J, Q=Q, r=r, dt=dt, dU=(U - U0)[it], dr=(rs - r0)[it], ref=ref,
                refs=refs, f=f)


N_GEN = (40, 24, 20)


def flux_scale(s):
    """The change one cell's face flux alone would make in a stage."""
    return np.abs(s["f"]).max() * np.abs(s["J"]).max() * s["dt"] * np.abs(s["M"][0]).max()


def test_generator_mesh_preserves_uniform_flow():
    """The generator's grid keeps a uniform flow exactly: the stage changes
    no bit of U or rho_k (the closed-form residual is zero up to the rounding
    of the cofactors' zero entries, ~1e-21 of the diagonal ones)."""
    s = oracle_stage(N_GEN, wavy=False)
    assert np.abs(s["ref"]).max() < 1e-12 * flux_scale(s)
    assert np.abs(s["dU"]).max() == 0.0, np.abs(s["dU"]).max()
    assert np.abs(s["dr"]).max() == 0.0


def test_wavy_mesh_change_is_the_metric_identity_residual():
    s = oracle_stage(N_GEN, wavy=True)
    ref, dU = s["ref"], s["dU"]
    # a non-trivial residual, small against the flux through one cell
    scale = flux_scale(s)
    assert 1e-8 * scale < np.abs(ref).max() < 0.1 * scale
    for n in range(5):
        err = np.abs(dU[..., n] - ref[..., n]).max() / np.abs(ref[..., n]).max()
        assert err < 1e-9, (n, err)
    err = np.abs(s["dr"] - s["refs"]).max() / np.abs(s["refs"]).max()
    assert err < 1e-9


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_freestream_gpu(prec):
    """FP64: the stage's change equals the oracle's to 1e-12 of the residual;
    the fused float stage: within float rounding of the fluxes it differences
    (1e-6 of the flux scale) of the closed-form residual, and exactly zero on
    the unperturbed generator mesh."""
    import flame3d_amd.abi as abi
    for wavy in (False, True):
        s = oracle_stage(N_GEN, wavy)
        print(f"\n[freestream] {prec} wavy={wavy} N={N_GEN}", flush=True)
        m = s["m"]
        ctx = abi.Context(*N_GEN, m, prec)
        ctx.set_metrics([to_torch(x) for x in s["M"]], to_torch(s["J"]))
        ctx.set_boundary((abi.BC_KEEP,) * 6)
        Qg, rg = to_torch(s["Q"]), to_torch(s["r"])
        U = ctx.zeros(5)
        ctx.prim2c(U, Qg)
        U0, r0 = U.clone(), rg.clone()
        ctx.rk3_stage(1, U, ctx.zeros(5), rg, ctx.zeros(m.Ns), Qg, s["dt"])
        torch.cuda.synchronize()
        it = (slice(NG, -NG),) * 3
        dU = to_np(U - U0)[it]
        dr = to_np(rg - r0)[it]
        ctx.close()
        if not wavy:
            print(f"  max |dU| {np.abs(dU).max():.3e}", flush=True)
            assert np.abs(dU).max() == 0.0 and np.abs(dr).max() == 0.0
            continue
        if prec == "fp64":
            # the residual is what is left after the face fluxes (of size
            # flux_scale) cancel, so rounding-level differences in those
            # fluxes are measured against the flux scale, not the residual
            scale = flux_scale(s)
            for n in range(5):
                d = np.abs(dU[..., n] - s["dU"][..., n]).max()
                e = d / scale
                print(f"  var {n}: |GPU - oracle| / flux scale {e:.2e} "
                      f"(/ residual {d / np.abs(s['dU'][..., n]).max():.2e})", flush=True)
                assert e < 1e-12
        else:
            scale = flux_scale(s)
            for n in range(5):
                e = np.abs(dU[..., n] - s["ref"][..., n]).max() / scale
                print(f"  var {n}: err vs closed form / flux scale {e:.2e}", flush=True)
                assert e < 1e-5
