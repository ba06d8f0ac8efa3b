"""Pins of the FP64 oracle by closed forms, invariants, brute force and
finite differences (no retyped formulas): each check is chosen so that a
dropped term, a wrong sign, a shifted stencil index or a transposed metric
would fail it."""
import math

import numpy as np
import pytest
import torch

from helpers import to_np

NG = 4


@pytest.fixture(scope="module")
def air():
    import flame3d_amd.mech as mm
    return mm.load("air5")


@pytest.fixture(scope="module")
def h2():
    import flame3d_amd.mech as mm
    return mm.load("h2_lidryer")


def affine_metrics(N, A):
    """Metrics of the affine grid x = A (i, j, k): constant, all nine non-zero."""
    import flame3d_amd.cases as cases
    sh = (N[2] + 2 * NG, N[1] + 2 * NG, N[0] + 2 * NG)
    c = lambda v: torch.full(sh, float(v), dtype=torch.float64)
    m, J = cases.metrics_from_jacobian((c(A[0][0]), c(A[0][1]), c(A[0][2])),
                                       (c(A[1][0]), c(A[1][1]), c(A[1][2])),
                                       (c(A[2][0]), c(A[2][1]), c(A[2][2])))
    return [to_np(x) for x in m], to_np(J)


A_SKEW = [[1.0e-3, 0.2e-3, -0.1e-3], [0.15e-3, 0.9e-3, 0.05e-3], [-0.1e-3, 0.1e-3, 1.1e-3]]


def random_Q(orc, mech, seed=1):
    import flame3d_amd.cases as cases
    rng = np.random.default_rng(seed)
    sh = orc.shape
    T = torch.tensor(rng.uniform(800, 1500, sh[::-1]))
    P = torch.tensor(rng.uniform(5e4, 2e5, sh[::-1]))
    Y = torch.tensor(rng.uniform(0.1, 1.0, (mech.Ns,) + sh[::-1]))
    Y = Y / Y.sum(0, keepdim=True)
    vel = [torch.tensor(rng.uniform(-900, 900, sh[::-1])) for _ in range(3)]
    Q, r = cases.state_from_TPY(mech, T, P, Y, *vel)
    return to_np(Q), to_np(r)


# ---------------------------------------------------------------- splitting
def test_flux_split_sums_to_physical_flux(air):
    """F+ + F- = (rho Un, rho u Un + p Ax, ..., (E + p) Un) for the
    J^-1-weighted metric (Ax, Ay, Az) (Steger-Warming consistency)."""
    from oracle import Oracle
    N = (6, 5, 4)
    orc = Oracle(*N, air)
    Q, _ = random_Q(orc, air)
    M, _ = affine_metrics(N, A_SKEW)
    for d in range(3):
        Ax, Ay, Az = M[3 * d:3 * d + 3]
        Fp, Fm = orc.flux_split(Q, Ax, Ay, Az)
        rho, u, v, w, p, ei = Q[..., 0], Q[..., 1], Q[..., 2], Q[..., 3], Q[..., 4], Q[..., 6]
        Un = Ax * u + Ay * v + Az * w
        E = ei + 0.5 * rho * (u * u + v * v + w * w)
        F = np.stack([rho * Un, rho * u * Un + p * Ax, rho * v * Un + p * Ay,
                      rho * w * Un + p * Az, (E + p) * Un], -1)
        S = Fp + Fm
        for n in range(5):
            scale = np.max(np.abs(F[..., n])) + np.max(np.abs(Fp[..., n]))
            assert np.max(np.abs(S[..., n] - F[..., n])) < 1e-12 * scale, (d, n)


def test_flux_split_supersonic_one_sided(air):
    """Un > c |A|: all three eigenvalues positive, so F- = 0 and F+ = F."""
    from oracle import Oracle
    N = (4, 4, 4)
    orc = Oracle(*N, air)
    Q, _ = random_Q(orc, air)
    Q[..., 1] = 3000.0  # u well above c (~700 m/s) ...
    Q[..., 2] = Q[..., 3] = 0.0
    one = np.ones(orc.shape, order="F")
    Fp, Fm = orc.flux_split(Q, 1e-6 * one, 0 * one, 0 * one)
    assert np.all(Fm == 0.0)
    Fp2, Fm2 = orc.flux_split(Q, -1e-6 * one, 0 * one, 0 * one)
    assert np.all(Fp2 == 0.0)


def test_species_split(air):
    from oracle import Oracle
    N = (4, 4, 4)
    orc = Oracle(*N, air)
    Q, r = random_Q(orc, air)
    M, _ = affine_metrics(N, A_SKEW)
    Fp, Fm = orc.species_split(r, Q, *M[3:6])
    Un = M[3] * Q[..., 1] + M[4] * Q[..., 2] + M[5] * Q[..., 3]
    np.testing.assert_allclose(Fp + Fm, Un[..., None] * r, rtol=1e-14)
    assert np.all(Fp * Fm <= 0)
    assert np.all(Fp[Un < 0] == 0) and np.all(Fm[Un > 0] == 0)


# ------------------------------------------------------------ reconstruction
def _face_diff(orc, d, Fp, Fm, phi_val):
    phi = orc.zeros() + phi_val
    F = orc.weno(d, phi, Fp, Fm, 1)
    # cell derivative estimate: F(face i) - F(face i+1) in the sweep direction
    sl_lo = [slice(None)] * 3
    sl_hi = [slice(None)] * 3
    sl_lo[d] = slice(0, -1)
    sl_hi[d] = slice(1, None)
    return F[tuple(sl_lo) + (0,)] - F[tuple(sl_hi) + (0,)]


def _poly_field(orc, d, k):
    idx = np.arange(orc.shape[d], dtype=np.float64) - NG
    shape = [1, 1, 1]
    shape[d] = orc.shape[d]
    f = (idx ** k).reshape(shape) * np.ones(orc.shape)
    der = (k * idx ** (k - 1) if k > 0 else 0 * idx)[NG:-NG]
    return np.asfortranarray(f[..., None]), der


@pytest.mark.parametrize("d", [0, 1, 2])
def test_up7_polynomial_exactness(air, d):
    """The UP7 flux difference is exact for degree <= 5 and not for 6
    (UP7 = 0.7 UP5 + 0.3 CD6, formally 5th order): -dF/dxi at cell i equals
    F_{i-1/2} - F_{i+1/2} for Fp = x^k (Fm = 0) and for Fm = x^k (Fp = 0)."""
    from oracle import Oracle
    orc = Oracle(12, 12, 12, air)
    for k in range(0, 8):
        f, der = _poly_field(orc, d, k)
        for which in ("p", "m"):
            Fp = f if which == "p" else np.zeros_like(f, order="F")
            Fm = f if which == "m" else np.zeros_like(f, order="F")
            got = -_face_diff(orc, d, Fp, Fm, 0.0)
            shape = [1, 1, 1]
            shape[d] = -1
            exact = np.broadcast_to(der.reshape(shape), got.shape)
            err = np.max(np.abs(got - exact)) / max(1.0, np.max(np.abs(exact)))
            if k <= 5:
                assert err < 1e-9, (k, which, err)
            else:
                assert err > 1e-6, (k, which, err)


@pytest.mark.parametrize("d", [0, 1, 2])
def test_weno5z_and_nnd_polynomial_exactness(air, d):
    """WENO5-Z (sensor in [1e-3, 0.1)) is exact for degree <= 2 whatever its
    weights; NND (sensor >= 0.1) is exact for degree <= 1 and not for 2."""
    from oracle import Oracle
    orc = Oracle(12, 12, 12, air)
    for k in range(0, 4):
        f, der = _poly_field(orc, d, k)
        shape = [1, 1, 1]
        shape[d] = -1
        for which in ("p", "m"):
            Fp = f if which == "p" else np.zeros_like(f, order="F")
            Fm = f if which == "m" else np.zeros_like(f, order="F")
            for phi, kmax in ((0.01, 2), (0.5, 1)):
                got = -_face_diff(orc, d, Fp, Fm, phi)
                exact = np.broadcast_to(der.reshape(shape), got.shape)
                err = np.max(np.abs(got - exact)) / max(1.0, np.max(np.abs(exact)))
                if k <= kmax:
                    assert err < 1e-9, (phi, k, which, err)
                else:
                    assert err > 1e-6, (phi, k, which, err)


def test_weno_branch_thresholds_and_no_new_extrema(air):
    """Branch selection uses max(phi[i-1], phi[i]) against 1e-3 and 0.1; on a
    step the WENO5-Z and NND face values stay within the data range."""
    from oracle import Oracle
    orc = Oracle(12, 4, 4, air)
    idx = np.arange(orc.shape[0]) - NG
    f = np.where(idx < 6, 1.0, 2.0)[:, None, None, None] * np.ones(orc.shape + (1,))
    f = np.asfortranarray(f)
    z = np.zeros_like(f, order="F")
    res = {}
    for phi in (0.99e-3, 1e-3, 0.0999, 0.1):
        F = orc.weno(0, orc.zeros() + phi, f, z, 1)[..., 0]
        res[phi] = F
    # switching exactly at the thresholds (strict '<')
    assert not np.array_equal(res[0.99e-3], res[1e-3])
    assert np.array_equal(res[1e-3], res[0.0999])
    assert not np.array_equal(res[0.0999], res[0.1])
    for phi in (1e-3, 0.1):
        assert res[phi].min() >= 1.0 - 1e-12 and res[phi].max() <= 2.0 + 1e-12
    # the linear UP7 does overshoot at a step (it is not monotone)
    assert res[0.99e-3].max() > 2.0 + 1e-6 or res[0.99e-3].min() < 1.0 - 1e-6
    # max over the two neighbours: phi large only at the right cell still switches
    phi = orc.zeros()
    phi[8, :, :] = 0.5
    F = orc.weno(0, phi, f, z, 1)[..., 0]
    nnd = res[0.1]
    a = 8 - NG  # faces a (right cell 8) and a+1 (left cell 8) use NND
    assert np.array_equal(F[a], nnd[a]) and np.array_equal(F[a + 1], nnd[a + 1])
    assert np.array_equal(F[a - 1], res[0.99e-3][a - 1])


def test_shock_sensor(air):
    """Linear pressure gives phi = 0; a single-cell bump gives the closed form."""
    from oracle import Oracle
    orc = Oracle(6, 6, 6, air)
    Q = orc.zeros(7)
    i, j, k = np.meshgrid(*[np.arange(s) for s in orc.shape], indexing="ij")
    Q[..., 4] = 1e5 + 10 * i + 20 * j - 5 * k
    phi = orc.shock_sensor(Q)
    assert np.max(np.abs(phi)) < 1e-15
    Q[..., 4] = 1e5
    Q[7, 7, 7, 4] = 2e5
    phi = orc.shock_sensor(Q)
    # at the bump: three directions each |-1 + 4 - 1| / (1 + 4 + 1) (in 1e5)
    assert abs(phi[7, 7, 7] - 3 * 2.0 / 6.0) < 1e-15
    assert abs(phi[8, 7, 7] - 1.0 / 5.0) < 1e-15  # neighbour: |-2+2-1|/(2+2+1)


# ------------------------------------------------------------------ viscous
def test_viscous_linear_field_exact(air):
    """u, v, w, T linear in physical space on an affine (skewed) grid: the
    face gradients are exact, so Fv equals the closed-form stress and heat
    flux through the face (mu, lambda constant; Fh = 0)."""
    from oracle import Oracle
    N = (6, 6, 6)
    orc = Oracle(*N, air)
    A = np.array(A_SKEW)
    M, J = affine_metrics(N, A)
    idx = np.stack(np.meshgrid(*[np.arange(s) - NG for s in orc.shape], indexing="ij"), -1)
    X = idx @ A.T  # physical coordinates
    G = np.array([[10.0, -3.0, 2.0], [4.0, 7.0, -1.0], [-2.0, 5.0, 3.0]])  # d(u,v,w)/dx
    gT = np.array([300.0, -100.0, 50.0])
    Q = orc.zeros(7)
    Q[..., 1:4] = X @ G.T
    Q[..., 5] = 1000 + X @ gT
    mu, lam = 2e-5, 0.05
    div = np.trace(G)
    tau = mu * (G + G.T - 2.0 / 3.0 * div * np.eye(3))
    for d in range(3):
        Fh = orc.face_zeros(d, 3)
        Fv = orc.viscous_flux(d, Q, M, J, orc.zeros() + lam, orc.zeros() + mu, Fh)
        n = np.array([M[3 * d][0, 0, 0], M[3 * d + 1][0, 0, 0], M[3 * d + 2][0, 0, 0]])
        vel_face = 0.5 * (np.roll(Q[..., 1:4], 1, axis=d) + Q[..., 1:4])
        sl = [slice(NG, NG + N[0] + (d == 0)), slice(NG, NG + N[1] + (d == 1)),
              slice(NG, NG + N[2] + (d == 2))]
        vf = vel_face[tuple(sl)]
        np.testing.assert_allclose(Fv[..., 0:3], np.broadcast_to(tau @ n, Fv[..., 0:3].shape),
                                   rtol=1e-9, atol=1e-12 * np.abs(tau @ n).max())
        E = vf @ tau + lam * gT  # u_j tau_jk + lambda dT/dx_k
        np.testing.assert_allclose(Fv[..., 3], E @ n, rtol=1e-9, atol=1e-12 * np.abs(E @ n).max())


def test_species_viscous_linear_field(air):
    """Linear Y_k on an affine grid: V_k = 2 D rho grad Y_k (the reference's
    (D_{i-1} + D_i) rho without 1/2), minus Y_face * sum_k V_k; Fh = sum V_k h_k."""
    from oracle import Oracle
    N = (6, 6, 6)
    orc = Oracle(*N, air)
    A = np.array(A_SKEW)
    M, J = affine_metrics(N, A)
    idx = np.stack(np.meshgrid(*[np.arange(s) - NG for s in orc.shape], indexing="ij"), -1)
    X = idx @ A.T
    gY = np.array([[5.0, 1.0, -2.0], [-3.0, 2.0, 1.0], [0.0, 0.0, 0.0], [1.0, -1.0, 4.0],
                   [-3.0, -2.0, -3.0]])
    Yi = orc.zeros(5)
    base = np.array([0.1, 0.2, 0.3, 0.15, 0.25])
    Yi[...] = base + X @ gY.T
    Q = orc.zeros(7)
    Q[..., 0] = 1.2
    Q[..., 5] = 900.0
    Dk = np.array([1e-4, 2e-4, 3e-4, 4e-4, 5e-4])
    D = orc.zeros(5) + Dk
    h = orc.h_specs(900.0)
    for d in range(3):
        Fd, Fh = orc.spec_viscous_flux(d, Q, Yi, M, J, D)
        n = np.array([M[3 * d][0, 0, 0], M[3 * d + 1][0, 0, 0], M[3 * d + 2][0, 0, 0]])
        V = 2 * Dk[:, None] * 1.2 * gY  # [Ns, 3]
        Yf = 0.5 * (np.roll(Yi, 1, axis=d) + Yi)
        sl = [slice(NG, NG + N[0] + (d == 0)), slice(NG, NG + N[1] + (d == 1)),
              slice(NG, NG + N[2] + (d == 2))]
        Yf = Yf[tuple(sl)]
        Vc = V[None, None, None] - Yf[..., :, None] * V.sum(0)[None, None, None, None, :]
        np.testing.assert_allclose(Fd, Vc @ n, rtol=1e-9, atol=1e-15)
        np.testing.assert_allclose(Fh, np.einsum("...kj,k->...j", Vc, h), rtol=1e-9, atol=1e-9)


# --------------------------------------------------------------- divergence
def test_div_periodic_telescoping(air):
    """With periodic face fluxes (last face = first face) the divergence
    changes sum(U / (J dt)) by zero (global conservation to roundoff)."""
    from oracle import Oracle
    N = (5, 6, 7)
    orc = Oracle(*N, air)
    rng = np.random.default_rng(0)
    F = []
    for d in range(3):
        for nv in (5, 4):
            f = orc.face_zeros(d, nv)
            f[...] = rng.standard_normal(f.shape)
            sl0 = [slice(None)] * 3
            sl1 = [slice(None)] * 3
            sl0[d] = 0
            sl1[d] = -1
            f[tuple(sl1)] = f[tuple(sl0)]
            F.append(f)
    U = orc.zeros(5) + 1.0
    J = orc.zeros() + 2.0
    U0 = U.copy(order="F")
    orc.div(U, F[0], F[2], F[4], F[1], F[3], F[5], 0.1, J)
    inner = (slice(NG, -NG),) * 3
    dU = (U - U0)[inner]
    assert np.max(np.abs(dU.sum(axis=(0, 1, 2)))) < 1e-12
    assert np.max(np.abs(dU)) > 0.1
    assert np.all((U - U0)[:NG] == 0)  # ghosts untouched


# -------------------------------------------------------------- thermo etc.
def test_internal_energy_monatomic_closed_form(h2):
    """H atom (a1 = 2.5, others 0): e = 1.5 R T / W exactly per kg."""
    from oracle import Oracle
    orc = Oracle(1, 1, 1, h2)
    r = np.zeros(h2.Ns)
    iH = h2.index("H")
    r[iH] = 0.7
    for T in (400.0, 1500.0):
        e = orc.internal_energy(T, r)
        assert abs(e - 0.7 * 1.5 * h2.Ru * T / h2.mw[iH]) < 1e-12 * e
        cv = orc.cv(T, r)
        assert abs(cv - 0.7 * 1.5 * h2.Ru / h2.mw[iH]) < 1e-12 * cv


def test_getT_inverts_energy_and_extrapolates(h2):
    from oracle import Oracle
    orc = Oracle(1, 1, 1, h2)
    rng = np.random.default_rng(5)
    for _ in range(50):
        r = rng.uniform(0, 1, h2.Ns)
        T0 = rng.uniform(350, 4800)
        T, it = orc.get_T(orc.internal_energy(T0, r), r)
        assert abs(T - T0) < 1e-3 and it <= 30
    r = rng.uniform(0, 1, h2.Ns)
    # below min_temp: linear extrapolation with cv(min_temp)
    e = orc.internal_energy(300.0, r) - 1000.0 * orc.cv(300.0, r)
    T, it = orc.get_T(e, r)
    assert abs(T - (300.0 - 1000.0)) < 1e-9 and it == 0


def test_getT_warm_start_agrees_with_getT(h2):
    """GetT warm-started from a nearby temperature (the FP32 stage's c2Prim and
    the surrogate step use it) stops within the 1e-3 K tolerance of the same
    root as the reference start; outside [min_temp, max_temp] it takes the
    reference's extrapolation exactly; a guess outside the range falls back."""
    from oracle import Oracle
    orc = Oracle(1, 1, 1, h2)
    rng = np.random.default_rng(6)
    for _ in range(200):
        r = rng.uniform(0, 1, h2.Ns)
        T0 = rng.uniform(350, 4800)
        e = orc.internal_energy(T0, r)
        for guess in (T0 + rng.uniform(-30, 30), T0 + rng.uniform(-400, 400), 1000.0):
            guess = min(max(guess, 300.0), 5000.0)
            Tw = orc.get_T_from(e, r, guess)
            assert abs(Tw - T0) < 1e-3, (T0, guess, Tw)
            assert abs(Tw - orc.get_T(e, r)[0]) < 2e-3
    r = rng.uniform(0, 1, h2.Ns)
    e = orc.internal_energy(300.0, r) - 1000.0 * orc.cv(300.0, r)
    assert orc.get_T_from(e, r, 400.0) == orc.get_T(e, r)[0]  # extrapolation region
    e = orc.internal_energy(1500.0, r)
    assert orc.get_T_from(e, r, float("nan")) == orc.get_T(e, r)[0]
    assert orc.get_T_from(e, r, 9000.0) == orc.get_T(e, r)[0]
    # an energy inside the NASA-7 jump at 1000 K (H2-rich: ~0.2 K wide) has no
    # root; the reference start spends its 30 updates oscillating across the
    # breakpoint and ends on the cycle point the parity selects.  The warm
    # start detects the non-shrinking flip and returns that same point
    # (get_T_gap: the reference iteration cut short once the cycle repeats),
    # whatever the guess and wherever in the jump the energy lies
    for comp in ((0.3, 0.01), (0.02, 0.9), (0.1, 0.2)):
        r = np.zeros(h2.Ns)
        r[h2.index("H2")], r[h2.index("N2")] = comp
        r[h2.index("O2")] = 0.05
        elo, ehi = orc.internal_energy(1000.0 - 1e-9, r), orc.internal_energy(1000.0, r)
        gap = (ehi - elo) / orc.cv(1000.0, r)
        assert abs(gap) > 1e-3, gap
        for frac in (0.05, 0.3, 0.5, 0.77, 0.95):
            e = elo + frac * (ehi - elo)
            T_ref, it = orc.get_T(e, r)
            assert it == 30 and abs(T_ref - 1000.0) <= abs(gap) + 1e-9
            for guess in (999.7, 1000.0, 1000.3, 1003.0, 960.0):
                Tw = orc.get_T_from(e, r, guess)
                assert abs(Tw - T_ref) < 1e-6, (comp, frac, guess, Tw, T_ref)
            # maxiter parity: an odd update budget ends on the other side
            T29, it29 = orc.get_T(e, r, maxiter=29)
            assert it29 == 29 and abs(T29 - T_ref) > 0.5 * abs(gap)


def test_gibbs_equals_h_minus_s(h2):
    """g/RT from thermo.jl:163-191 equals h/RT - s/R of the NASA-7 fit."""
    from oracle import Oracle
    orc = Oracle(1, 1, 1, h2)
    for T in (500.0, 999.0, 1000.0, 2500.0):
        g = orc.gibbs(T)
        for n in range(h2.Ns):
            a = h2.coeffs_lo[:, n] if T < h2.coeffs_sep[n] else h2.coeffs_hi[:, n]
            hRT = a[0] + a[1] * T / 2 + a[2] * T ** 2 / 3 + a[3] * T ** 3 / 4 + a[4] * T ** 4 / 5 + a[5] / T
            sR = a[0] * math.log(T) + a[1] * T + a[2] * T ** 2 / 2 + a[3] * T ** 3 / 3 + a[4] * T ** 4 / 4 + a[6]
            assert abs(g[n] - (hRT - sR)) < 1e-11 * max(1.0, abs(g[n]))


def _conc(h2, T=1800.0, P=3e5, seed=2):
    rng = np.random.default_rng(seed)
    Y = rng.uniform(0.01, 1.0, h2.Ns)
    Y /= Y.sum()
    W = 1.0 / np.sum(Y / h2.mw)
    rho = P * W / (h2.Ru * T)
    return Y * rho / h2.mw * 1e-6


@pytest.mark.parametrize("mname", ["air5", "h2_lidryer"])
def test_rates_conserve_mass_and_elements(mname, air, h2):
    from oracle import Oracle
    m = air if mname == "air5" else h2
    orc = Oracle(1, 1, 1, m)
    sc = _conc(m, T=3000.0 if mname == "air5" else 1800.0)
    w = orc.production_rate(sc, 3000.0 if mname == "air5" else 1800.0, jac=False)
    assert abs(np.sum(w * m.mw)) < 1e-12 * np.sum(np.abs(w * m.mw))
    assert np.max(np.abs(w)) > 0


def test_equilibrium_detailed_balance(h2):
    """At Kc-consistent concentrations a reaction's net rate vanishes: scale
    a product concentration so q_f = q_r for H2 + O <=> OH + H (elementary)."""
    from oracle import Oracle
    orc = Oracle(1, 1, 1, h2)
    T = 1500.0
    g = orc.gibbs(T)
    iH2, iO, iOH, iH = (h2.index(s) for s in ("H2", "O", "OH", "H"))
    Kp = math.exp(g[iH2] + g[iO] - g[iOH] - g[iH])  # sgm = 0
    sc = np.zeros(h2.Ns)
    sc[iH2], sc[iO], sc[iH] = 1e-5, 2e-7, 3e-7
    sc[iOH] = Kp * sc[iH2] * sc[iO] / sc[iH]
    # isolate reaction 2 (index 1) with a one-reaction mechanism
    import copy
    m1 = copy.deepcopy(h2)
    keep = [1]
    for f in ("reaction_type", "sgm"):
        setattr(m1, f, getattr(h2, f)[keep])
    for f in ("vf", "vr", "ef"):
        setattr(m1, f, getattr(h2, f)[:, keep])
    for f in ("Arr", "loP", "Troe"):
        setattr(m1, f, getattr(h2, f)[:, keep])
    o1 = Oracle(1, 1, 1, m1)
    w = o1.production_rate(sc, T, jac=False)
    kf = h2.Arr[0, 1] * T ** h2.Arr[1, 1] * math.exp(-h2.Arr[2, 1] / T)
    assert np.max(np.abs(w)) < 1e-10 * kf * sc[iH2] * sc[iO]


@pytest.mark.parametrize("mname", ["air5", "h2_lidryer"])
def test_full_jacobian_matches_finite_differences(mname, air, h2):
    """Newton-mode Jacobian (exact mass action + [M] derivative) against
    central differences of wdot; for the H2 mechanism the two Troe reactions'
    F_cent(Pr) dependence is excluded by removing them."""
    import copy
    from oracle import Oracle
    m = air if mname == "air5" else h2
    if mname == "h2_lidryer":
        keep = [j for j in range(m.Nr) if m.reaction_type[j] != 3]
        m = copy.deepcopy(m)
        for f in ("reaction_type", "sgm"):
            setattr(m, f, getattr(h2, f)[keep])
        for f in ("vf", "vr", "ef", "Arr", "loP", "Troe"):
            setattr(m, f, getattr(h2, f)[:, keep])
    orc = Oracle(1, 1, 1, m)
    T = 3500.0 if mname == "air5" else 1800.0
    sc = _conc(m, T=T)
    w, A = orc.production_rate(sc, T, full_jac=True)
    Afd = np.zeros_like(A)
    for j in range(m.Ns):
        h = 1e-6 * sc[j]
        sp, sm = sc.copy(), sc.copy()
        sp[j] += h
        sm[j] -= h
        Afd[:, j] = (orc.production_rate(sp, T, jac=False) - orc.production_rate(sm, T, jac=False)) / (2 * h)
    err = np.max(np.abs(A - Afd)) / np.max(np.abs(Afd))
    assert err < 1e-6, err
    # the reference Jacobian (full_jac = 0) neglects the third-body terms
    _, Aref = orc.production_rate(sc, T, full_jac=False)
    assert np.max(np.abs(Aref - Afd)) / np.max(np.abs(Afd)) > 1e-6


def test_gauss_solve_vs_numpy():
    import flame3d_amd.mech as mm
    from oracle import Oracle
    orc = Oracle(1, 1, 1, mm.load("air5"))
    rng = np.random.default_rng(9)
    for n in (5, 9, 10):
        A = rng.standard_normal((n, n)) + n * np.eye(n)
        A[0, 0] = 1e-3  # forces a pivot exchange
        b = rng.standard_normal(n)
        x = orc.gauss_solve(A, b)
        np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-12, atol=1e-14)


def test_mixture_single_and_identical_species(air):
    """One species: mu = mu_1, lambda = lambda_1, D = 0.  Identical species
    split in two: the mixture values equal the pure ones (Wilke Phi = 1)."""
    import copy
    from oracle import Oracle
    m1 = copy.deepcopy(air)
    T, P = 1200.0, 1e5
    orc = Oracle(1, 1, 1, air)
    X = np.zeros(5)
    X[air.index("N2")] = 1.0
    lam, mu, D = orc.mixture_properties(T, P, X)
    iN2 = air.index("N2")
    lgT = math.log(T)
    v = np.polyval(air.visc_poly[::-1, iN2], lgT)
    c = np.polyval(air.conduct_poly[::-1, iN2], lgT)
    assert abs(mu - (T ** 0.25 * v) ** 2) < 1e-12 * mu
    assert abs(lam - math.sqrt(T) * c) < 1e-12 * lam
    assert abs(D[iN2]) < 1e-12
    # clone N2 into the O2 slot: same W and transport -> X split does not matter
    for arr in ("mw", "coeffs_sep"):
        getattr(m1, arr)[air.index("O2")] = getattr(air, arr)[iN2]
    for arr in ("coeffs_lo", "coeffs_hi", "visc_poly", "conduct_poly"):
        getattr(m1, arr)[:, air.index("O2")] = getattr(air, arr)[:, iN2]
    o1 = Oracle(1, 1, 1, m1)
    X2 = np.zeros(5)
    X2[iN2], X2[air.index("O2")] = 0.4, 0.6
    lam2, mu2, _ = o1.mixture_properties(T, P, X2)
    assert abs(mu2 - mu) < 1e-12 * mu and abs(lam2 - lam) < 1e-12 * lam


def test_mixture_diffusivity_binary_limit(air):
    """Mixture-averaged D_k = (1 - X_k) / (P sum_{l != k} X_l / D_kl + eps):
    in a binary mixture D_1 = x_2 / (P x_2 / D_12 + eps) (-> D_12 / P) for
    every composition, including the nearly pure limits where 1 - X_1 is
    computed as sum_{l != 1} X_l (a plain 1 - X_1 would cancel to rounding
    noise)."""
    from oracle import Oracle
    orc = Oracle(1, 1, 1, air)
    T, P = 1500.0, 2e5
    i, j = air.index("N2"), air.index("O2")
    lgT = math.log(T)
    D12 = T ** 1.5 * np.polyval(air.binarydiff_poly[::-1, i, j], lgT)
    for x2 in (0.5, 1e-3, 1e-9, 1e-14):
        X = np.zeros(5)
        X[i], X[j] = 1.0 - x2, x2
        _, _, D = orc.mixture_properties(T, P, X)
        eps = 2.220446049250313e-16
        for k, xo in ((i, x2), (j, 1.0 - x2)):
            exp = xo / (P * xo / D12 + eps)
            assert abs(D[k] - exp) < 1e-12 * exp, (x2, k, D[k], exp)
        assert abs(D[j] * P - D12) < 1e-9 * D12


def test_bug_compat_flag_reads_transposed_coefficients(h2):
    """eval_gpu_stiff's coeffs_lo[n, 6] (reactions.jl:282) reads linear index
    35 + n of the 7 x Ns array; with the flag the energy change differs."""
    import flame3d_amd.cases as cases
    from oracle import Oracle
    N = (4, 4, 4)
    orc = Oracle(*N, h2)
    Q, r = cases.chem_cells(N, h2, seed=3)
    U = cases.prim2c(Q)
    outs = []
    for bug in (0, 1):
        Qo, ro, Uo = to_np(Q), to_np(r), to_np(U)
        orc.chem_stiff(Uo, Qo, ro, 1e-7, bug_compat=bug)
        outs.append(Uo[..., 4].copy())
    assert not np.allclose(outs[0], outs[1])


def test_newton_accepts_nasa7_breakpoint_gap(h2):
    """Regression: states whose energy falls in the jump of the NASA-7
    polynomials at T_mid (fixture tests/golden/newton_nasa_gap.json) made the
    T update flip sign forever; the step is accepted once the species have
    converged (2 iterations, not a failure after every substep restart)."""
    import json
    import os
    from oracle import Oracle
    path = os.path.join(os.path.dirname(__file__), "golden", "newton_nasa_gap.json")
    cells = [[float.fromhex(x) for x in row] for row in json.load(open(path))["cells"]]
    orc = Oracle(1, 1, 1, h2)
    for row in cells:
        Q = np.zeros((9, 9, 9, 7), order="F")
        R = np.zeros((9, 9, 9, 9), order="F")
        U = np.zeros((9, 9, 9, 5), order="F")
        Q[:], R[:], U[:] = row[:7], row[7:16], row[16:21]
        it = orc.chem_newton(U, Q, R, 1e-7)
        assert 1 <= int(it.ravel()[0]) <= 4, it.ravel()[0]
        assert abs(Q[4, 4, 4, 5] - 1000.0) < 1.0  # T from GetT stays at the stream's 1000 K


def test_chemistry_exp_log_accuracy(h2):
    """The exp / log the chemistry rates use on the CPU and the GPU (restated
    from the library's f3d_exp / f3d_log, only +, -, *, / and exact scaling):
    within 1 ulp (exp) and 3 ulp (log) of the correctly rounded value over
    the arguments the kinetics produce (measured maxima), exact where the
    result is exact, IEEE special values."""
    from oracle import Oracle
    L = Oracle(1, 1, 1, h2).lib
    rng = np.random.default_rng(8)
    xs = np.concatenate([rng.uniform(-700, 700, 20000), rng.uniform(-50, 50, 20000),
                         rng.uniform(-1, 1, 5000), [-745.0, 709.7, 1e-300, -1e-300]])
    got = np.array([L.or_exp(float(x)) for x in xs])
    ref = np.exp(xs.astype(np.longdouble))
    ulp = np.spacing(ref.astype(np.float64))
    err = np.abs(got - ref.astype(np.float64))
    ok = ref > np.finfo(np.float64).tiny  # normal results
    assert np.max(err[ok] / ulp[ok]) <= 1.0
    ys = np.concatenate([10.0 ** rng.uniform(-300, 300, 20000), rng.uniform(0.5, 2.0, 20000),
                         rng.uniform(200, 6000, 5000), [5e-324, 1.7e308]])
    got = np.array([L.or_log(float(y)) for y in ys])
    ref = np.log(ys.astype(np.longdouble)).astype(np.float64)
    ulp = np.spacing(np.abs(ref))
    nz = np.abs(ref) > 1e-300
    assert np.max(np.abs(got - ref)[nz] / ulp[nz]) <= 3.0
    assert L.or_exp(0.0) == 1.0 and L.or_log(1.0) == 0.0
    for k in (-1000, -3, 1, 52, 1000):
        # the argument k ln2 is itself rounded: relative |k ln2| 2^-53 in the result
        assert L.or_exp(k * math.log(2.0)) == pytest.approx(
            2.0 ** k, rel=abs(k) * math.log(2.0) * 2.3e-16 + 4e-16)
        assert L.or_log(2.0 ** k) == pytest.approx(k * math.log(2.0), rel=2e-16, abs=0)
    assert math.isnan(L.or_exp(float("nan"))) and math.isnan(L.or_log(-1.0))
    assert L.or_exp(800.0) == math.inf and L.or_exp(-800.0) == 0.0
    assert L.or_log(0.0) == -math.inf and L.or_log(math.inf) == math.inf


@pytest.mark.parametrize("mname", ["h2_lidryer", "air5"])
def test_rates_temperature_derivative(mname, air, h2):
    """The analytic d wdot / dT of the backward-Euler Newton's Jacobian
    (Arrhenius, Gibbs-Helmholtz for K_c, the Troe falloff's Pr/(1+Pr) and F)
    equals the central difference of the rates (step 1e-5 T; Richardson
    extrapolation of two steps for the O(h^4) check), species by species,
    over states on both sides of the NASA-7 breakpoint.  A dropped Troe term
    or a wrong sign in d ln K_c / dT fails it."""
    from oracle import Oracle
    m = {"air5": air, "h2_lidryer": h2}[mname]
    orc = Oracle(1, 1, 1, m)
    rng = np.random.default_rng(12)
    temps = (700.0, 1200.0, 1800.0, 2600.0) if mname == "h2_lidryer" else (2500.0, 4500.0)
    for T in temps:
        for _ in range(3):
            Y = rng.uniform(0.01, 1.0, m.Ns)
            Y /= Y.sum()
            P = rng.uniform(1e4, 1e6)
            W = 1.0 / np.sum(Y / m.mw)
            rho = P * W / (m.Ru * T)
            sc = Y * rho / m.mw * 1e-6
            w, dw = orc.production_rate_dT(sc, T)
            fd = []
            for h in (1e-5 * T, 0.5e-5 * T):
                wp = orc.production_rate(sc, T + h, jac=False)
                wm = orc.production_rate(sc, T - h, jac=False)
                fd.append((wp - wm) / (2 * h))
            rich = (4 * fd[1] - fd[0]) / 3
            scale = np.max(np.abs(rich)) + 1e-300
            err = np.max(np.abs(dw - rich)) / scale
            assert err < 1e-6, (mname, T, err)
