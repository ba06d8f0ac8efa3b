"""WENO7-Z reconstruction (recon_kind = WENO7: the middle branch of the hybrid
switch; the reference's is WENO5-Z).  Pinned on the CPU by exact rational
arithmetic: the candidate stencils and linear weights combine to the
7th-order upwind stencil (-3, 25, -101, 319, 214, -38, 4)/420, whose flux
difference is exact for polynomials of degree <= 7; the smoothness indicators
equal the Jiang-Shu definition sum_l int (d^l p_k / dx^l)^2 (l = 1..3) of each
candidate cubic, derived here from scratch (the kernels evaluate them as
quadratic forms in first differences); the Z weights follow Castro, Costa &
Don (2011).  Then the GPU kernels (x, y, z sweeps, both precisions) against
the oracle with every face forced into the middle branch."""
import itertools
from fractions import Fraction as Fr

import numpy as np
import pytest
import torch

from helpers import rel_err, to_np, to_torch

LIN = [Fr(c, 420) for c in (-3, 25, -101, 319, 214, -38, 4)]


def _orc():
    import flame3d_amd.mech as mm
    from oracle import Oracle
    return Oracle(1, 1, 1, mm.load("air5"))


def _betas_exact():
    """Quadratic forms of the four smoothness indicators over v_{j-3..j+3}."""
    out = []
    for k in range(4):
        cells = list(range(-3 + k, 1 + k))
        n = 4
        M = [[(Fr(2 * c + 1, 2) ** (m + 1) - Fr(2 * c - 1, 2) ** (m + 1)) / (m + 1)
              for m in range(n)] for c in cells]
        A = [row[:] + [Fr(int(i == r)) for i in range(n)] for r, row in enumerate(M)]
        for col in range(n):  # Gauss-Jordan: polynomial coefficients from cell averages
            piv = next(r for r in range(col, n) if A[r][col] != 0)
            A[col], A[piv] = A[piv], A[col]
            pv = A[col][col]
            A[col] = [x / pv for x in A[col]]
            for r in range(n):
                if r != col and A[r][col] != 0:
                    f = A[r][col]
                    A[r] = [a - f * b for a, b in zip(A[r], A[col])]
        coef = [row[n:] for row in A]
        Q = [[Fr(0)] * 7 for _ in range(7)]
        d = coef
        for _ in range(3):
            d = [[m * x for x in d[m]] for m in range(1, len(d))]
            for m1, m2 in itertools.product(range(len(d)), repeat=2):
                e = m1 + m2
                I = (Fr(1, 2) ** (e + 1) - Fr(-1, 2) ** (e + 1)) / (e + 1)
                for a in range(4):
                    for b in range(4):
                        Q[cells[a] + 3][cells[b] + 3] += d[m1][a] * d[m2][b] * I
        out.append(np.array([[float(x) for x in r] for r in Q]))
    return out


def test_weno7_linear_limit_is_the_7th_order_stencil():
    orc = _orc()
    rng = np.random.default_rng(3)
    for _ in range(20):
        v = rng.normal(size=7)
        ref = sum(float(c) * x for c, x in zip(LIN, v))
        assert abs(orc.weno7z(v, eps=1e30) - ref) <= 1e-14 * np.abs(v).max()
    # the stencil's flux difference reproduces f'(0) for f = x^n, n <= 7 (not 8)
    for n in range(9):
        D = sum(c * (Fr(m - 3) ** n - Fr(m - 4) ** n) for m, c in enumerate(LIN))
        # cell averages of x^n are not point values: compare the flux difference
        # with the cell-average form, int_{-1/2}^{1/2} f'(x) dx = f(1/2) - f(-1/2)
        V = [(Fr(2 * c + 1, 2) ** (n + 1) - Fr(2 * c - 1, 2) ** (n + 1)) / (n + 1)
             for c in range(-4, 4)]  # averages of x^n over cells -4..3
        hp = sum(c * V[m + 1] for m, c in enumerate(LIN))   # face +1/2 from cells -3..3
        hm = sum(c * V[m] for m, c in enumerate(LIN))       # face -1/2 from cells -4..2
        exact = Fr(1, 2) ** n - Fr(-1, 2) ** n
        assert (hp - hm == exact) == (n <= 7), n


def test_weno7_smoothness_indicators_and_z_weights():
    orc = _orc()
    B = _betas_exact()
    d = [1, 12, 18, 4]
    C = [[-3, 13, -23, 25, 0, 0, 0], [0, 1, -5, 13, 3, 0, 0], [0, 0, -1, 7, 7, -1, 0],
         [0, 0, 0, 3, 13, -5, 1]]
    rng = np.random.default_rng(4)
    for trial in range(50):
        v = rng.normal(size=7) * (10.0 ** rng.uniform(-3, 3))
        if trial % 5 == 0:
            v[4:] += 5.0 * np.abs(v).max()  # a jump inside the stencil
        beta = [float(v @ Bk @ v) for Bk in B]
        tau = abs(beta[0] + 3 * beta[1] - 3 * beta[2] - beta[3])
        a = [d[k] * (1 + (tau / (1e-26 + beta[k])) ** 2) for k in range(4)]
        q = [float(np.dot(C[k], v)) / 12 for k in range(4)]
        ref = sum(a[k] * q[k] for k in range(4)) / sum(a)
        got = orc.weno7z(v)
        assert abs(got - ref) <= 1e-12 * np.abs(v).max(), (trial, got, ref)
    # linear data: equal indicators, tau7 = 0, the linear scheme exactly
    v = 0.3 + 0.7 * np.arange(7.0)
    assert abs(orc.weno7z(v) - sum(float(c) * x for c, x in zip(LIN, v))) < 1e-13
    # a jump: essentially non-oscillatory (stays within the data's range)
    v = np.array([0, 0, 0, 0, 1, 1, 1.0])
    assert -1e-3 <= orc.weno7z(v) <= 1 + 1e-3


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_weno7_gpu_parity(prec):
    import flame3d_amd.abi as abi
    import flame3d_amd.cases as cases
    import flame3d_amd.mech as mm
    from oracle import Oracle
    m = mm.load("h2_lidryer")
    N = (20, 12, 10)
    grid = cases.skewed_grid(*N)
    Q, rhoi = cases.random_state(grid, m, seed=9, shock=True)
    print(f"\n[weno7] {prec} size={N}", flush=True)
    orc = Oracle(*N, m, scheme=dict(phi1=-1.0, phi2=1e30, recon_kind=1))  # every face WENO7-Z
    ctx = abi.Context(*N, m, prec, scheme=dict(hybrid_phi1=-1.0, hybrid_phi2=1e30,
                                               recon_kind=abi.RECON_WENO7))
    Qn, rn = to_np(Q), to_np(rhoi)
    M = [to_np(x) for x in grid.metrics]
    phi = orc.shock_sensor(Qn)
    tol = 1e-12 if prec == "fp64" else 2e-4
    r = lambda a: to_torch(a, dtype=ctx.rdtype)
    for d in range(3):
        Fp, Fm = orc.flux_split(Qn, *M[3 * d:3 * d + 3])
        ref = orc.weno(d, phi, Fp, Fm, 5)
        F = ctx.face_zeros(d, 5)
        ctx.weno(d, F, r(phi), r(Fp), r(Fm), 5)
        Fpi, Fmi = orc.species_split(rn, Qn, *M[3 * d:3 * d + 3])
        refi = orc.weno(d, phi, Fpi, Fmi, m.Ns)
        Fi = ctx.face_zeros(d, m.Ns)
        ctx.weno(d, Fi, r(phi), r(Fpi), r(Fmi), m.Ns)
        torch.cuda.synchronize()
        for name, got, want in (("cons", F, ref), ("species", Fi, refi)):
            e = max(rel_err(to_np(got)[..., n], want[..., n]) for n in range(want.shape[-1]))
            print(f"  d={d} {name}: {e:.2e}", flush=True)
            assert e <= tol
    ctx.close()
