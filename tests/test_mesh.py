"""Mesh preprocessing (parse_mesh.jl): ghost nodes by reflection and the
CD6/CD2 cofactor metrics.

CPU: the oracle (oracle/mesh.py) against what the mathematics fixes --
  * a quadratic node map: CD6 and CD2 are exact for it, so the metric matrix
    M (rows xi, eta, zeta; columns x, y, z) must satisfy M A = det(A) I with
    A the analytic coordinate Jacobian and J det(A) = 1 (numpy's det, not the
    cofactor formulas), at every padded point;
  * an affine map: the reflected ghosts (edges and corners) continue it;
  * a map additive in i, j, k with nonlinear parts: the corner rule equals
    the two one-dimensional reflections added.
GPU: f3d_mesh_ghosts and f3d_mesh_metrics against the oracle, bit for bit,
on a wavy curvilinear block and on the reference generator's own mesh
family (x uniform then growing, asinh-clustered y and z)."""
import numpy as np
import pytest
import torch

from oracle import mesh

NG = 4


def _index(N):
    Nx, Ny, Nz = N
    K, J, I = np.meshgrid(np.arange(Nz + 2 * NG, dtype=np.float64),
                          np.arange(Ny + 2 * NG, dtype=np.float64),
                          np.arange(Nx + 2 * NG, dtype=np.float64), indexing="ij")
    return I, J, K


def _quad(I, J, K):
    """A quadratic map with every Jacobian entry nonzero, and its Jacobian."""
    x = 0.010 * I + 0.002 * J + 0.001 * K + 1e-4 * I * I + 2e-5 * J * K
    y = 0.001 * I + 0.012 * J + 0.002 * K + 3e-5 * I * J - 1e-4 * J * J
    z = 0.002 * I + 0.001 * J + 0.015 * K + 2e-5 * K * K + 1e-5 * I * K
    A = np.empty(I.shape + (3, 3))
    A[..., 0, 0] = 0.010 + 2e-4 * I
    A[..., 0, 1] = 0.002 + 2e-5 * K
    A[..., 0, 2] = 0.001 + 2e-5 * J
    A[..., 1, 0] = 0.001 + 3e-5 * J
    A[..., 1, 1] = 0.012 + 3e-5 * I - 2e-4 * J
    A[..., 1, 2] = 0.002 + 0 * I
    A[..., 2, 0] = 0.002 + 1e-5 * K
    A[..., 2, 1] = 0.001 + 0 * I
    A[..., 2, 2] = 0.015 + 4e-5 * K + 1e-5 * I
    return x, y, z, A


def test_quadratic_map_metrics_exact():
    I, J, K = _index((12, 10, 9))
    x, y, z, A = _quad(I, J, K)
    m, Jm = mesh.metrics(x, y, z)
    M = np.stack(m, axis=-1).reshape(I.shape + (3, 3))
    det = np.linalg.det(A)
    MA = M @ A
    eye = np.broadcast_to(np.eye(3), MA.shape)
    assert np.max(np.abs(MA - det[..., None, None] * eye)) <= 1e-12 * np.max(np.abs(det))
    assert np.max(np.abs(Jm * det - 1)) <= 1e-11
    # a cubic term along i: CD6 stays exact inside, CD2 is not in the three
    # outermost i layers (so those layers do use the one-sided formulas)
    A3 = A.copy()
    A3[..., 0, 0] += 3e-6 * I * I
    m3, _ = mesh.metrics(x + 1e-6 * I ** 3, y, z)
    M3 = np.stack(m3, axis=-1).reshape(I.shape + (3, 3))
    det3 = np.linalg.det(A3)
    res = np.abs(M3 @ A3 - det3[..., None, None] * eye).max(axis=(-1, -2)) / np.abs(det3).max()
    assert res[:, :, 3:-3].max() <= 1e-12 and res[:, :, :3].min() > 1e-6


def test_ghosts_continue_affine_map():
    N = (11, 9, 8)
    I, J, K = _index(N)
    ref = [0.01 * I + 0.003 * J - 0.002 * K + 0.5, -0.002 * I + 0.02 * J + 0.001 * K,
           0.004 * I - 0.001 * J + 0.03 * K - 1.0]
    xs = [np.full_like(I, np.nan) for _ in range(3)]
    it = (slice(NG, -NG),) * 3
    for a, r in zip(xs, ref):
        a[it] = r[it]
    mesh.ghosts(*xs)
    for a, r in zip(xs, ref):
        assert np.all(np.isfinite(a))
        assert np.max(np.abs(a - r)) < 1e-14


def test_corner_rule_adds_reflections():
    N = (7, 6, 5)
    Nx, Ny, Nz = N
    I, J, K = _index(N)
    f = lambda t: np.sin(0.3 * t) + 0.01 * t ** 3  # noqa: E731
    g = lambda t: np.cos(0.2 * t) * t  # noqa: E731
    h = lambda t: 0.5 * t + 0.02 * t ** 2  # noqa: E731
    v = f(I) + g(J) + h(K)
    w = np.full_like(v, np.nan)
    it = (slice(NG, -NG),) * 3
    w[it] = v[it]
    other = [np.zeros_like(v), np.zeros_like(v)]
    mesh.ghosts(w, *other)

    def refl(fun, t, n):  # 1-D reflection through the boundary nodes NG, n+NG-1
        b = np.where(t < NG, NG, n + NG - 1)
        inside = (t >= NG) & (t < n + NG)
        return np.where(inside, fun(t), 2 * fun(b) - fun(2 * b - t))

    exp = refl(f, I, Nx) + refl(g, J, Ny) + refl(h, K, Nz)
    assert np.max(np.abs(w - exp)) < 1e-12


def _wavy(N, seed=3):
    rng = np.random.default_rng(seed)
    I, J, K = _index(N)
    a = rng.uniform(0.5, 1.0, 6)
    x = 0.01 * I + 0.002 * a[0] * np.sin(0.4 * J) * np.cos(0.3 * K)
    y = 0.012 * J + 0.002 * a[1] * np.sin(0.35 * I + 0.2 * K)
    z = 0.015 * K + 0.002 * a[2] * np.cos(0.25 * I) * np.sin(0.3 * J)
    it = (slice(NG, -NG),) * 3
    out = []
    for v in (x, y, z):
        w = np.zeros_like(v)
        w[it] = v[it]
        out.append(w)
    return out


def _gpu_vs_oracle(N, coords):
    import flame3d_amd.abi as abi
    import flame3d_amd.cases as cases
    import flame3d_amd.mech as mm
    ctx = abi.Context(*N, mm.load("air5"), "fp64")
    xg = [torch.from_numpy(np.ascontiguousarray(c)).cuda() for c in coords]
    grid = cases.grid_from_coords(ctx, *xg)
    torch.cuda.synchronize()
    xo = [c.copy() for c in coords]
    mesh.ghosts(*xo)
    mo, Jo = mesh.metrics(*xo)
    for a, b in zip(xg, xo):
        assert np.array_equal(a.cpu().numpy(), b)
    for q in range(9):
        assert np.array_equal(grid.metrics[q].cpu().numpy(), mo[q]), q
    assert np.array_equal(grid.J.cpu().numpy(), Jo)
    ctx.close()
    return Jo


@pytest.mark.gpu
def test_mesh_gpu_matches_oracle_wavy():
    N = (24, 20, 12)
    print(f"\n[mesh] ghosts + metrics, wavy block {N}", flush=True)
    _gpu_vs_oracle(N, _wavy(N))


@pytest.mark.gpu
def test_mesh_gpu_matches_oracle_reference_mesh():
    """The reference generator's mesh family at 60 x 27 x 27 (x uniform over
    40 nodes then growing, asinh-clustered y and z), then its default
    300 x 127 x 127 with J > 0 at every padded point."""
    import flame3d_amd.cases as cases
    N = (60, 27, 27)
    print(f"\n[mesh] reference mesh family {N}", flush=True)
    coords = [c.numpy() for c in cases.parse_mesh_coords(*N, n_uniform=40)]
    _gpu_vs_oracle(N, coords)
    N = (300, 127, 127)
    print(f"[mesh] reference mesh defaults {N}", flush=True)
    coords = [c.numpy() for c in cases.parse_mesh_coords(*N)]
    Jo = _gpu_vs_oracle(N, coords)
    assert np.all(Jo > 0)
