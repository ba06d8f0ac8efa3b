"""FP64 numpy restatement of the reference's mesh preprocessing.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the library's
f3d_mesh_ghosts / f3d_mesh_metrics; never by the product path.

Arrays are padded blocks indexed [k, j, i] (x index fastest in memory, the
reference's column-major [i, j, k]), NG = 4 ghost layers.

  ghosts(x, y, z)   generateXYZ's ghost nodes (parse_mesh.jl:52-113)
  metrics(x, y, z)  computeMetrics (parse_mesh.jl:116-221): CD6 / CD2_L /
                    CD2_R differences along each index, J = 1/det, the nine
                    cofactor metrics in the order dxi/dx, dxi/dy, ..., dzeta/dz

Every expression keeps the reference's operation order, so a device
evaluation without fused multiply-adds matches it bit for bit.

Pinned by (tests/test_mesh.py): the differences are exact for polynomials of
degree <= 6 (CD6) and <= 2 (CD2), so a quadratic node map gives the analytic
cofactors and 1/det at every padded point to rounding, with every sign and
index of the cofactor formulas exercised by a map whose Jacobian has no zero
entry; an affine map's reflected ghosts continue it exactly (reflection and
corner rules), and a map additive in (i, j) has exact corners.
"""
from __future__ import annotations

import numpy as np

NG = 4


def ghosts(x, y, z, ng: int = NG):
    """Fill the ghost nodes of padded x, y, z in place (parse_mesh.jl:52-113)."""
    Nzt, Nyt, Nxt = x.shape
    Nx, Ny, Nz = Nxt - 2 * ng, Nyt - 2 * ng, Nzt - 2 * ng
    ki = slice(ng, Nz + ng)
    for v in (x, y, z):
        # x ghosts, interior j and k (parse_mesh.jl:53-63)
        for I in range(ng):
            v[ki, ng:Ny + ng, I] = 2 * v[ki, ng:Ny + ng, ng] - v[ki, ng:Ny + ng, 2 * ng - I]
        for I in range(Nx + ng, Nxt):
            b = Nx + ng - 1
            v[ki, ng:Ny + ng, I] = 2 * v[ki, ng:Ny + ng, b] - v[ki, ng:Ny + ng, 2 * b - I]
        # y ghosts, interior i and k (parse_mesh.jl:65-75)
        for J in range(ng):
            v[ki, J, ng:Nx + ng] = 2 * v[ki, ng, ng:Nx + ng] - v[ki, 2 * ng - J, ng:Nx + ng]
        for J in range(Ny + ng, Nyt):
            b = Ny + ng - 1
            v[ki, J, ng:Nx + ng] = 2 * v[ki, b, ng:Nx + ng] - v[ki, 2 * b - J, ng:Nx + ng]
        # corners (parse_mesh.jl:77-100): v[i, j] = v[i, jb] + v[ib, j] - v[ib, jb]
        for I in list(range(ng)) + list(range(Nx + ng, Nxt)):
            ib = ng if I < ng else Nx + ng - 1
            for J in list(range(ng)) + list(range(Ny + ng, Nyt)):
                jb = ng if J < ng else Ny + ng - 1
                v[ki, J, I] = v[ki, jb, I] + v[ki, J, ib] - v[ki, jb, ib]
        # z ghosts over the whole x-y extent (parse_mesh.jl:102-113)
        for K in range(ng):
            v[K] = 2 * v[ng] - v[2 * ng - K]
        for K in range(Nz + ng, Nzt):
            b = Nz + ng - 1
            v[K] = 2 * v[b] - v[2 * b - K]


def _d(v, axis):
    """d v / d(index) along `axis` (parse_mesh.jl:116-130, 155-207)."""
    a = np.moveaxis(v, axis, -1)
    n = a.shape[-1]
    out = np.empty_like(a)
    f = [a[..., m:n - 6 + m] for m in range(7)]  # f[m] = v[i - 3 + m] for i in 3..n-4
    out[..., 3:n - 3] = 1 / 60 * (f[6] - f[0]) - 3 / 20 * (f[5] - f[1]) + 3 / 4 * (f[4] - f[2])
    for i in range(3):  # CD2_L on v[i:i+3]
        out[..., i] = 2 * a[..., i + 1] - 0.5 * a[..., i + 2] - 1.5 * a[..., i]
    for i in range(n - 3, n):  # CD2_R on v[i-2:i+1]
        out[..., i] = -2 * a[..., i - 1] + 0.5 * a[..., i - 2] + 1.5 * a[..., i]
    return np.moveaxis(out, -1, axis)


def metrics(x, y, z):
    """(list of 9 cofactor metrics, J) at every padded point."""
    # axis 2 = i (xi), 1 = j (eta), 0 = k (zeta)
    xxi, yxi, zxi = _d(x, 2), _d(y, 2), _d(z, 2)
    xet, yet, zet = _d(x, 1), _d(y, 1), _d(z, 1)
    xze, yze, zze = _d(x, 0), _d(y, 0), _d(z, 0)
    J = 1 / (xxi * (yet * zze - yze * zet) - xet * (yxi * zze - yze * zxi)
             + xze * (yxi * zet - yet * zxi))
    m = [yet * zze - yze * zet,
         xze * zet - xet * zze,
         xet * yze - xze * yet,
         yze * zxi - yxi * zze,
         xxi * zze - xze * zxi,
         xze * yxi - xxi * yze,
         yxi * zet - yet * zxi,
         xet * zxi - xxi * zet,
         xxi * yet - xet * yxi]
    return m, J
