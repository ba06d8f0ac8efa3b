"""Shared helpers for the tests: layout conversion between the library's
torch layout (NV, Nz+8, Ny+8, Nx+8) and the oracle's numpy Fortran layout
(Nx+8, Ny+8, Nz+8, NV), and error norms."""
import numpy as np
import torch


def to_np(t: torch.Tensor) -> np.ndarray:
    """Device/host tensor -> writable Fortran-ordered float64 numpy array in
    the reference's [i, j, k, var] index order."""
    a = t.detach().cpu().numpy()
    return np.array(a.T, order="F", copy=True)


def to_torch(a: np.ndarray, device="cuda", dtype=torch.float64) -> torch.Tensor:
    return torch.from_numpy(np.ascontiguousarray(a.T)).to(device=device, dtype=dtype)


def rel_err(got, ref):
    """Normwise relative error max|got - ref| / max|ref| (0 if ref == 0)."""
    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    scale = np.max(np.abs(ref))
    d = np.max(np.abs(got - ref)) if ref.size else 0.0
    return 0.0 if scale == 0 else d / scale


def rel_err_per_var(got, ref):
    """Normwise relative error per variable (last axis)."""
    return [rel_err(got[..., n], ref[..., n]) for n in range(ref.shape[-1])]


def l2_linf(got, ref):
    d = np.asarray(got, dtype=np.float64) - np.asarray(ref, dtype=np.float64)
    return float(np.sqrt(np.mean(d * d))), float(np.max(np.abs(d)))


def branch_codes(phi: torch.Tensor, phi1: float, phi2: float, NG: int = 4):
    """Hybrid-switch branch of every interior face of the x, y and z sweeps
    from one stage's sensor phi ((Nz+8, Ny+8, Nx+8) device view): face
    i-1/2 takes phif = max(phi[i-1], phi[i]) and the branch 0 (upwind) for
    phif < phi1, 1 (WENO) for phif < phi2, else 2 (NND) (schemes.jl:96-100).
    Compared in phi's own precision, as the kernels compare."""
    Nzt, Nyt, Nxt = phi.shape
    Nx, Ny, Nz = Nxt - 2 * NG, Nyt - 2 * NG, Nzt - 2 * NG
    p = phi
    a1 = torch.tensor(phi1, dtype=p.dtype, device=p.device)
    a2 = torch.tensor(phi2, dtype=p.dtype, device=p.device)
    fx = torch.maximum(p[NG:NG + Nz, NG:NG + Ny, NG - 1:NG + Nx], p[NG:NG + Nz, NG:NG + Ny, NG:NG + Nx + 1])
    fy = torch.maximum(p[NG:NG + Nz, NG - 1:NG + Ny, NG:NG + Nx], p[NG:NG + Nz, NG:NG + Ny + 1, NG:NG + Nx])
    fz = torch.maximum(p[NG - 1:NG + Nz, NG:NG + Ny, NG:NG + Nx], p[NG:NG + Nz + 1, NG:NG + Ny, NG:NG + Nx])
    return [(f >= a1).to(torch.int8) + (f >= a2).to(torch.int8) for f in (fx, fy, fz)]


def flip_mask(codes_a, codes_b, radius: int = 2):
    """Interior cells (Nz, Ny, Nx) within `radius` cells of a face whose
    hybrid-switch branch differs between two runs at any stage (codes_a,
    codes_b: per-stage lists of branch_codes); also the number of such
    face-stage flips."""
    fx0 = codes_a[0][0]
    Nz, Ny, Nx1 = fx0.shape
    m = torch.zeros((Nz, Ny, Nx1 - 1), dtype=torch.bool, device=fx0.device)
    flips = 0
    for ca, cb in zip(codes_a, codes_b):
        for d, (x, y) in enumerate(zip(ca, cb)):
            f = x != y
            flips += int(f.sum())
            # a face between cells c-1 and c marks both
            if d == 0:
                m |= f[:, :, :-1] | f[:, :, 1:]
            elif d == 1:
                m |= f[:, :-1, :] | f[:, 1:, :]
            else:
                m |= f[:-1] | f[1:]
    if radius > 0:
        k = 2 * radius + 1
        m = torch.nn.functional.max_pool3d(m[None, None].float(), k, stride=1,
                                           padding=radius)[0, 0] > 0
    return m, flips
