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
