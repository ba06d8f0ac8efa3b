"""Restart files with the reference's checkpoint layout (solver.jl:437-465):
datasets "Q_h" [Nx_tot, Ny_tot, Nz_tot, Nprim, Nprocs] and "ρi_h"
[Nx_tot, Ny_tot, Nz_tot, Nspecs, Nprocs], Float64, column-major, whole
padded blocks (ghosts included), one slot per rank (here one).  The
reference writes them with HDF5; h5py is not in this image, so the same
datasets go into a NumPy .npz archive (no pickled objects), with the step
and time added.  Loading returns device tensors in the library's layout
(NV, Nz+8, Ny+8, Nx+8)."""
from __future__ import annotations

import numpy as np


def save_chk(path: str, Q, rhoi, step: int = 0, time: float = 0.0):
    """Write Q (7, Nz+8, Ny+8, Nx+8) and rhoi (Ns, ...) FP64 tensors."""
    q = Q.detach().cpu().numpy()
    r = rhoi.detach().cpu().numpy()
    # (NV, K, J, I) C-order -> [I, J, K, NV, 1] column-major: the same bytes
    # as the reference's array, reshaped with a trailing rank axis
    qh = np.asfortranarray(q.transpose(3, 2, 1, 0)[..., None])
    rh = np.asfortranarray(r.transpose(3, 2, 1, 0)[..., None])
    with open(path, "wb") as fh:
        np.savez(fh, **{"Q_h": qh, "ρi_h": rh, "step": np.int64(step), "time": np.float64(time)})


def load_chk(path: str, device: str = "cuda", rank: int = 0):
    """(Q, rhoi, step, time) from a file written by save_chk."""
    import torch
    z = np.load(path, allow_pickle=False)
    qh, rh = z["Q_h"][..., rank], z["ρi_h"][..., rank]
    Q = torch.from_numpy(np.ascontiguousarray(qh.transpose(3, 2, 1, 0))).to(device)
    r = torch.from_numpy(np.ascontiguousarray(rh.transpose(3, 2, 1, 0))).to(device)
    return Q, r, int(z["step"]), float(z["time"])
