"""Output formats.  Plot files (.vts, VTK XML structured grid, appended
raw) with the fields of the reference's plot output (solver.jl:394-434):
write, read back, and compare with the state they came from (Float32
rounding of the FP64 state, Int64 tag exact, column-major i-fastest layout,
the time as field data).  Restart files with the reference's checkpoint
datasets (solver.jl:437-465) round-trip bit for bit."""
import numpy as np
import torch

NG = 4


def test_vts_round_trip(tmp_path, pkg):
    import flame3d_amd.cases as cases
    import flame3d_amd.mech as mm
    import flame3d_amd.vtk as vtk
    m = mm.load("air5")
    N = (9, 7, 5)
    grid = cases.stretched_y_grid(*N, 1.0, 2.0, 0.5)
    Q, rhoi = cases.random_state(grid, m, seed=3)
    phi = torch.rand(grid.shape3, dtype=torch.float64)
    tag = torch.randint(0, 3, grid.shape3, dtype=torch.int64)
    path = str(tmp_path / "plt0-10.vts")
    vtk.write_vts(path, grid.x, grid.y, grid.z, Q, rhoi, m.species, phi=phi, tag=tag, time=1.25e-5)
    head = open(path, "rb").read(400).decode(errors="replace")
    assert head.startswith('<?xml version="1.0"?>') and 'type="StructuredGrid"' in head
    f, pts, t = vtk.read_vts(path)
    it = (slice(NG, -NG),) * 3
    assert t == 1.25e-5
    assert pts.shape == (N[2], N[1], N[0], 3)
    for c, v in enumerate((grid.x, grid.y, grid.z)):
        assert np.array_equal(pts[..., c], v[it].numpy().astype(np.float32))
    assert np.array_equal(f["rho"], Q[0][it].numpy().astype(np.float32))
    for c in range(3):
        assert np.array_equal(f["velocity"][..., c], Q[1 + c][it].numpy().astype(np.float32))
    assert np.array_equal(f["p"], Q[4][it].numpy().astype(np.float32))
    assert np.array_equal(f["T"], Q[5][it].numpy().astype(np.float32))
    assert np.array_equal(f["phi"], phi[it].numpy().astype(np.float32))
    for k, sp in enumerate(m.species):  # partial densities under "Y" + name, as the reference
        assert np.array_equal(f["Y" + sp], rhoi[k][it].numpy().astype(np.float32))
    assert f["tag"].dtype == np.int64 and np.array_equal(f["tag"], tag[it].numpy())
    # x index fastest in the payload: the first nx values of rho are the first row
    raw = f["rho"].reshape(-1)
    assert np.array_equal(raw[:N[0]], Q[0][NG, NG, NG:-NG].numpy().astype(np.float32))


def test_checkpoint_round_trip(tmp_path, pkg):
    """Restart files: the reference's dataset names and [Nx_tot, Ny_tot,
    Nz_tot, NV, Nprocs] column-major shapes (solver.jl:437-465), FP64 bit
    for bit back into the library's layout."""
    import flame3d_amd.cases as cases
    import flame3d_amd.checkpoint as chk
    import flame3d_amd.mech as mm
    m = mm.load("h2_lidryer")
    N = (9, 7, 5)
    grid = cases.stretched_y_grid(*N, 1.0, 2.0, 0.5)
    Q, rhoi = cases.random_state(grid, m, seed=5)
    path = str(tmp_path / "chk100.npz")
    chk.save_chk(path, Q, rhoi, step=100, time=2e-5)
    z = np.load(path, allow_pickle=False)
    assert z["Q_h"].shape == (N[0] + 8, N[1] + 8, N[2] + 8, 7, 1)
    assert z["ρi_h"].shape == (N[0] + 8, N[1] + 8, N[2] + 8, m.Ns, 1)
    assert z["Q_h"][5, 3, 2, 4, 0] == Q[4, 2, 3, 5].item()  # [i, j, k, var] indexing
    Q2, r2, step, t = chk.load_chk(path, device="cpu")
    assert step == 100 and t == 2e-5
    assert torch.equal(Q2, Q) and torch.equal(r2, rhoi)
