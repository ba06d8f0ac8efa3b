"""GPU parity of the FP64 chemistry kernels against the oracle.

H2/O2 reactors drawn from the 128^3 chemistry workload's distribution
(T ~ U[900, 2500] K, P ~ U[1, 10] atm, Z ~ Beta(2, 2)) on a small block, and
dissociating air at 3000-6000 K.  Bars:
  * eval_gpu_stiff / eval_gpu: 1e-12 normwise relative error on rho_k, U[5],
    p, T, e;
  * backward-Euler Newton: the integer iteration count of every cell equals
    the oracle's exactly, and the fields match to 1e-12.
"""
import numpy as np
import pytest
import torch

from helpers import rel_err, rel_err_per_var, to_np, to_torch

pytestmark = pytest.mark.gpu
SIZE = (16, 8, 8)


def _cells(mname):
    import flame3d_amd.cases as cases
    import flame3d_amd.mech as mm
    m = mm.load(mname)
    if mname == "h2_lidryer":
        Q, r = cases.chem_cells(SIZE, m, seed=12345)
    else:
        Q, r = cases.chem_cells_air(SIZE, m, seed=4)
    return m, Q, r


def _run(mname, mode, **kw):
    import flame3d_amd.abi as abi
    import flame3d_amd.cases as cases
    from oracle import Oracle
    m, Q, r = _cells(mname)
    U = cases.prim2c(Q)
    print(f"\n[chem] mech={mname} mode={mode} size={SIZE}", flush=True)
    orc = Oracle(*SIZE, m)
    Qo, ro, Uo = to_np(Q), to_np(r), to_np(U)
    ctx = abi.Context(*SIZE, m, "fp64")
    Qg, rg, Ug = Q.cuda(), r.cuda(), U.cuda()
    dt = kw.pop("dt", 1e-7)
    it_o = it_g = None
    if mode == "stiff":
        orc.chem_stiff(Uo, Qo, ro, dt)
        ctx.chem_stiff(Ug, Qg, rg, dt)
    elif mode == "explicit":
        orc.chem_explicit(Uo, Qo, ro, dt)
        ctx.chem_explicit(Ug, Qg, rg, dt)
    else:
        it_o = orc.chem_newton(Uo, Qo, ro, dt, **kw)
        it_g = torch.zeros((SIZE[2], SIZE[1], SIZE[0]), dtype=torch.int32, device="cuda")
        ctx.chem_newton(Ug, Qg, rg, dt, iters=it_g, **kw)
    torch.cuda.synchronize()
    ctx.close()
    return m, (Qo, ro, Uo, it_o), (to_np(Qg), to_np(rg), to_np(Ug),
                                   None if it_g is None else it_g.cpu().numpy().T)


def _compare(ref, got, tol=1e-12):
    Qo, ro, Uo, _ = ref
    Qg, rg, Ug, _ = got
    e_r = max(rel_err_per_var(rg, ro))
    e_U = rel_err(Ug[..., 4], Uo[..., 4])
    e_Q = max(rel_err(Qg[..., n], Qo[..., n]) for n in (4, 5, 6))
    same = np.array_equal(rg, ro) and np.array_equal(Ug[..., 4], Uo[..., 4])
    print(f"  rho_k {e_r:.2e}  U5 {e_U:.2e}  p/T/e {e_Q:.2e}  bit-identical {same}", flush=True)
    assert e_r <= tol and e_U <= tol and e_Q <= tol, (e_r, e_U, e_Q)


@pytest.mark.parametrize("mname", ["h2_lidryer", "air5"])
def test_chem_stiff(mname):
    m, ref, got = _run(mname, "stiff")
    # the reaction step changed the state (T_criteria = 500 K is below every cell)
    assert np.max(np.abs(ref[1] - to_np(_cells(mname)[2]))) > 0
    _compare(ref, got)


@pytest.mark.parametrize("mname", ["h2_lidryer", "air5"])
def test_chem_explicit(mname):
    _, ref, got = _run(mname, "explicit", dt=1e-9)
    _compare(ref, got)


@pytest.mark.parametrize("mname", ["h2_lidryer", "air5"])
def test_chem_newton_iteration_counts(mname):
    _, ref, got = _run(mname, "newton", maxiter=50, rtol=1e-10, ttol=1e-6)
    it_o, it_g = ref[3], got[3]
    hist = np.bincount(it_o.ravel())
    print(f"  Newton iterations histogram {dict(enumerate(hist.tolist()))}", flush=True)
    assert it_o.max() < 50, "Newton did not converge in some cells"
    assert np.array_equal(it_o, it_g), np.argwhere(it_o != it_g)[:10]
    _compare(ref, got)


def test_chem_newton_restart_mid_ignition():
    """Reactors mid-ignition at high pressure (nine Newton steps of 1e-7 s
    from the H2/O2 distribution, seed 5): one cell's single backward-Euler
    Newton leaves the admissible set (T < 0 on the second iterate) and
    restarts with substeps.  Iteration counts (restarts included) and fields
    must match the oracle exactly / to 1e-12, and no cell may fail."""
    import flame3d_amd.abi as abi
    import flame3d_amd.cases as cases
    import flame3d_amd.mech as mm
    from oracle import Oracle
    size = (16, 16, 8)
    print(f"\n[chem] newton restart size={size}", flush=True)
    m = mm.load("h2_lidryer")
    Q, r = cases.chem_cells(size, m, seed=5)
    U = cases.prim2c(Q)
    orc = Oracle(*size, m)
    Qo, ro, Uo = to_np(Q), to_np(r), to_np(U)
    for _ in range(9):
        orc.chem_newton(Uo, Qo, ro, 1e-7)
    Qg, rg, Ug = to_torch(Qo).cuda(), to_torch(ro).cuda(), to_torch(Uo).cuda()
    it_o = orc.chem_newton(Uo, Qo, ro, 1e-7)
    ctx = abi.Context(*size, m, "fp64")
    it_g = torch.zeros((size[2], size[1], size[0]), dtype=torch.int32, device="cuda")
    ctx.chem_newton(Ug, Qg, rg, 1e-7, iters=it_g)
    torch.cuda.synchronize()
    it_g = it_g.cpu().numpy().T
    assert it_o.min() >= 1 and it_o.max() >= 10
    assert np.array_equal(it_o, it_g), np.argwhere(it_o != it_g)[:10]
    _compare((Qo, ro, Uo, None), (to_np(Qg), to_np(rg), to_np(Ug), None))
    ctx.close()


def test_chem_newton_nasa7_breakpoint_gap():
    """The states of tests/golden/newton_nasa_gap.json (energy inside the
    NASA-7 jump at 1000 K) on the GPU: same iteration counts as the oracle,
    fields to 1e-12, no failed cell."""
    import json
    import os
    import flame3d_amd.abi as abi
    import flame3d_amd.mech as mm
    from oracle import Oracle
    print("\n[chem] newton at the NASA-7 breakpoint", flush=True)
    path = os.path.join(os.path.dirname(__file__), "golden", "newton_nasa_gap.json")
    cells = [[float.fromhex(x) for x in row] for row in json.load(open(path))["cells"]]
    m = mm.load("h2_lidryer")
    size = (4, 4, 4)  # the three states repeated over the block
    Q = np.zeros((12, 12, 12, 7), order="F")
    R = np.zeros((12, 12, 12, 9), order="F")
    U = np.zeros((12, 12, 12, 5), order="F")
    for i in range(4):
        for j in range(4):
            for k in range(4):
                row = cells[(i + j + k) % len(cells)]
                Q[4 + i, 4 + j, 4 + k] = row[:7]
                R[4 + i, 4 + j, 4 + k] = row[7:16]
                U[4 + i, 4 + j, 4 + k] = row[16:21]
    Qg, Rg, Ug = to_torch(Q), to_torch(R), to_torch(U)
    it_o = Oracle(*size, m).chem_newton(U, Q, R, 1e-7)
    ctx = abi.Context(*size, m, "fp64")
    it_g = torch.zeros((4, 4, 4), dtype=torch.int32, device="cuda")
    ctx.chem_newton(Ug, Qg, Rg, 1e-7, iters=it_g)
    torch.cuda.synchronize()
    it_g = it_g.cpu().numpy().T
    assert it_o.min() >= 1 and np.array_equal(it_o, it_g), (it_o.ravel(), it_g.ravel())
    _compare((Q, R, U, None), (to_np(Qg), to_np(Rg), to_np(Ug), None))
    ctx.close()


def test_chem_newton_full_workload_128():
    """The pointwise H2/O2 chemistry workload at its stated size: 128^3 =
    2,097,152 independent reactors (T ~ U[900, 2500] K, P ~ U[1, 10] atm,
    mixture fraction ~ Beta(2, 2), seed 12345), one backward-Euler step of
    dt = 1e-7 s by Newton with the dense 10x10 FP64 LU.  The integer
    iteration count of every reactor equals the oracle's (OpenMP over the host
    cores), restarted substeps included, and every field is bit-identical:
    both sides round each product and sum (no contraction) and use the same
    exp / log.  (With the vendor exp / log and contraction on the GPU, 19 of
    the reactors -- mid-ignition states that need substep restarts -- took
    another number of restarts than the oracle and ended elsewhere.)"""
    import time
    import flame3d_amd.abi as abi
    import flame3d_amd.cases as cases
    import flame3d_amd.mech as mm
    from oracle import Oracle
    N = (128, 128, 128)
    m = mm.load("h2_lidryer")
    print(f"\n[chem] newton, full workload {N} seed 12345", flush=True)
    Q, r = cases.chem_cells(N, m, seed=12345)
    U = cases.prim2c(Q)
    ctx = abi.Context(*N, m, "fp64")
    Qg, rg, Ug = Q.cuda(), r.cuda(), U.cuda()
    it_g = torch.zeros((N[2], N[1], N[0]), dtype=torch.int32, device="cuda")
    ctx.chem_newton(Ug, Qg, rg, 1e-7, iters=it_g)
    torch.cuda.synchronize()
    ctx.close()
    it_g = it_g.cpu().numpy().T
    Qg, rg, Ug = to_np(Qg), to_np(rg), to_np(Ug)
    orc = Oracle(*N, m)
    Qo, ro, Uo = to_np(Q), to_np(r), to_np(U)
    del Q, r, U
    t0 = time.time()
    it_o = orc.chem_newton(Uo, Qo, ro, 1e-7)
    print(f"  oracle: {time.time() - t0:.1f} s on {orc.num_threads()} threads", flush=True)
    mism = int(np.sum(it_o != it_g))
    vals, counts = np.unique(it_o, return_counts=True)
    print(f"  iterations per reactor {dict(zip(vals.tolist(), counts.tolist()))}; "
          f"mismatched counts {mism}", flush=True)
    assert it_o.min() >= 1  # every reactor converged, none was left unreacted
    assert mism == 0, np.argwhere(it_o != it_g)[:10]
    same = [np.array_equal(a, b) for a, b in ((rg, ro), (Ug, Uo), (Qg, Qo))]
    print(f"  bit-identical rho_k, U, Q: {same}", flush=True)
    assert all(same)


@pytest.mark.parametrize("mname", ["h2_lidryer", "air5"])
def test_kinetics_tables_and_runtime_parameters(mname):
    """The shipped mechanisms run on their compiled-in reaction tables
    (f3d_kinetics_source); a mechanism that differs in one rate parameter
    takes the run-time parameter block.  Both paths are checked against the
    oracle, with its own parameters each, to the Newton mode's bars (every
    iteration count equal, fields bit-identical)."""
    import dataclasses
    import flame3d_amd.abi as abi
    import flame3d_amd.cases as cases
    from oracle import Oracle
    m, Q, r = _cells(mname)
    want = abi.KIN_H2_LIDRYER if mname == "h2_lidryer" else abi.KIN_AIR5
    Arr = np.array(m.Arr, copy=True)
    Arr[0, 0] *= 1.5  # the first reaction's pre-exponential factor
    m2 = dataclasses.replace(m, Arr=Arr)
    for mech, kind in ((m, want), (m2, abi.KIN_RUNTIME)):
        print(f"\n[chem] kinetics source mech={mname} expect={kind}", flush=True)
        ctx = abi.Context(*SIZE, mech, "fp64")
        assert ctx.kinetics_source() == kind
        U = cases.prim2c(Q)
        Qg, rg, Ug = Q.cuda(), r.cuda(), U.cuda()
        it_g = torch.zeros((SIZE[2], SIZE[1], SIZE[0]), dtype=torch.int32, device="cuda")
        ctx.chem_newton(Ug, Qg, rg, 1e-7, iters=it_g)
        torch.cuda.synchronize()
        ctx.close()
        orc = Oracle(*SIZE, mech)
        Qo, ro, Uo = to_np(Q), to_np(r), to_np(U)
        it_o = orc.chem_newton(Uo, Qo, ro, 1e-7)
        assert np.array_equal(it_g.cpu().numpy().T, it_o)
        assert np.array_equal(to_np(rg), ro) and np.array_equal(to_np(Ug), Uo)
        assert np.array_equal(to_np(Qg), Qo)
