"""Benchmark: cell-updates/s per RK stage of the per-stage RHS with chemistry,
FP32 flux + learned chemistry vs the plain FP64 method, on one MI355X.

Workload: the reacting H2/air shock-laden mixing layer on a 256^3 curvilinear
block (tanh-stretched y, factor 2.5; Ma 1.5 air at 1200 K over H2 at 1000 K,
3 atm, normal shock): 9-species H2/O2 mechanism, hybrid UP7 /
WENO5-Z / NND with the pressure sensor, viscous + species diffusion, and
Strang-split chemistry (solver.jl:240-378): per time step, chemistry over
dt/2, the three RK3 stages, chemistry over dt/2 (dt = 2e-7 s at CFL 0.5).
Inputs are generated in HBM (synthetic, seeded).

  * FP64 plain: FP64 stage arithmetic and the FP64 backward-Euler Newton
    chemistry on every cell;
  * FP32 + NN (`value`): FP32 flux / viscous / divergence arithmetic (FP64
    state, FP64 e -> T Newton) and the BF16 MFMA MLP chemistry step with its
    gates and the FP64 Newton fallback.

Both start from the same state; after the timed steps the FP32+NN fields are
compared with the FP64-plain ones (the FP64 path matches the CPU oracle to
1e-12 in tests/).  Cell-updates per RK stage = cells x 3 / time step.

    python bench.py [--gpus 1] [--steps 20] [--warmup 5] [--size 256]
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, vendor figure
FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 (vector / matrix), vendor figure
BF16_PEAK_TFLOPS = 2500.0  # MI355X dense BF16 MFMA, vendor figure
MLP_FLOP_PER_CELL = 2 * (11 * 64 + 64 * 64 + 64 * 64 + 64 * 9)  # 11-64-64-64-9 forward
OPCOUNT_FILE = os.path.join(ROOT, "profiles", "chem_opcount.json")


def _package():
    if "flame3d_amd" in sys.modules:
        return sys.modules["flame3d_amd"]
    pkg_dir = os.path.join(ROOT, "flame3d.jl_amd")
    spec = importlib.util.spec_from_file_location(
        "flame3d_amd", os.path.join(pkg_dir, "__init__.py"), submodule_search_locations=[pkg_dir])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["flame3d_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# ---------------------------------------------------------------------------
# Algorithmic bytes per launch of each kernel: the compulsory reads and
# writes of its inputs/outputs once each (state in double; metrics and J in
# M bytes: double on the FP64 path, the float copies on the FP32 path;
# intermediates in the path's precision R bytes).  npts = padded points,
# nf = faces of the sweep, nc = interior cells.
# ---------------------------------------------------------------------------
def kernel_bytes(name, R, Ns, npts, nf, nc):
    D = 8
    F = 4
    M = R
    table = {
        "flux_split": npts * (6 * D + 3 * M + 10 * R),       # Q(6) + 3 metrics -> Fp, Fm (5 each)
        "species_split": npts * ((3 + Ns) * D + 3 * M + 2 * Ns * R),  # u,v,w, rho_k, 3 metrics
        "shock_sensor": npts * (D + R),                        # p -> phi
        "weno5": nf * 16 * R,                                  # Fp, Fm (10) + phi -> F (5)
        "weno_species": nf * (2 * Ns + 1 + Ns) * R,
        "viscous": nf * (4 * D + 10 * M + 9 * R),              # u,v,w,T; 9 metrics + J; lam, mu, Fh(3) -> Fv(4)
        "spec_viscous": nf * (2 * D + 10 * M + (3 * Ns + 3) * R),  # rho, T; metrics, J; Yi, D -> Fd, Fh
        "mixture": npts * ((3 + Ns) * D + (2 * Ns + 2) * R),
        "div": nc * (27 * R + 11 * D),                         # 3x(5+4) faces, J, U RMW
        "div_specs": nc * (6 * Ns * R + (1 + 2 * Ns) * D),
        "c2prim": nc * ((10 + 2 * Ns + 2) * D),                # U, rho_k -> Q(7), rho_k
        # fused float stage (unit: interior cell); the convective sweep reads
        # Q (rho, u, v, w, p, e), rho_k, the sweep's 3 metrics and phi and
        # stores (x) or updates (y, z: read + write) the 5 + Ns right-hand sides
        "sweep_conv": nc * ((6 + Ns) * D + 4 * F + (5 + Ns) * F * (1 + 2 / 3)),
        # viscous: u, v, w, T, 9 metrics + J, mu, lambda, rhs 1..4 updated;
        # species: rho, T, Y_k, D_k, 9 metrics + J, rhs 4 and 5.. updated
        "sweep_diff": nc * (4 * D + 12 * F + 8 * F + 2 * D + 2 * Ns * F + 10 * F
                            + 2 * (1 + Ns) * F),
        # rhs, J, U and rho_k read and written, Un and rho_n read (stages 2, 3:
        # 2/3 of the stages), Q written and T read
        "stage_end": nc * ((5 + Ns) * F + D + 2 * (5 + Ns) * D + (5 + Ns) * D * 2 / 3 + 8 * D),
    }
    return table[name]


# ---------------------------------------------------------------------------
# HBM traffic from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs
# ---------------------------------------------------------------------------
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_reacting_mixing_layer_256.json")


def _pmc_avg(path):
    import csv
    import re
    agg = {}
    for r in csv.DictReader(open(path)):
        short = re.sub(r"\(.*", "", r["Kernel_Name"]).strip()
        agg.setdefault(short, []).append(float(r["Counter_Value"]))
    return {k: sum(v) / len(v) for k, v in agg.items()}


def _pmc_multi(path):
    """{kernel: {counter: mean per dispatch}} of a multi-counter run."""
    import csv
    import re
    agg = {}
    for r in csv.DictReader(open(path)):
        short = re.sub(r"\(.*", "", r["Kernel_Name"]).strip()
        agg.setdefault(short, {}).setdefault(r["Counter_Name"], []).append(float(r["Counter_Value"]))
    return {k: {c: sum(v) / len(v) for c, v in cs.items()} for k, cs in agg.items()}


def valu_summary(d):
    """Per-kernel instruction counters per dispatch from the runs pmc_sq1 and
    pmc_sq2 under `d` (SQ counters, summed over the GPU by rocprofv3; one
    count per wave instruction): FP64 flop = 64 (ADD + MUL + 2 FMA) (the
    transcendental helpers TRANS_F64 listed apart), BF16 MFMA flop =
    512 MOPS, the share of the waves' lifetime spent issuing VALU
    instructions (ACTIVE_INST_VALU / WAVE_CYCLES) and waiting
    (WAIT_INST_ANY / WAVE_CYCLES)."""
    runs = {}
    for sub in ("pmc_sq1", "pmc_sq2"):
        path = os.path.join(d, sub, "pmc_counter_collection.csv")
        if os.path.exists(path):
            for k, cs in _pmc_multi(path).items():
                runs.setdefault(k, {}).update(cs)
    out = {}
    keep = ("SQ_INSTS_VALU", "SQ_INSTS_SALU", "SQ_WAVES", "SQ_INSTS_VALU_TRANS_F64",
            "SQ_INSTS_VMEM", "SQ_INSTS_LDS")
    for k, c in runs.items():
        if "f3d::" not in k:
            continue
        e = {x: c[x] for x in keep if x in c}
        if all(x in c for x in ("SQ_INSTS_VALU_ADD_F64", "SQ_INSTS_VALU_MUL_F64", "SQ_INSTS_VALU_FMA_F64")):
            e["fp64_flop"] = 64.0 * (c["SQ_INSTS_VALU_ADD_F64"] + c["SQ_INSTS_VALU_MUL_F64"] +
                                     2.0 * c["SQ_INSTS_VALU_FMA_F64"])
        if "SQ_INSTS_VALU_MFMA_MOPS_BF16" in c:
            e["bf16_mfma_flop"] = 512.0 * c["SQ_INSTS_VALU_MFMA_MOPS_BF16"]
        if c.get("SQ_WAVE_CYCLES"):
            if "SQ_ACTIVE_INST_VALU" in c:
                e["valu_active_per_wave_cycle"] = c["SQ_ACTIVE_INST_VALU"] / c["SQ_WAVE_CYCLES"]
            if "SQ_WAIT_INST_ANY" in c:
                e["wait_per_wave_cycle"] = c["SQ_WAIT_INST_ANY"] / c["SQ_WAVE_CYCLES"]
        out[k] = e
    return out


def pmc_summary(d):
    """Per-kernel HBM bytes per dispatch from the counter CSVs of four runs
    under `d` (pmc_fetch, pmc_write: the benchmark; copy_fetch, copy_write:
    --calibrate-copy, a 1 GiB device copy).  FETCH_SIZE / WRITE_SIZE are in
    KiB per dispatch; each is scaled by (known copy bytes / counter reading
    of the copy kernel).  With the SQ-counter runs present (valu_summary),
    their per-kernel instruction counts are added under "valu"."""
    f = _pmc_avg(os.path.join(d, "pmc_fetch", "pmc_counter_collection.csv"))
    w = _pmc_avg(os.path.join(d, "pmc_write", "pmc_counter_collection.csv"))
    cf = _pmc_avg(os.path.join(d, "copy_fetch", "pmc_counter_collection.csv"))
    cw = _pmc_avg(os.path.join(d, "copy_write", "pmc_counter_collection.csv"))
    copy = "__amd_rocclr_copyBuffer"
    nbytes = float(1 << 30)
    kf = nbytes / (cf[copy] * 1024.0)
    kw = nbytes / (cw[copy] * 1024.0)
    out = {"fetch_scale": kf, "write_scale": kw,
           "copy_counter_KiB": {"fetch": cf[copy], "write": cw[copy]}, "kernels": {}}
    for k in f:
        if "f3d::" not in k:
            continue
        out["kernels"][k] = {"fetch_bytes": f[k] * 1024.0 * kf,
                             "write_bytes": w.get(k, 0.0) * 1024.0 * kw}
    v = valu_summary(d)
    if v:
        out["valu"] = v
    return out


def traffic_for(symbols):
    """HBM bytes per launch (read + write), launch-weighted over `symbols`."""
    if not symbols or not os.path.exists(TRAFFIC_FILE):
        return None
    ks = json.load(open(TRAFFIC_FILE))["kernels"]
    tot = 0.0
    for sym, w in symbols:
        t = ks.get(sym)
        if t is None:
            return None
        tot += w * (t["fetch_bytes"] + t["write_bytes"])
    return tot


def kernel_symbols(name, Ns):
    """(rocprofv3 kernel symbol, share of the launches) of each timed kernel on
    the FP32 path: x sweeps use the per-face kernels, y and z sweeps the
    marching ones.  None where one symbol serves launches of different sizes
    (the WENO kernels run for 5 and for Ns variables)."""
    return {
        "flux_split": [("void f3d::k_flux_split<float, float>", 1.0)],
        "species_split": [("void f3d::k_species_split<float, float>", 1.0)],
        "viscous": [("void f3d::k_viscous<float, float>", 1 / 3),
                    ("void f3d::k_viscous_march<float>", 2 / 3)],
        "spec_viscous": [(f"void f3d::k_spec_viscous<{Ns}, float, float>", 1 / 3),
                         (f"void f3d::k_spec_viscous_march<{Ns}, float>", 2 / 3)],
        "shock_sensor": [("void f3d::k_shock_sensor<float>", 1.0)],
        "mixture": [(f"void f3d::k_mixture<{Ns}, float>", 1.0)],
        "sweep_conv": [(f"void f3d::k_sw_conv_x<{Ns}>", 1 / 3),
                       (f"void f3d::k_sw_conv_m<{Ns}>", 2 / 3)],
        "sweep_diff": [("f3d::k_sw_visc_x", 1 / 3), (f"void f3d::k_sw_spec_x<{Ns}>", 1 / 3),
                       ("f3d::k_sw_visc_m", 2 / 3), (f"void f3d::k_sw_spec_m<{Ns}>", 2 / 3)],
        "stage_end": [(f"void f3d::k_sw_stage_end<{Ns}>", 1.0)],
    }.get(name)


def time_kernels(ctx, torch, Q, rhoi, metrics, J, reps=10):
    """Average duration of one launch of each kernel (HIP events recorded on
    the stream the kernels are launched on), over its x, y and z sweeps, with
    the inputs a stage left behind; returns {name: (ms, launches per stage)}."""
    Ns = ctx.Ns
    w = ctx.workspace
    ax = "xyz"

    def split(d):
        return lambda: ctx.flux_split(Q, w("Fp"), w("Fm"), *metrics[3 * d:3 * d + 3])

    def ssplit(d):
        return lambda: ctx.species_split(rhoi, Q, w("Fp_i"), w("Fm_i"), *metrics[3 * d:3 * d + 3])

    def weno(d):
        return lambda: ctx.weno(d, w("F" + ax[d]), w("phi"), w("Fp"), w("Fm"), 5)

    def wenos(d):
        return lambda: ctx.weno(d, w("F" + ax[d] + "_i"), w("phi"), w("Fp_i"), w("Fm_i"), Ns)

    def visc(d):
        return lambda: ctx.viscous_flux(d, w("Fv" + ax[d]), Q, metrics, J, w("lambda"), w("mu"),
                                        w("Fh" + ax[d]))

    def svisc(d):
        return lambda: ctx.spec_viscous_flux(d, w("Fd" + ax[d]), Q, w("Yi"), metrics, J, w("D"),
                                             w("Fh" + ax[d]))

    calls = {
        "flux_split": [split(d) for d in range(3)],
        "species_split": [ssplit(d) for d in range(3)],
        "weno5": [weno(d) for d in range(3)],
        "weno_species": [wenos(d) for d in range(3)],
        "viscous": [visc(d) for d in range(3)],
        "spec_viscous": [svisc(d) for d in range(3)],
        "shock_sensor": [lambda: ctx.shock_sensor(w("phi"), Q)],
        "mixture": [lambda: ctx.mixture(Q, rhoi, w("Yi"), w("lambda"), w("mu"), w("D"))],
    }
    out = {}
    for name, fns in calls.items():
        for fn in fns:
            fn()
        torch.cuda.synchronize()
        # the fastest of five timed batches: a host stall while the launches
        # are enqueued (the kernels take ~0.3 ms) would otherwise count as GPU time
        best = float("inf")
        for _ in range(5):
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                for fn in fns:
                    fn()
            e1.record()
            torch.cuda.synchronize()
            best = min(best, e0.elapsed_time(e1) / (reps * len(fns)))
        out[name] = (best, len(fns))
    return out


def _best_of(torch, fns, reps):
    """Mean launch duration of `fns` (HIP events on the launch stream), the
    fastest of five timed batches of `reps` rounds (a host stall while the
    launches are enqueued would otherwise count as GPU time)."""
    for fn in fns:
        fn()
    torch.cuda.synchronize()
    best = float("inf")
    for _ in range(5):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            for fn in fns:
                fn()
        e1.record()
        torch.cuda.synchronize()
        best = min(best, e0.elapsed_time(e1) / (reps * len(fns)))
    return best


def time_fused_kernels(ctx, torch, U, Un, Q, rhoi, rhon, dt, reps=10):
    """Mean launch duration of each kernel of the fused float stage on the
    state a run left behind: {name: (ms, launches per stage)}.  The diffusive
    sweep of a direction is its viscous and species kernels together; the
    stage end runs on copies of the state (KRK = 2, with linComb)."""
    w = ctx.workspace
    rhs = ctx.rhs_zeros()
    out = {
        "mixture": (_best_of(torch, [lambda: ctx.mixture(Q, rhoi, w("Yi"), w("lambda"), w("mu"),
                                                          w("D"))], reps), 1),
        "shock_sensor": (_best_of(torch, [lambda: ctx.shock_sensor(w("phi"), Q)], reps), 1),
        "sweep_conv": (_best_of(torch, [
            (lambda d=d: ctx.sweep_conv(d, rhs, Q, rhoi, w("phi"), d > 0)) for d in range(3)],
            reps), 3),
        "sweep_diff": (_best_of(torch, [
            (lambda d=d: ctx.sweep_diff(d, rhs, Q, w("Yi"), w("lambda"), w("mu"), w("D")))
            for d in range(3)], reps), 3),
    }
    U2, Q2, r2 = U.clone(), Q.clone(), rhoi.clone()
    out["stage_end"] = (_best_of(torch, [lambda: ctx.sweep_stage_end(2, U2, Un, r2, rhon, Q2, rhs,
                                                                     dt)], reps), 1)
    del U2, Q2, r2, rhs
    return out


def calibrate_copy(torch, nbytes=1 << 30, reps=5):
    """A plain device copy of known size, for scaling the FETCH_SIZE /
    WRITE_SIZE counters (rocprofv3 --pmc) to bytes."""
    src = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "ms": ms,
                      "GBps": 2 * nbytes / (ms * 1e-3) / 1e9}))


def strang_step(ctx, st, dt, chem):
    """One time step as solver.jl:240-378: chemistry over dt/2 and ghost
    fill, the three RK3 stages (each ends with c2Prim and the fills), then
    chemistry over dt/2 and the fill."""
    U, Un, rhoi, rhon, Q = st
    chem(dt / 2)
    ctx.fill(Q, U, rhoi)
    for k in (1, 2, 3):
        ctx.rk3_stage(k, U, Un, rhoi, rhon, Q, dt)
    chem(dt / 2)
    ctx.fill(Q, U, rhoi)


def run_path(prec, case, mech, steps, warmup, torch, abi, kernel_reps, surrogate=None,
             chemistry=True, scheme=None):
    """Time `steps` full time steps of the path.  prec fp64: FP64 stage +
    FP64 backward-Euler Newton chemistry (the plain method); fp32 with a
    surrogate: FP32 flux arithmetic + the BF16 MLP chemistry step with Newton
    fallback.  Each chemistry call is also timed on its own (HIP events)."""
    N = case.grid.Nx, case.grid.Ny, case.grid.Nz
    ctx = abi.Context(*N, mech, prec, scheme=scheme)
    Q = case.Q.clone()
    rhoi = case.rhoi.clone()
    U = ctx.zeros(5)
    ctx.prim2c(U, Q)
    Un, rhon = ctx.zeros(5), ctx.zeros(mech.Ns)
    ctx.set_metrics(case.grid.metrics, case.grid.J)
    ctx.set_boundary(case.bc)
    st = (U, Un, rhoi, rhon, Q)
    chem_ms, fallback = [], []
    if not chemistry:
        def chem(h):
            pass
    elif surrogate is not None:
        surrogate(ctx)

        def chem(h):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            fallback.append(ctx.nn_chem(U, Q, rhoi, h, 50, 1e-10, 1e-6))
            e1.record()
            chem_ms.append((e0, e1))
    else:
        def chem(h):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            ctx.chem_newton(U, Q, rhoi, h, 50, 1e-10, 1e-6)
            e1.record()
            chem_ms.append((e0, e1))
    for s in range(warmup):
        strang_step(ctx, st, case.dt, chem)
    torch.cuda.synchronize()
    # restart from the initial state so both paths advance identically
    Q.copy_(case.Q)
    rhoi.copy_(case.rhoi)
    ctx.prim2c(U, Q)
    chem_ms.clear()
    fallback.clear()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record()
    for s in range(steps):
        strang_step(ctx, st, case.dt, chem)
    e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = e0.elapsed_time(e1)
    cms = [a.elapsed_time(b) for a, b in chem_ms] or [0.0]
    fused = prec == "fp32" and (scheme or {}).get("stage_kind", abi.STAGE_FUSED) == abi.STAGE_FUSED
    if fused:
        kt = time_fused_kernels(ctx, torch, U, Un, Q, rhoi, rhon, case.dt, kernel_reps)
    else:
        kt = time_kernels(ctx, torch, Q, rhoi, case.grid.metrics, case.grid.J, kernel_reps)
    extra = {}
    if chemistry and surrogate is not None:
        # the surrogate step's parts (network kernel, gate epilogue, scan +
        # compaction, read-back + Newton fallback), event-timed inside the
        # library on the launch stream, over two further calls, each from a
        # copy of the end state
        ctx.nn_timing(True)
        parts, nfb = [], []
        for _ in range(2):
            Uc, Qc, rc = U.clone(), Q.clone(), rhoi.clone()
            nfb.append(ctx.nn_chem(Uc, Qc, rc, case.dt / 2, 50, 1e-10, 1e-6))
            parts.append(ctx.nn_timing())
        ctx.nn_timing(False)
        extra["nn_parts_ms"] = [sum(p[k] for p in parts) / len(parts) for k in range(4)]
        extra["nn_parts_fallback"] = sum(nfb) / len(nfb)
        # the fallback's Newton iterations: the Newton step over every cell
        # from the same state counts them per cell (the fallback runs the same
        # kernel on the rejected cells, so their counts are the same)
        N3 = (N[2], N[1], N[0])
        st = torch.zeros(N3, dtype=torch.int8, device="cuda")
        iters = torch.zeros(N3, dtype=torch.int32, device="cuda")
        Uc, Qc, rc = U.clone(), Q.clone(), rhoi.clone()
        ctx.nn_chem(Uc, Qc, rc, case.dt / 2, 50, 1e-10, 1e-6, status=st)
        Uc, Qc, rc = U.clone(), Q.clone(), rhoi.clone()
        ctx.chem_newton(Uc, Qc, rc, case.dt / 2, 50, 1e-10, 1e-6, iters=iters)
        torch.cuda.synchronize()
        rej = st >= 2
        extra["fallback_iterations"] = int(iters.abs()[rej].to(torch.int64).sum())
        extra["fallback_cells"] = int(rej.sum())
        extra["nn_status_counts"] = {str(c): int((st == c).sum()) for c in range(9)}
        del Uc, Qc, rc, st, iters
    elif chemistry:
        # one more Newton call from copies of the end state with the
        # iteration counts returned (the FP64 work of the launch)
        Uc, Qc, rc = U.clone(), Q.clone(), rhoi.clone()
        iters = torch.zeros((N[2], N[1], N[0]), dtype=torch.int32, device="cuda")
        torch.cuda.synchronize()
        a0, a1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a0.record()
        ctx.chem_newton(Uc, Qc, rc, case.dt / 2, 50, 1e-10, 1e-6, iters=iters)
        a1.record()
        torch.cuda.synchronize()
        it = iters.abs().to(torch.int64)
        extra["newton_ms"] = a0.elapsed_time(a1)
        extra["newton_iterations"] = int(it.sum())
        extra["newton_cells"] = int((it > 0).sum())
        del Uc, Qc, rc, iters, it
    return ctx, dict(ms_total=ms, wall_s=wall, Q=Q, rhoi=rhoi, U=U, kernels=kt,
                     chem_ms_mean=sum(cms) / len(cms), chem_calls=len(cms),
                     fallback=fallback, **extra)


def field_errors(torch, a, b, names):
    out = {}
    sl = (slice(None),) + (slice(4, -4),) * 3
    A, B = a[sl], b[sl]
    for n, name in enumerate(names):
        d = (A[n] - B[n]).double()
        ref = B[n].double()
        scale = ref.abs().max().item() or 1.0
        out[name] = {"l2": math.sqrt(float((d * d).mean())) / scale,
                     "linf": float(d.abs().max()) / scale,
                     "linf_abs": float(d.abs().max())}
    return out


def state_errors(ra, rb, mech):
    """Mass-fraction and temperature differences of two runs' interiors."""
    sl = (slice(None),) + (slice(4, -4),) * 3
    Ya = ra["rhoi"][sl] / ra["Q"][0][sl[1:]].unsqueeze(0)
    Yb = rb["rhoi"][sl] / rb["Q"][0][sl[1:]].unsqueeze(0)
    dY = (Ya - Yb).abs()
    dT = (ra["Q"][5][sl[1:]] - rb["Q"][5][sl[1:]]).abs()
    return {"Y_linf": float(dY.max()), "Y_l2": float(dY.pow(2).mean().sqrt()),
            "T_linf_K": float(dT.max()), "T_l2_K": float(dT.pow(2).mean().sqrt()),
            "per_species_Y_linf": {sp: float(dY[k].max()) for k, sp in enumerate(mech.species)}}


ACC_BAR = {"T_K": 1.0, "Y": 1e-3}  # the accuracy bar of the FP32 + surrogate path


def branch_codes(torch, phi, phi1, phi2, NG=4):
    """Hybrid-switch branch (0 upwind, 1 WENO, 2 NND; schemes.jl:96-100) of
    every interior face of the x, y, z sweeps from one stage's sensor phi
    ((Nz+8, Ny+8, Nx+8)), face i-1/2 at max(phi[i-1], phi[i]), compared in
    phi's own precision as the kernels compare."""
    Nzt, Nyt, Nxt = phi.shape
    Nx, Ny, Nz = Nxt - 2 * NG, Nyt - 2 * NG, Nzt - 2 * NG
    p = phi
    a1 = torch.tensor(phi1, dtype=p.dtype, device=p.device)
    a2 = torch.tensor(phi2, dtype=p.dtype, device=p.device)
    fx = torch.maximum(p[NG:NG + Nz, NG:NG + Ny, NG - 1:NG + Nx], p[NG:NG + Nz, NG:NG + Ny, NG:NG + Nx + 1])
    fy = torch.maximum(p[NG:NG + Nz, NG - 1:NG + Ny, NG:NG + Nx], p[NG:NG + Nz, NG:NG + Ny + 1, NG:NG + Nx])
    fz = torch.maximum(p[NG - 1:NG + Nz, NG:NG + Ny, NG:NG + Nx], p[NG:NG + Nz + 1, NG:NG + Ny, NG:NG + Nx])
    return [(f >= a1).to(torch.int8) + (f >= a2).to(torch.int8) for f in (fx, fy, fz)]


def mark_flips(torch, m, ca, cb):
    """OR into the cell mask m (Nz, Ny, Nx) the two cells of every face whose
    branch differs between codes ca and cb; returns the number of such faces."""
    n = 0
    for d, (x, y) in enumerate(zip(ca, cb)):
        f = x != y
        n += int(f.sum())
        if d == 0:
            m |= f[:, :, :-1] | f[:, :, 1:]
        elif d == 1:
            m |= f[:, :-1, :] | f[:, 1:, :]
        else:
            m |= f[:-1] | f[1:]
    return n


def accuracy_decomposition(case, mech, steps, torch, abi, sur, marks=(1, 6), radius=2,
                           scheme=None, attach=None, legs=None):
    """Untimed runs of `steps` time steps from the same initial state, each
    compared after every step with FP64 plain (FP64 stage + FP64 Newton
    chemistry) by the interior's max |dT| and max_k |dY_k|:
      fp32_newton   FP32 stage + Newton chemistry   (the float stage's share)
      fp64_nn       FP64 stage + surrogate          (the surrogate's share)
      fp32_nn       FP32 stage + surrogate          (the benchmarked path)
      fp64_perturbed  FP64 plain from velocities perturbed by 1e-7 relative
                    (float rounding of the initial state): the flow's own
                    sensitivity to a perturbation of that size.
    The hybrid scheme's switch is discontinuous in the sensor, so a
    perturbation of any size flips the branch at the faces whose sensor sits
    at a threshold; every stage's branch codes are compared with the FP64
    plain run's, and T is also reported away from the cells within `radius`
    of a face flipped at any stage so far (T_linf_away_K).
    Returns per-step curves, the values at `marks` and at the last step, and
    the first step at which each path leaves the bar (in Linf, and away from
    the flips).  scheme: the contexts' scheme options; attach: loads the
    surrogate into a context (default: the H2/O2 network); legs: the names
    of the comparisons to run (default: all four)."""
    attach = attach or (lambda c: sur.attach(c))
    N = case.grid.Nx, case.grid.Ny, case.grid.Nz
    sl = (slice(4, -4),) * 3
    keep_all = steps <= 40

    def run(prec, nn, sink, perturb=0.0):
        ctx = abi.Context(*N, mech, prec, scheme=scheme)
        ctx.set_metrics(case.grid.metrics, case.grid.J)
        ctx.set_boundary(case.bc)
        Q, rhoi = case.Q.clone(), case.rhoi.clone()
        if perturb:
            g = torch.Generator(device="cuda").manual_seed(7)
            for n in (1, 2, 3):
                Q[n] *= 1 + perturb * torch.randn(Q[n].shape, generator=g, device="cuda",
                                                  dtype=torch.float64)
        U = ctx.zeros(5)
        ctx.prim2c(U, Q)
        Un, rn = ctx.zeros(5), ctx.zeros(mech.Ns)
        if nn:
            attach(ctx)
        phi1, phi2 = ctx.scheme["hybrid_phi1"], ctx.scheme["hybrid_phi2"]

        def half():
            if nn:
                ctx.nn_chem(U, Q, rhoi, case.dt / 2, 50, 1e-10, 1e-6)
            else:
                ctx.chem_newton(U, Q, rhoi, case.dt / 2, 50, 1e-10, 1e-6)
            ctx.fill(Q, U, rhoi)
        # f3d_time_step's sequence, with every stage's branch codes
        for s in range(1, steps + 1):
            half()
            codes = []
            for k in (1, 2, 3):
                ctx.rk3_stage(k, U, Un, rhoi, rn, Q, case.dt)
                codes.append(branch_codes(torch, ctx.workspace("phi"), phi1, phi2))
            half()
            sink(s, Q, rhoi, codes)
        torch.cuda.synchronize()
        ctx.close()

    def TY(Q, rhoi):
        return Q[5][sl].clone(), (rhoi[(slice(None),) + sl] / Q[0][sl].unsqueeze(0)).float()

    ref, ref_codes = {}, {}

    def keep(s, Q, rhoi, codes):
        ref_codes[s] = codes
        if keep_all or s in marks or s == steps:
            ref[s] = TY(Q, rhoi)
    log(f"[bench] accuracy: FP64 plain reference, {steps} steps")
    run("fp64", False, keep)
    out = {}
    for name, prec, nn, pert in (("fp32_newton", "fp32", False, 0.0), ("fp64_nn", "fp64", True, 0.0),
                                 ("fp32_nn", "fp32", True, 0.0),
                                 ("fp64_perturbed", "fp64", False, 1e-7)):
        if legs is not None and name not in legs:
            continue
        curve = {"step": [], "T_linf_K": [], "Y_linf": [], "cells_T_over_1K": [],
                 "face_flips": [], "cells_near_flips": [], "T_linf_away_K": [],
                 "Y_linf_away": []}
        near = torch.zeros(N[::-1], dtype=torch.bool, device="cuda")
        nflips = [0]

        def cmp(s, Q, rhoi, codes):
            for ca, cb in zip(ref_codes[s], codes):
                nflips[0] += mark_flips(torch, near, ca, cb)
            if s not in ref:
                return
            T, Y = TY(Q, rhoi)
            Tr, Yr = ref[s]
            dT = (T - Tr).abs()
            k = 2 * radius + 1
            m = torch.nn.functional.max_pool3d(near[None, None].float(), k, stride=1,
                                               padding=radius)[0, 0] > 0
            curve["step"].append(s)
            curve["T_linf_K"].append(float(dT.max()))
            curve["Y_linf"].append(float((Y - Yr).abs().max()))
            curve["cells_T_over_1K"].append(int((dT > 1.0).sum()))
            curve["face_flips"].append(nflips[0])
            curve["cells_near_flips"].append(int(m.sum()))
            curve["T_linf_away_K"].append(float(dT[~m].max()))
            curve["Y_linf_away"].append(float((Y - Yr).abs().amax(0)[~m].max()))
        log(f"[bench] accuracy: {name}")
        run(prec, nn, cmp, pert)
        first = next((s for s, t, y in zip(curve["step"], curve["T_linf_K"], curve["Y_linf"])
                      if t > ACC_BAR["T_K"] or y > ACC_BAR["Y"]), None)
        first_away = next((s for s, t, y in zip(curve["step"], curve["T_linf_away_K"],
                                                curve["Y_linf_away"])
                           if t > ACC_BAR["T_K"] or y > ACC_BAR["Y"]), None)
        at = {str(s): {"T_linf_K": curve["T_linf_K"][i], "Y_linf": curve["Y_linf"][i],
                       "cells_T_over_1K": curve["cells_T_over_1K"][i],
                       "T_linf_away_from_flips_K": curve["T_linf_away_K"][i],
                       "Y_linf_away_from_flips": curve["Y_linf_away"][i],
                       "face_flips": curve["face_flips"][i]}
              for i, s in enumerate(curve["step"]) if s in marks or s == steps}
        out[name] = {"at_step": at, "first_step_over_bar": first,
                     "first_step_over_bar_away_from_flips": first_away, "curve": curve}
        del near
    del ref, ref_codes
    torch.cuda.empty_cache()
    return {"bar": ACC_BAR, "steps": steps, "vs": "FP64 plain (FP64 stage + FP64 Newton chemistry)",
            "flip_radius_cells": radius, **out}


def cpu_baseline(mech, size, steps, min_seconds=10.0):
    """The FP64 oracle (OpenMP over the host cores) on a bounded sample of the
    same workload: a size^3 block of the reacting mixing layer, whole time
    steps (Newton chemistry over dt/2, three RK3 stages, Newton over dt/2), at
    least `steps` steps and at least `min_seconds` of CPU work."""
    sys.path.insert(0, ROOT)
    import numpy as np
    import flame3d_amd.cases as cases
    from oracle import Oracle

    case = cases.reacting_mixing_layer((size, size, size), mech, dev="cpu")

    def f(t):
        return np.array(t.numpy().T, order="F", copy=True)

    orc = Oracle(size, size, size, mech)
    Q, r = f(case.Q), f(case.rhoi)
    M, J = [f(m) for m in case.grid.metrics], f(case.grid.J)
    U = orc.prim2c(Q)
    Un, rn, ws = orc.zeros(5), orc.zeros(mech.Ns), orc.workspace()

    def step():
        orc.chem_newton(U, Q, r, case.dt / 2)
        orc.fill(case.bc, Q, U, r)
        for k in (1, 2, 3):
            orc.rk3_stage(k, U, Un, r, rn, Q, ws, M, J, case.dt, case.bc)
        orc.chem_newton(U, Q, r, case.dt / 2)
        orc.fill(case.bc, Q, U, r)

    step()  # warm-up
    t0 = time.perf_counter()
    done = 0
    while done < steps or time.perf_counter() - t0 < min_seconds:
        step()
        done += 1
    dt = time.perf_counter() - t0
    return {"value": 3 * size ** 3 * done / dt, "unit": "cell-updates/s", "cores": orc.num_threads(),
            "kind": "port",
            "sample": f"FP64 oracle (OpenMP, {orc.num_threads()} threads), reacting H2/air mixing "
                      f"layer {size}^3 block, {done} time steps (= {3 * done} RK3 stages with "
                      f"Strang Newton chemistry), {dt:.1f} s"}


REF_CHEM = os.path.join(ROOT, "oracle", "_ref", "chemical_test")


def cpu_baseline_reference(t_end="4e-3"):
    """The reference's own native chemistry, NN/H2/chemical_test.cpp compiled
    from its source (oracle/_ref, by oracle/Makefile): a 0-D H2/O2 reactor at
    1500 K and 10 atm advanced by its point-implicit step (rates, Jacobian,
    pivoted solve, GetT) with dt = 1e-8 s to t_end (4e5 steps, ~6 s on one
    core, its per-step output to /dev/null); reactor-steps per second on one
    core.  None where the binary is absent."""
    import subprocess
    if not os.path.exists(REF_CHEM):
        return None
    import tempfile
    with tempfile.TemporaryDirectory() as td:
        t0 = time.perf_counter()
        subprocess.run([REF_CHEM, "1e-8", t_end, "1500", "10"], cwd=td, check=True,
                       stdout=subprocess.DEVNULL)
        dt = time.perf_counter() - t0
    steps = round(float(t_end) / 1e-8)
    return {"value": steps / dt, "unit": "reactor-steps/s", "cores": 1, "kind": "reference",
            "sample": f"NN/H2/chemical_test.cpp (compiled from the reference's source), one "
                      f"0-D H2/O2 reactor, 1500 K / 10 atm, {steps} point-implicit steps of "
                      f"1e-8 s in {dt:.1f} s, per-step output discarded"}


def stage_roofline(kt, Ns):
    """HBM roofline of the stage kernel that takes the most time per stage
    (ms x launches) in the per-kernel table `kt` of one precision (traffic:
    the counters are measured on the reacting mixing layer only, so None)."""
    share = {k: v["ms"] * v["launches_per_stage"] for k, v in kt.items()}
    dom = max(share, key=share.get)
    dk = kt[dom]
    return {"bound": "hbm", "kernel": dom, "achieved": dk["GBps"], "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": dk["GBps"] / HBM_PEAK_GBS,
            "traffic": None, "alg_bytes_per_launch": dk["alg_bytes"],
            "avg_launch_ms": dk["ms"], "ms_per_stage": share[dom]}


def bench_air(a, torch, abi, cases, mm):
    """The shock-laden mixing layer of 5-species air, non-reacting, hybrid
    WENO5-Z / UP7 switched by the Ducros sensor at 0.1: one RK3 stage (the
    step) on the FP32 path against the FP64 path."""
    mech = mm.load("air5")
    N = (a.size, a.size, a.size)
    case = cases.mixing_layer(N, mech, dev="cuda")
    nc = N[0] * N[1] * N[2]
    stages = 3 * a.steps
    c64, r64 = run_path("fp64", case, mech, a.steps, a.warmup, torch, abi, a.kernel_reps,
                        chemistry=False, scheme=abi.DUCROS_SCHEME)
    c32, r32 = run_path("fp32", case, mech, a.steps, a.warmup, torch, abi, a.kernel_reps,
                        chemistry=False, scheme=abi.DUCROS_SCHEME)
    ms64, ms32 = r64["ms_total"] / stages, r32["ms_total"] / stages
    errs = field_errors(torch, r32["Q"], r64["Q"], ["rho", "u", "v", "w", "p", "T", "e"])
    npts = (N[0] + 8) * (N[1] + 8) * (N[2] + 8)
    nf = (N[0] + 1) * N[1] * N[2]
    ktab = {}
    for prec, res in (("fp32", r32), ("fp64", r64)):
        R = 4 if prec == "fp32" else 8
        for name, (ms, per) in res["kernels"].items():
            b = kernel_bytes(name, R, mech.Ns, npts, nf, nc)
            ktab.setdefault(prec, {})[name] = {
                "ms": ms, "launches_per_stage": per, "alg_bytes": b,
                "GBps": b / (ms * 1e-3) / 1e9, "frac_hbm": b / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
    print(json.dumps({
        "metric": "cell-updates/s per RK stage, FP32 flux vs FP64 (non-reacting air)",
        "value": nc / (ms32 * 1e-3), "unit": "cell-updates/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": 3 * ms32, "ms_per_stage": ms32,
        "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 flux, f64 state", "data": "synthetic",
        "config": {"workload": f"shock-laden mixing layer {N[0]}^3, 5-species air, non-reacting, "
                               "WENO5-Z/UP7 switched by the Ducros sensor at 0.1"},
        "roofline": stage_roofline(ktab["fp32"], mech.Ns),
        "fp64": {"value": nc / (ms64 * 1e-3), "ms_per_stage": ms64},
        "speedup_fp32_vs_fp64": ms64 / ms32, "error_fp32_vs_fp64": errs, "kernels": ktab}))
    c32.close()
    c64.close()


def bench_jet(a, torch, abi, cases, mm):
    """The supersonic hot-air jet (5-species air, 5000 K, Ma 1.5 into 300 K
    air) on a 2N x N x N stretched block, WENO7-Z in the middle branch of the
    hybrid switch, at the fixed dt of 1.2e-8 s the air surrogate was trained
    for (chemistry half-steps of 6e-9 s): FP64 stage + FP64 Newton air
    chemistry (the plain method), FP32 stage + the same Newton chemistry, and
    FP32 stage + the BF16 MLP air chemistry with Newton fallback."""
    import flame3d_amd.surrogate as sur
    mech = mm.load("air5")
    N = (2 * a.size, a.size, a.size)
    log(f"[bench] building the reacting jet case {N} on the GPU")
    case = cases.reacting_jet(N, mech, dev="cuda", dt=2 * sur.AIR_DT)
    nc = N[0] * N[1] * N[2]
    stages = 3 * a.steps
    scheme = {"recon_kind": abi.RECON_WENO7}
    c64, r64 = run_path("fp64", case, mech, a.steps, a.warmup, torch, abi, a.kernel_reps,
                        scheme=scheme)
    c64.close()
    ms64 = r64["ms_total"] / stages
    chem64 = r64["chem_ms_mean"]
    log(f"[bench] FP64 + Newton {ms64:.3f} ms/stage (chemistry {chem64:.2f} ms/call)")
    c32, r32 = run_path("fp32", case, mech, a.steps, a.warmup, torch, abi, a.kernel_reps,
                        scheme=scheme)
    c32.close()
    ms32 = r32["ms_total"] / stages
    chem32 = r32["chem_ms_mean"]
    log(f"[bench] FP32 + Newton {ms32:.3f} ms/stage (chemistry {chem32:.2f} ms/call)")
    cnn, rnn = run_path("fp32", case, mech, a.steps, a.warmup, torch, abi, a.kernel_reps,
                        scheme=scheme, surrogate=lambda c: sur.attach(c, sur.AIR_PATH))
    cnn.close()
    msnn = rnn["ms_total"] / stages
    nfb = rnn["fallback"]
    fb = sum(nfb) / (nc * len(nfb))
    log(f"[bench] FP32 + NN {msnn:.3f} ms/stage (chemistry {rnn['chem_ms_mean']:.2f} ms/call, "
        f"fallback {fb:.4f})")
    names = ["rho", "u", "v", "w", "p", "T", "e"]
    npts = (N[0] + 8) * (N[1] + 8) * (N[2] + 8)
    nf = (N[0] + 1) * N[1] * N[2]
    err_nn = state_errors(rnn, r64, mech)
    err_32 = state_errors(r32, r64, mech)
    fields = field_errors(torch, rnn["Q"], r64["Q"], names)
    del r64, r32
    accuracy = None
    if not a.no_accuracy:
        # the same steps untimed, every stage's hybrid-switch branches compared
        accuracy = accuracy_decomposition(
            case, mech, a.steps, torch, abi, sur, marks=(1, 10, 25), scheme=scheme,
            attach=lambda c: sur.attach(c, sur.AIR_PATH),
            legs=("fp32_newton", "fp32_nn", "fp64_perturbed"))
    ktab = {}
    for name, (ms, per) in rnn["kernels"].items():
        b = kernel_bytes(name, 4, mech.Ns, npts, nf, nc)
        ktab[name] = {"ms": ms, "launches_per_stage": per, "alg_bytes": b,
                      "GBps": b / (ms * 1e-3) / 1e9, "frac_hbm": b / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
    print(json.dumps({
        "metric": "cell-updates/s per RK stage, FP32+NN vs FP64 plain (reacting air jet)",
        "value": nc / (msnn * 1e-3), "unit": "cell-updates/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": 3 * msnn, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 flux + bf16 MLP + f64 chemistry/state",
        "data": "synthetic",
        "config": {"workload": f"hot air jet {N[0]}x{N[1]}x{N[2]}, 5-species air, 5000 K Ma 1.5 "
                               "into 300 K, WENO7-Z/UP7/NND, Strang chemistry", "dt_s": case.dt},
        "fp64_newton": {"value": nc / (ms64 * 1e-3), "ms_per_stage": ms64,
                        "chem_ms_per_call": chem64},
        "fp32_newton": {"value": nc / (ms32 * 1e-3), "ms_per_stage": ms32,
                        "chem_ms_per_call": chem32, "error_vs_fp64": err_32},
        "fp32_nn": {"ms_per_stage": msnn, "chem_ms_per_call": rnn["chem_ms_mean"],
                    "fallback_fraction": fb},
        "speedup_vs_fp64_plain": ms64 / msnn, "speedup_fp32_newton_vs_fp64": ms64 / ms32,
        "error_vs_fp64_plain": err_nn, "error_fields_vs_fp64_plain": fields,
        "accuracy": accuracy,
        "roofline": stage_roofline(ktab, mech.Ns), "kernels": {"fp32": ktab}}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20,
                    help="timed time steps (each = 3 RK3 stages + 2 chemistry half-steps)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--cpu-size", type=int, default=48)
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true",
                    help="skip the per-step accuracy decomposition (untimed runs)")
    ap.add_argument("--kernel-reps", type=int, default=10)
    ap.add_argument("--calibrate-copy", action="store_true")
    ap.add_argument("--workload", choices=("reacting", "air", "jet"), default="reacting",
                    help="air: the non-reacting 5-species air mixing layer, FP32 vs FP64 stage; "
                    "jet: the hot air jet at 2*size x size x size with Newton air chemistry")
    ap.add_argument("--pmc-summary", metavar="DIR",
                    help="write the per-kernel HBM traffic summary from rocprofv3 --pmc runs")
    ap.add_argument("--pmc-out", metavar="FILE", help="where --pmc-summary writes (default: "
                    "profiles/hbm_traffic_reacting_mixing_layer_256.json)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: this build measures one GPU's block only; --gpus {a.gpus} is not "
              "supported (the block decomposition over several GPUs is a design, not built)",
              file=sys.stderr)
        sys.exit(2)

    if a.pmc_summary:
        res = pmc_summary(a.pmc_summary)
        out = a.pmc_out or TRAFFIC_FILE
        os.makedirs(os.path.dirname(os.path.abspath(out)), exist_ok=True)
        with open(out, "w") as fh:
            json.dump(res, fh, indent=1)
        print(json.dumps(res, indent=1))
        return
    _package()
    import torch
    if a.calibrate_copy:
        calibrate_copy(torch)
        return
    import flame3d_amd.abi as abi
    import flame3d_amd.cases as cases
    import flame3d_amd.mech as mm
    import flame3d_amd.surrogate as sur

    if a.workload == "air":
        return bench_air(a, torch, abi, cases, mm)
    if a.workload == "jet":
        return bench_jet(a, torch, abi, cases, mm)
    mech = mm.load("h2_lidryer")
    N = (a.size, a.size, a.size)
    log(f"[bench] building the reacting mixing-layer case {N} on the GPU")
    case = cases.reacting_mixing_layer(N, mech, dev="cuda")
    nc = N[0] * N[1] * N[2]
    stages = 3 * a.steps
    log(f"[bench] dt = {case.dt:.3e} s; FP64 plain path (FP64 stage + Newton chemistry)")
    c64, r64 = run_path("fp64", case, mech, a.steps, a.warmup, torch, abi, a.kernel_reps)
    ms64 = r64["ms_total"] / stages
    log(f"[bench] FP64 plain {ms64:.3f} ms/stage (chemistry {r64['chem_ms_mean']:.2f} ms/call); "
        "FP32 + surrogate path")
    c32, r32 = run_path("fp32", case, mech, a.steps, a.warmup, torch, abi, a.kernel_reps,
                        surrogate=sur.attach)
    ms32 = r32["ms_total"] / stages
    log(f"[bench] FP32+NN {ms32:.3f} ms/stage (chemistry {r32['chem_ms_mean']:.2f} ms/call, "
        f"fallback {sum(r32['fallback']) / (nc * len(r32['fallback'])):.3f})")
    errs = field_errors(torch, r32["Q"], r64["Q"], ["rho", "u", "v", "w", "p", "T", "e"])
    acc = state_errors(r32, r64, mech)
    value = nc / (ms32 * 1e-3)
    c32.close()
    c64.close()

    # ---- per-kernel rooflines ------------------------------------------
    Ns = mech.Ns
    npts = (N[0] + 8) * (N[1] + 8) * (N[2] + 8)
    nf = (N[0] + 1) * N[1] * N[2]  # faces per sweep (equal for x, y, z on a cube)
    ktab = {}
    for prec, res in (("fp32", r32), ("fp64", r64)):
        R = 4 if prec == "fp32" else 8
        for name, (ms, per) in res["kernels"].items():
            b = kernel_bytes(name, R, Ns, npts, nf, nc)
            ktab.setdefault(prec, {})[name] = {
                "ms": ms, "launches_per_stage": per, "alg_bytes": b,
                "GBps": b / (ms * 1e-3) / 1e9, "frac_hbm": b / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
    # chemistry: FP64 work from the oracle's operation counts
    # (profiles/chem_opcount.json) over the event-timed launches
    ops = json.load(open(OPCOUNT_FILE))["mechanisms"]["h2_lidryer"]
    hot = int((case.Q[5][(slice(4, -4),) * 3] > 500.0).sum())
    net_ms, epi_ms, nn_scan_ms, nn_fb_ms = r32["nn_parts_ms"]
    nn_ms = net_ms + epi_ms
    nn_f64 = hot * ops["surrogate_epilogue"]["flop"]
    nn_bf16 = hot * MLP_FLOP_PER_CELL
    chem_tab = {
        "k_nn_net": {"ms": net_ms, "launches_per_step": 2, "cells": hot,
                     "bf16_mfma_flop": nn_bf16,
                     "bf16_mfma_frac": nn_bf16 / (net_ms * 1e-3) / 1e12 / BF16_PEAK_TFLOPS},
        "k_nn_pred+k_nn_gate": {"ms": epi_ms, "launches_per_step": 2, "cells": hot,
                                "fp64_flop": nn_f64,
                                "fp64_frac": nn_f64 / (epi_ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS},
        "k_nn_chem": {"ms": nn_ms, "launches_per_step": 2, "cells": hot,
                      "fp64_flop": nn_f64, "fp64_TFLOPs": nn_f64 / (nn_ms * 1e-3) / 1e12,
                      "bf16_mfma_flop": nn_bf16,
                      "bf16_mfma_TFLOPs": nn_bf16 / (nn_ms * 1e-3) / 1e12,
                      "bf16_mfma_frac": nn_bf16 / (nn_ms * 1e-3) / 1e12 / BF16_PEAK_TFLOPS},
        "nn_scan_compact": {"ms": nn_scan_ms, "launches_per_step": 2},
        "newton_fallback": {"ms": nn_fb_ms, "launches_per_step": 2,
                            "cells": r32["fallback_cells"],
                            "iterations": r32["fallback_iterations"],
                            "fp64_flop": r32["fallback_iterations"] * ops["newton_iteration"]["flop"]
                            + r32["fallback_cells"] * ops["newton_cell_overhead"]["flop"],
                            "gate_codes": r32["nn_status_counts"]},
        "k_chem_newton_fp64_plain": {
            "ms": r64["newton_ms"], "launches_per_step": 2, "cells": r64["newton_cells"],
            "iterations": r64["newton_iterations"],
            "fp64_flop": r64["newton_iterations"] * ops["newton_iteration"]["flop"]
            + r64["newton_cells"] * ops["newton_cell_overhead"]["flop"]},
    }
    kf = chem_tab["newton_fallback"]
    kf["fp64_TFLOPs"] = kf["fp64_flop"] / (kf["ms"] * 1e-3) / 1e12
    kf["fp64_frac"] = kf["fp64_TFLOPs"] / FP64_PEAK_TFLOPS
    kn = chem_tab["k_chem_newton_fp64_plain"]
    kn["fp64_TFLOPs"] = kn["fp64_flop"] / (kn["ms"] * 1e-3) / 1e12
    kn["fp64_frac"] = kn["fp64_TFLOPs"] / FP64_PEAK_TFLOPS
    chem_tab["k_nn_chem"]["fp64_frac"] = chem_tab["k_nn_chem"]["fp64_TFLOPs"] / FP64_PEAK_TFLOPS
    # the dominant kernel of the FP32 + surrogate step: largest time per step
    share = {k: v["ms"] * v["launches_per_stage"] * 3 for k, v in ktab["fp32"].items()}
    share["k_nn_chem"] = nn_ms * 2
    share["newton_fallback"] = nn_fb_ms * 2
    dom = max(share, key=share.get)
    if dom == "k_nn_chem":
        kc = chem_tab["k_nn_chem"]
        roofline = {"bound": "mfma", "kernel": "k_nn_net + k_nn_pred + k_nn_gate (FP64 epilogue bound)",
                    "achieved": kc["fp64_TFLOPs"], "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                    "frac": kc["fp64_frac"], "traffic": None,
                    "alg_fp64_flop_per_launch": nn_f64, "avg_launch_ms": nn_ms,
                    "bf16_mfma_frac": kc["bf16_mfma_frac"]}
    elif dom == "newton_fallback":
        roofline = {"bound": "mfma", "kernel": "k_chem<9, 2, KinCT<MechH2LiDryer>> (backward-Euler "
                    "Newton on the surrogate's rejected cells; FP64 vector, against the FP64 "
                    "peak; FP64 flop from the oracle's op count)", "achieved": kf["fp64_TFLOPs"],
                    "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": kf["fp64_frac"],
                    "traffic": None, "alg_fp64_flop_per_launch": kf["fp64_flop"],
                    "avg_launch_ms": nn_fb_ms}
    else:
        dk = ktab["fp32"][dom]
        roofline = {"bound": "hbm", "kernel": dom, "achieved": dk["GBps"], "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": dk["GBps"] / HBM_PEAK_GBS,
                    "traffic": traffic_for(kernel_symbols(dom, Ns)),
                    "alg_bytes_per_launch": dk["alg_bytes"], "avg_launch_ms": dk["ms"]}
    roofline["ms_per_step"] = share[dom]
    roofline["timing"] = ("HIP events on the launch stream; stage kernels: fastest of five "
                          "batches of launches on the run's end state; chemistry parts: mean of "
                          "two calls on copies of it")

    accuracy = None
    if not a.no_accuracy:
        accuracy = accuracy_decomposition(case, mech, a.steps, torch, abi, sur)

    cpu = cpu_ref = None
    if not a.no_cpu:
        log(f"[bench] CPU baseline: oracle {a.cpu_size}^3 x >= {a.cpu_steps} steps")
        cpu = cpu_baseline(mech, a.cpu_size, a.cpu_steps)
        cpu_ref = cpu_baseline_reference()

    nfb = r32["fallback"]
    out = {
        "metric": "cell-updates/s per RK stage (FP32+NN vs FP64 plain) at stated error; % HBM peak",
        "value": value, "unit": "cell-updates/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": ms32 * 3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 flux + bf16 MLP + f64 chemistry/state",
        "data": "synthetic",
        "config": {"workload": f"reacting H2/air shock-laden mixing layer {N[0]}x{N[1]}x{N[2]}, "
                               "tanh-stretched y (2.5), 9-species H2/O2 (Li-Dryer form), hybrid "
                               "UP7/WENO5-Z/NND, viscous + diffusion, Strang chemistry",
                   "cells": nc, "step": "one time step = chemistry dt/2 + 3 RK3 stages + "
                   "chemistry dt/2; value = cells x 3 / step time",
                   "dt_s": case.dt, "precision": "FP32 flux (fused sweeps), FP64 state, BF16 "
                   "MLP chemistry with FP64 Newton fallback"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "cpu_baseline_reference": cpu_ref,
        "fp64_plain": {"value": nc / (ms64 * 1e-3), "ms_per_stage": ms64,
                       "chem_ms_per_call": r64["chem_ms_mean"],
                       "stage": "FP64 kernel-granular (the reference's launch sequence)"},
        "fp32_nn": {"ms_per_stage": ms32, "chem_ms_per_call": r32["chem_ms_mean"],
                    "fallback_fraction": sum(nfb) / (nc * len(nfb))},
        "speedup_vs_fp64_plain": ms64 / ms32,
        "error_vs_fp64_plain": acc,
        "accuracy": accuracy,
        "error_fields_vs_fp64_plain": errs,
        "kernels": ktab,
        "chemistry_kernels": chem_tab,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
