"""Benchmark: cell-updates/s per RK stage of the per-stage RHS on one MI355X.

Workload: the shock-laden mixing layer on a 256^3 curvilinear block
(tanh-stretched y, factor 2.5), 5-species air, hybrid UP7 / WENO5-Z / NND
with the pressure sensor, viscous + species diffusion, non-reacting; inputs
generated in HBM (synthetic, seeded).  A step is one RK3 stage
(solver.jl:298-326) over the whole block: mixture, sensor, species and flow
convective sweeps, viscous fluxes, divergence, linComb, c2Prim, fills.

`value` is the FP32-flux path (FP32 flux / viscous / divergence arithmetic,
FP64 state, FP64 e->T Newton); the FP64 path is timed beside it, and the
FP32 fields after the timed stages are compared with the FP64 ones
(normwise relative L2 / Linf per field; the FP64 path itself matches the CPU
oracle to 1e-15 in tests/test_gpu_parity.py).

    python bench.py [--gpus 1] [--steps 30] [--warmup 3] [--size 256]
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
# writes of its inputs/outputs once each (state, metrics and J in double;
# intermediates in the path's precision R bytes).  npts = padded points,
# nf = faces of the sweep, nc = interior cells.
# ---------------------------------------------------------------------------
def kernel_bytes(name, R, Ns, npts, nf, nc):
    D = 8
    table = {
        "flux_split": npts * (9 * D + 10 * R),            # Q(6) + 3 metrics -> Fp, Fm (5 each)
        "species_split": npts * ((6 + Ns) * D + 2 * Ns * R),  # u,v,w + 3 metrics + rho_k -> 2Ns
        "shock_sensor": npts * (D + R),                    # p -> phi
        "weno5": nf * 16 * R,                              # Fp, Fm (10) + phi -> F (5)
        "weno_species": nf * (2 * Ns + 1 + Ns) * R,
        "viscous": nf * (14 * D + 9 * R),                  # u,v,w,T + 9 metrics + J; lam, mu, Fh(3) -> Fv(4)
        "spec_viscous": nf * (12 * D + (3 * Ns + 3) * R),  # rho, T + 9 metrics + J; Yi, D -> Fd, Fh
        "mixture": npts * ((3 + Ns) * D + (2 * Ns + 2) * R),
        "div": nc * (27 * R + 11 * D),                     # 3x(5+4) faces, J, U RMW
        "div_specs": nc * (6 * Ns * R + (1 + 2 * Ns) * D),
        "c2prim": nc * ((10 + 2 * Ns + 2) * D),            # U, rho_k -> Q(7), rho_k
    }
    return table[name]


# ---------------------------------------------------------------------------
# HBM traffic from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs
# ---------------------------------------------------------------------------
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_mixing_layer_256.json")


def _pmc_avg(path):
    import csv
    import re
    agg = {}
    for r in csv.DictReader(open(path)):
        short = re.sub(r"\(.*", "", r["Kernel_Name"]).strip()
        agg.setdefault(short, []).append(float(r["Counter_Value"]))
    return {k: sum(v) / len(v) for k, v in agg.items()}


def pmc_summary(d):
    """Per-kernel HBM bytes per dispatch from the counter CSVs of four runs
    under `d` (pmc_fetch, pmc_write: the benchmark; copy_fetch, copy_write:
    --calibrate-copy, a 1 GiB device copy).  FETCH_SIZE / WRITE_SIZE are in
    KiB per dispatch; each is scaled by (known copy bytes / counter reading
    of the copy kernel)."""
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
    return out


def traffic_for(kernel_symbol):
    if not os.path.exists(TRAFFIC_FILE):
        return None
    t = json.load(open(TRAFFIC_FILE))["kernels"].get(kernel_symbol)
    return None if t is None else t["fetch_bytes"] + t["write_bytes"]


# kernel symbol (as rocprofv3 prints it) of each timed kernel on the FP32 path
KERNEL_SYMBOL = {
    "flux_split": "void f3d::k_flux_split<float>", "species_split": "void f3d::k_species_split<float>",
    "weno5": "void f3d::k_weno<float>", "weno_species": "void f3d::k_weno<float>",
    "viscous": "void f3d::k_viscous<float>", "spec_viscous": "void f3d::k_spec_viscous<5, float>",
    "shock_sensor": "void f3d::k_shock_sensor<float>", "mixture": "void f3d::k_mixture<5, float>",
}


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
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            for fn in fns:
                fn()
        e1.record()
        torch.cuda.synchronize()
        out[name] = (e0.elapsed_time(e1) / (reps * len(fns)), len(fns))
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


def run_path(prec, case, mech, steps, warmup, torch, abi, kernel_reps):
    N = case.grid.Nx, case.grid.Ny, case.grid.Nz
    ctx = abi.Context(*N, mech, prec)
    Q = case.Q.clone()
    rhoi = case.rhoi.clone()
    U = ctx.zeros(5)
    ctx.prim2c(U, Q)
    Un, rhon = ctx.zeros(5), ctx.zeros(mech.Ns)
    ctx.set_metrics(case.grid.metrics, case.grid.J)
    ctx.set_boundary(case.bc)
    KRK = [1, 2, 3]
    for s in range(warmup):
        ctx.rk3_stage(KRK[s % 3], U, Un, rhoi, rhon, Q, case.dt)
    torch.cuda.synchronize()
    # restart from the initial state so both precisions advance identically
    Q.copy_(case.Q)
    rhoi.copy_(case.rhoi)
    ctx.prim2c(U, Q)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record()
    for s in range(steps):
        ctx.rk3_stage(KRK[s % 3], U, Un, rhoi, rhon, Q, case.dt)
    e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = e0.elapsed_time(e1)
    kt = time_kernels(ctx, torch, Q, rhoi, case.grid.metrics, case.grid.J, kernel_reps)
    return ctx, dict(ms_total=ms, wall_s=wall, Q=Q, rhoi=rhoi, U=U, kernels=kt)


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


def cpu_baseline(mech, size, stages, min_seconds=10.0):
    """The FP64 oracle (OpenMP over the host cores) on a bounded sample of the
    same workload: a size^3 mixing-layer block, at least `stages` RK3 stages
    and at least `min_seconds` of CPU work."""
    sys.path.insert(0, ROOT)
    import numpy as np
    import flame3d_amd.cases as cases
    from oracle import Oracle

    case = cases.mixing_layer((size, size, size), mech, dev="cpu")

    def f(t):
        return np.array(t.numpy().T, order="F", copy=True)

    orc = Oracle(size, size, size, mech)
    Q, r = f(case.Q), f(case.rhoi)
    M, J = [f(m) for m in case.grid.metrics], f(case.grid.J)
    U = orc.prim2c(Q)
    Un, rn, ws = orc.zeros(5), orc.zeros(mech.Ns), orc.workspace()
    orc.rk3_stage(1, U, Un, r, rn, Q, ws, M, J, case.dt, case.bc)  # warm-up
    t0 = time.perf_counter()
    done = 0
    while done < stages or time.perf_counter() - t0 < min_seconds:
        orc.rk3_stage(done % 3 + 1, U, Un, r, rn, Q, ws, M, J, case.dt, case.bc)
        done += 1
    dt = time.perf_counter() - t0
    stages = done
    return {"value": size ** 3 * stages / dt, "unit": "cell-updates/s", "cores": orc.num_threads(),
            "kind": "port",
            "sample": f"FP64 oracle (OpenMP, {orc.num_threads()} threads), mixing layer "
                      f"{size}^3 block, {stages} RK3 stages, {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--cpu-size", type=int, default=96)
    ap.add_argument("--cpu-stages", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--kernel-reps", type=int, default=10)
    ap.add_argument("--calibrate-copy", action="store_true")
    ap.add_argument("--pmc-summary", metavar="DIR",
                    help="write the per-kernel HBM traffic summary from rocprofv3 --pmc runs")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: this build measures one GPU's block only; --gpus {a.gpus} is not "
              "supported (the multi-GPU decomposition is design-only, see DESIGN.md)",
              file=sys.stderr)
        sys.exit(2)

    if a.pmc_summary:
        res = pmc_summary(a.pmc_summary)
        os.makedirs(os.path.dirname(TRAFFIC_FILE), exist_ok=True)
        with open(TRAFFIC_FILE, "w") as fh:
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

    mech = mm.load("air5")
    N = (a.size, a.size, a.size)
    log(f"[bench] building mixing-layer case {N} on the GPU")
    case = cases.mixing_layer(N, mech, dev="cuda")
    nc = N[0] * N[1] * N[2]
    log(f"[bench] dt = {case.dt:.3e} s; timing FP64 path")
    c64, r64 = run_path("fp64", case, mech, a.steps, a.warmup, torch, abi, a.kernel_reps)
    log(f"[bench] FP64 {r64['ms_total'] / a.steps:.3f} ms/stage; timing FP32 path")
    c32, r32 = run_path("fp32", case, mech, a.steps, a.warmup, torch, abi, a.kernel_reps)
    log(f"[bench] FP32 {r32['ms_total'] / a.steps:.3f} ms/stage")
    errs = field_errors(torch, r32["Q"], r64["Q"], ["rho", "u", "v", "w", "p", "T", "e"])
    errs.update(field_errors(torch, r32["rhoi"], r64["rhoi"], [f"rho_{s}" for s in mech.species]))

    ms32 = r32["ms_total"] / a.steps
    ms64 = r64["ms_total"] / a.steps
    value = nc / (ms32 * 1e-3)

    # per-kernel roofline (FP32 path); dominant = largest time per stage
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
    dom = max(ktab["fp32"], key=lambda k: ktab["fp32"][k]["ms"] * ktab["fp32"][k]["launches_per_stage"])
    dk = ktab["fp32"][dom]
    roofline = {"bound": "hbm", "kernel": dom, "achieved": dk["GBps"], "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": dk["GBps"] / HBM_PEAK_GBS,
                "traffic": traffic_for(KERNEL_SYMBOL[dom]),
                "alg_bytes_per_launch": dk["alg_bytes"], "avg_launch_ms": dk["ms"]}

    cpu = None
    if not a.no_cpu:
        log(f"[bench] CPU baseline: oracle {a.cpu_size}^3 x {a.cpu_stages} stages")
        cpu = cpu_baseline(mech, a.cpu_size, a.cpu_stages)

    out = {
        "metric": "cell-updates/s per RK stage (FP32+NN vs FP64 plain) at stated error; % HBM peak",
        "value": value, "unit": "cell-updates/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": ms32, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"shock-laden mixing layer {N[0]}x{N[1]}x{N[2]}, tanh-stretched y "
                               "(2.5), 5-species air non-reacting, hybrid UP7/WENO5-Z/NND, viscous",
                   "cells": nc, "step": "one RK3 stage", "precision": "FP32 flux, FP64 state"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "fp64_plain": {"value": nc / (ms64 * 1e-3), "ms_per_step": ms64},
        "speedup_fp32_vs_fp64": ms64 / ms32,
        "error_fp32_vs_fp64": errs,
        "kernels": ktab,
    }
    print(json.dumps(out))
    c32.close()
    c64.close()


if __name__ == "__main__":
    main()
