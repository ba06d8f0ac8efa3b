"""Train a chemistry surrogate on the GPU and save it.

    python flame3d.jl_amd/train_surrogate.py --out PATH [--mech h2_lidryer|air5]
        [--dt S] [--samples-shape 64 64 32] [--steps 2000] [--records 32]
        [--epochs 1500] [--batch 8192] [--lr 2e-3] [--label-mode 3] [--trace 1e-6]
        [--flow-steps 8]

5-species air (--mech air5): reactors of dissociating air (T ~ U[1500,
6000] K, P ~ U[0.02, 0.1] atm by default) and the hot supersonic jet as
flow data, chemistry step --dt (default 6e-9 s, the half-step of the jet
run at a fixed dt of 1.2e-8 s on its 512 x 256 x 256 block).

Data: trajectories of the library's FP64 backward-Euler Newton integrator
(f3d_chem_newton, dt = 1e-7 s) from H2/O2/N2 reactors with T ~ U[900, 2500] K,
P ~ U[1, 10] atm, mixture fraction ~ Beta(2, 2) plus pure streams
(NN/H2/gen_data.py pattern), the same reactors seeded with species traces,
and the integrator's steps inside the reacting H2/air mixing layer (another
seed and resolution than the benchmark's).
Training: the library's BF16 MFMA forward/backward and Adam kernels
(f3d_nn_train_step).  Then a held-out check on fresh reactors with a
different seed: surrogate step (with gates and Newton fallback) against the
pure Newton step.  Prints one JSON summary line.
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))


def _load_package():
    if "flame3d_amd" in sys.modules:
        return sys.modules["flame3d_amd"]
    spec = importlib.util.spec_from_file_location(
        "flame3d_amd", os.path.join(HERE, "__init__.py"), submodule_search_locations=[HERE])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["flame3d_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


AIR_T, AIR_P = (1500.0, 6000.0), (0.02, 0.1)
H2_T, H2_P = (900.0, 2500.0), (1.0, 10.0)


def evaluate(pkg, mech, path, shape=(64, 64, 64), seed=777, log=print, dt=None,
             T_range=None, P_atm=None, **gates):
    """Surrogate step vs pure Newton step on fresh reactors (held-out seed)."""
    import torch
    abi, cases, sur = pkg.abi, pkg.cases, pkg.surrogate
    air = "H2" not in mech.species
    dt = dt or (sur.AIR_DT if air else sur.DT)
    Nx, Ny, Nz = shape
    ctx = abi.Context(Nx, Ny, Nz, mech, precision="fp64")
    sur.attach(ctx, path, **gates)
    if air:
        Q, rhoi = cases.chem_cells_air(shape, mech, seed=seed, T_range=T_range or AIR_T,
                                       P_atm=P_atm or AIR_P)
        Q, rhoi = Q.cuda().contiguous(), rhoi.cuda().contiguous()
    else:
        Q, rhoi = cases.chem_cells(shape, mech, seed=seed, dev="cuda")
    U = torch.zeros((5,) + tuple(Q.shape[1:]), dtype=torch.float64, device="cuda")
    Qn, rn, Un = Q.clone(), rhoi.clone(), U.clone()
    iters = torch.zeros((Nz, Ny, Nx), dtype=torch.int32, device="cuda")
    ctx.chem_newton(Un, Qn, rn, dt, 50, 1e-10, 1e-6, iters=iters)
    status = torch.zeros((Nz, Ny, Nx), dtype=torch.int8, device="cuda")
    nfb = ctx.nn_chem(U, Q, rhoi, dt, 50, 1e-10, 1e-6, status=status)
    torch.cuda.synchronize()
    Y1, Y0 = sur.mass_fractions(Q, rhoi), sur.mass_fractions(Qn, rn)
    T1, T0 = sur.interior(Q)[:, 5], sur.interior(Qn)[:, 5]
    dY = (Y1 - Y0).abs()
    dT = (T1 - T0).abs()
    acc = (status.reshape(-1) == 1)
    it = iters.reshape(-1).double()
    codes = torch.bincount(status.reshape(-1).long(), minlength=9).tolist()
    res = {
        "cells": Nx * Ny * Nz, "status_counts": codes,
        "fallback_fraction": nfb / (Nx * Ny * Nz),
        "Y_linf": float(dY.max()), "Y_l2": float(torch.sqrt((dY ** 2).mean())),
        "T_linf": float(dT.max()), "T_l2": float(torch.sqrt((dT ** 2).mean())),
        "newton_iters_mean": float(it.mean()),
        "newton_iters_avoided_fraction": float(it[acc].sum() / it.sum()),
    }
    ctx.close()
    return res


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--out", required=True)
    ap.add_argument("--mech", choices=("h2_lidryer", "air5"), default="h2_lidryer")
    ap.add_argument("--dt", type=float, default=None,
                    help="chemistry step the network is trained for (default 1e-7 s for "
                    "H2/O2, 6e-9 s for air)")
    ap.add_argument("--T-range", type=float, nargs=2, default=None)
    ap.add_argument("--P-atm", type=float, nargs=2, default=None)
    ap.add_argument("--samples-shape", type=int, nargs=3, default=(64, 64, 32))
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--records", type=int, default=32)
    ap.add_argument("--epochs", type=int, default=1500)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--decay-every", type=int, default=500)
    ap.add_argument("--res-tol", type=float, default=3e-5)
    ap.add_argument("--res-T-tol", type=float, default=0.25)
    ap.add_argument("--dT-max", type=float, default=20.0)
    ap.add_argument("--chi-max", type=float, default=0.0)
    ap.add_argument("--chain-max", type=float, default=0.5)
    ap.add_argument("--label-mode", type=int, default=3)
    ap.add_argument("--label-scale", type=float, default=1e-6)
    ap.add_argument("--label-rel", type=float, default=0.1)
    ap.add_argument("--label-floor", type=float, default=1e-14)
    ap.add_argument("--y-floor", type=float, default=1e-12)
    ap.add_argument("--trace", type=float, default=1e-6,
                    help="extra reactors seeded with log-uniform species traces up to this (0: none)")
    ap.add_argument("--trace-shape", type=int, nargs=3, default=(64, 64, 16))
    ap.add_argument("--elem-tol", type=float, default=1e30)
    ap.add_argument("--seed", type=int, default=12345)
    ap.add_argument("--flow-steps", type=int, default=8,
                    help="time steps of the reacting mixing layer sampled (0: reactors only)")
    ap.add_argument("--flow-size", type=int, default=192)
    ap.add_argument("--flow-per-call", type=int, default=262144)
    ap.add_argument("--flow-seed", type=int, default=999)
    a = ap.parse_args(argv)

    pkg = _load_package()
    import flame3d_amd.abi  # noqa: F401
    import flame3d_amd.cases  # noqa: F401
    import flame3d_amd.mech as mm
    import flame3d_amd.surrogate as sur
    import torch

    def log(msg):
        print(msg, flush=True)

    mech = mm.load(a.mech)
    air = a.mech == "air5"
    dt = a.dt or (sur.AIR_DT if air else sur.DT)
    T_range = tuple(a.T_range or (AIR_T if air else H2_T))
    P_atm = tuple(a.P_atm or (AIR_P if air else H2_P))
    t0 = time.time()
    log(f"generating data: {a.samples_shape} reactors x {a.records} records over {a.steps} "
        f"steps of {dt:g} s, T {T_range} K, P {P_atm} atm")
    data = sur.generate(mech, tuple(a.samples_shape), a.steps, a.records, a.seed,
                        T_range=T_range, P_atm=P_atm, dt=dt, log=log)
    if a.trace > 0:
        log(f"generating trace-seeded reactors: {a.trace_shape}, traces up to {a.trace:g}")
        data = sur.concat(data, sur.generate(mech, tuple(a.trace_shape), a.steps, a.records,
                                             a.seed + 7, T_range=T_range, P_atm=P_atm,
                                             trace=a.trace, dt=dt, log=log))
    if a.flow_steps > 0:
        log(f"generating flow data: {a.flow_size}^3 flow block, {a.flow_steps} steps")
        data = sur.concat(data, sur.generate_flow(mech, (a.flow_size,) * 3, a.flow_steps,
                                                  a.flow_per_call, a.flow_seed, dt=dt,
                                                  log=log))
    t_gen = time.time() - t0
    M = data.X.shape[0]
    log(f"  {M} samples in {t_gen:.1f}s, mean Newton iterations {data.newton_iters:.2f}")
    cfg = sur.make_config(mech, data, label_mode=a.label_mode, res_tol=a.res_tol,
                           res_T_tol=a.res_T_tol, dT_max=a.dT_max, chi_max=a.chi_max,
                           chain_max=a.chain_max,
                           label_scale=a.label_scale, label_rel=a.label_rel,
                           label_floor=a.label_floor, y_floor=a.y_floor, elem_tol=a.elem_tol,
                           dt=dt)
    Xn, Yn = sur.normalize(cfg, data)
    del data
    # hold out 1/16 for validation
    g = torch.Generator(device="cuda").manual_seed(a.seed)
    perm = torch.randperm(M, generator=g, device="cuda")
    nval = (M // 16 // 64) * 64
    Xv, Yv = Xn[perm[:nval]].contiguous(), Yn[perm[:nval]].contiguous()
    Xt, Yt = Xn[perm[nval:]].contiguous(), Yn[perm[nval:]].contiguous()
    del Xn, Yn
    ctx = pkg.abi.Context(8, 8, 8, mech, precision="fp64")
    ctx.nn_init(cfg)
    ctx.nn_set_params(sur.init_params(a.seed, mech.Ns + 2, mech.Ns))
    t1 = time.time()
    hist = sur.train(ctx, Xt, Yt, a.epochs, a.batch, a.lr, 0.3, a.decay_every, a.seed, log=log)
    torch.cuda.synchronize()
    t_train = time.time() - t1
    val = sur.mse(ctx, Xv, Yv)
    params = ctx.nn_get_params()
    meta = {"samples": M, "epochs": a.epochs, "batch": a.batch, "lr": a.lr,
            "final_train_loss": hist[-1], "val_mse": val, "train_seconds": t_train,
            "generate_seconds": t_gen, "shape": list(a.samples_shape), "steps": a.steps,
            "records": a.records, "seed": a.seed, "label_mode": a.label_mode,
            "trace": a.trace, "flow_steps": a.flow_steps, "mech": a.mech, "dt": dt,
            "T_range": list(T_range), "P_atm": list(P_atm)}
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    sur.save(a.out, cfg, params, meta)
    ctx.close()
    log(f"saved {a.out}: val mse {val:.3e}, d_max {cfg.d_max:.3f}")
    ev = evaluate(pkg, mech, a.out, log=log, dt=dt, T_range=T_range, P_atm=P_atm)
    meta["heldout"] = ev
    sur.save(a.out, cfg, params, meta)
    print(json.dumps(meta), flush=True)


if __name__ == "__main__":
    main()
