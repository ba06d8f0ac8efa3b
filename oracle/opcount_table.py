"""FP64 operation counts of the chemistry kernels' work per cell, taken from
the oracle's own operation sequence (the chemistry templates of
f3d_oracle.cpp run on a counting scalar): TEST/MEASUREMENT INFRASTRUCTURE,
writes profiles/chem_opcount.json, which the benchmark reads as data.

  python oracle/opcount_table.py [--check]

Per mechanism (9-species H2/O2, 5-species air):
  * newton_iteration: one iteration of the backward-Euler Newton solve
    (rates with the exact Jacobian and the analytic dR/dT, the
    (Ns+1)^2 system, its pivoted solve, the update), mean over reactors of
    the 128^3 chemistry workload's distribution (T ~ U[900, 2500] K,
    P ~ U[1, 10] atm, Beta(2,2) mixture fraction; seed 12345) for H2/O2 and
    of the dissociating-air distribution for air, at the benchmarks' steps;
  * newton_cell_overhead: per cell outside the iterations (the energy E0,
    two e -> T updates of the final GetT);
  * rates: one production-rate evaluation without the Jacobian;
  * getT_update: one e -> T Newton update (InternalEnergy, CV);
  * surrogate_epilogue: the FP64 work per cell of the surrogate step's gates
    and energy bookkeeping, counted as one rate evaluation at the predicted
    state (the residual gate) plus two e -> T updates -- a lower bound (the
    chain gate's carrier-block derivatives and the projection are not
    counted).
flop = add/sub + mul + div; transcendental calls (exp, log, log10, pow) are
counted separately and not included in flop.
"""
from __future__ import annotations

import argparse
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
OUT = os.path.join(ROOT, "profiles", "chem_opcount.json")


def _flop(c):
    return c["add"] + c["mul"] + c["div"]


def _mean(cs, n=1.0):
    keys = ("add", "mul", "div", "trans", "sqrt")
    return {k: float(np.mean([c[k] for c in cs])) / n for k in keys}


def compute(n_cells=256, seed=12345):
    sys.path.insert(0, ROOT)
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import conftest  # noqa: F401  loads the flame3d_amd package
    import flame3d_amd.cases as cases
    import flame3d_amd.mech as mm
    from oracle import Oracle

    res = {}
    for name, h in (("h2_lidryer", 1e-7), ("air5", 6e-9)):
        m = mm.load(name)
        if name == "h2_lidryer":
            Qt, rt = cases.chem_cells((8, 8, 4), m, seed=seed, dev="cpu")
        else:
            Qt, rt = cases.chem_cells_air((8, 8, 4), m, seed=seed)
        Q = Qt[:, 4:-4, 4:-4, 4:-4].reshape(7, -1).numpy()
        r = rt[:, 4:-4, 4:-4, 4:-4].reshape(m.Ns, -1).numpy()
        orc = Oracle(1, 1, 1, m)
        per_it, tot_it = [], 0
        n = min(n_cells, Q.shape[1])
        for c in range(n):
            cnt = orc.newton_opcount(r[:, c].copy(), float(Q[5, c]), h)
            if cnt["iterations"] > 0:
                per_it.append({k: cnt[k] / cnt["iterations"] for k in cnt})
                tot_it += cnt["iterations"]
        c0 = n // 2
        rates = orc.rates_opcount(r[:, c0].copy(), float(Q[5, c0]), full_jac=False)
        gt = orc.getT_update_opcount(r[:, c0].copy(), float(Q[5, c0]), float(Q[6, c0]))
        e0 = {k: gt[k] for k in ("add", "mul", "div", "trans", "sqrt")}
        it = _mean(per_it)
        over = {k: 2 * gt[k] + e0[k] for k in e0}
        epi = {k: rates[k] + 2 * gt[k] for k in e0}
        res[name] = {
            "Ns": m.Ns, "Nr": m.Nr, "step_h_s": h, "cells_sampled": n,
            "mean_iterations": tot_it / max(1, len(per_it)),
            "newton_iteration": dict(it, flop=_flop(it)),
            "newton_cell_overhead": dict(over, flop=_flop(over)),
            "rates": dict({k: rates[k] for k in e0}, flop=_flop(rates)),
            "getT_update": dict(e0, flop=_flop(e0)),
            "surrogate_epilogue": dict(epi, flop=_flop(epi)),
        }
    return {"generator": "oracle/opcount_table.py", "flop": "add/sub + mul + div (FP64); "
            "transcendental calls counted separately", "mechanisms": res}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--check", action="store_true", help="compare with the committed table")
    a = ap.parse_args()
    res = compute()
    if a.check:
        old = json.load(open(OUT))
        assert old == json.loads(json.dumps(res)), "profiles/chem_opcount.json is stale"
        print("up to date")
        return
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    with open(OUT, "w") as fh:
        json.dump(res, fh, indent=1)
    print(json.dumps(res, indent=1))


if __name__ == "__main__":
    main()
