"""Write golden vectors for the 0-D H2/O2 point-implicit reactor.

Source: the reference's own std-only C++ integrator NN/H2/chemical_test.cpp
(Li-Dryer mechanism, main loop :914-1009), compiled from its source by
``make -C oracle ref`` into oracle/_ref/chemical_test and run here.  Its stdout
(12 significant digits per step: T at the start of the step, P, E, Y_H2,
Y_O2, Y_H2O after the step; rho and E0 with 18 digits) is parsed and a
subsample is stored as JSON next to this script.

    make -C oracle ref && python tests/golden/make_h2_0d.py
"""
import json
import os
import re
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
BIN = os.path.join(ROOT, "oracle", "_ref", "chemical_test")

CASES = {
    # name: (dt, t_end, T0, P0 [atm]) -- README 0-D case, an ignition that
    # crosses the 1000 K NASA-7 break, and a slow low-temperature case
    "h2_0d_1500K_10atm": ("1e-8", "1e-5", "1500", "10"),
    "h2_0d_1100K_1atm": ("1e-7", "2e-4", "1100", "1"),
    "h2_0d_980K_3atm": ("2e-7", "2e-4", "980", "3"),
}


def parse(text):
    rho = float(re.search(r"^rho = (\S+)", text, re.M).group(1))
    e0 = float(re.search(r"^E = (\S+)", text, re.M).group(1))
    blocks = text.split("---------------")[1:]
    rows = []
    for b in blocks:
        g = lambda pat: float(re.search(pat, b, re.M).group(1))
        rows.append([g(r"^Time = (\S+) s"), g(r"^T = (\S+) K"), g(r"^P = (\S+) Pa"),
                     g(r"^E = (\S+)"), g(r"^Mass frac of H2:\s+(\S+)"),
                     g(r"^Mass frac of O2:\s+(\S+)"), g(r"^Mass frac of H2O:\s+(\S+)")])
    return rho, e0, rows


def main():
    for name, args in CASES.items():
        with tempfile.TemporaryDirectory() as td:
            out = subprocess.run([BIN, *args], cwd=td, check=True, capture_output=True,
                                 text=True).stdout
        rho, e0, rows = parse(out)
        keep = [i for i in range(len(rows)) if i < 10 or i % 10 == 9 or i == len(rows) - 1]
        doc = {"source": "NN/H2/chemical_test.cpp (reference), compiled by oracle/Makefile",
               "args": {"dt": float(args[0]), "t_end": float(args[1]), "T0": float(args[2]),
                        "P0_atm": float(args[3])},
               "Y0": {"H2": 0.5, "O2": 0.5}, "rho": rho, "E0": e0, "nsteps": len(rows),
               "columns": ["time", "T", "P", "E", "Y_H2", "Y_O2", "Y_H2O"],
               "step_index": keep, "rows": [rows[i] for i in keep]}
        with open(os.path.join(HERE, name + ".json"), "w") as fh:
            json.dump(doc, fh, indent=0)
        print(name, len(rows), "steps,", len(keep), "kept")


if __name__ == "__main__":
    main()
