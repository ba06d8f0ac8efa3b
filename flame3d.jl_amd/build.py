"""Build libflame3d.so (gfx950) in-tree with hipcc.

    python flame3d.jl_amd/build.py [--jobs N]

Every .hip file under csrc/ is compiled to an object in _build/ (in
parallel), then linked into flame3d.jl_amd/libflame3d.so, after
mechanisms/gen_mech_tables.py has written csrc/f3d_mech_tables.h.  Flags:
  --offload-arch=gfx950            CDNA4 only, no other targets
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                   no packed FP32 instructions
  -fno-hip-fp32-correctly-rounded-divide-sqrt
                                   hardware float div/sqrt on the FP32 path
                                   (double div/sqrt stay IEEE)
  chem.hip also: -ffp-contract=off  the FP64 chemistry rounds every product
                                   and sum as the oracle does (with its own
                                   exp/log), so the two agree bit for bit
"""
from __future__ import annotations

import argparse
import glob
import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libflame3d.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-fno-hip-fp32-correctly-rounded-divide-sqrt", "-Wall", "-Wno-unused-function",
         "-Wno-unused-variable"]


FILE_FLAGS = {"chem.hip": ["-ffp-contract=off"]}


def build(jobs: int = 8, verbose: bool = False) -> str:
    # the shipped mechanisms' reaction tables (csrc/f3d_mech_tables.h)
    subprocess.run([sys.executable, os.path.join(HERE, "mechanisms", "gen_mech_tables.py")],
                   check=True)
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    objs = []
    procs = []
    t0 = time.time()
    for src in srcs:
        obj = os.path.join(BUILD, os.path.basename(src) + ".o")
        objs.append(obj)
        cmd = [HIPCC, *FLAGS, *FILE_FLAGS.get(os.path.basename(src), []), "-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        while len([p for p in procs if p[0].poll() is None]) >= jobs:
            time.sleep(0.1)
        procs.append((subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                                       text=True), src))
    failed = False
    for p, src in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            failed = True
            sys.stderr.write(f"--- {os.path.basename(src)} ---\n{out}\n")
        elif out.strip() and verbose:
            sys.stderr.write(out)
    if failed:
        raise RuntimeError("hipcc compilation failed")
    tmp = LIB + ".tmp"
    subprocess.run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs],
                   check=True)
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} in {time.time() - t0:.1f}s", flush=True)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    a = ap.parse_args()
    build(a.jobs, verbose=True)
