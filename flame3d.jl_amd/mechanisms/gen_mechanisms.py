"""Regenerate the packaged mechanism files from the reference's Cantera YAML
mechanisms (NN/Air/air.yaml, NN/H2/LiDryer.yaml).  The YAML files are read in
place; only the derived parameter arrays are written here, so the GPU host
needs no YAML and no reference tree.

    python flame3d.jl_amd/mechanisms/gen_mechanisms.py [/root/reference]
"""
import importlib.util
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
spec = importlib.util.spec_from_file_location("f3d_mech", os.path.join(HERE, "..", "mech.py"))
mech = importlib.util.module_from_spec(spec)
sys.modules["f3d_mech"] = mech
spec.loader.exec_module(mech)

ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
for name, rel in (("air5", "NN/Air/air.yaml"), ("h2_lidryer", "NN/H2/LiDryer.yaml")):
    m = mech.load_yaml(os.path.join(ref, rel))
    with open(os.path.join(HERE, name + ".json"), "w") as fh:
        fh.write(m.to_json())
    print(name, "Ns", m.Ns, "Nr", m.Nr, "T range", m.min_temp, m.max_temp)
