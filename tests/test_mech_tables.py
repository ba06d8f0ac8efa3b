"""The compile-time reaction tables of the shipped mechanisms
(flame3d.jl_amd/csrc/f3d_mech_tables.h) are current with the mechanism files
and hold exactly the arrays the library receives through f3d_kinetics
(mech.load's arrays, indexed as f3d_create copies them)."""
import ast
import importlib.util
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GEN = os.path.join(ROOT, "flame3d.jl_amd", "mechanisms", "gen_mech_tables.py")
HDR = os.path.join(ROOT, "flame3d.jl_amd", "csrc", "f3d_mech_tables.h")


def _gen():
    spec = importlib.util.spec_from_file_location("gen_mech_tables", GEN)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_header_is_current():
    assert open(HDR).read() == _gen().render()


def _tables(cls):
    text = open(HDR).read()
    body = text[text.index(f"struct {cls} {{"):]
    body = body[:body.index("\n};")]
    out = {}
    for name, val in re.findall(r"static constexpr (?:int|double) (\w+)\[[^=]*= (\{.*?\});", body):
        out[name] = np.array(ast.literal_eval(val.replace("{", "[").replace("}", "]")))
    return out


@pytest.mark.parametrize("name,cls", [("h2_lidryer", "MechH2LiDryer"), ("air5", "MechAir5")])
def test_tables_equal_mechanism_arrays(name, cls):
    import flame3d_amd.mech as mm
    m = mm.load(name)
    t = _tables(cls)
    # f3d_create: rtype[r], sgm[r], vf[r][n] = vf[n + Ns r], Arr[r][c] = Arr[c + 3 r], ...
    assert np.array_equal(t["rtype"], np.asarray(m.reaction_type))
    assert np.array_equal(t["sgm"], np.asarray(m.sgm))
    assert np.array_equal(t["vf"], np.asarray(m.vf).T)
    assert np.array_equal(t["vr"], np.asarray(m.vr).T)
    for k in ("Arr", "ef", "loP", "Troe"):
        assert np.array_equal(t[k], np.asarray(getattr(m, k), dtype=np.float64).T), k
