"""bench.py's map from the fused FP32 stage's timed parts to rocprofv3 kernel
symbols names kernels that the committed traffic profile of the reacting
mixing layer holds (a symbol spelled differently from the trace, e.g. with
a return type a non-template kernel's trace name lacks, would leave that
part without counter traffic)."""
import importlib.util
import json
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROFILE = os.path.join(ROOT, "profiles", "hbm_traffic_reacting_mixing_layer_256.json")


def _bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_fused_stage_symbols_are_in_the_profile():
    bench = _bench()
    kernels = json.load(open(PROFILE))["kernels"]
    for part in ("mixture", "shock_sensor", "sweep_conv", "sweep_diff", "stage_end"):
        syms = bench.kernel_symbols(part, 9)
        assert syms, part
        for sym, share in syms:
            assert sym in kernels, (part, sym)
            assert 0 < share <= 1
        # every launch of a stage part is one of its kernels: the x sweep takes
        # one of three directions, the y/z marches two
        if part.startswith("sweep"):
            assert abs(sum(w for _, w in syms) - (2.0 if part == "sweep_diff" else 1.0)) < 1e-12
