"""bench.py against the committed rocprofv3 summaries: every kernel the bench
line can name as dominant finds its measured HBM traffic, and the summaries
carry the commands that made them."""
import os
import re

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _hits(traffic, sym):
    if "<" in sym:
        return [v for k, v in traffic.items() if re.fullmatch(sym, k)]
    return [v for k, v in traffic.items() if k == sym or k.startswith(sym + "<")]


def test_main_stream_kernels_find_their_traffic():
    traffic = bench.measured_traffic()
    assert traffic, "profiles/hbm_traffic_dam_break_1m.txt has no table"
    for name in ("pbf_lambda", "pbf_delta", "pbf_delta_final", "pbf_build_list", "pbf_xsph",
                 "predict_count", "reorder", "mc_density_field"):
        hits = _hits(traffic, bench.KERNEL_SYMBOL[name])
        assert len(hits) == 1, (name, bench.KERNEL_SYMBOL[name], hits)
        assert hits[0] > 0


def test_profiles_name_their_commands():
    for f in ("kernel_trace_dam_break_1m.md", "hbm_traffic_dam_break_1m.txt",
              "sq_counters_dam_break_1m.md"):
        text = open(os.path.join(ROOT, "profiles", f)).read()
        assert "rocprofv3" in text and "bench.py" in text, f
