"""The algorithmic FP64 operation counts that bench.py prices its FP64
roofline with come from the oracle built with a counting scalar type
(oracle/flopcount.cpp). The counter is checked on a closed form (the SSP-RK3
combine: 2 + 5 + 5 operations per value over three stages, 8 values per
node) and bench.py's pinned constants against a fresh count."""
import json
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def counts():
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "flopcount"])
    out = subprocess.check_output([os.path.join(ROOT, "oracle", "_build", "flopcount")], text=True)
    return {r["p"]: r for r in json.loads(out)}


def test_rk_combine_closed_form(counts):
    for p in (1, 2, 3, 4):
        assert counts[p]["rk_combine"] == pytest.approx(8 * (2 + 5 + 5) / 3, abs=1e-9)


def test_counts_grow_with_degree(counts):
    # more EC pairs per node (p per node) and larger OEDG derivative work
    for key in ("residual", "post_process", "post_process_oedg"):
        assert [counts[p][key] for p in (1, 2, 3, 4)] == sorted(counts[p][key] for p in (1, 2, 3, 4))


def test_bench_constants_match_the_count(counts):
    import sys
    sys.path.insert(0, ROOT)
    import bench
    for p, c in bench.ALG_FLOPS_PER_NODE.items():
        assert c["residual"] == pytest.approx(counts[p]["residual"], abs=0.01)
        assert c["node_stage"] == pytest.approx(counts[p]["node_stage_rk3_oedg_final"], abs=0.01)
