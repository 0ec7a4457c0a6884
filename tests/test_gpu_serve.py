"""Serving loop (continuous batching): requests that arrive while others run
are admitted into the running batch as memory frees up; each result is the
same bit for bit as the request solved alone, and the memory budget holds at
every round."""
import numpy as np
import pytest

import ppesdg
import serve
import workloads as W

pytestmark = pytest.mark.gpu

# arrival round, request
STREAM = [
    (0, dict(prob="ot2pi", nx=12, ny=12, p=2, t_final=0.02)),
    (0, dict(prob="blast", nx=10, ny=10, p=1, t_final=2e-4)),
    (2, dict(prob="rotor", nx=9, ny=11, p=3, t_final=2e-3)),
    (3, dict(prob="ot2pi", nx=8, ny=10, p=4, max_steps=6)),
    (5, dict(prob="blast", nx=8, ny=8, p=3, t_final=1e-4)),
    (9, dict(prob="rotor", nx=10, ny=10, p=2, max_steps=4)),
]


def alone(r):
    dom, g, bc, bcs, cfl, oedg = W.PROBLEMS[r["prob"]]
    s = ppesdg.Solver(r["nx"], r["ny"], r["p"], dom, g, bc, bcs)
    s.init_problem(r["prob"])
    st = s.solve(r.get("t_final", -1.0), r.get("max_steps", 0), cfl, oedg, True)
    return s.download_state(0), st


def serve_rounds(budget):
    srv = serve.Server(budget_bytes=budget, keep_results=True)
    clock = {"r": 0}
    done, k, peak = [], 0, 0
    while k < len(STREAM) or srv.busy():
        while k < len(STREAM) and STREAM[k][0] <= clock["r"]:
            srv.submit(dict(STREAM[k][1], tag=k), clock["r"])
            k += 1
        done += srv.round(lambda: clock["r"])
        peak = max(peak, srv.used)
        assert srv.used <= budget
        clock["r"] += 1
    srv.close()
    return sorted(done, key=lambda r: r["tag"]), peak


@pytest.mark.parametrize("budget", [200e9, None])
def test_serving_results_equal_alone(budget):
    if budget is None:  # room for the two largest requests only
        sizes = sorted(ppesdg.instance_bytes(r["nx"], r["ny"], r["p"]) for _, r in STREAM)
        budget = sizes[-1] + sizes[-2]
    done, peak = serve_rounds(budget)
    assert [r["tag"] for r in done] == list(range(len(STREAM)))
    for r in done:
        u, st = alone(STREAM[r["tag"]][1])
        assert r["stats"]["state"] == "done"
        assert r["stats"]["steps"] == st["steps"] and r["stats"]["t"] == st["t"]
        assert np.array_equal(r["result"][:, :, :, 1:-1, 1:-1], u[:, :, :, 1:-1, 1:-1])
        assert r["t_admit"] >= r["t_submit"]


def test_serving_admits_while_running():
    """a request arriving later joins instances that are still running"""
    done, _ = serve_rounds(200e9)
    first = done[0]
    late = [r for r in done if r["t_submit"] > 0]
    assert any(r["t_admit"] < first["t_done"] for r in late)
