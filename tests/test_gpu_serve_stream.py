"""A prefix of the serving stream at its stated parameters, served for real
and checked against the FP64 oracle.

The stream: Poisson arrivals at 20 requests/s (numpy default_rng(2)) of
Orszag-Tang / rotor / blast solves, grid sides 64-512, P1-P4, t_final uniform
in [0.05, 0.5] (workloads.serving_stream). Its first 8 requests are served by
the continuous-batching loop (serve.run_stream: real-time arrivals, admission
under the 200 GB budget, one batched step per round) with their grids
reduced 8x per side and their end times scaled per problem (OT 1/20, rotor
1/100, blast 1/4000) so that the CPU oracle follows in seconds. Every served
result must match the oracle's solve of the same request: equal step count,
relative L1 density within 1e-10, every variable within 1e-10 normwise, and
density and pressure positive at every stage. Retry semantics follow
drive.f90:86-120 on both sides.
"""
import numpy as np
import pytest

import oracle as O
import ppesdg
import serve
import workloads as W
from helpers import interior, normwise, rel_l1

pytestmark = pytest.mark.gpu

SCALE = {"ot2pi": 1.0 / 20.0, "rotor": 1.0 / 100.0, "blast": 1.0 / 4000.0}


def prefix(n=8):
    reqs = W.serving_stream(rate=20.0, duration=120.0, seed=2, sides=(64, 128, 256, 512), degrees=(1, 2, 3, 4),
                            t_frac=(0.1, 1.0), horizon={"ot2pi": 0.5, "rotor": 0.5, "blast": 0.5})[:n]
    out = []
    for k, r in enumerate(reqs):
        q = dict(r, tag=k)
        q["nx"] = q["ny"] = r["nx"] // 8
        q["t_final"] = r["t_final"] * SCALE[r["prob"]]
        out.append(q)
    return out


def test_stream_prefix_matches_oracle():
    reqs = prefix()
    assert [r["prob"] for r in reqs].count("blast") >= 2 and len({r["p"] for r in reqs}) >= 3
    done, wall, rejected = serve.run_stream(reqs, 200e9, use_graph=True, keep_results=True)
    assert not rejected and len(done) == len(reqs)
    for r in sorted(done, key=lambda q: q["tag"]):
        dom, g, bc, bcs, cfl, oedg = W.PROBLEMS[r["prob"]]
        o = O.Oracle(r["nx"], r["ny"], r["p"], dom, g, bc, bcs)
        o.init_problem(r["prob"])
        so = o.run(r["t_final"], 0, cfl, oedg, True)
        got, ref = interior(r["result"]), interior(np.array(o.u))
        err, l1 = normwise(got, ref), rel_l1(got[0], ref[0])
        print(f"request {r['tag']} {r['prob']} {r['nx']}^2 P{r['p']} t_final {r['t_final']:.3e}: "
              f"steps {r['stats']['steps']}/{so['steps']}, retries {r['stats']['retries']}/{so['retries']}, "
              f"L1(rho) {l1:.2e}, normwise {err:.2e}, latency {r['t_done'] - r['arrival']:.3f} s")
        assert r["stats"]["state"] == "done" and so["status"] == 0
        assert r["stats"]["steps"] == so["steps"] and r["stats"]["retries"] == so["retries"]
        assert abs(r["stats"]["t"] - so["t"]) <= 1e-14 * so["t"]
        assert l1 <= 1e-10 and err <= 1e-10
        q = got
        prs = (g - 1.0) * (q[7] - 0.5 * ((q[1] ** 2 + q[2] ** 2 + q[3] ** 2) / q[0] + (q[4] ** 2 + q[5] ** 2 + q[6] ** 2)))
        assert q[0].min() > 0.0 and prs.min() > 0.0
        assert r["positivity"][0] > 0.0 and r["positivity"][1] > 0.0   # every stage of the solve
