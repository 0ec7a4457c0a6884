#!/usr/bin/env python3
"""Benchmark of the PP-ES nodal DG step for 2D ideal MHD on one MI355X.

Workload (the GPU-scale single-instance configuration): Orszag-Tang on
[0,2pi]^2 (gamma = 5/3) plus a seeded smooth perturbation (amplitude 1e-3 on
rho and, through a vector potential, on B; integer modes 0 < |k| <= 4, phases
from numpy default_rng(0)), P3 (4x4 GLL nodes), 1024 x 1024 elements, FP64,
SSP-RK3 with CFL 0.2, LDF + positivity limiter every stage, OEDG on the final
stage. A "step" is one SSP-RK3 step (three stages) of the whole instance.

Prints ONE JSON line. value = DOF-stage updates per second, where a DOF is one
GLL node of one element (8 FP64 variables) and a DOF-stage update is one node
advanced through one RK stage (residual + combine + post-process).

    python bench.py [--gpus 1] [--steps 200] [--warmup 5] [--no-cpu]
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "pp-es-dg-mhd_amd"))

METRIC = "DOF-stage updates/s, batched vs plain at matched L1 err; p50/p99 solve latency"
HBM_PEAK_GBS = 8000.0

# Algorithmic (compulsory) HBM bytes per GLL node for one launch of each
# kernel class: 8 FP64 variables = 64 B per node per state stream.
#   residual: read u, write ut (neighbour traces and ghosts not counted)
#   stage:    read u^n, ut (+ u^(s-1) in stages 2, 3), write u^(s): mean over
#             the three SSP-RK3 stages (cell averages, LDF, limiter fused)
#   oedg:     read u once (face indicators)
#   limit:    read u, write u (OEDG damping + limiter)
ALG_BYTES_PER_NODE = {"residual": 128.0, "stage": (192.0 + 256.0 + 256.0) / 3.0, "oedg": 64.0,
                      "limit": 128.0, "reduce": 0.0, "bc": 0.0}
# whole path, SSP-RK3 with OEDG on the final stage: 640 B per node-step
STEP_BYTES_PER_NODE_STAGE = 640.0 / 3.0


def perturbation_table(seed=0, kmax=4):
    ks = [(kx, ky) for kx in range(-kmax, kmax + 1) for ky in range(-kmax, kmax + 1)
          if 0 < kx * kx + ky * ky <= kmax * kmax]
    ph = np.random.default_rng(seed).uniform(0.0, 2.0 * math.pi, size=(2, len(ks)))
    return np.array([[kx, ky, ph[0, q], ph[1, q]] for q, (kx, ky) in enumerate(ks)])


def cpu_baseline(p, cfl):
    """The CPU oracle (plain-C restatement, OpenMP over elements) on a bounded
    sample of the same problem: 512 x 512 elements, start-up + 4 steps."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    cores = min(16, os.cpu_count() or 1)
    O.set_threads(cores)
    nx = 512
    steps = 4
    o = O.Oracle(nx, nx, p, (0.0, 2 * math.pi, 0.0, 2 * math.pi), 5.0 / 3.0, "periodic")
    o.init_problem("otpert", perturbation_table(0), 1e-3)
    t0 = time.perf_counter()
    s = o.run(-1.0, steps, cfl, 1, True)
    dt = time.perf_counter() - t0
    stage_equiv = 3 * s["steps"] + 1  # the start-up post-process + residual counts as one stage
    value = nx * nx * (p + 1) ** 2 * stage_equiv / dt
    return {"value": value, "unit": "DOF-stage updates/s", "cores": cores, "kind": "port",
            "sample": f"oracle/ppes_oracle.c, OT-2pi + perturbation P{p} {nx}x{nx}, start-up + {steps} RK3 steps "
                      f"({stage_equiv} stage-equivalents) in {dt:.1f} s"}


def load_traffic(kernel):
    """HBM bytes per launch from the committed rocprofv3 PMC summary, if any."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_p3_1024_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("bytes_per_launch")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--nx", type=int, default=1024)
    ap.add_argument("--p", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only a plain 1 GiB device-to-device copy (reference for HBM byte counters)")
    a = ap.parse_args()
    if a.calibrate_copy:
        import torch
        n = 1 << 27  # 2^27 doubles = 1 GiB
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 8 * n, "copy_bytes_written": 8 * n, "launches": 3}))
        return
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(instances shard across GPUs with no communication; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import ppesdg  # raises if the HIP library is missing: there is no fallback

    cfl = 0.2
    nx = ny = a.nx
    p = a.p
    nodes = nx * ny * (p + 1) ** 2
    s = ppesdg.Solver(nx, ny, p, (0.0, 2 * math.pi, 0.0, 2 * math.pi), 5.0 / 3.0, "periodic")
    s.init_problem("otpert", perturbation_table(0), 1e-3)
    s.solve_begin(use_ldf=True)
    if a.warmup:
        s.advance(-1.0, a.warmup, cfl, 1)
    s.sync()
    before = s.advance(-1.0, 0, cfl, 1)  # no-op: reads the cumulative counters
    s.set_timing(True)
    t0 = time.perf_counter()
    after = s.advance(-1.0, a.steps, cfl, 1)
    s.sync()
    elapsed = time.perf_counter() - t0
    stages = after["stages"] - before["stages"]
    value = nodes * stages / elapsed

    # per-kernel device time over the timed region (HIP events on the solver's stream)
    kt = {k: s.kernel_time(k) for k in ppesdg.KERNELS}
    s.set_timing(False)
    dominant = max((k for k in kt if kt[k][1] > 0), key=lambda k: kt[k][0])
    tot_ms, launches = kt[dominant]
    avg_s = tot_ms / launches / 1e3
    bytes_per_launch = ALG_BYTES_PER_NODE[dominant] * nodes
    achieved = bytes_per_launch / avg_s / 1e9
    traffic = load_traffic(dominant)

    # sanity of the result: exact positivity and mass conservation (periodic)
    u = s.download_state(0)[:, :, :, 1:-1, 1:-1]
    g = 5.0 / 3.0
    prs = (g - 1.0) * (u[7] - 0.5 * ((u[1] ** 2 + u[2] ** 2 + u[3] ** 2) / u[0] + (u[4] ** 2 + u[5] ** 2 + u[6] ** 2)))
    checks = {"min_rho": float(u[0].min()), "min_p": float(prs.min()), "retries": after["retries"],
              "pp_limited_cells": after["negcells"]}
    del u, prs

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "DOF-stage updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": elapsed / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"Orszag-Tang 2pi + seeded perturbation, P{p}, {nx}x{ny} elements, SSP-RK3, "
                               f"CFL {cfl}, LDF+PP every stage, OEDG final stage, plain single-instance path",
                   "nodes": nodes, "dof_per_node": 8, "path": "plain"},
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "avg_launch_ms": avg_s * 1e3, "alg_bytes_per_launch": bytes_per_launch},
        "step_roofline": {"alg_bytes_per_node_stage": STEP_BYTES_PER_NODE_STAGE,
                          "achieved_GBs": STEP_BYTES_PER_NODE_STAGE * value / 1e9,
                          "frac_hbm": STEP_BYTES_PER_NODE_STAGE * value / 1e9 / HBM_PEAK_GBS},
        "kernel_ms_per_step": {k: v[0] / a.steps for k, v in kt.items()},
        "checks": checks,
    }
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(p, cfl)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
