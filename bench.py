#!/usr/bin/env python3
"""Benchmark of the PP-ES nodal DG step for 2D ideal MHD on one MI355X.

Headline workload (`value`): 64 uneven instances drawn with numpy
default_rng(1) -- each one Orszag-Tang (2pi, periodic, CFL 0.2), rotor (fixed
ambient ghosts, CFL 0.2) or blast wave (transmissive, CFL 0.15, OEDG every
stage), grid side uniform in {32, 64, 128, 256}, degree uniform in P1-P4 --
FP64, SSP-RK3, LDF + positivity limiter, each instance with its own dt from
its own CFL condition. The batched path advances all of them together: one
launch per kernel class and polynomial degree per RK stage, the whole step
replayed from a captured hipGraph. A "step" is one RK step of every
instance. The same instances are then run one at a time through the plain
path (one instance per launch sequence, no graph) for the batched/plain ratio,
and the final states of both paths are compared (matched L1 error).

Also reported: the single-instance workload (Orszag-Tang 2pi + seeded
perturbation, P3, 1024 x 1024, 200 RK steps), per-kernel device times (HIP
events on the engine's stream), the HBM roofline of the dominant kernel and
the residual's algorithmic FP64 rate against the FP64 peak, and a CPU
baseline (the CPU oracle on a bounded sample of the mixed workload).

value = DOF-stage updates per second: a DOF is one GLL node of one element
(8 FP64 variables); a DOF-stage update is one node advanced through one RK
stage (residual + combine + post-process).

    python bench.py [--gpus 1] [--steps 100] [--warmup 3] [--no-cpu] [--no-single]
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

import workloads as W  # noqa: E402

METRIC = "DOF-stage updates/s, batched vs plain at matched L1 err; p50/p99 solve latency"
HBM_PEAK_GBS = 8000.0

# Algorithmic (compulsory) HBM bytes per GLL node for one launch of each
# kernel class: 8 FP64 variables = 64 B per node per state stream; means over
# the three launches of an SSP-RK3 step.
#   residual: read u, write ut (neighbour traces and ghosts not counted); the
#             launches of L(u1) and L(u2) also read u^n and store the next
#             stage's combined state in place of ut: 128, 192, 192
#   stage:    stage 1 reads u^n and ut, stages 2 and 3 the combined state;
#             each writes its stage (cell averages, LDF, limiter fused): 192,
#             128, 128
#   oedg:     read u once (face indicators)
#   limit:    read u, write u (OEDG damping + limiter)
ALG_BYTES_PER_NODE = {"residual": (128.0 + 192.0 + 192.0) / 3.0, "stage": (192.0 + 128.0 + 128.0) / 3.0,
                      "oedg": 64.0, "limit": 128.0, "reduce": 0.0, "bc": 0.0}
# whole path, SSP-RK3 with OEDG on the final stage: 640 B per node-step
STEP_BYTES_PER_NODE_STAGE = 640.0 / 3.0

# Algorithmic FP64 operations per GLL node (add, sub, mul, div, sqrt, log,
# exp each one; the reference's algorithm: each EC pair flux and each face
# flux evaluated once), counted by oracle/flopcount.cpp on an Orszag-Tang
# state: one residual, and one node-stage of SSP-RK3 with OEDG on the final
# stage (residual + combine + post-process).
ALG_FLOPS_PER_NODE = {
    1: {"residual": 788.19, "node_stage": 1177.85},
    2: {"residual": 796.79, "node_stage": 1347.11},
    3: {"residual": 884.47, "node_stage": 1659.74},
    4: {"residual": 1003.60, "node_stage": 2062.88},
}
FP64_PEAK_TFLOPS = 78.6  # MI355X dense FP64 (vector and matrix alike)


def add_instance(b, inst, max_steps):
    dom, g, bc, bcs, cfl, oedg = W.PROBLEMS[inst["prob"]]
    i = b.add(inst["nx"], inst["ny"], inst["p"], dom, g, bc, bcs)
    b.init_problem(i, inst["prob"])
    b.start(i, -1.0, max_steps, cfl, oedg, True)
    return i


def launches_per_step(kernel, mix):
    """launches of one kernel class per batch step (one per degree group)"""
    groups = len({m["p"] for m in mix})
    oedg_stages = 3 if any(W.PROBLEMS[m["prob"]][5] == 2 for m in mix) else 1
    return {"residual": 3, "stage": 3, "oedg": oedg_stages, "limit": oedg_stages, "bc": 3}.get(kernel, 0) * groups


def run_batched(ppesdg, mix, steps, warmup, tsteps):
    """Batched path: warmup steps, `steps` timed graph steps, then `tsteps`
    steps without the graph with per-launch HIP events (kernel times)."""
    b = ppesdg.Batch(3)
    total = warmup + 1 + steps + tsteps
    ids = [add_instance(b, m, total) for m in mix]
    if warmup > 0:
        b.step(warmup, True)
    tuned = b.autotune(5)  # residual workgroup size per degree (the plan is rebuilt with it)
    b.step(1, True)  # the first step of a new plan runs without the graph: keep it out of the timing
    t0 = time.perf_counter()
    done, active = b.step(steps, True)
    elapsed = time.perf_counter() - t0  # step() returns after the last step's status is on the host
    if done != steps or active != len(mix):
        print(f"# warning: {done} of {steps} steps, {active} of {len(mix)} instances active", file=sys.stderr)
    b.set_timing(True)
    if tsteps > 0:
        b.step(tsteps, False)
    kt = {k: b.kernel_time(k) for k in ppesdg.KERNELS}
    b.set_timing(False)
    stats = [b.status(i) for i in ids]
    return b, ids, elapsed, kt, stats, tuned


def run_plain(ppesdg, mix, steps, warmup, tsteps, batch, ids):
    """Plain path: every instance alone (one launch sequence each, no graph);
    the timed part is each instance's `steps` steps after its warmup. The
    final state is compared with the batched run of the same instance."""
    elapsed = 0.0
    max_rel_l1 = 0.0
    bitwise = True
    for inst, i in zip(mix, ids):
        dom, g, bc, bcs, cfl, oedg = W.PROBLEMS[inst["prob"]]
        s = ppesdg.Solver(inst["nx"], inst["ny"], inst["p"], dom, g, bc, bcs)
        s.init_problem(inst["prob"])
        s.solve_begin(True)
        s.advance(-1.0, warmup, cfl, oedg)
        s.autotune(5)
        s.advance(-1.0, 1, cfl, oedg)
        s.sync()
        t0 = time.perf_counter()
        s.advance(-1.0, steps, cfl, oedg)
        elapsed += time.perf_counter() - t0
        s.advance(-1.0, tsteps, cfl, oedg)
        a = s.download_state(0)[:, :, :, 1:-1, 1:-1]
        s.close()
        bb = batch.download_state(i, 0)[:, :, :, 1:-1, 1:-1]
        bitwise = bitwise and bool(np.array_equal(a, bb))
        max_rel_l1 = max(max_rel_l1, float(np.sum(np.abs(a[0] - bb[0])) / np.sum(np.abs(a[0]))))
    return elapsed, max_rel_l1, bitwise


def run_single(ppesdg, steps, warmup):
    """Orszag-Tang 2pi + seeded perturbation, P3, 1024 x 1024, plain path."""
    cfl, nx, p = 0.2, 1024, 3
    nodes = nx * nx * (p + 1) ** 2
    s = ppesdg.Solver(nx, nx, p, (0.0, 2 * math.pi, 0.0, 2 * math.pi), 5.0 / 3.0, "periodic")
    s.init_problem("otpert", W.perturbation_table(0), 1e-3)
    s.solve_begin(use_ldf=True)
    if warmup:
        s.advance(-1.0, warmup, cfl, 1)
    tuned = s.autotune(5)
    s.advance(-1.0, 1, cfl, 1)
    s.sync()
    before = s.advance(-1.0, 0, cfl, 1)
    s.set_timing(True)
    t0 = time.perf_counter()
    after = s.advance(-1.0, steps, cfl, 1)
    s.sync()
    elapsed = time.perf_counter() - t0
    kt = {k: s.kernel_time(k) for k in ppesdg.KERNELS}
    s.set_timing(False)
    stages = after["stages"] - before["stages"]
    u = s.download_state(0)[:, :, :, 1:-1, 1:-1]
    g = 5.0 / 3.0
    prs = (g - 1.0) * (u[7] - 0.5 * ((u[1] ** 2 + u[2] ** 2 + u[3] ** 2) / u[0] + (u[4] ** 2 + u[5] ** 2 + u[6] ** 2)))
    r = kt["residual"]
    res = {"workload": f"Orszag-Tang 2pi + seeded perturbation (amplitude 1e-3, |k| <= 4, seed 0), P{p}, "
                       f"{nx}x{nx} elements, SSP-RK3, CFL {cfl}, {steps} steps, plain path",
           "value": nodes * stages / elapsed, "ms_per_step": elapsed / steps * 1e3, "steps": steps,
           "kernel_ms_per_step": {k: v[0] / steps for k, v in kt.items()},
           "residual_avg_launch_ms": r[0] / max(r[1], 1),
           "residual_hbm_frac": ALG_BYTES_PER_NODE["residual"] * nodes / (r[0] / max(r[1], 1) / 1e3) / 1e9
           / HBM_PEAK_GBS,
           "residual_fp64_frac": ALG_FLOPS_PER_NODE[p]["residual"] * nodes / (r[0] / max(r[1], 1) / 1e3) / 1e12
           / FP64_PEAK_TFLOPS,
           "min_rho": float(u[0].min()), "min_p": float(prs.min()), "retries": after["retries"],
           "residual_workgroup_lanes": tuned, "oedg_face_workgroup_lanes": s.face_tb(),
           "stage_workgroup_lanes": {s.shape: s.tiles("stage")}, "limit_workgroup_lanes": {s.shape: s.tiles("limit")}}
    s.close()
    return res


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(mix, budget_s=15.0):
    """The CPU oracle (plain-C restatement, OpenMP over elements, threads
    bound close to cores) on a bounded sample of the mixed workload:
    instances in mix order, start-up + 6 RK3 steps each, until about
    budget_s seconds of CPU work. Uses every core this process may run on, up
    to the 16 a one-GPU share of the host holds. Returns the baseline and
    each sampled instance's final interior density (for the oracle match)."""
    # before the OpenMP runtime starts (it starts with the oracle's library)
    os.environ["OMP_PROC_BIND"] = "close"
    os.environ["OMP_PLACES"] = "cores"
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    cores = max(1, min(16, avail))
    O.set_threads(cores)
    work, spent, used = 0.0, 0.0, 0
    finals = []
    for inst in mix:
        if spent >= budget_s:
            break
        dom, g, bc, bcs, cfl, oedg = W.PROBLEMS[inst["prob"]]
        o = O.Oracle(inst["nx"], inst["ny"], inst["p"], dom, g, bc, bcs)
        o.init_problem(inst["prob"])
        t0 = time.perf_counter()
        s = o.run(-1.0, 6, cfl, oedg, True)
        spent += time.perf_counter() - t0
        work += W.nodes(inst) * (3 * s["steps"] + 1)  # the start-up counts as one stage
        finals.append((inst, np.array(o.u[0, :, :, 1:-1, 1:-1]), s["steps"]))
        used += 1
        del o
    base = {"value": work / spent, "unit": "DOF-stage updates/s", "cores": cores, "kind": "port",
            "cpu": cpu_model(), "omp": "OMP_PROC_BIND=close OMP_PLACES=cores",
            "sample": f"oracle/ppes_oracle.c on the first {used} instances of the mixed workload, "
                      f"start-up + 6 RK3 steps each, {spent:.1f} s"}
    return base, finals


def match_oracle(ppesdg, finals):
    """The sampled instances through the batched GPU path for the same 6 RK3
    steps: max relative L1 of the interior density against the oracle."""
    b = ppesdg.Batch(3)
    ids = [add_instance(b, inst, 6) for inst, _, _ in finals]
    b.step(0, True)
    worst = 0.0
    for (inst, rho_o, steps_o), i in zip(finals, ids):
        st = b.status(i)
        if st["steps"] != steps_o:
            worst = float("inf")
            continue
        rho = b.download_state(i, 0)[0, :, :, 1:-1, 1:-1]
        worst = max(worst, float(np.sum(np.abs(rho - rho_o)) / np.sum(np.abs(rho_o))))
    b.close()
    return worst


def run_serving(duration, rate, sides):
    """Scaled serving stream: Poisson arrivals (seed 2) of requests from the
    mixed workload's problems with the given grid sides, P1-P4 and t_final a
    uniform fraction in [0.1, 1] of a per-problem horizon (OT 0.05, rotor
    0.01, blast 0.001); continuous batching under a 200 GB budget against
    the plain path serving the same stream one request at a time (first come
    first served). The default stream (sides 32-256) brings more work than
    the batched server clears in its arrival window, so both run saturated.
    Throughput is work over the makespan (first arrival to last result) for
    both."""
    import serve
    reqs = W.serving_stream(rate=rate, duration=duration, seed=2, sides=sides,
                            horizon={"ot2pi": 0.05, "rotor": 0.01, "blast": 0.001})
    t_first = reqs[0]["arrival"]
    done, wall, rejected = serve.run_stream(reqs, 200e9, use_graph=True)
    lat = [r["t_done"] - r["t_submit"] for r in done]
    span = max(r["t_done"] for r in done) - t_first
    total_work = sum(serve.work(r) for r in done)
    seq, busy = serve.run_sequential(reqs)
    lat_seq = [r["t_done"] - r["arrival"] for r in seq]
    span_seq = max(r["t_done"] for r in seq) - t_first
    work_seq = sum(serve.work(r) for r in seq)
    out = {"requests": len(reqs), "finished": len(done), "rejected": len(rejected),
           "stream": f"Poisson {rate}/s for {duration} s (seed 2), sides {'/'.join(map(str, sides))}, P1-P4, "
                     f"t_final in [0.1, 1] x (OT 0.05, rotor 0.01, blast 0.001)",
           "batched": {"value": total_work / span, "makespan_s": span, **serve.percentiles(lat)},
           "plain_sequential": {"value": work_seq / span_seq, "makespan_s": span_seq, "busy_s": busy,
                                **serve.percentiles(lat_seq)},
           "unit": "DOF-stage updates/s over the makespan; latency (arrival to result) in s"}
    out["batched_over_plain"] = out["batched"]["value"] / out["plain_sequential"]["value"]
    return out


def run_serving_config5(window, cap, budget=200e9):
    """The serving stream at its stated parameters: Poisson arrivals at
    20 requests/s (seed 2) of the mixed workload's problems, grid sides
    64/128/256/512, P1-P4, t_final uniform in [0.05, 0.5] (every problem),
    continuous batching under a 200 GB device budget, against the plain path
    serving the same stream one request at a time. Bounded: arrivals over the
    first `window` seconds of the stream (the same draws as the 120 s stream,
    cut there) and `cap` seconds of wall time for each server; a request
    still running at the cap counts its progress as work, not as a result."""
    import serve
    reqs = W.serving_stream(rate=20.0, duration=window, seed=2, sides=(64, 128, 256, 512), degrees=(1, 2, 3, 4),
                            t_frac=(0.1, 1.0), horizon={"ot2pi": 0.5, "rotor": 0.5, "blast": 0.5})

    def say(msg):
        print("# " + msg, file=sys.stderr, flush=True)

    say(f"config 5: {len(reqs)} requests over {window} s, cap {cap} s per server")
    rep = {}
    done, wall, rejected = serve.run_stream(reqs, budget, use_graph=True, wall_cap=cap, progress=say, report=rep)
    lat = [r["t_done"] - r["arrival"] for r in done]
    work_done = sum(serve.work(r) for r in done)
    work_run = sum(W.nodes(r) * r["stats"]["stages"] for r in rep["running"])
    # work the admitted requests need in all: exact for the finished ones,
    # estimated from the last dt for those still running
    need_run = sum(W.nodes(r) * 3 * (r["t_final"] - r["stats"]["t"]) / max(r["stats"]["dt_last"], 1e-300)
                   for r in rep["running"])
    batched = {"finished": len(done), "running_at_cap": len(rep["running"]), "queued_at_cap": len(rep["queued"]),
               "not_yet_arrived_at_cap": len(rep["unsubmitted"]), "rejected": len(rejected),
               "admitted": rep["admitted"], "budget_high_water_GB": rep["high_water_bytes"] / 1e9,
               "rounds": rep["rounds"], "wall_s": wall,
               "served_dof_stage_updates": work_done + work_run,
               "value": (work_done + work_run) / wall,
               "admitted_work_remaining_est": need_run,
               **({"p50": serve.percentiles(lat)["p50"], "p99": serve.percentiles(lat)["p99"]} if lat else {})}
    say(f"config 5 batched: {batched['finished']} finished, {batched['value']:.3e} DOF-stage/s")
    seq, busy = serve.run_sequential(reqs, wall_cap=cap, progress=say)
    fin = [r for r in seq if "t_done" in r]
    lat_seq = [r["t_done"] - r["arrival"] for r in fin]
    work_seq = sum(W.nodes(r) * r["stats"]["stages"] for r in seq)
    plain = {"finished": len(fin), "started": len(seq), "busy_s": busy, "served_dof_stage_updates": work_seq,
             "value": work_seq / max(busy, 1e-9),
             **({"p50": serve.percentiles(lat_seq)["p50"], "p99": serve.percentiles(lat_seq)["p99"]} if lat_seq else {})}
    say(f"config 5 plain: {plain['finished']} finished, {plain['value']:.3e} DOF-stage/s")
    return {"stream": f"Poisson 20/s (seed 2), arrivals over the first {window} s of the 120 s stream, sides "
                      f"64/128/256/512, P1-P4, t_final uniform in [0.05, 0.5], 200 GB budget, {cap} s wall cap per server",
            "requests": len(reqs), "requested_GB": sum(ppes_bytes(r) for r in reqs) / 1e9,
            "batched": batched, "plain_sequential": plain,
            "batched_over_plain": batched["value"] / max(plain["value"], 1e-300),
            "unit": "DOF-stage updates/s over the capped run (progress of unfinished requests counted); "
                    "latency (arrival to result) in s, finished requests only"}


def ppes_bytes(r):
    import ppesdg
    return ppesdg.instance_bytes(r["nx"], r["ny"], r["p"], 3)


def load_traffic(kernel):
    """HBM bytes per launch from the committed rocprofv3 PMC summary, if any."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_mixed64_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("bytes_per_launch")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--timing-steps", type=int, default=10)
    ap.add_argument("--instances", type=int, default=64)
    ap.add_argument("--single-steps", type=int, default=200)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-single", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-duration", type=float, default=5.0)
    ap.add_argument("--serve-rate", type=float, default=40.0)
    ap.add_argument("--serve-sides", default="32,64,128,256")
    ap.add_argument("--serve-config5", action="store_true",
                    help="run only the serving stream at its stated parameters (bounded by the two options below)")
    ap.add_argument("--serve5-window", type=float, default=30.0)
    ap.add_argument("--serve5-cap", type=float, default=270.0)
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
              "(a batch would shard across GPUs by whole instances, with no communication)", file=sys.stderr)
        sys.exit(2)

    import ppesdg  # raises if the HIP library is missing: there is no fallback

    if a.serve_config5:
        print(json.dumps({"serving_config5": run_serving_config5(a.serve5_window, a.serve5_cap)}))
        return

    mix = W.mixed_batch(a.instances, seed=1)
    nodes_total = sum(W.nodes(m) for m in mix)
    print(f"# batched: {len(mix)} instances, {nodes_total} nodes", file=sys.stderr, flush=True)
    batch, ids, elapsed, kt, stats, tuned = run_batched(ppesdg, mix, a.steps, a.warmup, a.timing_steps)
    value = nodes_total * 3 * a.steps / elapsed
    retries = sum(s["retries"] for s in stats)
    failed = [s for s in stats if s["status"] != 0]

    # roofline of the dominant kernel class over the timing steps
    dominant = max((k for k in kt if kt[k][1] > 0), key=lambda k: kt[k][0])
    tot_ms, launches = kt[dominant]
    per_step_launches = launches_per_step(dominant, mix)
    alg_bytes_step = ALG_BYTES_PER_NODE[dominant] * nodes_total * (3 if dominant in ("residual", "stage") else 1)
    bytes_per_launch = alg_bytes_step / per_step_launches
    avg_s = tot_ms / launches / 1e3
    achieved = bytes_per_launch / avg_s / 1e9
    # the same kernel's algorithmic FP64 rate (residual: per-degree counts)
    resid_flops_step = 3 * sum(W.nodes(m) * ALG_FLOPS_PER_NODE[m["p"]]["residual"] for m in mix)
    resid_s_step = kt["residual"][0] / a.timing_steps / 1e3
    stage_flops = sum(W.nodes(m) * ALG_FLOPS_PER_NODE[m["p"]]["node_stage"] for m in mix) / nodes_total
    shapes = {f"P{m['p']}x{m['p']}" for m in mix}

    def present(d):
        return {k: v for k, v in d.items() if k in shapes}

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
        "config": {"workload": f"{len(mix)} uneven instances (numpy default_rng(1): Orszag-Tang 2pi / rotor / blast, "
                               f"sides 32-256, P1-P4), SSP-RK3, LDF + PP every stage, OEDG final stage (blast: every "
                               f"stage), per-instance CFL dt; batched path, hipGraph step",
                   "instances": len(mix), "nodes_total": nodes_total, "dof_per_node": 8, "path": "batched",
                   "retries": retries, "failed": len(failed), "residual_workgroup_lanes": tuned,
                   "oedg_face_workgroup_lanes": batch.face_tb(),
                   "stage_workgroup_lanes": present(batch.tiles("stage")),
                   "limit_workgroup_lanes": present(batch.tiles("limit"))},
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": load_traffic(dominant),
                     "avg_launch_ms": avg_s * 1e3, "alg_bytes_per_launch": bytes_per_launch,
                     "timing_steps": a.timing_steps},
        "roofline_fp64": {"kernel": "residual", "alg_flops_per_step": resid_flops_step,
                          "achieved": resid_flops_step / resid_s_step / 1e12, "peak": FP64_PEAK_TFLOPS,
                          "unit": "TFLOP/s", "frac": resid_flops_step / resid_s_step / 1e12 / FP64_PEAK_TFLOPS},
        "step_roofline": {"alg_bytes_per_node_stage": STEP_BYTES_PER_NODE_STAGE,
                          "achieved_GBs": STEP_BYTES_PER_NODE_STAGE * value / 1e9,
                          "frac_hbm": STEP_BYTES_PER_NODE_STAGE * value / 1e9 / HBM_PEAK_GBS,
                          "alg_flops_per_node_stage": stage_flops,
                          "achieved_TFLOPs": stage_flops * value / 1e12,
                          "frac_fp64": stage_flops * value / 1e12 / FP64_PEAK_TFLOPS},
        "kernel_ms_per_step": {k: v[0] / a.timing_steps for k, v in kt.items()},
    }
    print(f"# batched {value:.4e} DOF-stage/s", file=sys.stderr, flush=True)
    if not a.no_plain:
        p_elapsed, max_l1, bitwise = run_plain(ppesdg, mix, a.steps, a.warmup, a.timing_steps, batch, ids)
        pv = nodes_total * 3 * a.steps / p_elapsed
        out["plain"] = {"value": pv, "ms_per_step": p_elapsed / a.steps * 1e3,
                        "batched_over_plain": value / pv}
        out["matched_l1"] = {"max_rel_l1_rho_batched_vs_plain": max_l1, "batched_plain_bitwise_equal": bitwise}
        print(f"# plain {pv:.4e} DOF-stage/s", file=sys.stderr, flush=True)
    batch.close()
    if not a.no_single:
        out["single_instance"] = run_single(ppesdg, a.single_steps, 2)
        print("# single instance done", file=sys.stderr, flush=True)
    if not a.no_serve:
        out["serving"] = run_serving(a.serve_duration, a.serve_rate, tuple(int(x) for x in a.serve_sides.split(",")))
        print("# serving done", file=sys.stderr, flush=True)
    if not a.no_cpu:
        out["cpu_baseline"], finals = cpu_baseline(mix)
        # the headline match: every sampled solve against the FP64 oracle
        out.setdefault("matched_l1", {})
        out["matched_l1"]["max_rel_l1_rho_vs_oracle"] = match_oracle(ppesdg, finals)
        out["matched_l1"]["oracle_sample"] = f"the {len(finals)} cpu_baseline instances, 6 RK3 steps, batched path"
    print(json.dumps(out))


if __name__ == "__main__":
    main()
