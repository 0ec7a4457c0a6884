"""Summaries of rocprofv3 CSV output for the kernels of this library.

    python pp-es-dg-mhd_amd/profsum.py stats DIR [--title T] [--command C]
        per-kernel table of DIR/*kernel_stats.csv (--kernel-trace --stats)
    python pp-es-dg-mhd_amd/profsum.py traffic --fetch D1 --write D2 --copy-fetch D3 --copy-write D4
                                              --kernel residual [--command C]
        HBM bytes per launch of one kernel class in the mixed 64-instance
        workload, from FETCH_SIZE / WRITE_SIZE (--pmc, KiB per dispatch), each
        scaled by what the same counter reads for a plain 1 GiB device copy
        (bench.py --calibrate-copy); prints a JSON record and a Markdown table
    python pp-es-dg-mhd_amd/profsum.py fp64 --pmc D1 --trace D2 --kernel residual [--command C]
        FP64 operations per launch of one kernel class in the mixed workload:
        executed, from the SQ_INSTS_VALU_{ADD,MUL,FMA,TRANS}_F64 counters of a
        --pmc run (wave instructions x 64 lanes, an FMA two operations) and
        from SQ_INSTS_VALU_FLOPS_FP64, and algorithmic (bench.py's counts per
        node), with the rates over the launch durations of a --kernel-trace run
    python pp-es-dg-mhd_amd/profsum.py counters --pmc D [--title T] [--command C]
        per-launch averages of every counter of a --pmc run, per kernel

Launches in which every workgroup exits at once (the retry residual of a step
with no retry) are left out: their counters are under 2% of the largest
launch of the same kernel and grid.
"""
import argparse
import collections
import csv
import glob
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

CLASS_OF = {"k_residual": "residual", "k_stage_pp": "stage", "k_oedg_faces": "oedg", "k_oedg_limit": "limit",
            "k_reduce_global": "reduce", "k_bc": "bc"}
# algorithmic bytes per node and launch, means over the three SSP-RK3
# launches (the table of bench.py: the launches of L(u1), L(u2) also read u^n
# and store the next stage's combined state, which stages 2 and 3 read alone)
ALG_BYTES_PER_NODE = {"residual": (128.0 + 192.0 + 192.0) / 3.0, "stage": (192.0 + 128.0 + 128.0) / 3.0, "oedg": 64.0,
                      "limit": 128.0}
LAUNCHES_PER_STEP = {"residual": 3, "stage": 3}
# launches of a step that read about 64 and about 128 B per node: (launches
# of each per SSP-RK3 step, algorithmic bytes per node of each); the residual
# of u^n stores L(u^n), those of u1 and u2 the combined state; stage 1 reads
# u^n and L(u^n), stages 2 and 3 the combined state
STEP_KINDS = {"residual": ((1, 2), (128.0, 192.0)), "stage": ((2, 1), (128.0, 192.0))}


def concurrent_launches(tr):
    """indices of the launches of a kernel trace that ran beside another
    kernel (overlap over 10% of the shorter of the two; back-to-back launches
    of one stream touch at their ends)"""
    spans = sorted((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), i) for i, r in enumerate(tr))
    out = set()
    reach, reach_0, reach_i = -1, -1, -1
    for a0, a1, i in spans:
        if a0 < reach and min(a1, reach) - a0 > 0.1 * min(a1 - a0, reach - reach_0):
            out.add(i)
            out.add(reach_i)
        if a1 > reach:
            reach, reach_0, reach_i = a1, a0, i
    return out


def one_csv(d, pattern):
    files = glob.glob(os.path.join(d, "**", pattern), recursive=True)
    if not files:
        raise SystemExit(f"no {pattern} under {d}")
    return files[0]


def stats(d, title, command):
    rows = list(csv.DictReader(open(one_csv(d, "*kernel_stats.csv"))))
    out = [f"# {title}", "", f"Command: `{command}`", "",
           "| kernel | calls | avg (us) | min (us) | max (us) | total (ms) | % |", "|---|---|---|---|---|---|---|"]
    for r in rows:
        name = r["Name"].split("(")[0].replace("void ", "")
        out.append(f"| {name} | {r['Calls']} | {float(r['AverageNs']) / 1e3:.1f} | {float(r['MinNs']) / 1e3:.1f} | "
                   f"{float(r['MaxNs']) / 1e3:.1f} | {float(r['TotalDurationNs']) / 1e6:.2f} | "
                   f"{float(r['Percentage']):.1f} |")
    # per-class totals (all degrees): average launch duration as bench.py reports it
    cls = collections.defaultdict(lambda: [0.0, 0])
    for r in rows:
        base = r["Name"].split("<")[0].split("(")[0].replace("void ppes::", "").replace("ppes::", "")
        if base in CLASS_OF:
            cls[CLASS_OF[base]][0] += float(r["TotalDurationNs"])
            cls[CLASS_OF[base]][1] += int(r["Calls"])
    out += ["", "All launches of a class (start-up and single-instance launches, and retry launches whose",
            "workgroups exit at once, included):", "",
            "| class | launches | total (ms) | avg per launch (us) |", "|---|---|---|---|"]
    for k, (t, n) in cls.items():
        out.append(f"| {k} | {n} | {t / 1e6:.2f} | {t / n / 1e3:.1f} |")
    # full launches only: the largest grid of each kernel (one launch covers a
    # whole degree group) and not an immediate exit (over 5% of the longest),
    # and none that ran beside another kernel (the batched graph runs the
    # degree groups on parallel streams; bench.py times its kernels in the
    # sequential form)
    tr = list(csv.DictReader(open(one_csv(d, "*kernel_trace.csv"))))
    overlapped = concurrent_launches(tr)
    per = collections.defaultdict(list)
    for i, r in enumerate(tr):
        if i in overlapped:
            continue
        name = r["Kernel_Name"].split("(")[0].replace("void ", "")
        dur = int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
        per[name].append((int(r["Grid_Size_X"]) * int(r["Grid_Size_Y"]) * int(r["Grid_Size_Z"]), dur))
    out += ["", f"Full launches (largest grid of each kernel, duration over 5% of the longest, not overlapping",
            f"another kernel: {len(overlapped)} of {len(tr)} launches ran concurrently and are left out): the",
            "launches bench.py times; the class average is the mean over the degree groups, as each group",
            "has the same number of launches per step.", "",
            "| kernel | full launches | avg (us) |", "|---|---|---|"]
    # per (class, degree): the variant with the most full launches is the one
    # the steps run (the others are the autotuner's trial sizes)
    best = {}
    for name, v in sorted(per.items()):
        base = name.split("<")[0].replace("ppes::", "")
        if base not in CLASS_OF:
            continue
        g = max(x[0] for x in v)
        dmax = max(x[1] for x in v if x[0] == g)
        full = [x[1] for x in v if x[0] == g and x[1] > 0.05 * dmax]
        avg = sum(full) / len(full)
        out.append(f"| {name} | {len(full)} | {avg / 1e3:.1f} |")
        key = (CLASS_OF[base], name.split("<")[1].split(",")[0].rstrip(">") if "<" in name else "")
        if key not in best or len(full) > best[key][0]:
            best[key] = (len(full), avg)
    cls_full = collections.defaultdict(list)
    for (cls, _), (_, avg) in sorted(best.items()):
        cls_full[cls].append(avg)
    out += ["", "| class | avg per full launch over degree groups (us) |", "|---|---|"]
    for k, v in cls_full.items():
        out.append(f"| {k} | {sum(v) / len(v) / 1e3:.1f} |")
    return "\n".join(out) + "\n"


def load_counter(d):
    rows = list(csv.DictReader(open(one_csv(d, "*counter_collection.csv"))))
    per = collections.defaultdict(list)
    for r in rows:
        per[(r["Kernel_Name"].split("(")[0].replace("void ", ""), int(r["Grid_Size"]))].append(float(r["Counter_Value"]))
    return per


def copy_factor(d):
    per = load_counter(d)
    vals = [v for (name, _), vs in per.items() if "copyBuffer" in name for v in vs]
    return (1 << 30) / (max(vals) * 1024.0)


def traffic(a):
    import workloads as W
    mix = W.mixed_batch(64, seed=1)
    nodes_p = collections.Counter()
    for m in mix:
        nodes_p[m["p"]] += W.nodes(m)
    ff, fw = copy_factor(a.copy_fetch), copy_factor(a.copy_write)
    F, Wr = load_counter(a.fetch), load_counter(a.write)
    kname = {v: k for k, v in CLASS_OF.items()}[a.kernel]
    per_degree = {}
    step_read = step_write = 0.0
    for p in (1, 2, 3, 4):
        keys = [k for k in F if k[0] == f"ppes::{kname}<{p}>" or k[0].startswith(f"ppes::{kname}<{p},")]
        if not keys:
            continue
        # the variant the steps run: the most dispatches (the others are
        # autotuner trials), in each run on its own (their autotuners time
        # separately); its full launches (largest grid, counter over 2% of
        # the largest)
        def full(C):
            keys = [k for k in C if k[0] == f"ppes::{kname}<{p}>" or k[0].startswith(f"ppes::{kname}<{p},")]
            name = max({k[0] for k in keys}, key=lambda n: sum(len(C[k]) for k in keys if k[0] == n))
            v = C[(name, max(k[1] for k in keys if k[0] == name))]
            return [x * 1024.0 for x in v if x > 0.02 * max(v)]
        f = [x * ff for x in full(F)]
        w = [x * fw for x in full(Wr)]
        # every launch writes one state stream (64 B per node)
        wb = sum(w) / len(w)
        rec_p = {"launches_counted": len(f)}
        if a.kernel in STEP_KINDS:
            # launches of the two kinds of a step, told apart by the bytes they
            # read per node (about 64 or 128), weighted as a step runs them
            (w_lo, w_hi), (alg_lo, alg_hi) = STEP_KINDS[a.kernel]
            lo = [x for x in f if x < 96.0 * nodes_p[p]]
            hi = [x for x in f if x >= 96.0 * nodes_p[p]]
            kinds = [(k, v, wt, al) for k, v, wt, al in (("lo", lo, w_lo, alg_lo), ("hi", hi, w_hi, alg_hi)) if v]
            tw = sum(wt for _, _, wt, _ in kinds)
            rb = sum(wt * sum(v) / len(v) for _, v, wt, _ in kinds) / tw
            alg = sum(wt * al for _, _, wt, al in kinds) / tw * nodes_p[p]
            for k, v, _, al in kinds:
                rec_p[k] = {"launches": len(v), "read_bytes": sum(v) / len(v), "alg_bytes": al * nodes_p[p]}
        else:
            rb = sum(f) / len(f)
            alg = ALG_BYTES_PER_NODE[a.kernel] * nodes_p[p]
        per_degree[p] = dict({"read_bytes": rb, "write_bytes": wb, "alg_bytes": alg,
                              "traffic_over_alg": (rb + wb) / alg}, **rec_p)
        step_read += rb
        step_write += wb
    groups = len(per_degree)
    bytes_per_launch = (step_read + step_write) / groups
    rec = {"kernel": a.kernel, "workload": "mixed 64 instances (numpy default_rng(1)), one launch per degree group",
           "bytes_per_launch": bytes_per_launch, "read_bytes_per_launch": step_read / groups,
           "write_bytes_per_launch": step_write / groups,
           "alg_bytes_per_launch": sum(v["alg_bytes"] for v in per_degree.values()) / groups,
           "copy_calibration": {"fetch_factor": ff, "write_factor": fw,
                                "note": "bytes = counter KiB x 1024 x factor; factor = 1 GiB / counter of a 1 GiB copy"},
           "per_degree": per_degree, "command": a.command}
    return rec


def full_durations(d):
    """average duration (ns) of the full launches of each kernel in a
    --kernel-trace run (largest grid, over 5% of the longest, not
    overlapping another kernel)"""
    tr = list(csv.DictReader(open(one_csv(d, "*kernel_trace.csv"))))
    overlapped = concurrent_launches(tr)
    per = collections.defaultdict(list)
    for i, r in enumerate(tr):
        if i in overlapped:
            continue
        name = r["Kernel_Name"].split("(")[0].replace("void ", "")
        per[name].append((int(r["Grid_Size_X"]) * int(r["Grid_Size_Y"]) * int(r["Grid_Size_Z"]),
                          int(r["End_Timestamp"]) - int(r["Start_Timestamp"])))
    out = {}
    for name, v in per.items():
        g = max(x[0] for x in v)
        dmax = max(x[1] for x in v if x[0] == g)
        full = [x[1] for x in v if x[0] == g and x[1] > 0.05 * dmax]
        out[name] = sum(full) / len(full)
    return out


def fp64(a):
    import workloads as W
    sys.path.insert(0, os.path.dirname(HERE))
    from bench import ALG_FLOPS_PER_NODE, FP64_PEAK_TFLOPS
    mix = W.mixed_batch(64, seed=1)
    nodes_p = collections.Counter()
    for m in mix:
        nodes_p[m["p"]] += W.nodes(m)
    rows = list(csv.DictReader(open(one_csv(a.pmc, "*counter_collection.csv"))))
    per = collections.defaultdict(lambda: collections.defaultdict(float))
    grid = {}
    for r in rows:
        name = r["Kernel_Name"].split("(")[0].replace("void ", "")
        key = (name, int(r["Grid_Size"]), r["Dispatch_Id"])
        per[key][r["Counter_Name"]] += float(r["Counter_Value"])
        grid[name] = max(grid.get(name, 0), int(r["Grid_Size"]))
    dur = full_durations(a.trace)
    kname = {v: k for k, v in CLASS_OF.items()}[a.kernel]
    res = {}
    for p in (1, 2, 3, 4):
        names = [n for n in grid if n.startswith(f"ppes::{kname}<{p}, {p},")]
        if not names:
            continue
        # the variant the steps run: the most dispatches (the others are autotuner trials)
        name = max(names, key=lambda n: sum(1 for (m, _, _) in per if m == n))
        disp = [c for (n, g, _), c in per.items() if n == name and g == grid[name]]
        vmax = max(c["SQ_INSTS_VALU"] for c in disp)
        disp = [c for c in disp if c["SQ_INSTS_VALU"] > 0.02 * vmax]
        avg = lambda key: sum(c[key] for c in disp) / len(disp)
        executed = 64.0 * (avg("SQ_INSTS_VALU_ADD_F64") + avg("SQ_INSTS_VALU_MUL_F64") + avg("SQ_INSTS_VALU_TRANS_F64")
                           + 2.0 * avg("SQ_INSTS_VALU_FMA_F64"))
        alg = ALG_FLOPS_PER_NODE[p][a.kernel] * nodes_p[p]
        t = dur[name] * 1e-9
        res[p] = {"kernel": name, "launches_counted": len(disp), "avg_launch_us": t * 1e6, "nodes": nodes_p[p],
                  "alg_flops": alg, "executed_flops": executed, "flops_fp64_counter": avg("SQ_INSTS_VALU_FLOPS_FP64"),
                  "valu_insts": avg("SQ_INSTS_VALU"),
                  "alg_TFLOPs": alg / t / 1e12, "executed_TFLOPs": executed / t / 1e12,
                  "alg_frac": alg / t / 1e12 / FP64_PEAK_TFLOPS, "executed_frac": executed / t / 1e12 / FP64_PEAK_TFLOPS}
    tot_alg = sum(v["alg_flops"] for v in res.values())
    tot_exe = sum(v["executed_flops"] for v in res.values())
    tot_t = sum(v["avg_launch_us"] for v in res.values()) * 1e-6
    out = [f"# FP64 operations of the {a.kernel} kernel: mixed 64-instance workload", "",
           f"Command: `{a.command}`", "",
           "Executed operations = 64 x (ADD + MUL + TRANS + 2 FMA) FP64 wave instructions per launch",
           "(SQ_INSTS_VALU_*_F64); algorithmic = the reference algorithm's operation count per node",
           "(oracle/flopcount.cpp) x nodes; rates over the average duration of the same launches in a",
           f"--kernel-trace run; peak {FP64_PEAK_TFLOPS} TFLOP/s.", "",
           "| degree | avg launch (us) | alg. GFLOP | executed GFLOP | alg. TFLOP/s | executed TFLOP/s | "
           "alg. frac | executed frac |", "|---|---|---|---|---|---|---|---|"]
    for p, v in res.items():
        out.append(f"| P{p} | {v['avg_launch_us']:.1f} | {v['alg_flops'] / 1e9:.2f} | {v['executed_flops'] / 1e9:.2f} | "
                   f"{v['alg_TFLOPs']:.2f} | {v['executed_TFLOPs']:.2f} | {v['alg_frac']:.3f} | "
                   f"{v['executed_frac']:.3f} |")
    out.append(f"| all degrees | {tot_t * 1e6:.1f} | {tot_alg / 1e9:.2f} | {tot_exe / 1e9:.2f} | "
               f"{tot_alg / tot_t / 1e12:.2f} | {tot_exe / tot_t / 1e12:.2f} | "
               f"{tot_alg / tot_t / 1e12 / FP64_PEAK_TFLOPS:.3f} | {tot_exe / tot_t / 1e12 / FP64_PEAK_TFLOPS:.3f} |")
    out += ["", "```json", json.dumps({"per_degree": res}, indent=1), "```"]
    return "\n".join(out) + "\n"


def counters(a):
    """average of every counter over the full launches of each kernel class
    and degree in a --pmc run (largest grid, over 2% of the largest
    SQ_INSTS_VALU or first counter)"""
    rows = list(csv.DictReader(open(one_csv(a.pmc, "*counter_collection.csv"))))
    per = collections.defaultdict(lambda: collections.defaultdict(float))
    grid = {}
    for r in rows:
        name = r["Kernel_Name"].split("(")[0].replace("void ", "")
        base = name.split("<")[0].replace("ppes::", "")
        if base not in CLASS_OF:
            continue
        per[(name, int(r["Grid_Size"]), r["Dispatch_Id"])][r["Counter_Name"]] += float(r["Counter_Value"])
        grid[name] = max(grid.get(name, 0), int(r["Grid_Size"]))
    names = sorted(grid)
    cols = sorted({c for v in per.values() for c in v})
    key = "SQ_INSTS_VALU" if "SQ_INSTS_VALU" in cols else cols[0]
    out = [f"# {a.title}", "", f"Command: `{a.command}`", "",
           "Per-launch averages over the full launches (largest grid of each kernel, counter over 2% of its",
           "largest) of a --pmc run.", "", "| kernel | launches | " + " | ".join(cols) + " |",
           "|---|---|" + "---|" * len(cols)]
    for name in names:
        disp = [c for (n, g, _), c in per.items() if n == name and g == grid[name]]
        vmax = max(c.get(key, 0.0) for c in disp)
        disp = [c for c in disp if c.get(key, 0.0) > 0.02 * vmax] or disp
        out.append(f"| {name} | {len(disp)} | " +
                   " | ".join(f"{sum(c.get(k, 0.0) for c in disp) / len(disp):.4g}" for k in cols) + " |")
    return "\n".join(out) + "\n"


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("dir")
    s.add_argument("--title", default="rocprofv3 kernel statistics")
    s.add_argument("--command", default="")
    t = sub.add_parser("traffic")
    for k in ("fetch", "write", "copy_fetch", "copy_write"):
        t.add_argument("--" + k.replace("_", "-"), dest=k, required=True)
    t.add_argument("--kernel", default="residual")
    t.add_argument("--command", default="")
    cn = sub.add_parser("counters")
    cn.add_argument("--pmc", required=True)
    cn.add_argument("--title", default="PMC counters per launch")
    cn.add_argument("--command", default="")
    fl = sub.add_parser("fp64")
    fl.add_argument("--pmc", required=True)
    fl.add_argument("--trace", required=True)
    fl.add_argument("--kernel", default="residual")
    fl.add_argument("--command", default="")
    a = ap.parse_args()
    if a.cmd == "stats":
        sys.stdout.write(stats(a.dir, a.title, a.command))
    elif a.cmd == "fp64":
        sys.stdout.write(fp64(a))
    elif a.cmd == "counters":
        sys.stdout.write(counters(a))
    else:
        print(json.dumps(traffic(a), indent=1))


if __name__ == "__main__":
    main()
