"""Serial CPU timing of the reference against the C oracle (TEST
INFRASTRUCTURE: runs only where the reference is compiled, oracle/_ref).

The CPU-runnable configuration: Orszag-Tang on [0,2pi]^2, P2, 128 x 128,
SSP-RK3, CFL 0.2, OEDG on the final stage, LDF and the positivity limiter on,
to t = 0.5. Both run on one core:
  * the reference's own modules, compiled serially by flang from their
    sources (oracle/Makefile: with -fopenmp flang rejects DG_solver.f90:469),
    driven by oracle/ref_harness.f90 (mode=solve calls solve_execute);
  * oracle/ppes_oracle.c with one OpenMP thread.
Wall time of the whole process / call (set-up is under 1% of it). Prints one
JSON line with both times, the steps, and the oracle-to-reference speed ratio.

    python oracle/time_reference.py
"""
import json
import os
import platform
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "pp-es-dg-mhd_amd"))


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def main():
    import math

    import oracle as O
    harness = os.path.join(HERE, "_ref", "ref_harness")
    if not os.path.exists(harness):
        sys.exit("oracle/_ref/ref_harness is not built (make -C oracle ref)")
    args = dict(mode="solve", prob="ot2pi", nx=128, ny=128, p=2, tfinal=0.5, oedg=1, ldf=1, cfl=0.2)
    with tempfile.TemporaryDirectory() as d:
        cmd = [harness] + [f"{k}={v}" for k, v in args.items()] + [f"out={d}"]
        env = dict(os.environ, OMP_NUM_THREADS="1")
        t0 = time.perf_counter()
        subprocess.run(cmd, check=True, capture_output=True, cwd=d, env=env)
        t_ref = time.perf_counter() - t0
    O.set_threads(1)
    o = O.Oracle(128, 128, 2, (0.0, 2 * math.pi, 0.0, 2 * math.pi), 5.0 / 3.0, "periodic")
    o.init_problem("ot2pi")
    t0 = time.perf_counter()
    s = o.run(0.5, 0, 0.2, 1, True)
    t_orc = time.perf_counter() - t0
    work = 128 * 128 * 9 * (3 * s["steps"] + 1)
    print(json.dumps({"workload": "Orszag-Tang 2pi, P2, 128x128, SSP-RK3, CFL 0.2, t = 0.5", "cores": 1,
                      "cpu": cpu_model(), "steps": s["steps"], "reference_s": t_ref, "oracle_s": t_orc,
                      "reference_dof_stage_per_s": work / t_ref, "oracle_dof_stage_per_s": work / t_orc,
                      "oracle_over_reference_speed": t_ref / t_orc}))


if __name__ == "__main__":
    main()
