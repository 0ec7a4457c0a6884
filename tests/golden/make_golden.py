"""Write the golden fixtures from the reference solver (run in the build
container only: it needs /root/reference compiled by `make -C oracle ref`).

Every fixture is produced by oracle/_ref/ref_harness, which drives the
reference's own modules through their public API (see oracle/ref_harness.f90).
The fixtures are data only: inputs, outputs and the harness arguments.

    python tests/golden/make_golden.py [--long]

--long also runs the full-size Orszag-Tang case (P2, 128x128, t = 0.5),
which takes a few minutes on one core.
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")


def load_dump(d):
    out = {}
    for line in open(os.path.join(d, "manifest.txt")):
        f = line.split()
        if not f:
            continue
        shp = tuple(int(x) for x in f[2:])
        out[f[0]] = np.fromfile(os.path.join(d, f[0] + ".bin"), dtype="<f8").reshape(shp, order="F")
    return out


def run(args, keep=None):
    with tempfile.TemporaryDirectory() as d:
        cmd = [HARNESS] + [f"{k}={v}" for k, v in args.items()] + [f"out={d}"]
        # the reference writes its Tecplot files (entropy.plt, solution_0.plt)
        # into the working directory
        res = subprocess.run(cmd, capture_output=True, text=True, cwd=d)
        if res.returncode != 0:
            raise RuntimeError(f"{cmd} failed:\n{res.stdout}\n{res.stderr}")
        data = load_dump(d)
        data["_stdout"] = res.stdout
        for f in ("solution_0", "entropy"):
            path = os.path.join(d, f + ".plt")
            if args.get("plt") and os.path.exists(path):
                data["plt_" + f] = np.fromfile(path, dtype=np.uint8)
    if keep is not None:
        data = {k: v for k, v in data.items() if any(k == p or k.startswith(p) for p in keep) or k == "_stdout"}
    return data


def save(name, args, data):
    arrays = {k: v for k, v in data.items() if not k.startswith("_")}
    arrays["args_json"] = np.array(json.dumps(args))
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **arrays)
    size = os.path.getsize(os.path.join(HERE, name + ".npz"))
    print(f"{name}.npz  {size / 1024:.0f} KiB  ({len(arrays)} arrays)")


STAGE_CASES = {
    # name: harness arguments (first step dumped stage by stage, then the final state)
    "stages_ot2pi_p2_ldf0": dict(mode="stages", prob="ot2pi", nx=8, ny=6, p=2, nsteps=3, oedg=1, ldf=0, cfl=0.2, dumpsteps=1),
    "stages_ot2pi_p2": dict(mode="stages", prob="ot2pi", nx=8, ny=6, p=2, nsteps=3, oedg=1, ldf=1, cfl=0.2, dumpsteps=1),
    "stages_rotor_p3_ldf0": dict(mode="stages", prob="rotor", nx=7, ny=6, p=3, nsteps=3, oedg=2, ldf=0, cfl=0.15, dumpsteps=1),
    "stages_rotor_p3": dict(mode="stages", prob="rotor", nx=7, ny=6, p=3, nsteps=3, oedg=2, ldf=1, cfl=0.15, dumpsteps=1),
    "stages_blast_p1": dict(mode="stages", prob="blast", nx=9, ny=8, p=1, nsteps=3, oedg=2, ldf=1, cfl=0.15, dumpsteps=1),
    "stages_blast_p2_ldf0": dict(mode="stages", prob="blast", nx=8, ny=7, p=2, nsteps=3, oedg=2, ldf=0, cfl=0.15, dumpsteps=1),
    "stages_blast_p4": dict(mode="stages", prob="blast", nx=5, ny=4, p=4, nsteps=2, oedg=2, ldf=1, cfl=0.15, dumpsteps=1),
    "stages_otunit_p3": dict(mode="stages", prob="otunit", nx=6, ny=6, p=3, nsteps=3, oedg=1, ldf=1, cfl=0.6, dumpsteps=1),
    # the rest of the reference's catalogue (main.f90:57-190) with its boundary fillers
    "stages_briowu_p3_ldf0": dict(mode="stages", prob="briowu", nx=12, ny=2, p=3, nsteps=3, oedg=1, ldf=0, cfl=0.6, dumpsteps=1),
    "stages_cloud_p2_ldf0": dict(mode="stages", prob="cloud", nx=8, ny=8, p=2, nsteps=3, oedg=1, ldf=0, cfl=0.6, dumpsteps=1),
    "stages_jet_p2_ldf0": dict(mode="stages", prob="jet", nx=10, ny=6, p=2, nsteps=3, oedg=1, ldf=0, cfl=0.6, dumpsteps=1),
    "stages_yee_p2_ldf0": dict(mode="stages", prob="yee", nx=8, ny=8, p=2, nsteps=3, oedg=1, ldf=0, cfl=0.6, dumpsteps=1),
    "stages_alfven_p2_ldf0": dict(mode="stages", prob="alfven", nx=6, ny=5, p=2, nsteps=3, oedg=1, ldf=0, cfl=0.6, dumpsteps=1),
    "stages_wave1d_p3_ldf0": dict(mode="stages", prob="wave1d", nx=10, ny=2, p=3, nsteps=3, oedg=1, ldf=0, cfl=0.6, dumpsteps=1),
    "stages_khi_p2": dict(mode="stages", prob="khi", nx=6, ny=8, p=2, nsteps=3, oedg=1, ldf=1, cfl=0.6, dumpsteps=1),
    "stages_jet_p3": dict(mode="stages", prob="jet", nx=10, ny=4, p=3, nsteps=3, oedg=1, ldf=1, cfl=0.6, dumpsteps=1),
    "stages_cloud_p1": dict(mode="stages", prob="cloud", nx=10, ny=10, p=1, nsteps=3, oedg=1, ldf=1, cfl=0.6, dumpsteps=1),
    # the reference's 1D set-ups: degree 0 in y (main.f90:60-81)
    "stages_briowu_p3_py0_ldf0": dict(mode="stages", prob="briowu", nx=12, ny=1, p=3, py=0, nsteps=3, oedg=1, ldf=0, cfl=0.6, dumpsteps=1),
    "stages_wave1d_p3_py0_ldf0": dict(mode="stages", prob="wave1d", nx=10, ny=1, p=3, py=0, nsteps=3, oedg=1, ldf=0, cfl=0.6, dumpsteps=1),
    "stages_briowu_p2_py0": dict(mode="stages", prob="briowu", nx=16, ny=1, p=2, py=0, nsteps=3, oedg=1, ldf=1, cfl=0.6, dumpsteps=1),
    # degrees 5 and 6, the largest of the reference's tables (gauss.f90:296)
    "stages_ot2pi_p5_ldf0": dict(mode="stages", prob="ot2pi", nx=5, ny=4, p=5, nsteps=2, oedg=1, ldf=0, cfl=0.2, dumpsteps=1),
    "stages_rotor_p5": dict(mode="stages", prob="rotor", nx=5, ny=5, p=5, nsteps=2, oedg=1, ldf=1, cfl=0.2, dumpsteps=1),
    "stages_blast_p6": dict(mode="stages", prob="blast", nx=4, ny=4, p=6, nsteps=2, oedg=2, ldf=1, cfl=0.15, dumpsteps=1),
    "stages_blast_p6_ldf0": dict(mode="stages", prob="blast", nx=4, ny=3, p=6, nsteps=2, oedg=2, ldf=0, cfl=0.15, dumpsteps=1),
    "stages_wave1d_p5_py0_ldf0": dict(mode="stages", prob="wave1d", nx=8, ny=1, p=5, py=0, nsteps=3, oedg=1, ldf=0, cfl=0.6, dumpsteps=1),
    "stages_briowu_p6_py0": dict(mode="stages", prob="briowu", nx=10, ny=1, p=6, py=0, nsteps=3, oedg=1, ldf=1, cfl=0.6, dumpsteps=1),
    # ent_offset on (convex limiting against the cell entropy before LDF, DG_solver.f90:195-225,605-633)
    "stages_blast_p2_entoff_ldf0": dict(mode="stages", prob="blast", nx=8, ny=7, p=2, nsteps=3, oedg=2, ldf=0, cfl=0.15, dumpsteps=1, entoff=1),
    "stages_ot2pi_p3_entoff_ldf0": dict(mode="stages", prob="ot2pi", nx=6, ny=6, p=3, nsteps=3, oedg=1, ldf=0, cfl=0.6, dumpsteps=1, entoff=1),
    "stages_rotor_p2_entoff": dict(mode="stages", prob="rotor", nx=8, ny=8, p=2, nsteps=3, oedg=1, ldf=1, cfl=0.6, dumpsteps=1, entoff=1),
    "stages_khi_p3_entoff": dict(mode="stages", prob="khi", nx=6, ny=8, p=3, nsteps=3, oedg=1, ldf=1, cfl=0.6, dumpsteps=1, entoff=1),
}

END_CASES = {
    # reduced blast (config 3 shape: OEDG every stage, extrapolation BCs) to t = 0.002
    "end_blast_p2_40": dict(mode="stages", prob="blast", nx=40, ny=40, p=2, nsteps=0, tfinal=0.002, oedg=2, ldf=1, cfl=0.15, dumpsteps=0),
    # reference solve_execute itself (final-stage OEDG), rotor P2 32x32 to t = 0.01
    "end_rotor_p2_32": dict(mode="solve", prob="rotor", nx=32, ny=32, p=2, tfinal=0.01, oedg=1, ldf=1, cfl=0.2),
    # Orszag-Tang on [0,2pi]^2, P1 and P4 on small grids to t = 0.1
    # (also keeps the reference's Tecplot output of this run: solution_0.plt, entropy.plt)
    "end_ot2pi_p1_32": dict(mode="solve", prob="ot2pi", nx=32, ny=32, p=1, tfinal=0.1, oedg=1, ldf=1, cfl=0.2, plt=1),
    "end_ot2pi_p4_12": dict(mode="solve", prob="ot2pi", nx=12, ny=12, p=4, tfinal=0.1, oedg=1, ldf=1, cfl=0.2),
    "end_ot2pi_p6_8": dict(mode="solve", prob="ot2pi", nx=8, ny=8, p=6, tfinal=0.05, oedg=1, ldf=1, cfl=0.2),
    "end_rotor_p5_10": dict(mode="solve", prob="rotor", nx=10, ny=10, p=5, tfinal=0.005, oedg=1, ldf=1, cfl=0.2),
    # the other RK methods of solve_execute (drive.f90:132-174,213-272): RK1, RK2, RK5 (CFL_RK 1.508)
    "end_ot2pi_p2_16_rk1": dict(mode="solve", prob="ot2pi", nx=16, ny=16, p=2, tfinal=0.05, oedg=1, ldf=1, cfl=0.1, rk=1),
    "end_ot2pi_p2_16_rk2": dict(mode="solve", prob="ot2pi", nx=16, ny=16, p=2, tfinal=0.05, oedg=1, ldf=1, cfl=0.2, rk=2),
    "end_rotor_p3_16_rk5": dict(mode="solve", prob="rotor", nx=16, ny=16, p=3, tfinal=0.005, oedg=1, ldf=1, cfl=0.2, rk=5),
    # the smooth problems with their error functionals (main.f90:71-94), CFL 0.6, LDF off as there
    "end_alfven_p2_8": dict(mode="solve", prob="alfven", nx=8, ny=8, p=2, tfinal=0.05, oedg=1, ldf=0, cfl=0.6),
    "end_wave1d_p3_16_rk5": dict(mode="solve", prob="wave1d", nx=16, ny=2, p=3, tfinal=0.1, oedg=1, ldf=0, cfl=0.6, rk=5),
    # shock problems with their own boundary fillers
    "end_cloud_p2_16": dict(mode="solve", prob="cloud", nx=16, ny=16, p=2, tfinal=0.002, oedg=1, ldf=1, cfl=0.6),
    "end_jet_p2_20": dict(mode="solve", prob="jet", nx=20, ny=20, p=2, tfinal=1e-4, oedg=1, ldf=1, cfl=0.6),
    "end_briowu_p2_32": dict(mode="solve", prob="briowu", nx=32, ny=2, p=2, tfinal=0.02, oedg=1, ldf=1, cfl=0.6),
    "end_ot2pi_p2_16_entoff": dict(mode="solve", prob="ot2pi", nx=16, ny=16, p=2, tfinal=0.1, oedg=1, ldf=1, cfl=0.6, entoff=1),
    # main.f90 cases 0 and 1 in full: Brio-Wu and the smooth density wave, 480 x 1 elements, P3 in x, P0 in
    # y, SSP-RK(5,4), CFL 0.6 (the second with its error functional)
    "end_briowu_main0": dict(mode="solve", prob="briowu", nx=480, ny=1, p=3, py=0, tfinal=0.2, oedg=1, ldf=1, cfl=0.6, rk=5),
    "end_wave1d_main1": dict(mode="solve", prob="wave1d", nx=480, ny=1, p=3, py=0, tfinal=1.3, oedg=1, ldf=0, cfl=0.6, rk=5),
}

LONG_CASES = {
    # the CPU-runnable configuration: OT 2pi, P2, 128x128, CFL 0.2, t = 0.5
    "end_ot2pi_p2_128": dict(mode="solve", prob="ot2pi", nx=128, ny=128, p=2, tfinal=0.5, oedg=1, ldf=1, cfl=0.2),
}


def flux_pairs(seed=5, n=400):
    """Admissible state pairs (conservative) for the face-flux fixture:
    random states of every scale of the shipped problems, equal pairs,
    near-vacuum and strongly magnetised ones, mirrored pairs."""
    rng = np.random.default_rng(seed)
    out = []
    for q in range(n):
        pr = []
        for _ in range(2):
            rho = float(np.exp(rng.uniform(-6, 3)))
            v = rng.normal(0, [1.0, 1.0, 0.3]) * float(np.exp(rng.uniform(-3, 3)))
            b = rng.normal(0, 1.0, 3) * float(np.exp(rng.uniform(-3, 3)))
            p = float(np.exp(rng.uniform(-8, 7)))
            pr.append([rho, rho * v[0], rho * v[1], rho * v[2], b[0], b[1], b[2],
                       p / (5.0 / 3.0 - 1.0) + 0.5 * rho * float(v @ v) + 0.5 * float(b @ b)])
        if q % 10 == 0:
            pr[1] = list(pr[0])                     # equal states
        if q % 10 == 1:
            pr[1] = list(pr[0])
            pr[1][0] *= 1.0 + 1e-7                  # log-mean series branch
        if q % 10 == 2:
            pr[0][7] = 0.5 * (pr[0][1] ** 2 + pr[0][2] ** 2 + pr[0][3] ** 2) / pr[0][0] + \
                0.5 * (pr[0][4] ** 2 + pr[0][5] ** 2 + pr[0][6] ** 2) + 1e-13  # pressure at the floor
        out.append(pr)
    return np.array(out, dtype=np.float64)  # (n, 2, 8)


def run_fluxes():
    w = flux_pairs()
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "pairs.bin")
        np.ascontiguousarray(w).tofile(path)   # Fortran (8, 2, n) = C (n, 2, 8)
        data = run({"mode": "fluxes", "init": path, "gamma": 5.0 / 3.0})
    return {"pairs": w, "gamma": np.array(5.0 / 3.0), "fluxes": data["fluxes"], "_stdout": data["_stdout"]}


def steps_from_stdout(txt):
    for line in txt.splitlines():
        if line.strip().startswith("# time steps"):
            return int(line.split("=")[1])
    return -1


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--long", action="store_true")
    ap.add_argument("--only", default="")
    a = ap.parse_args()
    if not os.path.exists(HARNESS):
        sys.exit("build the reference harness first: make -C oracle ref")
    # tables of gauss.f90 and the LDF projector of LDF.f90, p = 1..6
    if not a.only or a.only == "tables":
        save("tables", {"mode": "tables"}, run({"mode": "tables"}))
    # the four face fluxes of DG_solver.f90 on fixed state pairs
    if not a.only or a.only == "fluxes":
        save("fluxes", {"mode": "fluxes", "gamma": 5.0 / 3.0}, run_fluxes())
    cases = dict(STAGE_CASES)
    cases.update(END_CASES)
    if a.long:
        cases.update(LONG_CASES)
    for name, args in cases.items():
        if a.only and a.only != name:
            continue
        if name.startswith("end_"):
            keep = ["u_final", "avg_final", "summary", "plt_", "grid_", "err"]
        else:
            keep = ["u_init", "avg_init", "ent_init", "ut_init", "dtlim_init", "s1_", "u_final", "avg_final", "summary"]
        data = run(args, keep)
        if args["mode"] == "solve":
            data["steps"] = np.array(steps_from_stdout(data["_stdout"]))
        if name.startswith("end_"):
            # interior density only keeps the fixture small
            data["rho_final"] = data.pop("u_final")[0, :, :, 1:-1, 1:-1]
            avg = data.pop("avg_final")
            data["rho_avg_final"] = avg[0, 1:-1, 1:-1]
            if args.get("plt"):
                data["avg_final"] = avg  # input of the Tecplot writer test
            else:
                data.pop("grid_x", None)
                data.pop("grid_y", None)
        save(name, args, data)


if __name__ == "__main__":
    main()
