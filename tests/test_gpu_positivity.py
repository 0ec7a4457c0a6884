"""The device's positivity record (ppes_positivity / ppes_batch_positivity).

Density and pressure must be positive at every GLL node after every stage's
post-process (the Zhang-Shu limiter, DG_solver.f90:528-603). The kernels that
store a limited state fold each stage's minimum density and pressure over
the interior nodes into a per-instance record (a wavefront minimum and one
atomic per wavefront). These tests pin that record:

* stage by stage through the per-call API (combine, BC, post-process,
  residual), it equals the running minimum the host computes from every
  post-processed stage it downloads, bit for bit, from a start-up state with
  inadmissible nodes that the limiter has to repair;
* an instance's record is the same alone and inside a batch, and it is
  reset when a solve starts.
"""
import numpy as np
import pytest

import ppesdg
from helpers import setup

pytestmark = pytest.mark.gpu
PP_EPS = 1e-11


def host_min(u, g):
    """min density and pressure over the interior nodes, pressure in the
    device's expression order (_macro_prs_, DG_solver.f90:125)."""
    q = u[:, :, :, 1:-1, 1:-1]
    prs = (g - 1.0) * (q[7] - 0.5 * ((q[1] * q[1] + q[2] * q[2] + q[3] * q[3]) / q[0] + (q[4] * q[4] + q[5] * q[5] + q[6] * q[6])))
    return float(q[0].min()), float(prs.min())


def stress(u, g, seed, n=4):
    """Inadmissible nodes inside admissible cells: n nodes with negative
    density, n with negative pressure (energy lowered by 1.5 p / (g - 1))."""
    rng = np.random.default_rng(seed)
    N, nx, ny = u.shape[1], u.shape[3] - 2, u.shape[4] - 2
    for _ in range(n):
        k, l = rng.integers(0, N, 2)
        u[0, k, l, rng.integers(1, nx + 1), rng.integers(1, ny + 1)] *= -0.05
        k, l = rng.integers(0, N, 2)
        i, j = rng.integers(1, nx + 1), rng.integers(1, ny + 1)
        q = u[:, k, l, i, j]
        prs = (g - 1.0) * (q[7] - 0.5 * ((q[1] ** 2 + q[2] ** 2 + q[3] ** 2) / q[0] + (q[4] ** 2 + q[5] ** 2 + q[6] ** 2)))
        u[7, k, l, i, j] -= 1.5 * prs / (g - 1.0)


@pytest.mark.parametrize("prob,p,oedg_every,cfl", [("blast", 2, True, 0.15), ("blast", 1, True, 0.15),
                                                    ("ot2pi", 3, False, 0.2), ("rotor", 4, False, 0.2)])
def test_record_equals_host_minimum_of_every_stage(prob, p, oedg_every, cfl):
    s = setup(prob)
    g = s["gamma"]
    sol = ppesdg.Solver(10, 9, p, s["domain"], g, s["bc"], s["bc_state"])
    sol.init_problem(prob)
    assert sol.positivity() == (np.inf, np.inf)   # nothing recorded yet
    u = sol.download_state(0)
    stress(u, g, 7)
    sol.upload_state(0, u)
    sol.apply_bc(0)
    r = sol.post_process(0, False, True, 0.0)       # start-up post-process (drive.f90:69)
    assert r["num_negcell"] > 0                     # the limiter acted on the stressed nodes
    run = host_min(sol.download_state(0), g)
    assert sol.positivity() == run
    sol.apply_bc(0)
    limited = 0
    for step in range(4):
        dtl = sol.semidiscretize(1, 0)
        dt = cfl * dtl
        for st in (1, 2, 3):
            sol.rk_combine(st, dt)
            sol.apply_bc(st)
            beta = (1.0, 0.25, 2.0 / 3.0)[st - 1]
            oedg = oedg_every or st == 3
            r = sol.post_process(st, oedg, True, beta * dt if oedg_every else dt)
            assert not r["neg_avg"], (step, st)
            limited += r["num_negcell"]
            sm = host_min(sol.download_state(st), g)
            run = (min(run[0], sm[0]), min(run[1], sm[1]))
            got = sol.positivity()
            print(f"{prob} P{p} step {step} stage {st}: stage min {sm}, record {got}")
            assert got == run, (step, st, got, run)
            sol.apply_bc(st)
            if st < 3:
                sol.semidiscretize(st + 1, st)
        # commit u^(n+1) = uRK(3) (drive.f90:304-311)
        sol.upload_state(0, sol.download_state(3))
    sol.close()
    assert run[0] >= PP_EPS and run[1] > 0.0


def test_record_alone_equals_in_batch_and_resets_on_start():
    cases = [("blast", 16, 14, 2, 0.15, 2), ("rotor", 12, 12, 3, 0.2, 1), ("ot2pi", 9, 11, 1, 0.2, 1)]
    b = ppesdg.Batch(3)
    ids = []
    for prob, nx, ny, p, cfl, oedg in cases:
        s = setup(prob)
        i = b.add(nx, ny, p, s["domain"], s["gamma"], s["bc"], s["bc_state"])
        b.init_problem(i, prob)
        b.start(i, -1.0, 6, cfl, oedg, True)
        ids.append(i)
    b.step(0, True)
    for (prob, nx, ny, p, cfl, oedg), i in zip(cases, ids):
        s = setup(prob)
        sol = ppesdg.Solver(nx, ny, p, s["domain"], s["gamma"], s["bc"], s["bc_state"])
        sol.init_problem(prob)
        sol.solve(-1.0, 6, cfl, oedg, True)
        alone = sol.positivity()
        assert alone == b.positivity(i), (prob, alone, b.positivity(i))
        assert alone[0] > 0.0 and alone[1] > 0.0
        # a new solve starts a new record: one step from the initial data on
        # the used solver records what it records on a fresh one
        sol.init_problem(prob)
        sol.solve(-1.0, 1, cfl, oedg, True)
        reused = sol.positivity()
        sol.close()
        fresh_sol = ppesdg.Solver(nx, ny, p, s["domain"], s["gamma"], s["bc"], s["bc_state"])
        fresh_sol.init_problem(prob)
        fresh_sol.solve(-1.0, 1, cfl, oedg, True)
        fresh = fresh_sol.positivity()
        fresh_sol.close()
        assert reused == fresh, (prob, reused, fresh)
        if prob == "blast":
            # the blast's density falls over its first steps, so a record that
            # was not reset would differ
            assert fresh[0] > alone[0], (fresh, alone)
    b.close()
