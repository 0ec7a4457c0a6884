"""The autotuner (ppes_batch_autotune / ppes_autotune): it picks a residual
and an OEDG-face workgroup size from the candidates for every degree present,
and the solves continue to the same bits as without it (an element's
arithmetic does not depend on how many elements share its workgroup, nor a
face's on how many faces share it)."""
import numpy as np
import pytest

import ppesdg
from helpers import interior, setup

pytestmark = pytest.mark.gpu

MIX = [("ot2pi", 12, 10, 1, 0.2, 1), ("rotor", 10, 12, 2, 0.2, 1), ("blast", 10, 10, 3, 0.15, 2),
       ("ot2pi", 8, 9, 4, 0.2, 1)]


def run_batch(tune):
    b = ppesdg.Batch(3)
    ids = []
    for prob, nx, ny, p, cfl, oedg in MIX:
        s = setup(prob)
        i = b.add(nx, ny, p, s["domain"], s["gamma"], s["bc"], s["bc_state"])
        b.init_problem(i, prob)
        b.start(i, -1.0, 8, cfl, oedg, True)
        ids.append(i)
    b.step(3, True)
    tuned = b.autotune(3) if tune else {}
    faces = b.face_tb()
    b.step(5, True)
    out = [interior(b.download_state(i, 0)) for i in ids]
    b.close()
    return out, tuned, faces


def test_batch_autotune_keeps_results():
    ref, _, faces0 = run_batch(False)
    got, tuned, faces = run_batch(True)
    print("tuned residual workgroups:", tuned, "OEDG face workgroups:", faces)
    assert faces0 == {}
    assert set(tuned) == {"P1x1", "P2x2", "P3x3", "P4x4"} and set(faces) == set(tuned)
    assert all(v in (64, 128, 256) for v in list(tuned.values()) + list(faces.values()))
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)


def test_solver_autotune_keeps_results():
    s = setup("ot2pi")
    out = []
    for tune in (False, True):
        g = ppesdg.Solver(16, 12, 3, s["domain"], s["gamma"], s["bc"], s["bc_state"])
        g.init_problem("ot2pi")
        g.solve_begin(True)
        g.advance(-1.0, 2, 0.2, 1)
        if tune:
            assert set(g.autotune(2)) == {"P3x3"} and set(g.face_tb()) == {"P3x3"}
        g.advance(-1.0, 3, 0.2, 1)
        out.append(interior(g.download_state(0)))
        g.close()
    assert np.array_equal(out[0], out[1])


def test_exact_mode_is_not_tuned():
    s = setup("ot2pi")
    g = ppesdg.Solver(8, 8, 2, s["domain"], s["gamma"], s["bc"], s["bc_state"])
    g.set_arith("exact")
    g.init_problem("ot2pi")
    g.solve_begin(True)
    g.advance(-1.0, 1, 0.2, 1)
    assert g.autotune(2) == {}
    g.close()
