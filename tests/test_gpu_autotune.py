"""The autotuner (ppes_batch_autotune / ppes_autotune): it picks a residual,
an OEDG-face, a stage and a limiter workgroup size from the candidates for
every degree present, and the solves continue to the same bits as without it
(an element's arithmetic does not depend on how many elements share its
workgroup, nor a face's on how many faces share it, and the stage and
limiter kernels form their partial sums per fixed chunk of elements). Every
stage and limiter tile, set by hand, gives the same bits too, entropy
included."""
import numpy as np
import pytest

import ppesdg
from helpers import interior, setup

pytestmark = pytest.mark.gpu

MIX = [("ot2pi", 12, 10, 1, 0.2, 1), ("rotor", 10, 12, 2, 0.2, 1), ("blast", 10, 10, 3, 0.15, 2),
       ("ot2pi", 8, 9, 4, 0.2, 1)]


# stage / limiter tiles built per degree (lanes per workgroup)
TILES = {1: (64, 128, 256, 512), 2: (64, 128, 256, 512), 3: (64, 128, 256, 512), 4: (128, 256, 512)}


def run_batch(tune, tiles=None):
    b = ppesdg.Batch(3)
    ids = []
    for prob, nx, ny, p, cfl, oedg in MIX:
        s = setup(prob)
        i = b.add(nx, ny, p, s["domain"], s["gamma"], s["bc"], s["bc_state"])
        b.init_problem(i, prob)
        b.start(i, -1.0, 8, cfl, oedg, True)
        ids.append(i)
    for kernel, by_p in (tiles or {}).items():
        for p, lanes in by_p.items():
            b.set_tile(kernel, p, p, lanes)
    b.step(3, True)
    tuned = b.autotune(3) if tune else {}
    faces = b.face_tb()
    chosen = {k: b.tiles(k) for k in ("stage", "limit")}
    b.step(5, True)
    out = [interior(b.download_state(i, 0)) for i in ids]
    ent = [b.status(i)["tot_ent"] for i in ids]
    b.close()
    return out, tuned, faces, chosen, ent


def test_batch_autotune_keeps_results():
    ref, _, faces0, chosen0, ent0 = run_batch(False)
    got, tuned, faces, chosen, ent = run_batch(True)
    print("tuned residual workgroups:", tuned, "OEDG face workgroups:", faces, "stage / limiter:", chosen)
    assert faces0 == {}
    assert all(chosen0[k][f"P{p}x{p}"] == 256 for k in chosen0 for p in TILES)
    assert set(tuned) == {"P1x1", "P2x2", "P3x3", "P4x4"} and set(faces) == set(tuned)
    assert all(v in (64, 128, 256) for v in list(tuned.values()) + list(faces.values()))
    for k in ("stage", "limit"):
        for p in TILES:
            assert chosen[k][f"P{p}x{p}"] in TILES[p]
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)
    assert ent == ent0


@pytest.mark.parametrize("pick", range(4))
def test_every_stage_and_limiter_tile_keeps_results(pick):
    """each stage tile with a different limiter tile: the same bits as the
    default 256-lane tiles, entropy totals included (partials per chunk)"""
    ref, _, _, _, ent0 = run_batch(False)
    stage = {p: TILES[p][min(pick, len(TILES[p]) - 1)] for p in TILES}
    limit = {p: TILES[p][::-1][min(pick, len(TILES[p]) - 1)] for p in TILES}
    got, _, _, chosen, ent = run_batch(False, {"stage": stage, "limit": limit})
    for p in TILES:
        assert chosen["stage"][f"P{p}x{p}"] == stage[p] and chosen["limit"][f"P{p}x{p}"] == limit[p]
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)
    assert ent == ent0


def test_tile_sizes_not_built_are_refused():
    b = ppesdg.Batch(3)
    try:
        for kernel, p, lanes in (("stage", 4, 64), ("limit", 1, 1024), ("stage", 2, 96), ("residual", 4, 512),
                                 ("oedg", 3, 512)):
            with pytest.raises(ppesdg.PPESError):
                b.set_tile(kernel, p, p, lanes)
    finally:
        b.close()


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
            assert g.tiles("stage") in TILES[3] and g.tiles("limit") in TILES[3]
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


# odd grids (face rows that a workgroup's faces straddle, a one-element-wide
# grid) and every-stage OEDG, whose earlier stages run the face sums over a
# sub-plan of the group
ODD = [("ot2pi", 7, 5, 2, 0.2, 1), ("blast", 13, 3, 4, 0.15, 2), ("rotor", 1, 9, 3, 0.2, 1),
       ("ot2pi", 31, 2, 1, 0.2, 1), ("blast", 9, 11, 2, 0.15, 2)]


def run_odd(tiles):
    b = ppesdg.Batch(3)
    ids = []
    for prob, nx, ny, p, cfl, oedg in ODD:
        s = setup(prob)
        i = b.add(nx, ny, p, s["domain"], s["gamma"], s["bc"], s["bc_state"])
        b.init_problem(i, prob)
        b.start(i, -1.0, 6, cfl, oedg, True)
        ids.append(i)
    for kernel, lanes in tiles.items():
        for p in (1, 2, 3, 4):
            if kernel != "residual" or lanes >= (p + 1) ** 2:
                b.set_tile(kernel, p, p, lanes)
    b.step(0, True)
    out = [(b.download_state(i, 0), b.status(i)) for i in ids]
    b.close()
    return out


@pytest.mark.parametrize("lanes", [64, 128, 256])
def test_every_face_and_residual_workgroup_size_keeps_results(lanes):
    """the OEDG face sums and the residual at 64, 128 and 256 lanes (their
    block maps in XCD order: row-interleaved face blocks, column-strip
    residual blocks) give the bits of the default sizes on odd grids"""
    ref = run_odd({})
    got = run_odd({"oedg": lanes, "residual": lanes})
    for (u0, s0), (u1, s1), m in zip(ref, got, ODD):
        assert np.array_equal(u0, u1), (m, lanes)
        assert s0 == s1, (m, lanes)
