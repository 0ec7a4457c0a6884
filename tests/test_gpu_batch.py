"""Batched instances (ppes_batch_*): every instance run inside a batch, with
or without a captured HIP graph, and admitted mid-run, gives bit for bit the
result of the same instance run alone through the single-solver API, and
matches the CPU oracle within the trajectory tolerance (1e-10 normwise)."""
import numpy as np
import pytest

import oracle as O
import ppesdg
from helpers import interior, normwise, rel_l1, setup

pytestmark = pytest.mark.gpu

# (problem, nx, ny, p, run parameters): uneven sizes, degrees, BCs, OEDG
# modes and end conditions
MIX = [
    ("ot2pi", 10, 9, 1, dict(max_steps=5, cfl=0.2, oedg_mode=1, use_ldf=True)),
    ("rotor", 12, 12, 2, dict(max_steps=7, cfl=0.2, oedg_mode=1, use_ldf=True)),
    ("blast", 10, 10, 3, dict(max_steps=4, cfl=0.15, oedg_mode=2, use_ldf=True)),
    ("ot2pi", 8, 8, 4, dict(t_final=0.02, cfl=0.2, oedg_mode=1, use_ldf=True)),
    ("rotor", 9, 11, 3, dict(max_steps=3, cfl=0.2, oedg_mode=0, use_ldf=True)),
    ("ot2pi", 12, 10, 2, dict(max_steps=6, cfl=0.2, oedg_mode=1, use_ldf=False)),
    ("blast", 7, 9, 1, dict(t_final=2e-4, cfl=0.15, oedg_mode=2, use_ldf=True)),
]


def run_alone(prob, nx, ny, p, kw):
    s = setup(prob)
    g = ppesdg.Solver(nx, ny, p, s["domain"], s["gamma"], s["bc"], s["bc_state"])
    g.init_problem(prob)
    st = g.solve(kw.get("t_final", -1.0), kw.get("max_steps", 0), kw["cfl"], kw["oedg_mode"], kw["use_ldf"])
    return interior(g.download_state(0)), st


def add_and_start(b, prob, nx, ny, p, kw):
    s = setup(prob)
    i = b.add(nx, ny, p, s["domain"], s["gamma"], s["bc"], s["bc_state"])
    b.init_problem(i, prob)
    b.start(i, kw.get("t_final", -1.0), kw.get("max_steps", 0), kw["cfl"], kw["oedg_mode"], kw["use_ldf"])
    return i


def drain(b, use_graph):
    for _ in range(10000):
        _, act = b.step(1, use_graph)
        if act == 0:
            return
    raise AssertionError("batch did not finish")


def assert_same(b, i, ref_u, ref_st):
    st = b.status(i)
    assert st["state"] == "done", st
    assert st["steps"] == ref_st["steps"] and st["retries"] == ref_st["retries"]
    assert st["t"] == ref_st["t"] and st["negcells"] == ref_st["negcells"]
    assert st["tot_ent"] == ref_st["tot_ent"]
    got = interior(b.download_state(i, 0))
    assert np.array_equal(got, ref_u)


@pytest.fixture(scope="module")
def alone():
    return [run_alone(*case) for case in MIX]


@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_equals_alone(alone, use_graph):
    b = ppesdg.Batch(3)
    ids = [add_and_start(b, *case) for case in MIX]
    drain(b, use_graph)
    for i, (u, st) in zip(ids, alone):
        assert_same(b, i, u, st)


def test_batch_admission_mid_run(alone):
    """Continuous batching: instances join while others are running, and a
    finished instance's slot is reused."""
    b = ppesdg.Batch(3)
    first = [add_and_start(b, *MIX[k]) for k in (0, 1, 2)]
    b.step(2, True)
    late = [add_and_start(b, *MIX[k]) for k in (3, 4)]
    b.step(3, True)
    drain(b, True)
    for i, k in zip(first + late, (0, 1, 2, 3, 4)):
        assert_same(b, i, *alone[k])
    b.remove(first[0])
    again = [add_and_start(b, *MIX[k]) for k in (5, 6)]
    drain(b, True)
    for i, k in zip(again, (5, 6)):
        assert_same(b, i, *alone[k])


def test_batch_matches_oracle_and_positivity():
    b = ppesdg.Batch(3)
    ids = [add_and_start(b, *case) for case in MIX]
    drain(b, True)
    for i, (prob, nx, ny, p, kw) in zip(ids, MIX):
        s = setup(prob)
        o = O.Oracle(nx, ny, p, s["domain"], s["gamma"], s["bc"], s["bc_state"])
        o.init_problem(prob)
        so = o.run(kw.get("t_final", -1.0), kw.get("max_steps", 0), kw["cfl"], kw["oedg_mode"], kw["use_ldf"])
        got = interior(b.download_state(i, 0))
        err = normwise(got, interior(o.u))
        print(f"batch {prob} {nx}x{ny} P{p}: normwise {err:.3e}, steps {b.status(i)['steps']}/{so['steps']}")
        assert b.status(i)["steps"] == so["steps"]
        assert err <= 1e-10
        q = got
        prs = (s["gamma"] - 1.0) * (q[7] - 0.5 * ((q[1] ** 2 + q[2] ** 2 + q[3] ** 2) / q[0] +
                                                   (q[4] ** 2 + q[5] ** 2 + q[6] ** 2)))
        assert q[0].min() > 0 and prs.min() > 0


def test_batch_memory_accounting():
    b = ppesdg.Batch(3)
    i = b.add(16, 12, 3, (0.0, 1.0, 0.0, 1.0), 1.4)
    assert b.memory(i) == ppesdg.instance_bytes(16, 12, 3)


def test_ent_offset_instance_alone_equals_in_mix():
    """An instance with ent_offset on gives the same bits alone and in a batch
    of instances without it."""
    s = setup("rotor")
    alone = ppesdg.Batch(3)
    i = alone.add(12, 10, 2, s["domain"], s["gamma"], s["bc"], s["bc_state"])
    alone.init_problem(i, "rotor")
    alone.start(i, -1.0, 5, 0.6, 1, True, ent_offset=True)
    alone.step(0)
    ref = alone.download_state(i, 0)
    mix = ppesdg.Batch(3)
    ids = []
    for k, p in enumerate([3, 2, 1]):
        j = mix.add(12 + k, 10, p, s["domain"], s["gamma"], s["bc"], s["bc_state"])
        mix.init_problem(j, "rotor")
        ids.append(j)
    j = mix.add(12, 10, 2, s["domain"], s["gamma"], s["bc"], s["bc_state"])
    mix.init_problem(j, "rotor")
    for q in ids:
        mix.start(q, -1.0, 5, 0.6, 1, True)
    mix.start(j, -1.0, 5, 0.6, 1, True, ent_offset=True)
    mix.step(0)
    assert np.array_equal(mix.download_state(j, 0), ref)


def test_degree_5_and_6_instances_alone_equal_in_mix():
    """P5 and P6 instances (36 and 49 nodes per element, the default
    workgroup sizes only) give the same bits alone and in a batch with every
    other degree, and match the oracle after the same steps."""
    import oracle as O
    s = setup("ot2pi")
    refs = {}
    for p in (5, 6):
        alone = ppesdg.Batch(3)
        i = alone.add(9, 8, p, s["domain"], s["gamma"], s["bc"], s["bc_state"])
        alone.init_problem(i, "ot2pi")
        alone.start(i, -1.0, 4, 0.2, 1, True)
        alone.step(0)
        refs[p] = alone.download_state(i, 0)
        alone.close()
        o = O.Oracle(9, 8, p, s["domain"], s["gamma"], s["bc"], s["bc_state"])
        o.init_problem("ot2pi")
        o.run(-1.0, 4, 0.2, 1, True)
        # P6 with OEDG: the conditioning of the order-6 jumps (test_oedg_conditioning_cpu.py)
        assert normwise(interior(refs[p]), interior(o.u)) <= (1e-10 if p == 5 else 1e-7)
        assert rel_l1(interior(refs[p])[0], interior(o.u)[0]) <= 1e-10
    mix = ppesdg.Batch(3)
    ids = {}
    for p in (1, 2, 3, 4, 5, 6):
        k = mix.add(9, 8, p, s["domain"], s["gamma"], s["bc"], s["bc_state"])
        mix.init_problem(k, "ot2pi")
        mix.start(k, -1.0, 4, 0.2, 1, True)
        ids[p] = k
    mix.step(1)
    mix.autotune(2)
    mix.step(0)
    for p in (5, 6):
        assert np.array_equal(mix.download_state(ids[p], 0), refs[p])
    mix.close()
