"""Parity of the HIP path (through the C-ABI) with the CPU oracle.

Tolerances (normwise per variable, max|gpu - oracle| / max|oracle|):
  residual, post-processed state      <= 1e-12  (one call)
  trajectories after a few RK steps   <= 1e-10
The device code keeps the reference's evaluation order and is built without
FP contraction, so most differences come only from last-bit differences of
log/exp and from the LDF operator's form (dense nodal vs factored modal).
"""
import numpy as np
import pytest

import oracle as O
import ppesdg
from helpers import interior, normwise, perturbation_table, rel_l1, setup, strips

pytestmark = pytest.mark.gpu

PROBS = ["ot2pi", "rotor", "blast"]
# every degree of the reference's tables (gauss.f90:296)
DEGREES = [1, 2, 3, 4, 5, 6]


def oedg_tol(p, base):
    """At degree 6 the OEDG indicators' order-6 jumps cancel by about five
    orders of magnitude, so a last-bit difference of the state (the device's
    LDF product sums in another order) comes out ~1e5 larger in the damping
    factor. The oracle itself, given a one-ulp change of B_x, moves by 1.6e-8
    normwise over four P6 steps (1.9e-14 at P4): test_oedg_conditioning_cpu.py.
    The tolerance for P6 with OEDG follows that measured conditioning; the
    L1 density checks keep 1e-10."""
    return base if p <= 5 else max(base, 1e-7)


def make_pair(prob, nx, ny, p, rk=3):
    s = setup(prob)
    o = O.Oracle(nx, ny, p, s["domain"], s["gamma"], s["bc"], s["bc_state"], rk=rk)
    g = ppesdg.Solver(nx, ny, p, s["domain"], s["gamma"], s["bc"], s["bc_state"], rk=rk)
    return o, g


def advanced_state(prob, nx, ny, p, nsteps, oedg_mode=1):
    """Oracle state after a few RK3 steps (a non-trivial, limited state)."""
    s = setup(prob)
    o = O.Oracle(nx, ny, p, s["domain"], s["gamma"], s["bc"], s["bc_state"])
    o.init_problem(prob)
    if nsteps:
        o.run(-1.0, nsteps, 0.2, oedg_mode, True)
    u = np.array(o.u, copy=True)
    o.apply_bc(u)
    return u


@pytest.mark.parametrize("p", DEGREES)
@pytest.mark.parametrize("prob", PROBS)
def test_init_problem_matches_oracle(prob, p):
    o, g = make_pair(prob, 7, 5, p)
    o.init_problem(prob)
    g.init_problem(prob)
    assert np.array_equal(interior(g.download_state(0)), interior(o.u))


def test_init_perturbed_ot_matches_oracle():
    pert = perturbation_table(0)
    o, g = make_pair("ot2pi", 9, 8, 3)
    o.init_problem("otpert", pert, 1e-3)
    g.init_problem("otpert", pert, 1e-3)
    assert normwise(interior(g.download_state(0)), interior(o.u)) <= 1e-15


@pytest.mark.parametrize("arith", ["exact", "fast"])
@pytest.mark.parametrize("p", DEGREES)
@pytest.mark.parametrize("prob", PROBS)
def test_residual_matches_oracle(prob, p, arith):
    nx, ny = 9, 7
    u = advanced_state(prob, nx, ny, p, 2)
    o, g = make_pair(prob, nx, ny, p)
    g.set_arith(arith)
    ut_ref = np.zeros(o.shape_ut, order="F")
    dtl_ref = o.semidiscretize(ut_ref, u)
    g.upload_state(0, u)
    dtl = g.semidiscretize(1, 0)
    ut = g.download_residual(1)
    err = normwise(ut, ut_ref)
    print(f"residual {prob} P{p} {arith}: normwise {err:.3e}, dt_lim {dtl} vs {dtl_ref}")
    # the fast form's few-ulp flux differences enter the volume sums through
    # D, whose entries grow like N^2: 1.1e-12 at P6 on the Orszag-Tang state
    assert err <= (1e-12 if arith == "exact" or p <= 4 else 2e-12)
    assert abs(dtl - dtl_ref) <= 1e-14 * dtl_ref


@pytest.mark.parametrize("p", DEGREES)
@pytest.mark.parametrize("prob", ["ot2pi", "blast"])
def test_residual_exact_mode_is_bitwise(prob, p):
    """In the exact mode the residual is the oracle's (and so the
    reference's) bit for bit on these states, where no log/exp last-bit
    difference reaches it."""
    nx, ny = 9, 7
    u = advanced_state(prob, nx, ny, p, 2)
    o, g = make_pair(prob, nx, ny, p)
    g.set_arith("exact")
    ut_ref = np.zeros(o.shape_ut, order="F")
    dtl_ref = o.semidiscretize(ut_ref, u)
    g.upload_state(0, u)
    dtl = g.semidiscretize(1, 0)
    ut = g.download_residual(1)
    print(f"exact residual {prob} P{p}: {np.count_nonzero(ut != ut_ref)} of {ut.size} values differ")
    assert np.array_equal(ut, ut_ref) and dtl == dtl_ref


@pytest.mark.parametrize("p", DEGREES)
@pytest.mark.parametrize("prob", PROBS)
@pytest.mark.parametrize("oedg", [0, 1])
def test_post_process_matches_oracle(prob, p, oedg):
    nx, ny = 8, 9
    u = advanced_state(prob, nx, ny, p, 2)
    o, g = make_pair(prob, nx, ny, p)
    ref = np.array(u, copy=True)
    avg_ref = np.zeros(o.shape_avg, order="F")
    r_ref = o.post_process(ref, avg_ref, oedg, True, 3e-3, True)
    g.upload_state(0, u)
    r = g.post_process(0, oedg, True, 3e-3, True)
    got = g.download_state(0)
    err = normwise(strips(got), strips(ref))
    print(f"post_process {prob} P{p} oedg={oedg}: {err:.3e}, ent {r['tot_ent']} vs {r_ref['tot_ent']}")
    assert not r["neg_avg"] and not r_ref["neg_avg"]
    assert err <= (oedg_tol(p, 1e-12) if oedg else 1e-12)
    assert r["num_negcell"] == r_ref["num_negcell"]
    assert abs(r["tot_ent"] - r_ref["tot_ent"]) <= 1e-12 * abs(r_ref["tot_ent"])
    avg = g.download_avg()
    assert normwise(strips(avg[:, None, None]), strips(avg_ref[:, None, None])) <= 1e-13


@pytest.mark.parametrize("p", [1, 3])
def test_positivity_limiter_matches_oracle(p):
    """Near-vacuum nodes inside admissible cells: the Zhang-Shu scaling fires."""
    nx, ny = 8, 8
    u = advanced_state("ot2pi", nx, ny, p, 1)
    N = p + 1
    rng = np.random.default_rng(5)
    for _ in range(6):
        i, j = rng.integers(1, nx + 1), rng.integers(1, ny + 1)
        k, l = rng.integers(0, N), rng.integers(0, N)
        u[0, k, l, i, j] *= -0.05          # negative density at one node
        i, j = rng.integers(1, nx + 1), rng.integers(1, ny + 1)
        u[7, k, l, i, j] *= 0.2            # energy below kinetic + magnetic
    o, g = make_pair("ot2pi", nx, ny, p)
    o.apply_bc(u)
    ref = np.array(u, copy=True)
    avg_ref = np.zeros(o.shape_avg, order="F")
    r_ref = o.post_process(ref, avg_ref, 1, True, 1e-2, True)
    g.upload_state(0, u)
    r = g.post_process(0, 1, True, 1e-2, True)
    assert r_ref["num_negcell"] > 0
    assert r["num_negcell"] == r_ref["num_negcell"]
    got = g.download_state(0)
    assert normwise(interior(got), interior(ref)) <= 1e-12
    # exact positivity of density and pressure at every node
    q = interior(got)
    prs = (5.0 / 3.0 - 1.0) * (q[7] - 0.5 * ((q[1] ** 2 + q[2] ** 2 + q[3] ** 2) / q[0] + (q[4] ** 2 + q[5] ** 2 + q[6] ** 2)))
    assert q[0].min() > 0 and prs.min() > 0


@pytest.mark.parametrize("p", DEGREES)
@pytest.mark.parametrize("prob,oedg_mode,cfl", [("ot2pi", 1, 0.2), ("rotor", 1, 0.2), ("blast", 2, 0.15)])
def test_solve_matches_oracle(prob, oedg_mode, cfl, p):
    nx, ny = 10, 9
    o, g = make_pair(prob, nx, ny, p)
    o.init_problem(prob)
    g.init_problem(prob)
    s_ref = o.run(-1.0, 4, cfl, oedg_mode, True)
    s = g.solve(-1.0, 4, cfl, oedg_mode, True)
    got = interior(g.download_state(0))
    ref = interior(o.u)
    err = normwise(got, ref)
    print(f"solve {prob} P{p}: normwise {err:.3e}, rel L1(rho) {rel_l1(got[0], ref[0]):.3e}, "
          f"steps {s['steps']}/{s_ref['steps']} retries {s['retries']}/{s_ref['retries']}")
    assert s["status"] == 0 and s_ref["status"] == 0
    assert s["steps"] == s_ref["steps"] and s["retries"] == s_ref["retries"]
    assert s["negcells"] == s_ref["negcells"]
    assert err <= oedg_tol(p, 1e-10)
    assert rel_l1(got[0], ref[0]) <= 1e-10
    # dt follows dt_lim, which carries the LDF round-off (at P6 lifted by the OEDG conditioning)
    assert abs(s["t"] - s_ref["t"]) <= (1e-12 if p <= 5 else 1e-10) * s_ref["t"]


@pytest.mark.parametrize("p", [5, 6])
@pytest.mark.parametrize("prob", ["ot2pi", "blast"])
def test_high_degree_oedg_without_ldf_matches_oracle(prob, p):
    """Degrees 5 and 6 with OEDG on and LDF off (the state then carries no
    last-bit difference from the LDF product's order): the exact device path
    follows the oracle to round-off, so the P6 tolerance above is the
    algorithm's conditioning, not a device error."""
    o, g = make_pair(prob, 9, 8, p)
    g.set_arith("exact")
    o.init_problem(prob)
    g.init_problem(prob)
    cfl, mode = (0.15, 2) if prob == "blast" else (0.2, 1)
    s_ref = o.run(-1.0, 3, cfl, mode, False)
    s = g.solve(-1.0, 3, cfl, mode, False)
    got, ref = interior(g.download_state(0)), interior(o.u)
    err = normwise(got, ref)
    # the oracle's own response to one ulp added to B_x, over the same steps:
    # the global OEDG means are summed per chunk here and serially there, and
    # the blast's pressure ratio of 1e4 lifts such a last-bit difference (the
    # oracle's spread: 3.7e-12 at P4, 7.7e-11 at P6)
    o2 = make_pair(prob, 9, 8, p)[0]
    o2.init_problem(prob)
    o2.u[4] = np.nextafter(o2.u[4], np.inf)
    o2.run(-1.0, 3, cfl, mode, False)
    spread = normwise(interior(o2.u), ref)
    print(f"{prob} P{p} OEDG without LDF: normwise {err:.3e} (oracle's one-ulp spread {spread:.3e})")
    assert s["steps"] == s_ref["steps"] and s["negcells"] == s_ref["negcells"]
    assert err <= max(1e-12, 10.0 * spread)


def test_solve_to_final_time_matches_oracle():
    """t_final clamp and the RK pseudo-time (drive.f90:115-120,207,312)."""
    o, g = make_pair("ot2pi", 12, 12, 2)
    o.init_problem("ot2pi")
    g.init_problem("ot2pi")
    s_ref = o.run(0.05, 0, 0.2, 1, True)
    s = g.solve(0.05, 0, 0.2, 1, True)
    assert s["steps"] == s_ref["steps"]
    # t advances through the RK pseudo-time (drive.f90:207,312), which can end
    # an ulp away from t_final, as in the reference
    assert s["t"] == s_ref["t"] and abs(s["t"] - 0.05) <= 1e-15
    assert rel_l1(interior(g.download_state(0))[0], interior(o.u)[0]) <= 1e-10


def test_mass_conservation_periodic():
    g = ppesdg.Solver(16, 16, 3, *[setup("ot2pi")[k] for k in ("domain", "gamma")])
    pert = perturbation_table(0)
    g.init_problem("otpert", pert, 1e-3)
    w = ppesdg.tables(3)["w"]
    geo = g.geometry()
    wt = 0.25 * geo["hx"] * geo["hy"] * np.outer(w, w)
    m0 = float(np.einsum("kl,klij->", wt, interior(g.download_state(0))[0]))
    g.solve(-1.0, 10, 0.2, 1, True)
    m1 = float(np.einsum("kl,klij->", wt, interior(g.download_state(0))[0]))
    assert abs(m1 - m0) <= 1e-13 * abs(m0)


@pytest.mark.parametrize("rk,cfl", [(1, 0.1), (2, 0.2), (5, 0.2)])
@pytest.mark.parametrize("prob", ["ot2pi", "rotor"])
def test_solve_other_rk_methods_match_oracle(prob, rk, cfl):
    """RK1, RK2 and RK5 stage combines, t_RK and CFL_RK (drive.f90:57-63,132-174,213-272)."""
    nx, ny, p = 9, 8, 2
    o, g = make_pair(prob, nx, ny, p, rk=rk)
    o.init_problem(prob)
    g.init_problem(prob)
    s_ref = o.run(-1.0, 4, cfl, 1, True)
    s = g.solve(-1.0, 4, cfl, 1, True)
    assert s["steps"] == s_ref["steps"] and s["retries"] == s_ref["retries"]
    assert normwise(interior(g.download_state(0)), interior(o.u)) <= 1e-10
    assert abs(s["t"] - s_ref["t"]) <= 1e-12 * s_ref["t"]


GOLDEN_END = ["end_blast_p2_40", "end_rotor_p2_32", "end_ot2pi_p1_32", "end_ot2pi_p4_12",
              "end_ot2pi_p2_16_rk1", "end_ot2pi_p2_16_rk2", "end_rotor_p3_16_rk5", "end_ot2pi_p2_128",
              "end_ot2pi_p6_8", "end_rotor_p5_10"]


@pytest.mark.parametrize("name", GOLDEN_END)
def test_solve_matches_reference_fixture(name):
    """The HIP path against the reference solver's own output at t_final
    (fixtures written by the flang build of the reference). end_ot2pi_p2_128
    is the CPU-runnable configuration: OT 2pi, P2, 128x128, CFL 0.2, t = 0.5."""
    import json
    import os
    z = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", name + ".npz"))
    a = json.loads(str(z["args_json"]))
    s = setup(a["prob"])
    g = ppesdg.Solver(a["nx"], a["ny"], a["p"], s["domain"], s["gamma"], s["bc"], s["bc_state"], rk=a.get("rk", 3))
    g.init_problem(a["prob"])
    st = g.solve(a["tfinal"], 0, a["cfl"], a["oedg"], a["ldf"])
    ref_steps = int(z["steps"]) if "steps" in z.files else int(z["summary"][1])
    rho = interior(g.download_state(0))[0]
    err = rel_l1(rho, z["rho_final"])
    print(f"{name}: steps {st['steps']}/{ref_steps}, rel L1(rho) {err:.3e}")
    assert st["status"] == 0 and st["steps"] == ref_steps
    assert err <= 1e-10
    assert rho.min() > 0


# ------------------------------------------------------------------ ent_offset

@pytest.mark.parametrize("p", DEGREES)
@pytest.mark.parametrize("oedg", [0, 1])
def test_post_process_ent_offset_matches_oracle(p, oedg):
    """ent_offset (DG_solver.f90:195-225,605-633) in one post_process call,
    with and without the switch. In one call the offset moves these states by
    round-off only (the LDF projection seldom raises a cell's entropy); its
    effect is pinned through whole solves below."""
    nx, ny = 8, 8
    u = advanced_state("rotor", nx, ny, p, 2)
    o, g = make_pair("rotor", nx, ny, p)
    for eo in (False, True):
        ref = np.array(u, copy=True)
        avg_ref = np.zeros(o.shape_avg, order="F")
        r_ref = o.post_process(ref, avg_ref, oedg, True, 3e-3, True, ent_offset=eo)
        g.upload_state(0, u)
        r = g.post_process(0, oedg, True, 3e-3, True, ent_offset=eo)
        got = g.download_state(0)
        assert normwise(strips(got), strips(ref)) <= 1e-12
        assert r["num_negcell"] == r_ref["num_negcell"]
        assert abs(r["tot_ent"] - r_ref["tot_ent"]) <= 1e-12 * abs(r_ref["tot_ent"])


@pytest.mark.parametrize("p", [2, 3, 4])
def test_solve_ent_offset_matches_oracle(p):
    """Rotor, CFL 0.6, 4 steps: the offset moves the state by 6e-10 to 1e-8
    normwise (oracle with vs without it), 10 to 100 times the tolerance."""
    o, g = make_pair("rotor", 10, 9, p)
    o.init_problem("rotor")
    g.init_problem("rotor")
    s_ref = o.run(-1.0, 4, 0.6, 1, True, ent_offset=True)
    s = g.solve(-1.0, 4, 0.6, 1, True, ent_offset=True)
    ref = np.array(interior(o.u), copy=True)
    o.init_problem("rotor")
    o.run(-1.0, 4, 0.6, 1, True, ent_offset=False)
    assert normwise(interior(o.u), ref) >= 5e-10
    assert s["steps"] == s_ref["steps"] and s["retries"] == s_ref["retries"]
    assert normwise(interior(g.download_state(0)), ref) <= 5e-11
    assert abs(s["tot_ent"] - s_ref["tot_ent"]) <= 1e-12 * abs(s_ref["tot_ent"])


def test_solve_ent_offset_matches_reference_fixture():
    import json
    import os
    z = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "end_ot2pi_p2_16_entoff.npz"))
    a = json.loads(str(z["args_json"]))
    s = setup(a["prob"])
    g = ppesdg.Solver(a["nx"], a["ny"], a["p"], s["domain"], s["gamma"], s["bc"], s["bc_state"])
    g.init_problem(a["prob"])
    st = g.solve(a["tfinal"], 0, a["cfl"], a["oedg"], a["ldf"], ent_offset=True)
    assert st["steps"] == int(z["steps"])
    assert rel_l1(interior(g.download_state(0))[0], z["rho_final"]) <= 1e-10


# --------------------------------------------------------- alternative fluxes

@pytest.mark.parametrize("arith", ["exact", "fast"])
@pytest.mark.parametrize("p", DEGREES)
@pytest.mark.parametrize("flux", ["pp_hll", "espp_lf", "ec_lf"])
def test_residual_alternative_flux_matches_oracle(flux, p, arith):
    """PP_HLL_flux_x / ESPP_LF_flux_x / EC_LF_flux_x at the face nodes
    (DG_solver.f90:1073-1497; the oracle's flux functions equal the
    reference's bit for bit on fixed state pairs). The exact mode agrees to
    1e-14 normwise (bit for bit but where a device log differs from glibc's
    in the last bit: one energy value at an element centre of the P4 case),
    the fast mode (fast volume terms) within 1e-12."""
    nx, ny = 9, 7
    u = advanced_state("blast", nx, ny, p, 2, oedg_mode=2)
    o, g = make_pair("blast", nx, ny, p)
    o.set_flux(flux)
    g.set_flux(flux)
    g.set_arith(arith)
    ut_ref = np.zeros(o.shape_ut, order="F")
    dtl_ref = o.semidiscretize(ut_ref, u)
    g.upload_state(0, u)
    dtl = g.semidiscretize(1, 0)
    ut = g.download_residual(1)
    if arith == "exact":
        assert normwise(ut, ut_ref) <= 1e-14
        assert np.sum(ut != ut_ref) <= 2
        assert dtl == dtl_ref
    else:
        assert normwise(ut, ut_ref) <= 1e-12
        assert abs(dtl - dtl_ref) <= 1e-14 * dtl_ref
    # the choice matters for the Lax-Friedrichs forms (the HLL forms agree
    # wherever the PP speeds bound the ES ones, as in this early blast state)
    if flux != "pp_hll":
        o.set_flux("espp_hll")
        ut_def = np.zeros(o.shape_ut, order="F")
        o.semidiscretize(ut_def, u)
        assert normwise(ut_def, ut_ref) > 1e-8


@pytest.mark.parametrize("flux", ["pp_hll", "espp_lf", "ec_lf"])
def test_solve_alternative_flux_matches_oracle(flux):
    o, g = make_pair("rotor", 10, 9, 2)
    o.set_flux(flux)
    g.set_flux(flux)
    o.init_problem("rotor")
    g.init_problem("rotor")
    s_ref = o.run(-1.0, 4, 0.2, 1, True)
    s = g.solve(-1.0, 4, 0.2, 1, True)
    assert s["steps"] == s_ref["steps"] and s["retries"] == s_ref["retries"]
    assert normwise(interior(g.download_state(0)), interior(o.u)) <= 1e-10


def test_alternative_flux_instance_alone_equals_in_mix():
    """A batch mixing face fluxes within one degree group: each instance
    gives the same bits as alone."""
    s = setup("blast")
    res = {}
    for flux in ("pp_hll", "ec_lf"):
        b = ppesdg.Batch(3)
        i = b.add(10, 8, 2, s["domain"], s["gamma"], s["bc"], s["bc_state"])
        b.set_flux(i, flux)
        b.init_problem(i, "blast")
        b.start(i, -1.0, 3, 0.15, 2, True)
        b.step(0)
        res[flux] = b.download_state(i, 0)
    b = ppesdg.Batch(3)
    ids = {}
    for k, flux in enumerate(("espp_hll", "pp_hll", "ec_lf")):
        i = b.add(10, 8, 2, s["domain"], s["gamma"], s["bc"], s["bc_state"])
        b.set_flux(i, flux)
        b.init_problem(i, "blast")
        b.start(i, -1.0, 3, 0.15, 2, True)
        ids[flux] = i
    b.step(0)
    for flux in ("pp_hll", "ec_lf"):
        assert np.array_equal(b.download_state(ids[flux], 0), res[flux])
