"""Pins the oracle's perturbed Orszag-Tang initial data (the P3 1024^2
benchmark workload) against its definition, built here independently.

Definition (the workload of the single-instance benchmark):
  * base state on [0, 2pi]^2, gamma = 5/3: rho = gamma^2, p = gamma,
    v = (-sin y, sin x, 0), B = (-sin y, sin 2x, 0);
  * rho += 1e-3 * sum_k cos(k.x + phi_k) / n;
  * B += curl(A_z e_z) = (dA_z/dy, -dA_z/dx, 0) with
    A_z = 1e-3 * sum_k sin(k.x + psi_k) / (|k| n);
  * k runs over the n = 48 integer vectors with 0 < |k|_2 <= 4; phi and psi
    are numpy.random.default_rng(0).uniform(0, 2pi, (2, n)), committed as
    tests/golden/otpert_phases_seed0.json.
Nodal sampling at the GLL nodes, then the primitive-to-conservative map
(problems.f90:25-51, MHD_pri2con at MHD_PDE.f90:58-70).

The construction below evaluates cos(k.x + phi) directly (the oracle uses
the separable expansion), takes B from the analytic curl, places the nodes
from numpy's Legendre roots and maps to conservative variables in numpy. It
shares nothing with the oracle but the phase table. A plausible mistake
(phi and psi swapped, the wrong sign of B_y, the wrong normalisation) is shown
to fail the same comparison.
"""
import json
import math
import os

import numpy as np
import pytest

import oracle as O
import workloads as W
from helpers import perturbation_table

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
AMP = 1e-3
GAMMA = 5.0 / 3.0


def phase_fixture():
    with open(os.path.join(GOLD, "otpert_phases_seed0.json")) as f:
        doc = json.load(f)
    return np.array([[r[0], r[1], float.fromhex(r[2]), float.fromhex(r[3])] for r in doc["rows"]])


def gll_nodes(p):
    """Roots of (1 - x^2) P_p'(x), ascending."""
    c = np.zeros(p + 1)
    c[p] = 1.0
    inner = np.polynomial.legendre.Legendre(c).deriv().roots() if p > 1 else np.array([])
    return np.concatenate(([-1.0], np.sort(np.real(inner)), [1.0]))


def node_coords(n, p, lo, hi):
    """x of GLL node k of element i: left + (i - 1/2) h + (h/2) xhat_k -> (p+1, n)."""
    h = (hi - lo) / n
    xh = gll_nodes(p)
    return lo + (np.arange(n)[None, :] + 0.5) * h + 0.5 * h * xh[:, None]


def perturbed_ot_conservative(x, y, table, amp=AMP, g=GAMMA, swap_phases=False, by_sign=-1.0, norm=None):
    """The definition at points (x, y) (same shape) -> conservative (8, ...)."""
    n = len(table) if norm is None else norm
    rho = np.full_like(x, g * g)
    pr = np.full_like(x, g)
    vx, vy, vz = -np.sin(y), np.sin(x), np.zeros_like(x)
    bx, by, bz = -np.sin(y), np.sin(2.0 * x), np.zeros_like(x)
    drho = np.zeros_like(x)
    dbx = np.zeros_like(x)
    dby = np.zeros_like(x)
    for kx, ky, phi, psi in table:
        if swap_phases:
            phi, psi = psi, phi
        kn = math.hypot(kx, ky)
        drho += np.cos(kx * x + ky * y + phi)
        # A_z = sin(k.x + psi) / |k|: dA/dy = ky cos(.)/|k|, dA/dx = kx cos(.)/|k|
        dadx_dady = np.cos(kx * x + ky * y + psi) / kn
        dbx += ky * dadx_dady
        dby += by_sign * kx * dadx_dady
    rho = rho + amp * drho / n
    bx = bx + amp * dbx / n
    by = by + amp * dby / n
    e = pr / (g - 1.0) + 0.5 * (rho * (vx ** 2 + vy ** 2 + vz ** 2) + (bx ** 2 + by ** 2 + bz ** 2))
    return np.stack([rho, rho * vx, rho * vy, rho * vz, bx, by, bz, e])


def vector_potential(x, y, table, amp=AMP):
    n = len(table)
    a = np.zeros_like(x)
    for kx, ky, _, psi in table:
        a += np.sin(kx * x + ky * y + psi) / math.hypot(kx, ky)
    return amp * a / n


def oracle_interior(nx, ny, p, table):
    dom = (0.0, 2 * math.pi, 0.0, 2 * math.pi)
    o = O.Oracle(nx, ny, p, dom, GAMMA, "periodic")
    o.init_problem("otpert", table, AMP)
    return o, np.array(o.u[:, :, :, 1:-1, 1:-1])


def independent_interior(nx, ny, p, table, **kw):
    xs = node_coords(nx, p, 0.0, 2 * math.pi)   # (N, nx)
    ys = node_coords(ny, p, 0.0, 2 * math.pi)   # (N, ny)
    # layout (v, k, l, i, j): x depends on (k, i), y on (l, j)
    X = np.broadcast_to(xs[:, None, :, None], (p + 1, p + 1, nx, ny))
    Y = np.broadcast_to(ys[None, :, None, :], (p + 1, p + 1, nx, ny))
    return perturbed_ot_conservative(np.array(X), np.array(Y), table, **kw)


def worst_rel(a, b):
    """max over variables of max|a_v - b_v| / max(max|b_v|, 1) (m_z and B_z are 0)."""
    return max(float(np.max(np.abs(a[v] - b[v]))) / max(float(np.max(np.abs(b[v]))), 1.0) for v in range(8))


def test_phase_fixture_is_the_seeded_draw():
    """The committed table is default_rng(0)'s draw, and both generators reproduce it bit for bit."""
    fix = phase_fixture()
    assert fix.shape == (48, 4)
    ks = {(int(a), int(b)) for a, b in fix[:, :2]}
    assert ks == {(a, b) for a in range(-4, 5) for b in range(-4, 5) if 0 < a * a + b * b <= 16}
    ph = np.random.default_rng(0).uniform(0.0, 2 * math.pi, size=(2, 48))
    assert np.array_equal(fix[:, 2], ph[0]) and np.array_equal(fix[:, 3], ph[1])
    assert np.array_equal(W.perturbation_table(0), fix)
    assert np.array_equal(perturbation_table(0), fix)


def test_independent_construction_is_divergence_free_curl():
    """The construction's B perturbation is the curl of A_z (central
    differences of the potential, step 1e-5) and has zero divergence; its
    density perturbation has zero mean over the periodic domain."""
    tab = phase_fixture()
    rng = np.random.default_rng(5)
    x = rng.uniform(0, 2 * math.pi, 200)
    y = rng.uniform(0, 2 * math.pi, 200)
    u = perturbed_ot_conservative(x, y, tab)
    base = perturbed_ot_conservative(x, y, tab, amp=0.0)
    dbx, dby = u[4] - base[4], u[5] - base[5]
    h = 1e-5
    dady = (vector_potential(x, y + h, tab) - vector_potential(x, y - h, tab)) / (2 * h)
    dadx = (vector_potential(x + h, y, tab) - vector_potential(x - h, y, tab)) / (2 * h)
    assert np.max(np.abs(dbx - dady)) < 1e-12 and np.max(np.abs(dby + dadx)) < 1e-12
    div = ((perturbed_ot_conservative(x + h, y, tab)[4] - perturbed_ot_conservative(x - h, y, tab)[4])
           + (perturbed_ot_conservative(x, y + h, tab)[5] - perturbed_ot_conservative(x, y - h, tab)[5])) / (2 * h)
    assert np.max(np.abs(div)) < 1e-8
    # mean of the density perturbation over a uniform periodic grid is 0 (|k| <= 4 < 64)
    g = np.linspace(0, 2 * math.pi, 64, endpoint=False)
    GX, GY = np.meshgrid(g, g, indexing="ij")
    d = perturbed_ot_conservative(GX, GY, tab)[0] - GAMMA ** 2
    assert abs(float(np.mean(d))) < 1e-17 and float(np.max(np.abs(d))) > 1e-4


@pytest.mark.parametrize("nx,ny,p", [(16, 12, 3), (9, 7, 2), (6, 5, 4)])
def test_oracle_otpert_matches_definition(nx, ny, p):
    """Oracle.init_problem("otpert") equals the independent construction to
    1e-14 relative per variable, at nodes the test places itself."""
    tab = phase_fixture()
    o, uo = oracle_interior(nx, ny, p, tab)
    # the oracle's node coordinates are the GLL nodes of the uniform grid
    assert np.max(np.abs(np.array(o.xx) - node_coords(nx, p, 0.0, 2 * math.pi))) < 1e-14
    assert np.max(np.abs(np.array(o.yy) - node_coords(ny, p, 0.0, 2 * math.pi))) < 1e-14
    ui = independent_interior(nx, ny, p, tab)
    assert worst_rel(uo, ui) <= 1e-14
    # the perturbation itself (state minus base) agrees to 1e-10 of its size,
    # beyond the rounding of the base state it is added to
    base = independent_interior(nx, ny, p, tab, amp=0.0)
    eps = np.finfo(np.float64).eps
    for v in (0, 4, 5, 7):
        d_o, d_i = uo[v] - base[v], ui[v] - base[v]
        tol = 1e-10 * np.max(np.abs(d_i)) + 4 * eps * np.max(np.abs(ui[v]))
        assert np.max(np.abs(d_o - d_i)) <= tol, v


@pytest.mark.parametrize("mistake", [dict(swap_phases=True), dict(by_sign=+1.0), dict(norm=47)])
def test_plausible_mistakes_fail_the_comparison(mistake):
    """phi and psi swapped, B_y with the wrong sign, or n off by one: the
    same comparison that passes above rejects each."""
    tab = phase_fixture()
    _, uo = oracle_interior(16, 12, 3, tab)
    wrong = independent_interior(16, 12, 3, tab, **mistake)
    assert worst_rel(uo, wrong) > 1e-7
