"""Shared helpers of the test suite: problem set-ups and comparison norms."""
import math

import numpy as np

TWO_PI = 2.0 * math.pi

# name -> (domain, gamma, bc, bc_state)
SETUPS = {
    "ot2pi": ((0.0, TWO_PI, 0.0, TWO_PI), 5.0 / 3.0, "periodic", None),
    "otunit": ((0.0, 1.0, 0.0, 1.0), 5.0 / 3.0, "periodic", None),
    "blast": ((-0.5, 0.5, -0.5, 0.5), 1.4, "extrap", None),
    "blastref": ((0.0, 1.0, 0.0, 1.0), 1.4, "periodic", None),
}


def rotor_bc_state(g=1.4):
    """Conservative ambient state of BC_rotor (problems.f90:597-654)."""
    bx = 5.0 / math.sqrt(4 * math.pi)
    return np.array([1.0, 0.0, 0.0, 0.0, bx, 0.0, 0.0, 1.0 / (g - 1.0) + 0.5 * (1.0 * 0.0 + bx * bx)])


SETUPS["rotor"] = ((0.0, 1.0, 0.0, 1.0), 1.4, "fixed", rotor_bc_state())
# the rest of the reference's catalogue (main.f90:57-190), on its domains
SETUPS["alfven"] = ((0.0, 1.0 / math.cos(math.pi / 6.0), 0.0, 1.0 / math.sin(math.pi / 6.0)), 5.0 / 3.0, "periodic", None)
SETUPS["khi"] = ((0.0, 1.0, -1.0, 1.0), 1.4, "periodic", None)
SETUPS["briowu"] = ((-1.0, 1.0, -0.5, 0.5), 5.0 / 3.0, "briowu", None)
SETUPS["wave1d"] = ((0.0, TWO_PI, -0.5, 0.5), 5.0 / 3.0, "periodic", None)
SETUPS["cloud"] = ((0.0, 1.0, 0.0, 1.0), 5.0 / 3.0, "cloud", None)
SETUPS["jet"] = ((0.0, 0.5, 0.0, 1.5), 1.4, "jet", None)
SETUPS["yee"] = ((-1.0, 1.0, -1.0, 1.0), 5.0 / 3.0, "extrap", None)


def setup(prob):
    dom, g, bc, st = SETUPS[prob]
    return dict(domain=dom, gamma=g, bc=bc, bc_state=st)


def normwise(a, b, zero_below=1e-6):
    """max over variables v of max|a_v - b_v| / max|b_v|.

    A variable that is zero in exact arithmetic (m_z and B_z in these 2D
    problems, or the density residual of a uniform density in a
    divergence-free flow) holds only round-off; when its max|b_v| is below
    `zero_below` times the largest variable's, it is measured against that
    largest scale instead, so that noise / noise is not read as an error."""
    a = np.asarray(a)
    b = np.asarray(b)
    scale = max(max(float(np.max(np.abs(b[v]))) for v in range(b.shape[0])), 1e-300)
    worst = 0.0
    for v in range(a.shape[0]):
        mv = float(np.max(np.abs(b[v])))
        den = mv if mv >= zero_below * scale else scale
        worst = max(worst, float(np.max(np.abs(a[v] - b[v]))) / den)
    return worst


def interior(u):
    return u[:, :, :, 1:-1, 1:-1]


def strips(u):
    """Interior plus ghost strips, corners zeroed (never written)."""
    w = np.array(u, copy=True)
    for ex in (0, -1):
        for ey in (0, -1):
            w[:, :, :, ex, ey] = 0.0
    return w


def rel_l1(a, b):
    return float(np.sum(np.abs(a - b)) / np.sum(np.abs(b)))


def perturbation_table(seed=0, kmax=4):
    """Integer wave vectors 0 < |k| <= kmax (kx, ky ascending) with phases
    phi, psi drawn as numpy.random.default_rng(seed).uniform(0, 2pi, (2, n))."""
    ks = [(kx, ky) for kx in range(-kmax, kmax + 1) for ky in range(-kmax, kmax + 1)
          if 0 < kx * kx + ky * ky <= kmax * kmax]
    ph = np.random.default_rng(seed).uniform(0.0, TWO_PI, size=(2, len(ks)))
    return np.array([[kx, ky, ph[0, q], ph[1, q]] for q, (kx, ky) in enumerate(ks)])
