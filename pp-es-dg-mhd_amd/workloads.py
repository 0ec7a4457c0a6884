"""Synthetic workloads of the PP-ES DG step (problem set-ups and mixes).

Problems (domain, gamma, boundary condition, CFL, OEDG mode):
  ot2pi  Orszag-Tang on [0,2pi]^2, gamma 5/3, periodic, CFL 0.2, OEDG on the
         final RK stage
  rotor  rotor of problems.f90:170-199 on [0,1]^2, gamma 1.4, fixed ambient
         state in the ghosts (BC_rotor, problems.f90:597-654), CFL 0.2, OEDG
         on the final RK stage
  blast  blast wave on [-0.5,0.5]^2, gamma 1.4, transmissive (BC_YeeSjogreen,
         problems.f90:386-423), CFL 0.15, OEDG on every RK stage
LDF and the positivity limiter are on in all of them.
"""
import math

import numpy as np

TWO_PI = 2.0 * math.pi


def rotor_bc_state(g=1.4):
    """Conservative ambient state of BC_rotor (problems.f90:597-654)."""
    bx = 5.0 / math.sqrt(4 * math.pi)
    return np.array([1.0, 0.0, 0.0, 0.0, bx, 0.0, 0.0, 1.0 / (g - 1.0) + 0.5 * (1.0 * 0.0 + bx * bx)])


# name -> (domain, gamma, bc, bc_state, cfl, oedg_mode)
PROBLEMS = {
    "ot2pi": ((0.0, TWO_PI, 0.0, TWO_PI), 5.0 / 3.0, "periodic", None, 0.2, 1),
    "rotor": ((0.0, 1.0, 0.0, 1.0), 1.4, "fixed", rotor_bc_state(), 0.2, 1),
    "blast": ((-0.5, 0.5, -0.5, 0.5), 1.4, "extrap", None, 0.15, 2),
}


def perturbation_table(seed=0, kmax=4):
    """Integer wave vectors 0 < |k| <= kmax (kx, ky ascending) with phases
    phi, psi drawn as numpy.random.default_rng(seed).uniform(0, 2pi, (2, n))."""
    ks = [(kx, ky) for kx in range(-kmax, kmax + 1) for ky in range(-kmax, kmax + 1)
          if 0 < kx * kx + ky * ky <= kmax * kmax]
    ph = np.random.default_rng(seed).uniform(0.0, TWO_PI, size=(2, len(ks)))
    return np.array([[kx, ky, ph[0, q], ph[1, q]] for q, (kx, ky) in enumerate(ks)])


def mixed_batch(n=64, seed=1, sides=(32, 64, 128, 256), degrees=(1, 2, 3, 4)):
    """Uneven instances: for each of n solves, in turn, the problem (uniform
    over ot2pi, rotor, blast), the grid side (uniform over `sides`) and the
    degree (uniform over `degrees`) drawn from numpy default_rng(seed)."""
    rng = np.random.default_rng(seed)
    names = list(PROBLEMS)
    out = []
    for _ in range(n):
        prob = names[int(rng.integers(len(names)))]
        side = int(sides[int(rng.integers(len(sides)))])
        p = int(degrees[int(rng.integers(len(degrees)))])
        out.append(dict(prob=prob, nx=side, ny=side, p=p))
    return out


# end time of each problem that a serving request's t_final is a fraction of
HORIZON = {"ot2pi": 0.5, "rotor": 0.295, "blast": 0.01}


def serving_stream(rate=20.0, duration=120.0, seed=2, sides=(64, 128, 256, 512), degrees=(1, 2, 3, 4),
                   t_frac=(0.1, 1.0), horizon=None):
    """Poisson arrivals (exponential inter-arrival times of mean 1/rate) over
    [0, duration); each request draws, in turn, its problem, side, degree and
    end time t_final = f * horizon[problem], f uniform in t_frac."""
    horizon = HORIZON if horizon is None else horizon
    rng = np.random.default_rng(seed)
    names = list(PROBLEMS)
    t = 0.0
    out = []
    while True:
        t += float(rng.exponential(1.0 / rate))
        if t >= duration:
            break
        prob = names[int(rng.integers(len(names)))]
        side = int(sides[int(rng.integers(len(sides)))])
        p = int(degrees[int(rng.integers(len(degrees)))])
        tf = float(rng.uniform(*t_frac)) * horizon[prob]
        out.append(dict(arrival=t, prob=prob, nx=side, ny=side, p=p, t_final=tf))
    return out


def nodes(inst):
    return inst["nx"] * inst["ny"] * (inst["p"] + 1) ** 2
