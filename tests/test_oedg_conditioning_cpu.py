"""Conditioning of the OEDG damping at high degree, on the oracle (CPU).

The OEDG indicators (DG_solver.f90:295-422) weigh the face jumps of every
normal and transverse derivative up to the element's degree. At degree 6 the
order-6 derivative of a smooth state is about 1e-4 of the state while the
repeated products with D that form it reach |D|^6 ~ 1e8 times the state:
the jump keeps only a few digits, and any last-bit change of the state moves
the damping factor by far more than the change itself. Measured here on the
oracle alone: one ulp added to B_x at every node moves the state after four
RK3 steps with OEDG by ~1e-8 normwise at P6 (~1e-14 at P4, ~1e-12 at P5), and
by ~1e-15 without OEDG. The GPU parity tests of P6 with OEDG use a tolerance
set from this measurement (the L1 density tolerance stays 1e-10)."""
import numpy as np
import pytest

import oracle as O
from helpers import interior, normwise, rel_l1, setup


def spread(p, oedg):
    s = setup("ot2pi")
    out = []
    for bump in (False, True):
        o = O.Oracle(10, 9, p, s["domain"], s["gamma"], s["bc"], s["bc_state"])
        o.init_problem("ot2pi")
        if bump:
            o.u[4] = np.nextafter(o.u[4], np.inf)
        o.run(-1.0, 4, 0.2, oedg, True)
        out.append(interior(np.array(o.u)))
    return normwise(out[1], out[0]), rel_l1(out[1][0], out[0][0])


@pytest.mark.parametrize("p,lo,hi", [(4, 0.0, 1e-12), (5, 0.0, 1e-11), (6, 1e-10, 1e-7)])
def test_one_ulp_sensitivity_with_oedg(p, lo, hi):
    nw, l1 = spread(p, 1)
    print(f"P{p}: one-ulp B_x change -> normwise {nw:.2e}, rel L1(rho) {l1:.2e}")
    assert lo <= nw <= hi
    assert l1 <= 1e-10


@pytest.mark.parametrize("p", [4, 6])
def test_one_ulp_sensitivity_without_oedg(p):
    nw, _ = spread(p, 0)
    assert nw <= 1e-14
