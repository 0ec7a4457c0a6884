"""The C-ABI library on the CPU: it loads, exports every entry point declared
in include/ppesdg.h, and its host-only table builders agree with the oracle
and the reference (no device calls here)."""
import os
import re

import numpy as np
import pytest

import oracle as O
import ppesdg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "ppesdg.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(ppes_[a-z_0-9]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = ppesdg.lib()
    names = declared_symbols()
    assert "ppes_create" in names and "ppes_solve" in names
    for name in names:
        assert hasattr(L, name), name


@pytest.mark.parametrize("p", [0, 1, 2, 3, 4, 5, 6])
def test_library_tables_equal_oracle(p):
    a, b = ppesdg.tables(p), O.tables(p)
    for k in ("x", "w", "D", "Pmn", "Pnm"):
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("p", [1, 2, 3, 4])
@pytest.mark.parametrize("hx,hy", [(0.1, 0.1), (0.05, 0.2)])
def test_ldf_nodal_operator_matches_modal_projector(p, hx, hy):
    """Dense nodal LDF operator = Pmn2 * diag(2/hy, 2/hx) * P_modal * diag(hy/2, hx/2) * Pnm2,
    with P_modal the oracle's (reference-pinned) projector."""
    A = ppesdg.ldf_nodal_operator(p, hx, hy)
    t = O.tables(p)
    n = p + 1
    W2 = np.kron(t["Pnm"], t["Pnm"])   # index a + N b <- k + N l (column-major kron order)
    V2 = np.kron(t["Pmn"], t["Pmn"])
    # numpy kron(A, B)[(i*n + j), ...] pairs the slow index with A; our node
    # index is k + N l (k fast), so use kron(P_l, P_k) = kron(Pnm, Pnm) with
    # the slow factor acting on l -- identical factors, so the order is immaterial
    Pm = O.ldf_modal_operator(p)
    NN = n * n
    Sin = np.diag(np.r_[np.full(NN, hy / 2), np.full(NN, hx / 2)])
    Sout = np.diag(np.r_[np.full(NN, 2 / hy), np.full(NN, 2 / hx)])
    Z = np.zeros((NN, NN))
    ref = Sout @ np.block([[V2, Z], [Z, V2]]) @ Pm @ np.block([[W2, Z], [Z, W2]]) @ Sin
    assert np.max(np.abs(A - ref)) <= 1e-13 * max(1.0, np.max(np.abs(ref)))
    # a constant field (in the nullspace of the divergence) is unchanged
    assert np.max(np.abs(A @ np.ones(2 * NN) - 1.0)) < 1e-13


def test_element_shapes_accepted():
    """The device takes py = px in 1..6 (every degree of the reference's
    tables) and the reference's 1D set-ups py = 0 (px in 1..6); other pairs
    are rejected before any allocation."""
    for p in (1, 2, 3, 4, 5, 6):
        assert ppesdg.instance_bytes(64, 64, p) > 0
        assert ppesdg.instance_bytes(480, 1, p, rk=5, py=0) > 0
    for p, py in ((3, 1), (2, 3), (0, 0), (7, 7), (7, 0), (6, 5)):
        assert ppesdg.instance_bytes(16, 16, p, py=py) == -1
    # a P3 x P0 element holds a quarter of a P3 element's nodes
    a, b = ppesdg.instance_bytes(480, 1, 3, rk=5, py=0), ppesdg.instance_bytes(480, 1, 3, rk=5)
    assert 3.0 < b / a < 4.5
