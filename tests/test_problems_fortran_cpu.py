"""The problem catalogue of the Fortran drop-in module (problems.f90
counterpart, host code): initial data through init_by_fun_pri /
init_by_fun_con, the boundary fillers and the error functionals, against the
CPU oracle, whose own versions are pinned to the reference's fixtures
(test_oracle_golden.py). Initial data and ghost strips bit for bit; the
error functionals to round-off of the sum.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from helpers import setup

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FORT = os.path.join(ROOT, "pp-es-dg-mhd_amd", "fortran")
LIBDIR = os.path.join(ROOT, "pp-es-dg-mhd_amd", "lib")
FLANG = "/opt/rocm/lib/llvm/bin/flang"

PROGRAM = """
program problems_check
	use MHD_PDE_m, only: set_gamma
	use problems
	use gauss_m
	implicit none
	integer :: u, what, prob, nx, ny, px, py
	real :: g, hx, hy, t, e(2), x1(1), w1(1), m1(1,1,5)
	real, allocatable :: xx(:,:), yy(:,:), uc(:,:,:,:,:)
	open(newunit=u, file='in.bin', access='stream', form='unformatted', status='old')
	read(u) what, prob, nx, ny, px, py
	read(u) g, hx, hy, t
	allocate(xx(px+1, nx), yy(py+1, ny), uc(8, px+1, py+1, 0:nx+1, 0:ny+1))
	read(u) xx, yy, uc
	close(u)
	call set_gamma(g)
	select case (what)
	case (0)  ! initial data
		select case (prob)
		case (6); init_pri_fun => smooth_alfven_init_fun
		case (7); init_pri_fun => KHI_init_pri
		case (8); init_pri_fun => Brio_Wu_init_fun
		case (9); init_pri_fun => smooth_1Dwave_init_fun
		case (10); init_pri_fun => cloud_shock_init_pri
		case (11); init_pri_fun => jet1_init_pri
		end select
		if (prob == 12) then
			init_con_fun => YeeSjogreenInitCon
			call init_by_fun_con(uc, xx, yy, nx, ny, px, py)
		else
			call init_by_fun_pri(uc, xx, yy, nx, ny, px, py)
		end if
	case (1)  ! boundary filler
		select case (prob)
		case (4); call BC_BrioWu(uc, xx, yy, nx, ny, px, py)
		case (5); call BC_cloud_shock(uc, xx, yy, nx, ny, px, py)
		case (6); call BC_jet1_cell_ext(uc, xx, yy, nx, ny, px, py)
		end select
	case (2)
		call compute_error_smooth_alfven(uc, hx, hy, xx, yy, nx, ny, px, py, e(1), e(2))
	case (3)
		call compute_error_smooth1D(uc, hx, hy, xx, yy, nx, ny, px, py, t, e(1), e(2))
	case (4)  ! the one-node tables of a degree-0 direction
		call gauss_lobatto(1, x1, w1)
		call gauss_lobatto_matrix(0, m1(:,:,1), m1(:,:,2), m1(:,:,3), m1(:,:,4), m1(:,:,5))
		uc = 0.0
		uc(1:5, 1, 1, 0, 0) = m1(1, 1, :)
		e = [x1(1), w1(1)]
	end select
	open(newunit=u, file='out.bin', access='stream', form='unformatted', status='replace')
	write(u) uc, e
	close(u)
end program problems_check
"""


@pytest.fixture(scope="module")
def prog(tmp_path_factory):
    if not os.path.exists(FLANG):
        pytest.skip("flang not available")
    if not os.path.exists(os.path.join(LIBDIR, "libppesdg.so")):
        pytest.skip("libppesdg.so not built")
    d = tmp_path_factory.mktemp("problems")
    srcs = [os.path.join(FORT, f) for f in ("utility.f90", "ppesdg_capi.f90", "gauss.f90", "MHD_PDE.f90",
                                            "problems.f90")]
    with open(d / "prog.f90", "w") as f:
        f.write(PROGRAM)
    subprocess.check_call([FLANG, "-fdefault-real-8", "-cpp", "-O2", "-ffp-contract=off", "-o", str(d / "check")] +
                          srcs + [str(d / "prog.f90"), "-L" + LIBDIR, "-lppesdg", "-Wl,-rpath," + LIBDIR], cwd=d)
    return d


def call(d, what, prob, o, s, u, t=0.0):
    """Run the check program on oracle instance o (set-up s) and state u."""
    nx, ny, p = o.nx, o.ny, o.p
    py = o.shape_u[2] - 1
    dom = s["domain"]
    with open(d / "in.bin", "wb") as f:
        np.array([what, prob, nx, ny, p, py], dtype="<i4").tofile(f)
        np.array([s["gamma"], (dom[1] - dom[0]) / nx, (dom[3] - dom[2]) / ny, t], dtype="<f8").tofile(f)
        for a in (o.xx, o.yy, u):
            np.asfortranarray(a, dtype="<f8").ravel(order="F").tofile(f)
    subprocess.check_call([str(d / "check")], cwd=d)
    out = np.fromfile(d / "out.bin", dtype="<f8")
    n = int(np.prod(o.shape_u))
    return out[:n].reshape(o.shape_u, order="F"), out[n:]


def make(prob, nx, ny, p, py=None):
    s = setup(prob)
    return O.Oracle(nx, ny, p, s["domain"], s["gamma"], s["bc"], s["bc_state"], py=py), s


@pytest.mark.parametrize("p", [1, 3])
@pytest.mark.parametrize("prob", ["alfven", "khi", "briowu", "wave1d", "cloud", "jet", "yee"])
def test_initial_data(prog, prob, p):
    o, s = make(prob, 9, 7, p)
    o.init_problem(prob)
    got, _ = call(prog, 0, O.PROB[prob], o, s, np.zeros(o.shape_u))
    assert np.array_equal(got[:, :, :, 1:-1, 1:-1], np.array(o.u)[:, :, :, 1:-1, 1:-1])


@pytest.mark.parametrize("p", [1, 2, 4])
@pytest.mark.parametrize("prob", ["briowu", "cloud", "jet"])
def test_boundary_fillers(prog, prob, p):
    o, s = make(prob, 20, 6, p)
    o.init_problem(prob)
    rng = np.random.default_rng(11)
    u = np.array(o.u, copy=True)
    u[:, :, :, 1:-1, 1:-1] += 0.05 * rng.standard_normal(u[:, :, :, 1:-1, 1:-1].shape)
    got, _ = call(prog, 1, O.BC[s["bc"]], o, s, u)
    ref = np.array(u, copy=True)
    o.apply_bc(ref)
    for ex in (0, -1):
        assert np.array_equal(got[:, :, :, ex, 1:-1], ref[:, :, :, ex, 1:-1])
    for ey in (0, -1):
        assert np.array_equal(got[:, :, :, 1:-1, ey], ref[:, :, :, 1:-1, ey])


def test_error_functionals(prog):
    o, s = make("alfven", 8, 6, 2)
    o.init_problem("alfven")
    u = np.array(o.u, copy=True)
    u[4:6] *= 1.01
    _, e = call(prog, 2, 0, o, s, u)
    ref = o.error_smooth_alfven(u)
    assert abs(e[0] - ref[0]) <= 1e-14 * ref[0] and e[1] == ref[1]
    o, s = make("wave1d", 12, 2, 3)
    o.init_problem("wave1d")
    u = np.array(o.u, copy=True)
    _, e = call(prog, 3, 0, o, s, u, t=0.05)
    ref = o.error_smooth1d(0.05, u)
    assert abs(e[0] - ref[0]) <= 1e-14 * ref[0] and e[1] == ref[1]


@pytest.mark.parametrize("p", [1, 3])
def test_degree_zero_direction(prog, p):
    """The reference's 1D set-ups (py = 0, main.f90 cases 0 and 1): the
    one-node Gauss-Lobatto rule and tables (x = 0, w = 2, M = 2, M^-1 = 1/2,
    D = 0, Pmn = Pnm = 1, gauss.f90:35-37,165-170), and the Brio-Wu data and
    filler and the 1D error functional on a P x P0 strip."""
    o, s = make("briowu", 16, 1, p, py=0)
    got, e = call(prog, 4, 0, o, s, np.zeros(o.shape_u))
    assert list(e) == [0.0, 2.0]
    assert list(got[:5, 0, 0, 0, 0]) == [2.0, 0.5, 0.0, 1.0, 1.0]
    o.init_problem("briowu")
    got, _ = call(prog, 0, O.PROB["briowu"], o, s, np.zeros(o.shape_u))
    assert np.array_equal(got[:, :, :, 1:-1, 1:-1], np.array(o.u)[:, :, :, 1:-1, 1:-1])
    u = np.array(o.u, copy=True)
    u[:, :, :, 1:-1, 1:-1] *= 1.0 + 0.01 * np.random.default_rng(5).standard_normal(u[:, :, :, 1:-1, 1:-1].shape)
    got, _ = call(prog, 1, O.BC["briowu"], o, s, u)
    ref = np.array(u, copy=True)
    o.apply_bc(ref)
    for sl in ((slice(None),) * 3 + (0, slice(1, -1)), (slice(None),) * 3 + (-1, slice(1, -1)),
               (slice(None),) * 3 + (slice(1, -1), 0), (slice(None),) * 3 + (slice(1, -1), -1)):
        assert np.array_equal(got[sl], ref[sl])
    o, s = make("wave1d", 24, 1, p, py=0)
    o.init_problem("wave1d")
    u = np.array(o.u, copy=True)
    _, e = call(prog, 3, 0, o, s, u, t=0.05)
    ref = o.error_smooth1d(0.05, u)
    assert abs(e[0] - ref[0]) <= 1e-14 * ref[0] and e[1] == ref[1]
