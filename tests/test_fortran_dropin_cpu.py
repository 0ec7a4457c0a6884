"""The Fortran drop-in modules as a reference caller sees them (host only).

* The reference's own program, main.f90, compiles and links against the
  drop-in modules (pp-es-dg-mhd_amd/_build/fmod, lib/libppesdg_f.a,
  lib/libppesdg.so) in place of the reference's modules. It is not run (it
  would start a GPU solve); skipped where the reference sources are absent.
* Pure callers compile: gauss_lobatto, gauss_lobatto_matrix,
  LDF_projector%apply / %div and the initial-data functions are pure, as in
  gauss.f90:17,134, LDF.f90:219,252 and problems.f90:74-259.
* The gauss_m helpers give the mathematics' values: the quadrature
  self-tests vanish to round-off, legendre / Clenshaw_Legendre / Horner_Taylor
  / interpolation_matrix / mat_fast_pow agree with numpy, and the table
  identity checks (test_lobatto_mat, test_legendre_mat) print round-off.
"""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "pp-es-dg-mhd_amd")
FMOD = os.path.join(PKG, "_build", "fmod")
LIBDIR = os.path.join(PKG, "lib")
FLANG = "/opt/rocm/lib/llvm/bin/flang"
REF_MAIN = os.path.join(os.environ.get("PPES_REFERENCE", "/root/reference"), "main.f90")
FFLAGS = ["-fdefault-real-8", "-cpp", "-O2", "-ffp-contract=off"]


def need_build():
    if not os.path.exists(FLANG):
        pytest.skip("flang not available")
    for f in (os.path.join(LIBDIR, "libppesdg_f.a"), os.path.join(LIBDIR, "libppesdg.so"),
              os.path.join(FMOD, "gauss_m.mod")):
        if not os.path.exists(f):
            pytest.skip(f"{f} not built (make -C pp-es-dg-mhd_amd)")


def link(d, obj, exe):
    subprocess.check_call([FLANG, "-o", exe, obj, os.path.join(LIBDIR, "libppesdg_f.a"), "-L" + LIBDIR, "-lppesdg",
                           "-Wl,-rpath," + LIBDIR], cwd=d)


def test_reference_main_builds_against_dropins(tmp_path):
    need_build()
    if not os.path.exists(REF_MAIN):
        pytest.skip("reference sources not present")
    subprocess.check_call([FLANG] + FFLAGS + ["-I", FMOD, "-c", REF_MAIN, "-o", "main.o"], cwd=tmp_path)
    link(tmp_path, "main.o", "main")
    assert os.path.getsize(tmp_path / "main") > 0


PROGRAM = """
module pure_callers
	use gauss_m
	use LDF_m
	use problems
	implicit none
contains
	pure function quad_sum(n) result(s)
		integer, intent(in) :: n
		real :: s, x(n), w(n)
		call gauss_lobatto(n, x, w)
		s = sum(w)
	end function quad_sum
	pure function mass_trace(p) result(s)
		integer, intent(in) :: p
		real :: s
		real, dimension(p+1,p+1) :: M, Minv, D, Pmn, Pnm
		call gauss_lobatto_matrix(p, M, Minv, D, Pmn, Pnm)
		s = sum(M)
	end function mass_trace
	pure subroutine project(proj, uu)
		type(LDF_projector), intent(in) :: proj
		real, intent(inout) :: uu(4,4,2)
		call proj%apply(uu)
	end subroutine project
	pure function divergence(proj, uu) result(d)
		type(LDF_projector), intent(in) :: proj
		real, intent(in) :: uu(4,4,2)
		real :: d(4,4)
		call proj%div(uu, d)
	end function divergence
	pure function sample(x, y) result(u)
		real, intent(in) :: x, y
		real :: u(8)
		u = OrszagTang_init_pri(x, y) + Rotor_init_pri(x, y) + Blast_init_pri(x, y) + smooth_alfven_init_fun(x, y) &
			+ KHI_init_pri(x, y) + Brio_Wu_init_fun(x, y) + smooth_1Dwave_init_fun(x, y) + cloud_shock_init_pri(x, y) &
			+ jet1_init_pri(x, y) + YeeSjogreenInitCon(x, y)
	end function sample
end module pure_callers

program dropin_check
	use pure_callers
	use utility, only: omp_para_output
	implicit none
	type(LDF_projector) :: proj
	real :: uu(4,4,2), v(4,4,2), A(3,3), P5(3,3), y, dy, xs(4), x0(4), x1(5), Im(5,4)
	real :: c(5), m(0:3), minv(0:3), tl(0:3), tr(0:3), Dm(0:3,0:3), xg(0:3), wg(0:3)
	real, dimension(0:3,0:3) :: M4, Mi4, D4, Pmn4, Pnm4
	integer :: n, k, i
	write(*,'(A,3ES25.16)') 'selftest ', gauss_lobatto_test(), gauss_lobatto_test([0.0, 2.0]), &
		gauss_lobatto_test_legendre()
	do n = 1, 7
		write(*,'(A,I3,2ES25.16)') 'quad ', n, quad_sum(n), mass_trace(n-1)
	end do
	xs = [-1.0, -0.3, 0.2, 1.0]
	do n = 0, 6
		do k = 1, 4
			call legendre(n, xs(k), y, dy)
			write(*,'(A,I3,3ES25.16)') 'legendre ', n, xs(k), y, dy
		end do
	end do
	c = [0.5, -1.25, 2.0, 0.75, -0.3]
	do k = 1, 4
		write(*,'(A,3ES25.16)') 'series ', xs(k), Clenshaw_Legendre(c, xs(k)), Horner_Taylor(c, xs(k))
	end do
	x0 = [-1.0, -0.2, 0.5, 1.0]
	x1 = [-0.9, -0.2, 0.1, 0.7, 1.0]
	Im = interpolation_matrix(x0, x1)
	write(*,'(A,20ES25.16)') 'interp ', Im
	A = reshape([0.5, -0.2, 0.1, 0.3, 0.9, -0.4, 0.0, 0.25, 0.6], [3,3])
	call mat_fast_pow(A, 5, P5)
	write(*,'(A,9ES25.16)') 'pow5 ', P5
	write(*,'(A,9ES25.16)') 'eye ', eye(3)
	! LDF: the projection is idempotent and its result divergence-free
	call proj%init(3, 3)
	do k = 1, 2
		do i = 1, 4
			do n = 1, 4
				uu(n,i,k) = sin(1.0*n + 2.0*i + 3.0*k)
			end do
		end do
	end do
	call project(proj, uu)
	v = uu
	call project(proj, v)
	write(*,'(A,2ES25.16)') 'ldf ', maxval(abs(v - uu)), maxval(abs(divergence(proj, uu)))
	write(*,'(A,8ES25.16)') 'pure_sample ', sample(0.3, 0.6)
	! the reference's public boundary data is assignable
	u_pri_L = [1.0, 2.0, 3.0, 4.0]
	u_pri_R = 0.5 * u_pri_L
	BC_pri_fun => Brio_Wu_init_fun
	write(*,'(A,L2,4ES25.16)') 'bcdata ', associated(BC_pri_fun), u_pri_R
	! table identity checks print (error, SBP error)
	call gauss_lobatto(4, xg, wg)
	call gauss_lobatto_matrix(3, M4, Mi4, D4, Pmn4, Pnm4)
	write(*,'(A)', advance='no') 'lobatto_mat '
	call test_lobatto_mat(3, xg, wg, M4, Mi4, D4, Pmn4, Pnm4)
	call gauss_legendre_matrix(3, m, minv, Dm, tl, tr)
	write(*,'(A)', advance='no') 'legendre_mat '
	call test_legendre_mat(3, m, Dm, tl, tr)
	open(newunit=k, file='omp.txt', status='replace')
	call omp_para_output(k)
	close(k)
end program dropin_check
"""


@pytest.fixture(scope="module")
def run(tmp_path_factory):
    need_build()
    d = tmp_path_factory.mktemp("dropin")
    (d / "check.f90").write_text(PROGRAM)
    subprocess.check_call([FLANG] + FFLAGS + ["-I", FMOD, "-c", "check.f90", "-o", "check.o"], cwd=d)
    link(d, "check.o", "check")
    out = subprocess.run([str(d / "check")], cwd=d, check=True, capture_output=True, text=True).stdout
    rows = {}
    for line in out.splitlines():
        key, *vals = line.split()
        rows.setdefault(key, []).append(vals)
    return rows, d


def nums(vals):
    return np.array([float(v) for v in vals])


def test_quadrature_selftests_vanish(run):
    rows, _ = run
    st = nums(rows["selftest"][0])
    assert np.all(st < 1e-13), st
    for vals in rows["quad"]:
        n, s, mt = int(vals[0]), float(vals[1]), float(vals[2])
        assert abs(s - 2.0) < 1e-14, (n, s)
        # sum of the nodal mass matrix = integral of (sum of Lagrange basis)^2 = 2
        assert abs(mt - 2.0) < 1e-13, (n, mt)


def test_legendre_values_and_derivatives(run):
    rows, _ = run
    L = np.polynomial.legendre
    for vals in rows["legendre"]:
        n, x, y, dy = int(vals[0]), float(vals[1]), float(vals[2]), float(vals[3])
        c = np.zeros(n + 1)
        c[n] = 1.0
        assert abs(y - L.legval(x, c)) < 1e-14, (n, x)
        assert abs(dy - L.legval(x, L.legder(c))) < 1e-13 * max(1.0, n * n), (n, x)


def test_series_interpolation_and_powers(run):
    rows, _ = run
    c = np.array([0.5, -1.25, 2.0, 0.75, -0.3])
    for vals in rows["series"]:
        x, cl, ho = nums(vals)
        assert abs(cl - np.polynomial.legendre.legval(x, c)) < 1e-14
        assert abs(ho - np.polynomial.polynomial.polyval(x, c)) < 1e-14
    x0 = np.array([-1.0, -0.2, 0.5, 1.0])
    x1 = np.array([-0.9, -0.2, 0.1, 0.7, 1.0])
    Im = nums(rows["interp"][0]).reshape((5, 4), order="F")
    for j in range(4):
        others = np.delete(x0, j)
        expect = np.prod([(x1 - xk) / (x0[j] - xk) for xk in others], axis=0)
        assert np.max(np.abs(Im[:, j] - expect)) < 1e-15
    A = np.array([0.5, -0.2, 0.1, 0.3, 0.9, -0.4, 0.0, 0.25, 0.6]).reshape((3, 3), order="F")
    P5 = nums(rows["pow5"][0]).reshape((3, 3), order="F")
    assert np.max(np.abs(P5 - np.linalg.matrix_power(A, 5))) < 1e-15
    assert np.array_equal(nums(rows["eye"][0]).reshape(3, 3), np.eye(3))


def test_ldf_apply_pure_idempotent_divergence_free(run):
    rows, _ = run
    idem, div = nums(rows["ldf"][0])
    assert idem < 1e-14 and div < 1e-13


def test_pure_initial_data_and_boundary_data(run):
    rows, d = run
    assert np.all(np.isfinite(nums(rows["pure_sample"][0])))
    vals = rows["bcdata"][0]
    assert vals[0] == "T" and np.array_equal(nums(vals[1:]), [0.5, 1.0, 1.5, 2.0])
    assert "OpenMP" in (d / "omp.txt").read_text() or "OMP_" in (d / "omp.txt").read_text()


def test_table_identity_checks_print_roundoff(run):
    rows, _ = run
    for key in ("lobatto_mat", "legendre_mat"):
        err, sbp = nums(rows[key][0])
        assert 0.0 <= err < 1e-12 and 0.0 <= sbp < 1e-12, (key, err, sbp)
