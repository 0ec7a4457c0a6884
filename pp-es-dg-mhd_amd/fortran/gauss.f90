! Gauss-Lobatto tables and the polynomial helpers of the reference's gauss_m
! (drop-in for gauss.f90:17-557). The nodal tables come from the library,
! which computes them in extended precision; the nodal mass matrix is
! M = Pnm^T diag(2/(2j+1)) Pnm (exact for products of degree-p Lagrange
! polynomials) and Minv = Pmn diag((2j+1)/2) Pmn^T its inverse. The
! self-tests (gauss_lobatto_test, gauss_lobatto_test_legendre,
! test_lobatto_mat, test_legendre_mat) measure the same identities as the
! reference's and print in the same form.
module gauss_m
	use ppesdg_capi
	implicit none
contains
	! n Gauss-Lobatto nodes and weights on [-1, 1] or on `interval` (gauss.f90:17-71)
	pure subroutine gauss_lobatto(n, x, w, interval)
		integer, intent(in) :: n
		real, intent(out) :: x(n), w(n)
		real, intent(in), optional :: interval(2)
		real :: D(n,n), Pmn(n,n), Pnm(n,n), half
		if (n < 1 .or. n > 7) error stop "gauss_lobatto: n must be in [1,7], otherwise not implemented."
		if (ppes_tables(n-1, x, w, D, Pmn, Pnm) /= PPES_OK) error stop 'gauss_lobatto: table failure'
		if (present(interval)) then
			half = 0.5*(interval(2)-interval(1))
			x = half * x + 0.5*(interval(1)+interval(2))
			w = half * w
		end if
	end subroutine gauss_lobatto

	! sum over n = 1..7 and monomial degrees d <= 2n-3 of the quadrature's
	! relative error |int x^d - sum w x^d| / (|int x^d| + 1) (gauss.f90:73-105)
	pure function gauss_lobatto_test(interval) result(out)
		real, intent(in), optional :: interval(2)
		real :: out
		real :: ab(2), exact
		real, allocatable :: x(:), w(:)
		integer :: n, d
		ab = [-1.0, 1.0]
		if (present(interval)) ab = interval
		out = 0.0
		do n = 1, 7
			allocate(x(n), w(n))
			call gauss_lobatto(n, x, w, ab)
			do d = 0, 2*n-3
				exact = (ab(2)**(d+1) - ab(1)**(d+1)) / (d+1)
				out = out + abs(exact - sum(w * x**d)) / (abs(exact) + 1.0)
			end do
			deallocate(x, w)
		end do
	end function gauss_lobatto_test

	! the same with Legendre polynomials: int P_d = 2 for d = 0, else 0 (gauss.f90:107-132)
	pure function gauss_lobatto_test_legendre() result(out)
		real :: out
		real :: exact
		real, allocatable :: x(:), w(:), pd(:)
		integer :: n, d
		out = 0.0
		do n = 1, 7
			allocate(x(n), w(n), pd(n))
			call gauss_lobatto(n, x, w)
			do d = 0, 2*n-3
				exact = merge(2.0, 0.0, d == 0)
				call legendre(d, x, pd)
				out = out + abs(exact - sum(w * pd)) / (abs(exact) + 1.0)
			end do
			deallocate(x, w, pd)
		end do
	end function gauss_lobatto_test_legendre

	! nodal mass matrix and inverse, differentiation matrix, and the
	! Legendre evaluation / interpolation pair of degree p (gauss.f90:134-331)
	pure subroutine gauss_lobatto_matrix(p, M, Minv, D, Pmn, Pnm)
		integer, intent(in) :: p
		real, dimension(p+1,p+1), intent(out) :: M, Minv, D, Pmn, Pnm
		real :: x(p+1), w(p+1), gm(p+1)
		integer :: j
		if (p < 0 .or. p > 6) error stop "gauss_lobatto_prep_table: p must be in [0,6], otherwise not implemented."
		if (ppes_tables(p, x, w, D, Pmn, Pnm) /= PPES_OK) error stop 'gauss_lobatto_matrix: table failure'
		do j = 1, p+1
			gm(j) = 2.0/(2*j-1)
		end do
		M = matmul(transpose(Pnm), matmul(diag(gm), Pnm))
		Minv = matmul(Pmn, matmul(diag(1.0/gm), transpose(Pmn)))
	contains
		pure function diag(v) result(a)
			real, intent(in) :: v(:)
			real :: a(size(v), size(v))
			integer :: i
			a = 0.0
			do i = 1, size(v)
				a(i,i) = v(i)
			end do
		end function diag
	end subroutine gauss_lobatto_matrix

	! n x n identity (gauss.f90:333-342)
	pure function eye(n) result(ans)
		integer, intent(in) :: n
		real, dimension(n,n) :: ans
		integer :: i
		ans = 0.0
		do i = 1, n
			ans(i,i) = 1.0
		end do
	end function eye

	! ans = mat**n by repeated squaring, ans = I for n <= 0 (gauss.f90:344-363)
	pure subroutine mat_fast_pow(mat, n, ans)
		real, dimension(:,:), intent(in) :: mat
		integer, value :: n
		real, dimension(size(mat,1), size(mat,2)), intent(out) :: ans
		real, dimension(size(mat,1), size(mat,2)) :: sq
		ans = eye(size(mat,1))
		sq = mat
		do while (n > 0)
			if (btest(n, 0)) ans = matmul(sq, ans)
			n = ishft(n, -1)
			if (n > 0) sq = matmul(sq, sq)
		end do
	end subroutine mat_fast_pow

	! identities of the nodal tables (gauss.f90:365-388): Pmn Pnm = I,
	! M Minv = I, D nilpotent of order p+1, and SBP M D + D^T M = diag(-1, 0, .., 0, 1);
	! prints the largest violation and the SBP one
	subroutine test_lobatto_mat(p, x, w, M, Minv, D, Pmn, Pnm)
		use, intrinsic :: iso_fortran_env, only: stdout => output_unit
		integer, intent(in) :: p
		real, dimension(0:p), intent(in) :: x, w
		real, dimension(0:p, 0:p), intent(in) :: M, Minv, D, Pmn, Pnm
		real, dimension(0:p, 0:p) :: Dp, sbp
		real :: err
		err = maxval(abs(matmul(Pmn, Pnm) - eye(p+1)))
		err = max(err, maxval(abs(matmul(M, Minv) - eye(p+1))))
		call mat_fast_pow(D, p+1, Dp)
		err = max(err, maxval(abs(Dp)))
		sbp = matmul(M, D) + matmul(transpose(D), M)
		sbp(0,0) = sbp(0,0) + 1.0
		sbp(p,p) = sbp(p,p) - 1.0
		err = max(err, maxval(abs(sbp)))
		write(stdout,*) err, maxval(abs(sbp))
	end subroutine test_lobatto_mat

	! Legendre polynomial P_n(x) by the three-term recurrence and, if asked,
	! P_n'(x) from P_k' = P_(k-2)' + (2k-1) P_(k-1) (gauss.f90:390-434)
	pure elemental subroutine legendre(n, x, y, dy)
		integer, intent(in) :: n
		real, intent(in) :: x
		real, intent(out) :: y
		real, intent(out), optional :: dy
		real :: pkm1, pkm2, dkm1, dkm2, dk
		integer :: k
		if (n < 0) then
			y = 0.0
			if (present(dy)) dy = 0.0
			return
		end if
		pkm1 = 1.0          ! P_0
		dkm2 = 0.0          ! P_(-1)'
		dkm1 = 0.0          ! P_0'
		if (n == 0) then
			y = pkm1
			if (present(dy)) dy = dkm1
			return
		end if
		pkm2 = pkm1
		pkm1 = x            ! P_1
		dk = 1.0            ! P_1'
		dkm2 = dkm1
		dkm1 = dk
		do k = 2, n
			y = ((2*k-1)*x*pkm1 - (k-1)*pkm2) / k
			dk = dkm2 + (2*k-1)*pkm1
			pkm2 = pkm1
			pkm1 = y
			dkm2 = dkm1
			dkm1 = dk
		end do
		y = pkm1
		if (present(dy)) dy = dkm1
	end subroutine legendre

	! modal Legendre tables of degree p (gauss.f90:436-475): mass m_j = 1/(j+1/2),
	! its inverse, traces P_j(-1) and P_j(1), D(i,j) = 2i+1 for j-i odd
	pure subroutine gauss_legendre_matrix(p, m, minv, D, tl, tr)
		integer, intent(in) :: p
		real, dimension(0:p), intent(out) :: m, minv, tl, tr
		real, dimension(0:p,0:p), intent(out) :: D
		integer :: i, j
		D = 0.0
		do j = 0, p
			m(j) = 1.0 / (j + 0.5)
			minv(j) = j + 0.5
			tr(j) = 1.0
			tl(j) = merge(1.0, -1.0, mod(j,2) == 0)
			do i = 0, j-1
				if (mod(j-i, 2) == 1) D(i,j) = 2.0*i + 1.0
			end do
		end do
	end subroutine gauss_legendre_matrix

	! identities of the modal tables (gauss.f90:477-502): D nilpotent of
	! order p+1 and SBP m_i D(i,j) + D(j,i) m_j = tr_i tr_j - tl_i tl_j
	subroutine test_legendre_mat(p, m, D, tl, tr)
		use, intrinsic :: iso_fortran_env, only: stdout => output_unit
		integer, intent(in) :: p
		real, dimension(0:p), intent(in) :: m, tl, tr
		real, dimension(0:p, 0:p), intent(in) :: D
		real, dimension(0:p, 0:p) :: Dp, sbp
		real :: err
		integer :: i, j
		call mat_fast_pow(D, p+1, Dp)
		err = maxval(abs(Dp))
		do j = 0, p
			do i = 0, p
				sbp(i,j) = m(i)*D(i,j) + D(j,i)*m(j) - (tr(i)*tr(j) - tl(i)*tl(j))
			end do
		end do
		err = max(err, maxval(abs(sbp)))
		write(stdout,*) err, maxval(abs(sbp))
	end subroutine test_legendre_mat

	! Lagrange interpolation from nodes x0 to points x1: mat(i,j) = l_j(x1(i))
	! (gauss.f90:504-522)
	pure function interpolation_matrix(x0, x1) result(mat)
		real, dimension(:), intent(in) :: x0, x1
		real, dimension(size(x1), size(x0)) :: mat
		integer :: i, j, k
		do j = 1, size(x0)
			do i = 1, size(x1)
				mat(i,j) = 1.0
				do k = 1, size(x0)
					if (k /= j) mat(i,j) = mat(i,j) * ((x1(i) - x0(k)) / (x0(j) - x0(k)))
				end do
			end do
		end do
	end function interpolation_matrix

	! sum_k list(k) x^(k-1) by Horner's rule (gauss.f90:524-537)
	pure function Horner_Taylor(list, x) result(ans)
		real, dimension(:), intent(in) :: list
		real, intent(in) :: x
		real :: ans
		integer :: k
		ans = 0.0
		do k = size(list), 1, -1
			ans = ans*x + list(k)
		end do
	end function Horner_Taylor

	! sum_k list(k) P_(k-1)(x) by Clenshaw's recurrence for the Legendre
	! three-term relation (gauss.f90:539-557): b_j = c_j + (2j+1)/(j+1) x b_(j+1)
	! - (j+1)/(j+2) b_(j+2), the sum is b_0
	pure function Clenshaw_Legendre(list, x) result(b0)
		real, dimension(:), intent(in) :: list
		real, intent(in) :: x
		real :: b0
		real :: b1, b2
		integer :: j
		b0 = 0.0
		b1 = 0.0
		do j = size(list)-1, 0, -1
			b2 = b1
			b1 = b0
			b0 = list(j+1) + (real(2*j+1)/(j+1))*x*b1 - (real(j+1)/(j+2))*b2
		end do
	end function Clenshaw_Legendre
end module gauss_m
