! Gauss-Lobatto tables (drop-in for gauss_lobatto / gauss_lobatto_matrix /
! gauss_legendre_matrix, gauss.f90:17-71,134-331,436-475). The nodal tables
! come from the library, which computes them in extended precision; the
! nodal mass matrix is M = Pnm^T diag(2/(2j+1)) Pnm (exact for products of
! degree-p Lagrange polynomials) and Minv its inverse M^-1 = Pmn diag((2j+1)/2) Pmn^T.
module gauss_m
	use ppesdg_capi
	implicit none
contains
	subroutine gauss_lobatto(n, x, w, interval)
		integer, intent(in) :: n
		real, intent(out) :: x(n), w(n)
		real, intent(in), optional :: interval(2)
		real :: D(n,n), Pmn(n,n), Pnm(n,n)
		if (n < 1 .or. n > 7) error stop "gauss_lobatto: n must be in [1,7], otherwise not implemented."
		if (ppes_tables(n-1, x, w, D, Pmn, Pnm) /= PPES_OK) error stop 'gauss_lobatto: table failure'
		if (present(interval)) then
			x = (0.5*(interval(2)-interval(1))) * x + 0.5*(interval(1)+interval(2))
			w = (0.5*(interval(2)-interval(1))) * w
		end if
	end subroutine gauss_lobatto

	subroutine gauss_lobatto_matrix(p, M, Minv, D, Pmn, Pnm)
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
end module gauss_m
