! Locally divergence-free projector on Legendre modal coefficients of
! (Bx, By) on the reference element (drop-in for type(LDF_projector),
! LDF.f90:94-269; apply and div are pure, as there). The library supplies the projector as one dense modal
! operator (the unique M-orthogonal projection onto divergence-free Q_p^2).
module LDF_m
	use ppesdg_capi
	implicit none
	private
	type, public :: LDF_projector
		private
		logical :: ready = .false.
		integer :: px = -1, py = -1
		real, allocatable :: op(:,:)
	contains
		procedure, pass :: init
		procedure, pass :: apply
		procedure, pass :: div
		final :: destroy
	end type LDF_projector
contains
	subroutine init(this, px, py)
		class(LDF_projector), intent(inout) :: this
		integer, intent(in) :: px, py
		integer :: m
		if (px /= py) error stop 'LDF_projector: px /= py is not supported'
		if (px < 1 .or. px > 6) error stop 'LDF_projector: error para in init'
		this%px = px
		this%py = py
		m = 2*(px+1)*(py+1)
		if (allocated(this%op)) deallocate(this%op)
		allocate(this%op(m, m))
		if (ppes_ldf_modal_operator(px, this%op) /= PPES_OK) error stop 'LDF_projector: init failed'
		this%ready = .true.
	end subroutine init

	pure subroutine apply(this, uu)
		class(LDF_projector), intent(in) :: this
		real, dimension(this%px+1, this%py+1, 2), intent(inout) :: uu
		if (.not. this%ready) error stop 'LDF_projector: not ready!'
		uu = reshape(matmul(this%op, reshape(uu, [size(uu)])), shape(uu))
	end subroutine apply

	pure subroutine div(this, uu, udiv)
		use gauss_m, only: gauss_legendre_matrix
		class(LDF_projector), intent(in) :: this
		real, dimension(this%px+1, this%py+1, 2), intent(in) :: uu
		real, dimension(this%px+1, this%py+1), intent(out) :: udiv
		real :: m(this%px+1), minv(this%px+1), tl(this%px+1), tr(this%px+1), Dm(this%px+1, this%px+1)
		if (.not. this%ready) error stop 'LDF_projector: not ready!'
		call gauss_legendre_matrix(this%px, m, minv, Dm, tl, tr)
		udiv = matmul(Dm, uu(:,:,1)) + matmul(uu(:,:,2), transpose(Dm))
	end subroutine div

	! release the operator (LDF.f90:262-269)
	subroutine destroy(this)
		type(LDF_projector), intent(inout) :: this
		if (allocated(this%op)) deallocate(this%op)
		this%ready = .false.
	end subroutine destroy
end module LDF_m
