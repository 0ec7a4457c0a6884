! Problem set-ups (drop-in for the plug-ins of problems.f90 used on the hot
! path): nodal point-sampling initialisation (init_by_fun_pri / _con,
! problems.f90:25-72), initial-data functions, and the boundary fillers
! BC_periodic, BC_YeeSjogreen and BC_rotor (problems.f90:312-348,386-423,
! 597-654). drive_m recognises these three fillers and runs them on the
! device; any other BC_handle runs on the host between device calls.
module problems
	use utility, only: pi => PI_math
	implicit none

	abstract interface
		function fun_type(x, y) result(u_pri)
			real, intent(in) :: x, y
			real, dimension(8) :: u_pri
		end function fun_type
	end interface

	procedure(fun_type), pointer, public :: init_pri_fun => null()
	procedure(fun_type), pointer, public :: init_con_fun => null()
contains

	subroutine init_by_fun_pri(u_con, xx, yy, nx, ny, px, py)
		use MHD_PDE_m, only: MHD_pri2con
		integer, intent(in) :: nx, ny, px, py
		real, intent(in) :: xx(px+1, nx), yy(py+1, ny)
		real, intent(inout) :: u_con(8, px+1, py+1, 0:nx+1, 0:ny+1)
		integer :: i, j, k, l
		if (.not. associated(init_pri_fun)) error stop "init_by_fun_pri: required function pointer 'init_pri_fun' not initialized!"
		do j = 1, ny
			do i = 1, nx
				do l = 1, py+1
					do k = 1, px+1
						u_con(:,k,l,i,j) = MHD_pri2con(init_pri_fun(xx(k,i), yy(l,j)))
					end do
				end do
			end do
		end do
	end subroutine init_by_fun_pri

	subroutine init_by_fun_con(u_con, xx, yy, nx, ny, px, py)
		integer, intent(in) :: nx, ny, px, py
		real, intent(in) :: xx(px+1, nx), yy(py+1, ny)
		real, intent(inout) :: u_con(8, px+1, py+1, 0:nx+1, 0:ny+1)
		integer :: i, j, k, l
		if (.not. associated(init_con_fun)) error stop "init_by_fun_con: required function pointer 'init_con_fun' not initialized!"
		do j = 1, ny
			do i = 1, nx
				do l = 1, py+1
					do k = 1, px+1
						u_con(:,k,l,i,j) = init_con_fun(xx(k,i), yy(l,j))
					end do
				end do
			end do
		end do
	end subroutine init_by_fun_con

	! Orszag-Tang vortex on [0,1]^2 (problems.f90:155-168)
	function OrszagTang_init_pri(x, y) result(u_pri)
		real, intent(in) :: x, y
		real, dimension(8) :: u_pri
		u_pri = [25.0/(36.0*pi), -sin(2*pi*y), sin(2*pi*x), 0.0, -sin(2*pi*y)/sqrt(4.0*pi), &
			sin(4*pi*x)/sqrt(4.0*pi), 0.0, 5.0/(12.0*pi)]
	end function OrszagTang_init_pri

	! Orszag-Tang vortex on [0,2pi]^2 with gamma from MHD_PDE_m: rho = gamma^2,
	! p = gamma, u = (-sin y, sin x, 0), B = (-sin y, sin 2x, 0)
	function OrszagTang2pi_init_pri(x, y) result(u_pri)
		use MHD_PDE_m, only: gamma
		real, intent(in) :: x, y
		real, dimension(8) :: u_pri
		u_pri = [gamma**2, -sin(y), sin(x), 0.0, -sin(y), sin(2.0*x), 0.0, gamma]
	end function OrszagTang2pi_init_pri

	! rotor (problems.f90:170-199)
	function Rotor_init_pri(x, y) result(u_pri)
		real, intent(in) :: x, y
		real, dimension(8) :: u_pri
		real, parameter :: r0 = 0.1, r1 = 0.115, u0 = 2.0
		real :: r, f
		r = sqrt((x-0.5)**2 + (y-0.5)**2)
		f = (r1-r)/(r1-r0)
		u_pri = 0.0
		if (r < r0) then
			u_pri(1:3) = [10.0, -(u0/r0)*(y-0.5), (u0/r0)*(x-0.5)]
		else if (r < r1) then
			u_pri(1:3) = [1.0 + 9*f, -(u0/r0)*(y-0.5)*f, (u0/r0)*(x-0.5)*f]
		else
			u_pri(1:3) = [1.0, 0.0, 0.0]
		end if
		u_pri(5) = 5.0/sqrt(4*pi)
		u_pri(8) = 1.0
	end function Rotor_init_pri

	! blast wave on [0,1]^2 (problems.f90:201-218)
	function Blast_init_pri(x, y) result(u_pri)
		real, intent(in) :: x, y
		real, dimension(8) :: u_pri
		u_pri = [1.0, 0.0, 0.0, 0.0, 100.0/sqrt(4*pi), 0.0, 0.0, 0.1]
		if ((x-0.5)**2 + (y-0.5)**2 < 0.01) u_pri(8) = 1000.0
	end function Blast_init_pri

	! blast wave on [-0.5,0.5]^2 with a diagonal field B = (100/sqrt(4pi))(1,1,0)/sqrt(2)
	function BlastDiag_init_pri(x, y) result(u_pri)
		real, intent(in) :: x, y
		real, dimension(8) :: u_pri
		real :: b0
		b0 = (100.0/sqrt(4.0*pi))/sqrt(2.0)
		u_pri = [1.0, 0.0, 0.0, 0.0, b0, b0, 0.0, 0.1]
		if (x**2 + y**2 < 0.01) u_pri(8) = 1000.0
	end function BlastDiag_init_pri

	! ------------------------------------------------------------ boundaries

	subroutine BC_periodic(u_con, xx, yy, nx, ny, px, py)
		integer, intent(in) :: nx, ny, px, py
		real, intent(in) :: xx(px+1, nx), yy(py+1, ny)
		real, intent(inout) :: u_con(8, px+1, py+1, 0:nx+1, 0:ny+1)
		u_con(:,:,:,0,1:ny) = u_con(:,:,:,nx,1:ny)
		u_con(:,:,:,nx+1,1:ny) = u_con(:,:,:,1,1:ny)
		u_con(:,:,:,1:nx,0) = u_con(:,:,:,1:nx,ny)
		u_con(:,:,:,1:nx,ny+1) = u_con(:,:,:,1:nx,1)
	end subroutine BC_periodic

	! zero-gradient extrapolation of the boundary trace
	subroutine BC_YeeSjogreen(u_con, xx, yy, nx, ny, px, py)
		integer, intent(in) :: nx, ny, px, py
		real, intent(in) :: xx(px+1, nx), yy(py+1, ny)
		real, intent(inout) :: u_con(8, px+1, py+1, 0:nx+1, 0:ny+1)
		integer :: k, l
		do l = 1, py+1
			do k = 1, px+1
				u_con(:,k,l,0,1:ny) = u_con(:,1,l,1,1:ny)
				u_con(:,k,l,nx+1,1:ny) = u_con(:,px+1,l,nx,1:ny)
				u_con(:,k,l,1:nx,0) = u_con(:,k,1,1:nx,1)
				u_con(:,k,l,1:nx,ny+1) = u_con(:,k,py+1,1:nx,ny)
			end do
		end do
	end subroutine BC_YeeSjogreen

	! the rotor's fixed ambient state
	function rotor_ghost_state() result(u_con_)
		use MHD_PDE_m, only: MHD_pri2con
		real, dimension(8) :: u_con_
		u_con_ = MHD_pri2con([1.0, 0.0, 0.0, 0.0, 5.0/sqrt(4*pi), 0.0, 0.0, 1.0])
	end function rotor_ghost_state

	subroutine BC_rotor(u_con, xx, yy, nx, ny, px, py)
		integer, intent(in) :: nx, ny, px, py
		real, intent(in) :: xx(px+1, nx), yy(py+1, ny)
		real, intent(inout) :: u_con(8, px+1, py+1, 0:nx+1, 0:ny+1)
		real :: s(8)
		integer :: i, j, k, l
		s = rotor_ghost_state()
		do l = 1, py+1
			do k = 1, px+1
				do j = 1, ny
					u_con(:,k,l,0,j) = s
					u_con(:,k,l,nx+1,j) = s
				end do
				do i = 1, nx
					u_con(:,k,l,i,0) = s
					u_con(:,k,l,i,ny+1) = s
				end do
			end do
		end do
	end subroutine BC_rotor
end module problems
