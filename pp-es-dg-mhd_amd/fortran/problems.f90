! Problem set-ups (drop-in for problems.f90): nodal point-sampling
! initialisation (init_by_fun_pri / _con, problems.f90:25-72), the shipped
! initial-data functions (:74-308), the boundary fillers BC_periodic,
! BC_BrioWu, BC_YeeSjogreen, BC_cloud_shock, BC_jet1_cell_ext and BC_rotor
! (:312-654) and the error functionals of the smooth problems (:659-742).
! drive_m recognises the shipped fillers and runs them on the device; any
! other BC_handle runs on the host between device calls.
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

	! left and right primitive states of a fixed-state boundary, and the
	! primitive function of a function-defined boundary (problems.f90:16-20):
	! public for programs that set them; no shipped filler reads them
	real, dimension(4), public :: u_pri_L, u_pri_R
	procedure(fun_type), pointer, public :: BC_pri_fun => null()
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
	pure function OrszagTang_init_pri(x, y) result(u_pri)
		real, intent(in) :: x, y
		real, dimension(8) :: u_pri
		u_pri = [25.0/(36.0*pi), -sin(2*pi*y), sin(2*pi*x), 0.0, -sin(2*pi*y)/sqrt(4.0*pi), &
			sin(4*pi*x)/sqrt(4.0*pi), 0.0, 5.0/(12.0*pi)]
	end function OrszagTang_init_pri

	! Orszag-Tang vortex on [0,2pi]^2 with gamma from MHD_PDE_m: rho = gamma^2,
	! p = gamma, u = (-sin y, sin x, 0), B = (-sin y, sin 2x, 0)
	pure function OrszagTang2pi_init_pri(x, y) result(u_pri)
		use MHD_PDE_m, only: gamma
		real, intent(in) :: x, y
		real, dimension(8) :: u_pri
		u_pri = [gamma**2, -sin(y), sin(x), 0.0, -sin(y), sin(2.0*x), 0.0, gamma]
	end function OrszagTang2pi_init_pri

	! rotor (problems.f90:170-199)
	pure function Rotor_init_pri(x, y) result(u_pri)
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
	pure function Blast_init_pri(x, y) result(u_pri)
		real, intent(in) :: x, y
		real, dimension(8) :: u_pri
		u_pri = [1.0, 0.0, 0.0, 0.0, 100.0/sqrt(4*pi), 0.0, 0.0, 0.1]
		if ((x-0.5)**2 + (y-0.5)**2 < 0.01) u_pri(8) = 1000.0
	end function Blast_init_pri

	! blast wave on [-0.5,0.5]^2 with a diagonal field B = (100/sqrt(4pi))(1,1,0)/sqrt(2)
	pure function BlastDiag_init_pri(x, y) result(u_pri)
		real, intent(in) :: x, y
		real, dimension(8) :: u_pri
		real :: b0
		b0 = (100.0/sqrt(4.0*pi))/sqrt(2.0)
		u_pri = [1.0, 0.0, 0.0, 0.0, b0, b0, 0.0, 0.1]
		if (x**2 + y**2 < 0.01) u_pri(8) = 1000.0
	end function BlastDiag_init_pri

	! smooth Alfven wave along the direction pi/6 (problems.f90:74-98)
	pure function smooth_alfven_init_fun(x, y) result(u_pri)
		real, intent(in) :: x, y
		real, dimension(8) :: u_pri
		real :: ca, sa, xp, vp, vz
		ca = cos(pi/6.0)
		sa = sin(pi/6.0)
		xp = x*ca + y*sa
		vp = 0.1*sin(2*pi*xp)
		vz = 0.1*cos(2*pi*xp)
		u_pri = [1.0, vp*(-sa), vp*ca, vz, 1.0*ca - vp*sa, 1.0*sa + vp*ca, vz, 0.1]
	end function smooth_alfven_init_fun

	! Kelvin-Helmholtz shear layers at y = +-0.5 (problems.f90:100-113)
	pure function KHI_init_pri(x, y) result(u_pri)
		real, intent(in) :: x, y
		real, dimension(8) :: u_pri
		u_pri = [1.0, 5.0*((tanh(20*(y + 0.5)) - tanh(20*(y - 0.5))) - 1.0), &
			(0.25*sin(2*pi*x))*(exp(-(100*(y + 0.5)**2)) - exp(-(100*(y - 0.5)**2))), 0.0, 1.0, 0.0, 0.0, 50.0]
	end function KHI_init_pri

	! Brio-Wu shock tube, discontinuity at x = 0 (problems.f90:115-138)
	pure function Brio_Wu_init_fun(x, y) result(u_pri)
		real, intent(in) :: x, y
		real, dimension(8) :: u_pri
		if (x < 0.0) then
			u_pri = [1.0, 0.0, 0.0, 0.0, 0.75, 1.0, 0.0, 1.0]
		else
			u_pri = [0.125, 0.0, 0.0, 0.0, 0.75, -1.0, 0.0, 0.1]
		end if
	end function Brio_Wu_init_fun

	! smooth density wave advected in x (problems.f90:140-153)
	pure function smooth_1Dwave_init_fun(x, y) result(u_pri)
		real, intent(in) :: x, y
		real, dimension(8) :: u_pri
		u_pri = [1.0 + 0.2*sin(x), 1.0, 0.0, 0.0, 0.5, 1.0, 1.5, 2.0]
	end function smooth_1Dwave_init_fun

	! shock (x < 0.6) hitting a dense cloud at (0.8, 0.5) (problems.f90:220-247)
	pure function cloud_shock_init_pri(x, y) result(u_pri)
		real, intent(in) :: x, y
		real, dimension(8) :: u_pri
		if (x < 0.6) then
			u_pri = [3.86859, 0.0, 0.0, 0.0, 0.0, 2.1826182, -2.1826182, 167.345]
		else
			u_pri = cloud_inflow_pri()
			if ((x - 0.8)**2 + (y - 0.5)**2 <= 0.15**2) u_pri(1) = 10.0
		end if
	end function cloud_shock_init_pri

	pure function cloud_inflow_pri() result(u_pri)
		real, dimension(8) :: u_pri
		u_pri = [1.0, -11.2536, 0.0, 0.0, 0.0, 0.56418958, 0.56418958, 1.0]
	end function cloud_inflow_pri

	! astrophysical jet: the ambient state (problems.f90:249-257)
	pure function jet1_init_pri(x, y) result(u_pri)
		use MHD_PDE_m, only: gamma
		real, intent(in) :: x, y
		real, dimension(8) :: u_pri
		u_pri = [0.1*gamma, 0.0, 0.0, 0.0, 0.0, sqrt(200.0), 0.0, 1.0]
	end function jet1_init_pri

	! 2D Riemann problem of Yee and Sjogreen, conservative (problems.f90:259-308)
	pure function YeeSjogreenInitCon(x, y) result(u_con)
		real, intent(in) :: x, y
		real, dimension(8) :: u_con
		if (y >= 0.0 .and. x >= 0.0) then
			u_con = [0.9308, 1.4557, -0.4633, 0.0575, 0.3501, 0.9830, 0.3050, 5.0838]
		else if (y >= 0.0) then
			u_con = [1.0304, 1.5774, -1.0455, -0.1016, 0.3501, 0.5078, 0.1576, 5.7813]
		else if (x <= 0.0) then
			u_con = [1.0000, 1.7500, -1.0000, 0.0000, 0.5642, 0.5078, 0.2539, 6.0000]
		else
			u_con = [1.8887, 0.2334, -1.7422, 0.0733, 0.5642, 0.9830, 0.4915, 12.999]
		end if
	end function YeeSjogreenInitCon

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
	pure function rotor_ghost_state() result(u_con_)
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

	! trace extrapolation in x, periodic in y
	subroutine BC_BrioWu(u_con, xx, yy, nx, ny, px, py)
		integer, intent(in) :: nx, ny, px, py
		real, intent(in) :: xx(px+1, nx), yy(py+1, ny)
		real, intent(inout) :: u_con(8, px+1, py+1, 0:nx+1, 0:ny+1)
		integer :: k
		do k = 1, px+1
			u_con(:,k,:,0,1:ny) = u_con(:,1,:,1,1:ny)
			u_con(:,k,:,nx+1,1:ny) = u_con(:,px+1,:,nx,1:ny)
		end do
		u_con(:,:,:,1:nx,0) = u_con(:,:,:,1:nx,ny)
		u_con(:,:,:,1:nx,ny+1) = u_con(:,:,:,1:nx,1)
	end subroutine BC_BrioWu

	! element copy on the left, bottom and top; the pre-shock inflow on the right
	subroutine BC_cloud_shock(u_con, xx, yy, nx, ny, px, py)
		use MHD_PDE_m, only: MHD_pri2con
		integer, intent(in) :: nx, ny, px, py
		real, intent(in) :: xx(px+1, nx), yy(py+1, ny)
		real, intent(inout) :: u_con(8, px+1, py+1, 0:nx+1, 0:ny+1)
		real :: s(8)
		integer :: k, l
		s = MHD_pri2con(cloud_inflow_pri())
		u_con(:,:,:,0,1:ny) = u_con(:,:,:,1,1:ny)
		do l = 1, py+1
			do k = 1, px+1
				u_con(:,k,l,nx+1,1:ny) = spread(s, 2, ny)
			end do
		end do
		u_con(:,:,:,1:nx,0) = u_con(:,:,:,1:nx,1)
		u_con(:,:,:,1:nx,ny+1) = u_con(:,:,:,1:nx,ny)
	end subroutine BC_cloud_shock

	pure function jet_nozzle_state() result(u_con_)
		use MHD_PDE_m, only: gamma, MHD_pri2con
		real, dimension(8) :: u_con_
		u_con_ = MHD_pri2con([gamma, 0.0, 800.0, 0.0, 0.0, sqrt(200.0), 0.0, 1.0])
	end function jet_nozzle_state

	! mirror symmetry at the left (normal momentum and field change sign),
	! element copy on the right and top, the nozzle inflow on bottom cells
	! 1..nx/10 and element copy on the rest of the bottom
	subroutine BC_jet1_cell_ext(u_con, xx, yy, nx, ny, px, py)
		integer, intent(in) :: nx, ny, px, py
		real, intent(in) :: xx(px+1, nx), yy(py+1, ny)
		real, intent(inout) :: u_con(8, px+1, py+1, 0:nx+1, 0:ny+1)
		real :: s(8), sgn(8)
		integer :: i, k, l
		s = jet_nozzle_state()
		sgn = [1.0, -1.0, 1.0, 1.0, -1.0, 1.0, 1.0, 1.0]
		do k = 1, px+1
			do l = 1, py+1
				do i = 1, ny
					u_con(:,k,l,0,i) = sgn * u_con(:,px+2-k,l,1,i)
				end do
			end do
		end do
		u_con(:,:,:,nx+1,1:ny) = u_con(:,:,:,nx,1:ny)
		do i = 1, nx
			if (i <= nx/10) then
				do l = 1, py+1
					do k = 1, px+1
						u_con(:,k,l,i,0) = s
					end do
				end do
			else
				u_con(:,:,:,i,0) = u_con(:,:,:,i,1)
			end if
		end do
		u_con(:,:,:,1:nx,ny+1) = u_con(:,:,:,1:nx,ny)
	end subroutine BC_jet1_cell_ext

	! --------------------------------------------------------- error functionals

	! L2 and max error of B_perp against the initial wave, serial order
	subroutine compute_error_smooth_alfven(u_con, hx, hy, xx, yy, nx, ny, px, py, err_l2, err_max)
		use gauss_m, only: gauss_lobatto
		integer, intent(in) :: nx, ny, px, py
		real, intent(in) :: hx, hy
		real, intent(in) :: xx(px+1, nx), yy(py+1, ny)
		real, dimension(8, px+1, py+1, 0:nx+1, 0:ny+1), intent(in) :: u_con
		real, intent(out) :: err_l2, err_max
		real :: qx(px+1), qwx(px+1), qy(py+1), qwy(py+1), d
		integer :: i, j, k, l
		call gauss_lobatto(px+1, qx, qwx)
		call gauss_lobatto(py+1, qy, qwy)
		err_l2 = 0.0
		err_max = 0.0
		do j = 1, ny
			do i = 1, nx
				do l = 1, py+1
					do k = 1, px+1
						d = (u_con(5,k,l,i,j)*(-sin(pi/6.0)) + u_con(6,k,l,i,j)*cos(pi/6.0)) &
							- 0.1*sin(2*pi*(xx(k,i)*cos(pi/6.0) + yy(l,j)*sin(pi/6.0)))
						err_l2 = err_l2 + (0.25*hx*hy*qwx(k)*qwy(l))*d**2
						err_max = max(err_max, abs(d))
					end do
				end do
			end do
		end do
		err_l2 = sqrt(err_l2)
	end subroutine compute_error_smooth_alfven

	! L2 and max error of the 8 conservative variables against the density
	! wave advected to time T, serial order
	subroutine compute_error_smooth1D(u_con, hx, hy, xx, yy, nx, ny, px, py, T, err_l2, err_max)
		use gauss_m, only: gauss_lobatto
		use MHD_PDE_m, only: MHD_pri2con
		integer, intent(in) :: nx, ny, px, py
		real, intent(in) :: hx, hy, T
		real, intent(in) :: xx(px+1, nx), yy(py+1, ny)
		real, dimension(8, px+1, py+1, 0:nx+1, 0:ny+1), intent(in) :: u_con
		real, intent(out) :: err_l2, err_max
		real :: qx(px+1), qwx(px+1), qy(py+1), qwy(py+1), d(8)
		integer :: i, j, k, l
		call gauss_lobatto(px+1, qx, qwx)
		call gauss_lobatto(py+1, qy, qwy)
		err_l2 = 0.0
		err_max = 0.0
		do j = 1, ny
			do i = 1, nx
				do l = 1, py+1
					do k = 1, px+1
						d = u_con(:,k,l,i,j) - MHD_pri2con([1.0 + 0.2*sin(xx(k,i) - T), 1.0, 0.0, 0.0, 0.5, 1.0, 1.5, 2.0])
						err_l2 = err_l2 + (0.25*hx*hy*qwx(k)*qwy(l))*sum(d**2)
						err_max = max(err_max, maxval(abs(d)))
					end do
				end do
			end do
		end do
		err_l2 = sqrt(err_l2)
	end subroutine compute_error_smooth1D
end module problems
