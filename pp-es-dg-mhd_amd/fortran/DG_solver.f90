! Spatial discretisation and limiters on the MI355X (drop-in for
! DG_solver_m, DG_solver.f90:1-123 and the public routines post_process :128
! and DG_semidiscretize :663).
!
! The public arrays keep the reference's names, shapes and meaning and live
! on the host; the computation runs on the device through the C-ABI. A whole
! solve (drive_m's solve_execute) keeps its state on the device and only
! fills these arrays at the end; the two routines below are the per-call
! form, which copies their array arguments in and out.
module DG_solver_m
	use, intrinsic :: iso_c_binding, only: c_ptr, c_null_ptr, c_associated, c_int, c_double
	use ppesdg_capi
	use LDF_m
	implicit none

	real, dimension(:,:,:), allocatable, public :: u_avg
	real, dimension(:,:,:,:,:), allocatable, public :: u_con
	real, dimension(:,:,:,:,:,:), allocatable, public :: ut_con
	real, dimension(:,:,:,:,:,:), allocatable, public :: uRK_con

	real, parameter, public :: PP_eps = 1.0e-11

	real, protected :: domain_left, domain_right, domain_top, domain_bottom
	integer, protected :: nx, ny, px, py
	integer, protected :: rknstg
	real, protected :: hx, hy
	real, dimension(:), allocatable, protected :: grid_x, grid_y
	real, dimension(:,:), allocatable, protected :: xx, yy

	real, dimension(:), save, allocatable, protected :: qx_x, qx_y, qw_x, qw_y
	real, dimension(:,:), save, allocatable, protected :: diff_x, diff_y
	real, dimension(:,:), save, allocatable, protected :: M_x, Minv_x, M_y, Minv_y
	real, dimension(:,:), save, allocatable, protected :: Pmn_x, Pmn_y, Pnm_x, Pnm_y

	! the device solver behind this module
	type(c_ptr), public, save :: ppes_handle = c_null_ptr

contains

	subroutine init_DG_solver(nx_in, ny_in, px_in, py_in, rknstg_in, domain_left_in, domain_right_in, &
			domain_bottom_in, domain_top_in)
		use gauss_m
		use MHD_PDE_m, only: gamma
		use, intrinsic :: ieee_arithmetic
		integer, intent(in) :: nx_in, ny_in, px_in, py_in, rknstg_in
		real, intent(in) :: domain_left_in, domain_right_in, domain_bottom_in, domain_top_in
		type(ppes_config) :: cfg
		real :: nan_
		integer(c_int) :: rc
		nan_ = ieee_value(1.0, ieee_signaling_nan)
		if (px_in /= py_in .and. py_in /= 0) error stop 'init_DG_solver: the device takes py = px or py = 0'
		nx = nx_in; ny = ny_in; px = px_in; py = py_in; rknstg = rknstg_in
		domain_left = domain_left_in; domain_right = domain_right_in
		domain_bottom = domain_bottom_in; domain_top = domain_top_in

		allocate(u_avg(8, 0:nx+1, 0:ny+1), source = nan_)
		allocate(u_con(8, px+1, py+1, 0:nx+1, 0:ny+1), source = nan_)
		allocate(ut_con(8, px+1, py+1, nx, ny, rknstg), source = nan_)
		allocate(uRK_con(8, px+1, py+1, 0:nx+1, 0:ny+1, rknstg), source = nan_)
		allocate(grid_x(nx+1), grid_y(ny+1), xx(px+1, nx), yy(py+1, ny))
		allocate(qx_x(px+1), qw_x(px+1), qx_y(py+1), qw_y(py+1))
		allocate(diff_x(px+1,px+1), diff_y(py+1,py+1), M_x(px+1,px+1), Minv_x(px+1,px+1), M_y(py+1,py+1), Minv_y(py+1,py+1))
		allocate(Pmn_x(px+1,px+1), Pnm_x(px+1,px+1), Pmn_y(py+1,py+1), Pnm_y(py+1,py+1))
		call gauss_lobatto(px+1, qx_x, qw_x)
		call gauss_lobatto(py+1, qx_y, qw_y)
		call gauss_lobatto_matrix(px, M_x, Minv_x, diff_x, Pmn_x, Pnm_x)
		call gauss_lobatto_matrix(py, M_y, Minv_y, diff_y, Pmn_y, Pnm_y)

		cfg%nx = nx; cfg%ny = ny; cfg%px = px; cfg%py = py; cfg%rk_stages = rknstg
		cfg%left = domain_left; cfg%right = domain_right; cfg%bottom = domain_bottom; cfg%top = domain_top
		cfg%gamma = gamma
		cfg%bc_kind = PPES_BC_PERIODIC
		cfg%bc_state = 0.0
		rc = ppes_create(cfg, ppes_handle)
		if (rc /= PPES_OK) then
			write(*, '(A, I0)') 'init_DG_solver: ppes_create failed, status ', rc
			error stop 'init_DG_solver: device solver creation failed'
		end if
		rc = ppes_geometry(ppes_handle, hx, hy, grid_x, grid_y, xx, yy)
	end subroutine init_DG_solver

	subroutine finalize_DG_solver
		integer(c_int) :: rc
		deallocate(u_avg, u_con, ut_con, uRK_con)
		deallocate(grid_x, grid_y, xx, yy)
		deallocate(qx_x, qw_x, qx_y, qw_y, diff_x, diff_y, M_x, Minv_x, M_y, Minv_y, Pmn_x, Pnm_x, Pmn_y, Pnm_y)
		if (c_associated(ppes_handle)) rc = ppes_destroy(ppes_handle)
		ppes_handle = c_null_ptr
	end subroutine finalize_DG_solver

	! Per-call post_process (DG_solver.f90:128-661) on caller arrays.
	subroutine post_process(u_con, u_avg, neg_avg, num_negcell, apply_oedg, apply_LDF, oedg_dt, ent_offset, &
			comput_ent, tot_ent)
		use MHD_PDE_m, only: gamma
		real, dimension(8, px+1, py+1, 0:nx+1, 0:ny+1), intent(inout) :: u_con
		real, dimension(8, 0:nx+1, 0:ny+1), intent(inout) :: u_avg
		logical, intent(out) :: neg_avg
		integer, intent(out) :: num_negcell
		logical, intent(in) :: apply_oedg, apply_LDF, ent_offset, comput_ent
		real, intent(in) :: oedg_dt
		real, intent(out) :: tot_ent
		integer(c_int) :: rc, neg, nneg
		rc = ppes_set_gamma(ppes_handle, gamma)
		rc = ppes_upload_state(ppes_handle, 1_c_int, u_con)
		if (rc /= PPES_OK) error stop 'post_process: upload failed'
		rc = ppes_post_process(ppes_handle, 1_c_int, merge(1, 0, apply_oedg), merge(1, 0, apply_LDF), oedg_dt, &
			merge(1, 0, ent_offset), &
			merge(1, 0, comput_ent), neg, nneg, tot_ent)
		if (rc == PPES_PP_FAIL) then
			write(*, *) 're-enter limiting due to FP error in pressure'
			error stop 123456789
		end if
		if (rc /= PPES_OK .and. rc /= PPES_NEG_AVG) error stop 'post_process: device failure'
		neg_avg = neg /= 0
		num_negcell = nneg
		rc = ppes_download_state(ppes_handle, 1_c_int, u_con)
		rc = ppes_download_avg(ppes_handle, u_avg)
	end subroutine post_process

	! Per-call DG_semidiscretize (DG_solver.f90:663-892) on caller arrays.
	subroutine DG_semidiscretize(ut_con, dt_lim, u_con)
		use MHD_PDE_m, only: gamma
		real, dimension(8, px+1, py+1, 0:nx+1, 0:ny+1), intent(in) :: u_con
		real, dimension(8, px+1, py+1, nx, ny), intent(out) :: ut_con
		real, intent(out) :: dt_lim
		integer(c_int) :: rc
		rc = ppes_set_gamma(ppes_handle, gamma)
		rc = ppes_upload_state(ppes_handle, 1_c_int, u_con)
		if (rc /= PPES_OK) error stop 'DG_semidiscretize: upload failed'
		rc = ppes_semidiscretize(ppes_handle, 1_c_int, 1_c_int, dt_lim)
		if (rc /= PPES_OK) error stop 'DG_semidiscretize: device failure'
		rc = ppes_download_residual(ppes_handle, 1_c_int, ut_con)
	end subroutine DG_semidiscretize
end module DG_solver_m
