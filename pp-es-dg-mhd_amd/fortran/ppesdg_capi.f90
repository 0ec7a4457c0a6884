! ISO_C_BINDING interface to libppesdg.so (include/ppesdg.h).
!
! The modules DG_solver_m, drive_m, MHD_PDE_m, LDF_m, gauss_m, problems and
! io of this directory keep the reference's names and call through these
! bindings; this module holds only the C-ABI declarations.
module ppesdg_capi
	use, intrinsic :: iso_c_binding
	implicit none

	integer(c_int), parameter :: PPES_OK = 0, PPES_NEG_AVG = 1, PPES_DT_COLLAPSE = 2, PPES_PP_FAIL = 3, &
		PPES_BAD_ARG = 4, PPES_HIP_ERROR = 5, PPES_OOM = 6, PPES_UNSUPPORTED = 7
	integer(c_int), parameter :: PPES_BC_PERIODIC = 0, PPES_BC_EXTRAP = 1, PPES_BC_FIXED = 2, PPES_BC_HOST = 3
	integer(c_int), parameter :: PPES_BC_BRIOWU = 4, PPES_BC_CLOUD = 5, PPES_BC_JET = 6

	type, bind(C) :: ppes_config
		integer(c_int) :: nx, ny, px, py, rk_stages
		real(c_double) :: left, right, bottom, top, gamma
		integer(c_int) :: bc_kind
		real(c_double) :: bc_state(8)
	end type ppes_config

	type, bind(C) :: ppes_run_stats
		real(c_double) :: t, dt_last, dt_lim, tot_ent
		integer(c_long_long) :: steps, retries, negcells, stages
		integer(c_int) :: status, pad
	end type ppes_run_stats

	interface
		integer(c_int) function ppes_create(cfg, out) bind(C, name='ppes_create')
			import :: c_int, c_ptr, ppes_config
			type(ppes_config), intent(in) :: cfg
			type(c_ptr), intent(out) :: out
		end function ppes_create

		integer(c_int) function ppes_destroy(s) bind(C, name='ppes_destroy')
			import :: c_int, c_ptr
			type(c_ptr), value :: s
		end function ppes_destroy

		integer(c_int) function ppes_set_gamma(s, g) bind(C, name='ppes_set_gamma')
			import :: c_int, c_ptr, c_double
			type(c_ptr), value :: s
			real(c_double), value :: g
		end function ppes_set_gamma

		integer(c_int) function ppes_set_bc(s, kind, state) bind(C, name='ppes_set_bc')
			import :: c_int, c_ptr, c_double
			type(c_ptr), value :: s
			integer(c_int), value :: kind
			real(c_double), intent(in) :: state(8)
		end function ppes_set_bc

		integer(c_int) function ppes_geometry(s, hx, hy, grid_x, grid_y, xx, yy) bind(C, name='ppes_geometry')
			import :: c_int, c_ptr, c_double
			type(c_ptr), value :: s
			real(c_double), intent(out) :: hx, hy
			real(c_double), intent(out) :: grid_x(*), grid_y(*), xx(*), yy(*)
		end function ppes_geometry

		integer(c_int) function ppes_upload_state(s, slot, u) bind(C, name='ppes_upload_state')
			import :: c_int, c_ptr, c_double
			type(c_ptr), value :: s
			integer(c_int), value :: slot
			real(c_double), intent(in) :: u(*)
		end function ppes_upload_state

		integer(c_int) function ppes_download_state(s, slot, u) bind(C, name='ppes_download_state')
			import :: c_int, c_ptr, c_double
			type(c_ptr), value :: s
			integer(c_int), value :: slot
			real(c_double), intent(out) :: u(*)
		end function ppes_download_state

		integer(c_int) function ppes_upload_residual(s, slot, ut) bind(C, name='ppes_upload_residual')
			import :: c_int, c_ptr, c_double
			type(c_ptr), value :: s
			integer(c_int), value :: slot
			real(c_double), intent(in) :: ut(*)
		end function ppes_upload_residual

		integer(c_int) function ppes_download_residual(s, slot, ut) bind(C, name='ppes_download_residual')
			import :: c_int, c_ptr, c_double
			type(c_ptr), value :: s
			integer(c_int), value :: slot
			real(c_double), intent(out) :: ut(*)
		end function ppes_download_residual

		integer(c_int) function ppes_download_avg(s, u_avg) bind(C, name='ppes_download_avg')
			import :: c_int, c_ptr, c_double
			type(c_ptr), value :: s
			real(c_double), intent(out) :: u_avg(*)
		end function ppes_download_avg

		integer(c_int) function ppes_apply_bc(s, slot) bind(C, name='ppes_apply_bc')
			import :: c_int, c_ptr
			type(c_ptr), value :: s
			integer(c_int), value :: slot
		end function ppes_apply_bc

		integer(c_int) function ppes_post_process(s, slot, apply_oedg, apply_ldf, oedg_dt, ent_offset, comput_ent, &
				neg_avg, num_negcell, tot_ent) bind(C, name='ppes_post_process')
			import :: c_int, c_ptr, c_double
			type(c_ptr), value :: s
			integer(c_int), value :: slot, apply_oedg, apply_ldf, ent_offset, comput_ent
			real(c_double), value :: oedg_dt
			integer(c_int), intent(out) :: neg_avg, num_negcell
			real(c_double), intent(out) :: tot_ent
		end function ppes_post_process

		integer(c_int) function ppes_semidiscretize(s, ut_slot, u_slot, dt_lim) bind(C, name='ppes_semidiscretize')
			import :: c_int, c_ptr, c_double
			type(c_ptr), value :: s
			integer(c_int), value :: ut_slot, u_slot
			real(c_double), intent(out) :: dt_lim
		end function ppes_semidiscretize

		integer(c_int) function ppes_rk_combine(s, stage, dt) bind(C, name='ppes_rk_combine')
			import :: c_int, c_ptr, c_double
			type(c_ptr), value :: s
			integer(c_int), value :: stage
			real(c_double), value :: dt
		end function ppes_rk_combine

		integer(c_int) function ppes_set_ent_offset(s, on) bind(C, name='ppes_set_ent_offset')
			import :: c_int, c_ptr
			type(c_ptr), value :: s
			integer(c_int), value :: on
		end function ppes_set_ent_offset

		integer(c_int) function ppes_solve_begin(s, use_ldf) bind(C, name='ppes_solve_begin')
			import :: c_int, c_ptr
			type(c_ptr), value :: s
			integer(c_int), value :: use_ldf
		end function ppes_solve_begin

		integer(c_int) function ppes_advance(s, t_final, nsteps, cfl, oedg_mode, stats) bind(C, name='ppes_advance')
			import :: c_int, c_ptr, c_double, c_long_long, ppes_run_stats
			type(c_ptr), value :: s
			real(c_double), value :: t_final, cfl
			integer(c_long_long), value :: nsteps
			integer(c_int), value :: oedg_mode
			type(ppes_run_stats), intent(out) :: stats
		end function ppes_advance

		! host-only table builders without side effects: pure, so that the
		! reference's pure routines (gauss_lobatto, LDF_projector%apply) stay pure
		pure integer(c_int) function ppes_tables(p, x, w, D, Pmn, Pnm) bind(C, name='ppes_tables')
			import :: c_int, c_double
			integer(c_int), value :: p
			real(c_double), intent(out) :: x(*), w(*), D(*), Pmn(*), Pnm(*)
		end function ppes_tables

		pure integer(c_int) function ppes_ldf_modal_operator(p, op) bind(C, name='ppes_ldf_modal_operator')
			import :: c_int, c_double
			integer(c_int), value :: p
			real(c_double), intent(out) :: op(*)
		end function ppes_ldf_modal_operator
	end interface
end module ppesdg_capi
