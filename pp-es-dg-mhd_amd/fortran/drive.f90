! Time integration driver on the MI355X (drop-in for drive_m, drive.f90:1-343).
!
! solve_execute keeps the reference's interface and plug-in pointers. The
! initial data is produced on the host by Init_handle and uploaded once; the
! whole SSP-RK loop (stage combines, BCs, post-processing, residuals, the dt
! rule, retries on an inadmissible average, the commit) then runs on the
! device, one host synchronisation per step, which returns (t, entropy) for
! the entropy log. A device boundary kind is chosen when BC_handle points at
! one of the six shipped fillers (BC_periodic, BC_YeeSjogreen, BC_rotor,
! BC_BrioWu, BC_cloud_shock, BC_jet1_cell_ext); any other BC_handle is called
! on the host between device calls (slower, same result).
module drive_m
	implicit none

	abstract interface
		subroutine Init_type(u_con, xx, yy, nx, ny, px, py)
			integer, intent(in) :: nx, ny, px, py
			real, intent(in) :: xx(px+1, nx), yy(py+1, ny)
			real, intent(inout) :: u_con(8, px+1, py+1, 0:nx+1, 0:ny+1)
		end subroutine Init_type

		subroutine BC_type(u_con, xx, yy, nx, ny, px, py)
			integer, intent(in) :: nx, ny, px, py
			real, intent(in) :: xx(px+1, nx), yy(py+1, ny)
			real, intent(inout) :: u_con(8, px+1, py+1, 0:nx+1, 0:ny+1)
		end subroutine BC_type
	end interface

	procedure(Init_type), pointer, public :: Init_handle => null()
	procedure(BC_type), pointer, public :: BC_handle => null()

	! run summary of the last solve_execute
	integer, public, save :: last_time_step_num = 0, last_retry_num = 0
	real, public, save :: last_t = 0.0, last_tot_ent = 0.0

	public :: Init_type, BC_type
	public :: solve_execute

contains

	subroutine solve_execute(final_T, CFL, use_oedg, use_LDF, ent_offset)
		use, intrinsic :: iso_c_binding
		use utility, only: stdout, stderr
		use ppesdg_capi
		use MHD_PDE_m, only: gamma
		use DG_solver_m
		use problems, only: BC_periodic, BC_YeeSjogreen, BC_rotor, rotor_ghost_state, BC_BrioWu, BC_cloud_shock, &
			BC_jet1_cell_ext
		use io, only: save_tecbin_entlog
		real, intent(in) :: final_T, CFL
		logical, intent(in) :: use_oedg, use_LDF, ent_offset
		integer(c_int) :: rc, kind, oedg_mode
		real(c_double) :: state(8)
		type(ppes_run_stats) :: st
		integer(c_long_long) :: prev_steps
		real, allocatable :: my_log(:,:)
		integer :: log_count
		real :: tic, toc
		character(len=128) :: filename, zonename

		if (.not. associated(Init_handle)) error stop "drive: solve_execute: procedure pointer 'Init_handle' not initialized!"
		if (.not. associated(BC_handle)) error stop "drive: solve_execute: procedure pointer 'BC_handle' not initialized!"
		if (rknstg /= 1 .and. rknstg /= 2 .and. rknstg /= 3 .and. rknstg /= 5) error stop 'unsupported yet'

		state = 0.0
		if (associated(BC_handle, BC_periodic)) then
			kind = PPES_BC_PERIODIC
		else if (associated(BC_handle, BC_YeeSjogreen)) then
			kind = PPES_BC_EXTRAP
		else if (associated(BC_handle, BC_rotor)) then
			kind = PPES_BC_FIXED
			state = rotor_ghost_state()
		else if (associated(BC_handle, BC_BrioWu)) then
			kind = PPES_BC_BRIOWU
		else if (associated(BC_handle, BC_cloud_shock)) then
			kind = PPES_BC_CLOUD
		else if (associated(BC_handle, BC_jet1_cell_ext)) then
			kind = PPES_BC_JET
		else
			call solve_execute_host_bc(final_T, CFL, use_oedg, use_LDF, ent_offset)
			return
		end if
		oedg_mode = merge(1, 0, use_oedg)

		call Init_handle(u_con, xx, yy, nx, ny, px, py)
		rc = ppes_set_gamma(ppes_handle, gamma)
		rc = ppes_set_bc(ppes_handle, kind, state)
		rc = ppes_upload_state(ppes_handle, 0_c_int, u_con)
		if (rc /= PPES_OK) error stop 'drive: solve_execute: upload failed'
		rc = ppes_set_ent_offset(ppes_handle, merge(1, 0, ent_offset))
		rc = ppes_solve_begin(ppes_handle, merge(1, 0, use_LDF))
		if (rc == PPES_NEG_AVG) error stop 'Error: Initial value has negative cell average!'
		if (rc /= PPES_OK) error stop 'drive: solve_execute: device start-up failed'
		! no steps: reads the start-up's entropy and dt_lim
		rc = ppes_advance(ppes_handle, -1.0_c_double, 0_c_long_long, real(CFL, c_double), oedg_mode, st)
		write(stdout, "(A, ES14.7, A, ES14.7, A, ES14.7, A, ES14.7)") 't=', 0.0, ', ent=', st%tot_ent, &
			', dt_lim=', st%dt_lim, ', time per it:', 0.0
		allocate(my_log(2, 64))
		log_count = 1
		my_log(:, 1) = [0.0, st%tot_ent]

		call cpu_time(tic)
		do
			prev_steps = st%steps
			rc = ppes_advance(ppes_handle, real(final_T, c_double), 1_c_long_long, real(CFL, c_double), oedg_mode, st)
			if (rc == PPES_DT_COLLAPSE) then
				write(stderr, *) "Computation failed: time step converges to 0! dt_lim=", st%dt_lim
				error stop 1
			else if (rc == PPES_PP_FAIL) then
				write(stderr, *) 're-enter limiting due to FP error in pressure'
				error stop 123456789
			else if (rc /= PPES_OK) then
				error stop 'drive: solve_execute: device failure'
			end if
			if (st%steps == prev_steps) exit
			log_count = log_count + 1
			if (log_count > size(my_log, 2)) call grow(my_log)
			my_log(:, log_count) = [st%t, st%tot_ent]
			if (mod(st%steps, 256_c_long_long) == 0) then
				call cpu_time(toc)
				write(stdout, "(A, ES14.7, A, ES14.7, A, ES14.7, A, ES14.7)") 't=', st%t, ', ent=', st%tot_ent, &
					', dt_lim=', st%dt_lim, ', time per it:', (toc - tic) / st%steps
			end if
		end do
		call cpu_time(toc)
		write(stdout, "(A, ES14.7, A, ES14.7, A, ES14.7, A, ES14.7)") 't=', st%t, ', ent=', st%tot_ent, &
			', dt_lim=', st%dt_lim, ', time per it:', (toc - tic) / max(st%steps, 1_c_long_long)

		rc = ppes_download_state(ppes_handle, 0_c_int, u_con)
		rc = ppes_download_avg(ppes_handle, u_avg)
		rc = ppes_download_residual(ppes_handle, 1_c_int, ut_con(:,:,:,:,:,1))
		last_time_step_num = int(st%steps)
		last_retry_num = int(st%retries)
		last_t = st%t
		last_tot_ent = st%tot_ent
		write(stdout, '(A)') "drive: solve_execute: Solving completed."
		write(stdout, '(A, g0)') "# time steps = ", st%steps
		write(stdout, '(A, g0)') "# re-trys = ", st%retries
		write(stdout, '(A, ES11.4, A)') "total CPU time = ", toc - tic, " s"
		write(filename, '(A)') 'entropy.plt'
		write(zonename, '(I0, 1X, I0, 1X, I0, 1X, I0)') nx, ny, px, py
		call save_tecbin_entlog(my_log, log_count, filename, zonename)
	end subroutine solve_execute

	subroutine grow(a)
		real, allocatable, intent(inout) :: a(:,:)
		real, allocatable :: t(:,:)
		allocate(t(size(a,1), 2*size(a,2)))
		t(:, 1:size(a,2)) = a
		call move_alloc(t, a)
	end subroutine grow

	! The loop with an arbitrary host BC_handle: every stage's ghost layer is
	! filled on the host (download, BC_handle, upload) around the device
	! post-process and residual (drive.f90:83-320).
	subroutine solve_execute_host_bc(final_T, CFL, use_oedg, use_LDF, ent_offset)
		use, intrinsic :: iso_c_binding
		use utility, only: stdout, stderr
		use ppesdg_capi
		use MHD_PDE_m, only: gamma
		use DG_solver_m
		real, intent(in) :: final_T, CFL
		logical, intent(in) :: use_oedg, use_LDF, ent_offset
		integer(c_int) :: rc, neg, nneg, s, slot_next
		real :: t, dt, dt_lim, tot_ent, t_RK(5), CFL_RK, beta
		logical :: flag, try_again, rk_final
		integer :: try_num, steps, retries
		real, allocatable :: w(:,:,:,:,:)

		CFL_RK = merge(1.0, 1.508, rknstg <= 3)
		allocate(w(8, px+1, py+1, 0:nx+1, 0:ny+1))
		rc = ppes_set_gamma(ppes_handle, gamma)
		rc = ppes_set_bc(ppes_handle, PPES_BC_HOST, [(0.0, s = 1, 8)])
		call Init_handle(u_con, xx, yy, nx, ny, px, py)
		t = 0.0
		call BC_handle(u_con, xx, yy, nx, ny, px, py)
		rc = ppes_upload_state(ppes_handle, 0_c_int, u_con)
		rc = ppes_post_process(ppes_handle, 0_c_int, 0_c_int, merge(1, 0, use_LDF), 0.0, 0_c_int, 1_c_int, neg, nneg, tot_ent)
		if (neg /= 0) error stop 'Error: Initial value has negative cell average!'
		rc = ppes_download_state(ppes_handle, 0_c_int, u_con)
		call BC_handle(u_con, xx, yy, nx, ny, px, py)
		rc = ppes_upload_state(ppes_handle, 0_c_int, u_con)
		rc = ppes_semidiscretize(ppes_handle, 1_c_int, 0_c_int, dt_lim)
		flag = .true.
		steps = 0
		retries = 0
		do while (flag)
			steps = steps + 1
			try_num = 0
			try_again = .true.
			do while (try_again)
				try_again = .false.
				try_num = try_num + 1
				if (try_num > 10 .or. dt_lim < 1e-14) then
					write(stderr, *) "Computation failed: time step converges to 0! dt_lim=", dt_lim
					error stop 1
				else if (try_num > 1) then
					retries = retries + 1
					rc = ppes_semidiscretize(ppes_handle, 1_c_int, 0_c_int, dt_lim)
				end if
				if (try_num <= 1) then
					dt = CFL_RK * CFL * dt_lim
				else
					dt = ((sqrt(1.25)-0.5)**(try_num-1)) * dt
				end if
				if (t + dt >= final_T) then
					dt = final_T - t
					flag = .false.
				else
					flag = .true.
				end if
				do s = 1, rknstg
					rc = ppes_rk_combine(ppes_handle, s, dt)
					call rk_time(rknstg, s, t, t_RK, dt, beta)
					rk_final = s == rknstg
					rc = ppes_download_state(ppes_handle, s, w)
					call BC_handle(w, xx, yy, nx, ny, px, py)
					rc = ppes_upload_state(ppes_handle, s, w)
					rc = ppes_post_process(ppes_handle, s, merge(1, 0, use_oedg .and. rk_final), merge(1, 0, use_LDF), &
						real(dt, c_double), merge(1, 0, ent_offset), merge(1, 0, rk_final), neg, nneg, tot_ent)
					if (rc == PPES_PP_FAIL) error stop 123456789
					if (neg /= 0) then
						write(stderr, '(A, g0, A, g0)') "RKstg ", s, ": Inadmissible average! Restart RK with reduced dt. cur_try_num=", try_num
						try_again = .true.
						exit
					end if
					rc = ppes_download_state(ppes_handle, s, w)
					call BC_handle(w, xx, yy, nx, ny, px, py)
					rc = ppes_upload_state(ppes_handle, s, w)
					slot_next = mod(s, rknstg) + 1
					rc = ppes_semidiscretize(ppes_handle, slot_next, s, dt_lim)
				end do
			end do
			! commit the interior only (drive.f90:304-311)
			rc = ppes_download_state(ppes_handle, rknstg, w)
			u_con(:,:,:,1:nx,1:ny) = w(:,:,:,1:nx,1:ny)
			rc = ppes_upload_state(ppes_handle, 0_c_int, u_con)
			t = t_RK(rknstg)
		end do
		rc = ppes_download_avg(ppes_handle, u_avg)
		last_time_step_num = steps
		last_retry_num = retries
		last_t = t
		last_tot_ent = tot_ent
		write(stdout, '(A)') "drive: solve_execute: Solving completed."
		write(stdout, '(A, g0)') "# time steps = ", steps
		write(stdout, '(A, g0)') "# re-trys = ", retries
	contains
		subroutine rk_time(S, st, t, tRK, dt, beta)
			integer, intent(in) :: S
			integer(c_int), intent(in) :: st
			real, intent(in) :: t, dt
			real, intent(inout) :: tRK(5)
			real, intent(out) :: beta
			select case (S)
			case (1)
				tRK(1) = t + dt*1.0; beta = 1.0
			case (2)
				if (st == 1) then
					tRK(1) = t + dt*1.0; beta = 1.0
				else
					tRK(2) = 0.5*t + 0.5*(tRK(1) + dt*1.0); beta = 0.5
				end if
			case (3)
				select case (st)
				case (1); tRK(1) = t + dt*1.0; beta = 1.0
				case (2); tRK(2) = 0.75*t + 0.25*(tRK(1) + dt*1.0); beta = 0.25
				case default; tRK(3) = (1.0/3.0)*t + (2.0/3.0)*(tRK(2) + dt*1.0); beta = 2.0/3.0
				end select
			case default
				select case (st)
				case (1); tRK(1) = t + (0.391752226571890*dt)*1.0; beta = 0.391752226571890
				case (2); tRK(2) = 0.444370493651235*t + 0.555629506348765*tRK(1) + (0.368410593050371*dt)*1.0
					beta = 0.368410593050371
				case (3); tRK(3) = 0.620101851488403*t + 0.379898148511597*tRK(2) + (0.251891774271694*dt)*1.0
					beta = 0.251891774271694
				case (4); tRK(4) = 0.178079954393132*t + 0.821920045606868*tRK(3) + (0.544974750228521*dt)*1.0
					beta = 0.544974750228521
				case default
					tRK(5) = 0.517231671970585*tRK(2) + 0.096059710526147*tRK(3) + (0.063692468666290*dt)*1.0 &
						+ 0.386708617503269*tRK(4) + (0.226007483236906*dt)*1.0
					beta = 0.063692468666290 + 0.226007483236906
				end select
			end select
		end subroutine rk_time
	end subroutine solve_execute_host_bc
end module drive_m
