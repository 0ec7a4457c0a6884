! Golden-vector harness for the reference solver (test infrastructure only).
!
! This program is builder-written: it `use`s the reference modules compiled
! from their own sources (DG_solver_m, drive_m, problems, MHD_PDE_m, LDF_m,
! gauss_m) and drives them through their public API only:
!   init_DG_solver / post_process / DG_semidiscretize / finalize_DG_solver
!   (DG_solver.f90:47,128,663,109), solve_execute (drive.f90:32), the plug-in
!   pointers Init_handle / BC_handle (drive.f90:24-25), init_by_fun_pri
!   / init_by_fun_con (problems.f90:25,53), the shipped initial data and
!   boundary fillers (problems.f90:74-654), the error functionals
!   (problems.f90:659-742), LDF_projector%init/apply (LDF.f90:114,219).
! Its step loop restates drive.f90:83-320 so that every RK stage can be
! dumped; mode=solve calls solve_execute itself for end-to-end checks.
!
! Output: one raw little-endian binary file per array in `out=` plus a
! manifest line "name dtype n1 n2 ..." (Fortran order) per file.
module harness_ics
	use utility, only: pi => PI_math
	use MHD_PDE_m, only: gamma
	implicit none
	real :: pert_amp = 0.0
contains
	! Orszag-Tang vortex on [0,2pi]^2: rho=gamma^2, p=gamma,
	! u=(-sin y, sin x, 0), B=(-sin y, sin 2x, 0)
	function ot2pi_pri(x, y) result(u_pri)
		real, intent(in) :: x, y
		real, dimension(8) :: u_pri
		u_pri(1) = gamma**2
		u_pri(2) = -sin(y)
		u_pri(3) = sin(x)
		u_pri(4) = 0.0
		u_pri(5) = -sin(y)
		u_pri(6) = sin(2.0*x)
		u_pri(7) = 0.0
		u_pri(8) = gamma
	end function ot2pi_pri

	! MHD blast wave on [-0.5,0.5]^2 (gamma = 1.4): rho=1, u=0,
	! p=1000 for r<0.1 else 0.1, B=(100/sqrt(4pi))(1,1,0)/sqrt(2)
	function blastc_pri(x, y) result(u_pri)
		real, intent(in) :: x, y
		real, dimension(8) :: u_pri
		real :: b0
		b0 = (100.0/sqrt(4.0*pi))/sqrt(2.0)
		u_pri(1) = 1.0
		u_pri(2) = 0.0
		u_pri(3) = 0.0
		u_pri(4) = 0.0
		u_pri(5) = b0
		u_pri(6) = b0
		u_pri(7) = 0.0
		if (x**2 + y**2 < 0.01) then
			u_pri(8) = 1000.0
		else
			u_pri(8) = 0.1
		end if
	end function blastc_pri
end module harness_ics

module harness_io
	implicit none
	character(len=512) :: outdir = '.'
	integer :: man_unit = -1
contains
	subroutine open_manifest()
		open(newunit=man_unit, file=trim(outdir)//'/manifest.txt', status='replace', action='write')
	end subroutine open_manifest

	subroutine dump_real(name, a, shp)
		character(len=*), intent(in) :: name
		real, intent(in) :: a(*)
		integer, intent(in) :: shp(:)
		integer :: u, n, i
		n = product(shp)
		open(newunit=u, file=trim(outdir)//'/'//trim(name)//'.bin', access='stream', form='unformatted', status='replace')
		write(u) a(1:n)
		close(u)
		write(man_unit, '(A,1X,A)', advance='no') trim(name), 'f8'
		do i = 1, size(shp)
			write(man_unit, '(1X,I0)', advance='no') shp(i)
		end do
		write(man_unit, *)
		flush(man_unit)
	end subroutine dump_real

	subroutine dump_scalar(name, v)
		character(len=*), intent(in) :: name
		real, intent(in) :: v
		real :: a(1)
		a(1) = v
		call dump_real(name, a, [1])
	end subroutine dump_scalar
end module harness_io

module harness_fileinit
	implicit none
	character(len=512) :: init_path = ''
contains
	! Init_type plug-in that reads the interior nodal conservative state
	! (8, px+1, py+1, nx, ny) from a raw binary file.
	subroutine init_from_file(u_con, xx, yy, nx, ny, px, py)
		integer, intent(in) :: nx, ny, px, py
		real, intent(in) :: xx(px+1, nx), yy(py+1, ny)
		real, intent(inout) :: u_con(8, px+1, py+1, 0:nx+1, 0:ny+1)
		real, allocatable :: buf(:,:,:,:,:)
		integer :: u
		allocate(buf(8, px+1, py+1, nx, ny))
		open(newunit=u, file=trim(init_path), access='stream', form='unformatted', status='old')
		read(u) buf
		close(u)
		u_con(:,:,:,1:nx,1:ny) = buf
	end subroutine init_from_file
end module harness_fileinit

program ref_harness
	use utility, only: pi => PI_math
	use gauss_m
	use LDF_m
	use MHD_PDE_m
	use DG_solver_m
	use drive_m
	use io, only: save_tecbin_finite_volume
	use problems
	use harness_ics
	use harness_io
	use harness_fileinit
	implicit none

	character(len=64) :: mode = 'stages', prob = 'ot2pi', bc = 'auto'
	integer :: nx_ = 8, ny_ = -1, p_ = 2, py_ = -1, nsteps = 1, oedg = 1, ldf = 1, dumpsteps = 1000000, rk_ = 3, plt = 0, entoff = 0
	real :: cfl = 0.2, tfinal = -1.0, oedg_dt_call = 0.01
	real :: left, right, bottom, top, gam
	real :: err2(2)
	logical :: have_init_file

	call parse_args()
	call execute_command_line('mkdir -p '//trim(outdir))
	call open_manifest()

	if (trim(mode) == 'tables') then
		call dump_tables()
		stop
	end if
	if (trim(mode) == 'fluxes') then
		call dump_fluxes()
		stop
	end if

	call setup_problem()
	if (rk_ /= 3 .and. trim(mode) /= 'solve') error stop 'harness: rk /= 3 needs mode=solve'
	call init_DG_solver(nx_, ny_, p_, py_, rk_, left, right, bottom, top)
	call dump_real('xx', xx, [p_+1, nx_])
	call dump_real('grid_x', grid_x, [nx_+1])
	call dump_real('grid_y', grid_y, [ny_+1])
	call dump_real('yy', yy, [py_+1, ny_])

	select case (trim(mode))
	case ('stages')
		call run_stages()
	case ('solve')
		if (tfinal <= 0.0) error stop 'harness: mode=solve needs tfinal'
		call solve_execute(tfinal, cfl, oedg >= 1, ldf == 1, entoff == 1)
		call dump_real('u_final', u_con, [8, p_+1, py_+1, nx_+2, ny_+2])
		call dump_real('avg_final', u_avg, [8, nx_+2, ny_+2])
		! the reference's Tecplot writer on the final averages (main.f90:67); its
		! solve_execute has written entropy.plt (drive.f90:338-341)
		if (plt == 1) call save_tecbin_finite_volume(u_avg, nx_, ny_, grid_x, grid_y, 0, tfinal)
		! the error functionals of the smooth problems (main.f90:78-80,91-93)
		if (trim(prob) == 'alfven') then
			call compute_error_smooth_alfven(u_con, hx, hy, xx, yy, nx, ny, px, py, err2(1), err2(2))
			call dump_real('err', err2, [2])
		else if (trim(prob) == 'wave1d') then
			call compute_error_smooth1D(u_con, hx, hy, xx, yy, nx, ny, px, py, tfinal, err2(1), err2(2))
			call dump_real('err', err2, [2])
		end if
	case ('calls')
		call run_calls()
	case default
		error stop 'harness: unknown mode'
	end select
	call finalize_DG_solver
	close(man_unit)

contains

	subroutine parse_args()
		integer :: i, n, eq
		character(len=600) :: arg, key, val
		n = command_argument_count()
		do i = 1, n
			call get_command_argument(i, arg)
			eq = index(arg, '=')
			if (eq <= 0) cycle
			key = arg(1:eq-1)
			val = arg(eq+1:)
			select case (trim(key))
			case ('mode'); mode = val
			case ('prob'); prob = val
			case ('bc'); bc = val
			case ('out'); outdir = val
			case ('init'); init_path = val
			case ('nx'); read(val, *) nx_
			case ('ny'); read(val, *) ny_
			case ('p'); read(val, *) p_
			case ('py'); read(val, *) py_
			case ('rk'); read(val, *) rk_
			case ('plt'); read(val, *) plt
			case ('entoff'); read(val, *) entoff
			case ('nsteps'); read(val, *) nsteps
			case ('dumpsteps'); read(val, *) dumpsteps
			case ('oedg'); read(val, *) oedg
			case ('ldf'); read(val, *) ldf
			case ('cfl'); read(val, *) cfl
			case ('tfinal'); read(val, *) tfinal
			case ('oedg_dt'); read(val, *) oedg_dt_call
			case ('gamma'); read(val, *) gam
			case ('left'); read(val, *) left
			case ('right'); read(val, *) right
			case ('bottom'); read(val, *) bottom
			case ('top'); read(val, *) top
			end select
		end do
		if (ny_ < 0) ny_ = nx_
		if (py_ < 0) py_ = p_
		have_init_file = len_trim(init_path) > 0
	end subroutine parse_args

	subroutine setup_problem()
		select case (trim(prob))
		case ('ot2pi')
			left = 0.0; right = 2*pi; bottom = 0.0; top = 2*pi
			call set_gamma(5.0/3.0)
			init_pri_fun => ot2pi_pri
			Init_handle => init_by_fun_pri
			if (trim(bc) == 'auto') bc = 'periodic'
		case ('otunit')
			left = 0.0; right = 1.0; bottom = 0.0; top = 1.0
			call set_gamma(5.0/3.0)
			init_pri_fun => OrszagTang_init_pri
			Init_handle => init_by_fun_pri
			if (trim(bc) == 'auto') bc = 'periodic'
		case ('rotor')
			left = 0.0; right = 1.0; bottom = 0.0; top = 1.0
			call set_gamma(1.4)
			init_pri_fun => Rotor_init_pri
			Init_handle => init_by_fun_pri
			if (trim(bc) == 'auto') bc = 'rotor'
		case ('blast')
			left = -0.5; right = 0.5; bottom = -0.5; top = 0.5
			call set_gamma(1.4)
			init_pri_fun => blastc_pri
			Init_handle => init_by_fun_pri
			if (trim(bc) == 'auto') bc = 'extrap'
		case ('blastref')
			left = 0.0; right = 1.0; bottom = 0.0; top = 1.0
			call set_gamma(1.4)
			init_pri_fun => Blast_init_pri
			Init_handle => init_by_fun_pri
			if (trim(bc) == 'auto') bc = 'periodic'
		! the rest of the reference's problem catalogue (main.f90:57-190)
		case ('alfven')
			left = 0.0; right = 1.0/cos(pi/6.0); bottom = 0.0; top = 1.0/sin(pi/6.0)
			call set_gamma(5.0/3.0)
			init_pri_fun => smooth_alfven_init_fun
			Init_handle => init_by_fun_pri
			if (trim(bc) == 'auto') bc = 'periodic'
		case ('khi')
			left = 0.0; right = 1.0; bottom = -1.0; top = 1.0
			call set_gamma(1.4)
			init_pri_fun => KHI_init_pri
			Init_handle => init_by_fun_pri
			if (trim(bc) == 'auto') bc = 'periodic'
		case ('briowu')
			left = -1.0; right = 1.0; bottom = -0.5; top = 0.5
			call set_gamma(5.0/3.0)
			init_pri_fun => Brio_Wu_init_fun
			Init_handle => init_by_fun_pri
			if (trim(bc) == 'auto') bc = 'briowu'
		case ('wave1d')
			left = 0.0; right = 2*pi; bottom = -0.5; top = 0.5
			call set_gamma(5.0/3.0)
			init_pri_fun => smooth_1Dwave_init_fun
			Init_handle => init_by_fun_pri
			if (trim(bc) == 'auto') bc = 'periodic'
		case ('cloud')
			left = 0.0; right = 1.0; bottom = 0.0; top = 1.0
			call set_gamma(5.0/3.0)
			init_pri_fun => cloud_shock_init_pri
			Init_handle => init_by_fun_pri
			if (trim(bc) == 'auto') bc = 'cloud'
		case ('jet')
			left = 0.0; right = 0.5; bottom = 0.0; top = 1.5
			call set_gamma(1.4)
			init_pri_fun => jet1_init_pri
			Init_handle => init_by_fun_pri
			if (trim(bc) == 'auto') bc = 'jet'
		case ('yee')
			left = -1.0; right = 1.0; bottom = -1.0; top = 1.0
			call set_gamma(5.0/3.0)
			init_con_fun => YeeSjogreenInitCon
			Init_handle => init_by_fun_con
			if (trim(bc) == 'auto') bc = 'extrap'
		case ('file')
			! domain and gamma come from the command line
			call set_gamma(gam)
			Init_handle => init_from_file
			if (trim(bc) == 'auto') bc = 'periodic'
		case default
			error stop 'harness: unknown prob'
		end select
		select case (trim(bc))
		case ('periodic'); BC_handle => BC_periodic
		case ('extrap'); BC_handle => BC_YeeSjogreen
		case ('rotor'); BC_handle => BC_rotor
		case ('briowu'); BC_handle => BC_BrioWu
		case ('cloud'); BC_handle => BC_cloud_shock
		case ('jet'); BC_handle => BC_jet1_cell_ext
		case default; error stop 'harness: unknown bc'
		end select
	end subroutine setup_problem

	! Restatement of the solve_execute loop (drive.f90:57-320) for SSP-RK3,
	! dumping every stage of the first `dumpsteps` steps. oedg: 0 = off,
	! 1 = final stage with oedg_dt = dt (drive.f90:280), 2 = every stage with
	! oedg_dt = beta_s*dt (the commented variant at drive.f90:283).
	subroutine run_stages()
		real :: t, dt, dt_lim, tot_ent, oedg_beta
		real :: t_RK(3)
		integer :: i, j, s, step, try_num, retry_num, num_negcell, total_negcell
		logical :: neg_avg, flag, try_again, rk_final, do_oedg
		character(len=64) :: tag
		integer, parameter :: max_try_num = 10
		integer :: shp_u(5), shp_ut(5), shp_avg(3)

		shp_u = [8, p_+1, py_+1, nx_+2, ny_+2]
		shp_ut = [8, p_+1, py_+1, nx_, ny_]
		shp_avg = [8, nx_+2, ny_+2]

		call Init_handle(u_con, xx, yy, nx, ny, px, py)
		t = 0.0
		call dump_real('u_init_raw', u_con(:,:,:,1:nx,1:ny), shp_ut)
		call BC_handle(u_con, xx, yy, nx, ny, px, py)
		call post_process(u_con, u_avg, neg_avg, num_negcell, apply_oedg=.false., apply_LDF=(ldf==1), oedg_dt=0.0, ent_offset=.false., comput_ent=.true., tot_ent=tot_ent)
		if (neg_avg) error stop 'harness: initial negative average'
		call dump_real('u_init_pp', u_con, shp_u)
		call dump_real('avg_init', u_avg, shp_avg)
		call dump_scalar('ent_init', tot_ent)
		call BC_handle(u_con, xx, yy, nx, ny, px, py)
		call DG_semidiscretize(ut_con(:,:,:,:,:,1), dt_lim, u_con)
		call dump_real('ut_init', ut_con(:,:,:,:,:,1), shp_ut)
		call dump_scalar('dtlim_init', dt_lim)

		flag = .true.
		step = 0
		retry_num = 0
		total_negcell = 0
		do while (flag .and. (tfinal > 0.0 .or. step < nsteps))
			step = step + 1
			try_num = 0
			try_again = .true.
			do while (try_again)
				try_again = .false.
				try_num = try_num + 1
				if (try_num > max_try_num .or. dt_lim < 1e-14) error stop 'harness: dt collapse'
				if (try_num > 1) then
					retry_num = retry_num + 1
					call DG_semidiscretize(ut_con(:,:,:,:,:,1), dt_lim, u_con)
				end if
				if (try_num <= 1) then
					dt = cfl * dt_lim
				else
					dt = ((sqrt(1.25)-0.5)**(try_num-1)) * dt
				end if
				if (tfinal > 0.0) then
					if (t + dt >= tfinal) then
						dt = tfinal - t
						flag = .false.
					else
						flag = .true.
					end if
				end if
				do s = 1, 3
					select case (s)
					case (1)
						do j = 1, ny
							do i = 1, nx
								uRK_con(:,:,:,i,j,1) = u_con(:,:,:,i,j) + dt*ut_con(:,:,:,i,j,1)
							end do
						end do
						t_RK(1) = t + dt*1.0
						oedg_beta = 1.0
						rk_final = .false.
					case (2)
						do j = 1, ny
							do i = 1, nx
								uRK_con(:,:,:,i,j,2) = 0.75*u_con(:,:,:,i,j) + 0.25*(uRK_con(:,:,:,i,j,1) + dt*ut_con(:,:,:,i,j,2))
							end do
						end do
						t_RK(2) = 0.75*t + 0.25*(t_RK(1) + dt*1.0)
						oedg_beta = 0.25
						rk_final = .false.
					case (3)
						do j = 1, ny
							do i = 1, nx
								uRK_con(:,:,:,i,j,3) = (1.0/3.0)*u_con(:,:,:,i,j) + (2.0/3.0)*(uRK_con(:,:,:,i,j,2) + dt*ut_con(:,:,:,i,j,3))
							end do
						end do
						t_RK(3) = (1.0/3.0)*t + (2.0/3.0)*(t_RK(2) + dt*1.0)
						oedg_beta = 2.0/3.0
						rk_final = .true.
					end select
					call BC_handle(uRK_con(:,:,:,:,:,s), xx, yy, nx, ny, px, py)
					if (oedg == 2) then
						do_oedg = .true.
						call post_process(uRK_con(:,:,:,:,:,s), u_avg, neg_avg, num_negcell, apply_oedg=.true., apply_LDF=(ldf==1), oedg_dt=(oedg_beta*dt), ent_offset=(entoff==1), comput_ent=rk_final, tot_ent=tot_ent)
					else
						do_oedg = (oedg == 1) .and. rk_final
						call post_process(uRK_con(:,:,:,:,:,s), u_avg, neg_avg, num_negcell, apply_oedg=do_oedg, apply_LDF=(ldf==1), oedg_dt=dt, ent_offset=(entoff==1), comput_ent=rk_final, tot_ent=tot_ent)
					end if
					if (step <= dumpsteps) then
						write(tag, '(A,I0,A,I0,A,I0)') 's', step, '_', s, '_t', try_num
						call dump_real(trim(tag)//'_u', uRK_con(:,:,:,:,:,s), shp_u)
						call dump_real(trim(tag)//'_avg', u_avg, shp_avg)
						call dump_real(trim(tag)//'_info', [dt, merge(1.0, 0.0, neg_avg), real(num_negcell), tot_ent], [4])
					end if
					if (neg_avg) then
						try_again = .true.
						exit
					end if
					total_negcell = total_negcell + num_negcell
					call BC_handle(uRK_con(:,:,:,:,:,s), xx, yy, nx, ny, px, py)
					call DG_semidiscretize(ut_con(:,:,:,:,:,mod(s,3)+1), dt_lim, uRK_con(:,:,:,:,:,s))
					if (step <= dumpsteps) then
						call dump_real(trim(tag)//'_ut', ut_con(:,:,:,:,:,mod(s,3)+1), shp_ut)
						call dump_scalar(trim(tag)//'_dtlim', dt_lim)
					end if
				end do
			end do
			do j = 1, ny
				do i = 1, nx
					u_con(:,:,:,i,j) = uRK_con(:,:,:,i,j,3)
				end do
			end do
			t = t_RK(3)
		end do
		call dump_real('u_final', u_con, shp_u)
		call dump_real('avg_final', u_avg, shp_avg)
		call dump_real('summary', [t, real(step), real(retry_num), real(total_negcell), tot_ent, dt_lim], [6])
	end subroutine run_stages

	! Single calls on a state read from file: one DG_semidiscretize, then one
	! post_process with the OEDG and LDF switches given on the command line.
	subroutine run_calls()
		real :: dt_lim, tot_ent
		integer :: num_negcell
		logical :: neg_avg
		call Init_handle(u_con, xx, yy, nx, ny, px, py)
		call BC_handle(u_con, xx, yy, nx, ny, px, py)
		call dump_real('c_u_in', u_con, [8, p_+1, py_+1, nx_+2, ny_+2])
		call DG_semidiscretize(ut_con(:,:,:,:,:,1), dt_lim, u_con)
		call dump_real('c_ut', ut_con(:,:,:,:,:,1), [8, p_+1, py_+1, nx_, ny_])
		call dump_scalar('c_dtlim', dt_lim)
		call post_process(u_con, u_avg, neg_avg, num_negcell, apply_oedg=(oedg>=1), apply_LDF=(ldf==1), oedg_dt=oedg_dt_call, ent_offset=.false., comput_ent=.true., tot_ent=tot_ent)
		call dump_real('c_u_pp', u_con, [8, p_+1, py_+1, nx_+2, ny_+2])
		call dump_real('c_avg', u_avg, [8, nx_+2, ny_+2])
		call dump_real('c_info', [merge(1.0, 0.0, neg_avg), real(num_negcell), tot_ent], [3])
	end subroutine run_calls

	! The four x-fluxes (ESPP_HLL_flux_x, PP_HLL_flux_x, ESPP_LF_flux_x,
	! EC_LF_flux_x, DG_solver.f90:894-1497) on the state pairs of file `init`
	! (n pairs of (w_l(8), w_r(8)), gamma from the command line): for each
	! pair and flux, nflux(8), nBn, sl, sr.
	subroutine dump_fluxes()
		real, allocatable :: w(:,:,:), res(:,:,:)
		integer :: u, n, i
		open(newunit=u, file=trim(init_path), access='stream', form='unformatted', status='old')
		inquire(unit=u, size=n)
		n = n / (8 * 16)
		allocate(w(8, 2, n), res(11, 4, n))
		read(u) w
		close(u)
		call set_gamma(gam)
		do i = 1, n
			call ESPP_HLL_flux_x(w(:,1,i), w(:,2,i), gamma, res(1:8,1,i), res(9,1,i), res(10,1,i), res(11,1,i))
			call PP_HLL_flux_x(w(:,1,i), w(:,2,i), gamma, res(1:8,2,i), res(9,2,i), res(10,2,i), res(11,2,i))
			call ESPP_LF_flux_x(w(:,1,i), w(:,2,i), gamma, res(1:8,3,i), res(9,3,i), res(10,3,i), res(11,3,i))
			call EC_LF_flux_x(w(:,1,i), w(:,2,i), gamma, res(1:8,4,i), res(9,4,i), res(10,4,i), res(11,4,i))
		end do
		call dump_real('fluxes', res, [11, 4, n])
	end subroutine dump_fluxes

	! 1D tables for p = 1..6 and the LDF modal projector for p = 1..6
	! (applied to every unit vector, giving the dense 2(p+1)^2 operator).
	subroutine dump_tables()
		integer :: p, n, j
		real, allocatable :: x(:), w(:), M(:,:), Minv(:,:), D(:,:), Pmn(:,:), Pnm(:,:), uu(:,:,:), op(:,:)
		type(LDF_projector), allocatable :: proj
		character(len=32) :: tag
		do p = 1, 6
			n = p + 1
			allocate(x(n), w(n), M(n,n), Minv(n,n), D(n,n), Pmn(n,n), Pnm(n,n), uu(n,n,2), op(2*n*n, 2*n*n))
			call gauss_lobatto(n, x, w)
			call gauss_lobatto_matrix(p, M, Minv, D, Pmn, Pnm)
			write(tag, '(A,I0)') 'p', p
			call dump_real(trim(tag)//'_x', x, [n])
			call dump_real(trim(tag)//'_w', w, [n])
			call dump_real(trim(tag)//'_M', M, [n, n])
			call dump_real(trim(tag)//'_Minv', Minv, [n, n])
			call dump_real(trim(tag)//'_D', D, [n, n])
			call dump_real(trim(tag)//'_Pmn', Pmn, [n, n])
			call dump_real(trim(tag)//'_Pnm', Pnm, [n, n])
			allocate(proj)
			call proj%init(p, p)
			do j = 1, 2*n*n
				uu = 0.0
				uu(mod(j-1, n)+1, mod((j-1)/n, n)+1, (j-1)/(n*n)+1) = 1.0
				call proj%apply(uu)
				op(:, j) = reshape(uu, [2*n*n])
			end do
			call dump_real(trim(tag)//'_ldf', op, [2*n*n, 2*n*n])
			deallocate(proj)
			deallocate(x, w, M, Minv, D, Pmn, Pnm, uu, op)
		end do
	end subroutine dump_tables
end program ref_harness
