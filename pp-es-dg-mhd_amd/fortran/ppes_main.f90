! Example host program in the reference's style (cf. main.f90:53-184): it
! binds the plug-ins, calls init_DG_solver / solve_execute /
! save_tecbin_finite_volume / finalize_DG_solver of the drop-in modules, and
! writes the final state for checking. Every problem of main.f90 is here,
! with its domain, gamma and filler; sizes, degrees, RK stages and t_final
! come from the command line (the 1D problems take py = 0 by default, as
! main.f90 cases 0 and 1 do).
!
!   ppes_main prob=ot2pi nx=128 ny=128 p=2 [py=2] [rk=3] tfinal=0.5 cfl=0.2 [oedg=1] [ldf=1] [out=final.bin]
program ppes_main
	use utility, only: pi => PI_math
	use MHD_PDE_m
	use DG_solver_m
	use drive_m
	use problems
	use io
	implicit none
	character(len=64) :: prob = 'ot2pi'
	character(len=256) :: out = ''
	integer :: nx_ = 32, ny_ = -1, p_ = 2, py_ = -1, oedg = 1, ldf = 1, rk = 3
	real :: tfinal = 0.1, cfl = 0.2, left, right, bottom, top, err_l2, err_max
	integer :: u
	logical :: one_d = .false.
	call parse_args()
	select case (trim(prob))
	case ('briowu')
		left = -1.0; right = 1.0; bottom = -0.5; top = 0.5
		Init_handle => init_by_fun_pri
		init_pri_fun => Brio_Wu_init_fun
		BC_handle => BC_BrioWu
		one_d = .true.
	case ('wave1d')
		left = 0.0; right = 2*pi; bottom = -0.5; top = 0.5
		Init_handle => init_by_fun_pri
		init_pri_fun => smooth_1Dwave_init_fun
		BC_handle => BC_periodic
		one_d = .true.
	case ('alfven')
		left = 0.0; right = 1.0/cos(pi/6.0); bottom = 0.0; top = 1.0/sin(pi/6.0)
		Init_handle => init_by_fun_pri
		init_pri_fun => smooth_alfven_init_fun
		BC_handle => BC_periodic
	case ('yee')
		left = -1.0; right = 1.0; bottom = -1.0; top = 1.0
		Init_handle => init_by_fun_con
		init_con_fun => YeeSjogreenInitCon
		BC_handle => BC_YeeSjogreen
	case ('blast_periodic')
		left = 0.0; right = 1.0; bottom = 0.0; top = 1.0
		call set_gamma(1.4)
		Init_handle => init_by_fun_pri
		init_pri_fun => Blast_init_pri
		BC_handle => BC_periodic
	case ('cloud')
		left = 0.0; right = 1.0; bottom = 0.0; top = 1.0
		Init_handle => init_by_fun_pri
		init_pri_fun => cloud_shock_init_pri
		BC_handle => BC_cloud_shock
	case ('jet')
		left = 0.0; right = 0.5; bottom = 0.0; top = 1.5
		call set_gamma(1.4)
		Init_handle => init_by_fun_pri
		init_pri_fun => jet1_init_pri
		BC_handle => BC_jet1_cell_ext
	case ('khi')
		left = 0.0; right = 1.0; bottom = -1.0; top = 1.0
		call set_gamma(1.4)
		Init_handle => init_by_fun_pri
		init_pri_fun => KHI_init_pri
		BC_handle => BC_periodic
	case ('ot2pi')
		left = 0.0; right = 2*pi; bottom = 0.0; top = 2*pi
		call set_gamma(5.0/3.0)
		Init_handle => init_by_fun_pri
		init_pri_fun => OrszagTang2pi_init_pri
		BC_handle => BC_periodic
	case ('otunit')
		left = 0.0; right = 1.0; bottom = 0.0; top = 1.0
		Init_handle => init_by_fun_pri
		init_pri_fun => OrszagTang_init_pri
		BC_handle => BC_periodic
	case ('rotor')
		left = 0.0; right = 1.0; bottom = 0.0; top = 1.0
		call set_gamma(1.4)
		Init_handle => init_by_fun_pri
		init_pri_fun => Rotor_init_pri
		BC_handle => BC_rotor
	case ('blast')
		left = -0.5; right = 0.5; bottom = -0.5; top = 0.5
		call set_gamma(1.4)
		Init_handle => init_by_fun_pri
		init_pri_fun => BlastDiag_init_pri
		BC_handle => BC_YeeSjogreen
	case ('blast_hostbc')
		! the same blast, with the boundary filler called on the host
		left = -0.5; right = 0.5; bottom = -0.5; top = 0.5
		call set_gamma(1.4)
		Init_handle => init_by_fun_pri
		init_pri_fun => BlastDiag_init_pri
		BC_handle => extrap_host
	case default
		error stop 'ppes_main: unknown prob'
	end select
	if (one_d .and. ny_ < 0) ny_ = 1
	if (ny_ < 0) ny_ = nx_
	if (py_ < 0) py_ = merge(0, p_, one_d)
	call init_DG_solver(nx_, ny_, p_, py_, rk, left, right, bottom, top)
	print *, 'Problem ', trim(prob), ', nx=', nx, ', ny=', ny, ', px=', px, ', py=', py
	call solve_execute(tfinal, cfl, oedg == 1, ldf == 1, .false.)
	select case (trim(prob))
	case ('wave1d')
		call compute_error_smooth1D(u_con, hx, hy, xx, yy, nx, ny, px, py, tfinal, err_l2, err_max)
		print '(A, ES24.16)', ' l2_err = ', err_l2
		print '(A, ES24.16)', ' max_err = ', err_max
	case ('alfven')
		call compute_error_smooth_alfven(u_con, hx, hy, xx, yy, nx, ny, px, py, err_l2, err_max)
		print '(A, ES24.16)', ' l2_err = ', err_l2
		print '(A, ES24.16)', ' max_err = ', err_max
	case default
		call save_tecbin_finite_volume(u_avg, nx, ny, grid_x, grid_y, 0, tfinal)
	end select
	if (len_trim(out) > 0) then
		open(newunit=u, file=trim(out), access='stream', form='unformatted', status='replace')
		write(u) real(last_time_step_num, 8), real(last_retry_num, 8), last_t, last_tot_ent
		write(u) u_con(:,:,:,1:nx,1:ny)
		close(u)
	end if
	call finalize_DG_solver
	print '(A)', 'Program MAIN terminated successfully.'
contains
	subroutine parse_args()
		integer :: i, eq
		character(len=300) :: arg, key, val
		do i = 1, command_argument_count()
			call get_command_argument(i, arg)
			eq = index(arg, '=')
			if (eq <= 0) cycle
			key = arg(1:eq-1)
			val = arg(eq+1:)
			select case (trim(key))
			case ('prob'); prob = val
			case ('out'); out = val
			case ('nx'); read(val, *) nx_
			case ('ny'); read(val, *) ny_
			case ('p'); read(val, *) p_
			case ('py'); read(val, *) py_
			case ('rk'); read(val, *) rk
			case ('oedg'); read(val, *) oedg
			case ('ldf'); read(val, *) ldf
			case ('tfinal'); read(val, *) tfinal
			case ('cfl'); read(val, *) cfl
			end select
		end do
	end subroutine parse_args

	! a user-supplied BC (not one of the device kinds): trace extrapolation
	subroutine extrap_host(u_con, xx, yy, nx, ny, px, py)
		integer, intent(in) :: nx, ny, px, py
		real, intent(in) :: xx(px+1, nx), yy(py+1, ny)
		real, intent(inout) :: u_con(8, px+1, py+1, 0:nx+1, 0:ny+1)
		call BC_YeeSjogreen(u_con, xx, yy, nx, ny, px, py)
	end subroutine extrap_host
end program ppes_main
