! Tecplot binary (.plt, "#!TDV191") output of cell-average primitives and of
! the (t, total entropy) log (drop-in for save_tecbin_finite_volume and
! save_tecbin_entlog, io.f90:4-204). The byte layout follows the published
! Tecplot binary format for one ordered zone, as the reference writes it:
! header (magic, byte order, file type, title, variable names), one zone
! header, the end-of-header marker 357.0, then the data section with FP32
! values and FP64 per-variable min/max.
module io
	use, intrinsic :: iso_fortran_env, only: int32, real32, real64
	implicit none
	private
	public :: save_tecbin_finite_volume, save_tecbin_entlog
contains

	subroutine put_string(u, s)
		integer, intent(in) :: u
		character(len=*), intent(in) :: s
		integer :: i
		do i = 1, len_trim(s)
			write(u) int(iachar(s(i:i)), int32)
		end do
		write(u) 0_int32
	end subroutine put_string

	! header of a one-zone ordered file; loc(k) = 0 (node) or 1 (cell centre)
	subroutine put_header(u, title, names, zone, time, loc, ni, nj)
		integer, intent(in) :: u, ni, nj
		character(len=*), intent(in) :: title, zone
		character(len=*), intent(in) :: names(:)
		real(real64), intent(in) :: time
		integer, intent(in) :: loc(:)
		integer :: k
		write(u) "#!TDV191"
		write(u) 1_int32, 0_int32
		call put_string(u, title)
		write(u) int(size(names), int32)
		do k = 1, size(names)
			call put_string(u, names(k))
		end do
		write(u) 299.0_real32
		call put_string(u, zone)
		write(u) -1_int32, -2_int32
		write(u) time
		write(u) -1_int32, 0_int32, 1_int32
		do k = 1, size(loc)
			write(u) int(loc(k), int32)
		end do
		write(u) 0_int32, 0_int32
		write(u) int(ni, int32), int(nj, int32), 1_int32
		write(u) 0_int32
		write(u) 357.0_real32
		write(u) 299.0_real32
		do k = 1, size(names)
			write(u) 1_int32
		end do
		write(u) 0_int32, 0_int32, -1_int32
	end subroutine put_header

	subroutine save_tecbin_finite_volume(u_avg, nx, ny, grid_x, grid_y, id, time)
		use MHD_PDE_m, only: MHD_con2pri
		integer, intent(in) :: nx, ny, id
		real, intent(in) :: grid_x(nx+1), grid_y(ny+1), time
		real, intent(in) :: u_avg(8, 0:nx+1, 0:ny+1)
		character(len=100) :: filename, zonename
		real(real32), allocatable :: xy(:,:,:), var(:,:,:)
		real(real64) :: vmin(8), vmax(8)
		integer :: u, ios, i, j, l
		write(filename, '(A, "_", g0, ".plt")') 'solution', id
		write(zonename, '(A, "_", g0)') 'zone', id
		open(newunit=u, file=trim(filename), form='unformatted', status='replace', access='stream', iostat=ios)
		if (ios /= 0) error stop 'save_result_tecbin: failed to open file!'
		call put_header(u, "MHD result", [character(len=8) :: "x", "y", "Density", "Vx", "Vy", "Vz", "Bx", "By", &
			"Bz", "Pressure"], trim(zonename), real(time, real64), [0, 0, 1, 1, 1, 1, 1, 1, 1, 1], nx+1, ny+1)
		! cell-centred arrays of an ordered zone are dimensioned (maxI, maxJ-1):
		! the last I column is padding and is written as zero
		allocate(xy(nx+1, ny+1, 2), var(nx+1, ny, 8))
		var = 0.0
		vmin = huge(1.0_real64)
		vmax = -huge(1.0_real64)
		do j = 1, ny+1
			do i = 1, nx+1
				xy(i,j,1) = real(grid_x(i), real32)
				xy(i,j,2) = real(grid_y(j), real32)
			end do
		end do
		do j = 1, ny
			do i = 1, nx
				var(i,j,:) = real(MHD_con2pri(u_avg(:,i,j)), real32)
				do l = 1, 8
					vmin(l) = min(vmin(l), real(var(i,j,l), real64))
					vmax(l) = max(vmax(l), real(var(i,j,l), real64))
				end do
			end do
		end do
		write(u) real(minval(grid_x), real64), real(maxval(grid_x), real64)
		write(u) real(minval(grid_y), real64), real(maxval(grid_y), real64)
		do l = 1, 8
			write(u) vmin(l), vmax(l)
		end do
		write(u) xy
		write(u) var
		close(u)
	end subroutine save_tecbin_finite_volume

	subroutine save_tecbin_entlog(my_log, n, filename, zonename)
		integer, intent(in) :: n
		real, intent(in) :: my_log(:,:)
		character(len=*), intent(in) :: filename, zonename
		integer :: u, ios, i
		open(newunit=u, file=trim(filename), form='unformatted', status='replace', access='stream', iostat=ios)
		if (ios /= 0) error stop 'save_result_tecbin: failed to open file!'
		call put_header(u, "MHD entropy log", [character(len=13) :: "t", "total entropy"], zonename, 0.0_real64, &
			[0, 0], n, 1)
		do i = 1, 2
			write(u) real(minval(my_log(i, 1:n)), real64), real(maxval(my_log(i, 1:n)), real64)
		end do
		do i = 1, 2
			write(u) real(my_log(i, 1:n), real32)
		end do
		close(u)
	end subroutine save_tecbin_entlog
end module io
