! Constants and run-information helpers used by the host modules
! (counterpart of utility.f90:1-59 of the reference).
module utility
	use, intrinsic :: iso_fortran_env, only: stdin => input_unit, stdout => output_unit, stderr => error_unit
	implicit none
	real, parameter :: PI_math = 3.14159265358979323846264338327950288419716939937510582097494459230781640628620899862803482534211706798
contains
	subroutine compiler_para_output(file_unit)
		use, intrinsic :: iso_fortran_env, only: compiler_version, compiler_options
		integer, intent(in) :: file_unit
		write(file_unit, '(A)') 'compiler version: '//compiler_version()
		write(file_unit, '(A)') 'compiler options: '//compiler_options()
	end subroutine compiler_para_output

	! run-time parallel settings (utility.f90:21-59): the host side of this
	! build is serial Fortran over the GPU library, so it reports whether the
	! host was built with OpenMP and, if so, its thread settings
	subroutine omp_para_output(file_unit)
#ifdef _OPENMP
		use omp_lib
#endif
		integer, intent(in) :: file_unit
#ifdef _OPENMP
		write(file_unit, '(A, g0)') 'OMP_version     = ', _OPENMP
		write(file_unit, '(A, g0)') 'OMP_max_threads = ', omp_get_max_threads()
		write(file_unit, '(A, g0)') 'OMP_num_procs   = ', omp_get_num_procs()
		write(file_unit, '(A, g0)') 'OMP_proc_bind   = ', omp_get_proc_bind()
#else
		write(file_unit, '(A)') 'OpenMP is not enabled (host code is serial; the step runs on the GPU).'
#endif
	end subroutine omp_para_output
end module utility
