! Constants and run-information helpers used by the host modules
! (counterpart of utility.f90:1-19 of the reference).
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
end module utility
