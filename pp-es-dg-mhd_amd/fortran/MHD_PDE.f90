! Ratio of specific heats and the MHD primitive <-> conservative maps
! (drop-in for MHD_PDE.f90:1-10,44-70). gamma is global to the host program as
! in the reference; the device solvers receive it when a solve starts.
module MHD_PDE_m
	implicit none
	real, protected, save :: gamma = 5.0/3.0
contains
	subroutine set_gamma(gamma_)
		real, intent(in) :: gamma_
		if (gamma_ /= gamma) print '(A)', 'MHD_PDE: Warning: gamma is changed!'
		gamma = gamma_
	end subroutine set_gamma

	pure function MHD_con2pri(uc) result(up)
		real, dimension(8), intent(in) :: uc
		real, dimension(8) :: up
		real :: inv_rho
		up(1) = uc(1)
		up(2) = uc(2)/uc(1)
		up(3) = uc(3)/uc(1)
		up(4) = uc(4)/uc(1)
		up(5:7) = uc(5:7)
		up(8) = (gamma-1.0)*(uc(8) - 0.5*((uc(2)**2 + uc(3)**2 + uc(4)**2)/uc(1) + (uc(5)**2 + uc(6)**2 + uc(7)**2)))
	end function MHD_con2pri

	pure function MHD_pri2con(up) result(uc)
		real, dimension(8), intent(in) :: up
		real, dimension(8) :: uc
		uc(1) = up(1)
		uc(2:4) = up(1)*up(2:4)
		uc(5:7) = up(5:7)
		uc(8) = up(8)/(gamma-1.0) + 0.5*(up(1)*(up(2)**2 + up(3)**2 + up(4)**2) + (up(5)**2 + up(6)**2 + up(7)**2))
	end function MHD_pri2con
end module MHD_PDE_m
