! adj_driver.f90 -- the device adjoint called from Fortran through the C-ABI
! of include/wpic_adj.h: the 2D Weibel beam problem from the seeded state,
! a Bz perturbation, a recorded run with checkpoints, its adjoint, and the
! gradients of the final magnetic energy with respect to drift, electron
! temperature and perturbation amplitude.
!
! Usage: adj_driver nx ny n_ppc nsteps every udrift uth_e amp
! Prints: W_B dW/dudrift dW/duth_e dW/damp
program adj_driver
  use iso_c_binding
  use wpic_capi
  implicit none
  type(wpic_params2d) :: p
  type(c_ptr) :: ctx
  integer :: nx, ny, n_ppc, nsteps, every
  real(c_double) :: udrift, uth_e, amp, wb, g(3), wpe, wpi, pi
  character(len=64) :: arg
  integer :: k

  call get_command_argument(1, arg); read(arg, *) nx
  call get_command_argument(2, arg); read(arg, *) ny
  call get_command_argument(3, arg); read(arg, *) n_ppc
  call get_command_argument(4, arg); read(arg, *) nsteps
  call get_command_argument(5, arg); read(arg, *) every
  call get_command_argument(6, arg); read(arg, *) udrift
  call get_command_argument(7, arg); read(arg, *) uth_e
  call get_command_argument(8, arg); read(arg, *) amp

  ! q, r of 2d/proj/weibel/app.f90:292-303 (sigma_e = 0, mass ratio 25,
  ! omega_pe = 0.1), ions first
  pi = 4d0*atan(1d0)
  wpe = 0.1d0
  wpi = wpe/sqrt(25d0)
  p%ndim = 6; p%nsp = 2
  p%nxgs = 2; p%nxge = 1 + nx; p%nygs = 2; p%nyge = 1 + ny
  p%delx = 1d0; p%delt = 0.5d0; p%c = 1d0; p%gfac = 0.501d0
  p%q = 0d0; p%r = 1d0
  p%r(1) = 25d0; p%r(2) = 1d0
  p%q(1) = +sqrt(p%r(1)/(4d0*pi*n_ppc))*wpi
  p%q(2) = -sqrt(p%r(2)/(4d0*pi*n_ppc))*wpe
  p%bc = WPIC_BC_PERIODIC; p%mode = WPIC_MODE_FAST
  p%capacity = 2_c_long_long*n_ppc*nx*ny

  call chk(wpic_create2d(ctx, p))
  call chk(wpic_set_cg_tol2d(ctx, 1d-13))
  call chk(wpic_ic_beams2d(ctx, 20261202_c_long_long, n_ppc, udrift, uth_e, 0.01d0, 0d0))
  call chk(wpic_perturb_bz2d(ctx, amp, 1))
  call chk(wpic_adj_forward2d(ctx, nsteps, every, wb))
  call chk(wpic_adj_backward2d(ctx, 1d0))
  call chk(wpic_adj_param_grads2d(ctx, udrift, uth_e, 1, g))
  write(*, '(4es26.17)') wb, (g(k), k = 1, 3)
  call chk(wpic_destroy(ctx))

contains

  subroutine chk(rc)
    integer(c_int), intent(in) :: rc
    if (rc /= 0) then
      write(6,*) wpic_error_text()
      stop 1
    end if
  end subroutine chk

end program adj_driver
