! wpic_capi.f90 -- ISO_C_BINDING interfaces to the C-ABI of include/wpic.h,
! include/wpic_adj.h and include/wpic_io.h.
module wpic_capi
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: WPIC_BC_PERIODIC = 0, WPIC_BC_WALLX = 1, WPIC_BC_SHOCK = 2
  integer(c_int), parameter :: WPIC_MODE_FAST = 0, WPIC_MODE_PARITY = 1, WPIC_MODE_MIXED = 2

  type, bind(C) :: wpic_params2d
     integer(c_int) :: ndim, nsp
     integer(c_int) :: nxgs, nxge, nygs, nyge
     real(c_double) :: delx, delt, c, gfac
     real(c_double) :: q(4), r(4)
     integer(c_int) :: bc, mode
     integer(c_long_long) :: capacity
  end type wpic_params2d

  interface
     integer(c_int) function wpic_create2d(ctx, p) bind(C, name='wpic_create2d')
       import :: c_int, c_ptr, wpic_params2d
       type(c_ptr), intent(out) :: ctx
       type(wpic_params2d), intent(in) :: p
     end function wpic_create2d

     integer(c_int) function wpic_destroy(ctx) bind(C, name='wpic_destroy')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function wpic_destroy

     type(c_ptr) function wpic_last_error() bind(C, name='wpic_last_error')
       import :: c_ptr
     end function wpic_last_error

     integer(c_int) function wpic_upload_up2d(ctx, up, np2, cumcnt, np) bind(C, name='wpic_upload_up2d')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: up(*)
       integer(c_int), intent(in) :: np2(*), cumcnt(*)
       integer(c_int), value :: np
     end function wpic_upload_up2d

     integer(c_int) function wpic_download_up2d(ctx, up, np2, cumcnt, np) bind(C, name='wpic_download_up2d')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(inout) :: up(*)
       integer(c_int), intent(out) :: np2(*), cumcnt(*)
       integer(c_int), value :: np
     end function wpic_download_up2d

     integer(c_int) function wpic_upload_gp2d(ctx, gp, np2, np) bind(C, name='wpic_upload_gp2d')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: gp(*)
       integer(c_int), intent(in) :: np2(*)
       integer(c_int), value :: np
     end function wpic_upload_gp2d

     integer(c_int) function wpic_download_gp2d(ctx, gp, np2, np) bind(C, name='wpic_download_gp2d')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(inout) :: gp(*)
       integer(c_int), intent(out) :: np2(*)
       integer(c_int), value :: np
     end function wpic_download_gp2d

     integer(c_int) function wpic_upload_fields2d(ctx, uf) bind(C, name='wpic_upload_fields2d')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: uf(*)
     end function wpic_upload_fields2d

     integer(c_int) function wpic_download_fields2d(ctx, uf) bind(C, name='wpic_download_fields2d')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: uf(*)
     end function wpic_download_fields2d

     integer(c_int) function wpic_mom_accl2d(ctx) bind(C, name='wpic_mom_accl2d')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function wpic_mom_accl2d

     integer(c_int) function wpic_mom_nvt2d(ctx) bind(C, name='wpic_mom_nvt2d')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function wpic_mom_nvt2d

     integer(c_int) function wpic_bc_mom2d(ctx) bind(C, name='wpic_bc_mom2d')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function wpic_bc_mom2d

     integer(c_int) function wpic_moments2d(ctx) bind(C, name='wpic_moments2d')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function wpic_moments2d

     integer(c_int) function wpic_download_mom2d(ctx, mom) bind(C, name='wpic_download_mom2d')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: mom(*)
     end function wpic_download_mom2d

     integer(c_int) function wpic_upload_mom2d(ctx, mom) bind(C, name='wpic_upload_mom2d')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: mom(*)
     end function wpic_upload_mom2d

     integer(c_int) function wpic_paraio_init2d(io, ndim, np, nsp, nxgs, nxge, nygs, nyge, nys, nye, nproc, &
          nrank, delx, delt, c, q, r, dir) bind(C, name='wpic_paraio_init2d')
       import :: c_int, c_ptr, c_double, c_char
       type(c_ptr), intent(out) :: io
       integer(c_int), value :: ndim, np, nsp, nxgs, nxge, nygs, nyge, nys, nye, nproc, nrank
       real(c_double), value :: delx, delt, c
       real(c_double), intent(in) :: q(*), r(*)
       character(kind=c_char), intent(in) :: dir(*)
     end function wpic_paraio_init2d

     integer(c_int) function wpic_paraio_finalize(io) bind(C, name='wpic_paraio_finalize')
       import :: c_int, c_ptr
       type(c_ptr), value :: io
     end function wpic_paraio_finalize

     integer(c_int) function wpic_paraio_output(io, up, uf, np2, nxs, nxe, it, filename) &
          bind(C, name='wpic_paraio_output')
       import :: c_int, c_ptr, c_double, c_char
       type(c_ptr), value :: io
       real(c_double), intent(in) :: up(*), uf(*)
       integer(c_int), intent(in) :: np2(*)
       integer(c_int), value :: nxs, nxe, it
       character(kind=c_char), intent(in) :: filename(*)
     end function wpic_paraio_output

     integer(c_int) function wpic_paraio_input(io, up, uf, np2, nxs, nxe, it, filename) &
          bind(C, name='wpic_paraio_input')
       import :: c_int, c_ptr, c_double, c_char
       type(c_ptr), value :: io
       real(c_double), intent(out) :: up(*), uf(*)
       integer(c_int), intent(out) :: np2(*)
       integer(c_int), intent(out) :: nxs, nxe, it
       character(kind=c_char), intent(in) :: filename(*)
     end function wpic_paraio_input

     integer(c_int) function wpic_paraio_param(io, n0, wpe, wpi, wge, wgi, vti, vte, filename) &
          bind(C, name='wpic_paraio_param')
       import :: c_int, c_ptr, c_double, c_char
       type(c_ptr), value :: io
       integer(c_int), value :: n0
       real(c_double), value :: wpe, wpi, wge, wgi, vti, vte
       character(kind=c_char), intent(in) :: filename(*)
     end function wpic_paraio_param

     integer(c_int) function wpic_paraio_mom(io, mom, uf, it) bind(C, name='wpic_paraio_mom')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: io
       real(c_double), intent(inout) :: mom(*)
       real(c_double), intent(in) :: uf(*)
       integer(c_int), value :: it
     end function wpic_paraio_mom

     integer(c_int) function wpic_paraio_ptcl(io, up, uf, np2, it) bind(C, name='wpic_paraio_ptcl')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: io
       real(c_double), intent(in) :: up(*), uf(*)
       integer(c_int), intent(in) :: np2(*)
       integer(c_int), value :: it
     end function wpic_paraio_ptcl

     integer(c_int) function wpic_paraio_orb(io, up, uf, np2, it) bind(C, name='wpic_paraio_orb')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: io
       real(c_double), intent(in) :: up(*), uf(*)
       integer(c_int), intent(in) :: np2(*)
       integer(c_int), value :: it
     end function wpic_paraio_orb

     integer(c_int) function wpic_particle_solv2d(ctx) bind(C, name='wpic_particle_solv2d')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function wpic_particle_solv2d

     integer(c_int) function wpic_bc_injection2d(ctx, u0) bind(C, name='wpic_bc_injection2d')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: u0
     end function wpic_bc_injection2d

     integer(c_int) function wpic_set_nxe2d(ctx, nxe) bind(C, name='wpic_set_nxe2d')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: nxe
     end function wpic_set_nxe2d

     integer(c_int) function wpic_relocate2d(ctx, seed, it, n0, v0, v_thi, v_the, l_damp, b0, theta_bn, phi_bn) &
          bind(C, name='wpic_relocate2d')
       import :: c_int, c_ptr, c_double, c_long_long
       type(c_ptr), value :: ctx
       integer(c_long_long), value :: seed
       integer(c_int), value :: it, n0
       real(c_double), value :: v0, v_thi, v_the, l_damp, b0, theta_bn, phi_bn
     end function wpic_relocate2d

     integer(c_int) function wpic_particle_solv_vay2d(ctx) bind(C, name='wpic_particle_solv_vay2d')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function wpic_particle_solv_vay2d

     integer(c_int) function wpic_field_fdtd_i2d(ctx) bind(C, name='wpic_field_fdtd_i2d')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function wpic_field_fdtd_i2d

     integer(c_int) function wpic_bc_particle_x2d(ctx) bind(C, name='wpic_bc_particle_x2d')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function wpic_bc_particle_x2d

     integer(c_int) function wpic_bc_particle_y2d(ctx) bind(C, name='wpic_bc_particle_y2d')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function wpic_bc_particle_y2d

     integer(c_int) function wpic_sort_bucket2d(ctx) bind(C, name='wpic_sort_bucket2d')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function wpic_sort_bucket2d

     integer(c_int) function wpic_step2d(ctx, nsteps) bind(C, name='wpic_step2d')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: nsteps
     end function wpic_step2d

     integer(c_int) function wpic_energy2d(ctx, out) bind(C, name='wpic_energy2d')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: out(3)
     end function wpic_energy2d

     integer(c_int) function wpic_ic_beams2d(ctx, seed, n_ppc, udrift, uth_e, uth_i, b0) &
          bind(C, name='wpic_ic_beams2d')
       import :: c_int, c_ptr, c_double, c_long_long
       type(c_ptr), value :: ctx
       integer(c_long_long), value :: seed
       integer(c_int), value :: n_ppc
       real(c_double), value :: udrift, uth_e, uth_i, b0
     end function wpic_ic_beams2d

     ! adjoint (include/wpic_adj.h)
     integer(c_int) function wpic_set_cg_tol2d(ctx, err) bind(C, name='wpic_set_cg_tol2d')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: err
     end function wpic_set_cg_tol2d

     integer(c_int) function wpic_perturb_bz2d(ctx, amp, mode) bind(C, name='wpic_perturb_bz2d')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: amp
       integer(c_int), value :: mode
     end function wpic_perturb_bz2d

     integer(c_int) function wpic_adj_forward2d(ctx, nsteps, every, wb) bind(C, name='wpic_adj_forward2d')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: nsteps, every
       real(c_double), intent(out) :: wb
     end function wpic_adj_forward2d

     integer(c_int) function wpic_adj_backward2d(ctx, seed) bind(C, name='wpic_adj_backward2d')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: seed
     end function wpic_adj_backward2d

     integer(c_int) function wpic_adj_param_grads2d(ctx, udrift, uth_e, bz_mode, out) &
          bind(C, name='wpic_adj_param_grads2d')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: udrift, uth_e
       integer(c_int), value :: bz_mode
       real(c_double), intent(out) :: out(3)
     end function wpic_adj_param_grads2d
  end interface

contains

  ! the library's message for the last failing call, as a Fortran string
  function wpic_error_text() result(msg)
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: s(:)
    integer :: n
    p = wpic_last_error()
    call c_f_pointer(p, s, [4096])
    n = 0
    do while (n < 4096)
       if (s(n+1) == c_null_char) exit
       n = n + 1
    end do
    allocate(character(len=n) :: msg)
    msg = transfer(s(1:n), msg)
  end function wpic_error_text

end module wpic_capi
