! wpic_capi3d.f90 -- ISO_C_BINDING interfaces to the 3D C-ABI of
! include/wpic3d.h (the entry points the 3d/common shim modules call).
module wpic_capi3d
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: WPIC_MODE_FAST = 0, WPIC_MODE_PARITY = 1, WPIC_MODE_MIXED = 2
  integer(c_int), parameter :: WPIC_PUSHER_BORIS = 0, WPIC_PUSHER_VAY = 1

  type, bind(C) :: wpic_params3d
     integer(c_int) :: ndim, nsp
     integer(c_int) :: nxgs, nxge, nygs, nyge, nzgs, nzge
     real(c_double) :: delx, delt, c, gfac
     real(c_double) :: q(4), r(4)
     integer(c_int) :: mode
     integer(c_long_long) :: capacity
  end type wpic_params3d

  interface
     integer(c_int) function wpic_create3d(ctx, p) bind(C, name='wpic_create3d')
       import :: c_int, c_ptr, wpic_params3d
       type(c_ptr), intent(out) :: ctx
       type(wpic_params3d), intent(in) :: p
     end function wpic_create3d

     integer(c_int) function wpic_destroy3d(ctx) bind(C, name='wpic_destroy3d')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function wpic_destroy3d

     type(c_ptr) function wpic_last_error() bind(C, name='wpic_last_error')
       import :: c_ptr
     end function wpic_last_error

     integer(c_int) function wpic_upload_up3d(ctx, up, np2, cumcnt, np) bind(C, name='wpic_upload_up3d')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: up(*)
       integer(c_int), intent(in) :: np2(*), cumcnt(*)
       integer(c_int), value :: np
     end function wpic_upload_up3d

     integer(c_int) function wpic_download_up3d(ctx, up, np2, cumcnt, np) bind(C, name='wpic_download_up3d')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(inout) :: up(*)
       integer(c_int), intent(out) :: np2(*), cumcnt(*)
       integer(c_int), value :: np
     end function wpic_download_up3d

     integer(c_int) function wpic_upload_gp3d(ctx, gp, np2, np) bind(C, name='wpic_upload_gp3d')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: gp(*)
       integer(c_int), intent(in) :: np2(*)
       integer(c_int), value :: np
     end function wpic_upload_gp3d

     integer(c_int) function wpic_download_gp3d(ctx, gp, np2, np) bind(C, name='wpic_download_gp3d')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(inout) :: gp(*)
       integer(c_int), intent(out) :: np2(*)
       integer(c_int), value :: np
     end function wpic_download_gp3d

     integer(c_int) function wpic_upload_fields3d(ctx, uf) bind(C, name='wpic_upload_fields3d')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: uf(*)
     end function wpic_upload_fields3d

     integer(c_int) function wpic_download_fields3d(ctx, uf) bind(C, name='wpic_download_fields3d')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: uf(*)
     end function wpic_download_fields3d

     integer(c_int) function wpic_particle_solv3d(ctx) bind(C, name='wpic_particle_solv3d')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function wpic_particle_solv3d

     integer(c_int) function wpic_particle_solv_vay3d(ctx) bind(C, name='wpic_particle_solv_vay3d')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function wpic_particle_solv_vay3d

     integer(c_int) function wpic_field_fdtd_i3d(ctx) bind(C, name='wpic_field_fdtd_i3d')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function wpic_field_fdtd_i3d

     integer(c_int) function wpic_bc_particle_x3d(ctx) bind(C, name='wpic_bc_particle_x3d')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function wpic_bc_particle_x3d

     integer(c_int) function wpic_bc_particle_yz3d(ctx) bind(C, name='wpic_bc_particle_yz3d')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function wpic_bc_particle_yz3d

     integer(c_int) function wpic_sort_bucket3d(ctx) bind(C, name='wpic_sort_bucket3d')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function wpic_sort_bucket3d

     integer(c_int) function wpic_mom_accl3d(ctx) bind(C, name='wpic_mom_accl3d')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function wpic_mom_accl3d

     integer(c_int) function wpic_mom_nvt3d(ctx) bind(C, name='wpic_mom_nvt3d')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function wpic_mom_nvt3d

     integer(c_int) function wpic_bc_mom3d(ctx) bind(C, name='wpic_bc_mom3d')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function wpic_bc_mom3d

     integer(c_int) function wpic_download_mom3d(ctx, mom) bind(C, name='wpic_download_mom3d')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: mom(*)
     end function wpic_download_mom3d

     integer(c_int) function wpic_upload_mom3d(ctx, mom) bind(C, name='wpic_upload_mom3d')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: mom(*)
     end function wpic_upload_mom3d
  end interface

contains

  ! the library's last error message as a Fortran string
  function wpic_error_text3d() result(msg)
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
  end function wpic_error_text3d

end module wpic_capi3d
