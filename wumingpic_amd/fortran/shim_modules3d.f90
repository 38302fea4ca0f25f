! shim_modules3d.f90 -- the reference's 3D module entry points, same names and
! argument lists, implemented on the GPU through the C-ABI of include/wpic3d.h:
!   particle:  particle__init, particle__solv, particle__solv_vay
!                                                  (3d/common/particle.f90:18,52,236)
!   field:     field__init, field__fdtd_i          (3d/common/field.f90:22,70)
!   sort:      sort__init, sort__bucket            (3d/common/sort.f90:16,40)
!   mom_calc:  mom_calc__init, mom_calc__accl, mom_calc__nvt
!                                                  (3d/common/mom_calc.f90:16,49,219)
!   boundary_periodic:
!              __init, __particle_x, __particle_yz, __dfield, __curre, __phi,
!              __mom                               (3d/common/boundary_periodic.f90:25,68,104,
!                                                   458,676,981,1102)
! The module names are the 3d/common ones, which repeat the 2D names: these
! modules are built into their own module directory and archive
! (libwpic_fortran3d.a), which a 3D app links instead of 3d/common's objects.
! The boundary procedures are accepted by field__fdtd_i for source
! compatibility and applied on the device.

module particle
  use iso_c_binding
  use wpic_capi3d
  use wpic_shim_state3d
  implicit none
  private
  public :: particle__init, particle__solv, particle__solv_vay

contains

  subroutine particle__init(ndim_in, np_in, nsp_in, nxgs_in, nxge_in, nygs_in, nyge_in, nzgs_in, nzge_in, &
                            nys_in, nye_in, nzs_in, nze_in, delx_in, delt_in, c_in, q_in, r_in)
    integer, intent(in) :: ndim_in, np_in, nsp_in
    integer, intent(in) :: nxgs_in, nxge_in, nygs_in, nyge_in, nzgs_in, nzge_in, nys_in, nye_in, nzs_in, nze_in
    real(8), intent(in) :: delx_in, delt_in, c_in, q_in(nsp_in), r_in(nsp_in)
    call shim_set_grid(ndim_in, np_in, nsp_in, nxgs_in, nxge_in, nygs_in, nyge_in, nzgs_in, nzge_in, &
                       nys_in, nye_in, nzs_in, nze_in)
    prm%delx = delx_in
    prm%delt = delt_in
    prm%c = c_in
    prm%q = 0
    prm%r = 1
    prm%q(1:nsp_in) = q_in
    prm%r(1:nsp_in) = r_in
    have_particle = .true.
  end subroutine particle__init

  subroutine particle__solv(gp, up, uf, cumcnt, nxs, nxe)
    integer, intent(in)  :: nxs, nxe
    integer, intent(in)  :: cumcnt(s_nxgs:s_nxge+1, s_nys:s_nye, s_nzs:s_nze, s_nsp)
    real(8), intent(in)  :: up(s_ndim, s_np, s_nys:s_nye, s_nzs:s_nze, s_nsp)
    real(8), intent(in)  :: uf(6, s_nxgs-2:s_nxge+2, s_nys-2:s_nye+2, s_nzs-2:s_nze+2)
    real(8), intent(out) :: gp(s_ndim, s_np, s_nys:s_nye, s_nzs:s_nze, s_nsp)
    call push(gp, up, uf, cumcnt, nxe, .false.)
  end subroutine particle__solv

  ! the Vay kick (3d/common/particle.f90:236)
  subroutine particle__solv_vay(gp, up, uf, cumcnt, nxs, nxe)
    integer, intent(in)  :: nxs, nxe
    integer, intent(in)  :: cumcnt(s_nxgs:s_nxge+1, s_nys:s_nye, s_nzs:s_nze, s_nsp)
    real(8), intent(in)  :: up(s_ndim, s_np, s_nys:s_nye, s_nzs:s_nze, s_nsp)
    real(8), intent(in)  :: uf(6, s_nxgs-2:s_nxge+2, s_nys-2:s_nye+2, s_nzs-2:s_nze+2)
    real(8), intent(out) :: gp(s_ndim, s_np, s_nys:s_nye, s_nzs:s_nze, s_nsp)
    call push(gp, up, uf, cumcnt, nxe, .true.)
  end subroutine particle__solv_vay

  subroutine push(gp, up, uf, cumcnt, nxe, vay)
    integer, intent(in)  :: nxe
    integer, intent(in)  :: cumcnt(s_nxgs:s_nxge+1, s_nys:s_nye, s_nzs:s_nze, s_nsp)
    real(8), intent(in)  :: up(s_ndim, s_np, s_nys:s_nye, s_nzs:s_nze, s_nsp)
    real(8), intent(in)  :: uf(6, s_nxgs-2:s_nxge+2, s_nys-2:s_nye+2, s_nzs-2:s_nze+2)
    real(8), intent(out) :: gp(s_ndim, s_np, s_nys:s_nye, s_nzs:s_nze, s_nsp)
    logical, intent(in)  :: vay
    integer(c_int) :: np2(s_nys:s_nye, s_nzs:s_nze, s_nsp)
    if (.not. have_particle) then
       write(6,*) 'Initialize first by calling particle__init()'
       stop
    end if
    ! columns are sorted: the last cell offset is the column count (sort.f90:64-75)
    np2 = cumcnt(nxe+1, :, :, :)
    call shim_ctx(np2)
    if (.not. resident) then
       call shim_check(wpic_upload_up3d(ctx, up, np2, cumcnt, s_np))
       call shim_check(wpic_upload_fields3d(ctx, uf))
    end if
    if (vay) then
       call shim_check(wpic_particle_solv_vay3d(ctx))
    else
       call shim_check(wpic_particle_solv3d(ctx))
    end if
    if (.not. resident) then
       call shim_check(wpic_download_gp3d(ctx, gp, np2, s_np))
       ! the id slot is copied over the whole capacity (particle.f90:227-231)
       if (s_ndim == 7) gp(7, :, :, :, :) = up(7, :, :, :, :)
    end if
  end subroutine push

end module particle


module field
  use iso_c_binding
  use wpic_capi3d
  use wpic_shim_state3d
  implicit none
  private
  public :: field__init, field__fdtd_i

contains

  subroutine field__init(ndim_in, np_in, nsp_in, nxgs_in, nxge_in, nygs_in, nyge_in, nzgs_in, nzge_in, &
                         nys_in, nye_in, nzs_in, nze_in, mnpr_in, ncomw_in, opsum_in, nerr_in, &
                         delx_in, delt_in, c_in, q_in, r_in, gfac_in)
    integer, intent(in) :: ndim_in, np_in, nsp_in
    integer, intent(in) :: nxgs_in, nxge_in, nygs_in, nyge_in, nzgs_in, nzge_in, nys_in, nye_in, nzs_in, nze_in
    integer, intent(in) :: mnpr_in, ncomw_in, opsum_in, nerr_in
    real(8), intent(in) :: delx_in, delt_in, c_in, q_in(nsp_in), r_in(nsp_in), gfac_in
    call shim_set_grid(ndim_in, np_in, nsp_in, nxgs_in, nxge_in, nygs_in, nyge_in, nzgs_in, nzge_in, &
                       nys_in, nye_in, nzs_in, nze_in)
    prm%delx = delx_in
    prm%delt = delt_in
    prm%c = c_in
    prm%gfac = gfac_in
    have_field = .true.
  end subroutine field__init

  subroutine field__fdtd_i(uf, up, gp, cumcnt, nxs, nxe, set_boundary_dfield, set_boundary_curre, set_boundary_phi)
    interface
       subroutine set_boundary_dfield(df, nxs, nxe, nys, nye, nzs, nze, nxgs, nxge)
         real(8), intent(inout) :: df(6, nxgs-2:nxge+2, nys-2:nye+2, nzs-2:nze+2)
         integer, intent(in)    :: nxs, nxe, nys, nye, nzs, nze, nxgs, nxge
       end subroutine set_boundary_dfield
       subroutine set_boundary_curre(uj, nxs, nxe, nys, nye, nzs, nze, nxgs, nxge)
         real(8), intent(inout) :: uj(3, nxgs-2:nxge+2, nys-2:nye+2, nzs-2:nze+2)
         integer, intent(in)    :: nxs, nxe, nys, nye, nzs, nze, nxgs, nxge
       end subroutine set_boundary_curre
       subroutine set_boundary_phi(phi, nxs, nxe, nys, nye, nzs, nze, l)
         real(8), intent(inout) :: phi(nxs-1:nxe+1, nys-1:nye+1, nzs-1:nze+1)
         integer, intent(in)    :: nxs, nxe, nys, nye, nzs, nze, l
       end subroutine set_boundary_phi
    end interface
    integer, intent(in)    :: nxs, nxe
    integer, intent(in)    :: cumcnt(s_nxgs:s_nxge+1, s_nys:s_nye, s_nzs:s_nze, s_nsp)
    real(8), intent(in)    :: gp(s_ndim, s_np, s_nys:s_nye, s_nzs:s_nze, s_nsp)
    real(8), intent(in)    :: up(s_ndim, s_np, s_nys:s_nye, s_nzs:s_nze, s_nsp)
    real(8), intent(inout) :: uf(6, s_nxgs-2:s_nxge+2, s_nys-2:s_nye+2, s_nzs-2:s_nze+2)
    integer(c_int) :: np2(s_nys:s_nye, s_nzs:s_nze, s_nsp)
    if (.not. have_field) then
       write(6,*) 'Initialize first by calling field__init()'
       stop
    end if
    np2 = cumcnt(nxe+1, :, :, :)
    call shim_ctx(np2)
    if (.not. resident) then
       call shim_check(wpic_upload_up3d(ctx, up, np2, cumcnt, s_np))
       call shim_check(wpic_upload_gp3d(ctx, gp, np2, s_np))
       call shim_check(wpic_upload_fields3d(ctx, uf))
    end if
    call shim_check(wpic_field_fdtd_i3d(ctx))
    if (.not. resident) call shim_check(wpic_download_fields3d(ctx, uf))
  end subroutine field__fdtd_i

end module field


module sort
  use iso_c_binding
  use wpic_capi3d
  use wpic_shim_state3d
  implicit none
  private
  public :: sort__init, sort__bucket

contains

  subroutine sort__init(ndim_in, np_in, nsp_in, nxgs_in, nxge_in, nygs_in, nyge_in, nzgs_in, nzge_in, &
                        nys_in, nye_in, nzs_in, nze_in)
    integer, intent(in) :: ndim_in, np_in, nsp_in
    integer, intent(in) :: nxgs_in, nxge_in, nygs_in, nyge_in, nzgs_in, nzge_in, nys_in, nye_in, nzs_in, nze_in
    call shim_set_grid(ndim_in, np_in, nsp_in, nxgs_in, nxge_in, nygs_in, nyge_in, nzgs_in, nzge_in, &
                       nys_in, nye_in, nzs_in, nze_in)
  end subroutine sort__init

  subroutine sort__bucket(gp, up, cumcnt, np2, nxs, nxe)
    integer, intent(in)  :: nxs, nxe
    integer, intent(in)  :: np2(s_nys:s_nye, s_nzs:s_nze, s_nsp)
    integer, intent(out) :: cumcnt(s_nxgs:s_nxge+1, s_nys:s_nye, s_nzs:s_nze, s_nsp)
    real(8), intent(in)  :: up(s_ndim, s_np, s_nys:s_nye, s_nzs:s_nze, s_nsp)
    real(8), intent(out) :: gp(s_ndim, s_np, s_nys:s_nye, s_nzs:s_nze, s_nsp)
    integer(c_int) :: n2(s_nys:s_nye, s_nzs:s_nze, s_nsp)
    call shim_ctx(np2)
    if (.not. resident) call shim_check(wpic_upload_gp3d(ctx, up, np2, s_np))
    call shim_check(wpic_sort_bucket3d(ctx))
    if (.not. resident) call shim_check(wpic_download_up3d(ctx, gp, n2, cumcnt, s_np))
  end subroutine sort__bucket

end module sort


module mom_calc
  use iso_c_binding
  use wpic_capi3d
  use wpic_shim_state3d
  implicit none
  private
  public :: mom_calc__init, mom_calc__accl, mom_calc__nvt

contains

  ! mom_calc__init takes particle__init's arguments; its delt*0.5 is applied
  ! on the device from the context's delt
  subroutine mom_calc__init(ndim_in, np_in, nsp_in, nxgs_in, nxge_in, nygs_in, nyge_in, nzgs_in, nzge_in, &
                            nys_in, nye_in, nzs_in, nze_in, delx_in, delt_in, c_in, q_in, r_in)
    integer, intent(in) :: ndim_in, np_in, nsp_in
    integer, intent(in) :: nxgs_in, nxge_in, nygs_in, nyge_in, nzgs_in, nzge_in, nys_in, nye_in, nzs_in, nze_in
    real(8), intent(in) :: delx_in, delt_in, c_in, q_in(nsp_in), r_in(nsp_in)
    call shim_set_grid(ndim_in, np_in, nsp_in, nxgs_in, nxge_in, nygs_in, nyge_in, nzgs_in, nzge_in, &
                       nys_in, nye_in, nzs_in, nze_in)
    if (have_particle) then
       if (prm%delt /= delt_in .or. prm%c /= c_in .or. prm%delx /= delx_in) then
          write(6,*) 'wpic shim: mom_calc__init must use particle__init''s delx, delt and c'
          stop
       end if
    end if
    have_mom = .true.
  end subroutine mom_calc__init

  subroutine mom_calc__accl(gp, up, uf, cumcnt, nxs, nxe)
    integer, intent(in)  :: nxs, nxe
    integer, intent(in)  :: cumcnt(s_nxgs:s_nxge+1, s_nys:s_nye, s_nzs:s_nze, s_nsp)
    real(8), intent(in)  :: up(s_ndim, s_np, s_nys:s_nye, s_nzs:s_nze, s_nsp)
    real(8), intent(in)  :: uf(6, s_nxgs-2:s_nxge+2, s_nys-2:s_nye+2, s_nzs-2:s_nze+2)
    real(8), intent(out) :: gp(s_ndim, s_np, s_nys:s_nye, s_nzs:s_nze, s_nsp)
    integer(c_int) :: np2(s_nys:s_nye, s_nzs:s_nze, s_nsp)
    if (.not. have_mom) then
       write(6,*) 'Initialize first by calling mom_calc__init()'
       stop
    end if
    np2 = cumcnt(nxe+1, :, :, :)
    call shim_ctx(np2)
    if (.not. resident) then
       call shim_check(wpic_upload_up3d(ctx, up, np2, cumcnt, s_np))
       call shim_check(wpic_upload_fields3d(ctx, uf))
    end if
    call shim_check(wpic_mom_accl3d(ctx))
    if (.not. resident) call shim_check(wpic_download_gp3d(ctx, gp, np2, s_np))
  end subroutine mom_calc__accl

  subroutine mom_calc__nvt(mom, up, np2)
    integer, intent(in)    :: np2(s_nys:s_nye, s_nzs:s_nze, s_nsp)
    real(8), intent(in)    :: up(s_ndim, s_np, s_nys:s_nye, s_nzs:s_nze, s_nsp)
    real(8), intent(inout) :: mom(7, s_nxgs-1:s_nxge+1, s_nys-1:s_nye+1, s_nzs-1:s_nze+1, s_nsp)
    if (.not. have_mom) then
       write(6,*) 'Initialize first by calling mom_calc__init()'
       stop
    end if
    call shim_ctx(np2)
    ! resident: the device's gp (from mom_calc__accl) is the argument
    if (.not. resident) call shim_check(wpic_upload_gp3d(ctx, up, np2, s_np))
    call shim_check(wpic_mom_nvt3d(ctx))
    call shim_check(wpic_download_mom3d(ctx, mom))
  end subroutine mom_calc__nvt

end module mom_calc


module boundary_periodic
  use iso_c_binding
  use wpic_capi3d
  use wpic_shim_state3d
  implicit none
  private
  public :: boundary_periodic__init, boundary_periodic__particle_x, boundary_periodic__particle_yz
  public :: boundary_periodic__dfield, boundary_periodic__curre, boundary_periodic__phi, boundary_periodic__mom

contains

  subroutine boundary_periodic__init(ndim_in, np_in, nsp_in, nxgs_in, nxge_in, nygs_in, nyge_in, nzgs_in, &
                                     nzge_in, nys_in, nye_in, nzs_in, nze_in, jup_in, jdown_in, kup_in, kdown_in, &
                                     mnpi_in, mnpr_in, ncomw_in, nerr_in, nstat_in, delx_in, delt_in, c_in)
    integer, intent(in) :: ndim_in, np_in, nsp_in
    integer, intent(in) :: nxgs_in, nxge_in, nygs_in, nyge_in, nzgs_in, nzge_in, nys_in, nye_in, nzs_in, nze_in
    integer, intent(in) :: jup_in, jdown_in, kup_in, kdown_in, mnpi_in, mnpr_in, ncomw_in, nerr_in, nstat_in(:)
    real(8), intent(in) :: delx_in, delt_in, c_in
    call shim_set_grid(ndim_in, np_in, nsp_in, nxgs_in, nxge_in, nygs_in, nyge_in, nzgs_in, nzge_in, &
                       nys_in, nye_in, nzs_in, nze_in)
  end subroutine boundary_periodic__init

  subroutine boundary_periodic__particle_x(up, np2)
    integer, intent(in)    :: np2(s_nys:s_nye, s_nzs:s_nze, s_nsp)
    real(8), intent(inout) :: up(s_ndim, s_np, s_nys:s_nye, s_nzs:s_nze, s_nsp)
    integer(c_int) :: n2(s_nys:s_nye, s_nzs:s_nze, s_nsp)
    call shim_ctx(np2)
    if (.not. resident) call shim_check(wpic_upload_gp3d(ctx, up, np2, s_np))
    call shim_check(wpic_bc_particle_x3d(ctx))
    if (.not. resident) call shim_check(wpic_download_gp3d(ctx, up, n2, s_np))
  end subroutine boundary_periodic__particle_x

  subroutine boundary_periodic__particle_yz(up, np2)
    integer, intent(inout) :: np2(s_nys:s_nye, s_nzs:s_nze, s_nsp)
    real(8), intent(inout) :: up(s_ndim, s_np, s_nys:s_nye, s_nzs:s_nze, s_nsp)
    integer(c_int) :: n2(s_nys:s_nye, s_nzs:s_nze, s_nsp)
    call shim_ctx(np2)
    if (.not. resident) call shim_check(wpic_upload_gp3d(ctx, up, np2, s_np))
    call shim_check(wpic_bc_particle_yz3d(ctx))
    if (.not. resident) then
       call shim_check(wpic_download_gp3d(ctx, up, n2, s_np))
       np2 = n2
    end if
  end subroutine boundary_periodic__particle_yz

  ! The three boundary procedures are handed to field__fdtd_i, which applies
  ! the same boundary on the device; on host arrays they are not provided.
  subroutine boundary_periodic__dfield(df, nxs, nxe, nys, nye, nzs, nze, nxgs, nxge)
    integer, intent(in)    :: nxs, nxe, nys, nye, nzs, nze, nxgs, nxge
    real(8), intent(inout) :: df(6, nxgs-2:nxge+2, nys-2:nye+2, nzs-2:nze+2)
    write(6,*) 'boundary_periodic__dfield: applied on the device inside field__fdtd_i'
    stop
  end subroutine boundary_periodic__dfield

  subroutine boundary_periodic__curre(uj, nxs, nxe, nys, nye, nzs, nze, nxgs, nxge)
    integer, intent(in)    :: nxs, nxe, nys, nye, nzs, nze, nxgs, nxge
    real(8), intent(inout) :: uj(3, nxgs-2:nxge+2, nys-2:nye+2, nzs-2:nze+2)
    write(6,*) 'boundary_periodic__curre: applied on the device inside field__fdtd_i'
    stop
  end subroutine boundary_periodic__curre

  subroutine boundary_periodic__phi(phi, nxs, nxe, nys, nye, nzs, nze, l)
    integer, intent(in)    :: nxs, nxe, nys, nye, nzs, nze, l
    real(8), intent(inout) :: phi(nxs-1:nxe+1, nys-1:nye+1, nzs-1:nze+1)
    write(6,*) 'boundary_periodic__phi: applied on the device inside field__fdtd_i'
    stop
  end subroutine boundary_periodic__phi

  ! the moment folds on the device's moments; mirror mode sends the host
  ! array up and brings the folded one back
  subroutine boundary_periodic__mom(mom)
    real(8), intent(inout) :: mom(7, s_nxgs-1:s_nxge+1, s_nys-1:s_nye+1, s_nzs-1:s_nze+1, s_nsp)
    if (.not. c_associated(ctx)) then
       write(6,*) 'boundary_periodic__mom: no device state (call mom_calc__accl first)'
       stop
    end if
    if (.not. resident) call shim_check(wpic_upload_mom3d(ctx, mom))
    call shim_check(wpic_bc_mom3d(ctx))
    call shim_check(wpic_download_mom3d(ctx, mom))
  end subroutine boundary_periodic__mom

end module boundary_periodic


! shim controls beyond the reference's modules (transfer policy, mode)
module wpic_shim3d
  use iso_c_binding
  use wpic_capi3d
  use wpic_shim_state3d
  implicit none
  private
  public :: wpic_shim_resident, wpic_shim_mode, wpic_shim_push, wpic_shim_pull, wpic_shim_finalize

contains

  ! resident mode: entry points move no host data; wpic_shim_push loads the
  ! state once, wpic_shim_pull brings it back for host readers
  subroutine wpic_shim_resident(on)
    logical, intent(in) :: on
    resident = on
  end subroutine wpic_shim_resident

  ! copy the app's sorted particles (with cumcnt) and fields to the device
  subroutine wpic_shim_push(up, uf, np2, cumcnt)
    real(8), intent(in) :: up(s_ndim, s_np, s_nys:s_nye, s_nzs:s_nze, s_nsp)
    real(8), intent(in) :: uf(6, s_nxgs-2:s_nxge+2, s_nys-2:s_nye+2, s_nzs-2:s_nze+2)
    integer, intent(in) :: np2(s_nys:s_nye, s_nzs:s_nze, s_nsp)
    integer, intent(in) :: cumcnt(s_nxgs:s_nxge+1, s_nys:s_nye, s_nzs:s_nze, s_nsp)
    call shim_ctx(np2)
    call shim_check(wpic_upload_up3d(ctx, up, np2, cumcnt, s_np))
    call shim_check(wpic_upload_fields3d(ctx, uf))
  end subroutine wpic_shim_push

  ! arithmetic mode: WPIC_MODE_FAST or WPIC_MODE_PARITY (before first use)
  subroutine wpic_shim_mode(mode)
    integer(c_int), intent(in) :: mode
    s_mode = mode
  end subroutine wpic_shim_mode

  ! copy the device's sorted particles, cell offsets and fields to the host
  subroutine wpic_shim_pull(up, uf, np2, cumcnt)
    real(8), intent(inout) :: up(s_ndim, s_np, s_nys:s_nye, s_nzs:s_nze, s_nsp)
    real(8), intent(out)   :: uf(6, s_nxgs-2:s_nxge+2, s_nys-2:s_nye+2, s_nzs-2:s_nze+2)
    integer, intent(out)   :: np2(s_nys:s_nye, s_nzs:s_nze, s_nsp)
    integer, intent(out)   :: cumcnt(s_nxgs:s_nxge+1, s_nys:s_nye, s_nzs:s_nze, s_nsp)
    call shim_check(wpic_download_up3d(ctx, up, np2, cumcnt, s_np))
    call shim_check(wpic_download_fields3d(ctx, uf))
  end subroutine wpic_shim_pull

  subroutine wpic_shim_finalize()
    integer(c_int) :: rc
    if (c_associated(ctx)) rc = wpic_destroy3d(ctx)
    ctx = c_null_ptr
  end subroutine wpic_shim_finalize

end module wpic_shim3d
