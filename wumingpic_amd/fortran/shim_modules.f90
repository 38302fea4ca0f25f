! shim_modules.f90 -- the reference's 2D module entry points, same names and
! argument lists, implemented on the GPU through the C-ABI:
!   particle:  particle__init, particle__solv, particle__solv_vay
!                                                  (2d/common/particle.f90:18,48,182)
!   field:     field__init, field__fdtd_i          (2d/common/field.f90:22,66)
!   sort:      sort__init, sort__bucket            (2d/common/sort.f90:16,36)
!   mom_calc:  mom_calc__init, mom_calc__accl, mom_calc__nvt
!                                                  (2d/common/mom_calc.f90:18,48,167)
!   paraio:    paraio__init, __finalize, __output, __input, __param, __mom,
!              __ptcl, __orb                       (2d/common/paraio.f90:37-741)
!   boundary_periodic / boundary_reconnection / boundary_shock:
!              *__init, *__particle_x, *__particle_y, *__mom and the boundary
!              procedures handed to field__fdtd_i  (boundary_periodic.f90:26-636,
!              2d/proj/reconnection/boundary_reconnection.f90:26-650,
!              2d/proj/shock/boundary_shock.f90:27-693, with __injection)
! An app built against these modules instead of the reference's runs its
! per-step loop on one MI355X.  The boundary procedures are accepted by
! field__fdtd_i for source compatibility and applied on the device according
! to the boundary module whose __init was called.

module particle
  use iso_c_binding
  use wpic_capi
  use wpic_shim_state
  implicit none
  private
  public :: particle__init, particle__solv, particle__solv_vay

contains

  subroutine particle__init(ndim_in, np_in, nsp_in, nxgs_in, nxge_in, nygs_in, nyge_in, nys_in, nye_in, &
                            delx_in, delt_in, c_in, q_in, r_in)
    integer, intent(in) :: ndim_in, np_in, nsp_in
    integer, intent(in) :: nxgs_in, nxge_in, nygs_in, nyge_in, nys_in, nye_in
    real(8), intent(in) :: delx_in, delt_in, c_in, q_in(nsp_in), r_in(nsp_in)
    call shim_set_grid(ndim_in, np_in, nsp_in, nxgs_in, nxge_in, nygs_in, nyge_in, nys_in, nye_in)
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
    integer, intent(in)  :: cumcnt(s_nxgs:s_nxge+1, s_nys:s_nye, s_nsp)
    real(8), intent(in)  :: up(s_ndim, s_np, s_nys:s_nye, s_nsp)
    real(8), intent(in)  :: uf(6, s_nxgs-2:s_nxge+2, s_nys-2:s_nye+2)
    real(8), intent(out) :: gp(s_ndim, s_np, s_nys:s_nye, s_nsp)
    call push(gp, up, uf, cumcnt, nxe, .false.)
  end subroutine particle__solv

  ! the Vay kick (particle.f90:182-315)
  subroutine particle__solv_vay(gp, up, uf, cumcnt, nxs, nxe)
    integer, intent(in)  :: nxs, nxe
    integer, intent(in)  :: cumcnt(s_nxgs:s_nxge+1, s_nys:s_nye, s_nsp)
    real(8), intent(in)  :: up(s_ndim, s_np, s_nys:s_nye, s_nsp)
    real(8), intent(in)  :: uf(6, s_nxgs-2:s_nxge+2, s_nys-2:s_nye+2)
    real(8), intent(out) :: gp(s_ndim, s_np, s_nys:s_nye, s_nsp)
    call push(gp, up, uf, cumcnt, nxe, .true.)
  end subroutine particle__solv_vay

  subroutine push(gp, up, uf, cumcnt, nxe, vay)
    integer, intent(in)  :: nxe
    integer, intent(in)  :: cumcnt(s_nxgs:s_nxge+1, s_nys:s_nye, s_nsp)
    real(8), intent(in)  :: up(s_ndim, s_np, s_nys:s_nye, s_nsp)
    real(8), intent(in)  :: uf(6, s_nxgs-2:s_nxge+2, s_nys-2:s_nye+2)
    real(8), intent(out) :: gp(s_ndim, s_np, s_nys:s_nye, s_nsp)
    logical, intent(in)  :: vay
    integer(c_int) :: np2(s_nys:s_nye, s_nsp), cc(s_nxgs:s_nxge+1, s_nys:s_nye, s_nsp)
    if (.not. have_particle) then
       write(6,*) 'Initialize first by calling particle__init()'
       stop
    end if
    ! rows are sorted: the last cell offset is the row count (sort.f90:64-69)
    np2 = cumcnt(nxe+1, :, :)
    call shim_ctx(np2)
    call shim_edge(nxe)
    if (.not. resident) then
       call shim_cumcnt(cumcnt, nxe, cc)
       call shim_check(wpic_upload_up2d(ctx, up, np2, cc, s_np))
       call shim_check(wpic_upload_fields2d(ctx, uf))
    end if
    if (vay) then
       call shim_check(wpic_particle_solv_vay2d(ctx))
    else
       call shim_check(wpic_particle_solv2d(ctx))
    end if
    if (.not. resident) then
       call shim_check(wpic_download_gp2d(ctx, gp, np2, s_np))
       ! the id slot is copied over the whole capacity (particle.f90:173-177)
       if (s_ndim == 6) gp(6, :, :, :) = up(6, :, :, :)
    end if
  end subroutine push

end module particle


module field
  use iso_c_binding
  use wpic_capi
  use wpic_shim_state
  implicit none
  private
  public :: field__init, field__fdtd_i

contains

  subroutine field__init(ndim_in, np_in, nsp_in, nxgs_in, nxge_in, nygs_in, nyge_in, nys_in, nye_in, &
                         mnpr_in, ncomw_in, opsum_in, nerr_in, delx_in, delt_in, c_in, q_in, r_in, gfac_in)
    integer, intent(in) :: ndim_in, np_in, nsp_in
    integer, intent(in) :: nxgs_in, nxge_in, nygs_in, nyge_in, nys_in, nye_in
    integer, intent(in) :: mnpr_in, ncomw_in, opsum_in, nerr_in
    real(8), intent(in) :: delx_in, delt_in, c_in, q_in(nsp_in), r_in(nsp_in), gfac_in
    call shim_set_grid(ndim_in, np_in, nsp_in, nxgs_in, nxge_in, nygs_in, nyge_in, nys_in, nye_in)
    prm%delx = delx_in
    prm%delt = delt_in
    prm%c = c_in
    prm%gfac = gfac_in
    have_field = .true.
  end subroutine field__init

  subroutine field__fdtd_i(uf, up, gp, cumcnt, nxs, nxe, set_boundary_dfield, set_boundary_curre, set_boundary_phi)
    interface
       subroutine set_boundary_dfield(df, nxs, nxe, nys, nye, nxgs, nxge)
         real(8), intent(inout) :: df(6, nxgs-2:nxge+2, nys-2:nye+2)
         integer, intent(in)    :: nxs, nxe, nys, nye, nxgs, nxge
       end subroutine set_boundary_dfield
       subroutine set_boundary_curre(uj, nxs, nxe, nys, nye, nxgs, nxge)
         real(8), intent(inout) :: uj(3, nxgs-2:nxge+2, nys-2:nye+2)
         integer, intent(in)    :: nxs, nxe, nys, nye, nxgs, nxge
       end subroutine set_boundary_curre
       subroutine set_boundary_phi(phi, nxs, nxe, nys, nye, l)
         real(8), intent(inout) :: phi(nxs-1:nxe+1, nys-1:nye+1)
         integer, intent(in)    :: nxs, nxe, nys, nye, l
       end subroutine set_boundary_phi
    end interface
    integer, intent(in)    :: nxs, nxe
    integer, intent(in)    :: cumcnt(s_nxgs:s_nxge+1, s_nys:s_nye, s_nsp)
    real(8), intent(in)    :: gp(s_ndim, s_np, s_nys:s_nye, s_nsp)
    real(8), intent(in)    :: up(s_ndim, s_np, s_nys:s_nye, s_nsp)
    real(8), intent(inout) :: uf(6, s_nxgs-2:s_nxge+2, s_nys-2:s_nye+2)
    integer(c_int) :: np2(s_nys:s_nye, s_nsp), cc(s_nxgs:s_nxge+1, s_nys:s_nye, s_nsp)
    if (.not. have_field) then
       write(6,*) 'Initialize first by calling field__init()'
       stop
    end if
    np2 = cumcnt(nxe+1, :, :)
    call shim_ctx(np2)
    call shim_edge(nxe)
    if (.not. resident) then
       call shim_cumcnt(cumcnt, nxe, cc)
       call shim_check(wpic_upload_up2d(ctx, up, np2, cc, s_np))
       call shim_check(wpic_upload_gp2d(ctx, gp, np2, s_np))
       call shim_check(wpic_upload_fields2d(ctx, uf))
    end if
    call shim_check(wpic_field_fdtd_i2d(ctx))
    if (.not. resident) call shim_check(wpic_download_fields2d(ctx, uf))
  end subroutine field__fdtd_i

end module field


module sort
  use iso_c_binding
  use wpic_capi
  use wpic_shim_state
  implicit none
  private
  public :: sort__init, sort__bucket

contains

  subroutine sort__init(ndim_in, np_in, nsp_in, nxgs_in, nxge_in, nygs_in, nyge_in, nys_in, nye_in)
    integer, intent(in) :: ndim_in, np_in, nsp_in
    integer, intent(in) :: nxgs_in, nxge_in, nygs_in, nyge_in, nys_in, nye_in
    call shim_set_grid(ndim_in, np_in, nsp_in, nxgs_in, nxge_in, nygs_in, nyge_in, nys_in, nye_in)
  end subroutine sort__init

  subroutine sort__bucket(gp, up, cumcnt, np2, nxs, nxe)
    integer, intent(in)  :: nxs, nxe
    integer, intent(in)  :: np2(s_nys:s_nye, s_nsp)
    integer, intent(out) :: cumcnt(s_nxgs:s_nxge+1, s_nys:s_nye, s_nsp)
    real(8), intent(in)  :: up(s_ndim, s_np, s_nys:s_nye, s_nsp)
    real(8), intent(out) :: gp(s_ndim, s_np, s_nys:s_nye, s_nsp)
    integer(c_int) :: n2(s_nys:s_nye, s_nsp)
    call shim_ctx(np2)
    call shim_edge(nxe)
    if (.not. resident) call shim_check(wpic_upload_gp2d(ctx, up, np2, s_np))
    call shim_check(wpic_sort_bucket2d(ctx))
    if (.not. resident) then
       call shim_check(wpic_download_up2d(ctx, gp, n2, cumcnt, s_np))
    end if
  end subroutine sort__bucket

end module sort


module mom_calc
  use iso_c_binding
  use wpic_capi
  use wpic_shim_state
  implicit none
  private
  public :: mom_calc__init, mom_calc__accl, mom_calc__nvt

contains

  ! mom_calc__init takes particle__init's arguments; its delt*0.5 is applied
  ! on the device from the context's delt
  subroutine mom_calc__init(ndim_in, np_in, nsp_in, nxgs_in, nxge_in, nygs_in, nyge_in, nys_in, nye_in, &
                            delx_in, delt_in, c_in, q_in, r_in)
    integer, intent(in) :: ndim_in, np_in, nsp_in
    integer, intent(in) :: nxgs_in, nxge_in, nygs_in, nyge_in, nys_in, nye_in
    real(8), intent(in) :: delx_in, delt_in, c_in, q_in(nsp_in), r_in(nsp_in)
    call shim_set_grid(ndim_in, np_in, nsp_in, nxgs_in, nxge_in, nygs_in, nyge_in, nys_in, nye_in)
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
    integer, intent(in)  :: cumcnt(s_nxgs:s_nxge+1, s_nys:s_nye, s_nsp)
    real(8), intent(in)  :: up(s_ndim, s_np, s_nys:s_nye, s_nsp)
    real(8), intent(in)  :: uf(6, s_nxgs-2:s_nxge+2, s_nys-2:s_nye+2)
    real(8), intent(out) :: gp(s_ndim, s_np, s_nys:s_nye, s_nsp)
    integer(c_int) :: np2(s_nys:s_nye, s_nsp), cc(s_nxgs:s_nxge+1, s_nys:s_nye, s_nsp)
    if (.not. have_mom) then
       write(6,*) 'Initialize first by calling mom_calc__init()'
       stop
    end if
    np2 = cumcnt(nxe+1, :, :)
    call shim_ctx(np2)
    call shim_edge(nxe)
    if (.not. resident) then
       call shim_cumcnt(cumcnt, nxe, cc)
       call shim_check(wpic_upload_up2d(ctx, up, np2, cc, s_np))
       call shim_check(wpic_upload_fields2d(ctx, uf))
    end if
    call shim_check(wpic_mom_accl2d(ctx))
    if (.not. resident) call shim_check(wpic_download_gp2d(ctx, gp, np2, s_np))
  end subroutine mom_calc__accl

  subroutine mom_calc__nvt(mom, up, np2)
    integer, intent(in)  :: np2(s_nys:s_nye, s_nsp)
    real(8), intent(in)  :: up(s_ndim, s_np, s_nys:s_nye, s_nsp)
    real(8), intent(out) :: mom(7, s_nxgs-1:s_nxge+1, s_nys-1:s_nye+1, s_nsp)
    if (.not. have_mom) then
       write(6,*) 'Initialize first by calling mom_calc__init()'
       stop
    end if
    call shim_ctx(np2)
    ! resident: the device's gp (from mom_calc__accl) is the argument
    if (.not. resident) call shim_check(wpic_upload_gp2d(ctx, up, np2, s_np))
    call shim_check(wpic_mom_nvt2d(ctx))
    call shim_check(wpic_download_mom2d(ctx, mom))
  end subroutine mom_calc__nvt

end module mom_calc


module boundary_periodic
  use iso_c_binding
  use wpic_capi
  use wpic_shim_state
  implicit none
  private
  public :: boundary_periodic__init, boundary_periodic__dfield, boundary_periodic__particle_x
  public :: boundary_periodic__particle_y, boundary_periodic__curre, boundary_periodic__phi, boundary_periodic__mom

contains

  subroutine boundary_periodic__init(ndim_in, np_in, nsp_in, nxgs_in, nxge_in, nygs_in, nyge_in, nys_in, nye_in, &
                                     nup_in, ndown_in, mnpi_in, mnpr_in, ncomw_in, nerr_in, nstat_in, &
                                     delx_in, delt_in, c_in)
    integer, intent(in) :: ndim_in, np_in, nsp_in
    integer, intent(in) :: nxgs_in, nxge_in, nygs_in, nyge_in, nys_in, nye_in
    integer, intent(in) :: nup_in, ndown_in, mnpi_in, mnpr_in, ncomw_in, nerr_in, nstat_in(:)
    real(8), intent(in) :: delx_in, delt_in, c_in
    call shim_set_grid(ndim_in, np_in, nsp_in, nxgs_in, nxge_in, nygs_in, nyge_in, nys_in, nye_in)
    prm%bc = WPIC_BC_PERIODIC
    have_bc = .true.
  end subroutine boundary_periodic__init

  subroutine boundary_periodic__particle_x(up, np2)
    integer, intent(in)    :: np2(s_nys:s_nye, s_nsp)
    real(8), intent(inout) :: up(s_ndim, s_np, s_nys:s_nye, s_nsp)
    integer(c_int) :: n2(s_nys:s_nye, s_nsp)
    call shim_ctx(np2)
    if (.not. resident) call shim_check(wpic_upload_gp2d(ctx, up, np2, s_np))
    call shim_check(wpic_bc_particle_x2d(ctx))
    if (.not. resident) call shim_check(wpic_download_gp2d(ctx, up, n2, s_np))
  end subroutine boundary_periodic__particle_x

  subroutine boundary_periodic__particle_y(up, np2)
    integer, intent(inout) :: np2(s_nys:s_nye, s_nsp)
    real(8), intent(inout) :: up(s_ndim, s_np, s_nys:s_nye, s_nsp)
    integer(c_int) :: n2(s_nys:s_nye, s_nsp)
    call shim_ctx(np2)
    if (.not. resident) call shim_check(wpic_upload_gp2d(ctx, up, np2, s_np))
    call shim_check(wpic_bc_particle_y2d(ctx))
    if (.not. resident) then
       call shim_check(wpic_download_gp2d(ctx, up, n2, s_np))
       np2 = n2
    end if
  end subroutine boundary_periodic__particle_y

  ! The three boundary procedures are handed to field__fdtd_i, which applies
  ! the same boundary on the device; on host arrays they are not provided.
  subroutine boundary_periodic__dfield(df, nxs, nxe, nys, nye, nxgs, nxge)
    integer, intent(in)    :: nxs, nxe, nys, nye, nxgs, nxge
    real(8), intent(inout) :: df(6, nxgs-2:nxge+2, nys-2:nye+2)
    write(6,*) 'boundary_periodic__dfield: applied on the device inside field__fdtd_i'
    stop
  end subroutine boundary_periodic__dfield

  subroutine boundary_periodic__curre(uj, nxs, nxe, nys, nye, nxgs, nxge)
    integer, intent(in)    :: nxs, nxe, nys, nye, nxgs, nxge
    real(8), intent(inout) :: uj(3, nxgs-2:nxge+2, nys-2:nye+2)
    write(6,*) 'boundary_periodic__curre: applied on the device inside field__fdtd_i'
    stop
  end subroutine boundary_periodic__curre

  subroutine boundary_periodic__phi(phi, nxs, nxe, nys, nye, l)
    integer, intent(in)    :: nxs, nxe, nys, nye, l
    real(8), intent(inout) :: phi(nxs-1:nxe+1, nys-1:nye+1)
    write(6,*) 'boundary_periodic__phi: applied on the device inside field__fdtd_i'
    stop
  end subroutine boundary_periodic__phi

  ! the moment folds on the device's moments; mirror mode sends the host
  ! array up and brings the folded one back
  subroutine boundary_periodic__mom(mom)
    real(8), intent(inout) :: mom(7, s_nxgs-1:s_nxge+1, s_nys-1:s_nye+1, s_nsp)
    if (.not. c_associated(ctx)) then
       write(6,*) 'boundary_periodic__mom: no device state (call mom_calc__accl first)'
       stop
    end if
    if (.not. resident) call shim_check(wpic_upload_mom2d(ctx, mom))
    call shim_check(wpic_bc_mom2d(ctx))
    call shim_check(wpic_download_mom2d(ctx, mom))
  end subroutine boundary_periodic__mom

end module boundary_periodic


module boundary_reconnection
  use iso_c_binding
  use wpic_capi
  use wpic_shim_state
  implicit none
  private
  public :: boundary_reconnection__init, boundary_reconnection__dfield, boundary_reconnection__particle_x
  public :: boundary_reconnection__particle_y, boundary_reconnection__curre, boundary_reconnection__phi, boundary_reconnection__mom

contains

  subroutine boundary_reconnection__init(ndim_in, np_in, nsp_in, nxgs_in, nxge_in, nygs_in, nyge_in, nys_in, &
                                         nye_in, nup_in, ndown_in, mnpi_in, mnpr_in, ncomw_in, nerr_in, nstat_in, &
                                         delx_in, delt_in, c_in)
    integer, intent(in) :: ndim_in, np_in, nsp_in
    integer, intent(in) :: nxgs_in, nxge_in, nygs_in, nyge_in, nys_in, nye_in
    integer, intent(in) :: nup_in, ndown_in, mnpi_in, mnpr_in, ncomw_in, nerr_in, nstat_in(:)
    real(8), intent(in) :: delx_in, delt_in, c_in
    call shim_set_grid(ndim_in, np_in, nsp_in, nxgs_in, nxge_in, nygs_in, nyge_in, nys_in, nye_in)
    prm%bc = WPIC_BC_WALLX
    have_bc = .true.
  end subroutine boundary_reconnection__init

  subroutine boundary_reconnection__particle_x(up, np2, nxs, nxe)
    integer, intent(in)    :: nxs, nxe
    integer, intent(in)    :: np2(s_nys:s_nye, s_nsp)
    real(8), intent(inout) :: up(s_ndim, s_np, s_nys:s_nye, s_nsp)
    integer(c_int) :: n2(s_nys:s_nye, s_nsp)
    call shim_ctx(np2)
    call shim_edge(nxe)
    if (.not. resident) call shim_check(wpic_upload_gp2d(ctx, up, np2, s_np))
    call shim_check(wpic_bc_particle_x2d(ctx))
    if (.not. resident) call shim_check(wpic_download_gp2d(ctx, up, n2, s_np))
  end subroutine boundary_reconnection__particle_x

  subroutine boundary_reconnection__particle_y(up, np2)
    integer, intent(inout) :: np2(s_nys:s_nye, s_nsp)
    real(8), intent(inout) :: up(s_ndim, s_np, s_nys:s_nye, s_nsp)
    integer(c_int) :: n2(s_nys:s_nye, s_nsp)
    call shim_ctx(np2)
    if (.not. resident) call shim_check(wpic_upload_gp2d(ctx, up, np2, s_np))
    call shim_check(wpic_bc_particle_y2d(ctx))
    if (.not. resident) then
       call shim_check(wpic_download_gp2d(ctx, up, n2, s_np))
       np2 = n2
    end if
  end subroutine boundary_reconnection__particle_y

  subroutine boundary_reconnection__dfield(df, nxs, nxe, nys, nye, nxgs, nxge)
    integer, intent(in)    :: nxs, nxe, nys, nye, nxgs, nxge
    real(8), intent(inout) :: df(6, nxgs-2:nxge+2, nys-2:nye+2)
    write(6,*) 'boundary_reconnection__dfield: applied on the device inside field__fdtd_i'
    stop
  end subroutine boundary_reconnection__dfield

  subroutine boundary_reconnection__curre(uj, nxs, nxe, nys, nye, nxgs, nxge)
    integer, intent(in)    :: nxs, nxe, nys, nye, nxgs, nxge
    real(8), intent(inout) :: uj(3, nxgs-2:nxge+2, nys-2:nye+2)
    write(6,*) 'boundary_reconnection__curre: applied on the device inside field__fdtd_i'
    stop
  end subroutine boundary_reconnection__curre

  subroutine boundary_reconnection__phi(phi, nxs, nxe, nys, nye, l)
    integer, intent(in)    :: nxs, nxe, nys, nye, l
    real(8), intent(inout) :: phi(nxs-1:nxe+1, nys-1:nye+1)
    write(6,*) 'boundary_reconnection__phi: applied on the device inside field__fdtd_i'
    stop
  end subroutine boundary_reconnection__phi

  ! the moment folds on the device's moments; mirror mode sends the host
  ! array up and brings the folded one back
  subroutine boundary_reconnection__mom(mom)
    real(8), intent(inout) :: mom(7, s_nxgs-1:s_nxge+1, s_nys-1:s_nye+1, s_nsp)
    if (.not. c_associated(ctx)) then
       write(6,*) 'boundary_reconnection__mom: no device state (call mom_calc__accl first)'
       stop
    end if
    if (.not. resident) call shim_check(wpic_upload_mom2d(ctx, mom))
    call shim_check(wpic_bc_mom2d(ctx))
    call shim_check(wpic_download_mom2d(ctx, mom))
  end subroutine boundary_reconnection__mom

end module boundary_reconnection


! the shock app's boundaries (2d/proj/shock/boundary_shock.f90) at the active
! edge nxe each call passes; the field boundaries are applied on the device
module boundary_shock
  use iso_c_binding
  use wpic_capi
  use wpic_shim_state
  implicit none
  private
  public :: boundary_shock__init, boundary_shock__dfield, boundary_shock__particle_x
  public :: boundary_shock__particle_y, boundary_shock__curre, boundary_shock__phi, boundary_shock__mom
  public :: boundary_shock__injection

contains

  subroutine boundary_shock__init(ndim_in, np_in, nsp_in, nxgs_in, nxge_in, nygs_in, nyge_in, nys_in, &
                                         nye_in, nup_in, ndown_in, mnpi_in, mnpr_in, ncomw_in, nerr_in, nstat_in, &
                                         delx_in, delt_in, c_in)
    integer, intent(in) :: ndim_in, np_in, nsp_in
    integer, intent(in) :: nxgs_in, nxge_in, nygs_in, nyge_in, nys_in, nye_in
    integer, intent(in) :: nup_in, ndown_in, mnpi_in, mnpr_in, ncomw_in, nerr_in, nstat_in(:)
    real(8), intent(in) :: delx_in, delt_in, c_in
    call shim_set_grid(ndim_in, np_in, nsp_in, nxgs_in, nxge_in, nygs_in, nyge_in, nys_in, nye_in)
    prm%bc = WPIC_BC_SHOCK
    have_bc = .true.
  end subroutine boundary_shock__init

  subroutine boundary_shock__particle_x(up, np2, nxs, nxe)
    integer, intent(in)    :: nxs, nxe
    integer, intent(in)    :: np2(s_nys:s_nye, s_nsp)
    real(8), intent(inout) :: up(s_ndim, s_np, s_nys:s_nye, s_nsp)
    integer(c_int) :: n2(s_nys:s_nye, s_nsp)
    call shim_ctx(np2)
    call shim_edge(nxe)
    if (.not. resident) call shim_check(wpic_upload_gp2d(ctx, up, np2, s_np))
    call shim_check(wpic_bc_particle_x2d(ctx))
    if (.not. resident) call shim_check(wpic_download_gp2d(ctx, up, n2, s_np))
  end subroutine boundary_shock__particle_x

  ! the injection edge moving with the inflow (boundary_shock.f90:255-297),
  ! with nxs = nxgs
  subroutine boundary_shock__injection(up, np2, nxs, nxe, u0)
    integer, intent(in)    :: nxs, nxe
    integer, intent(in)    :: np2(s_nys:s_nye, s_nsp)
    real(8), intent(inout) :: up(s_ndim, s_np, s_nys:s_nye, s_nsp)
    real(8), intent(in)    :: u0
    integer(c_int) :: n2(s_nys:s_nye, s_nsp)
    call shim_ctx(np2)
    call shim_edge(nxe)
    if (.not. resident) call shim_check(wpic_upload_gp2d(ctx, up, np2, s_np))
    call shim_check(wpic_bc_injection2d(ctx, u0))
    if (.not. resident) call shim_check(wpic_download_gp2d(ctx, up, n2, s_np))
  end subroutine boundary_shock__injection

  subroutine boundary_shock__particle_y(up, np2)
    integer, intent(inout) :: np2(s_nys:s_nye, s_nsp)
    real(8), intent(inout) :: up(s_ndim, s_np, s_nys:s_nye, s_nsp)
    integer(c_int) :: n2(s_nys:s_nye, s_nsp)
    call shim_ctx(np2)
    if (.not. resident) call shim_check(wpic_upload_gp2d(ctx, up, np2, s_np))
    call shim_check(wpic_bc_particle_y2d(ctx))
    if (.not. resident) then
       call shim_check(wpic_download_gp2d(ctx, up, n2, s_np))
       np2 = n2
    end if
  end subroutine boundary_shock__particle_y

  subroutine boundary_shock__dfield(df, nxs, nxe, nys, nye, nxgs, nxge)
    integer, intent(in)    :: nxs, nxe, nys, nye, nxgs, nxge
    real(8), intent(inout) :: df(6, nxgs-2:nxge+2, nys-2:nye+2)
    write(6,*) 'boundary_shock__dfield: applied on the device inside field__fdtd_i'
    stop
  end subroutine boundary_shock__dfield

  subroutine boundary_shock__curre(uj, nxs, nxe, nys, nye, nxgs, nxge)
    integer, intent(in)    :: nxs, nxe, nys, nye, nxgs, nxge
    real(8), intent(inout) :: uj(3, nxgs-2:nxge+2, nys-2:nye+2)
    write(6,*) 'boundary_shock__curre: applied on the device inside field__fdtd_i'
    stop
  end subroutine boundary_shock__curre

  subroutine boundary_shock__phi(phi, nxs, nxe, nys, nye, l)
    integer, intent(in)    :: nxs, nxe, nys, nye, l
    real(8), intent(inout) :: phi(nxs-1:nxe+1, nys-1:nye+1)
    write(6,*) 'boundary_shock__phi: applied on the device inside field__fdtd_i'
    stop
  end subroutine boundary_shock__phi

  ! the moment folds on the device's moments; mirror mode sends the host
  ! array up and brings the folded one back
  subroutine boundary_shock__mom(mom)
    real(8), intent(inout) :: mom(7, s_nxgs-1:s_nxge+1, s_nys-1:s_nye+1, s_nsp)
    if (.not. c_associated(ctx)) then
       write(6,*) 'boundary_shock__mom: no device state (call mom_calc__accl first)'
       stop
    end if
    if (.not. resident) call shim_check(wpic_upload_mom2d(ctx, mom))
    call shim_check(wpic_bc_mom2d(ctx))
    call shim_check(wpic_download_mom2d(ctx, mom))
  end subroutine boundary_shock__mom

end module boundary_shock


! resident-mode controls (no reference counterpart)
module wpic_shim
  use iso_c_binding
  use wpic_capi
  use wpic_shim_state
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

  ! copy the app's sorted particles (with cumcnt) and fields to the device;
  ! nxe: the shock's active edge while its box expands (cumcnt is read up to
  ! nxe+1, as the reference's sort__bucket writes it)
  subroutine wpic_shim_push(up, uf, np2, cumcnt, nxe)
    real(8), intent(in) :: up(s_ndim, s_np, s_nys:s_nye, s_nsp)
    real(8), intent(in) :: uf(6, s_nxgs-2:s_nxge+2, s_nys-2:s_nye+2)
    integer, intent(in) :: np2(s_nys:s_nye, s_nsp)
    integer, intent(in) :: cumcnt(s_nxgs:s_nxge+1, s_nys:s_nye, s_nsp)
    integer, intent(in), optional :: nxe
    integer(c_int) :: cc(s_nxgs:s_nxge+1, s_nys:s_nye, s_nsp)
    call shim_ctx(np2)
    if (present(nxe)) then
       call shim_edge(nxe)
       call shim_cumcnt(cumcnt, nxe, cc)
    else
       cc = cumcnt
    end if
    call shim_check(wpic_upload_up2d(ctx, up, np2, cc, s_np))
    call shim_check(wpic_upload_fields2d(ctx, uf))
  end subroutine wpic_shim_push

  ! arithmetic mode: WPIC_MODE_FAST or WPIC_MODE_PARITY (before first use)
  subroutine wpic_shim_mode(mode)
    integer(c_int), intent(in) :: mode
    s_mode = mode
  end subroutine wpic_shim_mode

  ! copy the device's sorted particles, cell offsets and fields to the host
  subroutine wpic_shim_pull(up, uf, np2, cumcnt)
    real(8), intent(inout) :: up(s_ndim, s_np, s_nys:s_nye, s_nsp)
    real(8), intent(out)   :: uf(6, s_nxgs-2:s_nxge+2, s_nys-2:s_nye+2)
    integer, intent(out)   :: np2(s_nys:s_nye, s_nsp)
    integer, intent(out)   :: cumcnt(s_nxgs:s_nxge+1, s_nys:s_nye, s_nsp)
    call shim_check(wpic_download_up2d(ctx, up, np2, cumcnt, s_np))
    call shim_check(wpic_download_fields2d(ctx, uf))
  end subroutine wpic_shim_pull

  subroutine wpic_shim_finalize()
    integer(c_int) :: rc
    if (c_associated(ctx)) rc = wpic_destroy(ctx)
    ctx = c_null_ptr
  end subroutine wpic_shim_finalize

end module wpic_shim


! the raw + JSON files of 2d/common/paraio.f90 at one rank, written by the
! host C++ of include/wpic_io.h (byte-identical files); host arrays as the
! reference passes them
module paraio
  use iso_c_binding
  use wpic_capi
  implicit none
  private
  public :: paraio__init, paraio__finalize, paraio__output, paraio__input
  public :: paraio__param, paraio__mom, paraio__ptcl, paraio__orb

  type(c_ptr), save :: io = c_null_ptr
  integer, save :: s_ndim, s_np, s_nsp, s_nxgs, s_nxge, s_nys, s_nye

contains

  subroutine io_check(rc)
    integer(c_int), intent(in) :: rc
    if (rc /= 0) then
       write(0,*) wpic_error_text()
       stop
    end if
  end subroutine io_check

  function cstr(s) result(o)
    character(len=*), intent(in) :: s
    character(kind=c_char, len=len_trim(s)+1) :: o
    o = trim(s) // c_null_char
  end function cstr

  subroutine paraio__init(ndim_in, np_in, nsp_in, nxgs_in, nxge_in, nygs_in, nyge_in, nys_in, nye_in, &
       nproc_in, nrank_in, delx_in, delt_in, c_in, q_in, r_in, dir_in)
    integer, intent(in) :: ndim_in, np_in, nsp_in
    integer, intent(in) :: nxgs_in, nxge_in, nygs_in, nyge_in, nys_in, nye_in
    integer, intent(in) :: nproc_in, nrank_in
    real(8), intent(in) :: delx_in, delt_in, c_in, q_in(nsp_in), r_in(nsp_in)
    character(len=*), intent(in) :: dir_in
    s_ndim = ndim_in; s_np = np_in; s_nsp = nsp_in
    s_nxgs = nxgs_in; s_nxge = nxge_in; s_nys = nys_in; s_nye = nye_in
    call io_check(wpic_paraio_init2d(io, ndim_in, np_in, nsp_in, nxgs_in, nxge_in, nygs_in, nyge_in, &
         nys_in, nye_in, nproc_in, nrank_in, delx_in, delt_in, c_in, q_in, r_in, cstr(dir_in)))
  end subroutine paraio__init

  subroutine paraio__finalize()
    integer(c_int) :: rc
    if (c_associated(io)) rc = wpic_paraio_finalize(io)
    io = c_null_ptr
  end subroutine paraio__finalize

  subroutine paraio__output(up, uf, np2, nxs, nxe, it, filename)
    integer, intent(in) :: np2(s_nys:s_nye, s_nsp), nxs, nxe, it
    real(8), intent(in) :: up(s_ndim, s_np, s_nys:s_nye, s_nsp)
    real(8), intent(in) :: uf(6, s_nxgs-2:s_nxge+2, s_nys-2:s_nye+2)
    character(len=*), intent(in) :: filename
    call io_check(wpic_paraio_output(io, up, uf, np2, nxs, nxe, it, cstr(filename)))
  end subroutine paraio__output

  subroutine paraio__input(up, uf, np2, nxs, nxe, it, filename)
    real(8), intent(out) :: up(s_ndim, s_np, s_nys:s_nye, s_nsp)
    real(8), intent(out) :: uf(6, s_nxgs-2:s_nxge+2, s_nys-2:s_nye+2)
    integer, intent(out) :: np2(s_nys:s_nye, s_nsp)
    integer, intent(out) :: nxs, nxe, it
    character(len=*), intent(in) :: filename
    call io_check(wpic_paraio_input(io, up, uf, np2, nxs, nxe, it, cstr(filename)))
  end subroutine paraio__input

  subroutine paraio__param(n0, wpe, wpi, wge, wgi, vti, vte, filename)
    integer, intent(in) :: n0
    real(8), intent(in) :: wpe, wpi, wge, wgi, vti, vte
    character(len=*), intent(in) :: filename
    call io_check(wpic_paraio_param(io, n0, wpe, wpi, wge, wgi, vti, vte, cstr(filename)))
  end subroutine paraio__param

  subroutine paraio__mom(mom, uf, it)
    integer, intent(in) :: it
    real(8), intent(in) :: uf(6, s_nxgs-2:s_nxge+2, s_nys-2:s_nye+2)
    real(8), intent(inout) :: mom(7, s_nxgs-1:s_nxge+1, s_nys-1:s_nye+1, s_nsp)
    call io_check(wpic_paraio_mom(io, mom, uf, it))
  end subroutine paraio__mom

  subroutine paraio__ptcl(up, uf, np2, it)
    integer, intent(in) :: it
    integer, intent(in) :: np2(s_nys:s_nye, s_nsp)
    real(8), intent(in) :: up(s_ndim, s_np, s_nys:s_nye, s_nsp)
    real(8), intent(in) :: uf(6, s_nxgs-2:s_nxge+2, s_nys-2:s_nye+2)
    call io_check(wpic_paraio_ptcl(io, up, uf, np2, it))
  end subroutine paraio__ptcl

  subroutine paraio__orb(up, uf, np2, it)
    integer, intent(in) :: it
    integer, intent(in) :: np2(s_nys:s_nye, s_nsp)
    real(8), intent(in) :: up(s_ndim, s_np, s_nys:s_nye, s_nsp)
    real(8), intent(in) :: uf(6, s_nxgs-2:s_nxge+2, s_nys-2:s_nye+2)
    call io_check(wpic_paraio_orb(io, up, uf, np2, it))
  end subroutine paraio__orb

end module paraio
