! wpic_shim_state3d.f90 -- the one device context shared by the 3D shim
! modules particle, field, sort, mom_calc and boundary_periodic (the 3d/common
! modules keep their state in module variables set by each *__init; here the
! inits fill one parameter block and the context is created on first use).
!
! Transfer policy, as the 2D shim's:
!   mirror   (default) every entry point uploads its intent(in) arrays and
!            downloads its intent(out) arrays;
!   resident (call wpic_shim_resident(.true.)) no entry point moves host
!            data: load the state once with wpic_shim_push, and call
!            wpic_shim_pull before any host reader.
module wpic_shim_state3d
  use iso_c_binding
  use wpic_capi3d
  implicit none

  type(c_ptr), save :: ctx = c_null_ptr
  type(wpic_params3d), save :: prm
  logical, save :: have_particle = .false., have_field = .false., have_mom = .false.
  logical, save :: resident = .false.
  integer, save :: s_ndim, s_np, s_nsp, s_nxgs, s_nxge, s_nygs, s_nyge, s_nzgs, s_nzge
  integer, save :: s_nys, s_nye, s_nzs, s_nze
  integer(c_int), save :: s_mode = WPIC_MODE_FAST

contains

  subroutine shim_check(rc)
    integer(c_int), intent(in) :: rc
    if (rc /= 0) then
       write(6,*) wpic_error_text3d()
       stop
    end if
  end subroutine shim_check

  subroutine shim_set_grid(ndim_in, np_in, nsp_in, nxgs_in, nxge_in, nygs_in, nyge_in, nzgs_in, nzge_in, &
                           nys_in, nye_in, nzs_in, nze_in)
    integer, intent(in) :: ndim_in, np_in, nsp_in
    integer, intent(in) :: nxgs_in, nxge_in, nygs_in, nyge_in, nzgs_in, nzge_in, nys_in, nye_in, nzs_in, nze_in
    s_ndim = ndim_in
    s_np = np_in
    s_nsp = nsp_in
    s_nxgs = nxgs_in
    s_nxge = nxge_in
    s_nygs = nygs_in
    s_nyge = nyge_in
    s_nzgs = nzgs_in
    s_nzge = nzge_in
    s_nys = nys_in
    s_nye = nye_in
    s_nzs = nzs_in
    s_nze = nze_in
    if (nys_in /= nygs_in .or. nye_in /= nyge_in .or. nzs_in /= nzgs_in .or. nze_in /= nzge_in) then
       write(6,*) 'wpic shim: single rank only (nys = nygs, nye = nyge, nzs = nzgs, nze = nzge)'
       stop
    end if
    prm%ndim = ndim_in
    prm%nsp = nsp_in
    prm%nxgs = nxgs_in
    prm%nxge = nxge_in
    prm%nygs = nygs_in
    prm%nyge = nyge_in
    prm%nzgs = nzgs_in
    prm%nzge = nzge_in
  end subroutine shim_set_grid

  ! create the context on first use; capacity = largest species count seen
  ! in np2 (the periodic boundaries conserve every species' count)
  subroutine shim_ctx(np2)
    integer, intent(in) :: np2(s_nys:s_nye, s_nzs:s_nze, s_nsp)
    integer :: isp
    integer(c_long_long) :: cap
    if (c_associated(ctx)) return
    if (.not. have_particle) then
       write(6,*) 'Initialize first by calling particle__init()'
       stop
    end if
    if (.not. have_field) then
       write(6,*) 'Initialize first by calling field__init()'
       stop
    end if
    cap = 1
    do isp = 1, s_nsp
       cap = max(cap, int(sum(np2(:, :, isp)), c_long_long))
    end do
    prm%capacity = cap
    prm%mode = s_mode
    call shim_check(wpic_create3d(ctx, prm))
  end subroutine shim_ctx

end module wpic_shim_state3d
