! wpic_shim_state.f90 -- the one device context shared by the shim modules
! particle, field, sort and boundary_* (the reference keeps its state in
! module variables set by each *__init; here the inits fill one parameter
! block and the context is created on first use).
!
! Transfer policy:
!   mirror   (default) every entry point uploads its intent(in) arrays and
!            downloads its intent(out) arrays, so the host arrays hold exactly
!            what the reference's would after each call;
!   resident (call wpic_shim_resident(.true.)) no entry point moves host
!            data: load the state once with wpic_shim_push, and call
!            wpic_shim_pull before any host reader (io__*, energy_history,
!            save_restart).
module wpic_shim_state
  use iso_c_binding
  use wpic_capi
  implicit none

  type(c_ptr), save :: ctx = c_null_ptr
  type(wpic_params2d), save :: prm
  logical, save :: have_particle = .false., have_field = .false., have_bc = .false., have_mom = .false.
  logical, save :: resident = .false.
  integer, save :: s_ndim, s_np, s_nsp, s_nxgs, s_nxge, s_nygs, s_nyge, s_nys, s_nye
  integer(c_int), save :: s_mode = WPIC_MODE_FAST

contains

  subroutine shim_check(rc)
    integer(c_int), intent(in) :: rc
    if (rc /= 0) then
       write(6,*) wpic_error_text()
       stop
    end if
  end subroutine shim_check

  subroutine shim_set_grid(ndim_in, np_in, nsp_in, nxgs_in, nxge_in, nygs_in, nyge_in, nys_in, nye_in)
    integer, intent(in) :: ndim_in, np_in, nsp_in, nxgs_in, nxge_in, nygs_in, nyge_in, nys_in, nye_in
    s_ndim = ndim_in
    s_np = np_in
    s_nsp = nsp_in
    s_nxgs = nxgs_in
    s_nxge = nxge_in
    s_nygs = nygs_in
    s_nyge = nyge_in
    s_nys = nys_in
    s_nye = nye_in
    if (nys_in /= nygs_in .or. nye_in /= nyge_in) then
       write(6,*) 'wpic shim: single rank only (nys = nygs, nye = nyge)'
       stop
    end if
    prm%ndim = ndim_in
    prm%nsp = nsp_in
    prm%nxgs = nxgs_in
    prm%nxge = nxge_in
    prm%nygs = nygs_in
    prm%nyge = nyge_in
  end subroutine shim_set_grid

  ! the active right edge nxe an entry point is passed (the shock's box grows
  ! from nxgs + n_x_ini); the context takes it before the call
  subroutine shim_edge(nxe)
    integer, intent(in) :: nxe
    call shim_check(wpic_set_nxe2d(ctx, int(nxe, c_int)))
  end subroutine shim_edge

  ! cumcnt as the device takes it: the entries beyond nxe+1, which the
  ! reference's sort__bucket leaves as they were, carry the row's total
  subroutine shim_cumcnt(cumcnt, nxe, cc)
    integer, intent(in) :: cumcnt(s_nxgs:s_nxge+1, s_nys:s_nye, s_nsp), nxe
    integer(c_int), intent(out) :: cc(s_nxgs:s_nxge+1, s_nys:s_nye, s_nsp)
    integer :: i
    cc = cumcnt
    do i = nxe + 2, s_nxge + 1
       cc(i, :, :) = cumcnt(nxe + 1, :, :)
    end do
  end subroutine shim_cumcnt

  ! create the context on first use; capacity = largest species count seen
  ! in np2 (particle counts are conserved by the periodic and wall boundaries)
  subroutine shim_ctx(np2)
    integer, intent(in) :: np2(s_nys:s_nye, s_nsp)
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
       cap = max(cap, int(sum(np2(:, isp)), c_long_long))
    end do
    prm%capacity = cap
    prm%mode = s_mode
    if (.not. have_bc) prm%bc = WPIC_BC_PERIODIC
    call shim_check(wpic_create2d(ctx, prm))
  end subroutine shim_ctx

end module wpic_shim_state
