! shim_step2d_driver.f90 -- the reference app's per-step call sequence
! (2d/proj/weibel/app.f90:100-107 for periodic boxes,
! 2d/proj/reconnection/app.f90:100-106 for x walls, 2d/proj/shock/app.f90:
! 111-117 for the shock's box with its active edge nxe) written against the
! shim modules, reading and writing the app's arrays as raw files.
!
! Usage: shim_step2d_driver <case-dir>
!   params.txt  nsteps bc mode resident ndim np nsp nxgs nxge nygs nyge
!               delx delt c gfac
!               q(1:nsp)
!               r(1:nsp)
!               nxe u0          (bc = 2 only)
!   in:  up.bin uf.bin cumcnt.bin np2.bin     out: up.out uf.out cumcnt.out np2.out
!   and, after the loop, the app's moment output sequence
!   (2d/proj/weibel/app.f90:121-124): mom.out, then io__mom and save_restart's
!   io__output through the paraio module: <nsteps>_mom.{raw,json},
!   restart.{raw,json}
program shim_step2d_driver
  use iso_c_binding, only: c_int
  use wpic_capi, only: WPIC_MODE_FAST, WPIC_MODE_PARITY
  use wpic_shim
  use particle
  use field
  use sort
  use mom_calc
  use paraio
  use boundary_periodic, only: bcp_init => boundary_periodic__init, bcp_x => boundary_periodic__particle_x, &
       bcp_y => boundary_periodic__particle_y, bcp_dfield => boundary_periodic__dfield, &
       bcp_curre => boundary_periodic__curre, bcp_phi => boundary_periodic__phi, bcp_mom => boundary_periodic__mom
  use boundary_reconnection, only: bcr_init => boundary_reconnection__init, &
       bcr_x => boundary_reconnection__particle_x, bcr_y => boundary_reconnection__particle_y, &
       bcr_dfield => boundary_reconnection__dfield, bcr_curre => boundary_reconnection__curre, &
       bcr_phi => boundary_reconnection__phi, bcr_mom => boundary_reconnection__mom
  use boundary_shock, only: bcs_init => boundary_shock__init, bcs_inj => boundary_shock__injection, &
       bcs_y => boundary_shock__particle_y, bcs_dfield => boundary_shock__dfield, &
       bcs_curre => boundary_shock__curre, bcs_phi => boundary_shock__phi, bcs_mom => boundary_shock__mom
  implicit none
  character(len=512) :: dir
  integer :: nsteps, bc, mode, res, ndim, np, nsp, nxgs, nxge, nygs, nyge, nys, nye, nxs, nxe, u, it
  integer :: nstat(1)
  real(8) :: delx, delt, c, gfac, u0
  real(8), allocatable :: q(:), r(:), up(:,:,:,:), gp(:,:,:,:), uf(:,:,:), mom(:,:,:,:)
  integer, allocatable :: cumcnt(:,:,:), np2(:,:)

  call get_command_argument(1, dir)
  open(newunit=u, file=trim(dir)//'/params.txt', status='old', action='read')
  read(u,*) nsteps, bc, mode, res, ndim, np, nsp, nxgs, nxge, nygs, nyge
  allocate(q(nsp), r(nsp))
  read(u,*) delx, delt, c, gfac
  read(u,*) q
  read(u,*) r
  nys = nygs
  nye = nyge
  nxs = nxgs
  nxe = nxge
  u0 = 0.0d0
  if (bc == 2) read(u,*) nxe, u0
  close(u)
  nstat = 0

  allocate(up(ndim,np,nys:nye,nsp), gp(ndim,np,nys:nye,nsp))
  allocate(uf(6,nxgs-2:nxge+2,nys-2:nye+2))
  allocate(cumcnt(nxgs:nxge+1,nys:nye,nsp), np2(nys:nye,nsp))
  allocate(mom(7,nxgs-1:nxge+1,nys-1:nye+1,nsp))
  gp = 0.0d0
  call read_r8(trim(dir)//'/up.bin', up, size(up))
  call read_r8(trim(dir)//'/uf.bin', uf, size(uf))
  call read_i4(trim(dir)//'/cumcnt.bin', cumcnt, size(cumcnt))
  call read_i4(trim(dir)//'/np2.bin', np2, size(np2))

  if (mode == 1) call wpic_shim_mode(WPIC_MODE_PARITY)
  if (bc == 0) then
     call bcp_init(ndim, np, nsp, nxgs, nxge, nygs, nyge, nys, nye, 0, 0, 0, 0, 0, 0, nstat, delx, delt, c)
  else if (bc == 1) then
     call bcr_init(ndim, np, nsp, nxgs, nxge, nygs, nyge, nys, nye, 0, 0, 0, 0, 0, 0, nstat, delx, delt, c)
  else
     call bcs_init(ndim, np, nsp, nxgs, nxge, nygs, nyge, nys, nye, 0, 0, 0, 0, 0, 0, nstat, delx, delt, c)
  end if
  call particle__init(ndim, np, nsp, nxgs, nxge, nygs, nyge, nys, nye, delx, delt, c, q, r)
  call field__init(ndim, np, nsp, nxgs, nxge, nygs, nyge, nys, nye, 0, 0, 0, 0, delx, delt, c, q, r, gfac)
  call sort__init(ndim, np, nsp, nxgs, nxge, nygs, nyge, nys, nye)
  call mom_calc__init(ndim, np, nsp, nxgs, nxge, nygs, nyge, nys, nye, delx, delt, c, q, r)
  if (res == 1) then
     call wpic_shim_resident(.true.)
     call wpic_shim_push(up, uf, np2, cumcnt, nxe)
  end if

  do it = 1, nsteps
     call particle__solv(gp, up, uf, cumcnt, nxs, nxe)
     if (bc == 0) then
        call field__fdtd_i(uf, up, gp, cumcnt, nxs, nxe, bcp_dfield, bcp_curre, bcp_phi)
        call bcp_x(gp, np2)
        call bcp_y(gp, np2)
     else if (bc == 1) then
        call bcr_x(gp, np2, nxs, nxe)
        call field__fdtd_i(uf, up, gp, cumcnt, nxs, nxe, bcr_dfield, bcr_curre, bcr_phi)
        call bcr_y(gp, np2)
     else
        call bcs_inj(gp, np2, nxs, nxe, u0)
        call field__fdtd_i(uf, up, gp, cumcnt, nxs, nxe, bcs_dfield, bcs_curre, bcs_phi)
        call bcs_y(gp, np2)
     end if
     call sort__bucket(up, gp, cumcnt, np2, nxs, nxe)
  end do
  call mom_calc__accl(gp, up, uf, cumcnt, nxs, nxe)
  call mom_calc__nvt(mom, gp, np2)
  if (bc == 0) then
     call bcp_mom(mom)
  else if (bc == 1) then
     call bcr_mom(mom)
  else
     call bcs_mom(mom)
  end if
  if (res == 1) call wpic_shim_pull(up, uf, np2, cumcnt)
  call wpic_shim_finalize()

  call write_r8(trim(dir)//'/up.out', up, size(up))
  call write_r8(trim(dir)//'/uf.out', uf, size(uf))
  call write_i4(trim(dir)//'/cumcnt.out', cumcnt, size(cumcnt))
  call write_i4(trim(dir)//'/np2.out', np2, size(np2))
  call write_r8(trim(dir)//'/mom.out', mom, size(mom))

  call paraio__init(ndim, np, nsp, nxgs, nxge, nygs, nyge, nys, nye, 1, 0, delx, delt, c, q, r, trim(dir)//'/')
  call paraio__mom(mom, uf, nsteps)
  call paraio__output(up, uf, np2, nxs, nxe, nsteps, 'restart')
  call paraio__finalize()

contains

  subroutine read_r8(fn, a, n)
    character(len=*), intent(in) :: fn
    integer, intent(in) :: n
    real(8), intent(out) :: a(n)
    integer :: v
    open(newunit=v, file=fn, access='stream', form='unformatted', status='old', action='read')
    read(v) a
    close(v)
  end subroutine read_r8

  subroutine read_i4(fn, a, n)
    character(len=*), intent(in) :: fn
    integer, intent(in) :: n
    integer, intent(out) :: a(n)
    integer :: v
    open(newunit=v, file=fn, access='stream', form='unformatted', status='old', action='read')
    read(v) a
    close(v)
  end subroutine read_i4

  subroutine write_r8(fn, a, n)
    character(len=*), intent(in) :: fn
    integer, intent(in) :: n
    real(8), intent(in) :: a(n)
    integer :: v
    open(newunit=v, file=fn, access='stream', form='unformatted', status='replace', action='write')
    write(v) a
    close(v)
  end subroutine write_r8

  subroutine write_i4(fn, a, n)
    character(len=*), intent(in) :: fn
    integer, intent(in) :: n
    integer, intent(in) :: a(n)
    integer :: v
    open(newunit=v, file=fn, access='stream', form='unformatted', status='replace', action='write')
    write(v) a
    close(v)
  end subroutine write_i4

end program shim_step2d_driver
