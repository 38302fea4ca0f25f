! shim_step3d_driver.f90 -- the reference 3D app's per-step call sequence
! (3d/proj/weibel/app.f90:103-108) and its moment output sequence
! (app.f90:122-124) written against the 3D shim modules, reading and writing
! the app's arrays as raw files.
!
! Usage: shim_step3d_driver <case-dir>
!   params.txt  nsteps mode resident ndim np nsp nxgs nxge nygs nyge nzgs nzge
!               delx delt c gfac
!               q(1:nsp)
!               r(1:nsp)
!   in:  up.bin uf.bin cumcnt.bin np2.bin
!   out: up.out uf.out cumcnt.out np2.out mom.out
program shim_step3d_driver
  use iso_c_binding, only: c_int
  use wpic_capi3d, only: WPIC_MODE_FAST, WPIC_MODE_PARITY
  use wpic_shim3d
  use particle
  use field
  use sort
  use mom_calc
  use boundary_periodic, only: bc__init => boundary_periodic__init, bc__particle_x => boundary_periodic__particle_x, &
       bc__particle_yz => boundary_periodic__particle_yz, bc__dfield => boundary_periodic__dfield, &
       bc__curre => boundary_periodic__curre, bc__phi => boundary_periodic__phi, bc__mom => boundary_periodic__mom
  implicit none
  character(len=512) :: dir
  integer :: nsteps, mode, res, ndim, np, nsp, nxgs, nxge, nygs, nyge, nzgs, nzge
  integer :: nys, nye, nzs, nze, nxs, nxe, u, it
  integer :: nstat(1)
  real(8) :: delx, delt, c, gfac
  real(8), allocatable :: q(:), r(:), up(:,:,:,:,:), gp(:,:,:,:,:), uf(:,:,:,:), mom(:,:,:,:,:)
  integer, allocatable :: cumcnt(:,:,:,:), np2(:,:,:)

  call get_command_argument(1, dir)
  open(newunit=u, file=trim(dir)//'/params.txt', status='old', action='read')
  read(u,*) nsteps, mode, res, ndim, np, nsp, nxgs, nxge, nygs, nyge, nzgs, nzge
  allocate(q(nsp), r(nsp))
  read(u,*) delx, delt, c, gfac
  read(u,*) q
  read(u,*) r
  close(u)
  nys = nygs
  nye = nyge
  nzs = nzgs
  nze = nzge
  nxs = nxgs
  nxe = nxge
  nstat = 0

  allocate(up(ndim,np,nys:nye,nzs:nze,nsp), gp(ndim,np,nys:nye,nzs:nze,nsp))
  allocate(uf(6,nxgs-2:nxge+2,nys-2:nye+2,nzs-2:nze+2))
  allocate(cumcnt(nxgs:nxge+1,nys:nye,nzs:nze,nsp), np2(nys:nye,nzs:nze,nsp))
  allocate(mom(7,nxgs-1:nxge+1,nys-1:nye+1,nzs-1:nze+1,nsp))
  gp = 0.0d0
  call read_r8(trim(dir)//'/up.bin', up, size(up))
  call read_r8(trim(dir)//'/uf.bin', uf, size(uf))
  call read_i4(trim(dir)//'/cumcnt.bin', cumcnt, size(cumcnt))
  call read_i4(trim(dir)//'/np2.bin', np2, size(np2))

  if (mode == 1) call wpic_shim_mode(WPIC_MODE_PARITY)
  ! the init order of 3d/proj/weibel/app.f90:341-357
  call bc__init(ndim, np, nsp, nxgs, nxge, nygs, nyge, nzgs, nzge, nys, nye, nzs, nze, 0, 0, 0, 0, 0, 0, 0, 0, &
                nstat, delx, delt, c)
  call particle__init(ndim, np, nsp, nxgs, nxge, nygs, nyge, nzgs, nzge, nys, nye, nzs, nze, delx, delt, c, q, r)
  call field__init(ndim, np, nsp, nxgs, nxge, nygs, nyge, nzgs, nzge, nys, nye, nzs, nze, 0, 0, 0, 0, &
                   delx, delt, c, q, r, gfac)
  call sort__init(ndim, np, nsp, nxgs, nxge, nygs, nyge, nzgs, nzge, nys, nye, nzs, nze)
  call mom_calc__init(ndim, np, nsp, nxgs, nxge, nygs, nyge, nzgs, nzge, nys, nye, nzs, nze, delx, delt, c, q, r)
  if (res == 1) then
     call wpic_shim_resident(.true.)
     call wpic_shim_push(up, uf, np2, cumcnt)
  end if

  do it = 1, nsteps
     call particle__solv(gp, up, uf, cumcnt, nxs, nxe)
     call field__fdtd_i(uf, up, gp, cumcnt, nxs, nxe, bc__dfield, bc__curre, bc__phi)
     call bc__particle_x(gp, np2)
     call bc__particle_yz(gp, np2)
     call sort__bucket(up, gp, cumcnt, np2, nxs, nxe)
  end do
  call mom_calc__accl(gp, up, uf, cumcnt, nxs, nxe)
  call mom_calc__nvt(mom, gp, np2)
  call bc__mom(mom)
  if (res == 1) call wpic_shim_pull(up, uf, np2, cumcnt)
  call wpic_shim_finalize()

  call write_r8(trim(dir)//'/up.out', up, size(up))
  call write_r8(trim(dir)//'/uf.out', uf, size(uf))
  call write_i4(trim(dir)//'/cumcnt.out', cumcnt, size(cumcnt))
  call write_i4(trim(dir)//'/np2.out', np2, size(np2))
  call write_r8(trim(dir)//'/mom.out', mom, size(mom))

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

end program shim_step3d_driver
