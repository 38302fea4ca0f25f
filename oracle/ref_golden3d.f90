! ref_golden3d.f90 -- golden-vector driver for the reference's own 3D
! `particle`, `sort` and `mom_calc` modules (compiled from the reference sources by
! oracle/Makefile into oracle/_ref/).  TEST INFRASTRUCTURE ONLY.
!
! Usage: ref_golden3d <case-dir>
!   <case-dir>/params.txt   mode ndim np nsp nxgs nxge nygs nyge nzgs nzge
!                           nys nye nzs nze / delx delt c / q(1:nsp) / r(1:nsp)
!   mode 1 (push):  in  up.bin uf.bin cumcnt.bin      out gp.bin
!   mode 2 (sort):  in  up.bin np2.bin                out gp.bin cumcnt.bin
!   mode 4 (vay):   in  up.bin uf.bin cumcnt.bin      out gp.bin
!   mode 3 (mom):   in  up.bin uf.bin cumcnt.bin np2.bin
!                   out gp.bin (mom_calc__accl) mom.bin (mom_calc__nvt of that gp)
! Arrays are raw little-endian stream files in the reference's own layout.
program ref_golden3d
  use particle, only: particle__init, particle__solv, particle__solv_vay
  use sort, only: sort__init, sort__bucket
  use mom_calc, only: mom_calc__init, mom_calc__accl, mom_calc__nvt
  implicit none
  character(len=512) :: dir
  integer :: mode, ndim, np, nsp, nxgs, nxge, nygs, nyge, nzgs, nzge, nys, nye, nzs, nze, u
  real(8) :: delx, delt, c
  real(8), allocatable :: q(:), r(:), up(:,:,:,:,:), gp(:,:,:,:,:), uf(:,:,:,:)
  real(8), allocatable :: mom(:,:,:,:,:)
  integer, allocatable :: cumcnt(:,:,:,:), np2(:,:,:)

  call get_command_argument(1, dir)
  open(newunit=u, file=trim(dir)//'/params.txt', status='old', action='read')
  read(u,*) mode, ndim, np, nsp, nxgs, nxge, nygs, nyge, nzgs, nzge, nys, nye, nzs, nze
  allocate(q(nsp), r(nsp))
  read(u,*) delx, delt, c
  read(u,*) q
  read(u,*) r
  close(u)

  allocate(up(ndim,np,nys:nye,nzs:nze,nsp), gp(ndim,np,nys:nye,nzs:nze,nsp))
  allocate(cumcnt(nxgs:nxge+1,nys:nye,nzs:nze,nsp), np2(nys:nye,nzs:nze,nsp))
  gp = 0.0d0
  call read_r8(trim(dir)//'/up.bin', up, size(up))

  if (mode == 1) then
     allocate(uf(6,nxgs-2:nxge+2,nys-2:nye+2,nzs-2:nze+2))
     call read_r8(trim(dir)//'/uf.bin', uf, size(uf))
     call read_i4(trim(dir)//'/cumcnt.bin', cumcnt, size(cumcnt))
     call particle__init(ndim,np,nsp,nxgs,nxge,nygs,nyge,nzgs,nzge,nys,nye,nzs,nze,delx,delt,c,q,r)
     call particle__solv(gp,up,uf,cumcnt,nxgs,nxge)
     call write_r8(trim(dir)//'/gp.bin', gp, size(gp))
  else if (mode == 4) then
     allocate(uf(6,nxgs-2:nxge+2,nys-2:nye+2,nzs-2:nze+2))
     call read_r8(trim(dir)//'/uf.bin', uf, size(uf))
     call read_i4(trim(dir)//'/cumcnt.bin', cumcnt, size(cumcnt))
     call particle__init(ndim,np,nsp,nxgs,nxge,nygs,nyge,nzgs,nzge,nys,nye,nzs,nze,delx,delt,c,q,r)
     call particle__solv_vay(gp,up,uf,cumcnt,nxgs,nxge)
     call write_r8(trim(dir)//'/gp.bin', gp, size(gp))
  else if (mode == 3) then
     allocate(uf(6,nxgs-2:nxge+2,nys-2:nye+2,nzs-2:nze+2))
     allocate(mom(7,nxgs-1:nxge+1,nys-1:nye+1,nzs-1:nze+1,nsp))
     call read_r8(trim(dir)//'/uf.bin', uf, size(uf))
     call read_i4(trim(dir)//'/cumcnt.bin', cumcnt, size(cumcnt))
     call read_i4(trim(dir)//'/np2.bin', np2, size(np2))
     call mom_calc__init(ndim,np,nsp,nxgs,nxge,nygs,nyge,nzgs,nzge,nys,nye,nzs,nze,delx,delt,c,q,r)
     call mom_calc__accl(gp,up,uf,cumcnt,nxgs,nxge)
     call mom_calc__nvt(mom,gp,np2)
     call write_r8(trim(dir)//'/gp.bin', gp, size(gp))
     call write_r8(trim(dir)//'/mom.bin', mom, size(mom))
  else
     call read_i4(trim(dir)//'/np2.bin', np2, size(np2))
     call sort__init(ndim,np,nsp,nxgs,nxge,nygs,nyge,nzgs,nzge,nys,nye,nzs,nze)
     call sort__bucket(gp,up,cumcnt,np2,nxgs,nxge)
     call write_r8(trim(dir)//'/gp.bin', gp, size(gp))
     call write_i4(trim(dir)//'/cumcnt.bin', cumcnt, size(cumcnt))
  end if

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

end program ref_golden3d
