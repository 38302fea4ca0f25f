! ref_paraio2d.f90 -- the single-rank file pairs of the reference's 2D paraio
! routines, written through the reference's OWN jsonio module and
! JSON-Fortran library (utils/iocore/jsonio.f90, utils/json, compiled from
! the reference sources by oracle/Makefile into oracle/_ref/).
! TEST INFRASTRUCTURE ONLY.
!
! paraio.f90 itself uses MPI and MPI-IO, which the image lacks; this driver
! restates each routine's call sequence (2d/common/paraio.f90:93-228,
! 431-550, 555-713, 746-938, 956-1002) with the MPI-IO writes at one rank
! replaced by plain stream writes at the same offsets (rank 0's atomic writes,
! mpiio.f90:272-294, and the collective writes of whole arrays, 440-663).
!
! Usage: ref_paraio2d <case-dir> <routine>     routine: output|param|mom|ptcl|orb
!   <case-dir>/params.txt   ndim np nsp nxgs nxge nygs nyge it nxs nxe n0
!                           delx delt c wpe wpi wge wgi vti vte
!                           q(1:nsp)
!                           r(1:nsp)
!   in: up.bin uf.bin np2.bin mom.bin (raw, the reference's layouts)
!   out: the file pair of the routine, named as paraio names it, in <case-dir>/
program ref_paraio2d
  use iso_fortran_env, only: int64
  use jsonio
  implicit none
  integer, parameter :: MOK = 8
  character(len=512) :: dir
  character(len=16) :: routine
  integer :: ndim, np, nsp, nxgs, nxge, nygs, nyge, nys, nye, it, nxs, nxe, n0, u, nproc
  real(8) :: delx, delt, c, wpe, wpi, wge, wgi, vti, vte
  real(8), allocatable :: q(:), r(:), up(:,:,:,:), uf(:,:,:), mom(:,:,:,:)
  integer, allocatable :: np2(:,:)
  type(json_core) :: json
  type(json_value), pointer :: root, p
  integer(int64) :: disp
  integer :: fu

  call get_command_argument(1, dir)
  call get_command_argument(2, routine)
  dir = trim(dir) // '/'
  open(newunit=u, file=trim(dir)//'params.txt', status='old', action='read')
  read(u,*) ndim, np, nsp, nxgs, nxge, nygs, nyge, it, nxs, nxe, n0
  allocate(q(nsp), r(nsp))
  read(u,*) delx, delt, c, wpe, wpi, wge, wgi, vti, vte
  read(u,*) q
  read(u,*) r
  close(u)
  nys = nygs
  nye = nyge
  nproc = 1
  allocate(up(ndim,np,nys:nye,nsp), uf(6,nxgs-2:nxge+2,nys-2:nye+2))
  allocate(mom(7,nxgs-1:nxge+1,nys-1:nye+1,nsp), np2(nys:nye,nsp))
  call read_r8(trim(dir)//'up.bin', up, size(up))
  call read_r8(trim(dir)//'uf.bin', uf, size(uf))
  call read_r8(trim(dir)//'mom.bin', mom, size(mom))
  call read_i4(trim(dir)//'np2.bin', np2, size(np2))

  select case (trim(routine))
  case ('output')
     call do_output('restart')
  case ('param')
     call do_param('param')
  case ('mom')
     call do_mom()
  case ('ptcl')
     call do_particles(0, '_ptcl')
  case ('orb')
     call do_particles(1, '_orb')
  case default
     write(0,*) 'unknown routine ', trim(routine)
     stop 1
  end select

contains

  subroutine begin_file(name)
    character(len=*), intent(in) :: name
    integer :: endian
    integer(1) :: b(4) = (/1_1, 0_1, 0_1, 0_1/)
    open(newunit=fu, file=trim(dir)//trim(name)//'.raw', access='stream', form='unformatted', &
         status='replace', action='write')
    disp = 0
    call json%initialize()
    call json%create_object(root, 'root')
    endian = transfer(b, endian)
    call json%create_object(p, 'meta')
    call json%add(root, p)
    call json%add(p, 'endian', endian)
    call json%add(p, 'rawfile', trim(name)//'.raw')
    call json%create_object(p, 'attribute')
    call json%add(root, p)
  end subroutine begin_file

  subroutine end_file(name)
    character(len=*), intent(in) :: name
    call json%print(root, trim(dir)//trim(name)//'.json')
    call json%destroy()
    close(fu)
  end subroutine end_file

  subroutine att_i4(v, name)
    integer, intent(in) :: v
    character(len=*), intent(in) :: name
    call jsonio_put_attribute(json, p, v, name, disp, '')
    write(fu, pos=disp+1) v
    disp = disp + 4
  end subroutine att_i4

  subroutine att_r8(v, name)
    real(8), intent(in) :: v
    character(len=*), intent(in) :: name
    call jsonio_put_attribute(json, p, v, name, disp, '')
    write(fu, pos=disp+1) v
    disp = disp + 8
  end subroutine att_r8

  subroutine att_r8v(v, name)
    real(8), intent(in) :: v(:)
    character(len=*), intent(in) :: name
    call jsonio_put_attribute(json, p, v, name, disp, '')
    write(fu, pos=disp+1) v
    disp = disp + 8*size(v)
  end subroutine att_r8v

  ! paraio.f90:956-1002
  subroutine put_metadata()
    call att_i4(ndim, 'ndim')
    call att_i4(np, 'np')
    call att_i4(nxgs, 'nxgs')
    call att_i4(nxge, 'nxge')
    call att_i4(nygs, 'nygs')
    call att_i4(nyge, 'nyge')
    call att_i4(nsp, 'nsp')
    call att_i4(nproc, 'nproc')
    call att_r8(delx, 'delx')
    call att_r8(delt, 'delt')
    call att_r8(c, 'c')
    call att_r8v(r, 'r')
    call att_r8v(q, 'q')
  end subroutine put_metadata

  subroutine dset_r8(a, name, nd, gshape, desc)
    real(8), intent(in) :: a(:)
    character(len=*), intent(in) :: name, desc
    integer, intent(in) :: nd, gshape(:)
    integer(int64) :: dsize
    dsize = 8_int64 * size(a, kind=int64)
    call jsonio_put_metadata(json, p, name, 'f8', disp, dsize, nd, gshape(1:nd), desc)
    write(fu, pos=disp+1) a
    disp = disp + dsize
  end subroutine dset_r8

  subroutine dset_i4(a, name, nd, gshape, desc)
    integer, intent(in) :: a(:)
    character(len=*), intent(in) :: name, desc
    integer, intent(in) :: nd, gshape(:)
    integer(int64) :: dsize
    dsize = 4_int64 * size(a, kind=int64)
    call jsonio_put_metadata(json, p, name, 'i4', disp, dsize, nd, gshape(1:nd), desc)
    write(fu, pos=disp+1) a
    disp = disp + dsize
  end subroutine dset_i4

  ! paraio__output, paraio.f90:93-228
  subroutine do_output(name)
    character(len=*), intent(in) :: name
    integer :: nyl
    call begin_file(name)
    call att_i4(MOK, 'dummy_attribute')
    call att_i4(it, 'it')
    call att_i4(nxs, 'nxs')
    call att_i4(nxe, 'nxe')
    call put_metadata()
    call json%create_object(p, 'dataset')
    call json%add(root, p)
    nyl = nye - nys + 1
    call dset_r8(reshape(up, (/size(up)/)), 'up', 5, (/ndim, np, nyl, nsp, nproc/), 'particles')
    call dset_i4(reshape(np2, (/size(np2)/)), 'np2', 3, (/nyl, nsp, nproc/), 'number of active particles')
    call dset_r8(reshape(uf, (/size(uf)/)), 'uf', 4, (/6, size(uf,2), size(uf,3), nproc/), &
         'electromagnetic fields including ghost cells')
    call end_file(name)
  end subroutine do_output

  ! paraio__param, paraio.f90:431-550
  subroutine do_param(name)
    character(len=*), intent(in) :: name
    real(8) :: vai, vae
    vai = c * wgi/wpi
    vae = c * wge/wpe
    call begin_file(name)
    call att_i4(MOK, 'dummy_attribute')
    call att_i4(nxge - nxgs + 1, 'nx')
    call att_i4(nyge - nygs + 1, 'ny')
    call att_i4(np, 'np')
    call att_r8(delx, 'delx')
    call att_r8(delt, 'delt')
    call att_r8(c, 'c')
    call att_r8v(r, 'r')
    call att_r8v(q, 'q')
    call att_r8(wpe, 'wpe')
    call att_r8(wge, 'wge')
    call att_r8(wpi, 'wpi')
    call att_r8(wgi, 'wgi')
    call att_r8(vti, 'vti')
    call att_r8(vte, 'vte')
    call att_r8(vai, 'vai')
    call att_r8(vae, 'vae')
    call att_i4(n0, 'n0')
    call end_file(name)
  end subroutine do_param

  ! paraio__mom, paraio.f90:555-713
  subroutine do_mom()
    character(len=64) :: name
    integer :: i, j, nxg, nyg
    real(8) :: tmp(1:6,nxgs:nxge,nys:nye)
    write(name,'(i7.7, "_mom")') it
    call begin_file(name)
    call att_i4(MOK, 'dummy_attribute')
    call att_i4(it, 'it')
    call put_metadata()
    call json%create_object(p, 'dataset')
    call json%add(root, p)
    nxg = nxge - nxgs + 1
    nyg = nyge - nygs + 1
    call dset_r8(reshape(mom(1:1,nxgs:nxge,nys:nye,1:nsp), (/nxg*nyg*nsp/)), 'den', 3, (/nxg, nyg, nsp/), 'density')
    mom(2,:,:,:) = mom(2,:,:,:) / mom(1,:,:,:)
    mom(3,:,:,:) = mom(3,:,:,:) / mom(1,:,:,:)
    mom(4,:,:,:) = mom(4,:,:,:) / mom(1,:,:,:)
    call dset_r8(reshape(mom(2:4,nxgs:nxge,nys:nye,1:nsp), (/3*nxg*nyg*nsp/)), 'vel', 4, (/3, nxg, nyg, nsp/), &
         'velocity')
    mom(5,:,:,:) = mom(5,:,:,:) / mom(1,:,:,:)
    mom(6,:,:,:) = mom(6,:,:,:) / mom(1,:,:,:)
    mom(7,:,:,:) = mom(7,:,:,:) / mom(1,:,:,:)
    call dset_r8(reshape(mom(5:7,nxgs:nxge,nys:nye,1:nsp), (/3*nxg*nyg*nsp/)), 'temp', 4, (/3, nxg, nyg, nsp/), &
         'temperature')
    do j=nys,nye
       do i=nxgs,nxge
          tmp(1,i,j) = (uf(1,i,j)+uf(1,i,j+1)) / 2
          tmp(2,i,j) = (uf(2,i,j)+uf(2,i+1,j)) / 2
          tmp(3,i,j) = (uf(3,i,j)+uf(3,i+1,j)+uf(3,i+1,j)+uf(3,i+1,j+1)) / 4
          tmp(4,i,j) = (uf(4,i,j)+uf(4,i+1,j)) / 2
          tmp(5,i,j) = (uf(5,i,j)+uf(5,i,j+1)) / 2
          tmp(6,i,j) = uf(6,i,j)
       enddo
    enddo
    call dset_r8(reshape(tmp, (/6*nxg*nyg/)), 'uf', 3, (/6, nxg, nyg/), 'electromagnetic field')
    call end_file(name)
    ! the divided moments, for the caller to compare with its in-place result
    open(newunit=u, file=trim(dir)//'mom_divided.bin', access='stream', form='unformatted', status='replace')
    write(u) mom
    close(u)
  end subroutine do_mom

  ! write_particle, paraio.f90:746-893, with get_particle_count :897-938
  subroutine do_particles(mode, suffix)
    integer, intent(in) :: mode
    character(len=*), intent(in) :: suffix
    character(len=64) :: name, nm, desc
    integer :: isp, i, j, k, nxg, nyg
    integer(int64) :: cnt, pid, dsize
    integer(int64) :: gshape8(2)
    real(8), allocatable :: buf(:)
    write(name,'(i7.7, a)') it, trim(suffix)
    call begin_file(name)
    call att_i4(MOK, 'dummy_attribute')
    call att_i4(it, 'it')
    call put_metadata()
    call json%create_object(p, 'dataset')
    call json%add(root, p)
    nxg = nxge - nxgs + 1
    nyg = nyge - nygs + 1
    call dset_r8(reshape(uf(1:6,nxgs:nxge,nys:nye), (/6*nxg*nyg/)), 'uf', 3, (/6, nxg, nyg/), &
         'electromagnetic field')
    allocate(buf(ndim*np*(nye-nys+1)))
    do isp = 1, nsp
       cnt = 0
       do j = nys, nye
          do i = 1, np2(j,isp)
             pid = transfer(up(ndim,i,j,isp), 1_8)
             if (mode == 1 .and. pid <= 0) cycle
             do k = 1, ndim
                buf(cnt*ndim + k) = up(k,i,j,isp)
             end do
             cnt = cnt + 1
          end do
       end do
       write(desc, '("particle species #", i2.2)') isp
       write(nm, '("up", i2.2)') isp
       gshape8(1) = ndim
       gshape8(2) = cnt
       dsize = 8 * ndim * cnt
       call jsonio_put_metadata(json, p, trim(nm), 'f8', disp, dsize, 2_8, gshape8, trim(desc))
       if (cnt > 0) write(fu, pos=disp+1) buf(1:ndim*cnt)
       disp = disp + dsize
    end do
    call end_file(name)
  end subroutine do_particles

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

end program ref_paraio2d
