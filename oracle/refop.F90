! Fixture dumper for the umbrella order parameters of mc2: drives the
! reference's own q6local2 (Q6Local.F90) and triangle (Triangle.F90)
! modules, built unmodified from /root/reference by oracle/ref.mk, on a given
! pair list and writes what they compute as text.
!
! Test infrastructure only; never part of the product.
!
! stdin (list-directed):
!   nsite npair
!   npair lines: i j ox oy oz          (1-based molecule indices)
!   nsite lines: x y z                 (molecule centres)
!   21 lines: the reference histogram of the triangle order parameter
! stdout:
!   'Q6' / per molecule: q6 nbond / sumq6 op       (sq6system_prepareall)
!   'TRI' / 21 bins / op                            (sTriangle_prepareall)
!   'LOGTRI' / 21 bins / op                         (the logarithmic form)
! then, if stdin holds them, ncase lines: target dx dy dz (1-based target);
! for each case the order parameters after that trial displacement:
!   'DQ6' op   (sq6system_Differentiate, then sq6system_Reject)
!   'DTRI' op  (sTriangle_Differentiate of the linear form, then Reject)
program refop
  use interaction_module
  use vector_module
  use q6local2
  use triangle
  use distrib_module
  implicit none
  integer :: nsite, npair, i, k, ncase, c, target, ios
  real(kind=8) :: op, ddx, ddy, ddz
  type(sInteraction) :: ww
  type(vector3), allocatable :: coord(:)
  type(sq6system), pointer :: q6s
  type(sTriangle), pointer :: tri, tri1
  type(sq6diff) :: q6d
  type(sHistogram), pointer :: refhist
  character(len=*), parameter :: f2 = '(es26.17e3,1x,i6)', f1 = '(es26.17e3)'

  read(*,*) nsite, npair
  call interaction_initialize(ww, nsite, -1)
  ww%Npair = npair
  do k = 1, npair
     read(*,*) ww%pair_i(k), ww%pair_j(k), ww%ox(k), ww%oy(k), ww%oz(k)
  enddo
  allocate(coord(nsite))
  do i = 1, nsite
     read(*,*) coord(i)%vec(1), coord(i)%vec(2), coord(i)%vec(3)
  enddo
  nullify(refhist)
  call new(refhist, 21, 0.1d0, -1d0)
  do i = 1, 21
     read(*,*) refhist%bins(i)
  enddo

  call new(q6s)
  call sq6system_prepareall(q6s, nsite, ww, coord, op)
  write(*,'(a)') 'Q6'
  do i = 1, nsite
     write(*,f2) q6s%q6(i), q6s%nbond(i)
  enddo
  write(*,'(2es26.17e3)') q6s%sumq6, op

  call new(tri, .false.)
  call sTriangle_prepareall(tri, nsite, ww, coord, refhist, op)
  write(*,'(a)') 'TRI'
  do i = 1, 21
     write(*,f1) tri%hist%bins(i)
  enddo
  write(*,f1) op
  tri1 => tri

  call new(tri, .true.)
  call sTriangle_prepareall(tri, nsite, ww, coord, refhist, op)
  write(*,'(a)') 'LOGTRI'
  do i = 1, 21
     write(*,f1) tri%hist%bins(i)
  enddo
  write(*,f1) op

  read(*,*,iostat=ios) ncase
  if (ios /= 0) ncase = 0
  do c = 1, ncase
     read(*,*) target, ddx, ddy, ddz
     call sq6system_Differentiate(q6s, target, ddx, ddy, ddz, nsite, q6d, op)
     write(*,'(a)') 'DQ6'
     write(*,f1) op
     call sq6system_Reject(q6s, q6d, target, nsite)
     call sTriangle_Differentiate(tri1, target, ddx, ddy, ddz, refhist, op)
     write(*,'(a)') 'DTRI'
     write(*,f1) op
     call sTriangle_Reject(tri1)
  enddo
end program refop
