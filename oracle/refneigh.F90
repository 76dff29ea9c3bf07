! Fixture dumper for the neighbour arrays: drives the reference's own
! neighbors_module (built unmodified from /root/reference by oracle/ref.mk) on
! a given pair list and writes the arrays it builds as text, so that the
! oracle's C restatement and the GPU path can be pinned against the genuine
! code.
!
! Test infrastructure only; never part of the product.
!
! stdin (list-directed):
!   nsite npair
!   npair lines: i j ox oy oz          (1-based molecule indices)
!   nsite lines: x y z                 (molecule centres)
!   ncase
!   ncase lines: maxnei order radius cmaxnei corder cradius target dx dy dz
!     radius, cradius <= 0 mean "argument absent"
! For each case, to stdout:
!   'ALL2'  then per molecule: nnear / near dx dy dz dd per entry
!           (neighbors_all2)
!   'COMP'  the same for neighbor1_compress2 of every molecule's array
!   'DISP'  nmodified / per modified array: index nnear / entries
!           (neighbors_displacement of molecule `target` by (dx,dy,dz))
program refneigh
  use interaction_module
  use vector_module
  use neighbors_module
  implicit none
  integer :: nsite, npair, ncase, c, i, k, t
  integer :: maxnei, order, cmaxnei, corder, target, nmod
  real(kind=8) :: radius, cradius, ddx, ddy, ddz
  type(sInteraction) :: ww
  type(vector3), allocatable :: coord(:)
  type(sNeighbors), allocatable :: nei(:), cnei(:), newnei(:)
  integer, allocatable :: newidx(:)

  read(*,*) nsite, npair
  call interaction_initialize(ww, nsite, -1)
  ww%Npair = npair
  do k = 1, npair
     read(*,*) ww%pair_i(k), ww%pair_j(k), ww%ox(k), ww%oy(k), ww%oz(k)
  enddo
  allocate(coord(nsite), nei(nsite), cnei(nsite), newnei(nsite + 160), newidx(nsite + 160))
  do i = 1, nsite
     read(*,*) coord(i)%vec(1), coord(i)%vec(2), coord(i)%vec(3)
  enddo
  read(*,*) ncase
  do c = 1, ncase
     read(*,*) maxnei, order, radius, cmaxnei, corder, cradius, target, ddx, ddy, ddz
     if (radius > 0d0) then
        call neighbors_all2(nsite, nei, ww, coord, maxnei, order, radius)
     else
        call neighbors_all2(nsite, nei, ww, coord, maxnei, order)
     endif
     write(*,'(a)') 'ALL2'
     do i = 1, nsite
        call dump(nei(i))
     enddo
     do i = 1, nsite
        if (cradius > 0d0) then
           call neighbor1_compress2(cnei(i), nei(i), cmaxnei, corder, cradius)
        else
           call neighbor1_compress2(cnei(i), nei(i), cmaxnei, corder)
        endif
     enddo
     write(*,'(a)') 'COMP'
     do i = 1, nsite
        call dump(cnei(i))
     enddo
     call neighbors_displacement(nei, target, ddx, ddy, ddz, nmod, newnei, newidx)
     write(*,'(a)') 'DISP'
     write(*,*) nmod
     do t = 1, nmod
        write(*,*) newidx(t)
        call dump(newnei(t))
     enddo
  enddo

contains

  subroutine dump(nb)
    type(sNeighbors), intent(IN) :: nb
    integer :: j
    write(*,*) nb%nnear
    do j = 1, nb%nnear
       write(*,'(i8,4es26.17e3)') nb%near(j), nb%dx(j), nb%dy(j), nb%dz(j), nb%dd(j)
    enddo
  end subroutine dump
end program refneigh
