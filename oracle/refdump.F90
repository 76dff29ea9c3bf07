! Fixture dumper: drives the reference's own modules (built unmodified from
! /root/reference by oracle/ref.mk) on one configuration and writes the
! intermediate results of one interaction step as text, so that the oracle's
! C restatement can be pinned against the genuine code.
!
! Test infrastructure only; never part of the product.
!
! stdin (list-directed):
!   id                    8-char model id, e.g. 'TIP4P   '
!   n L rc taper req      molecules, box edge, cutoff, switch width, VOXA cell
!   n lines: x y z qa qb qc qd      (@NX4A order)
! Files written to the current directory:
!   pairs.txt     npair0 npair / k i j ox oy oz eratio fratio
!   force_std.txt StdForce: nsite / fx fy fz per site / ep per pair / vrmat
!   force_fast.txt custom kernel (TIP4P, ST2) in the same layout
!   rigid.txt     per molecule: F(3) tau_body(3) from rigid_collectforce (StdForce)
!   epot.txt      PairPotential (energy-only path) per pair
!   sites.txt     site coordinates after rigid_setsiteposition
!   cells.txt     ndiv(3), cell of each molecule, homo stencil n/n0 and offsets
program refdump
  use common_module
  use physconst_module
  use vector_module
  use box_module
  use grid_module
  use interaction_module
  use cutoff_module
  use site_module
  use mol_module
  use rigid_module
  use standard_interaction_module
  use rigid_setup_module
  use interaction2_module
  implicit none
  character(len=8) :: id
  integer :: n, i, k, s, err
  real(kind=8) :: bl, rc, taper, req
  type(sBox) :: box
  type(sGrid) :: g
  type(sAddress) :: addr
  type(sInteraction) :: ww
  type(sCutoff) :: co
  type(sSite) :: site
  type(sMol) :: mol
  type(sRigid) :: rigid
  type(sStdInt) :: si
  type(vector3), allocatable :: com(:)
  real(kind=8), allocatable :: ep1(:)
  real(kind=8) :: vrmat(3,3)
  character(len=*), parameter :: f1 = '(es26.17e3)'

  read(*,*) id
  read(*,*) n, bl, rc, taper, req
  call new(site, MAXsite_total)
  err = rigid_setup(rigid, id, .false., mol, si)
  if (err /= 0) stop 'unknown id'
  call Mol_allocate(mol, site, n, RIGID_MODE)
  mol%id = id
  do i = 1, n
     read(*,*) rigid%mol(i)%com%vec(1), rigid%mol(i)%com%vec(2), rigid%mol(i)%com%vec(3), &
          rigid%mol(i)%quat(1)%vec(1), rigid%mol(i)%quat(1)%vec(2), &
          rigid%mol(i)%quat(1)%vec(3), rigid%mol(i)%quat(1)%vec(4)
  enddo
  ! @BXLA
  box%size%vec(:) = bl
  box%invsize%vec(:) = 1d0 / bl
  box%mode = BOX_ORTHO
  ! @RCOA
  call cutoff_initialize(co, rc - taper, rc)
  co%mode = CUTOFF_POLY
  ! @VOXA
  call new(g)
  g%requested%vec(1:3) = req
  g%mode = GRID_ADAPTIVE
  call interaction_initialize(ww, n, -1)
  allocate(com(n))

  call rigid_setsiteposition(rigid, mol, site)
  do i = 1, n
     com(i) = rigid%mol(i)%com
  enddo
  call Grid_Update(g, box, 0d0)
  call Interaction_Reset(ww)
  call Address_Assign(addr, box, g, n, com)
  call Grid_NeighborList(g, g%homo, ww, addr, addr)
  call Interaction_Compress2(ww, .true., box, co%out**2, com, com, .false.)
  call CutOff_Dimmer2(co, ww, com, com)

  open(10, file='cells.txt')
  write(10,*) g%ndivx, g%ndivy, g%ndivz, g%homo%n, g%homo%n0
  do k = 1, g%homo%n
     write(10,*) g%homo%x(k), g%homo%y(k), g%homo%z(k)
  enddo
  do k = 1, g%homo%n0
     write(10,*) g%homo%x0(k), g%homo%y0(k), g%homo%z0(k)
  enddo
  do i = 1, n
     write(10,*) addr%cell(i)
  enddo
  close(10)

  open(10, file='sites.txt')
  write(10,*) mol%nsite
  do k = 1, n * mol%nsite
     write(10,'(3es26.17e3)') site%x(k), site%y(k), site%z(k)
  enddo
  close(10)

  open(10, file='pairs.txt')
  write(10,*) ww%npair0, ww%npair
  do k = 1, ww%npair
     write(10,'(3i8,5es26.17e3)') k, ww%pair_i(k), ww%pair_j(k), ww%ox(k), ww%oy(k), ww%oz(k), &
          ww%eratio(k), ww%fratio(k)
  enddo
  close(10)

  allocate(ep1(ww%npair))
  ! generic StdForce path (the reference default, @FAST 0)
  call site_resetforce(site)
  call force(ww, mol, mol, site, ep1, vrmat, .false., si, si)
  call dumpforce('force_std.txt')
  call rigid_collectforce(rigid, mol, site)
  open(10, file='rigid.txt')
  do i = 1, n
     write(10,'(6es26.17e3)') rigid%mol(i)%forcex, rigid%mol(i)%forcey, rigid%mol(i)%forcez, &
          rigid%mol(i)%torquex, rigid%mol(i)%torquey, rigid%mol(i)%torquez
  enddo
  close(10)
  ! hand-unrolled custom kernel (@FAST 1) where one exists
  if (id == 'TIP4P   ' .or. id == 'ST2_____') then
     call site_resetforce(site)
     call force(ww, mol, mol, site, ep1, vrmat, .true., si, si)
     call dumpforce('force_fast.txt')
  endif
  ! energy-only path used by the MC trial
  call PairPotential(ww, mol, mol, site, ep1, .false., si, si)
  open(10, file='epot.txt')
  do k = 1, ww%npair
     write(10,f1) ep1(k)
  enddo
  close(10)

contains

  subroutine dumpforce(fname)
    character(len=*), intent(in) :: fname
    integer :: kk, a, b
    open(11, file=fname)
    write(11,*) mol%nsite, ww%npair
    do kk = 1, n * mol%nsite
       write(11,'(3es26.17e3)') site%fx(kk), site%fy(kk), site%fz(kk)
    enddo
    do kk = 1, ww%npair
       write(11,f1) ep1(kk)
    enddo
    do a = 1, 3
       write(11,'(3es26.17e3)') (vrmat(a,b), b = 1, 3)
    enddo
    close(11)
  end subroutine dumpforce
end program refdump
