! g6mc: the mc2 Monte Carlo driver on the GPU path.  Reads mc2's own input
! files and writes mc2's output file, so a run set up for the reference's
! mc2 runs here unchanged and the two outputs can be compared.
!
! Command line: g6mc config-file.  The configuration file is mc2's: @MPIN
! nprocs followed by one input file name per node (mc2.F90:190-195,
! mc2util.F90:239-255); one node is run (nprocs 1).  The node's input file
! holds ReadParams' tags (mc2util.F90:119-300): @SEED seed | @MCLP loops |
! @INNR inner loops | @NLOG interval | @SNPO interval | @TEMP T |
! @DELT n, then n lines delta rtratio | @RCOA rc taper | @BXLA L |
! @BOX3 L L L | @VOXA | @VOXN | @BOOK | @NCMP 1 | @ID08 model |
! @NX4A n, then n lines x y z a b c d | the umbrellas: @HIST (nbin 21,
! width, minimum, then the 21 counts of the reference histogram), @REPQ
! center coeffi (a TRIANGLE umbrella), @GEUM kind, then center coeffi
! (kinds TRIANGLE, LOGTRIAN, Q6LOCAL_, EPOP____; at most 4).  Reading stops
! after the @NX4A of the last component, as ReadParams stops.  Tags it does
! not know are passed over, as ReadParams passes them over.  A saved random
! stream (@MTRN) and more than one node stop the run: those paths are not
! on the GPU.
!
! The run is mc2's main loop (mc2.F90:379-899): loop = 0, 1, ...; each loop
! is 2N single-molecule trials (g6_mc_trials, the decisions of mc2's trial
! loop, with the umbrellas' change in every trial's energy difference), then
! the configuration block when mod(loop, snpo) = 0 and the @MC2O line (loop,
! energy in internal units, 1/beta, then op, center and coeffi of each
! umbrella, mc2.F90:876-884) when mod(loop, nlog) = 0.  The
! output file is the node's input file name with ".out" appended; at the end
! it gets mc2's @CONT block: @DELT, @TEMP, @BOX3 and the final @NX4A
! (mc2.F90:1009-1025), without the saved random stream (@MTRN).
program g6mc
  use, intrinsic :: iso_c_binding
  use g6gpu
  implicit none
  character(len=256) :: cfg, infile
  character(len=5) :: tag
  character(len=8) :: lastid = '        ', id = '        '
  ! defaults of sSystem_new (mc2util.F90:94-117): T 400 K, 100 inner loops
  real(c_double) :: temp = 400d0, rc = -1d0, taper = 0d0, box = 0d0, box3(3), voxa(3) = 0d0
  real(c_double) :: epsum, book_margin
  integer :: nloop = 0, ninner = 100, nlog = 0, snpo = 0, nprocs = 0, ncompo_in = 0
  integer :: ncompo = 0, nd = 0, i, ios, loop, outer, inner, nouter, ninnerloop, voxn(3)
  integer :: book_interval
  integer(c_int32_t) :: seed = 5678        ! DEFAULT_SEED (mc2.F90:62)
  integer(c_int) :: model = -1, n = 0
  integer(c_int64_t) :: h
  integer(c_long_long) :: nacc
  real(c_double), allocatable :: com(:,:), quat(:,:,:), dx(:,:,:), w(:,:,:)
  real(c_double), allocatable :: delta(:), rtratio(:), com1(:,:), q1(:,:)
  integer(c_int32_t) :: ukind(4) = 0
  real(c_double) :: ucenter(4) = 0d0, ucoeffi(4) = 0d0, uop(4) = 0d0, refhist(21) = 0d0
  integer :: numb = 0
  logical :: have_hist = .false.
  character(len=100), parameter :: fmt0 = '(99(e25.17))', fmt1 = '(i11,1x,99(e25.17))'
  integer, parameter :: OUT = 13

  if (command_argument_count() < 1) then
     write(0, '(a)') 'usage: g6mc config-file'
     error stop 2
  end if
  call get_command_argument(1, cfg)
  infile = ''
  call read_params(trim(cfg))
  if (nprocs > 1) then
     write(0, '(a)') 'g6mc: runs one node (@MPIN 1)'
     error stop 2
  end if
  if (len_trim(infile) > 0) call read_params(trim(infile))
  if (model < 0 .or. n < 2 .or. box <= 0d0 .or. rc <= 0d0) then
     write(0, '(a)') 'g6mc: the input needs @ID08, @NX4A, @BXLA (or @BOX3) and @RCOA'
     error stop 2
  end if
  if (.not. allocated(delta)) then
     allocate(delta(n), rtratio(n))
     delta = 1d-1                          ! mc2.F90:365-370
     rtratio = 1d0
  else if (nd /= n) then
     write(0, '(a)') 'g6mc: @DELT and @NX4A disagree on the molecule count'
     error stop 2
  end if
  if (voxa(1) <= 0d0) voxa = rc

  call g6_check(g6_init(h, model, n, box, rc, taper, voxa(1), 0.002d0, G6_FP64, 0d0), 'g6_init')
  call g6_check(g6_upload_state(h, com, quat, dx, w), 'g6_upload_state')
  call g6_check(g6_mc_begin(h, temp, seed, delta, rtratio), 'g6_mc_begin')
  if (numb > 0) then
     if (any(ukind(1:numb) == G6_UMB_TRIANGLE .or. ukind(1:numb) == G6_UMB_LOGTRIAN) &
          .and. .not. have_hist) then
        write(0, '(a)') 'g6mc: a triangle umbrella needs @HIST before it'
        error stop 2
     end if
     call g6_check(g6_mc_set_umbrella(h, int(numb, c_int), ukind, ucenter, ucoeffi, refhist), &
          'g6_mc_set_umbrella')
  end if
  allocate(com1(n, 3), q1(n, 4))

  open(OUT, file=trim(infile) // '.out', status='replace')
  loop = 0
  nouter = (nloop - 1) / ninner + 1
  ninnerloop = nloop
  if (ninner < nloop) ninnerloop = ninner
  do outer = 1, nouter
     do inner = 1, ninnerloop
        call g6_check(g6_mc_trials(h, int(2 * n, c_int), nacc), 'g6_mc_trials')
        call g6_check(g6_mc_get(h, com1, q1, delta, rtratio, epsum), 'g6_mc_get')
        if (numb > 0) call g6_check(g6_mc_get_umbrella(h, uop, ucenter, ucoeffi), &
             'g6_mc_get_umbrella')
        if (snpo > 0) then
           if (mod(loop, snpo) == 0) then
              write(OUT, '("@NCMP")')
              write(OUT, *) 1
              call write_config()
           end if
        end if
        if (nlog > 0) then
           if (mod(loop, nlog) == 0) then
              write(OUT, '("@MC2O")')
              write(OUT, fmt1) loop, epsum, 1d0 / (1d0 / temp), &
                   (uop(i), ucenter(i), ucoeffi(i), i = 1, numb)
           end if
        end if
        loop = loop + 1
     end do
  end do
  write(OUT, '("@CONT")')
  write(OUT, '("@DELT")')
  write(OUT, *) n
  do i = 1, n
     write(OUT, fmt0) delta(i), rtratio(i)
  end do
  write(OUT, '("@TEMP")')
  write(OUT, *) temp
  call write_config()
  close(OUT)
  call g6_check(g6_finalize(h), 'g6_finalize')

contains

  ! ReadParams (mc2util.F90:119-300) for the tags above
  subroutine read_params(path)
    character(len=*), intent(in) :: path
    integer :: u, k
    real(c_double) :: w, wmin
    character(len=8) :: uid
    open(newunit=u, file=path, status='old', action='read')
    do
       read(u, '(a5)', iostat=ios) tag
       if (ios /= 0) exit
       select case (tag)
       case ('@DELT')
          read(u, *) nd
          if (allocated(delta)) deallocate(delta, rtratio)
          allocate(delta(nd), rtratio(nd))
          do k = 1, nd
             read(u, *) delta(k), rtratio(k)
          end do
       case ('@SEED')
          read(u, *) i
          seed = int(i, c_int32_t)
       case ('@INNR')
          read(u, *) ninner
       case ('@SNPO')
          read(u, *) snpo
       case ('@NLOG')
          read(u, *) nlog
       case ('@MCLP')
          read(u, *) nloop
       case ('@TEMP')
          read(u, *) temp
       case ('@MPIN')
          read(u, *) nprocs
          read(u, *) infile
       case ('@ID08')
          read(u, *) lastid
       case ('@NCMP')
          read(u, *) ncompo_in
          if (ncompo_in > 1) then
             write(0, '(a)') 'g6mc: one component (@NCMP 1) on the GPU path'
             error stop 2
          end if
       case ('@NX4A')
          call read_nx4a(u)
          ncompo = ncompo + 1
          if (ncompo == ncompo_in) exit
       case ('@RCOA')
          read(u, *) rc, taper
       case ('@BXLA')
          read(u, *) box
       case ('@BOX3')
          read(u, *) box3
          if (box3(2) /= box3(1) .or. box3(3) /= box3(1)) then
             write(0, '(a)') 'g6mc: @BOX3 with unequal edges: the GPU path runs cubic boxes'
             error stop 2
          end if
          box = box3(1)
       case ('@VOXA')
          read(u, *) voxa
       case ('@VOXN')
          read(u, *) voxn
       case ('@BOOK')
          read(u, *) book_interval, book_margin
       case ('@HIST')                          ! histogram_load
          read(u, *) k, w, wmin
          if (k /= 21) then
             write(0, '(a)') 'g6mc: @HIST needs the 21 bins of the triangle histogram'
             error stop 2
          end if
          do k = 1, 21
             read(u, *) refhist(k)
          end do
          have_hist = .true.
       case ('@REPQ')                          ! a TRIANGLE umbrella
          call add_umbrella(G6_UMB_TRIANGLE, u)
       case ('@GEUM')
          read(u, *) uid
          select case (uid)
          case ('TRIANGLE')
             call add_umbrella(G6_UMB_TRIANGLE, u)
          case ('LOGTRIAN')
             call add_umbrella(G6_UMB_LOGTRIAN, u)
          case ('Q6LOCAL_')
             call add_umbrella(G6_UMB_Q6LOCAL, u)
          case ('EPOP____')
             call add_umbrella(G6_UMB_EPOP, u)
          case default
             write(0, '(a,a)') 'g6mc: umbrella not on the GPU path: ', uid
             error stop 2
          end select
       case ('@MTRN', '@RXPR', '@NRXG')
          write(0, '(a,a)') 'g6mc: not on the GPU path: ', tag
          error stop 2
       case default
       end select
    end do
    close(u)
  end subroutine read_params

  ! umbrella_ReadParam (Umbrella.F90:18-23): center coeffi
  subroutine add_umbrella(kind, u)
    integer(c_int), intent(in) :: kind
    integer, intent(in) :: u
    if (numb == 4) then
       write(0, '(a)') 'g6mc: at most 4 umbrellas on the GPU path'
       error stop 2
    end if
    numb = numb + 1
    ukind(numb) = kind
    read(u, *) ucenter(numb), ucoeffi(numb)
  end subroutine add_umbrella

  subroutine read_nx4a(u)
    integer, intent(in) :: u
    integer :: k
    select case (lastid)
    case ('TIP4P   ')
       model = G6_TIP4P
    case ('TIP5P   ')
       model = G6_TIP5P
    case ('NVDE____')
       model = G6_NVDE
    case ('OPLSMEOH')
       model = G6_OPLSMEOH
    case default
       write(0, '(a,a)') 'g6mc: @ID08 model not on the GPU MC path: ', lastid
       error stop 2
    end select
    id = lastid
    read(u, *) n
    if (allocated(com)) deallocate(com, quat, dx, w)
    allocate(com(n, 3), quat(n, 4, 5), dx(n, 3, 5), w(n, 3, 5))
    quat = 0d0
    dx = 0d0
    w = 0d0
    do k = 1, n
       read(u, *) com(k, 1), com(k, 2), com(k, 3), quat(k, 1, 1), quat(k, 2, 1), &
            quat(k, 3, 1), quat(k, 4, 1)
    end do
  end subroutine read_nx4a

  ! Box_WriteBOX3 (BoxModule.F90:127-135), Rigid_WriteNX4A
  ! (RigidModule.F90:865-884)
  subroutine write_config()
    integer :: j
    write(OUT, '("@BOX3")')
    write(OUT, *) box, box, box
    write(OUT, '("@ID08")')
    write(OUT, '(a8)') id
    write(OUT, '("@NX4A")')
    write(OUT, *) n
    write(OUT, '(7(e17.10,1x))') (com1(j, 1), com1(j, 2), com1(j, 3), q1(j, 1), q1(j, 2), &
         q1(j, 3), q1(j, 4), j = 1, n)
  end subroutine write_config
end program g6mc
