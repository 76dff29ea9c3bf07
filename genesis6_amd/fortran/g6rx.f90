! g6rx: replica-exchange MD on the GPU path, the replica branch of Mixture
! (Mixture.F90:1478-1660 with MDReplica.F90:32-82).  All replicas are
! systems of this one process, stepped together as one ladder (each MD
! kernel launched once for all of them), with node 0's exchanges between
! intervals.
!
! Standard input: a Mixture deck (the tags of g6md: @DTPS @MDLP @NLOG @RCOA
! @BXLA @VOXA @VSET @FAST @ID08 @NX4A) plus the MDReplica tags @MPIN nprocs
! followed by one file name per node, @SEED seed (the exchange stream),
! @INNR interval (steps between exchanges), @NRXG trials per exchange
! (default 30), @RXPR nnode, then 0-based links i j ending with a negative
! i (the exchange graph; default the complete graph, Mixture.F90:1127-1133).
! Every replica starts from the deck's configuration.  Each node's file
! holds its thermostat (@NOS2 T q dump z0..z4 zsum) and may repeat the
! MDReplica tags (node 0's stream and graph decide, as in the reference).
! Command line: g6rx [fp64|fp32] [skin].
!
! Output per node, in <node file>.log: @MDRT with the node's temperature at
! the start and after every exchange that changes it, and the @LOGD rows of
! Mixture's layout every @NLOG steps.  The reference re-balances the
! interval per node by wall-clock time (Mixture.F90:1510-1565); here every
! interval is @INNR steps, so that a run is reproducible.
program g6rx
  use, intrinsic :: iso_c_binding
  use g6gpu
  implicit none
  integer, parameter :: MAXREP = 256
  character(len=256) :: line, tag, arg
  character(len=256) :: nodefile(MAXREP)
  real(c_double) :: dt = 0.002d0, rc = 8.5d0, taper = 0.5d0, box = 0d0, voxa(3) = 0d0
  real(c_double) :: vtemp = -1d0, skin = 0d0, nose(9, MAXREP), temp(MAXREP)
  real(c_double), allocatable, target :: th(:,:)
  integer :: nsteps = 0, nlog = 0, ios, i, k, done, chunk, nprocs = 0, ninner = 0
  integer :: nexchange = 30, nlinks = -1, li(4096), lj(4096), nbond, intercept
  integer(c_int32_t) :: rseed = 5678, vseed = 0, nacc
  integer(c_int) :: model = -1, prec = G6_FP64, n = 0, fast = 0
  integer(c_int64_t) :: hs(MAXREP), hl, hx
  integer(c_int32_t), allocatable :: ba(:), bb(:)
  logical :: has_nose(MAXREP) = .false.
  real(c_double), allocatable :: com(:,:), quat(:,:,:), dx(:,:,:), w(:,:,:)
  real(8), parameter :: I2J = 10d0, X2KX = 1d-3
  character(len=100), parameter :: fmt0 = '(99(e25.17))'

  if (command_argument_count() >= 1) then
     call get_command_argument(1, arg)
     if (trim(arg) == 'fp32') prec = G6_FP32
  end if
  if (command_argument_count() >= 2) then
     call get_command_argument(2, arg)
     read(arg, *) skin
  end if

  call read_deck(5, .true., 0)
  if (nprocs < 2 .or. nprocs > MAXREP) then
     write(0, '(a)') 'g6rx: the deck needs @MPIN with 2 or more node files'
     error stop 2
  end if
  if (model < 0 .or. n < 2 .or. box <= 0d0) then
     write(0, '(a)') 'g6rx: the deck needs @ID08, @NX4A and @BXLA'
     error stop 2
  end if
  do k = 1, nprocs
     open(20, file=trim(nodefile(k)), status='old', action='read')
     call read_deck(20, .false., k)
     close(20)
     if (.not. has_nose(k)) then
        write(0, '(a,a)') 'g6rx: a node needs its thermostat (@NOS2): ', trim(nodefile(k))
        error stop 2
     end if
  end do
  if (ninner < 2) then
     write(0, '(a)') 'g6rx: @INNR of 2 steps or more'
     error stop 2
  end if
  if (voxa(1) <= 0d0) voxa = rc

  do k = 1, nprocs
     call g6_check(g6_init(hs(k), model, n, box, rc, taper, voxa(1), dt, prec, skin), 'g6_init')
     call g6_check(g6_set_fast(hs(k), fast), 'g6_set_fast')
     call g6_check(g6_upload_state(hs(k), com, quat, dx, w), 'g6_upload_state')
     temp(k) = nose(1, k)
     nose(2, k) = nose(2, k) * dt          ! @NOS2 mass, times dt (Mixture.F90:1206)
     call g6_check(g6_set_nose(hs(k), 1_c_int, nose(:, k)), 'g6_set_nose')
     if (vtemp >= 0d0) call g6_check(g6_vreset(hs(k), vtemp, vseed), 'g6_vreset')
  end do
  call g6_check(g6_ladder_create(hl, int(nprocs, c_int), hs), 'g6_ladder_create')
  call exchange_bonds()
  call g6_check(g6_rx_create(hx, int(nprocs, c_int), hs, int(nbond, c_int), ba, bb, &
       int(nexchange, c_int), rseed), 'g6_rx_create')
  allocate(th(G6_THERMO_LEN, nprocs))
  do k = 1, nprocs
     open(100 + k, file=trim(nodefile(k)) // '.log', status='replace')
     write(100 + k, '(a)') '@MDRT'
     write(100 + k, fmt0) temp(k)
  end do

  ! Mixture's loop (loop = 1..nloop, Mixture.F90:1478-1660, 2429): at the
  ! top of the iteration loop = intercept (ninnerloop, 2 ninnerloop, ...)
  ! the exchange, on the energies of the step just done; after each step
  ! the log row when mod(nstep, nlog) = 0 (nstep = loop - 1)
  done = 0
  intercept = ninner
  do while (done < nsteps)
     chunk = nsteps - done
     call next_event(nlog)
     if (intercept - 1 > done) chunk = min(chunk, intercept - 1 - done)
     call g6_check(g6_ladder_step(hl, int(chunk, c_int)), 'g6_ladder_step')
     done = done + chunk
     call g6_check(g6_ladder_thermo(hl, th), 'g6_ladder_thermo')
     if (nlog > 0) then
        if (modulo(done - 1, nlog) == 0) then
           do k = 1, nprocs
              call write_logd(100 + k, done - 1, th(:, k))
           end do
        end if
     end if
     if (done == intercept - 1 .and. intercept <= nsteps) then
        intercept = intercept + ninner
        call g6_check(g6_rx_exchange(hx, c_loc(th), nacc), 'g6_rx_exchange')
        do k = 1, nprocs
           call g6_check(g6_get_nose(hs(k), nose(:, k)), 'g6_get_nose')
           if (nose(1, k) /= temp(k)) then
              temp(k) = nose(1, k)
              write(100 + k, '(a)') '@MDRT'
              write(100 + k, fmt0) temp(k)
           end if
        end do
     end if
  end do
  do k = 1, nprocs
     close(100 + k)
  end do
  call g6_check(g6_rx_free(hx), 'g6_rx_free')
  call g6_check(g6_ladder_free(hl), 'g6_ladder_free')
  do k = 1, nprocs
     call g6_check(g6_finalize(hs(k)), 'g6_finalize')
  end do

contains

  ! the main deck (unit 5, main) or node k's file
  subroutine read_deck(u, main, k)
    integer, intent(in) :: u, k
    logical, intent(in) :: main
    integer :: j, a, b
    do
       read(u, '(a)', iostat=ios) line
       if (ios /= 0) exit
       tag = adjustl(line)
       if (len_trim(tag) == 0) cycle
       select case (tag(1:5))
       case ('@MPIN')
          read(u, *) nprocs
          if (nprocs > MAXREP) then
             write(0, '(a)') 'g6rx: more nodes than the driver holds'
             error stop 2
          end if
          do j = 1, nprocs
             read(u, '(a)') nodefile(j)
             nodefile(j) = adjustl(nodefile(j))
          end do
       case ('@SEED')
          read(u, *) j
          if (main .or. k == 1) rseed = int(j, c_int32_t)
       case ('@INNR')
          read(u, *) j
          if (main .or. k == 1) ninner = j
       case ('@NRXG')
          read(u, *) j
          if (main .or. k == 1) nexchange = j
       case ('@RXPR')
          ! graph_read (Graph.F90:112-127): nnode, then 0-based links
          read(u, *) j
          nlinks = 0
          do
             read(u, *) a, b
             if (a < 0) exit
             nlinks = nlinks + 1
             li(nlinks) = a
             lj(nlinks) = b
          end do
       case ('@TEMP')
          read(u, *) temp(max(k, 1))
       case ('@NOS2')
          if (main) then
             write(0, '(a)') 'g6rx: @NOS2 belongs in the node files'
             error stop 2
          end if
          read(u, *) nose(:, k)
          has_nose(k) = .true.
       case ('@DTPS')
          read(u, *) dt
       case ('@MDLP')
          read(u, *) nsteps
       case ('@NLOG')
          read(u, *) nlog
       case ('@FAST')
          read(u, *) j
          fast = 0
          if (j /= 0) fast = 1
       case ('@RCOA')
          read(u, *) rc, taper
       case ('@BXLA')
          read(u, *) box
       case ('@VOXA')
          read(u, *) voxa
       case ('@VSET')
          read(u, *) vtemp, vseed
       case ('@ID08')
          read(u, '(a)') line
          select case (trim(adjustl(line)))
          case ('TIP4P')
             model = G6_TIP4P
          case ('TIP5P')
             model = G6_TIP5P
          case ('ST2', 'ST2_____')
             model = G6_ST2
          case ('SPC_E')
             model = G6_SPCE
          case ('NVDE____')
             model = G6_NVDE
          case ('OPLSMEOH')
             model = G6_OPLSMEOH
          case default
             write(0, '(a,a)') 'g6rx: @ID08 model not on the GPU path: ', trim(line)
             error stop 2
          end select
       case ('@NX4A')
          read(u, *) n
          if (allocated(com)) deallocate(com, quat, dx, w)
          allocate(com(n, 3), quat(n, 4, 5), dx(n, 3, 5), w(n, 3, 5))
          quat = 0d0
          dx = 0d0
          w = 0d0
          do j = 1, n
             read(u, *) com(j, 1), com(j, 2), com(j, 3), quat(j, 1, 1), quat(j, 2, 1), &
                  quat(j, 3, 1), quat(j, 4, 1)
          end do
       case default
          write(0, '(a,a)') 'g6rx: tag not handled: ', trim(tag)
          error stop 2
       end select
    end do
  end subroutine read_deck

  ! graph2pair (Graph.F90:34-52) of the @RXPR graph, or of graph_perfect
  ! (Graph.F90:93-110): node i's neighbours in the order they were linked
  subroutine exchange_bonds()
    integer :: nnei(MAXREP), nei(MAXREP, MAXREP), i, j, q
    nnei = 0
    if (nlinks < 0) then
       do i = 1, nprocs
          do j = i + 1, nprocs
             call link(i, j)
             call link(j, i)
          end do
       end do
    else
       do q = 1, nlinks
          call link(li(q) + 1, lj(q) + 1)
          call link(lj(q) + 1, li(q) + 1)
       end do
    end if
    nbond = sum(nnei(1:nprocs))
    allocate(ba(nbond), bb(nbond))
    q = 0
    do i = 1, nprocs
       do j = 1, nnei(i)
          q = q + 1
          ba(q) = int(i, c_int32_t)
          bb(q) = int(nei(j, i), c_int32_t)
       end do
    end do
  contains
    subroutine link(a, b)
      integer, intent(in) :: a, b
      nnei(a) = nnei(a) + 1
      nei(nnei(a), a) = b
    end subroutine link
  end subroutine exchange_bonds

  subroutine next_event(every)
    integer, intent(in) :: every
    if (every > 0) chunk = min(chunk, 1 + modulo(-done, every))
  end subroutine next_event

  ! @LOGD columns of Mixture.F90:2453-2471 (as g6md writes them)
  subroutine write_logd(u, step, t)
    integer, intent(in) :: u, step
    real(c_double), intent(in) :: t(G6_THERMO_LEN)
    real(8) :: ep, ek
    ep = t(1)
    ek = t(2) + t(3)
    write(u, '(a)') '@LOGD'
    write(u, '(i10,30(1x,E24.17))') step, t(4), t(26), t(27), ep * I2J * X2KX, &
         ek * I2J * X2KX, ep * I2J / dble(n) * X2KX, ek * I2J / dble(n) * X2KX, &
         t(28) * X2KX, t(29) * X2KX, t(31) * X2KX, t(32) * X2KX, t(14), 0d0, t(14), 0d0
  end subroutine write_logd
end program g6rx
