! g6md: the MD driver on the GPU path.  Reads a Mixture input deck on
! standard input (the tags below, with Mixture's payloads) and writes @LOGD
! lines in Mixture's layout (Mixture.F90:2429-2490), so a deck that runs the
! reference runs here unchanged and the two logs can be compared line by line.
!
! Tags read: @DTPS dt | @MDLP steps | @NLOG interval | @RCOA rc taper |
! @BXLA L | @BOX3 Lx Ly Lz (an orthorhombic box) | @VOXA vx vy vz |
! @VOXN nx ny nz | @BOOK interval margin | @VSET T seed |
! @NOS2 T q dump z0..z4 zsum | @NOSE T q dump z0..z4 (the older form) |
! @PCN4 mode pex mass (Andersen barostat: mode 1 orthorhombic, every edge
! following the volume; mode 2 orz, the z edge alone) | @PCN2 (refused, as
! the reference refuses it) | @FAST | @SNPO | @SNPV |
! @ID08 model | @NX4A n, then n lines x y z a b c d | @NCMP 1 or 2.
! Two @ID08 + @NX4A blocks make a mixture of two rigid components (as
! Mixture.F90:636-712 adds a component per block): both step together on
! the GPU (g6_mixture_*), the @LOGD line then carries the pair energy of each
! pair set and the kinetic energy of each component, and the restart the
! two components; a mixture takes neither the barostat nor snapshots here.
! Command line: g6md [fp64|fp32] [skin]   (default fp64, skin 0: the
! reference's list rebuilt every step).  Without a skin on the command line,
! @BOOK interval margin (interval > 0) makes margin the skin: the rows then
! hold every pair within rc + margin and are rebuilt whenever a molecule has
! moved margin/2, so the interacting set is every pair within rc at every
! step: the run is the reference's unbooked run (the reference evaluates a
! list booked every interval steps, and its run departs from its unbooked
! one).
! @VOXN fixes the reference's cell counts; the GPU MD path's interacting set
! is every pair within rc whatever the grid (the reference's fixed-grid
! stencil covers them all only when the cells allow it, e.g. one cell), and
! the cell edge L/nx is passed on as the requested edge.
!
! Restart files are Mixture's: a binary fort.12 in the working directory,
! if present, is read first (the records Mixture.F90:400-577 reads: @RCOA
! @BOX3 @BOOK @VOXA @NLOG @DTPS @ASTP @NOS2 @PCN5 @TEST @MDLP @NCMP @TIE4
! @FIXC @ID08 @WTG2), the text deck on standard input may then override any tag,
! and the run continues from the step count @ASTP with the full Gear state
! of @WTG2.  At the end the state is written to fort.13 with the records
! and layout Mixture writes there (Mixture.F90:2530-2598,
! RigidModule.F90:801-828): angular velocities in external units (w dt).
program g6md
  use, intrinsic :: iso_c_binding
  use g6gpu
  implicit none
  character(len=256) :: line, tag, arg
  real(c_double) :: dt = 0.002d0, rc = 8.5d0, taper = 0.5d0, voxa(3) = 0d0
  real(c_double) :: vtemp = -1d0, skin = 0d0, nose(9), th(G6_THERMO_LEN)
  real(c_double) :: book_margin = 0d0, box3(3) = 0d0, bxla
  integer :: book_interval = 0, voxn(3) = 0
  logical :: skin_given = .false.
  integer :: nsteps = 0, nlog = 0, ios, i, done, chunk, target
  integer :: nstep0 = 0, snpo = 0, snpv = 0
  logical :: restart, fixc = .false.
  integer(c_int) :: amode = 0
  real(c_double) :: apex = 0d0, amass = 0d0, astate(8)
  real(c_double), target :: av(6)
  logical :: has_av = .false.
  integer(c_int) :: model = -1, prec = G6_FP64, n = 0, fast = 0
  integer(c_int32_t) :: vseed = 0
  integer(c_int64_t) :: h, h2 = 0, hm = 0
  ! the second component of a mixture; idm: the model of the last @ID08
  integer(c_int) :: model2 = -1, n2 = 0, idm = -1
  integer :: ncomp = 0, rcomp = 0
  logical :: mix
  real(c_double) :: parts(5)
  real(c_double), allocatable :: com2(:,:), quat2(:,:,:), dx2(:,:,:), w2(:,:,:)
  logical :: has_nose = .false.
  real(c_double), allocatable :: com(:,:), quat(:,:,:), dx(:,:,:), w(:,:,:)
  real(8), parameter :: I2J = 10d0, X2KX = 1d-3

  if (command_argument_count() >= 1) then
     call get_command_argument(1, arg)
     if (trim(arg) == 'fp32') prec = G6_FP32
  end if
  if (command_argument_count() >= 2) then
     call get_command_argument(2, arg)
     read(arg, *) skin
     skin_given = .true.
  end if

  inquire(file='fort.12', exist=restart)
  if (restart) call read_restart()

  do
     read(5, '(a)', iostat=ios) line
     if (ios /= 0) exit
     tag = adjustl(line)
     if (len_trim(tag) == 0) cycle
     select case (tag(1:5))
     case ('@DTPS')
        read(5, *) dt
     case ('@MDLP')
        read(5, *) nsteps
     case ('@NLOG')
        read(5, *) nlog
     case ('@FAST')
        read(5, *) i
        fast = 0
        if (i /= 0) fast = 1
     case ('@SNPO')
        read(5, *) snpo
     case ('@SNPV')
        read(5, *) snpv
     case ('@RCOA')
        read(5, *) rc, taper
     case ('@BXLA')
        read(5, *) bxla
        box3 = bxla
     case ('@BOX3')
        ! Box_ReadBOX3 (BoxModule.F90:109-116): three edges
        read(5, *) box3
     case ('@VOXA')
        read(5, *) voxa
        voxn = 0
     case ('@VOXN')
        ! Grid_ReadVOXN (GridModule.F90:139-159): fixed cell counts
        read(5, *) voxn
     case ('@BOOK')
        ! Book_ReadBOOK (BookModule.F90:48-64): interval and margin
        read(5, *) book_interval, book_margin
     case ('@NOSE')
        ! Nose_ReadNOSE (NoseModule.F90:90-103): no zetasum, and q in the
        ! older unit, times 21.8852721616497249235**2
        read(5, *) nose(1:8)
        nose(9) = 0d0
        nose(2) = nose(2) * 21.8852721616497249235d0**2
        has_nose = .true.
     case ('@PCN2')
        ! Andersen_ReadPCN2 (AndersenModule.F90:118-132) stops the run
        write(0, '(a)') '@PCN2 is not supported any more. Please use @PCN4 instead.'
        error stop 2
     case ('@VSET')
        read(5, *) vtemp, vseed
     case ('@NOS2')
        read(5, *) nose
        has_nose = .true.
     case ('@PCN4')
        read(5, *) amode, apex, amass
        has_av = .false.
     case ('@ID08')
        read(5, '(a)') line
        select case (trim(adjustl(line)))
        case ('TIP4P')
           idm = G6_TIP4P
        case ('TIP5P')
           idm = G6_TIP5P
        case ('ST2', 'ST2_____')
           idm = G6_ST2
        case ('SPC_E')
           idm = G6_SPCE
        case ('NVDE____')
           idm = G6_NVDE
        case ('OPLSMEOH')
           idm = G6_OPLSMEOH
        case default
           write(0, '(a,a)') 'g6md: @ID08 model not on the GPU path: ', trim(line)
           error stop 2
        end select
     case ('@NCMP')
        read(5, *) i
        if (i < 1 .or. i > 2) then
           write(0, '(a)') 'g6md: @NCMP: one or two components on the GPU path'
           error stop 2
        end if
     case ('@NX4A')
        ncomp = ncomp + 1
        if (ncomp == 1) then
           read(5, *) n
           if (idm >= 0) model = idm
           call read_nx4a(n, com, quat, dx, w)
        else if (ncomp == 2 .and. idm >= 0) then
           read(5, *) n2
           model2 = idm
           call read_nx4a(n2, com2, quat2, dx2, w2)
        else
           write(0, '(a)') 'g6md: @NX4A: at most two components, each after its @ID08'
           error stop 2
        end if
        idm = -1
     case default
        write(0, '(a,a)') 'g6md: tag not handled: ', trim(tag)
        error stop 2
     end select
  end do
  if (idm >= 0 .and. ncomp == 0) model = idm
  mix = ncomp == 2 .or. (ncomp == 0 .and. rcomp == 2)
  if (mix .and. (amode /= 0 .or. snpo > 0 .or. snpv > 0)) then
     write(0, '(a)') 'g6md: a mixture takes neither @PCN4 nor snapshots on the GPU path'
     error stop 2
  end if
  if (model < 0 .or. n < 2 .or. any(box3 <= 0d0)) then
     write(0, '(a)') 'g6md: the deck needs @ID08, @NX4A and @BXLA'
     error stop 2
  end if
  if (voxn(1) > 0) voxa = box3 / voxn
  if (voxa(1) <= 0d0) voxa = rc
  if (.not. skin_given .and. book_interval > 0) skin = book_margin

  call g6_check(g6_init_box3(h, model, n, box3, rc, taper, voxa, dt, prec, skin), 'g6_init_box3')
  ! @FAST (default off, Mixture.F90:357, 732-736): StdForce or the custom kernels
  call g6_check(g6_set_fast(h, fast), 'g6_set_fast')
  call g6_check(g6_upload_state(h, com, quat, dx, w), 'g6_upload_state')
  if (mix) then
     call g6_check(g6_init_box3(h2, model2, n2, box3, rc, taper, voxa, dt, prec, skin), &
          'g6_init_box3')
     call g6_check(g6_set_fast(h2, fast), 'g6_set_fast')
     call g6_check(g6_upload_state(h2, com2, quat2, dx2, w2), 'g6_upload_state')
  end if
  if (has_nose) then
     nose(2) = nose(2) * dt        ! @NOS2 mass, times dt (Mixture.F90:1206)
     call g6_check(g6_set_nose(h, 1_c_int, nose), 'g6_set_nose')
  end if
  if (mix) then
     call g6_check(g6_mixture_create(hm, h, h2), 'g6_mixture_create')
     if (vtemp >= 0d0) call g6_check(g6_mixture_vreset(hm, vtemp, vseed), 'g6_mixture_vreset')
  else
     if (vtemp >= 0d0) call g6_check(g6_vreset(h, vtemp, vseed), 'g6_vreset')
  end if
  if (amode /= 0) then
     if (has_av) then
        call g6_check(g6_set_andersen(h, amode, apex, amass, c_loc(av)), 'g6_set_andersen')
     else
        call g6_check(g6_set_andersen(h, amode, apex, amass, c_null_ptr), 'g6_set_andersen')
     end if
  end if

  ! Mixture runs @MDLP more steps from the count nstep (0, or @ASTP of a
  ! restart) and, after the step that starts at nstep and before nstep is
  ! advanced, writes the coordinate snapshot when mod(nstep, snpo) = 0, the
  ! coordinate and velocity snapshot when mod(nstep, snpv) = 0 and the log
  ! row when mod(nstep, nlog) = 0, in that order (Mixture.F90:1513,
  ! 2338-2385, 2429-2490, 2524)
  done = 0
  do while (done < nsteps)
     chunk = nsteps - done
     call next_event(nlog)
     call next_event(snpo)
     call next_event(snpv)
     if (mix) then
        call g6_check(g6_mixture_step(hm, int(chunk, c_int)), 'g6_mixture_step')
     else
        call g6_check(g6_step(h, int(chunk, c_int)), 'g6_step')
     end if
     done = done + chunk
     if (due(snpo) .or. due(snpv)) then
        call g6_check(g6_download_state(h, com, quat, dx, w), 'g6_download_state')
        call g6_check(g6_get_box3(h, box3), 'g6_get_box3')
        if (due(snpo)) call write_snapshot(nstep0 + done - 1, .false.)
        if (due(snpv)) call write_snapshot(nstep0 + done - 1, .true.)
     end if
     if (due(nlog)) then
        if (mix) then
           call g6_check(g6_mixture_thermo(hm, th, parts), 'g6_mixture_thermo')
           call write_logd2(nstep0 + done - 1, th)
        else
           call g6_check(g6_get_thermo(h, th), 'g6_get_thermo')
           call write_logd(nstep0 + done - 1, th)
        end if
     end if
  end do
  call g6_check(g6_download_state(h, com, quat, dx, w), 'g6_download_state')
  if (has_nose) call g6_check(g6_get_nose(h, nose), 'g6_get_nose')
  if (amode /= 0) call g6_check(g6_get_andersen(h, astate), 'g6_get_andersen')
  call g6_check(g6_get_box3(h, box3), 'g6_get_box3')
  if (mix) then
     call g6_check(g6_download_state(h2, com2, quat2, dx2, w2), 'g6_download_state')
     call g6_check(g6_mixture_free(hm), 'g6_mixture_free')
     call g6_check(g6_finalize(h2), 'g6_finalize')
  end if
  call g6_check(g6_finalize(h), 'g6_finalize')
  call write_restart()

contains

  ! the binary restart of a previous run (Mixture's fort.13 format)
  subroutine read_restart()
    character(len=5) :: btag
    character(len=8) :: id
    integer :: ib, nt, nb, m, c, k
    real(8) :: dbl, d9(9)
    open(12, file='fort.12', form='unformatted', access='sequential', status='old')
    do
       read(12, iostat=ios) btag
       if (ios /= 0) exit
       select case (btag)
       case ('@RCOA')
          read(12) rc, taper             ! co%out, co%out - co%in
       case ('@BOX3')
          read(12) box3
       case ('@BXLA')
          read(12) bxla
          box3 = bxla
       case ('@BOOK')
          read(12) book_interval, book_margin
       case ('@VOXN')
          read(12) voxn
       case ('@NOSE')
          read(12) d9(1:8)               ! Nose_ReadBinaryNOSE (NoseModule.F90:105-118)
          nose(1:8) = d9(1:8)
          nose(9) = 0d0
          nose(2) = nose(2) * 21.8852721616497249235d0**2
          has_nose = .true.
       case ('@VOXA')
          read(12) voxa
       case ('@NLOG')
          read(12) nlog
       case ('@DTPS')
          read(12) dt
       case ('@ASTP')
          read(12) nstep0
       case ('@NOS2')
          read(12) d9
          nose = d9
          has_nose = .true.
       case ('@PCN5')
          read(12) amode, apex, amass, av
          has_av = .true.
       case ('@TEST')
       case ('@MDLP')
          read(12) nsteps
       case ('@NCMP')
          read(12) nt
          if (nt < 1 .or. nt > 2) then
             write(0, '(a)') 'g6md: restart with more than two components'
             error stop 2
          end if
       case ('@TIE4')
          read(12) nt, nb
          if (nt /= 0) then
             write(0, '(a)') 'g6md: restart with tied molecules (@TIE4) is not on the GPU path'
             error stop 2
          end if
       case ('@FIXC')
          read(12) fixc
       case ('@ID08')
          read(12) id
          select case (id)
          case ('TIP4P   ')
             idm = G6_TIP4P
          case ('TIP5P   ')
             idm = G6_TIP5P
          case ('ST2_____')
             idm = G6_ST2
          case ('SPC_E   ')
             idm = G6_SPCE
          case ('NVDE____')
             idm = G6_NVDE
          case ('OPLSMEOH')
             idm = G6_OPLSMEOH
          case default
             write(0, '(a,a)') 'g6md: @ID08 model not on the GPU path: ', id
             error stop 2
          end select
       case ('@WTG2')
          read(12) m
          rcomp = rcomp + 1
          if (rcomp == 1) then
             n = m
             model = idm
             call read_wtg2(n, com, quat, dx, w)
          else if (rcomp == 2) then
             n2 = m
             model2 = idm
             call read_wtg2(n2, com2, quat2, dx2, w2)
          else
             write(0, '(a)') 'g6md: restart with more than two components'
             error stop 2
          end if
          idm = -1
       case default
          write(0, '(a,a)') 'g6md: restart record not handled: ', btag
          error stop 2
       end select
    end do
    close(12)
  end subroutine read_restart

  ! the binary restart of this run, records and order as Mixture writes them
  ! the reference's 8-character id of the model
  character(len=8) function model_id(mdl)
    integer(c_int), intent(in) :: mdl
    select case (mdl)
    case (G6_TIP4P)
       model_id = 'TIP4P   '
    case (G6_TIP5P)
       model_id = 'TIP5P   '
    case (G6_SPCE)
       model_id = 'SPC_E   '
    case (G6_NVDE)
       model_id = 'NVDE____'
    case (G6_OPLSMEOH)
       model_id = 'OPLSMEOH'
    case default
       model_id = 'ST2_____'
    end select
  end function model_id

  subroutine write_restart()
    character(len=8) :: id
    integer :: c, k
    id = model_id(model)
    open(13, file='fort.13', form='unformatted', access='sequential', status='replace')
    write(13) '@RCOA'
    write(13) rc, taper
    write(13) '@BOX3'
    write(13) box3
    write(13) '@BOOK'
    write(13) 1, skin
    write(13) '@VOXA'
    write(13) voxa
    write(13) '@NLOG'
    write(13) nlog
    write(13) '@DTPS'
    write(13) dt
    write(13) '@ASTP'
    write(13) nstep0 + nsteps
    if (has_nose) then
       nose(2) = nose(2) / dt          ! back to the @NOS2 mass
       write(13) '@NOS2'
       write(13) nose
    end if
    if (amode /= 0) then               ! Andersen_WriteBinaryPCN5
       write(13) '@PCN5'
       write(13) amode, astate
    end if
    write(13) '@TEST'
    write(13) '@MDLP'
    write(13) nsteps
    ! @NCMP, one @TIE4 per component, then @FIXC @ID08 @WTG2 per component
    ! (Mixture.F90:2562-2580)
    write(13) '@NCMP'
    if (mix) then
       write(13) 2
       write(13) '@TIE4'
       write(13) 0, 2
    else
       write(13) 1
    end if
    write(13) '@TIE4'
    write(13) 0, 2
    write(13) '@FIXC'
    write(13) fixc
    write(13) '@ID08'
    write(13) id
    write(13) '@WTG2'
    write(13) n
    write(13) (com(:, c), c = 1, 3), ((dx(:, c, k), k = 1, 5), c = 1, 3), &
         ((quat(:, c, k), k = 1, 5), c = 1, 4), ((w(:, c, k) * dt, k = 1, 5), c = 1, 3)
    if (mix) then
       write(13) '@FIXC'
       write(13) fixc
       write(13) '@ID08'
       write(13) model_id(model2)
       write(13) '@WTG2'
       write(13) n2
       write(13) (com2(:, c), c = 1, 3), ((dx2(:, c, k), k = 1, 5), c = 1, 3), &
            ((quat2(:, c, k), k = 1, 5), c = 1, 4), ((w2(:, c, k) * dt, k = 1, 5), c = 1, 3)
    end if
    close(13)
  end subroutine write_restart

  ! shortens chunk so that the run stops after the next step k (1-based in
  ! this run) with mod(nstep0 + k - 1, every) = 0
  subroutine next_event(every)
    integer, intent(in) :: every
    if (every > 0) then
       target = done + 1 + modulo(-(nstep0 + done), every)
       chunk = min(chunk, target - done)
    end if
  end subroutine next_event

  ! whether the step just finished is one of every's output steps
  logical function due(every)
    integer, intent(in) :: every
    due = .false.
    if (every > 0) due = modulo(nstep0 + done - 1, every) == 0
  end function due

  ! the snapshot blocks of Mixture.F90:2338-2385 for one rigid component:
  ! @ASTP, @BOX3 (Box_WriteBOX3, BoxModule.F90:127-135), @NCMP, then
  ! @FIXC + Rigid_WriteNX4A (RigidModule.F90:865-884) or Rigid_WriteWTG3
  ! (RigidModule.F90:830-863: velocity dx/dt in A/ps, angular velocity in
  ! internal units)
  subroutine write_snapshot(step, velocities)
    integer, intent(in) :: step
    logical, intent(in) :: velocities
    integer :: j
    write(6, '("@ASTP")')
    write(6, *) step
    write(6, '("@BOX3")')
    write(6, *) box3(1), box3(2), box3(3)
    write(6, '("@NCMP")')
    write(6, *) 1
    if (.not. velocities) then
       write(6, '("@FIXC")')
       write(6, *) fixc
    end if
    write(6, '("@ID08")')
    write(6, '(a8)') model_id(model)
    if (velocities) then
       write(6, '("@WTG3")')
       write(6, *) n
       do j = 1, n
          write(6, '(13(e17.10,1x))') com(j, 1), com(j, 2), com(j, 3), quat(j, 1, 1), &
               quat(j, 2, 1), quat(j, 3, 1), quat(j, 4, 1), dx(j, 1, 1) / dt, dx(j, 2, 1) / dt, &
               dx(j, 3, 1) / dt, w(j, 1, 1), w(j, 2, 1), w(j, 3, 1)
       end do
    else
       write(6, '("@NX4A")')
       write(6, *) n
       write(6, '(7(e17.10,1x))') (com(j, 1), com(j, 2), com(j, 3), quat(j, 1, 1), &
            quat(j, 2, 1), quat(j, 3, 1), quat(j, 4, 1), j = 1, n)
    end if
  end subroutine write_snapshot

  ! n lines x y z a b c d of an @NX4A block into a component's state
  ! (Rigid_ReadNX4A: order-1 quaternions, higher orders and velocities zero)
  subroutine read_nx4a(nn, c, q, d, ww)
    integer(c_int), intent(in) :: nn
    real(c_double), allocatable, intent(inout) :: c(:,:), q(:,:,:), d(:,:,:), ww(:,:,:)
    integer :: j
    if (allocated(c)) deallocate(c, q, d, ww)
    allocate(c(nn, 3), q(nn, 4, 5), d(nn, 3, 5), ww(nn, 3, 5))
    q = 0d0
    d = 0d0
    ww = 0d0
    do j = 1, nn
       read(5, *) c(j, 1), c(j, 2), c(j, 3), q(j, 1, 1), q(j, 2, 1), q(j, 3, 1), q(j, 4, 1)
    end do
  end subroutine read_nx4a

  ! the Gear state of an @WTG2 record of the restart into a component's state
  subroutine read_wtg2(nn, c, q, d, ww)
    integer(c_int), intent(in) :: nn
    real(c_double), allocatable, intent(inout) :: c(:,:), q(:,:,:), d(:,:,:), ww(:,:,:)
    integer :: a, k
    if (allocated(c)) deallocate(c, q, d, ww)
    allocate(c(nn, 3), q(nn, 4, 5), d(nn, 3, 5), ww(nn, 3, 5))
    read(12) (c(:, a), a = 1, 3), ((d(:, a, k), k = 1, 5), a = 1, 3), &
         ((q(:, a, k), k = 1, 5), a = 1, 4), ((ww(:, a, k), k = 1, 5), a = 1, 3)
    ww = ww / dt                     ! rigid_toInternalUnit (RigidModule.F90:1639-1652)
  end subroutine read_wtg2

  ! @LOGD columns of Mixture.F90:2453-2471 for two components: T, P, V, ep
  ! of the pair sets 1-1, 2-2, 1-2, ek of each component, ep/N, ek/N over
  ! all molecules (kJ/mol), nose_ep, nose_ek, andersen_ep, andersen_ek (0),
  ! H, exsum (0), H, emonsum (0)
  subroutine write_logd2(step, t)
    integer, intent(in) :: step
    real(c_double), intent(in) :: t(G6_THERMO_LEN)
    real(8) :: ep, ek, nm
    ep = t(1)
    ek = t(2) + t(3)
    nm = dble(n + n2)
    write(6, '(a)') '@LOGD'
    write(6, '(i10,30(1x,E24.17))') step, t(4), t(26), t(27), parts(1) * I2J * X2KX, &
         parts(2) * I2J * X2KX, parts(3) * I2J * X2KX, parts(4) * I2J * X2KX, &
         parts(5) * I2J * X2KX, ep * I2J / nm * X2KX, ek * I2J / nm * X2KX, &
         t(28) * X2KX, t(29) * X2KX, t(31) * X2KX, t(32) * X2KX, t(14), 0d0, t(14), 0d0
  end subroutine write_logd2

  ! @LOGD columns of Mixture.F90:2453-2471 for one component: T, P, V, ep,
  ! ek, ep/N, ek/N (kJ/mol), nose_ep, nose_ek, andersen_ep, andersen_ek, H,
  ! exsum (0), H, emonsum (0)
  subroutine write_logd(step, t)
    integer, intent(in) :: step
    real(c_double), intent(in) :: t(G6_THERMO_LEN)
    real(8) :: ep, ek
    ep = t(1)
    ek = t(2) + t(3)
    write(6, '(a)') '@LOGD'
    write(6, '(i10,30(1x,E24.17))') step, t(4), t(26), t(27), ep * I2J * X2KX, &
         ek * I2J * X2KX, ep * I2J / dble(n) * X2KX, ek * I2J / dble(n) * X2KX, &
         t(28) * X2KX, t(29) * X2KX, t(31) * X2KX, t(32) * X2KX, t(14), 0d0, t(14), 0d0
  end subroutine write_logd
end program g6md
