! g6md: the MD driver on the GPU path.  Reads a Mixture input deck on
! standard input (the tags below, with Mixture's payloads) and writes @LOGD
! lines in Mixture's layout (Mixture.F90:2429-2490), so a deck that runs the
! reference runs here unchanged and the two logs can be compared line by line.
!
! Tags read: @DTPS dt | @MDLP steps | @NLOG interval | @RCOA rc taper |
! @BXLA L | @VOXA vx vy vz | @VSET T seed | @NOS2 T q dump z0..z4 zsum |
! @ID08 model | @NX4A n, then n lines x y z a b c d.
! Command line: g6md [fp64|fp32] [skin]   (default fp64, skin 0: the
! reference's list rebuilt every step).
program g6md
  use, intrinsic :: iso_c_binding
  use g6gpu
  implicit none
  character(len=256) :: line, tag, arg
  real(c_double) :: dt = 0.002d0, rc = 8.5d0, taper = 0.5d0, box = 0d0, voxa(3) = 0d0
  real(c_double) :: vtemp = -1d0, skin = 0d0, nose(9), th(G6_THERMO_LEN)
  integer :: nsteps = 0, nlog = 0, ios, i, done, chunk, target
  integer(c_int) :: model = -1, prec = G6_FP64, n = 0
  integer(c_int32_t) :: vseed = 0
  integer(c_int64_t) :: h
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
  end if

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
     case ('@RCOA')
        read(5, *) rc, taper
     case ('@BXLA')
        read(5, *) box
     case ('@VOXA')
        read(5, *) voxa
     case ('@VSET')
        read(5, *) vtemp, vseed
     case ('@NOS2')
        read(5, *) nose
        has_nose = .true.
     case ('@ID08')
        read(5, '(a)') line
        select case (trim(adjustl(line)))
        case ('TIP4P')
           model = G6_TIP4P
        case ('TIP5P')
           model = G6_TIP5P
        case ('ST2', 'ST2_____')
           model = G6_ST2
        case default
           write(0, '(a,a)') 'g6md: @ID08 model not on the GPU path: ', trim(line)
           error stop 2
        end select
     case ('@NX4A')
        read(5, *) n
        allocate(com(n, 3), quat(n, 4, 5), dx(n, 3, 5), w(n, 3, 5))
        quat = 0d0
        dx = 0d0
        w = 0d0
        do i = 1, n
           read(5, *) com(i, 1), com(i, 2), com(i, 3), quat(i, 1, 1), quat(i, 2, 1), &
                quat(i, 3, 1), quat(i, 4, 1)
        end do
     case default
        write(0, '(a,a)') 'g6md: tag not handled: ', trim(tag)
        error stop 2
     end select
  end do
  if (model < 0 .or. n < 2 .or. box <= 0d0) then
     write(0, '(a)') 'g6md: the deck needs @ID08, @NX4A and @BXLA'
     error stop 2
  end if
  if (voxa(1) <= 0d0) voxa(1) = rc

  call g6_check(g6_init(h, model, n, box, rc, taper, voxa(1), dt, prec, skin), 'g6_init')
  call g6_check(g6_upload_state(h, com, quat, dx, w), 'g6_upload_state')
  if (has_nose) then
     nose(2) = nose(2) * dt        ! @NOS2 mass, times dt (Mixture.F90:1206)
     call g6_check(g6_set_nose(h, 1_c_int, nose), 'g6_set_nose')
  end if
  if (vtemp >= 0d0) call g6_check(g6_vreset(h, vtemp, vseed), 'g6_vreset')

  ! Mixture counts nstep from 0 and logs when mod(nstep, nlog) = 0 after
  ! the step numbered nstep + 1 (Mixture.F90:360, 2429-2490, 2524): the
  ! log lines come after steps 1, nlog + 1, 2 nlog + 1, ...
  done = 0
  do while (done < nsteps)
     chunk = nsteps - done
     if (nlog > 0) then
        target = (done / nlog) * nlog + 1
        if (target <= done) target = target + nlog
        chunk = min(chunk, target - done)
     end if
     call g6_check(g6_step(h, int(chunk, c_int)), 'g6_step')
     done = done + chunk
     if (nlog > 0) then
        if (mod(done - 1, nlog) == 0) then
           call g6_check(g6_get_thermo(h, th), 'g6_get_thermo')
           call write_logd(done - 1, th)
        end if
     end if
  end do
  call g6_check(g6_download_state(h, com, quat, dx, w), 'g6_download_state')
  call g6_check(g6_finalize(h), 'g6_finalize')

contains

  ! @LOGD columns of Mixture.F90:2453-2471 for one component and no Andersen
  ! barostat: T, P, V, ep, ek, ep/N, ek/N (kJ/mol), nose_ep, nose_ek, 0, 0,
  ! H, exsum (0), H, emonsum (0)
  subroutine write_logd(step, t)
    integer, intent(in) :: step
    real(c_double), intent(in) :: t(G6_THERMO_LEN)
    real(8) :: ep, ek
    ep = t(1)
    ek = t(2) + t(3)
    write(6, '(a)') '@LOGD'
    write(6, '(i10,30(1x,E24.17))') step, t(4), t(26), t(27), ep * I2J * X2KX, &
         ek * I2J * X2KX, ep * I2J / dble(n) * X2KX, ek * I2J / dble(n) * X2KX, &
         t(28) * X2KX, t(29) * X2KX, 0d0, 0d0, t(14), 0d0, t(14), 0d0
  end subroutine write_logd
end program g6md
