! ISO_C_BINDING interface to libg6gpu (include/g6gpu.h): one interface per
! entry point, argument for argument.  Arrays are passed as assumed-size
! dummies, so a caller hands over its own contiguous arrays (component-major
! layouts such as com(n,3), quat(n,4,5), dx(n,3,5) are the C layouts).
! Every function returns 0 or a negative G6_E* code.
module g6gpu
  use, intrinsic :: iso_c_binding
  implicit none
  integer(c_int), parameter :: G6_TIP4P = 0, G6_TIP5P = 1, G6_ST2 = 2, G6_SPCE = 3, &
       G6_NVDE = 4, G6_OPLSMEOH = 5
  integer(c_int), parameter :: G6_FP64 = 0, G6_FP32 = 1
  integer(c_int), parameter :: G6_EHIP = -1, G6_EARG = -2, G6_ECAPACITY = -3, &
       G6_ENODEVICE = -4
  integer, parameter :: G6_THERMO_LEN = 32
  integer(c_int), parameter :: G6_LADDER_MAP_XCD = 0, G6_LADDER_MAP_FLAT = 1
  integer(c_int), parameter :: G6_UMB_TRIANGLE = 1, G6_UMB_Q6LOCAL = 2, G6_UMB_EPOP = 3, &
       G6_UMB_LOGTRIAN = 4
  integer(c_int), parameter :: G6_ORDER_NONE = 0, G6_ORDER_BY_DISTANCE = 1, &
       G6_ORDER_BY_LABEL = 2
  integer, parameter :: G6_NEAR = 151

  interface
     integer(c_int) function g6_init(h, model, n, box_l, rc, taper, voxa, dt, prec, skin) &
          bind(C, name="g6_init")
       import :: c_int, c_int64_t, c_double
       integer(c_int64_t), intent(out) :: h
       integer(c_int), value :: model, n, prec
       real(c_double), value :: box_l, rc, taper, voxa, dt, skin
     end function g6_init

     integer(c_int) function g6_init_box3(h, model, n, box3, rc, taper, voxa3, dt, prec, skin) &
          bind(C, name="g6_init_box3")
       import :: c_int, c_int64_t, c_double
       integer(c_int64_t), intent(out) :: h
       integer(c_int), value :: model, n, prec
       real(c_double), intent(in) :: box3(3), voxa3(3)
       real(c_double), value :: rc, taper, dt, skin
     end function g6_init_box3

     integer(c_int) function g6_finalize(h) bind(C, name="g6_finalize")
       import :: c_int, c_int64_t
       integer(c_int64_t), value :: h
     end function g6_finalize

     integer(c_int) function g6_upload_state(h, com, quat, dx, w) bind(C, name="g6_upload_state")
       import :: c_int, c_int64_t, c_double
       integer(c_int64_t), value :: h
       real(c_double), intent(in) :: com(*), quat(*), dx(*), w(*)
     end function g6_upload_state

     integer(c_int) function g6_download_state(h, com, quat, dx, w) &
          bind(C, name="g6_download_state")
       import :: c_int, c_int64_t, c_double
       integer(c_int64_t), value :: h
       real(c_double), intent(out) :: com(*), quat(*), dx(*), w(*)
     end function g6_download_state

     integer(c_int) function g6_set_fast(h, fast) bind(C, name="g6_set_fast")
       import :: c_int, c_int64_t
       integer(c_int64_t), value :: h
       integer(c_int), value :: fast
     end function g6_set_fast

     integer(c_int) function g6_set_nose(h, active, nose) bind(C, name="g6_set_nose")
       import :: c_int, c_int64_t, c_double
       integer(c_int64_t), value :: h
       integer(c_int), value :: active
       real(c_double), intent(in) :: nose(9)
     end function g6_set_nose

     integer(c_int) function g6_get_nose(h, nose) bind(C, name="g6_get_nose")
       import :: c_int, c_int64_t, c_double
       integer(c_int64_t), value :: h
       real(c_double), intent(out) :: nose(9)
     end function g6_get_nose

     integer(c_int) function g6_vreset(h, temp, seed) bind(C, name="g6_vreset")
       import :: c_int, c_int64_t, c_double, c_int32_t
       integer(c_int64_t), value :: h
       real(c_double), value :: temp
       integer(c_int32_t), value :: seed   ! uint32_t on the C side
     end function g6_vreset

     integer(c_int) function g6_step(h, nsteps) bind(C, name="g6_step")
       import :: c_int, c_int64_t
       integer(c_int64_t), value :: h
       integer(c_int), value :: nsteps
     end function g6_step

     integer(c_int) function g6_get_thermo(h, thermo) bind(C, name="g6_get_thermo")
       import :: c_int, c_int64_t, c_double
       integer(c_int64_t), value :: h
       real(c_double), intent(out) :: thermo(*)
     end function g6_get_thermo

     integer(c_int) function g6_eval_forces(h, force, torque, thermo) &
          bind(C, name="g6_eval_forces")
       import :: c_int, c_int64_t, c_double
       integer(c_int64_t), value :: h
       real(c_double), intent(out) :: force(*), torque(*), thermo(*)
     end function g6_eval_forces

     integer(c_int) function g6_eval_forces_noalloc(h, force, torque, thermo) &
          bind(C, name="g6_eval_forces_noalloc")
       import :: c_int, c_int64_t, c_double
       integer(c_int64_t), value :: h
       real(c_double), intent(out) :: force(*), torque(*), thermo(*)
     end function g6_eval_forces_noalloc

     integer(c_int) function g6_pairlist(h, com, cap, npair0, npair, pair_i, pair_j, ox, oy, oz) &
          bind(C, name="g6_pairlist")
       import :: c_int, c_int64_t, c_int32_t, c_double
       integer(c_int64_t), value :: h, cap
       real(c_double), intent(in) :: com(*)
       integer(c_int64_t), intent(out) :: npair0, npair
       integer(c_int32_t), intent(out) :: pair_i(*), pair_j(*)
       real(c_double), intent(out) :: ox(*), oy(*), oz(*)
     end function g6_pairlist

     integer(c_int) function g6_cells(h, com, cell, ndiv) bind(C, name="g6_cells")
       import :: c_int, c_int64_t, c_int32_t, c_double
       integer(c_int64_t), value :: h
       real(c_double), intent(in) :: com(*)
       integer(c_int32_t), intent(out) :: cell(*), ndiv
     end function g6_cells

     integer(c_int) function g6_cutoff(h, npair, pair_i, pair_j, com, ox, oy, oz, eratio, fratio) &
          bind(C, name="g6_cutoff")
       import :: c_int, c_int64_t, c_int32_t, c_double
       integer(c_int64_t), value :: h, npair
       integer(c_int32_t), intent(in) :: pair_i(*), pair_j(*)
       real(c_double), intent(in) :: com(*), ox(*), oy(*), oz(*)
       real(c_double), intent(out) :: eratio(*), fratio(*)
     end function g6_cutoff

     integer(c_int) function g6_force(h, npair, pair_i, pair_j, ox, oy, oz, eratio, fratio, &
          x, y, z, fx, fy, fz, ep, vrmat) bind(C, name="g6_force")
       import :: c_int, c_int64_t, c_int32_t, c_double
       integer(c_int64_t), value :: h, npair
       integer(c_int32_t), intent(in) :: pair_i(*), pair_j(*)
       real(c_double), intent(in) :: ox(*), oy(*), oz(*), eratio(*), fratio(*), x(*), y(*), z(*)
       real(c_double), intent(inout) :: fx(*), fy(*), fz(*)
       real(c_double), intent(out) :: ep(*), vrmat(9)
     end function g6_force

     integer(c_int) function g6_pair_energy(h, npair, pair_i, pair_j, ox, oy, oz, eratio, &
          x, y, z, ep) bind(C, name="g6_pair_energy")
       import :: c_int, c_int64_t, c_int32_t, c_double
       integer(c_int64_t), value :: h, npair
       integer(c_int32_t), intent(in) :: pair_i(*), pair_j(*)
       real(c_double), intent(in) :: ox(*), oy(*), oz(*), eratio(*), x(*), y(*), z(*)
       real(c_double), intent(out) :: ep(*)
     end function g6_pair_energy

     integer(c_int) function g6_site_positions(h, com, quat1, x, y, z) &
          bind(C, name="g6_site_positions")
       import :: c_int, c_int64_t, c_double
       integer(c_int64_t), value :: h
       real(c_double), intent(in) :: com(*), quat1(*)
       real(c_double), intent(out) :: x(*), y(*), z(*)
     end function g6_site_positions

     integer(c_int) function g6_collect(h, quat1, fx, fy, fz, force, torque) &
          bind(C, name="g6_collect")
       import :: c_int, c_int64_t, c_double
       integer(c_int64_t), value :: h
       real(c_double), intent(in) :: quat1(*), fx(*), fy(*), fz(*)
       real(c_double), intent(out) :: force(*), torque(*)
     end function g6_collect

     integer(c_int) function g6_neighbors(h, nsite, npair, pair_i, pair_j, ox, oy, oz, &
          coord, maxnei, order, radius, nnear, near, dx, dy, dz, dd) &
          bind(C, name="g6_neighbors")
       import :: c_int, c_int64_t, c_int32_t, c_double
       integer(c_int64_t), value :: h, npair
       integer(c_int), value :: nsite, maxnei, order
       real(c_double), value :: radius
       integer(c_int32_t), intent(in) :: pair_i(*), pair_j(*)
       real(c_double), intent(in) :: ox(*), oy(*), oz(*), coord(*)
       integer(c_int32_t), intent(out) :: nnear(*), near(*)
       real(c_double), intent(out) :: dx(*), dy(*), dz(*), dd(*)
     end function g6_neighbors

     integer(c_int) function g6_neighbors_compress(h, nrow, nnear_in, near_in, dx_in, &
          dy_in, dz_in, dd_in, maxnei, order, radius, nnear, near, dx, dy, dz, dd) &
          bind(C, name="g6_neighbors_compress")
       import :: c_int, c_int64_t, c_int32_t, c_double
       integer(c_int64_t), value :: h
       integer(c_int), value :: nrow, maxnei, order
       real(c_double), value :: radius
       integer(c_int32_t), intent(in) :: nnear_in(*), near_in(*)
       real(c_double), intent(in) :: dx_in(*), dy_in(*), dz_in(*), dd_in(*)
       integer(c_int32_t), intent(out) :: nnear(*), near(*)
       real(c_double), intent(out) :: dx(*), dy(*), dz(*), dd(*)
     end function g6_neighbors_compress

     integer(c_int) function g6_neighbors_displacement(h, nsite, nnear_in, near_in, dx_in, &
          dy_in, dz_in, dd_in, order, target, ddx, ddy, ddz, nmodified, index, nnear, near, &
          dx, dy, dz, dd) bind(C, name="g6_neighbors_displacement")
       import :: c_int, c_int64_t, c_int32_t, c_double
       integer(c_int64_t), value :: h
       integer(c_int), value :: nsite, order, target
       real(c_double), value :: ddx, ddy, ddz
       integer(c_int32_t), intent(in) :: nnear_in(*), near_in(*)
       real(c_double), intent(in) :: dx_in(*), dy_in(*), dz_in(*), dd_in(*)
       integer(c_int32_t), intent(out) :: nmodified, index(*), nnear(*), near(*)
       real(c_double), intent(out) :: dx(*), dy(*), dz(*), dd(*)
     end function g6_neighbors_displacement

     integer(c_int) function g6_q6local(h, nsite, npair, pair_i, pair_j, ox, oy, oz, coord, &
          q6, nbond, sumq6, op) bind(C, name="g6_q6local")
       import :: c_int, c_int64_t, c_int32_t, c_double
       integer(c_int64_t), value :: h, npair
       integer(c_int), value :: nsite
       integer(c_int32_t), intent(in) :: pair_i(*), pair_j(*)
       real(c_double), intent(in) :: ox(*), oy(*), oz(*), coord(*)
       real(c_double), intent(out) :: q6(*), sumq6, op
       integer(c_int32_t), intent(out) :: nbond(*)
     end function g6_q6local

     integer(c_int) function g6_triangle(h, nsite, npair, pair_i, pair_j, ox, oy, oz, coord, &
          refhist, logarithmic, hist, op) bind(C, name="g6_triangle")
       import :: c_int, c_int64_t, c_int32_t, c_double
       integer(c_int64_t), value :: h, npair
       integer(c_int), value :: nsite, logarithmic
       integer(c_int32_t), intent(in) :: pair_i(*), pair_j(*)
       real(c_double), intent(in) :: ox(*), oy(*), oz(*), coord(*), refhist(21)
       real(c_double), intent(out) :: hist(21), op
     end function g6_triangle

     ! the order parameters after a trial displacement of molecule target
     ! (0-based), as sq6system_Differentiate / sTriangle_Differentiate
     integer(c_int) function g6_q6local_displacement(h, nsite, npair, pair_i, pair_j, ox, oy, &
          oz, coord, target, dx, dy, dz, q6, nbond, sumq6, op) &
          bind(C, name="g6_q6local_displacement")
       import :: c_int, c_int64_t, c_int32_t, c_double
       integer(c_int64_t), value :: h, npair
       integer(c_int), value :: nsite, target
       integer(c_int32_t), intent(in) :: pair_i(*), pair_j(*)
       real(c_double), intent(in) :: ox(*), oy(*), oz(*), coord(*)
       real(c_double), value :: dx, dy, dz
       real(c_double), intent(out) :: q6(*), sumq6, op
       integer(c_int32_t), intent(out) :: nbond(*)
     end function g6_q6local_displacement

     integer(c_int) function g6_triangle_displacement(h, nsite, npair, pair_i, pair_j, ox, oy, &
          oz, coord, target, dx, dy, dz, refhist, logarithmic, hist, op) &
          bind(C, name="g6_triangle_displacement")
       import :: c_int, c_int64_t, c_int32_t, c_double
       integer(c_int64_t), value :: h, npair
       integer(c_int), value :: nsite, target, logarithmic
       integer(c_int32_t), intent(in) :: pair_i(*), pair_j(*)
       real(c_double), intent(in) :: ox(*), oy(*), oz(*), coord(*), refhist(21)
       real(c_double), value :: dx, dy, dz
       real(c_double), intent(out) :: hist(21), op
     end function g6_triangle_displacement

     integer(c_int) function g6_set_andersen(h, mode, pex, mass, v) &
          bind(C, name="g6_set_andersen")
       import :: c_int, c_int64_t, c_double, c_ptr
       integer(c_int64_t), value :: h
       integer(c_int), value :: mode
       real(c_double), value :: pex, mass
       type(c_ptr), value :: v               ! c_null_ptr, or c_loc of v0..v5
     end function g6_set_andersen

     integer(c_int) function g6_get_andersen(h, state8) bind(C, name="g6_get_andersen")
       import :: c_int, c_int64_t, c_double
       integer(c_int64_t), value :: h
       real(c_double), intent(out) :: state8(8)
     end function g6_get_andersen

     integer(c_int) function g6_get_box(h, box_l) bind(C, name="g6_get_box")
       import :: c_int, c_int64_t, c_double
       integer(c_int64_t), value :: h
       real(c_double), intent(out) :: box_l
     end function g6_get_box

     integer(c_int) function g6_get_box3(h, edges) bind(C, name="g6_get_box3")
       import :: c_int, c_int64_t, c_double
       integer(c_int64_t), value :: h
       real(c_double), intent(out) :: edges(3)
     end function g6_get_box3

     integer(c_int) function g6_lattice_start(n, box_l, seed, com, quat1) &
          bind(C, name="g6_lattice_start")
       import :: c_int, c_int32_t, c_double
       integer(c_int), value :: n
       real(c_double), value :: box_l
       integer(c_int32_t), value :: seed
       real(c_double), intent(out) :: com(*), quat1(*)
     end function g6_lattice_start

     integer(c_int) function g6_step_timed(h, nsteps, t) bind(C, name="g6_step_timed")
       import :: c_int, c_int64_t, c_double
       integer(c_int64_t), value :: h
       integer(c_int), value :: nsteps
       real(c_double), intent(out) :: t(5)
     end function g6_step_timed

     integer(c_int) function g6_time_force(h, nlaunch, ms) bind(C, name="g6_time_force")
       import :: c_int, c_int64_t, c_double
       integer(c_int64_t), value :: h
       integer(c_int), value :: nlaunch
       real(c_double), intent(out) :: ms
     end function g6_time_force

     integer(c_int) function g6_md_kernel(h, info) bind(C, name="g6_md_kernel")
       import :: c_int, c_int64_t, c_int32_t
       integer(c_int64_t), value :: h
       integer(c_int32_t), intent(out) :: info(6)
     end function g6_md_kernel

     integer(c_int) function g6_mc_begin(h, temp, seed, delta, rtratio) bind(C, name="g6_mc_begin")
       import :: c_int, c_int64_t, c_int32_t, c_double
       integer(c_int64_t), value :: h
       real(c_double), value :: temp
       integer(c_int32_t), value :: seed
       real(c_double), intent(in) :: delta(*), rtratio(*)
     end function g6_mc_begin

     integer(c_int) function g6_mc_trials(h, ntrial, naccept) bind(C, name="g6_mc_trials")
       import :: c_int, c_int64_t, c_long_long
       integer(c_int64_t), value :: h
       integer(c_int), value :: ntrial
       integer(c_long_long), intent(out) :: naccept
     end function g6_mc_trials

     integer(c_int) function g6_mc_get(h, com, quat1, delta, rtratio, epsum) bind(C, name="g6_mc_get")
       import :: c_int, c_int64_t, c_double
       integer(c_int64_t), value :: h
       real(c_double), intent(out) :: com(*), quat1(*), delta(*), rtratio(*), epsum
     end function g6_mc_get

     integer(c_int) function g6_mc_pairinfo(h, ep, epsum) bind(C, name="g6_mc_pairinfo")
       import :: c_int, c_int64_t, c_double
       integer(c_int64_t), value :: h
       real(c_double), intent(out) :: ep(*), epsum
     end function g6_mc_pairinfo

     integer(c_int) function g6_mc_energy(h, epsum) bind(C, name="g6_mc_energy")
       import :: c_int, c_int64_t, c_double
       integer(c_int64_t), value :: h
       real(c_double), intent(out) :: epsum
     end function g6_mc_energy

     ! umbrellas of the trial loop (@GEUM): kinds G6_UMB_*, refhist(21)
     integer(c_int) function g6_mc_set_umbrella(h, numbrella, kind, center, coeffi, refhist) &
          bind(C, name="g6_mc_set_umbrella")
       import :: c_int, c_int32_t, c_int64_t, c_double
       integer(c_int64_t), value :: h
       integer(c_int), value :: numbrella
       integer(c_int32_t), intent(in) :: kind(*)
       real(c_double), intent(in) :: center(*), coeffi(*), refhist(*)
     end function g6_mc_set_umbrella

     integer(c_int) function g6_mc_get_umbrella(h, op, center, coeffi) &
          bind(C, name="g6_mc_get_umbrella")
       import :: c_int, c_int64_t, c_double
       integer(c_int64_t), value :: h
       real(c_double), intent(out) :: op(*), center(*), coeffi(*)
     end function g6_mc_get_umbrella

     integer(c_int) function g6_mc_set_temperature(h, temp) bind(C, name="g6_mc_set_temperature")
       import :: c_int, c_int64_t, c_double
       integer(c_int64_t), value :: h
       real(c_double), value :: temp
     end function g6_mc_set_temperature

     ! replica ladders and exchange groups (the replica branch of Mixture.F90)
     integer(c_int) function g6_ladder_create(hl, nsys, hs) bind(C, name="g6_ladder_create")
       import :: c_int, c_int64_t
       integer(c_int64_t), intent(out) :: hl
       integer(c_int), value :: nsys
       integer(c_int64_t), intent(in) :: hs(*)
     end function g6_ladder_create

     integer(c_int) function g6_ladder_free(hl) bind(C, name="g6_ladder_free")
       import :: c_int, c_int64_t
       integer(c_int64_t), value :: hl
     end function g6_ladder_free

     integer(c_int) function g6_ladder_step(hl, nsteps) bind(C, name="g6_ladder_step")
       import :: c_int, c_int64_t
       integer(c_int64_t), value :: hl
       integer(c_int), value :: nsteps
     end function g6_ladder_step

     integer(c_int) function g6_ladder_set_map(hl, mode) bind(C, name="g6_ladder_set_map")
       import :: c_int, c_int64_t
       integer(c_int64_t), value :: hl
       integer(c_int), value :: mode
     end function g6_ladder_set_map

     integer(c_int) function g6_ladder_time_force(hl, nlaunch, ms) &
          bind(C, name="g6_ladder_time_force")
       import :: c_int, c_int64_t, c_double
       integer(c_int64_t), value :: hl
       integer(c_int), value :: nlaunch
       real(c_double), intent(out) :: ms
     end function g6_ladder_time_force

     integer(c_int) function g6_ladder_thermo(hl, thermo) bind(C, name="g6_ladder_thermo")
       import :: c_int, c_int64_t, c_double
       integer(c_int64_t), value :: hl
       real(c_double), intent(out) :: thermo(*)
     end function g6_ladder_thermo

     integer(c_int) function g6_set_temperature(h, temp) bind(C, name="g6_set_temperature")
       import :: c_int, c_int64_t, c_double
       integer(c_int64_t), value :: h
       real(c_double), value :: temp
     end function g6_set_temperature

     integer(c_int) function g6_rx_create(hx, nrep, hs, nbond, a, b, nexchange, seed) &
          bind(C, name="g6_rx_create")
       import :: c_int, c_int64_t, c_int32_t
       integer(c_int64_t), intent(out) :: hx
       integer(c_int), value :: nrep, nbond, nexchange
       integer(c_int64_t), intent(in) :: hs(*)
       integer(c_int32_t), intent(in) :: a(*), b(*)
       integer(c_int32_t), value :: seed
     end function g6_rx_create

     ! mixtures of two rigid components (g6_mixture_* in g6gpu.h)
     integer(c_int) function g6_mixture_create(hm, ha, hb) bind(C, name="g6_mixture_create")
       import :: c_int, c_int64_t
       integer(c_int64_t), intent(out) :: hm
       integer(c_int64_t), value :: ha, hb
     end function g6_mixture_create

     integer(c_int) function g6_mixture_free(hm) bind(C, name="g6_mixture_free")
       import :: c_int, c_int64_t
       integer(c_int64_t), value :: hm
     end function g6_mixture_free

     integer(c_int) function g6_mixture_vreset(hm, temp, seed) bind(C, name="g6_mixture_vreset")
       import :: c_int, c_int64_t, c_double, c_int32_t
       integer(c_int64_t), value :: hm
       real(c_double), value :: temp
       integer(c_int32_t), value :: seed   ! uint32_t on the C side
     end function g6_mixture_vreset

     integer(c_int) function g6_mixture_step(hm, nsteps) bind(C, name="g6_mixture_step")
       import :: c_int, c_int64_t
       integer(c_int64_t), value :: hm
       integer(c_int), value :: nsteps
     end function g6_mixture_step

     integer(c_int) function g6_mixture_thermo(hm, thermo, parts) bind(C, name="g6_mixture_thermo")
       import :: c_int, c_int64_t, c_double
       integer(c_int64_t), value :: hm
       real(c_double), intent(out) :: thermo(*), parts(*)
     end function g6_mixture_thermo

     integer(c_int) function g6_rx_free(hx) bind(C, name="g6_rx_free")
       import :: c_int, c_int64_t
       integer(c_int64_t), value :: hx
     end function g6_rx_free

     integer(c_int) function g6_rx_trials(hx, energy, beta, naccept) bind(C, name="g6_rx_trials")
       import :: c_int, c_int64_t, c_int32_t, c_double
       integer(c_int64_t), value :: hx
       real(c_double), intent(in) :: energy(*)
       real(c_double), intent(inout) :: beta(*)
       integer(c_int32_t), intent(out) :: naccept
     end function g6_rx_trials

     integer(c_int) function g6_rx_trials_umbrella(hx, energy, beta, numbrella, op, center, &
          coeffi, naccept) bind(C, name="g6_rx_trials_umbrella")
       import :: c_int, c_int32_t, c_int64_t, c_double
       integer(c_int64_t), value :: hx
       real(c_double), intent(in) :: energy(*), op(*)
       real(c_double), intent(inout) :: beta(*), center(*), coeffi(*)
       integer(c_int), value :: numbrella
       integer(c_int32_t), intent(out) :: naccept
     end function g6_rx_trials_umbrella

     integer(c_int) function g6_rx_nodes(hx, node_in_charge) bind(C, name="g6_rx_nodes")
       import :: c_int, c_int64_t, c_int32_t
       integer(c_int64_t), value :: hx
       integer(c_int32_t), intent(out) :: node_in_charge(*)
     end function g6_rx_nodes

     ! thermo: c_loc of the records (node order) or c_null_ptr
     integer(c_int) function g6_rx_exchange(hx, thermo, naccept) bind(C, name="g6_rx_exchange")
       import :: c_int, c_int64_t, c_int32_t, c_ptr
       integer(c_int64_t), value :: hx
       type(c_ptr), value :: thermo
       integer(c_int32_t), intent(out) :: naccept
     end function g6_rx_exchange

     ! NPT Monte Carlo with collective moves
     integer(c_int) function g6_npt_begin(h, temp, press_bar, seed, disp_max, rot_max_deg, &
          dlnv_max, p_volume, voxa) bind(C, name="g6_npt_begin")
       import :: c_int, c_int64_t, c_double
       integer(c_int64_t), value :: h, seed
       real(c_double), value :: temp, press_bar, disp_max, rot_max_deg, dlnv_max, p_volume, voxa
     end function g6_npt_begin

     integer(c_int) function g6_npt_run(h, count, hm, k, stats) bind(C, name="g6_npt_run")
       import :: c_int, c_int64_t, c_long_long
       integer(c_int64_t), value :: h, hm
       integer(c_int), value :: count, k
       integer(c_long_long), intent(out) :: stats(*)
     end function g6_npt_run

     integer(c_int) function g6_npt_get(h, com, quat1, box_l, energy) bind(C, name="g6_npt_get")
       import :: c_int, c_int64_t, c_double
       integer(c_int64_t), value :: h
       real(c_double), intent(out) :: com(*), quat1(*), box_l, energy
     end function g6_npt_get

     integer(c_int) function g6_npt_energy(h, energy) bind(C, name="g6_npt_energy")
       import :: c_int, c_int64_t, c_double
       integer(c_int64_t), value :: h
       real(c_double), intent(out) :: energy
     end function g6_npt_energy

     integer(c_int) function g6_npt_features(h, x, e) bind(C, name="g6_npt_features")
       import :: c_int, c_int64_t, c_double, c_float
       integer(c_int64_t), value :: h
       real(c_float), intent(out) :: x(*)
       real(c_double), intent(out) :: e(*)
     end function g6_npt_features

     ! the learned energy surrogate (arrays float32: real(c_float))
     integer(c_int) function g6_mlp_create(hm) bind(C, name="g6_mlp_create")
       import :: c_int, c_int64_t
       integer(c_int64_t), intent(out) :: hm
     end function g6_mlp_create

     integer(c_int) function g6_mlp_free(hm) bind(C, name="g6_mlp_free")
       import :: c_int, c_int64_t
       integer(c_int64_t), value :: hm
     end function g6_mlp_free

     ! params, norm: c_loc of real(c_float) arrays or c_null_ptr
     integer(c_int) function g6_mlp_set_params(hm, params, norm) bind(C, name="g6_mlp_set_params")
       import :: c_int, c_int64_t, c_ptr
       integer(c_int64_t), value :: hm
       type(c_ptr), value :: params, norm
     end function g6_mlp_set_params

     integer(c_int) function g6_mlp_get_params(hm, params, norm) bind(C, name="g6_mlp_get_params")
       import :: c_int, c_int64_t, c_ptr
       integer(c_int64_t), value :: hm
       type(c_ptr), value :: params, norm
     end function g6_mlp_get_params

     integer(c_int) function g6_mlp_forward(hm, m, x, y) bind(C, name="g6_mlp_forward")
       import :: c_int, c_int64_t, c_float
       integer(c_int64_t), value :: hm, m
       real(c_float), intent(in) :: x(*)
       real(c_float), intent(out) :: y(*)
     end function g6_mlp_forward

     integer(c_int) function g6_mlp_loss_grad(hm, m, x, t, loss, grad) &
          bind(C, name="g6_mlp_loss_grad")
       import :: c_int, c_int64_t, c_float
       integer(c_int64_t), value :: hm, m
       real(c_float), intent(in) :: x(*), t(*)
       real(c_float), intent(out) :: loss, grad(*)
     end function g6_mlp_loss_grad

     integer(c_int) function g6_mlp_train(hm, m, x, t, epochs, batch, lr, seed, loss_per_epoch) &
          bind(C, name="g6_mlp_train")
       import :: c_int, c_int64_t, c_float, c_int32_t
       integer(c_int64_t), value :: hm, m
       real(c_float), intent(in) :: x(*), t(*)
       integer(c_int), value :: epochs, batch
       real(c_float), value :: lr
       integer(c_int32_t), value :: seed
       real(c_float), intent(out) :: loss_per_epoch(*)
     end function g6_mlp_train

     integer(c_int) function g6_model_nsite(model) bind(C, name="g6_model_nsite")
       import :: c_int
       integer(c_int), value :: model
     end function g6_model_nsite
  end interface

contains

  ! stops with the reference's behaviour on an error (Error.F90:12-21: die)
  subroutine g6_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    if (rc /= 0) then
       write(0, '(a,a,a,i0)') 'g6gpu: ', what, ' failed with code ', rc
       error stop 1
    end if
  end subroutine g6_check
end module g6gpu
