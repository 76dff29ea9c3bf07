"""ctypes binding of the g6gpu C ABI (include/g6gpu.h).

Mirrors the reference's module entry points on the Python side: the method
names follow the reference routines they stand for (force, pair_energy,
pairlist, collect, site_positions, step).  There is no CPU fallback: if the
HIP library is missing or no device is present, every call raises.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "lib", "libg6gpu.so")
TIP4P, TIP5P, ST2, SPCE, NVDE, OPLSMEOH = 0, 1, 2, 3, 4, 5
FP64, FP32 = 0, 1
THERMO_LEN = 32
ORDER_NONE, ORDER_BY_DISTANCE, ORDER_BY_LABEL = 0, 1, 2
NEAR = 151          # entries of one neighbour array (MAXNEAR + 1)
ERRORS = {-1: "HIP runtime error", -2: "invalid argument",
          -3: "pair capacity exceeded", -4: "no HIP device"}

_lib = None
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i64p = C.POINTER(C.c_int64)
_i64a = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_H = C.c_int64


class G6Error(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise G6Error("libg6gpu.so is not built: run python -m genesis6_amd.build")
        L = C.CDLL(LIB)
        L.g6_init.argtypes = [C.POINTER(_H), C.c_int, C.c_int, C.c_double, C.c_double,
                              C.c_double, C.c_double, C.c_double, C.c_int, C.c_double]
        L.g6_finalize.argtypes = [_H]
        L.g6_upload_state.argtypes = [_H, _dp, _dp, _dp, _dp]
        L.g6_download_state.argtypes = [_H, _dp, _dp, _dp, _dp]
        L.g6_set_nose.argtypes = [_H, C.c_int, _dp]
        L.g6_get_nose.argtypes = [_H, _dp]
        L.g6_vreset.argtypes = [_H, C.c_double, C.c_uint32]
        L.g6_step.argtypes = [_H, C.c_int]
        L.g6_step_timed.argtypes = [_H, C.c_int, _dp]
        L.g6_get_thermo.argtypes = [_H, _dp]
        L.g6_eval_forces.argtypes = [_H, _dp, _dp, _dp]
        L.g6_eval_forces_noalloc.argtypes = [_H, _dp, _dp, _dp]
        L.g6_pairlist.argtypes = [_H, _dp, C.c_int64, _i64p, _i64p, _ip, _ip, _dp, _dp, _dp]
        L.g6_cells.argtypes = [_H, _dp, _ip, _ip]
        L.g6_cutoff.argtypes = [_H, C.c_int64, _ip, _ip, _dp, _dp, _dp, _dp, _dp, _dp]
        L.g6_force.argtypes = [_H, C.c_int64, _ip, _ip, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                               _dp, _dp, _dp, _dp, _dp, _dp]
        L.g6_pair_energy.argtypes = [_H, C.c_int64, _ip, _ip, _dp, _dp, _dp, _dp, _dp, _dp,
                                     _dp, _dp]
        L.g6_site_positions.argtypes = [_H, _dp, _dp, _dp, _dp, _dp]
        L.g6_collect.argtypes = [_H, _dp, _dp, _dp, _dp, _dp, _dp]
        L.g6_q6local.argtypes = [_H, C.c_int, C.c_int64, _ip, _ip, _dp, _dp, _dp, _dp, _dp, _ip,
                                 C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.g6_triangle.argtypes = [_H, C.c_int, C.c_int64, _ip, _ip, _dp, _dp, _dp, _dp, _dp,
                                  C.c_int, _dp, C.POINTER(C.c_double)]
        _d = C.c_double
        L.g6_q6local_displacement.argtypes = [_H, C.c_int, C.c_int64, _ip, _ip, _dp, _dp, _dp,
                                              _dp, C.c_int, _d, _d, _d, _dp, _ip,
                                              C.POINTER(_d), C.POINTER(_d)]
        L.g6_triangle_displacement.argtypes = [_H, C.c_int, C.c_int64, _ip, _ip, _dp, _dp, _dp,
                                               _dp, C.c_int, _d, _d, _d, _dp, C.c_int, _dp,
                                               C.POINTER(_d)]
        L.g6_set_fast.argtypes = [_H, C.c_int]
        L.g6_time_force.argtypes = [_H, C.c_int, C.POINTER(C.c_double)]
        L.g6_ladder_time_force.argtypes = [_H, C.c_int, C.POINTER(C.c_double)]
        L.g6_md_kernel.argtypes = [_H, _ip]
        L.g6_set_andersen.argtypes = [_H, C.c_int, C.c_double, C.c_double, C.c_void_p]
        L.g6_get_andersen.argtypes = [_H, _dp]
        L.g6_get_box.argtypes = [_H, C.POINTER(C.c_double)]
        L.g6_get_box3.argtypes = [_H, _dp]
        L.g6_init_box3.argtypes = [C.POINTER(_H), C.c_int, C.c_int, _dp, C.c_double, C.c_double,
                                   _dp, C.c_double, C.c_int, C.c_double]
        tab = [_ip, _ip, _dp, _dp, _dp, _dp]
        L.g6_neighbors.argtypes = [_H, C.c_int, C.c_int64, _ip, _ip, _dp, _dp, _dp, _dp,
                                   C.c_int, C.c_int, C.c_double] + tab
        L.g6_neighbors_compress.argtypes = [_H, C.c_int] + tab + [
            C.c_int, C.c_int, C.c_double] + tab
        L.g6_neighbors_displacement.argtypes = [_H, C.c_int] + tab + [
            C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, _ip, _ip] + tab
        L.g6_lattice_start.argtypes = [C.c_int, C.c_double, C.c_uint32, _dp, _dp]
        L.g6_model_nsite.argtypes = [C.c_int]
        L.g6_mc_begin.argtypes = [_H, C.c_double, C.c_uint32, _dp, _dp]
        L.g6_mc_trials.argtypes = [_H, C.c_int, C.POINTER(C.c_longlong)]
        L.g6_mc_get.argtypes = [_H, _dp, _dp, _dp, _dp, _dp]
        L.g6_mc_pairinfo.argtypes = [_H, _dp, _dp]
        L.g6_mc_energy.argtypes = [_H, _dp]
        L.g6_mc_set_umbrella.argtypes = [_H, C.c_int, _ip, _dp, _dp, _dp]
        L.g6_mc_get_umbrella.argtypes = [_H, _dp, _dp, _dp]
        L.g6_mc_set_temperature.argtypes = [_H, C.c_double]
        L.g6_ladder_create.argtypes = [C.POINTER(_H), C.c_int, _i64a]
        L.g6_ladder_free.argtypes = [_H]
        L.g6_ladder_step.argtypes = [_H, C.c_int]
        L.g6_ladder_thermo.argtypes = [_H, _dp]
        L.g6_ladder_set_map.argtypes = [_H, C.c_int]
        L.g6_set_temperature.argtypes = [_H, C.c_double]
        L.g6_rx_create.argtypes = [C.POINTER(_H), C.c_int, C.c_void_p, C.c_int, _ip, _ip,
                                   C.c_int, C.c_uint32]
        L.g6_rx_free.argtypes = [_H]
        L.g6_mixture_create.argtypes = [C.POINTER(_H), _H, _H]
        L.g6_mixture_free.argtypes = [_H]
        L.g6_mixture_vreset.argtypes = [_H, C.c_double, C.c_uint32]
        L.g6_mixture_step.argtypes = [_H, C.c_int]
        L.g6_mixture_thermo.argtypes = [_H, _dp, _dp]
        L.g6_rx_trials.argtypes = [_H, _dp, _dp, C.POINTER(C.c_int32)]
        L.g6_rx_trials_umbrella.argtypes = [_H, _dp, _dp, C.c_int, _dp, _dp, _dp,
                                            C.POINTER(C.c_int32)]
        L.g6_rx_nodes.argtypes = [_H, _ip]
        L.g6_rx_exchange.argtypes = [_H, C.c_void_p, C.POINTER(C.c_int32)]
        L.g6_npt_begin.argtypes = [_H, C.c_double, C.c_double, C.c_uint64, C.c_double,
                                   C.c_double, C.c_double, C.c_double, C.c_double]
        L.g6_npt_run.argtypes = [_H, C.c_int, _H, C.c_int, _i64a]
        L.g6_npt_get.argtypes = [_H, _dp, _dp, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.g6_npt_energy.argtypes = [_H, C.POINTER(C.c_double)]
        L.g6_npt_features.argtypes = [_H, _fp, _dp]
        L.g6_mlp_create.argtypes = [C.POINTER(_H)]
        L.g6_mlp_free.argtypes = [_H]
        L.g6_mlp_set_params.argtypes = [_H, C.c_void_p, C.c_void_p]
        L.g6_mlp_get_params.argtypes = [_H, C.c_void_p, C.c_void_p]
        L.g6_mlp_forward.argtypes = [_H, C.c_int64, _fp, _fp]
        L.g6_mlp_loss_grad.argtypes = [_H, C.c_int64, _fp, _fp, C.POINTER(C.c_float), _fp]
        L.g6_mlp_train.argtypes = [_H, C.c_int64, _fp, _fp, C.c_int, C.c_int, C.c_float,
                                   C.c_uint32, _fp]
        _lib = L
    return _lib


def check(rc, what):
    if rc != 0:
        raise G6Error("%s failed: %d (%s)" % (what, rc, ERRORS.get(rc, "reference error code")))


def _c(a, dt=np.float64):
    return np.ascontiguousarray(a, dtype=dt)


def lattice_start(n, L, seed):
    """Synthetic start: returns com (n,3) and quat1 (n,4)."""
    com = np.zeros(3 * n); q = np.zeros(4 * n)
    check(lib().g6_lattice_start(n, L, seed, com, q), "g6_lattice_start")
    return com.reshape(3, n).T.copy(), q.reshape(4, n).T.copy()


class System:
    """One rigid-water system resident on the GPU (a g6_handle)."""

    def __init__(self, model, n, L, rc, taper, dt, voxa=None, prec=FP64, skin=0.0):
        """L: the cubic box edge, or three edges of an orthorhombic box
        (@BOX3); voxa: the requested cell edge (default rc), or three."""
        self.model, self.n, self.L, self.rc, self.taper, self.dt = model, n, L, rc, taper, dt
        self.nsite = lib().g6_model_nsite(model)
        h = _H(0)
        if np.ndim(L) == 0:
            check(lib().g6_init(C.byref(h), model, n, float(L), rc, taper, voxa if voxa else rc,
                                dt, prec, skin), "g6_init")
        else:
            v = np.broadcast_to(np.asarray(voxa if voxa else rc, np.float64), (3,))
            check(lib().g6_init_box3(C.byref(h), model, n, _c(L), rc, taper, _c(v), dt, prec,
                                     skin), "g6_init_box3")
        self.h = h.value

    def close(self):
        if self.h:
            lib().g6_finalize(self.h)
            self.h = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- state ----------------------------------------------------------
    def upload(self, com, quat1, dx=None, w=None, quat=None):
        """com (n,3); quat1 (n,4) order-1 quaternions (higher orders zero,
        as Rigid_ReadNX4A leaves them) unless the full quat (5,4,n) is given;
        dx, w: Nordsieck (5,3,n) or None for zero."""
        n = self.n
        q = np.zeros((5, 4, n)) if quat is None else _c(quat).reshape(5, 4, n)
        if quat is None:
            q[0] = _c(quat1).T
        dxa = np.zeros((5, 3, n)) if dx is None else _c(dx).reshape(5, 3, n)
        wa = np.zeros((5, 3, n)) if w is None else _c(w).reshape(5, 3, n)
        check(lib().g6_upload_state(self.h, _c(com.T).ravel(), _c(q).ravel(),
                                    _c(dxa).ravel(), _c(wa).ravel()), "g6_upload_state")

    def download(self):
        n = self.n
        com = np.zeros(3 * n); q = np.zeros(20 * n); dx = np.zeros(15 * n); w = np.zeros(15 * n)
        check(lib().g6_download_state(self.h, com, q, dx, w), "g6_download_state")
        return (com.reshape(3, n).T.copy(), q.reshape(5, 4, n), dx.reshape(5, 3, n),
                w.reshape(5, 3, n))

    def set_fast(self, fast):
        """@FAST: True for the custom TIP4P/ST2 kernels, False for StdForce."""
        check(lib().g6_set_fast(self.h, 1 if fast else 0), "g6_set_fast")

    def set_nose(self, temp, q_nos2, dump=0.0, zeta=(0, 0, 0, 0, 0), zetasum=0.0,
                 active=True):
        """@NOS2 semantics: q_nos2 is the tag's mass; internally q*dt."""
        z = np.array([temp, q_nos2 * self.dt, dump] + list(zeta) + [zetasum], float)
        check(lib().g6_set_nose(self.h, 1 if active else 0, z), "g6_set_nose")

    def nose(self):
        z = np.zeros(9)
        check(lib().g6_get_nose(self.h, z), "g6_get_nose")
        return z

    def vreset(self, temp, seed):
        check(lib().g6_vreset(self.h, temp, seed), "g6_vreset")

    def step(self, nsteps=1):
        check(lib().g6_step(self.h, nsteps), "g6_step")

    def set_temperature(self, temp):
        """Thermostat retarget of a replica exchange: velocities scaled by
        sqrt(temp / nose temp), nose temp = temp, zeta0..zeta4 = 0."""
        check(lib().g6_set_temperature(self.h, temp), "g6_set_temperature")

    def set_andersen(self, pex, mass, v=None, mode=1):
        """Andersen barostat (@PCN4 mode pex mass; v = v0..v5 of a saved
        state, or None to start one); mode 0 switches it off."""
        vp = None
        if v is not None:
            self._av = np.ascontiguousarray(v, np.float64)
            vp = self._av.ctypes.data_as(C.c_void_p)
        check(lib().g6_set_andersen(self.h, mode, pex, mass, vp), "g6_set_andersen")

    def andersen_state(self):
        """pex, mass (per molecule), v0..v5."""
        st = np.zeros(8)
        check(lib().g6_get_andersen(self.h, st), "g6_get_andersen")
        return st

    def box(self):
        L = C.c_double(0)
        check(lib().g6_get_box(self.h, C.byref(L)), "g6_get_box")
        return L.value

    def box3(self):
        e = np.zeros(3)
        check(lib().g6_get_box3(self.h, e), "g6_get_box3")
        return e

    def md_kernel(self):
        """(tiles, cells per axis, workgroups per tile, waves, slots, capacity
        code that moved the system to the rows) of the MD force kernel
        (g6_md_kernel)."""
        info = np.zeros(6, np.int32)
        check(lib().g6_md_kernel(self.h, info), "g6_md_kernel")
        return tuple(int(x) for x in info)

    def time_force(self, nlaunch=200):
        """Mean ms per launch of the force kernel alone (g6_time_force)."""
        ms = C.c_double(0)
        check(lib().g6_time_force(self.h, nlaunch, C.byref(ms)), "g6_time_force")
        return ms.value

    def step_timed(self, nsteps):
        t = np.zeros(5)
        check(lib().g6_step_timed(self.h, nsteps, t), "g6_step_timed")
        return t

    # ---- Monte Carlo (mc2 trial loop, PairInfo) ----
    def mc_begin(self, temp, seed, delta, rtratio):
        n = self.n
        d = np.ascontiguousarray(np.broadcast_to(np.asarray(delta, float), (n,)))
        r = np.ascontiguousarray(np.broadcast_to(np.asarray(rtratio, float), (n,)))
        check(lib().g6_mc_begin(self.h, temp, seed, d, r), "g6_mc_begin")

    def mc_trials(self, ntrial):
        acc = C.c_longlong(0)
        check(lib().g6_mc_trials(self.h, ntrial, C.byref(acc)), "g6_mc_trials")
        return acc.value

    def mc_state(self):
        n = self.n
        com = np.zeros(3 * n); q = np.zeros(4 * n); d = np.zeros(n); r = np.zeros(n)
        e = np.zeros(1)
        check(lib().g6_mc_get(self.h, com, q, d, r, e), "g6_mc_get")
        return com.reshape(3, n).T.copy(), q.reshape(4, n).T.copy(), d, r, float(e[0])

    def mc_pairinfo(self):
        n = self.n
        ep = np.zeros(n * n); e = np.zeros(1)
        check(lib().g6_mc_pairinfo(self.h, ep, e), "g6_mc_pairinfo")
        return ep.reshape(n, n), float(e[0])

    # umbrella kinds of @GEUM (mc2util.F90:174-191) on the device path
    UMB = {"TRIANGLE": 1, "Q6LOCAL_": 2, "EPOP____": 3, "LOGTRIAN": 4}

    def mc_set_umbrella(self, kinds, center, coeffi, refhist=None):
        """Umbrellas coeffi (op - center)^2 on the order parameters named by
        kinds (TRIANGLE, LOGTRIAN, Q6LOCAL_, EPOP____); each mc_trials call
        is then one mc2 loop."""
        k = np.array([self.UMB[x] for x in kinds], np.int32)
        ref = _c(np.zeros(21) if refhist is None else refhist)
        check(lib().g6_mc_set_umbrella(self.h, len(k), k, _c(center), _c(coeffi), ref),
              "g6_mc_set_umbrella")

    def mc_umbrella(self):
        """(op, center, coeffi) per umbrella: op after the last mc_trials."""
        nb = 4
        op, c, k = np.zeros(nb), np.full(nb, np.nan), np.full(nb, np.nan)
        check(lib().g6_mc_get_umbrella(self.h, op, c, k), "g6_mc_get_umbrella")
        m = int(np.sum(~np.isnan(c)))
        return op[:m], c[:m], k[:m]

    def mc_set_temperature(self, temp):
        check(lib().g6_mc_set_temperature(self.h, temp), "g6_mc_set_temperature")

    def mc_energy(self):
        e = np.zeros(1)
        check(lib().g6_mc_energy(self.h, e), "g6_mc_energy")
        return float(e[0])

    # ---- NPT Monte Carlo with collective moves ----
    def npt_begin(self, temp, press_bar, seed, disp_max=0.15, rot_max_deg=8.0, dlnv_max=0.005,
                  p_volume=0.1, voxa=None):
        check(lib().g6_npt_begin(self.h, temp, press_bar, seed, disp_max, rot_max_deg, dlnv_max,
                                 p_volume, voxa if voxa else self.rc), "g6_npt_begin")

    def npt_run(self, count, mlp=None, K=16):
        """count plain moves (mlp None) or delayed-acceptance rounds of K
        proposals; returns the cumulative counters (dict)."""
        st = np.zeros(10, np.int64)
        check(lib().g6_npt_run(self.h, count, mlp.h if mlp is not None else 0, K, st),
              "g6_npt_run")
        keys = ["proposals", "accepted", "volume_proposals", "volume_accepted",
                "surrogate_passes", "exact_evaluations", "rounds", "void", "next_index",
                "volume_surrogate_passes"]
        return dict(zip(keys, (int(v) for v in st)))

    def npt_state(self):
        n = self.n
        com = np.zeros(3 * n); q = np.zeros(4 * n); L = C.c_double(0); U = C.c_double(0)
        check(lib().g6_npt_get(self.h, com, q, C.byref(L), C.byref(U)), "g6_npt_get")
        return com.reshape(3, n).T.copy(), q.reshape(4, n).copy(), L.value, U.value

    def npt_energy(self):
        U = C.c_double(0)
        check(lib().g6_npt_energy(self.h, C.byref(U)), "g6_npt_energy")
        return U.value

    def npt_features(self):
        n = self.n
        x = np.zeros(32 * n, np.float32); e = np.zeros(n)
        check(lib().g6_npt_features(self.h, x, e), "g6_npt_features")
        return x.reshape(n, 32), e

    def eval_forces(self):
        """Molecule forces (n,3), body torques (n,3) and the thermo record of
        the current state through the MD kernels (state unchanged)."""
        n = self.n
        F = np.zeros(3 * n); T = np.zeros(3 * n); th = np.zeros(THERMO_LEN)
        check(lib().g6_eval_forces(self.h, F, T, th), "g6_eval_forces")
        return F.reshape(3, n).T.copy(), T.reshape(3, n).T.copy(), th

    def eval_forces_noalloc(self):
        """eval_forces through g6_eval_forces_noalloc (no device allocation)."""
        n = self.n
        F = np.zeros(3 * n); T = np.zeros(3 * n); th = np.zeros(THERMO_LEN)
        check(lib().g6_eval_forces_noalloc(self.h, F, T, th), "g6_eval_forces_noalloc")
        return F.reshape(3, n).T.copy(), T.reshape(3, n).T.copy(), th

    def thermo(self):
        t = np.zeros(THERMO_LEN)
        check(lib().g6_get_thermo(self.h, t), "g6_get_thermo")
        return t

    # ---- module-seam entry points ----------------------------------------
    def pairlist(self, com, cap=None):
        n = self.n
        cap = cap or n * 800
        pi = np.zeros(cap, np.int32); pj = np.zeros(cap, np.int32)
        ox = np.zeros(cap); oy = np.zeros(cap); oz = np.zeros(cap)
        np0 = C.c_int64(0); np_ = C.c_int64(0)
        check(lib().g6_pairlist(self.h, _c(com.T).ravel(), cap, C.byref(np0), C.byref(np_),
                                pi, pj, ox, oy, oz), "g6_pairlist")
        k = np_.value
        return np0.value, pi[:k].copy(), pj[:k].copy(), np.stack([ox[:k], oy[:k], oz[:k]], 1)

    def cells(self, com):
        cell = np.zeros(self.n, np.int32); nd = np.zeros(1, np.int32)
        check(lib().g6_cells(self.h, _c(com.T).ravel(), cell, nd), "g6_cells")
        return cell, int(nd[0])

    def cutoff(self, pi, pj, com, shifts):
        k = len(pi)
        er = np.zeros(max(k, 1)); fr = np.zeros(max(k, 1))
        sh = _c(shifts).reshape(-1, 3)
        check(lib().g6_cutoff(self.h, k, _c(pi, np.int32), _c(pj, np.int32), _c(com.T).ravel(),
                              _c(sh[:, 0]), _c(sh[:, 1]), _c(sh[:, 2]), er, fr), "g6_cutoff")
        return er[:k], fr[:k]

    def site_positions(self, com, quat1):
        m = self.n * self.nsite
        x = np.zeros(m); y = np.zeros(m); z = np.zeros(m)
        check(lib().g6_site_positions(self.h, _c(com.T).ravel(), _c(quat1.T).ravel(), x, y, z),
              "g6_site_positions")
        return np.stack([x, y, z], 1)

    def force(self, pi, pj, shifts, er, fr, sites):
        k = len(pi)
        m = sites.shape[0]
        f = [np.zeros(m) for _ in range(3)]
        ep = np.zeros(max(k, 1)); vr = np.zeros(9)
        sh = _c(shifts).reshape(-1, 3)
        check(lib().g6_force(self.h, k, _c(pi, np.int32), _c(pj, np.int32), _c(sh[:, 0]),
                             _c(sh[:, 1]), _c(sh[:, 2]), _c(er), _c(fr), _c(sites[:, 0]),
                             _c(sites[:, 1]), _c(sites[:, 2]), *f, ep, vr), "g6_force")
        return np.stack(f, 1), ep[:k], vr.reshape(3, 3)

    def pair_energy(self, pi, pj, shifts, er, sites):
        k = len(pi)
        ep = np.zeros(max(k, 1))
        sh = _c(shifts).reshape(-1, 3)
        check(lib().g6_pair_energy(self.h, k, _c(pi, np.int32), _c(pj, np.int32), _c(sh[:, 0]),
                                   _c(sh[:, 1]), _c(sh[:, 2]), _c(er), _c(sites[:, 0]),
                                   _c(sites[:, 1]), _c(sites[:, 2]), ep), "g6_pair_energy")
        return ep[:k]

    def collect(self, quat1, forces):
        n = self.n
        F = np.zeros(3 * n); T = np.zeros(3 * n)
        check(lib().g6_collect(self.h, _c(quat1.T).ravel(), _c(forces[:, 0]), _c(forces[:, 1]),
                               _c(forces[:, 2]), F, T), "g6_collect")
        return F.reshape(3, n).T.copy(), T.reshape(3, n).T.copy()

    # ---- neighbour arrays (Neighbor.F90) ----
    def neighbors(self, nsite, pi, pj, shifts, coord, maxnei, order, radius=None):
        """neighbors_all2 over a pair list: a NeighborTable of nsite arrays."""
        k = len(pi)
        out = NeighborTable.empty(nsite)
        sh = _c(shifts).reshape(-1, 3)
        check(lib().g6_neighbors(self.h, nsite, k, _c(pi, np.int32), _c(pj, np.int32),
                                 _c(sh[:, 0]), _c(sh[:, 1]), _c(sh[:, 2]),
                                 _c(np.asarray(coord).T).ravel(), maxnei, order,
                                 radius if radius else 0.0, *out.arrays()), "g6_neighbors")
        return out

    def neighbors_compress(self, tab, maxnei, order, radius=None):
        """neighbor1_compress2 of every array of tab."""
        out = NeighborTable.empty(len(tab.nnear))
        check(lib().g6_neighbors_compress(self.h, len(tab.nnear), *tab.arrays(), maxnei, order,
                                          radius if radius else 0.0, *out.arrays()),
              "g6_neighbors_compress")
        return out

    def neighbors_displacement(self, tab, order, target, delta):
        """neighbors_displacement: (molecule indices, NeighborTable) of the
        arrays that molecule `target` moving by delta changes."""
        out = NeighborTable.empty(NEAR + 1)
        nmod = np.zeros(1, np.int32)
        idx = np.zeros(NEAR + 1, np.int32)
        check(lib().g6_neighbors_displacement(self.h, len(tab.nnear), *tab.arrays(), order,
                                              target, float(delta[0]), float(delta[1]),
                                              float(delta[2]), nmod, idx, *out.arrays()),
              "g6_neighbors_displacement")
        m = int(nmod[0])
        return idx[:m].copy(), out.rows(m)


def _pairs(pi, pj, shifts, coord):
    sh = _c(shifts).reshape(-1, 3)
    return [len(pi), _c(pi, np.int32), _c(pj, np.int32), _c(sh[:, 0]), _c(sh[:, 1]),
            _c(sh[:, 2]), _c(np.asarray(coord).T).ravel()]


def q6local(s, pi, pj, shifts, coord):
    """Q6 local bond order of mc2's umbrella (sq6system_prepareall) on the
    device of System s: (q6[n], nbond[n], sumq6, op)."""
    n = len(coord)
    q6 = np.zeros(n); nb = np.zeros(n, np.int32)
    sm, op = C.c_double(0), C.c_double(0)
    check(lib().g6_q6local(s.h, n, *_pairs(pi, pj, shifts, coord), q6, nb, C.byref(sm),
                           C.byref(op)), "g6_q6local")
    return q6, nb, sm.value, op.value


def triangle(s, pi, pj, shifts, coord, refhist, logarithmic=False):
    """The triangle order parameter (sTriangle_prepareall): (hist[21], op)."""
    n = len(coord)
    h = np.zeros(21)
    op = C.c_double(0)
    check(lib().g6_triangle(s.h, n, *_pairs(pi, pj, shifts, coord), _c(refhist),
                            int(logarithmic), h, C.byref(op)), "g6_triangle")
    return h, op.value


def q6local_displacement(s, pi, pj, shifts, coord, target, delta):
    """sq6system_Differentiate's order parameter after moving molecule
    target (0-based) by delta (3,): (q6[n], nbond[n], sumq6, op)."""
    n = len(coord)
    q6 = np.zeros(n); nb = np.zeros(n, np.int32)
    sm, op = C.c_double(0), C.c_double(0)
    check(lib().g6_q6local_displacement(s.h, n, *_pairs(pi, pj, shifts, coord), int(target),
                                        *[float(v) for v in delta], q6, nb, C.byref(sm),
                                        C.byref(op)), "g6_q6local_displacement")
    return q6, nb, sm.value, op.value


def triangle_displacement(s, pi, pj, shifts, coord, target, delta, refhist, logarithmic=False):
    """sTriangle_Differentiate's order parameter after moving molecule
    target by delta: (hist[21], op)."""
    h = np.zeros(21)
    op = C.c_double(0)
    check(lib().g6_triangle_displacement(s.h, len(coord), *_pairs(pi, pj, shifts, coord),
                                         int(target), *[float(v) for v in delta], _c(refhist),
                                         int(logarithmic), h, C.byref(op)),
          "g6_triangle_displacement")
    return h, op.value


class NeighborTable:
    """A table of neighbour arrays: nnear (rows,), near/dx/dy/dz/dd
    (rows, NEAR); near holds 0-based molecule indices."""

    def __init__(self, nnear, near, dx, dy, dz, dd):
        self.nnear, self.near = nnear, near
        self.dx, self.dy, self.dz, self.dd = dx, dy, dz, dd

    @classmethod
    def empty(cls, rows):
        return cls(np.zeros(rows, np.int32), np.zeros((rows, NEAR), np.int32),
                   *[np.zeros((rows, NEAR)) for _ in range(4)])

    def arrays(self):
        return [_c(self.nnear, np.int32), _c(self.near, np.int32), _c(self.dx), _c(self.dy),
                _c(self.dz), _c(self.dd)]

    def rows(self, m):
        return NeighborTable(*[a[:m].copy() for a in (self.nnear, self.near, self.dx, self.dy,
                                                      self.dz, self.dd)])

    def entries(self, r):
        """(near, dx, dy, dz, dd) of the filled part of array r."""
        c = int(self.nnear[r])
        return (self.near[r, :c], self.dx[r, :c], self.dy[r, :c], self.dz[r, :c],
                self.dd[r, :c])


class Ladder:
    """Several systems of one model and precision stepped together: each MD
    kernel is launched once for all of them (a g6_ladder)."""

    def __init__(self, systems):
        self.systems = list(systems)
        hs = np.array([s.h for s in self.systems], dtype=np.int64)
        h = _H(0)
        check(lib().g6_ladder_create(C.byref(h), len(hs), hs), "g6_ladder_create")
        self.h = h.value

    def step(self, nsteps):
        check(lib().g6_ladder_step(self.h, nsteps), "g6_ladder_step")

    MAP_XCD, MAP_FLAT = 0, 1

    def set_map(self, mode):
        check(lib().g6_ladder_set_map(self.h, mode), "g6_ladder_set_map")

    def time_force(self, nlaunch=100):
        """Mean ms per launch of the ladder's force kernel alone
        (g6_ladder_time_force)."""
        ms = C.c_double()
        check(lib().g6_ladder_time_force(self.h, nlaunch, C.byref(ms)), "g6_ladder_time_force")
        return ms.value

    def thermo(self):
        """Thermo records of the last g6_ladder_step, (nsys, THERMO_LEN)."""
        t = np.zeros(len(self.systems) * THERMO_LEN)
        check(lib().g6_ladder_thermo(self.h, t), "g6_ladder_thermo")
        return t.reshape(len(self.systems), THERMO_LEN)

    def close(self):
        if self.h:
            lib().g6_ladder_free(self.h)
            self.h = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def rxpr_bonds(nnode, links):
    """Directed exchange bonds (a, b), 1-based, in the order graph2pair
    (Graph.F90:34-52) lists the graph that graph_read (Graph.F90:112-127)
    builds from an @RXPR deck of 0-based links (i, j); links=None gives
    graph_perfect (Graph.F90:93-110), the default without @RXPR."""
    nei = [[] for _ in range(nnode)]
    if links is None:
        for i in range(nnode):
            for j in range(i + 1, nnode):
                nei[i].append(j)
                nei[j].append(i)
    else:
        for i, j in links:
            nei[i].append(j)
            nei[j].append(i)
    a = [i + 1 for i in range(nnode) for _ in nei[i]]
    b = [j + 1 for i in range(nnode) for j in nei[i]]
    return np.array(a, np.int32), np.array(b, np.int32)


class Mixture:
    """Two rigid components stepped together (a g6_mixture: the reference's
    @NCMP 2 run, Mixture.F90:1833-2290): the first component's Nose
    thermostat drives both, and the pairs between them are StdForce's."""

    def __init__(self, first, second):
        self.first, self.second = first, second
        h = _H(0)
        check(lib().g6_mixture_create(C.byref(h), first.h, second.h), "g6_mixture_create")
        self.h = h.value

    def vreset(self, temp, seed):
        check(lib().g6_mixture_vreset(self.h, float(temp), int(seed)), "g6_mixture_vreset")

    def step(self, nsteps):
        check(lib().g6_mixture_step(self.h, int(nsteps)), "g6_mixture_step")

    def thermo(self):
        """(thermo record of the whole system, parts): parts = ep inside the
        first, inside the second, between them; ek of each component."""
        t = np.zeros(THERMO_LEN)
        p = np.zeros(5)
        check(lib().g6_mixture_thermo(self.h, t, p), "g6_mixture_thermo")
        return t, p

    def close(self):
        if self.h:
            lib().g6_mixture_free(self.h)
            self.h = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ReplicaExchange:
    """An exchange group (a g6_rx): the node-0 state of the reference's
    replica exchange, i.e. its MT19937 stream and node_in_charge."""

    def __init__(self, nrep, bonds, nexchange, seed, systems=None):
        self.nrep = nrep
        self.systems = list(systems) if systems is not None else None
        a, b = (_c(x, np.int32) for x in bonds)
        hs = None
        if self.systems is not None:
            self._hs = np.array([s.h for s in self.systems], dtype=np.int64)
            hs = self._hs.ctypes.data_as(C.c_void_p)
        h = _H(0)
        check(lib().g6_rx_create(C.byref(h), nrep, hs, len(a), a, b, nexchange, seed),
              "g6_rx_create")
        self.h = h.value

    def trials(self, energy, beta):
        """One exchange loop on given energies and betas (node order);
        returns the new betas and the number of accepted swaps."""
        beta = _c(beta).copy()
        acc = C.c_int32(0)
        check(lib().g6_rx_trials(self.h, _c(energy), beta, C.byref(acc)), "g6_rx_trials")
        return beta, acc.value

    def trials_umbrella(self, energy, beta, op, center, coeffi):
        """The exchange loop weighted by umbrellas: op, center, coeffi are
        (nrep, numbrella) in node order; returns the new betas, centers,
        coefficients and the number of accepted swaps."""
        beta = _c(beta).copy()
        op = _c(op)
        nb = op.shape[1] if op.ndim == 2 else 0
        c, k = _c(center).copy(), _c(coeffi).copy()
        acc = C.c_int32(0)
        check(lib().g6_rx_trials_umbrella(self.h, _c(energy), beta, nb, op.ravel(), c.ravel(),
                                          k.ravel(), C.byref(acc)), "g6_rx_trials_umbrella")
        return beta, c, k, acc.value

    def nodes(self):
        nic = np.zeros(self.nrep, np.int32)
        check(lib().g6_rx_nodes(self.h, nic), "g6_rx_nodes")
        return nic

    def exchange(self, thermo=None):
        """The exchange of Mixture.F90:1570-1653 between the group's
        systems; thermo (nrep, THERMO_LEN) from Ladder.thermo or None."""
        ptr = None
        if thermo is not None:
            thermo = _c(thermo)
            ptr = thermo.ctypes.data_as(C.c_void_p)
        acc = C.c_int32(0)
        check(lib().g6_rx_exchange(self.h, ptr, C.byref(acc)), "g6_rx_exchange")
        return acc.value

    def close(self):
        if self.h:
            lib().g6_rx_free(self.h)
            self.h = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


MLP_NIN, MLP_NPARAM, MLP_NNORM = 32, 6337, 66


def _vp(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Mlp:
    """The per-molecule energy network on the GPU (a g6_mlp)."""

    def __init__(self, params=None, norm=None):
        h = _H(0)
        check(lib().g6_mlp_create(C.byref(h)), "g6_mlp_create")
        self.h = h.value
        if params is not None or norm is not None:
            self.set(params, norm)

    def set(self, params=None, norm=None):
        p = None if params is None else _c(params, np.float32)
        n = None if norm is None else _c(norm, np.float32)
        if p is not None and p.size != MLP_NPARAM or n is not None and n.size != MLP_NNORM:
            raise G6Error("g6_mlp_set_params: wrong sizes")
        check(lib().g6_mlp_set_params(self.h, _vp(p), _vp(n)), "g6_mlp_set_params")

    def get(self):
        p = np.zeros(MLP_NPARAM, np.float32); n = np.zeros(MLP_NNORM, np.float32)
        check(lib().g6_mlp_get_params(self.h, _vp(p), _vp(n)), "g6_mlp_get_params")
        return p, n

    def forward(self, x):
        x = _c(x, np.float32).reshape(-1, MLP_NIN)
        y = np.zeros(len(x), np.float32)
        check(lib().g6_mlp_forward(self.h, len(x), x, y), "g6_mlp_forward")
        return y

    def loss_grad(self, x, t):
        x = _c(x, np.float32).reshape(-1, MLP_NIN)
        g = np.zeros(MLP_NPARAM, np.float32); loss = C.c_float(0)
        check(lib().g6_mlp_loss_grad(self.h, len(x), x, _c(t, np.float32), C.byref(loss), g),
              "g6_mlp_loss_grad")
        return loss.value, g

    def train(self, x, t, epochs, batch, lr, seed):
        x = _c(x, np.float32).reshape(-1, MLP_NIN)
        out = np.zeros(epochs, np.float32)
        check(lib().g6_mlp_train(self.h, len(x), x, _c(t, np.float32), epochs, batch, lr, seed,
                                 out), "g6_mlp_train")
        return out

    def close(self):
        if self.h:
            lib().g6_mlp_free(self.h)
            self.h = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def mlp_init_params(seed):
    """Glorot-uniform weights and zero biases for a new network."""
    rng = np.random.default_rng(seed)
    a1 = np.sqrt(6.0 / (MLP_NIN + 64)); a2 = np.sqrt(6.0 / 128.0); a3 = np.sqrt(6.0 / 65.0)
    return np.concatenate([rng.uniform(-a1, a1, 64 * MLP_NIN), np.zeros(64),
                           rng.uniform(-a2, a2, 64 * 64), np.zeros(64),
                           rng.uniform(-a3, a3, 64), [0.0]]).astype(np.float32)
