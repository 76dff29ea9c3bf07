"""ctypes front end of the CPU oracle (oracle/liboracle.so).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.  Every wrapped
function restates a reference routine; the citations live in g6_oracle.h.
"""
import ctypes as C
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
TIP4P, TIP5P, ST2, SPCE, NVDE, OPLSMEOH = 0, 1, 2, 3, 4, 5
MODEL_IDS = {TIP4P: "TIP4P   ", TIP5P: "TIP5P   ", ST2: "ST2_____", SPCE: "SPC_E   ",
             NVDE: "NVDE____", OPLSMEOH: "OPLSMEOH"}
MAXSITE = 7         # G6O_MAXSITE

_lib = None


def build(quiet=True):
    """Compile the oracle (gcc, no FMA contraction) next to its source."""
    import subprocess
    cmd = ["gcc", "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-o", LIB,
           os.path.join(HERE, "g6_oracle.c"), "-lm"]
    subprocess.run(cmd, check=True, capture_output=quiet)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = C.CDLL(LIB)
        _declare(_lib)
    return _lib


_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class Model(C.Structure):
    _fields_ = [("nsite", C.c_int)] + [
        (k, C.c_double * MAXSITE) for k in ("q", "eps", "sig", "bx", "by", "bz")] + [
        ("Ixx", C.c_double), ("Iyy", C.c_double), ("Izz", C.c_double),
        ("mass", C.c_double)]


class MD(C.Structure):
    _fields_ = [("model", C.c_int), ("n", C.c_int), ("fast", C.c_int),
                ("nose_active", C.c_int), ("L", C.c_double),
                ("rc", C.c_double), ("taper", C.c_double), ("req", C.c_double),
                ("dt", C.c_double), ("nose", C.c_double * 9),
                ("com", C.POINTER(C.c_double)), ("dx", C.POINTER(C.c_double)),
                ("w", C.POINTER(C.c_double)), ("quat", C.POINTER(C.c_double)),
                ("ep", C.c_double), ("ekt", C.c_double), ("ekr", C.c_double),
                ("temperature", C.c_double), ("vir", C.c_double * 9),
                ("npair", C.c_long), ("andersen", C.c_int), ("anders", C.c_double * 8),
                ("pressure", C.c_double), ("volume", C.c_double),
                ("andersen_ep", C.c_double), ("andersen_ek", C.c_double)]


NEAR = 151   # G6O_NEAR


class NbTab(C.Structure):
    _fields_ = [("nnear", C.POINTER(C.c_int)), ("near", C.POINTER(C.c_int))] + [
        (k, C.POINTER(C.c_double)) for k in ("dx", "dy", "dz", "dd")]


class Neighbors:
    """A table of neighbour arrays (numpy): nnear (rows,), near/dx/dy/dz/dd
    (rows, NEAR), 0-based labels."""

    def __init__(self, rows):
        self.nnear = np.zeros(rows, np.int32)
        self.near = np.zeros((rows, NEAR), np.int32)
        self.dx, self.dy, self.dz, self.dd = [np.zeros((rows, NEAR)) for _ in range(4)]

    def c(self):
        ip, dp = C.POINTER(C.c_int), C.POINTER(C.c_double)
        return NbTab(self.nnear.ctypes.data_as(ip), self.near.ctypes.data_as(ip),
                     *[a.ctypes.data_as(dp) for a in (self.dx, self.dy, self.dz, self.dd)])

    @classmethod
    def of(cls, nnear, near, dx, dy, dz, dd):
        t = cls(len(nnear))
        t.nnear[:] = nnear
        for name, a in (("near", near), ("dx", dx), ("dy", dy), ("dz", dz), ("dd", dd)):
            getattr(t, name)[:] = a
        return t


def neighbors_all2(nsite, pi, pj, shifts, coord, maxnei, order, radius=None):
    """neighbors_all2 (Neighbor.F90:91-159); returns (Neighbors, overflow)."""
    t = Neighbors(nsite)
    sh = np.ascontiguousarray(shifts, np.float64).reshape(-1, 3)
    err = lib().g6o_neighbors_all2(
        nsite, len(pi), np.ascontiguousarray(pi, np.int32), np.ascontiguousarray(pj, np.int32),
        np.ascontiguousarray(sh[:, 0]), np.ascontiguousarray(sh[:, 1]),
        np.ascontiguousarray(sh[:, 2]), np.ascontiguousarray(coord, np.float64).ravel(),
        maxnei, order, radius if radius else 0.0, t.c())
    return t, bool(err)


def neighbors_compress(tab, maxnei, order, radius=None):
    """neighbor1_compress2 (Neighbor.F90:195-215) of every array of tab."""
    t = Neighbors(len(tab.nnear))
    lib().g6o_neighbors_compress(len(tab.nnear), tab.c(), maxnei, order,
                                 radius if radius else 0.0, t.c())
    return t


def neighbors_displacement(tab, order, target, delta):
    """neighbors_displacement (Neighbor.F90:401-457): (indices, Neighbors)."""
    t = Neighbors(NEAR + 1)
    idx = np.zeros(NEAR + 1, np.int32)
    nm = C.c_int(0)
    lib().g6o_neighbors_displacement(tab.c(), order, target, float(delta[0]), float(delta[1]),
                                     float(delta[2]), C.byref(nm), idx, t.c())
    m = nm.value
    out = Neighbors(m)
    for name in ("nnear", "near", "dx", "dy", "dz", "dd"):
        getattr(out, name)[:] = getattr(t, name)[:m]
    return idx[:m].copy(), out


def q6local(com, pi, pj, shifts):
    """sq6system_prepareall: (q6[n], nbond[n], sumq6, op, overflow)."""
    n = len(com)
    sh = np.ascontiguousarray(shifts, np.float64).reshape(-1, 3)
    q6 = np.zeros(n); nb = np.zeros(n, np.int32)
    s, op = C.c_double(0), C.c_double(0)
    err = lib().g6o_q6local(n, len(pi), np.ascontiguousarray(pi, np.int32),
                            np.ascontiguousarray(pj, np.int32), np.ascontiguousarray(sh[:, 0]),
                            np.ascontiguousarray(sh[:, 1]), np.ascontiguousarray(sh[:, 2]),
                            np.ascontiguousarray(com, np.float64).ravel(), q6, nb, C.byref(s),
                            C.byref(op))
    return q6, nb, s.value, op.value, bool(err)


def triangle(com, pi, pj, shifts, refhist, logarithmic=False):
    """sTriangle_prepareall: (hist[21], op)."""
    n = len(com)
    sh = np.ascontiguousarray(shifts, np.float64).reshape(-1, 3)
    h = np.zeros(21)
    op = C.c_double(0)
    err = lib().g6o_triangle(n, len(pi), np.ascontiguousarray(pi, np.int32),
                             np.ascontiguousarray(pj, np.int32), np.ascontiguousarray(sh[:, 0]),
                             np.ascontiguousarray(sh[:, 1]), np.ascontiguousarray(sh[:, 2]),
                             np.ascontiguousarray(com, np.float64).ravel(),
                             np.ascontiguousarray(refhist, np.float64), int(logarithmic), h,
                             C.byref(op))
    assert not err
    return h, op.value


def _declare(L):
    L.g6o_model_get.argtypes = [C.c_int, C.POINTER(Model)]
    L.g6o_constants.argtypes = [_dp]
    L.g6o_grid_ndiv.argtypes = [C.c_double, C.c_double]
    L.g6o_grid_ndiv.restype = C.c_int
    L.g6o_stencil_homo.argtypes = [_ip, _ip, _ip, _ip, C.POINTER(C.c_int), _ip, _ip, _ip]
    L.g6o_stencil_homo.restype = C.c_int
    L.g6o_address_assign.argtypes = [C.c_int, _dp, C.c_double, _ip, _ip]
    L.g6o_pairlist.argtypes = [C.c_int, _dp, C.c_double, C.c_double, C.c_double,
                               C.c_long, C.POINTER(C.c_long), _ip, _ip, _dp, _dp, _dp]
    L.g6o_pairlist.restype = C.c_long
    L.g6o_pairlist_alltoall.argtypes = [C.c_int, _dp, C.c_double, C.c_double,
                                        C.c_long, _ip, _ip, _dp, _dp, _dp]
    L.g6o_pairlist_alltoall.restype = C.c_long
    L.g6o_cutoff_dimmer.argtypes = [C.c_long, _ip, _ip, _dp, _dp, _dp, _dp,
                                    C.c_double, C.c_double, _dp, _dp]
    L.g6o_cutoff_func.argtypes = [C.c_double, C.c_double, C.c_double,
                                  C.POINTER(C.c_double), C.POINTER(C.c_double)]
    fsig = [C.c_long, _ip, _ip, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
            _dp, _dp, _dp]
    L.g6o_force_tip4p.argtypes = fsig
    L.g6o_force_st2.argtypes = fsig
    L.g6o_force_std.argtypes = [C.c_int] + fsig
    L.g6o_energy_std.argtypes = [C.c_int, C.c_long, _ip, _ip, _dp, _dp, _dp, _dp,
                                 _dp, _dp, _dp, _dp]
    L.g6o_setsites.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
    L.g6o_collect.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
    L.g6o_predict.argtypes = [C.c_int, _dp, _dp, _dp, _dp]
    L.g6o_correct.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, _dp, _dp,
                              _dp, _dp, _dp, _dp]
    L.g6o_ek.argtypes = [C.c_int, C.c_int, C.c_double, _dp, _dp,
                         C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.g6o_nose_predict.argtypes = [_dp]
    L.g6o_nose_correct.argtypes = [_dp, C.c_double]
    L.g6o_md_step.argtypes = [C.POINTER(MD)]
    L.g6o_md_step.restype = C.c_long
    L.g6o_md_hamiltonian.argtypes = [C.POINTER(MD)]
    L.g6o_md_hamiltonian.restype = C.c_double
    L.g6o_mt_seed.argtypes = [C.c_uint32]
    L.g6o_mt_next.restype = C.c_uint32
    L.g6o_random_getnext.restype = C.c_double
    L.g6o_vreset.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_uint32,
                             _dp, _dp]
    L.g6o_mc_pairinfo.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                  _dp, _dp, _dp]
    L.g6o_mc_pairinfo.restype = C.c_double
    L.g6o_mc_trials.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                C.c_double, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                C.POINTER(C.c_long), C.c_int]
    L.g6o_mc_trials.restype = C.c_long
    L.g6o_mc_trials_umbrella.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                         C.c_double, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                         C.POINTER(C.c_long), C.c_int, C.c_int, _ip, _dp, _dp,
                                         _dp, _dp]
    L.g6o_mc_trials_umbrella.restype = C.c_long
    L.g6o_philox10.argtypes = [C.POINTER(C.c_uint32), C.c_uint32, C.c_uint32]
    L.g6o_npt_run.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                              C.c_double, C.c_double, C.c_uint64, C.c_double, C.c_double,
                              C.c_double, C.c_double, C.POINTER(C.c_double), _dp, _dp,
                              C.POINTER(C.c_double), C.POINTER(C.c_longlong), C.c_int,
                              np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS"),
                              _ip]
    L.g6o_q6local.argtypes = [C.c_int, C.c_long, _ip, _ip, _dp, _dp, _dp, _dp, _dp, _ip,
                              C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.g6o_triangle.argtypes = [C.c_int, C.c_long, _ip, _ip, _dp, _dp, _dp, _dp, _dp, C.c_int,
                               _dp, C.POINTER(C.c_double)]
    L.g6o_q6_qlm.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double,
                             C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.g6o_neighbors_all2.argtypes = [C.c_int, C.c_long, _ip, _ip, _dp, _dp, _dp, _dp,
                                     C.c_int, C.c_int, C.c_double, NbTab]
    L.g6o_neighbors_compress.argtypes = [C.c_int, NbTab, C.c_int, C.c_int, C.c_double, NbTab]
    L.g6o_neighbors_displacement.argtypes = [NbTab, C.c_int, C.c_int, C.c_double, C.c_double,
                                             C.c_double, C.POINTER(C.c_int), _ip, NbTab]


def model(mid):
    m = Model()
    lib().g6o_model_get(mid, C.byref(m))
    return m


def constants():
    out = np.zeros(16)
    lib().g6o_constants(out)
    return out


def stencil(ndiv):
    nd = np.array([ndiv] * 3, dtype=np.int32)
    a = [np.zeros(27, np.int32) for _ in range(6)]
    n0 = C.c_int()
    n = lib().g6o_stencil_homo(nd, a[0], a[1], a[2], C.byref(n0), a[3], a[4], a[5])
    return (np.stack(a[:3], 1)[:n], np.stack(a[3:], 1)[:n0.value])


def address_assign(com, L, ndiv):
    n = com.shape[0]
    cell = np.zeros(n, np.int32)
    lib().g6o_address_assign(n, np.ascontiguousarray(com, np.float64), L,
                             np.array([ndiv] * 3, np.int32), cell)
    return cell


def pairlist(com, L, rc, req=None, alltoall=False):
    """Returns (npair0, pi, pj, shifts[np,3]) in the reference's order."""
    com = np.ascontiguousarray(com, np.float64)
    n = com.shape[0]
    cap = max(16, n * 800)
    pi = np.zeros(cap, np.int32); pj = np.zeros(cap, np.int32)
    o = [np.zeros(cap) for _ in range(3)]
    np0 = C.c_long(0)
    if alltoall:
        k = lib().g6o_pairlist_alltoall(n, com, L, rc, cap, pi, pj, *o)
        np0.value = n * (n - 1) // 2
    else:
        k = lib().g6o_pairlist(n, com, L, rc, req if req else rc, cap,
                               C.byref(np0), pi, pj, *o)
    assert k >= 0
    return np0.value, pi[:k].copy(), pj[:k].copy(), np.stack([x[:k] for x in o], 1)


def cutoff(pi, pj, com, shifts, rin, rout):
    k = len(pi)
    er = np.zeros(max(k, 1)); fr = np.zeros(max(k, 1))
    sh = np.ascontiguousarray(shifts.T)
    lib().g6o_cutoff_dimmer(k, pi, pj, np.ascontiguousarray(com), sh[0].copy(),
                            sh[1].copy(), sh[2].copy(), rin, rout, er, fr)
    return er[:k], fr[:k]


def setsites(mid, com, quat):
    """quat: (4, n) array of quat(1) components. Returns sites (n*ns,3), tmat, intra."""
    m = model(mid)
    n = com.shape[0]
    ns = m.nsite
    x = np.zeros(n * ns); y = np.zeros(n * ns); z = np.zeros(n * ns)
    t = np.zeros(9 * n); intra = np.zeros(3 * ns * n)
    q = np.zeros((5, 4, n)); q[0] = quat
    lib().g6o_setsites(mid, n, np.ascontiguousarray(com), q.ravel(), x, y, z, t, intra)
    return np.stack([x, y, z], 1), t, intra


def force(mid, kind, pi, pj, shifts, er, fr, sites):
    """kind: 'std' (StdForce), 'fast' (custom TIP4P/ST2 kernel)."""
    k = len(pi)
    ns = sites.shape[0]
    f = [np.zeros(ns) for _ in range(3)]
    ep = np.zeros(max(k, 1)); vr = np.zeros(9)
    sh = [np.ascontiguousarray(shifts[:, a]) for a in range(3)]
    s = [np.ascontiguousarray(sites[:, a]) for a in range(3)]
    args = [k, pi, pj, *sh, er, fr, *s, *f, ep, vr]
    if kind == "std":
        lib().g6o_force_std(mid, *args)
    elif mid == TIP4P:
        lib().g6o_force_tip4p(*args)
    elif mid == ST2:
        lib().g6o_force_st2(*args)
    else:
        raise ValueError("no custom kernel for this model")
    return np.stack(f, 1), ep[:k], vr.reshape(3, 3)


def energy(mid, pi, pj, shifts, er, sites):
    k = len(pi)
    ep = np.zeros(max(k, 1))
    sh = [np.ascontiguousarray(shifts[:, a]) for a in range(3)]
    s = [np.ascontiguousarray(sites[:, a]) for a in range(3)]
    lib().g6o_energy_std(mid, k, pi, pj, *sh, er, *s, ep)
    return ep[:k]


def collect(mid, forces, tmat, intra):
    n = tmat.size // 9
    F = np.zeros(3 * n); T = np.zeros(3 * n)
    f = [np.ascontiguousarray(forces[:, a]) for a in range(3)]
    lib().g6o_collect(mid, n, *f, tmat, intra, F, T)
    return F.reshape(n, 3), T.reshape(n, 3)


class MDRun:
    """The reference MD step (Mixture.F90 order) on numpy-owned state.

    State arrays: com (n,3); quat (5,4,n); dx (5,3,n); w (5,3,n), exactly the
    reference's Nordsieck layout with dx(1)=v*dt and w in rad/ps.
    """

    def __init__(self, mid, com, quat, L, rc, taper, dt, req=None, fast=False,
                 nose=None):
        n = com.shape[0]
        self.n = n
        self.com = np.ascontiguousarray(com, np.float64).copy()
        self.quat = np.zeros((5, 4, n)); self.quat[0] = quat
        self.dx = np.zeros((5, 3, n)); self.w = np.zeros((5, 3, n))
        self.s = MD()
        self.s.model = mid; self.s.n = n; self.s.fast = int(fast)
        self.s.L = L; self.s.rc = rc; self.s.taper = taper
        self.s.req = req if req else rc; self.s.dt = dt
        if nose is not None:
            # nose = (T0 [K], Q as in @NOS2, dump): q internal = Q*dt (Mixture.F90:1206)
            self.s.nose_active = 1
            for i, v in enumerate([nose[0], nose[1] * dt, nose[2]] + [0.0] * 6):
                self.s.nose[i] = v
        self._bind()

    def _bind(self):
        P = C.POINTER(C.c_double)
        self.s.com = self.com.ctypes.data_as(P)
        self.s.dx = self.dx.ctypes.data_as(P)
        self.s.w = self.w.ctypes.data_as(P)
        self.s.quat = self.quat.ctypes.data_as(P)

    def set_andersen(self, pex, mass):
        """Orthorhombic Andersen barostat from @PCN4 (mode 1, pex atm, mass
        per molecule): internal mass = mass * dof / dt^2 (Mixture.F90:1224-1228),
        v0 = the box volume (Andersen_Initialize)."""
        dof = 6.0 * self.n
        self.s.andersen = 1
        a = [pex, mass * dof / self.s.dt ** 2, (self.s.L * self.s.L) * self.s.L, 0, 0, 0, 0, 0]
        for k, v in enumerate(a):
            self.s.anders[k] = v

    def vreset(self, temp, seed):
        dx1 = np.zeros(3 * self.n); w1 = np.zeros(3 * self.n)
        lib().g6o_vreset(self.s.model, self.n, self.s.dt, temp, seed, dx1, w1)
        self.dx[0] = dx1.reshape(3, self.n)
        self.w[0] = w1.reshape(3, self.n)

    def step(self, nsteps=1):
        for _ in range(nsteps):
            k = lib().g6o_md_step(C.byref(self.s))
            assert k >= 0
        return self

    def hamiltonian(self):
        return lib().g6o_md_hamiltonian(C.byref(self.s))

    def set_temperature(self, temp):
        """After an accepted exchange (Mixture.F90:1630-1653): velocities
        times dsqrt(temp / nose%temp) (Rigid_ScaleVelocity,
        RigidModule.F90:621-633), nose%temp = temp, zeta0..zeta4 = 0."""
        ratio = math.sqrt(temp / self.s.nose[0])
        self.dx[0] *= ratio
        self.w[0] *= ratio
        self.s.nose[0] = temp
        for z in range(3, 8):
            self.s.nose[z] = 0.0

    @property
    def nose(self):
        return np.array(self.s.nose[:])


def mt_stream(seed, count):
    lib().g6o_mt_seed(seed)
    return np.array([lib().g6o_mt_next() for _ in range(count)], dtype=np.uint64)


def getnext_stream(seed, count):
    """The Random_GetNext values (RandomModule.F90:123-130) of an MT19937
    seeded with seed."""
    lib().g6o_mt_seed(seed)
    f = lib().g6o_random_getnext
    return np.array([f() for _ in range(count)], dtype=np.float64)


def mc_pairinfo(mid, com, quat, L, rc, taper):
    """PairInfo matrix (n, n) and its sum; com (n,3), quat (4,n) order 1."""
    com = np.ascontiguousarray(com, np.float64)
    n = com.shape[0]
    ep = np.zeros(n * n)
    e = lib().g6o_mc_pairinfo(mid, n, L, rc, taper, com, np.ascontiguousarray(quat, np.float64),
                              ep)
    return ep.reshape(n, n), e


class MCRun:
    """The mc2 trial loop on the oracle: state (com (n,3), quat (4,n), delta,
    rtratio, PairInfo, epsum) and the Random_GetNext stream of a seed."""

    def __init__(self, mid, com, quat, L, rc, taper, temp, seed, delta, rtratio):
        self.mid, self.L, self.rc, self.taper, self.temp = mid, L, rc, taper, temp
        self.com = np.array(com, np.float64)
        self.quat = np.array(quat, np.float64)
        n = self.n = self.com.shape[0]
        self.delta = np.array(np.broadcast_to(np.asarray(delta, float), (n,)))
        self.rtr = np.array(np.broadcast_to(np.asarray(rtratio, float), (n,)))
        self.ep, e = mc_pairinfo(mid, self.com, self.quat, L, rc, taper)
        self.ep = np.ascontiguousarray(self.ep.ravel())
        self.epsum = np.array([e])
        self.seed = seed
        self.stream = np.zeros(0)
        self.pos = 0
        self.nacc = 0

    def _ensure(self, count):
        if len(self.stream) < count:
            self.stream = getnext_stream(self.seed, max(count, 2 * len(self.stream)))

    # umbrella kinds of @GEUM (mc2util.F90:174-191)
    UMB = {"TRIANGLE": 1, "Q6LOCAL_": 2, "EPOP____": 3, "LOGTRIAN": 4}

    def set_umbrella(self, kinds, center, coeffi, refhist=None):
        """Umbrellas coeffi (op - center)^2 on the order parameters named by
        kinds (TRIANGLE, LOGTRIAN, EPOP____); refhist[21] for the triangle
        kinds.  Each trials() call is then one mc2 loop."""
        self.umb = np.array([self.UMB[k] for k in kinds], np.int32)
        self.center = np.array(center, np.float64)
        self.coeffi = np.array(coeffi, np.float64)
        self.refhist = np.zeros(21) if refhist is None else np.array(refhist, np.float64)
        self.op = np.zeros(len(kinds))

    def trials(self, ntrial):
        self._ensure(self.pos + 14 * ntrial + 2)
        p = C.c_long(self.pos)
        if getattr(self, "umb", None) is not None and len(self.umb):
            acc = lib().g6o_mc_trials_umbrella(
                self.mid, self.n, self.L, self.rc, self.taper, self.temp, self.com, self.quat,
                self.delta, self.rtr, self.ep, self.epsum, self.stream, C.byref(p), ntrial,
                len(self.umb), self.umb, self.center, self.coeffi, self.refhist, self.op)
            assert acc >= 0, "umbrella setup failed"
        else:
            acc = lib().g6o_mc_trials(self.mid, self.n, self.L, self.rc, self.taper, self.temp,
                                      self.com, self.quat, self.delta, self.rtr, self.ep,
                                      self.epsum, self.stream, C.byref(p), ntrial)
        self.pos = p.value
        self.nacc += acc
        return self.nacc


class RxRun:
    """Node 0 of the reference's replica-exchange MD: the exchange loop of
    Mixture.F90:1605-1622 with replica_exchange_trial (Replica.F90:7-73,
    numbrella = 0), its Random_GetNext stream (@SEED) and node_in_charge
    (identity at start, Mixture.F90:1483).  Bonds a, b are 1-based, in
    graph2pair order.  A short loop per exchange: plain Python."""

    I2J, J2K = 10.0, 0.120273  # PhysConst.F90:48, 56

    def __init__(self, nrep, a, b, nexchange, seed):
        self.a, self.b = [int(x) for x in a], [int(x) for x in b]
        self.nexchange, self.seed = nexchange, seed
        self.nic = list(range(1, nrep + 1))
        self.u, self.pos = getnext_stream(seed, 1024), 0

    def _getnext(self):
        if self.pos >= len(self.u):
            self.u = getnext_stream(self.seed, 2 * len(self.u))
        v = float(self.u[self.pos])
        self.pos += 1
        return v

    def trials(self, energy, beta, op=None, center=None, coeffi=None):
        """One exchange loop; with op, center, coeffi ((nrep, numbrella),
        node order) the umbrella-weighted form (Replica.F90:30-49): returns
        (beta, acc), or (beta, center, coeffi, acc) with umbrellas."""
        beta = [float(x) for x in beta]
        umb = op is not None
        if umb:
            op = [list(map(float, r)) for r in op]
            center = [list(map(float, r)) for r in center]
            coeffi = [list(map(float, r)) for r in coeffi]
        U = lambda k, c, o: k * ((o - c) * (o - c))  # Umbrella.F90:11-16
        nb, acc = len(self.a), 0
        for _ in range(self.nexchange):
            # k = random_getnext * nbond + 1 (Fortran truncation); a draw of
            # exactly 1 would index past the last bond: clamped
            k = min(int(self._getnext() * nb + 1), nb)
            j, jj = self.a[k - 1], self.b[k - 1]
            n0, n1 = self.nic[j - 1], self.nic[jj - 1]
            a0, a1 = float(energy[n0 - 1]), float(energy[n1 - 1])
            if umb:
                for u in range(len(op[0])):
                    a0 = a0 + U(coeffi[n0 - 1][u], center[n0 - 1][u], op[n0 - 1][u])
                    a1 = a1 + U(coeffi[n1 - 1][u], center[n1 - 1][u], op[n1 - 1][u])
            acc0 = -(a0 * beta[n0 - 1] + a1 * beta[n1 - 1]) * self.I2J * self.J2K
            a0, a1 = float(energy[n0 - 1]), float(energy[n1 - 1])
            if umb:
                for u in range(len(op[0])):
                    a0 = a0 + U(coeffi[n1 - 1][u], center[n1 - 1][u], op[n0 - 1][u])
                    a1 = a1 + U(coeffi[n0 - 1][u], center[n0 - 1][u], op[n1 - 1][u])
            acc1 = -(a0 * beta[n1 - 1] + a1 * beta[n0 - 1]) * self.I2J * self.J2K
            accept = acc0 < acc1 or self._getnext() < math.exp(acc1 - acc0)
            if accept:
                beta[n0 - 1], beta[n1 - 1] = beta[n1 - 1], beta[n0 - 1]
                if umb:
                    center[n0 - 1], center[n1 - 1] = center[n1 - 1], center[n0 - 1]
                    coeffi[n0 - 1], coeffi[n1 - 1] = coeffi[n1 - 1], coeffi[n0 - 1]
                self.nic[j - 1], self.nic[jj - 1] = n1, n0
                acc += 1
        if umb:
            return np.array(beta), np.array(center), np.array(coeffi), acc
        return np.array(beta), acc


def philox10(ctr, key):
    """Philox4x32-10 of a 4-word counter and a 2-word key."""
    c = (C.c_uint32 * 4)(*[int(v) & 0xFFFFFFFF for v in ctr])
    lib().g6o_philox10(c, int(key[0]) & 0xFFFFFFFF, int(key[1]) & 0xFFFFFFFF)
    return [int(v) for v in c]


class NptRun:
    """The NPT chain of plain moves (volume and block) on the oracle, from a
    configuration com (n,3), quat (4,n) of order 1 in a box L."""

    def __init__(self, mid, com, quat, L, rc, taper, voxa, temp, press_bar, seed, disp_max,
                 rot_max_deg, dlnv_max, p_volume, energy=None):
        self.mid, self.rc, self.taper, self.voxa = mid, rc, taper, voxa
        self.temp, self.press, self.seed = temp, press_bar, seed
        self.disp, self.rot, self.dlnv, self.pvol = disp_max, rot_max_deg, dlnv_max, p_volume
        self.com = np.array(com, np.float64)
        self.quat = np.array(quat, np.float64)
        self.L = float(L)
        if energy is None:
            _, energy = mc_pairinfo(mid, self.com, self.quat, self.L, rc, taper)
        self.U = float(energy)
        self.index = 0
        self.counts = np.zeros(5, np.int64)

    def run(self, nmoves):
        n = len(self.com)
        com = np.ascontiguousarray(self.com)
        quat = np.ascontiguousarray(self.quat)
        Lc, Uc, ic = C.c_double(self.L), C.c_double(self.U), C.c_longlong(self.index)
        acc = np.zeros(nmoves, np.int32)
        lib().g6o_npt_run(self.mid, n, self.rc, self.taper, self.voxa, self.temp, self.press,
                          self.seed, self.disp, self.rot, self.dlnv, self.pvol, C.byref(Lc),
                          com, quat, C.byref(Uc), C.byref(ic), nmoves, self.counts, acc)
        self.com, self.quat, self.L, self.U, self.index = com, quat, Lc.value, Uc.value, ic.value
        return acc
