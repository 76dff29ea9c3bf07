"""The cell-tile force kernel (each molecule pair evaluated once, force and
torque sums in 64-bit fixed point) against the FP64 oracle, at the sizes it
runs: every model on 1024 molecules and 4096 TIP4P with the tiles forced on,
and the 65,536-molecule TIP5P system (the workload the tiles serve by
default) at full size.  Bars as for the rows: FP32 pairs <= 1e-5 relative RMS
force and torque error, interacting pair counts exact; plus determinism
(repeat runs, graph replay and eager launches bitwise equal) and agreement
with the rows of both directions along a trajectory."""
import os
from contextlib import contextmanager

import numpy as np
import pytest

from genesis6_amd import g6
from oracle import pyoracle as O
from tests.golden.configs import lattice_start

pytestmark = pytest.mark.gpu
RC, TAPER, DT = 9.0, 0.5, 0.002


def relrms(a, b):
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


@contextmanager
def env(**kw):
    old = {k: os.environ.get(k) for k in kw}
    os.environ.update({k: str(v) for k, v in kw.items()})
    try:
        yield
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def oracle_forces(mid, kind, com, q, L):
    np0, pi, pj, sh = O.pairlist(com, L, RC, RC)
    er, fr = O.cutoff(pi, pj, com, sh, RC - TAPER, RC)
    sites, tm, intra = O.setsites(mid, com, q.T.copy())
    Fs, ep, vr = O.force(mid, kind, pi, pj, sh, er, fr, sites)
    Fm, Tm = O.collect(mid, Fs, tm, intra)
    return Fm, Tm, ep, vr, len(pi)


def check_forces(s, Fm, Tm, ep, vr, npair, tol=1e-5):
    F, T, th = s.eval_forces()
    assert relrms(F, Fm) <= tol and relrms(T, Tm) <= tol, (relrms(F, Fm), relrms(T, Tm))
    assert int(th[14]) == npair
    assert abs(th[0] - ep.sum()) <= tol * np.abs(ep).sum()
    assert relrms(th[4:13].reshape(3, 3), vr) <= tol
    # Newton's third law: the pair forces cancel in the total up to the FP32
    # rounding of the i-side quad sums
    assert np.linalg.norm(F.sum(0)) <= 1e-5 * np.linalg.norm(F)


# (model, oracle kernel, @FAST)
MODELS = [(g6.TIP4P, "fast", True), (g6.TIP5P, "std", True), (g6.SPCE, "std", True),
          (g6.NVDE, "std", True), (g6.ST2, "fast", True), (g6.ST2, "std", False),
          (g6.OPLSMEOH, "std", True)]


@pytest.mark.parametrize("mid,kind,fast", MODELS)
@pytest.mark.parametrize("skin", [0.0, 0.8])
def test_tile_forces_every_model(mid, kind, fast, skin):
    # water at 1.0 g/cm3; methanol at 0.79 g/cm3 (M 32.04)
    n, L = 1024, (41.008 if mid == g6.OPLSMEOH else 31.289)
    print("tiles model=%d fast=%d skin=%g" % (mid, fast, skin), flush=True)
    com, q = lattice_start(n, L, 20261202 + 7)
    ref = oracle_forces(mid, kind, com, q, L)
    with env(G6_TILE_MIN=0):
        s = g6.System(mid, n, L, RC, TAPER, DT, prec=g6.FP32, skin=skin)
    if not fast:
        s.set_fast(False)
    assert s.md_kernel()[0] == 1 and s.md_kernel()[1] == int(L / (RC + skin))
    s.upload(com, q)
    check_forces(s, *ref)
    s.close()


@pytest.mark.parametrize("skin", [0.0, 1.0])
def test_tile_forces_c2_size(skin):
    """4096 TIP4P (the C2 workload) through the tiles, forced on."""
    n, L = 4096, 49.669
    print("tiles c2 skin=%g" % skin, flush=True)
    com, q = lattice_start(n, L, 20261202 + 2)
    ref = oracle_forces(g6.TIP4P, "fast", com, q, L)
    with env(G6_TILE_MIN=0):
        s = g6.System(g6.TIP4P, n, L, RC, TAPER, DT, prec=g6.FP32, skin=skin)
    assert s.md_kernel()[0] == 1
    s.upload(com, q)
    check_forces(s, *ref)
    s.close()


def test_c3_tip5p_65536_forces_and_steps():
    """65,536 TIP5P at 1.0 g/cm3 (L = 125.157 A), rc 9 A: FP32 through the
    tiles (the default at this size) and FP64 through the rows against the
    oracle's forces, then three Nose NVT steps of each against the oracle's
    MD steps from the same Maxwell-Boltzmann start."""
    n, L, T = 65536, 125.157, 275.0
    com, q = lattice_start(n, L, 20261202 + 3)
    print("c3 oracle forces", flush=True)
    ref = oracle_forces(g6.TIP5P, "std", com, q, L)
    s32 = g6.System(g6.TIP5P, n, L, RC, TAPER, DT, prec=g6.FP32, skin=1.0)
    assert s32.md_kernel()[0] == 1
    s32.upload(com, q)
    print("c3 fp32 tiles", flush=True)
    check_forces(s32, *ref)
    s64 = g6.System(g6.TIP5P, n, L, RC, TAPER, DT, prec=g6.FP64)
    assert s64.md_kernel()[0] == 0
    s64.upload(com, q)
    print("c3 fp64 rows", flush=True)
    check_forces(s64, *ref, tol=1e-12)
    print("c3 oracle steps", flush=True)
    md = O.MDRun(O.TIP5P, com, q.T.copy(), L, RC, TAPER, DT, fast=False, nose=(T, T * 2500.0, 0.0))
    md.vreset(T, 31)
    temps = []
    for k in range(3):
        md.step(1)
        temps.append(md.s.temperature)
    for s, tol in ((s64, 1e-10), (s32, 1e-6)):
        s.upload(com, q)
        s.set_nose(T, T * 2500.0)
        s.vreset(T, 31)
        got = []
        for k in range(3):
            s.step(1)
            got.append(s.thermo()[3])
        assert np.allclose(got, temps, rtol=tol, atol=0), (got, temps)
    cg, _, _, _ = s64.download()
    assert np.abs(cg - md.com).max() < 1e-9
    s32.close()
    s64.close()


def _run(steps, graph=True, tiles=True, seed=5):
    n, L, T = 1024, 31.289, 300.0
    with env(G6_TILE_MIN=0, G6_TILES=1 if tiles else 0, G6_GRAPH_STEPS=50 if graph else 0):
        s = g6.System(g6.TIP4P, n, L, RC, TAPER, DT, prec=g6.FP32, skin=1.0)
        assert s.md_kernel()[0] == (1 if tiles else 0)
        com, q = lattice_start(n, L, seed)
        s.upload(com, q)
        s.set_nose(T, T * 2500.0)
        s.vreset(T, seed)
        ths = []
        for _ in range(steps // 50):
            s.step(50)
            ths.append(s.thermo().copy())
        st = s.download()
    s.close()
    return st, np.array(ths)


def test_tile_steps_are_deterministic():
    """Integer sums make the forces independent of the order of the LDS
    atomics; every other sum has a fixed order.  Two graph runs and an eager
    run of 100 steps (rows rebuilt along the way) end in the same bits."""
    print("tiles determinism", flush=True)
    a, tha = _run(100)
    b, thb = _run(100)
    c, thc = _run(100, graph=False)
    for x, y, z in zip(a, b, c):
        assert np.array_equal(x, y) and np.array_equal(x, z)
    assert np.array_equal(tha, thb) and np.array_equal(tha, thc)


def test_tile_trajectory_tracks_the_rows():
    """200 Nose NVT steps of 1024 TIP4P: tiles and rows (both FP32 pairs,
    summed in different orders, so the trajectories part slowly as chaotic
    ones do) agree on temperature, potential energy, the conserved H,
    pressure and the interacting pair count."""
    print("tiles vs rows trajectory", flush=True)
    _, tt = _run(200, tiles=True)
    _, tr = _run(200, tiles=False)
    for col, tol in ((3, 1e-3), (0, 1e-3), (13, 1e-4), (14, 1e-3)):
        assert np.allclose(tt[:, col], tr[:, col], rtol=tol, atol=0), (col, tt[:, col], tr[:, col])
    assert np.allclose(tt[:, 25], tr[:, 25], rtol=0, atol=0.02 * np.abs(tr[:, 25]).max() + 10.0)


def test_capacity_overflow_restores_the_state():
    """A g6_step call that runs out of a capacity leaves the state it was
    given.  On the rows the call reports the error; a tile system reruns the
    call on the rows (and stays on them), ending where a rows-only run ends,
    bit for bit."""
    n, L, T = 1024, 31.289, 300.0
    com, q = lattice_start(n, L, 9)

    def start(**kw):
        with env(**kw):
            s = g6.System(g6.TIP4P, n, L, RC, TAPER, DT, prec=g6.FP32, skin=1.0)
        s.upload(com, q)
        s.set_nose(T, T * 2500.0)
        s.vreset(T, 9)
        return s

    print("capacity: rows with 8 residents per cell", flush=True)
    s = start(G6_TILES=0, G6_CELL_CAP=8)
    before = s.download()
    with pytest.raises(g6.G6Error):
        s.step(60)
    after = s.download()
    for x, y in zip(before, after):
        assert np.array_equal(x, y)
    s.close()
    print("capacity: tiles with 64 slots", flush=True)
    a = start(G6_TILE_MIN=0, G6_TILE_SLOTS=64)
    assert a.md_kernel()[0] == 1
    b = start(G6_TILES=0)
    with env(G6_TILE_RETRY=100):
        a.step(60)
        # code 4: the neighbourhood ran past the slots
        assert a.md_kernel()[0] == 0 and a.md_kernel()[5] == 4
        a.step(60)
        # step 120 >= 100: the call tries the tiles again, runs out of the
        # slots again and reruns on the rows
        a.step(60)
        assert a.md_kernel()[0] == 0
    b.step(180)
    for x, y in zip(a.download(), b.download()):
        assert np.array_equal(x, y)
    # the rerun at step 120 starts from fresh rows: only the listed-pair
    # count [15] depends on when rows were last rebuilt
    ta, tb = a.thermo(), b.thermo()
    assert np.array_equal(np.delete(ta, 15), np.delete(tb, 15))
    a.close()
    b.close()


def test_c3_tiles_return_after_the_lattice_start():
    """65,536 TIP5P with a 1.8 A skin (11 cells of 11.4 A per axis, 49
    molecules per cell): the lattice start fills some tile neighbourhoods past
    the slots for its first steps, the system runs on the rows, and it takes
    the tiles again once the lattice has melted.  The tile forces then agree
    with the rows' on the same state."""
    n, Lb, T = 65536, 125.157, 275.0
    com, q = lattice_start(n, Lb, 20261205)
    s = g6.System(g6.TIP5P, n, Lb, RC, TAPER, DT, prec=g6.FP32, skin=1.8)
    assert s.md_kernel()[0] == 1
    s.upload(com, q)
    s.set_nose(T, T * 2500.0)
    s.vreset(T, 20261205)
    print("c3 skin 1.8: lattice start", flush=True)
    s.step(10)
    first = s.md_kernel()
    print("after 10 steps", first, flush=True)
    s.step(1000)
    s.step(200)
    back = s.md_kernel()
    print("after 1210 steps", back, s.thermo()[[0, 3, 14]], flush=True)
    assert back[0] == 1
    com2, q2, dx2, w2 = s.download()
    F, Tq, th = s.eval_forces()
    with env(G6_TILES=0):
        r = g6.System(g6.TIP5P, n, Lb, RC, TAPER, DT, prec=g6.FP32, skin=1.8)
    r.upload(com2, None, dx2, w2, quat=q2)
    Fr, Tr, thr = r.eval_forces()
    assert relrms(F, Fr) <= 2e-5 and relrms(Tq, Tr) <= 2e-5
    assert int(th[14]) == int(thr[14])
    s.close()
    r.close()
