"""Neighbour arrays on the GPU (g6_neighbors, g6_neighbors_compress,
g6_neighbors_displacement) against the reference's own arrays
(tests/golden/neighbors_*) and against the oracle, bit for bit."""
import json
import os
import time

import numpy as np
import pytest

from genesis6_amd import g6
from oracle import pyoracle as po
from tests.golden.make_neighbors import digest

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def vec(t):
    return np.stack([t.dx, t.dy, t.dz, t.dd], -1)


def same(t, nn, near, v):
    assert np.array_equal(t.nnear, nn)
    for i in range(len(nn)):
        c = nn[i]
        assert np.array_equal(t.near[i, :c], near[i, :c]), i
        assert np.array_equal(vec(t)[i, :c].view(np.int64), v[i, :c].view(np.int64)), i


def system(n):
    """A handle for the device and stream (water density)."""
    L = max((n / 0.0334) ** (1.0 / 3.0), 20.0)
    return g6.System(g6.TIP4P, n, L, 9.0, 0.5, 0.002)


def gpu_case(s, coord, pi, pj, sh, c):
    maxnei, order, radius, cmax, cord, crad, target = (int(c[0]), int(c[1]), c[2], int(c[3]),
                                                       int(c[4]), c[5], int(c[6]))
    t = s.neighbors(len(coord), pi, pj, sh, coord, maxnei, order, radius)
    ct = s.neighbors_compress(t, cmax, cord, crad)
    idx, dt = s.neighbors_displacement(t, order, target, c[7:10])
    return t, ct, idx, dt


def test_lattice_matches_reference_arrays(golden):
    g = golden("neighbors_lattice_64")
    print("lattice-64 cases", len(g["cases"]), flush=True)
    s = system(64)
    try:
        for ci, c in enumerate(g["cases"]):
            t, ct, idx, dt = gpu_case(s, g["coord"], g["pi"], g["pj"], g["shift"], c)
            same(t, g["all2_nnear_%d" % ci], g["all2_near_%d" % ci], g["all2_vec_%d" % ci])
            same(ct, g["comp_nnear_%d" % ci], g["comp_near_%d" % ci], g["comp_vec_%d" % ci])
            assert np.array_equal(idx, g["disp_index_%d" % ci])
            same(dt, g["disp_nnear_%d" % ci], g["disp_near_%d" % ci], g["disp_vec_%d" % ci])
    finally:
        s.close()


def test_tip4p_216_matches_reference_digests(golden):
    with open(os.path.join(GOLD, "neighbors_tip4p_216.json")) as f:
        ref = json.load(f)
    g = golden(ref["source"])
    print("tip4p-216 pairs", len(g["pi"]), flush=True)
    s = system(216)
    try:
        for c, d in zip(ref["cases"], ref["digests"]):
            t, ct, idx, dt = gpu_case(s, g["com"], g["pi"], g["pj"], g["shift"], np.array(c))
            assert digest(t.nnear, t.near, vec(t)) == d["all2"]
            assert digest(ct.nnear, ct.near, vec(ct)) == d["comp"]
            assert idx.tolist() == d["disp_index"]
            assert digest(dt.nnear, dt.near, vec(dt)) == d["disp"]
    finally:
        s.close()


@pytest.mark.parametrize("order,maxnei,radius", [(2, 0, 6.0), (1, 4, None), (1, 0, 5.0),
                                                 (0, 0, 3.5), (2, 12, None)])
def test_4096_matches_oracle(order, maxnei, radius):
    n, L = 4096, 49.7
    com, _ = g6.lattice_start(n, L, 5)
    com = com + np.random.default_rng(1).normal(0, 0.3, com.shape)
    com = np.mod(com, L)
    _, pi, pj, sh = po.pairlist(com, L, 6.5)
    print("4096: pairs", len(pi), "order", order, "maxnei", maxnei, "radius", radius, flush=True)
    s = system(n)
    try:
        t0 = time.perf_counter()
        t = s.neighbors(n, pi, pj, sh, com, maxnei, order, radius)
        t1 = time.perf_counter()
        o, over = po.neighbors_all2(n, pi, pj, sh, com, maxnei, order, radius)
        t2 = time.perf_counter()
        print("  g6_neighbors %.2f ms (with copies), oracle %.2f ms"
              % (1e3 * (t1 - t0), 1e3 * (t2 - t1)))
        assert not over
        same(t, o.nnear, o.near, vec(o))
        for cmax, cord, crad in ((4, 1, None), (0, 2, 3.0), (0, 0, 4.0)):
            ct = s.neighbors_compress(t, cmax, cord, crad)
            oc = po.neighbors_compress(o, cmax, cord, crad)
            same(ct, oc.nnear, oc.near, vec(oc))
        for target in (0, 17, 4095):
            delta = (0.1, -0.2, 0.05)
            idx, dt = s.neighbors_displacement(t, order, target, delta)
            oidx, od = po.neighbors_displacement(o, order, target, delta)
            assert np.array_equal(idx, oidx)
            same(dt, od.nnear, od.near, vec(od))
    finally:
        s.close()


def test_edges():
    s = system(160)
    try:
        n = 160
        coord = np.zeros((n, 3))
        coord[:, 0] = np.arange(n) * 0.01
        pi = np.zeros(n - 1, np.int32)
        pj = np.arange(1, n, dtype=np.int32)
        sh = np.zeros((n - 1, 3))
        # 159 neighbours of molecule 0: more than one array holds
        with pytest.raises(RuntimeError):
            s.neighbors(n, pi, pj, sh, coord, 0, g6.ORDER_NONE)
        t = s.neighbors(n, pi, pj, sh, coord, 5, g6.ORDER_BY_DISTANCE)
        assert t.nnear[0] == 5 and t.near[0, :5].tolist() == [1, 2, 3, 4, 5]
        assert np.all(t.nnear[1:] == 1)
        # no pairs
        e = s.neighbors(n, pi[:0], pj[:0], sh[:0], coord, 0, g6.ORDER_BY_LABEL)
        assert np.all(e.nnear == 0)
        # label out of range
        with pytest.raises(RuntimeError):
            s.neighbors(10, pi, pj, sh, coord, 0, g6.ORDER_NONE)
        # a target with no neighbours: only its own array
        idx, d = s.neighbors_displacement(e, g6.ORDER_BY_LABEL, 3, (1.0, 0.0, 0.0))
        assert idx.tolist() == [3] and d.nnear.tolist() == [0]
    finally:
        s.close()
