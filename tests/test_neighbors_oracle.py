"""The neighbour-array oracle (oracle/g6_oracle.c: neighbors_all2,
neighbor1_compress2, neighbors_displacement of Neighbor.F90) against the
reference's own arrays (tests/golden/neighbors_*, written by
tests/golden/make_neighbors.py from the unmodified reference routines), and
against an independent construction: each array is the stable sort of its
candidates by (key, insertion order), cut to the limit."""
import json
import os

import numpy as np
import pytest

from oracle import pyoracle as po
from tests.golden.make_neighbors import digest

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def vec(t):
    return np.stack([t.dx, t.dy, t.dz, t.dd], -1)


def same(t, nn, near, v):
    assert np.array_equal(t.nnear, nn)
    for i in range(len(nn)):
        c = nn[i]
        assert np.array_equal(t.near[i, :c], near[i, :c]), i
        assert np.array_equal(vec(t)[i, :c].view(np.int64), v[i, :c].view(np.int64)), i


def run_case(coord, pi, pj, sh, c):
    maxnei, order, radius, cmax, cord, crad, target = (int(c[0]), int(c[1]), c[2], int(c[3]),
                                                       int(c[4]), c[5], int(c[6]))
    t, over = po.neighbors_all2(len(coord), pi, pj, sh, coord, maxnei, order, radius)
    assert not over
    ct = po.neighbors_compress(t, cmax, cord, crad)
    idx, dt = po.neighbors_displacement(t, order, target, c[7:10])
    return t, ct, idx, dt


def test_lattice_matches_reference_arrays(golden):
    g = golden("neighbors_lattice_64")
    for ci, c in enumerate(g["cases"]):
        t, ct, idx, dt = run_case(g["coord"], g["pi"], g["pj"], g["shift"], c)
        same(t, g["all2_nnear_%d" % ci], g["all2_near_%d" % ci], g["all2_vec_%d" % ci])
        same(ct, g["comp_nnear_%d" % ci], g["comp_near_%d" % ci], g["comp_vec_%d" % ci])
        assert np.array_equal(idx, g["disp_index_%d" % ci])
        same(dt, g["disp_nnear_%d" % ci], g["disp_near_%d" % ci], g["disp_vec_%d" % ci])


def test_tip4p_216_matches_reference_digests(golden):
    with open(os.path.join(GOLD, "neighbors_tip4p_216.json")) as f:
        ref = json.load(f)
    g = golden(ref["source"])
    for c, d in zip(ref["cases"], ref["digests"]):
        t, ct, idx, dt = run_case(g["com"], g["pi"], g["pj"], g["shift"], np.array(c))
        assert digest(t.nnear, t.near, vec(t)) == d["all2"]
        assert int(t.nnear.sum()) == d["all2_total"]
        assert digest(ct.nnear, ct.near, vec(ct)) == d["comp"]
        assert idx.tolist() == d["disp_index"]
        assert digest(dt.nnear, dt.near, vec(dt)) == d["disp"]


def lexsort_all2(n, pi, pj, sh, coord, maxnei, order, radius):
    """Each molecule's candidates stably sorted by (key, pair index)."""
    cand = [[] for _ in range(n)]
    for k in range(len(pi)):
        i, j = pi[k], pj[k]
        d = coord[i] - coord[j]
        d = d - sh[k]
        dd = d[0] * d[0] + d[1] * d[1] + d[2] * d[2]
        if radius and radius * radius < dd:
            continue
        cand[i].append((k, j, -d, dd))
        cand[j].append((k, i, d, dd))
    out = []
    for m in range(n):
        c = cand[m]
        if order == po_ORDER["dist"]:
            c = sorted(c, key=lambda e: (e[3], e[0]))
        elif order == po_ORDER["label"]:
            c = sorted(c, key=lambda e: (e[1], e[0]))
        if order != po_ORDER["none"] and maxnei > 0:
            c = c[:maxnei]
        out.append(c)
    return out


po_ORDER = {"none": 0, "dist": 1, "label": 2}


@pytest.mark.parametrize("order", [0, 1, 2])
@pytest.mark.parametrize("maxnei", [0, 1, 3, 7])
def test_insertion_equals_stable_sort_with_ties(order, maxnei):
    """Integer coordinates make many equal distances; the insertion order
    must decide them, and a limit must keep exactly the first maxnei."""
    rng = np.random.default_rng(7 + order + 10 * maxnei)
    n = 40
    coord = rng.integers(0, 4, (n, 3)).astype(np.float64)
    pairs = [(i, j) for i in range(n) for j in range(i + 1, n) if rng.random() < 0.3]
    pairs += [(i, j) for (i, j) in pairs[:20]]          # repeated pairs (images)
    rng.shuffle(pairs)
    pi = np.array([p[0] for p in pairs], np.int32)
    pj = np.array([p[1] for p in pairs], np.int32)
    sh = rng.integers(-1, 2, (len(pairs), 3)).astype(np.float64) * 4.0
    for radius in (None, 4.5):
        t, over = po.neighbors_all2(n, pi, pj, sh, coord, maxnei, order, radius)
        assert not over
        ref = lexsort_all2(n, pi, pj, sh, coord, maxnei, order, radius)
        for m in range(n):
            c = ref[m]
            assert t.nnear[m] == len(c)
            assert t.near[m, :len(c)].tolist() == [e[1] for e in c]
            got = vec(t)[m, :len(c)]
            want = np.array([[*e[2], e[3]] for e in c]).reshape(-1, 4)
            assert np.array_equal(got, want)
        # compress under another order: the stable sort of the array's
        # entries by (key, position)
        for cord, cmax, crad in ((1, 2, None), (2, 0, 3.0), (0, 0, 2.5)):
            ct = po.neighbors_compress(t, cmax, cord, crad)
            for m in range(n):
                e = [(p, t.near[m, p], t.dd[m, p]) for p in range(t.nnear[m])
                     if not (crad and crad * crad < t.dd[m, p])]
                if cord == 1:
                    e.sort(key=lambda x: (x[2], x[0]))
                elif cord == 2:
                    e.sort(key=lambda x: (x[1], x[0]))
                if cord != 0 and cmax > 0:
                    e = e[:cmax]
                assert ct.nnear[m] == len(e)
                assert ct.near[m, :len(e)].tolist() == [x[1] for x in e]


def test_displacement_moves_both_sides():
    """Moving the target by delta and rebuilding from scratch gives the same
    vectors as neighbors_displacement, for the arrays it returns."""
    rng = np.random.default_rng(3)
    n = 30
    coord = rng.random((n, 3)) * 8.0
    pairs = [(i, j) for i in range(n) for j in range(i + 1, n)]
    pi = np.array([p[0] for p in pairs], np.int32)
    pj = np.array([p[1] for p in pairs], np.int32)
    sh = np.zeros((len(pairs), 3))
    target, delta = 4, np.array([0.25, -0.5, 0.125])  # exact binary steps
    for order in (0, 1, 2):
        t, _ = po.neighbors_all2(n, pi, pj, sh, coord, 0, order)
        idx, dt = po.neighbors_displacement(t, order, target, delta)
        assert idx[0] == target and len(idx) == n      # all pairs: everyone lists it
        moved = coord.copy()
        moved[target] += delta
        t2, _ = po.neighbors_all2(n, pi, pj, sh, moved, 0, order)
        for r, m in enumerate(idx):
            c = dt.nnear[r]
            assert c == t2.nnear[m]
            a = sorted(zip(dt.near[r, :c], dt.dx[r, :c], dt.dy[r, :c], dt.dz[r, :c]))
            b = sorted(zip(t2.near[m, :c], t2.dx[m, :c], t2.dy[m, :c], t2.dz[m, :c]))
            assert np.allclose(np.array(a), np.array(b), atol=1e-12)
            if order == 1:
                assert np.all(np.diff(dt.dd[r, :c]) >= 0)
            if order == 2:
                assert np.all(np.diff(dt.near[r, :c]) >= 0)


def test_overflow_is_reported():
    n = 160
    coord = np.zeros((n, 3))
    pi = np.zeros(n - 1, np.int32)
    pj = np.arange(1, n, dtype=np.int32)
    _, over = po.neighbors_all2(n, pi, pj, np.zeros((n - 1, 3)), coord, 0, 0)
    assert over
    t, over = po.neighbors_all2(n, pi, pj, np.zeros((n - 1, 3)), coord, 5, 1)
    assert not over and t.nnear[0] == 5 and t.near[0, :5].tolist() == [1, 2, 3, 4, 5]
