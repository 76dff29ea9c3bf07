"""The umbrella order parameters of mc2 on the GPU (g6_q6local, g6_triangle)
against the reference's own values (tests/golden/order_*.npz) and the
oracle."""
import math
import time

import numpy as np
import pytest

from genesis6_amd import g6
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def system(n, L):
    return g6.System(g6.TIP4P, n, L, 9.0, 0.5, 0.002)


@pytest.mark.parametrize("name", ["order_tip4p_216", "order_tip4p_1024"])
def test_trial_displacements_match_reference(golden, name):
    """g6_q6local_displacement / g6_triangle_displacement against what the
    reference's sq6system_Differentiate and sTriangle_Differentiate return for
    eight trial moves, and against the oracle's full evaluation of the
    displaced centres; in the one move where two bonded neighbours enter the
    target's shell together the reference's incremental Q6 counts their bond
    twice, and the GPU gives the full evaluation there."""
    from tests.test_order_oracle import double_counted_cases
    g = golden(name)
    deviating = set(double_counted_cases(g))
    s = system(len(g["com"]), float(g["L"]))
    try:
        for c, t in enumerate(g["disp_target"]):
            print(name, "move", c, "target", int(t), flush=True)
            d = g["disp_delta"][c]
            _, _, _, op = g6.q6local_displacement(s, g["pi"], g["pj"], g["shift"], g["com"], t, d)
            _, o = g6.triangle_displacement(s, g["pi"], g["pj"], g["shift"], g["com"], t, d,
                                            g["refhist"])
            com = g["com"].copy()
            com[t] += d
            _, _, _, op_o, _ = po.q6local(com, g["pi"], g["pj"], g["shift"])
            assert abs(op - op_o) <= 1e-14 * abs(op_o)
            if c not in deviating:
                assert abs(op - g["disp_op_q6"][c]) <= 1e-13 * abs(g["disp_op_q6"][c])
            assert abs(o - g["disp_op_tri"][c]) <= 1e-14 * abs(g["disp_op_tri"][c])
    finally:
        s.close()


@pytest.mark.parametrize("name", ["order_tip4p_216", "order_tip4p_1024"])
def test_order_parameters_match_reference(golden, name):
    g = golden(name)
    print(name, "pairs", len(g["pi"]), flush=True)
    s = system(len(g["com"]), float(g["L"]))
    try:
        q6, nb, sm, op = g6.q6local(s, g["pi"], g["pj"], g["shift"], g["com"])
        h, o = g6.triangle(s, g["pi"], g["pj"], g["shift"], g["com"], g["refhist"])
        hl, ol = g6.triangle(s, g["pi"], g["pj"], g["shift"], g["com"], g["refhist"], True)
    finally:
        s.close()
    np.testing.assert_array_equal(nb, g["nbond"])
    err = np.abs(q6 - g["q6"]).max() / np.abs(g["q6"]).max()
    print("  q6 max rel %.2e, bit-equal %.3f, op %.17g vs %.17g" % (err, (q6 == g["q6"]).mean(),
                                                                 op, g["op_q6"]))
    assert err <= 1e-14
    assert abs(op - g["op_q6"]) <= 1e-14 * abs(g["op_q6"])
    np.testing.assert_array_equal(h, g["hist_tri"])
    np.testing.assert_array_equal(hl, g["hist_logtri"])
    assert abs(o - g["op_tri"]) <= 1e-14 * abs(g["op_tri"])
    assert abs(ol - g["op_logtri"]) <= 1e-14 * abs(g["op_logtri"])


def test_order_parameters_4096_match_oracle():
    n, L = 4096, 49.7
    com, _ = g6.lattice_start(n, L, 9)
    com = np.mod(com + np.random.default_rng(4).normal(0, 0.25, com.shape), L)
    _, pi, pj, sh = po.pairlist(com, L, 6.5)
    refhist = np.linspace(1.0, 3.0, 21)
    s = system(n, L)
    try:
        t0 = time.perf_counter()
        q6, nb, sm, op = g6.q6local(s, pi, pj, sh, com)
        t1 = time.perf_counter()
        h, o = g6.triangle(s, pi, pj, sh, com, refhist)
        t2 = time.perf_counter()
    finally:
        s.close()
    oq6, onb, osm, oop, err = po.q6local(com, pi, pj, sh)
    t3 = time.perf_counter()
    oh, oo = po.triangle(com, pi, pj, sh, refhist)
    t4 = time.perf_counter()
    print("4096: q6 %.1f ms, triangle %.1f ms (with copies); oracle %.1f, %.1f ms"
          % (1e3 * (t1 - t0), 1e3 * (t2 - t1), 1e3 * (t3 - t2), 1e3 * (t4 - t3)))
    assert not err
    np.testing.assert_array_equal(nb, onb)
    assert np.allclose(q6, oq6, rtol=1e-14, atol=0, equal_nan=True)
    np.testing.assert_array_equal(h, oh)
    assert abs(o - oo) <= 1e-14 * abs(oo)


def test_more_than_maxbond_bonds_is_reported():
    """Eleven neighbours within 3 A of one molecule: the reference's bond
    arrays (MAXBOND = 10) would overflow; the call reports it."""
    n = 12
    com = np.zeros((n, 3)) + 30.0
    for k in range(1, n):
        t = 2 * math.pi * k / 11
        com[k] += 2.5 * np.array([math.cos(t), math.sin(t), 0.3 * (-1) ** k])
    _, pi, pj, sh = po.pairlist(com, 60.0, 6.5)
    s = system(n, 60.0)
    try:
        with pytest.raises(g6.G6Error):
            g6.q6local(s, pi, pj, sh, com)
    finally:
        s.close()
