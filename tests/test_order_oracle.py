"""The umbrella order parameters of mc2 on the oracle (g6o_q6local,
g6o_triangle) against the reference's own (tests/golden/order_*.npz,
written by tests/golden/make_order.py from the unmodified Q6Local and
Triangle modules), and against what the mathematics fixes."""
import ctypes as C
import math

import numpy as np
import pytest

from oracle import pyoracle as po


@pytest.mark.parametrize("name", ["order_tip4p_216", "order_tip4p_1024"])
def test_q6_and_triangle_match_reference(golden, name):
    g = golden(name)
    q6, nb, s, op, err = po.q6local(g["com"], g["pi"], g["pj"], g["shift"])
    assert not err
    np.testing.assert_array_equal(nb, g["nbond"])
    np.testing.assert_array_equal(q6, g["q6"])
    assert s == g["sumq6"] and op == g["op_q6"]
    h, o = po.triangle(g["com"], g["pi"], g["pj"], g["shift"], g["refhist"])
    np.testing.assert_array_equal(h, g["hist_tri"])
    assert o == g["op_tri"]
    h, o = po.triangle(g["com"], g["pi"], g["pj"], g["shift"], g["refhist"], True)
    np.testing.assert_array_equal(h, g["hist_logtri"])
    assert o == g["op_logtri"]


@pytest.mark.parametrize("name", ["order_tip4p_216", "order_tip4p_1024"])
def test_trial_displacement_matches_the_incremental_reference(golden, name):
    """sq6system_Differentiate / sTriangle_Differentiate: the reference
    updates its tables incrementally for a trial displacement of one
    molecule; the full evaluation on the displaced centres (same pair list,
    same image shifts) gives the same order parameters (Q6 within rounding
    of the displaced separations, the triangle histogram exactly)."""
    g = golden(name)
    deviating = set(double_counted_cases(g))
    assert deviating == ({5} if name == "order_tip4p_216" else set())
    for c, t in enumerate(g["disp_target"]):
        com = g["com"].copy()
        com[t] += g["disp_delta"][c]
        _, _, _, op, err = po.q6local(com, g["pi"], g["pj"], g["shift"])
        assert not err
        if c in deviating:
            # the reference's incremental sum counts the bond of the two
            # entering neighbours twice: it differs from the full evaluation
            assert abs(op - g["disp_op_q6"][c]) > 1e-6 * abs(op)
        else:
            assert abs(op - g["disp_op_q6"][c]) <= 1e-13 * abs(g["disp_op_q6"][c]), c
        _, o = po.triangle(com, g["pi"], g["pj"], g["shift"], g["refhist"])
        assert abs(o - g["disp_op_tri"][c]) <= 1e-13 * abs(g["disp_op_tri"][c]), c
        # the displacement changes the order parameter
        assert op != g["op_q6"] and o != g["op_tri"]


def double_counted_cases(g):
    """The trial displacements in which two molecules within 3 A of each
    other both enter (or both leave) the target's 5 A shell: there
    sq6system_Differentiate's shell loops (Q6Local.F90:379-411) add (or
    remove) their bond from both sides."""
    L, com = float(g["L"]), g["com"]

    def dist(a, b):
        d = a - b
        d -= L * np.round(d / L)
        return np.sqrt((d * d).sum(-1))
    out = []
    for c, t in enumerate(g["disp_target"]):
        new = com.copy()
        new[t] += g["disp_delta"][c]
        old5 = set(np.where(dist(com, com[t]) < 5.0)[0])
        new5 = set(np.where(dist(new, new[t]) < 5.0)[0])
        for shell in (sorted(new5 - old5), sorted(old5 - new5)):
            if any(dist(com[a], com[b]) < 3.0 for a in shell for b in shell if a < b):
                out.append(c)
                break
    return out


def qlm(m, x, y, z):
    a, b = C.c_double(), C.c_double()
    po.lib().g6o_q6_qlm(m, x, y, z, C.byref(a), C.byref(b))
    return complex(a.value, b.value)


def test_qlm_closed_forms():
    """For x, y exact in single precision the reference's rounding of
    cmplx(x, y) changes nothing, and qlm is the textbook Y_6m: the m = 0 and
    m = 6 closed forms, the sum rule sum_m |Y_6m|^2 = 13 / (4 pi), and
    Y_6,-m = (-1)^m conj(Y_6m)."""
    x, y = 0.5, -0.25
    z = math.sqrt(1.0 - x * x - y * y)
    st2 = x * x + y * y
    y60 = math.sqrt(13 / math.pi) / 32 * (231 * z ** 6 - 315 * z ** 4 + 105 * z ** 2 - 5)
    assert abs(qlm(0, x, y, z) - y60) < 1e-14
    e6 = complex(x, y) ** 6 / st2 ** 3
    y66 = math.sqrt(3003 / math.pi) / 64 * st2 ** 3 * e6
    assert abs(qlm(6, x, y, z) - y66) < 1e-14
    total = sum(abs(qlm(m, x, y, z)) ** 2 for m in range(-6, 7))
    assert abs(total - 13 / (4 * math.pi)) < 1e-13
    for m in range(1, 7):
        assert abs(qlm(-m, x, y, z) - (-1) ** m * qlm(m, x, y, z).conjugate()) < 1e-15


def test_q6_of_an_octahedron_and_the_tetrahedral_angle():
    """A molecule with six neighbours at 2.9 A on a generically turned
    octahedron (no bond along z, where the reference's qlm divides by zero):
    its twelve directed bonds give q6 = the Steinhardt Q6 of the simple-cubic
    shell, 1/(2 sqrt 2), to the single-precision rounding of cmplx(x, y).
    Four neighbours at the tetrahedral angle put their six angles in the bin
    of cos = -1/3."""
    c, s_ = math.cos(0.7), math.sin(0.7)
    R = np.array([[c, -s_, 0], [s_, c, 0], [0, 0, 1]]) @ np.array(
        [[1, 0, 0], [0, math.cos(0.4), -math.sin(0.4)], [0, math.sin(0.4), math.cos(0.4)]])
    v = np.array([(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]) @ R.T
    com = np.vstack([[0, 0, 0], 2.9 * v]) + 30.0
    _, pi, pj, sh = po.pairlist(com, 60.0, 6.5)
    q6, nb, s, op, err = po.q6local(com, pi, pj, sh)
    assert not err and nb[0] == 12
    assert abs(q6[0] - 1 / (2 * math.sqrt(2))) < 1e-6
    v = np.array([(1, 1, 1), (1, -1, -1), (-1, 1, -1), (-1, -1, 1)]) / math.sqrt(3) * 2.8
    com = np.vstack([[0, 0, 0], v]) + 30.0
    _, pi, pj, sh = po.pairlist(com, 60.0, 5.5)
    h, o = po.triangle(com, pi, pj, sh, np.ones(21))
    assert h[int(round((-1 / 3 + 1) / 0.1))] >= 6        # bin of cos = -1/3 (0-based)
