"""Umbrella MC on the GPU (g6_mc_set_umbrella): the trial loop with the
triangle and energy umbrellas of mc2 (mc2.F90:684-771) against the
reference program's own output (tests/golden/mc2_umbrella_tip4p_216.npz,
from the unmodified mc2) and against the oracle's umbrella trial loop."""
import numpy as np
import pytest

from genesis6_amd import g6
from genesis6_amd.g6 import lattice_start
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", [0, 1, 2, 3])
def test_gpu_umbrella_mc_equals_reference_mc2(golden, case):
    """Same seed and start as mc2: every decision is mc2's (step sizes and
    rotation ratios bit for bit), and after each sweep of 2N trials the
    energy and each order parameter equal what mc2 prints.  Cases 2 and 3
    carry the Q6 umbrella, whose sums the reference updates through every
    trial of a sweep, rejected ones included."""
    r = golden("mc2_umbrella_tip4p_216")
    g = golden(str(r["start_%d" % case]))
    seed, temp, sweeps = r["case_%d" % case]
    n = len(g["com"])
    print("umbrella mc2 case", case, flush=True)
    s = g6.System(g6.TIP4P, n, float(g["L"]), float(g["rc"]), float(g["taper"]), 0.002)
    s.upload(g["com"], g["quat"])
    s.mc_begin(float(temp), int(seed), 0.1, 1.0)
    kinds = [str(k) for k in r["kinds_%d" % case]]
    s.mc_set_umbrella(kinds, r["center_%d" % case], r["coeffi_%d" % case], r["refhist"])
    want = r["mc2o_%d" % case]
    for sweep in range(int(sweeps)):
        s.mc_trials(2 * n)
        op, center, coeffi = s.mc_umbrella()
        e = s.mc_state()[4]
        assert abs(e - want[sweep, 1]) <= 1e-11 * abs(want[sweep, 1]), sweep
        for k in range(len(kinds)):
            assert abs(op[k] - want[sweep, 3 + 3 * k]) <= 1e-11 * abs(want[sweep, 3 + 3 * k]), \
                (sweep, k, op[k], want[sweep, 3 + 3 * k])
        assert np.array_equal(center, r["center_%d" % case])
        assert np.array_equal(coeffi, r["coeffi_%d" % case])
    c, q, d, rt, _ = s.mc_state()
    assert np.array_equal(d, r["delta_%d" % case])
    assert np.array_equal(rt, r["rtratio_%d" % case])
    assert np.allclose(c, r["com_%d" % case], rtol=0, atol=2e-9 * 20)
    assert np.allclose(q, r["quat_%d" % case], rtol=0, atol=1e-9)
    s.close()


def test_gpu_umbrella_mc_equals_oracle_three_umbrellas(golden):
    """512 molecules, three umbrellas at once (TRIANGLE, LOGTRIAN and
    EPOP____), four loops of 2N trials: the oracle's decisions (accept
    counts, configuration and step sizes bit for bit) and order parameters."""
    r = golden("mc2_umbrella_tip4p_216")
    n, L = 512, 25.088
    com, q = lattice_start(n, L, 77)
    kinds = ["TRIANGLE", "LOGTRIAN", "EPOP____"]
    center, coeffi = [0.4, -1.5, -1.0e5], [4.0e3, 5.0e2, 5.0e-7]
    print("umbrella mc 512, three umbrellas", flush=True)
    s = g6.System(g6.TIP4P, n, L, 9.0, 0.5, 0.002)
    s.upload(com, q)
    s.mc_begin(270.0, 2024, 0.25, 1.0)
    s.mc_set_umbrella(kinds, center, coeffi, r["refhist"])
    o = O.MCRun(O.TIP4P, com, q.T.copy(), L, 9.0, 0.5, 270.0, 2024, 0.25, 1.0)
    o.set_umbrella(kinds, center, coeffi, r["refhist"])
    for loop in range(4):
        acc = s.mc_trials(2 * n)
        assert acc == o.trials(2 * n), loop
        op = s.mc_umbrella()[0]
        assert np.allclose(op, o.op, rtol=1e-11, atol=0), (loop, op, o.op)
    c, qq, d, rt, e = s.mc_state()
    np.testing.assert_array_equal(c, o.com)
    np.testing.assert_array_equal(qq.T, o.quat)
    np.testing.assert_array_equal(d, o.delta)
    np.testing.assert_array_equal(rt, o.rtr)
    assert abs(e - o.epsum[0]) <= 1e-10 * abs(e)
    # the umbrellas removed: plain trials again, still the oracle's
    s.mc_set_umbrella([], [], [])
    o.umb = None
    assert s.mc_trials(500) == o.trials(500)
    np.testing.assert_array_equal(s.mc_state()[2], o.delta)
    s.close()


def test_gpu_umbrella_capacity_error_keeps_the_state():
    """More than 64 molecules within 5 A of one molecule: the call stops
    before any trial (G6_ECAPACITY) and the configuration, step sizes and
    random stream are as before: the next plain call equals a fresh run."""
    rng = np.random.default_rng(5)
    n, L = 80, 30.0
    com = 10.0 + rng.uniform(0.0, 2.8, size=(n, 3))   # all pairs within 5 A
    q = np.tile([1.0, 0.0, 0.0, 0.0], (n, 1))
    print("umbrella capacity case", flush=True)
    s = g6.System(g6.TIP4P, n, L, 9.0, 0.5, 0.002)
    s.upload(com, q)
    s.mc_begin(300.0, 11, 0.1, 1.0)
    s.mc_set_umbrella(["TRIANGLE"], [0.0], [1.0], np.ones(21))
    before = s.mc_state()
    with pytest.raises(g6.G6Error, match="capacity"):
        s.mc_trials(2 * n)
    after = s.mc_state()
    for a, b in zip(before, after):
        np.testing.assert_array_equal(a, b)
    s.mc_set_umbrella([], [], [])
    acc = s.mc_trials(50)
    t = g6.System(g6.TIP4P, n, L, 9.0, 0.5, 0.002)
    t.upload(com, q)
    t.mc_begin(300.0, 11, 0.1, 1.0)
    assert acc == t.mc_trials(50)
    np.testing.assert_array_equal(s.mc_state()[0], t.mc_state()[0])
    s.close()
    t.close()


def test_gpu_q6_umbrella_mc_equals_oracle_1024(golden):
    """1024 molecules of a liquid (the centres of order_tip4p_1024.npz, seeded
    orientations), a Q6 and a triangle umbrella, two loops of 2N trials:
    the oracle's decisions and order parameters.  The Q6 sums are updated
    incrementally over 4096 trials on both sides."""
    g = golden("order_tip4p_1024")
    r = golden("mc2_umbrella_tip4p_216")
    com, L = g["com"], float(g["L"])
    n = len(com)
    rng = np.random.default_rng(1024)
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1)[:, None]
    q[q[:, 0] < 0] *= -1
    kinds = ["Q6LOCAL_", "TRIANGLE"]
    center, coeffi = [0.3, 0.5], [5.0e6, 3.0e3]
    print("q6 umbrella mc 1024", flush=True)
    s = g6.System(g6.TIP4P, n, L, 9.0, 0.5, 0.002)
    s.upload(com, q)
    s.mc_begin(300.0, 61, 0.15, 1.0)
    s.mc_set_umbrella(kinds, center, coeffi, r["refhist"])
    o = O.MCRun(O.TIP4P, com, q.T.copy(), L, 9.0, 0.5, 300.0, 61, 0.15, 1.0)
    o.set_umbrella(kinds, center, coeffi, r["refhist"])
    for loop in range(2):
        acc = s.mc_trials(2 * n)
        assert acc == o.trials(2 * n), loop
        op = s.mc_umbrella()[0]
        assert np.allclose(op, o.op, rtol=1e-12, atol=0), (loop, op, o.op)
    c, qq, d, rt, e = s.mc_state()
    np.testing.assert_array_equal(c, o.com)
    np.testing.assert_array_equal(d, o.delta)
    assert abs(e - o.epsum[0]) <= 1e-10 * abs(e)
    s.close()
