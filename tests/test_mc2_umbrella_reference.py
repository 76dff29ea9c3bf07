"""The oracle's umbrella MC against the reference program mc2 itself
(tests/golden/mc2_umbrella_tip4p_216.npz, written by
tests/golden/make_mc2_umbrella.py from the unmodified mc2.F90).

mc2 adds sum(U(newop) - U(op)) with U = coeffi (op - center)^2 to every
trial's energy change (mc2.F90:684-706), the order parameters being the
triangle histogram's divergence from a reference histogram (TRIANGLE,
LOGTRIAN: Triangle.F90), the local bond order Q6 updated incrementally over
the trials of a loop (Q6LOCAL_: Q6Local.F90, whose sums the oracle must
reproduce through rejected trials too) and the energy (EPOP____).  Same seed and start:
every decision must be mc2's, so the step sizes and rotation ratios are
equal bit for bit, and the order parameters after each sweep equal what
mc2 prints.  The umbrellas are strong enough to steer the chain: without
them (planted in the oracle) the decisions differ."""
import numpy as np
import pytest

from oracle import pyoracle as po


def run_case(golden, r, c, coeffi_scale=1.0, refhist=None):
    g = golden(str(r["start_%d" % c]))
    seed, temp, sweeps = r["case_%d" % c]
    n = len(g["com"])
    run = po.MCRun(po.TIP4P, g["com"], g["quat"].T.copy(), float(g["L"]), float(g["rc"]),
                   float(g["taper"]), float(temp), int(seed), 0.1, 1.0)
    run.set_umbrella([str(k) for k in r["kinds_%d" % c]], r["center_%d" % c],
                     r["coeffi_%d" % c] * coeffi_scale,
                     r["refhist"] if refhist is None else refhist)
    rows = []
    for _ in range(int(sweeps)):
        run.trials(2 * n)                 # one mc2 sweep: ntrial = 2 * nmol
        rows.append([run.epsum[0]] + list(run.op))
    return run, np.array(rows)


@pytest.mark.parametrize("case", [0, 1, 2, 3])
def test_oracle_replays_reference_mc2_umbrella(golden, case):
    r = golden("mc2_umbrella_tip4p_216")
    run, rows = run_case(golden, r, case)
    want = r["mc2o_%d" % case]
    nu = len(r["kinds_%d" % case])
    assert np.allclose(rows[:, 0], want[:, 1], rtol=1e-13, atol=0)
    for k in range(nu):
        assert np.allclose(rows[:, 1 + k], want[:, 3 + 3 * k], rtol=1e-12, atol=0), k
        assert np.all(want[:, 4 + 3 * k] == r["center_%d" % case][k])
    assert np.array_equal(run.delta, r["delta_%d" % case])
    assert np.array_equal(run.rtr, r["rtratio_%d" % case])
    assert np.allclose(run.com, r["com_%d" % case], rtol=0, atol=2e-9 * 20)
    assert np.allclose(run.quat.T, r["quat_%d" % case], rtol=0, atol=1e-9)


@pytest.mark.parametrize("case", [0, 1, 2, 3])
def test_umbrella_steers_the_decisions(golden, case):
    r = golden("mc2_umbrella_tip4p_216")
    run, _ = run_case(golden, r, case, coeffi_scale=0.0)
    assert not np.array_equal(run.delta, r["delta_%d" % case])


def test_reference_histogram_enters(golden):
    r = golden("mc2_umbrella_tip4p_216")
    flat = np.roll(r["refhist"], 1)
    _, rows = run_case(golden, r, 1, refhist=flat)
    assert not np.allclose(rows[:, 1], r["mc2o_1"][:, 3], rtol=1e-6)
