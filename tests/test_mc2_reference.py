"""The oracle's mc2 trial loop against the reference program mc2 itself
(tests/golden/mc2_tip4p_216.npz, written by tests/golden/make_mc2.py from
the unmodified mc2.F90).  Same seed and start: every accept/reject decision
must be the reference's, so the step sizes and rotation ratios, which move
by a factor 1.05 on every decision, must be equal bit for bit, the energy
after each sweep equal to rounding and the final configuration equal to the
ten digits mc2 prints.  A changed draw order (the two arguments of each
random_normal, mc2.F90:595-600), a dropped term or a wrong sign in the
energy difference changes the decisions within a few trials."""
import numpy as np
import pytest

from oracle import pyoracle as po


@pytest.mark.parametrize("case", [0, 1])
def test_oracle_replays_reference_mc2(golden, case):
    g = golden("tip4p_216")
    r = golden("mc2_tip4p_216")
    seed, temp, sweeps = r["cases"][case]
    n = len(g["com"])
    run = po.MCRun(po.TIP4P, g["com"], g["quat"].T.copy(), float(g["L"]), float(g["rc"]),
                   float(g["taper"]), float(temp), int(seed), 0.1, 1.0)
    eps = []
    for _ in range(int(sweeps)):
        run.trials(2 * n)                 # one mc2 sweep: ntrial = 2 * nmol
        eps.append(run.epsum[0])
    want = r["epsum_%d" % case]
    assert np.allclose(eps, want, rtol=1e-13, atol=0), (eps, want)
    assert np.array_equal(run.delta, r["delta_%d" % case])
    assert np.array_equal(run.rtr, r["rtratio_%d" % case])
    assert np.allclose(run.com, r["com_%d" % case], rtol=0, atol=2e-9 * 20)
    assert np.allclose(run.quat.T, r["quat_%d" % case], rtol=0, atol=1e-9)
