"""Energy conservation (NVE drift) of the GPU MD path against the reference
driver's own, on 1000 steps of 216 TIP4P from an equilibrated start
(tests/golden/md_tip4p_216_nve1000.npz, written by tests/golden/make_nve.py
from the reference Mixture run).  The reference drifts (Gear-5 with
quaternions at 2 fs): H grows by about 25 kJ/mol per ps.  The bar is "no
worse than the reference": the drift rate (least-squares slope of H over
steps 100-999) and the largest excursion |H - H0| may exceed the
reference's by at most 10 %, in FP64 and FP32-pair modes, with and without
the pair-list skin.  FP64 also follows the reference's H step by step."""
import numpy as np
import pytest

from genesis6_amd import g6

pytestmark = pytest.mark.gpu


def drift(H, dt):
    t = np.arange(len(H)) * dt
    return np.polyfit(t[100:], H[100:], 1)[0], np.abs(H - H[0]).max()


@pytest.mark.parametrize("prec,skin", [(g6.FP64, 0.0), (g6.FP64, 1.0), (g6.FP32, 0.0),
                                       (g6.FP32, 1.0)])
def test_nve_drift_no_worse_than_reference(golden, prec, skin):
    g = golden("md_tip4p_216_nve1000")
    n, dt = len(g["com"]), float(g["dt"])
    print("nve-1000 prec", prec, "skin", skin, flush=True)
    s = g6.System(g6.TIP4P, n, float(g["L"]), float(g["rc"]), float(g["taper"]), dt,
                  prec=prec, skin=skin)
    try:
        s.upload(g["com"], g["quat"])
        s.vreset(300.0, int(g["seed"]))
        H = []
        for _ in range(int(g["nsteps"])):
            s.step(1)
            H.append(s.thermo()[13])
    finally:
        s.close()
    H = np.array(H)
    ref = g["logd"][:, 12]
    rate, exc = drift(H, dt)
    rrate, rexc = drift(ref, dt)
    dev = np.abs(H - ref) / np.abs(ref)
    print("  drift %.1f J/mol/ps (reference %.1f), max |H-H0| %.1f (reference %.1f), "
          "max rel |H-H_ref| %.2e, at step 100 %.2e" % (rate, rrate, exc, rexc, dev.max(),
                                                        dev[100]))
    assert abs(rate) <= 1.1 * abs(rrate)
    assert exc <= 1.1 * rexc
    if prec == g6.FP64:
        assert dev[:200].max() <= 1e-9
