"""Constant-pressure MD with the Andersen barostat on the GPU against the
reference MD driver (tests/golden/md_tip4p_216_npt.npz, written by
tests/golden/make_npt_md.py): the Fortran driver g6md runs the reference's
decks and logs its @LOGD rows; the pair-list skin keeps the interacting set
while the box moves; the restart carries the barostat (@PCN5)."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from genesis6_amd import g6
from oracle import pyoracle as O
from tests.golden.configs import parse_logd
from tests.test_gpu_restart import same_restart

pytestmark = pytest.mark.gpu
EXE = os.path.abspath(os.path.join(os.path.dirname(__file__), "..", "genesis6_amd", "lib",
                                   "g6md"))


def run_g6md(deck, *args):
    with tempfile.TemporaryDirectory() as d:
        r = subprocess.run([EXE] + list(args), input=deck, capture_output=True, text=True,
                           timeout=120, cwd=d)
        assert r.returncode == 0, r.stderr
        with open(os.path.join(d, "fort.13"), "rb") as f:
            return parse_logd(r.stdout), f.read()


@pytest.mark.parametrize("name", ["npt", "nph"])
def test_andersen_log_matches_reference_driver(golden, name):
    f = golden("md_tip4p_216_npt")
    print("g6md fp64", name, flush=True)
    got, f13 = run_g6md(str(f["deck_" + name]), "fp64")
    ref = f["logd_" + name]
    assert got.shape == ref.shape
    np.testing.assert_array_equal(got[:, 0], ref[:, 0])
    # T, P, V, ep, ek, Nose, Andersen energies, H
    for col, tol in [(1, 1e-10), (2, 1e-6), (3, 1e-12), (4, 1e-9), (5, 1e-10), (10, 1e-12),
                     (11, 1e-6), (12, 1e-10)]:
        err = np.abs(got[:, col] - ref[:, col]).max() / (np.abs(ref[:, col]).max() + 1e-300)
        print("  column %d: %.2e" % (col, err))
        assert err <= tol, (col, err)
    if name == "npt":
        same_restart(f13, f["fort13_npt"].tobytes(), 216)


@pytest.mark.parametrize("mid", [g6.SPCE, g6.NVDE, g6.TIP5P])
def test_andersen_other_models_follow_the_oracle(golden, mid):
    """The barostat with the other rigid models (the Andersen step does not
    depend on the model): 30 FP64 steps of @PCN4 1 1.0 1e-4 with the Nose
    thermostat against the oracle's step (pinned bit for bit to the
    reference driver for TIP4P): T, P, V and H at every step."""
    g = golden("md_tip4p_216_nve1000")
    L, rc, taper, dt = float(g["L"]), float(g["rc"]), float(g["taper"]), float(g["dt"])
    print("andersen model", mid, flush=True)
    s = g6.System(mid, 216, L, rc, taper, dt, prec=g6.FP64)
    md = O.MDRun(mid, g["com"], g["quat"].T.copy(), L, rc, taper, dt, nose=(300.0, 7.5e5, 0.0))
    try:
        s.upload(g["com"], g["quat"])
        s.set_nose(300.0, 7.5e5)
        s.vreset(300.0, 4246)
        s.set_andersen(1.0, 1e-4)
        md.set_andersen(1.0, 1e-4)
        md.vreset(300.0, 4246)
        for k in range(30):
            s.step(1)
            md.step()
            th = s.thermo()
            assert abs(th[3] - md.s.temperature) <= 1e-11 * md.s.temperature, k
            assert abs(th[25] - md.s.pressure) <= 1e-8 * max(1.0, abs(md.s.pressure)), k
            assert abs(th[26] - md.s.volume) <= 1e-13 * md.s.volume, k
            assert abs(th[13] - md.hamiltonian()) <= 1e-11 * abs(md.hamiltonian()), k
        assert abs(s.box() - md.s.L) <= 1e-13 * md.s.L
    finally:
        s.close()


def test_skin_keeps_the_interacting_set_while_the_box_moves(golden):
    """FP64 with a 1 A skin against no skin under a light barostat (the box
    edge moves by about 0.1 A in 300 steps): the same interacting pairs at
    every step, so the same trajectory bit for bit."""
    g = golden("md_tip4p_216_nve1000")
    runs = []
    for skin in (0.0, 1.0):
        s = g6.System(g6.TIP4P, 216, float(g["L"]), float(g["rc"]), float(g["taper"]),
                      float(g["dt"]), prec=g6.FP64, skin=skin)
        try:
            s.upload(g["com"], g["quat"])
            s.set_nose(300.0, 7.5e5)
            s.vreset(300.0, 4246)
            s.set_andersen(1.0, 2e-6)
            rows = []
            for _ in range(30):
                s.step(10)
                th = s.thermo()
                rows.append([th[14], th[26], th[0], th[13]])
            runs.append((np.array(rows), s.download()[0], s.box()))
        finally:
            s.close()
    (a, ca, La), (b, cb, Lb) = runs
    print("box %.6f -> %.6f, pairs %d..%d" % (float(g["L"]), La, a[:, 0].min(), a[:, 0].max()))
    assert abs(La - float(g["L"])) > 0.05           # the box did move
    np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(ca, cb)
    assert La == Lb


def test_fp32_follows_fp64_under_the_barostat(golden):
    """FP32 pair arithmetic with the skin against FP64: the volume and the
    temperature follow the FP64 run over 50 steps."""
    f = golden("md_tip4p_216_npt")
    a, _ = run_g6md(str(f["deck_npt"]), "fp64")
    b, _ = run_g6md(str(f["deck_npt"]), "fp32", "1.0")
    assert np.abs(b[:, 3] / a[:, 3] - 1).max() < 1e-6
    assert np.abs(b[:, 1] / a[:, 1] - 1).max() < 1e-4
