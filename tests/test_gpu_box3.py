"""Orthorhombic boxes on the GPU (g6_init_box3): MD with a fixed box and
with both modes of the Andersen barostat against the reference MD driver
(tests/golden/md_tip4p_432_box3.npz, written by tests/golden/make_box3_md.py:
432 TIP4P in a 37.252 x 18.626 x 18.626 A box, reference grid 4 x 2 x 2).
The Fortran driver g6md runs the reference's decks (@BOX3 with three edges,
@PCN4 mode 1 or 2) and logs its @LOGD rows; the restart carries the three
edges, which in the orz mode keep x and y and move z alone."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from genesis6_amd import g6
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


@pytest.mark.parametrize("name", ["nvt", "orz", "ortho"])
def test_box3_log_matches_reference_driver(golden, name):
    f = golden("md_tip4p_432_box3")
    print("g6md fp64 box3", name, flush=True)
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
    if name == "orz":
        same_restart(f13, f["fort13_orz"].tobytes(), 432)


def box3_start(golden):
    g = golden("md_tip4p_216_nve1000")
    L = float(g["L"])
    com = np.concatenate([g["com"], g["com"] + np.array([L, 0.0, 0.0])])
    quat = np.concatenate([g["quat"], g["quat"]])
    return g, com, quat, L


def test_orz_moves_z_alone(golden):
    """The orz barostat: x and y stay at their edges bit for bit, z takes
    the volume (v0 / (x y)); the orthorhombic mode scales all three."""
    g, com, quat, L = box3_start(golden)
    edges = {}
    for mode in (1, 2):
        s = g6.System(g6.TIP4P, 432, [2 * L, L, L], float(g["rc"]), float(g["taper"]),
                      float(g["dt"]), prec=g6.FP64)
        s.upload(com, quat)
        s.set_nose(300.0, 7.5e5)
        s.vreset(300.0, 4246)
        s.set_andersen(1.0, 1e-4, mode=mode)
        s.step(20)
        edges[mode] = s.box3()
        vol = s.thermo()[26]
        assert abs(np.prod(edges[mode]) - vol) <= 1e-12 * vol
        s.close()
    assert edges[2][0] == 2 * L and edges[2][1] == L and edges[2][2] != L
    r = edges[1] / np.array([2 * L, L, L])
    assert r[0] != 1.0 and abs(r[0] - r[1]) < 1e-15 and abs(r[0] - r[2]) < 1e-15


@pytest.mark.parametrize("prec", [g6.FP64, g6.FP32])
def test_equal_edges_are_the_cube(golden, prec):
    """g6_init_box3 with three equal edges runs the very step of g6_init:
    20 Nose steps, every thermo value and the state bit for bit."""
    g = golden("md_tip4p_216_nve1000")
    L = float(g["L"])
    out = []
    for box in (L, [L, L, L]):
        s = g6.System(g6.TIP4P, 216, box, float(g["rc"]), float(g["taper"]), float(g["dt"]),
                      prec=prec, skin=1.0)
        s.upload(g["com"], g["quat"])
        s.set_nose(300.0, 7.5e5)
        s.vreset(300.0, 4246)
        s.step(20)
        out.append((s.thermo(), s.download()[0]))
        s.close()
    np.testing.assert_array_equal(out[0][0], out[1][0])
    np.testing.assert_array_equal(out[0][1], out[1][1])


def test_non_cubic_boxes_stay_off_the_cubic_paths(golden):
    """Tiles, ladders, Monte Carlo and NPT Monte Carlo take cubic boxes: a
    non-cubic handle runs the rows and the others refuse it."""
    g, com, quat, L = box3_start(golden)
    s = g6.System(g6.TIP4P, 432, [2 * L, L, L], 8.5, 0.5, 0.002, prec=g6.FP32, skin=1.0)
    s.upload(com, quat)
    assert s.md_kernel()[0] == 0
    with pytest.raises(g6.G6Error, match="invalid argument"):
        s.mc_begin(300.0, 1, 0.1, 1.0)
    with pytest.raises(g6.G6Error, match="invalid argument"):
        g6.Ladder([s])
    s.close()
