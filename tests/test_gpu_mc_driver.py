"""The Fortran MC driver g6mc reads the reference mc2's own input files
(configuration file with @MPIN, node input) and writes mc2's output file;
on TIP4P-216 it writes what mc2 itself wrote for the same input: the energy
of every @MC2O line, the final step sizes and rotation ratios (@DELT) and
the final configuration (@NX4A).  With umbrellas (@HIST, @GEUM) its @MC2O
lines carry each order parameter with its center and coefficient, as mc2's
do, and equal mc2's."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from tests.golden.configs import mc2_files, parse_mc2_out, parse_mc2o_rows

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "genesis6_amd", "lib", "g6mc")


@pytest.mark.parametrize("case", [0, 1])
def test_g6mc_writes_what_mc2_writes(golden, case):
    g = golden("tip4p_216")
    ref = golden("mc2_tip4p_216")
    seed, temp, sweeps = ref["cases"][case]
    cfg, deck = mc2_files(g["com"], g["quat"], float(g["L"]), float(g["rc"]), float(g["taper"]),
                          int(seed), float(temp), int(sweeps))
    print("g6mc case", case, flush=True)
    n = len(g["com"])
    with tempfile.TemporaryDirectory() as d:
        with open(os.path.join(d, "mc2.cfg"), "w") as f:
            f.write(cfg)
        with open(os.path.join(d, "node0.in"), "w") as f:
            f.write(deck)
        r = subprocess.run([EXE, "mc2.cfg"], cwd=d, capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr
        text = open(os.path.join(d, "node0.in.out")).read()
    eps, delta, rtr, com, quat = parse_mc2_out(text, n)
    e_ref = ref["epsum_%d" % case]
    assert len(eps) == len(e_ref)
    assert np.abs(eps - e_ref).max() <= 1e-13 * np.abs(e_ref).max()
    # every decision the same: step sizes and ratios change by 1.05 on each
    assert np.array_equal(delta, ref["delta_%d" % case])
    assert np.array_equal(rtr, ref["rtratio_%d" % case])
    # the configuration to the ten digits both print
    assert np.array_equal(com, ref["com_%d" % case]) and np.array_equal(quat, ref["quat_%d" % case])


def run_g6mc(cfg, deck):
    with tempfile.TemporaryDirectory() as d:
        with open(os.path.join(d, "mc2.cfg"), "w") as f:
            f.write(cfg)
        with open(os.path.join(d, "node0.in"), "w") as f:
            f.write(deck)
        r = subprocess.run([EXE, "mc2.cfg"], cwd=d, capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr
        return open(os.path.join(d, "node0.in.out")).read()


@pytest.mark.parametrize("case", [0, 1, 2, 3])
def test_g6mc_umbrella_deck_writes_what_mc2_writes(golden, case):
    ref = golden("mc2_umbrella_tip4p_216")
    g = golden(str(ref["start_%d" % case]))
    seed, temp, sweeps = ref["case_%d" % case]
    umb = list(zip([str(k) for k in ref["kinds_%d" % case]], ref["center_%d" % case],
                   ref["coeffi_%d" % case]))
    cfg, deck = mc2_files(g["com"], g["quat"], float(g["L"]), float(g["rc"]), float(g["taper"]),
                          int(seed), float(temp), int(sweeps), umb, ref["refhist"])
    print("g6mc umbrella case", case, flush=True)
    text = run_g6mc(cfg, deck)
    rows = parse_mc2o_rows(text)
    want = ref["mc2o_%d" % case]
    assert rows.shape == want.shape
    assert np.array_equal(rows[:, [0, 2]], want[:, [0, 2]])          # loop, 1/beta
    assert np.abs(rows[:, 1] - want[:, 1]).max() <= 1e-13 * np.abs(want[:, 1]).max()
    for k in range(len(umb)):
        o, w = rows[:, 3 + 3 * k], want[:, 3 + 3 * k]
        assert np.abs(o - w).max() <= 1e-11 * np.abs(w).max(), (k, o, w)
        assert np.array_equal(rows[:, 4 + 3 * k:6 + 3 * k], want[:, 4 + 3 * k:6 + 3 * k])
    _, delta, rtr, com, quat = parse_mc2_out(text, len(g["com"]))
    assert np.array_equal(delta, ref["delta_%d" % case])
    assert np.array_equal(rtr, ref["rtratio_%d" % case])
    assert np.array_equal(com, ref["com_%d" % case]) and np.array_equal(quat, ref["quat_%d" % case])
