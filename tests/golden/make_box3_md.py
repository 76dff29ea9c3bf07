"""Writes tests/golden/md_tip4p_432_box3.npz: MD in an orthorhombic box run
by the reference MD driver (Mixture), with a fixed box and with both modes
of the Andersen barostat.

Start: the equilibrated 216-molecule liquid of md_tip4p_216_nve1000.npz
and its copy shifted by one edge along x, 432 molecules in a box of
2L x L x L (37.252 x 18.626 x 18.626 A); rc 8.5 A switched over 0.5 A,
@VOXA rc, so the reference grid has 4 x 2 x 2 cells; dt 2 fs; @VSET 300 K;
the Nose thermostat at 300 K.  Decks (@BOX3 with the three edges):
"nvt" 30 steps; "orz" 50 steps with @PCN4 2 1.0 1e-4 (the z edge follows
the volume); "ortho" 50 steps with @PCN4 1 1.0 1e-4 (every edge follows
it).  The fixture keeps each deck and its @LOGD rows, and the fort.13
restarts of "orz".  The Mixture binary is built on demand and removed
afterwards.

    python tests/golden/make_box3_md.py
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, HERE)
from configs import mixture_deck, parse_logd  # noqa: E402
from make_golden import REF, run  # noqa: E402


def start():
    g = np.load(os.path.join(HERE, "md_tip4p_216_nve1000.npz"))
    L = float(g["L"])
    com = np.concatenate([g["com"], g["com"] + np.array([L, 0.0, 0.0])])
    quat = np.concatenate([g["quat"], g["quat"]])
    return g, com, quat, (2.0 * L, L, L)


def deck(nsteps, pcn4=None):
    g, com, quat, box = start()
    d = mixture_deck(com, quat, box[0], float(g["rc"]), float(g["taper"]), float(g["dt"]),
                     nsteps, 4246, (300.0, 7.5e5, 0.0))
    d = d.replace("@BXLA\n%r\n" % box[0], "@BOX3\n%r %r %r\n" % box, 1)
    if pcn4:
        d = d.replace("@ID08", "@PCN4\n%s\n@ID08" % pcn4, 1)
    return d


def main():
    decks = {"nvt": deck(30), "orz": deck(50, "2 1.0 0.0001"), "ortho": deck(50, "1 1.0 0.0001")}
    assert all("@BOX3" in d and "@BXLA" not in d for d in decks.values())
    exe = os.path.join(REF, "Mixture")
    built = not os.path.exists(exe)
    if built:
        subprocess.run(["make", "-f", os.path.join(ROOT, "oracle", "ref.mk"), "mixture"],
                       cwd=ROOT, check=True, capture_output=True)
    out = {}
    try:
        for name, dk in decks.items():
            with tempfile.TemporaryDirectory() as d:
                L = parse_logd(run(exe, d, dk))
                out["deck_" + name] = dk
                out["logd_" + name] = L
                if name == "orz":
                    with open(os.path.join(d, "fort.13"), "rb") as f:
                        out["fort13_orz"] = np.frombuffer(f.read(), np.uint8)
                print(name, L.shape, "T %.3f -> %.3f" % (L[0, 1], L[-1, 1]),
                      "V %.4f -> %.4f" % (L[0, 3], L[-1, 3]))
    finally:
        if built:
            os.remove(exe)
    np.savez_compressed(os.path.join(HERE, "md_tip4p_432_box3.npz"), **out)


if __name__ == "__main__":
    main()
