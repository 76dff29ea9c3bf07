"""Writes tests/golden/restart_tip4p_216.npz: the reference MD driver's
binary restart file (fort.13, Mixture.F90:2530-2598) and a resumed run.

Run A: the deck of 216 TIP4P from the equilibrated start of
md_tip4p_216_nve1000.npz, Nose NVT 300 K, @VSET 300 K, 10 steps, @NLOG 1;
the fixture keeps the deck, the @LOGD rows and the fort.13 bytes.
Run B: fort.12 = run A's fort.13 and a deck of one tag (@MDLP 10): the
reference resumes from the restart for 10 more steps; the fixture keeps the
@LOGD rows and its fort.13 bytes.  The Mixture binary is built on demand
(`make -f oracle/ref.mk mixture`) and removed afterwards.

    python tests/golden/make_restart.py
"""
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, HERE)
from configs import mixture_deck, parse_logd  # noqa: E402
from make_golden import REF, run  # noqa: E402


def main():
    g = np.load(os.path.join(HERE, "md_tip4p_216_nve1000.npz"))
    deck_a = mixture_deck(g["com"], g["quat"], float(g["L"]), float(g["rc"]), float(g["taper"]),
                          float(g["dt"]), 10, 4244, (300.0, 7.5e5, 0.0))
    deck_b = "@MDLP\n10\n"
    exe = os.path.join(REF, "Mixture")
    built = not os.path.exists(exe)
    if built:
        subprocess.run(["make", "-f", os.path.join(ROOT, "oracle", "ref.mk"), "mixture"],
                       cwd=ROOT, check=True, capture_output=True)
    try:
        with tempfile.TemporaryDirectory() as d:
            out_a = run(exe, d, deck_a)
            with open(os.path.join(d, "fort.13"), "rb") as f:
                a13 = f.read()
            shutil.copy(os.path.join(d, "fort.13"), os.path.join(d, "fort.12"))
            os.remove(os.path.join(d, "fort.13"))
            out_b = run(exe, d, deck_b)
            with open(os.path.join(d, "fort.13"), "rb") as f:
                b13 = f.read()
    finally:
        if built:
            os.remove(exe)
    la, lb = parse_logd(out_a), parse_logd(out_b)
    print("run A", la.shape, "run B", lb.shape, "steps", lb[0, 0], "..", lb[-1, 0])
    np.savez_compressed(os.path.join(HERE, "restart_tip4p_216.npz"), deck_a=deck_a,
                        deck_b=deck_b, logd_a=la, logd_b=lb,
                        fort13_a=np.frombuffer(a13, np.uint8),
                        fort13_b=np.frombuffer(b13, np.uint8))


if __name__ == "__main__":
    main()
