"""Writes tests/golden/md_tip4p_216_npt.npz: constant-pressure MD with the
reference's Andersen barostat (@PCN4, orthorhombic: the cubic box follows
the volume), run by the reference MD driver (Mixture).

Two decks on the equilibrated start of md_tip4p_216_nve1000.npz (216
TIP4P, rc 8.5 A switched over 0.5 A, dt 2 fs, @VSET 300 K), 50 steps,
@NLOG 1, barostat at 1 atm with mass 1e-4 per molecule: "npt" with the Nose
thermostat at 300 K, "nph" without.  The fixture keeps each deck, the @LOGD
rows (T, P, V, the barostat's energies and H) and the npt run's fort.13 (with
@PCN5).  The Mixture binary is built on demand and removed afterwards.

    python tests/golden/make_npt_md.py
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

PCN4 = "@PCN4\n1 1.0 0.0001\n"


def deck(g, nose):
    d = mixture_deck(g["com"], g["quat"], float(g["L"]), float(g["rc"]), float(g["taper"]),
                     float(g["dt"]), 50, 4246, nose)
    return d.replace("@ID08", PCN4 + "@ID08", 1)


def main():
    g = np.load(os.path.join(HERE, "md_tip4p_216_nve1000.npz"))
    decks = {"npt": deck(g, (300.0, 7.5e5, 0.0)), "nph": deck(g, None)}
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
                if name == "npt":
                    with open(os.path.join(d, "fort.13"), "rb") as f:
                        out["fort13_npt"] = np.frombuffer(f.read(), np.uint8)
                print(name, L.shape, "V %.4f -> %.4f" % (L[0, 3], L[-1, 3]))
    finally:
        if built:
            os.remove(exe)
    np.savez_compressed(os.path.join(HERE, "md_tip4p_216_npt.npz"), **out)


if __name__ == "__main__":
    main()
