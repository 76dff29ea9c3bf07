"""Writes tests/golden/md_mixture_tip4p_meoh.npz: a two-component MD run
(TIP4P water and OPLS methanol, @NCMP 2) by the reference MD driver
(Mixture).

Start: a simple cubic lattice of 8 x 8 x 8 sites, spacing 3.625 A, in a
box of 29 A; the 64 sites with even lattice indices on all three axes hold
methanol, the other 448 water; orientations are uniform random quaternions
(normalised Gaussian 4-vectors, seed 7).  rc 8.5 A switched over 0.5 A,
@VOXA rc, so the reference grid has 3 x 3 x 3 cells; dt 2 fs;
@VSET 300 K; the Nose thermostat at 300 K; 30 steps, one @LOGD row each;
the same deck once more with @FAST 1 (the reference's custom TIP4P,
methanol and TIP4P x methanol kernels, Interaction2.F90:19-59).
The deck is the single-component deck of mixture_deck with a second @ID08
OPLSMEOH / @NX4A block after the water's (Mixture.F90:636-712 adds a
component for each @ID08 + @NX4A pair).  The Mixture binary is built on
demand and removed afterwards.

    python tests/golden/make_mixture_md.py
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

L, NSIDE, RC, TAPER, DT, NSTEPS, SEED = 29.0, 8, 8.5, 0.5, 0.002, 30, 4246


def start():
    a = L / NSIDE
    rng = np.random.default_rng(7)
    water, meoh = [], []
    for i in range(NSIDE):
        for j in range(NSIDE):
            for k in range(NSIDE):
                c = [(i + 0.5) * a, (j + 0.5) * a, (k + 0.5) * a]
                q = rng.standard_normal(4)
                q = q / np.linalg.norm(q)
                (meoh if i % 2 == 0 and j % 2 == 0 and k % 2 == 0 else water).append(
                    list(c) + list(q))
    return np.array(water), np.array(meoh)


def deck():
    water, meoh = start()
    d = mixture_deck(water[:, :3], water[:, 3:], L, RC, TAPER, DT, NSTEPS, SEED,
                     (300.0, 7.5e5, 0.0))
    d += "@ID08\nOPLSMEOH\n@NX4A\n%d\n" % len(meoh)
    d += "\n".join(" ".join("%.17e" % v for v in r) for r in meoh) + "\n"
    return d, water, meoh


def main():
    dk, water, meoh = deck()
    exe = os.path.join(REF, "Mixture")
    built = not os.path.exists(exe)
    if built:
        subprocess.run(["make", "-f", os.path.join(ROOT, "oracle", "ref.mk"), "mixture"],
                       cwd=ROOT, check=True, capture_output=True)
    fast = dk.replace("@ID08", "@FAST\n1\n@ID08", 1)
    try:
        with tempfile.TemporaryDirectory() as d:
            text = run(exe, d, dk)
        with tempfile.TemporaryDirectory() as d:
            text_fast = run(exe, d, fast)
    finally:
        if built:
            os.remove(exe)
    logd = parse_logd(text)
    print(logd.shape, "T %.3f -> %.3f" % (logd[0, 1], logd[-1, 1]),
          "ep(ww, mm, wm) %s" % logd[0, 4:7], "H %.6f -> %.6f" % (logd[0, -4], logd[-1, -4]))
    logd_fast = parse_logd(text_fast)
    print("@FAST 1: largest relative difference per column from StdForce's run",
          (np.abs(logd_fast - logd).max(0) / (np.abs(logd).max(0) + 1e-300)).max())
    np.savez_compressed(os.path.join(HERE, "md_mixture_tip4p_meoh.npz"), deck=dk, logd=logd,
                        deck_fast=fast, logd_fast=logd_fast,
                        water=water, meoh=meoh, L=L, rc=RC, taper=TAPER, dt=DT, seed=SEED,
                        nsteps=NSTEPS)


if __name__ == "__main__":
    main()
