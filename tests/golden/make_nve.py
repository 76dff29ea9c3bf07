"""Writes tests/golden/md_tip4p_216_nve1000.npz: the energy conservation of
the reference MD driver (Mixture) over 1000 NVE steps of 216 TIP4P (C1: L
18.626 A, rc 8.5 A switched over 0.5 A, dt 2 fs) from an equilibrated
start with @VSET 300 K velocities.

The start is the lattice start of the other C1 fixtures after 4000 steps of
Nose NVT at 300 K on the oracle (CPU); the fixture keeps it (com, quat) and
the reference's @LOGD rows (column 13 is the conserved H in J/mol,
Mixture.F90:2431-2468).  The Mixture binary is built on demand by
`make -f oracle/ref.mk mixture` and removed afterwards.

    python tests/golden/make_nve.py
"""
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)
from configs import lattice_start  # noqa: E402
from make_golden import REF, mixture  # noqa: E402
from oracle import pyoracle as po  # noqa: E402

N, L, RC, TAPER, DT, SEED, STEPS = 216, 18.626, 8.5, 0.5, 0.002, 4243, 1000


def equilibrated():
    com, q = lattice_start(N, L, 20261202 + 1)
    md = po.MDRun(po.TIP4P, com, q.T.copy(), L, RC, TAPER, DT, nose=(300.0, 7.5e5, 0.0))
    md.vreset(300.0, 4242)
    md.step(4000)
    q = md.quat[0].T.copy()
    q /= np.sqrt((q * q).sum(1))[:, None]
    return md.com.copy(), q


def main():
    com, q = equilibrated()
    built = not os.path.exists(os.path.join(REF, "Mixture"))
    if built:
        subprocess.run(["make", "-f", os.path.join(ROOT, "oracle", "ref.mk"), "mixture"],
                       cwd=ROOT, check=True, capture_output=True)
    try:
        logd, _, _, _, _ = mixture(com, q, L, RC, TAPER, DT, STEPS, SEED, None)
    finally:
        if built:
            os.remove(os.path.join(REF, "Mixture"))
    H = logd[:, 12]
    print("steps", len(logd), "H0 %.6f  max|H-H0| %.6f  T %.1f..%.1f"
          % (H[0], np.abs(H - H[0]).max(), logd[:, 1].min(), logd[:, 1].max()))
    np.savez_compressed(os.path.join(HERE, "md_tip4p_216_nve1000.npz"), com=com, quat=q, L=L,
                        rc=RC, taper=TAPER, dt=DT, seed=SEED, nsteps=STEPS, logd=logd)


if __name__ == "__main__":
    main()
