"""Writes tests/golden/md_tip5p_216.npz: 20 NVE and 20 Nose NVT steps of
216 TIP5P by the reference MD driver (Mixture, @ID08 TIP5P, StdForce:
Mixture.F90:1984, Interaction2.F90:67) from the start of tip5p_216.npz
(L 18.70 A, rc 8.5 A switched over 0.5 A, dt 2 fs): @LOGD rows and the
final @WTG2 state of each.  The Mixture binary is built on demand by
`make -f oracle/ref.mk mixture` and removed afterwards.

    python tests/golden/make_tip5p_md.py
"""
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, HERE)
from make_golden import REF, mixture  # noqa: E402

DT, SEED, STEPS = 0.002, 4242, 20


def main():
    g = np.load(os.path.join(HERE, "tip5p_216.npz"))
    com, q, L, rc, taper = g["com"], g["quat"], float(g["L"]), float(g["rc"]), float(g["taper"])
    built = not os.path.exists(os.path.join(REF, "Mixture"))
    if built:
        subprocess.run(["make", "-f", os.path.join(ROOT, "oracle", "ref.mk"), "mixture"],
                       cwd=ROOT, check=True, capture_output=True)
    res = {}
    try:
        for key, nose in [("nve", None), ("nvt", (300.0, 7.5e5, 0.0))]:
            logd, com_f, dxyz, quat, wxyz = mixture(com, q, L, rc, taper, DT, STEPS, SEED, nose,
                                                    model="TIP5P")
            res.update({key + "_logd": logd, key + "_com_final": com_f,
                        key + "_dxyz_final": dxyz, key + "_quat_final": quat,
                        key + "_w_final_ext": wxyz})
            print(key, logd.shape, "H0 %.6f  max|H-H0| %.3e" %
                  (logd[0, 12], np.abs(logd[:, 12] - logd[0, 12]).max()))
    finally:
        if built:
            os.remove(os.path.join(REF, "Mixture"))
    np.savez_compressed(os.path.join(HERE, "md_tip5p_216.npz"), com=com, quat=q, L=L, rc=rc,
                        taper=taper, dt=DT, seed=SEED, nsteps=STEPS,
                        nose=np.array([300.0, 7.5e5, 0.0]), **res)


if __name__ == "__main__":
    main()
