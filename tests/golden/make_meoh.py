"""Writes the OPLS methanol fixtures (the united-atom three-site methanol
of OPLSMeOHModule.F90 through StdForce, the reference driver's path without
@FAST: RigidSetup.F90:43-57, Mixture.F90:1984, Interaction2.F90:67):

  meoh_216.npz     one force evaluation of 216 methanol molecules by refdump
                   (our dumper over the reference's unmodified modules): pair
                   list, StdForce site forces, per-pair energies, virial,
                   molecular forces and torques
  md_meoh_216.npz  20 NVE and 20 Nose NVT steps of the reference MD driver
                   (Mixture, @ID08 OPLSMEOH) from the same start: @LOGD rows
                   and the final @WTG2 state

Start: a seeded lattice start of 216 molecules at 0.79 g/cm3 (M 32.04
g/mol, L = 24.409 A), rc 8.5 A switched over 0.5 A, dt 2 fs, T 300 K.  The
Mixture binary is built on demand by `make -f oracle/ref.mk mixture` and
removed afterwards.

    python tests/golden/make_meoh.py
"""
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, HERE)
from configs import lattice_start  # noqa: E402
from make_golden import REF, mixture, refdump  # noqa: E402

N, RC, TAPER, DT, SEED, STEPS = 216, 8.5, 0.5, 0.002, 4242, 20
L = round((N * 32.04 / (6.0221e23 * 0.79) * 1e24) ** (1.0 / 3.0), 3)
MEOH = 5


def main():
    if not os.path.exists(os.path.join(REF, "refdump")):
        sys.exit("build the reference first: make -f oracle/ref.mk")
    com, q = lattice_start(N, L, 20261202 + N + MEOH)
    out = refdump(MEOH, com, q, L, RC, TAPER, RC)
    np.savez_compressed(os.path.join(HERE, "meoh_216.npz"), com=com, quat=q, model=MEOH, L=L,
                        rc=RC, taper=TAPER, req=RC, **out)
    print("meoh_216 L", L, "npair", len(out["pi"]), "npair0", out["npair0"])
    built = not os.path.exists(os.path.join(REF, "Mixture"))
    if built:
        subprocess.run(["make", "-f", os.path.join(ROOT, "oracle", "ref.mk"), "mixture"],
                       cwd=ROOT, check=True, capture_output=True)
    res = {}
    try:
        for key, nose in [("nve", None), ("nvt", (300.0, 7.5e5, 0.0))]:
            logd, com_f, dxyz, quat, wxyz = mixture(com, q, L, RC, TAPER, DT, STEPS, SEED, nose,
                                                    model="OPLSMEOH")
            res.update({key + "_logd": logd, key + "_com_final": com_f,
                        key + "_dxyz_final": dxyz, key + "_quat_final": quat,
                        key + "_w_final_ext": wxyz})
            print(key, logd.shape, "T0 %.2f H0 %.6f  max|H-H0| %.3e" %
                  (logd[0, 1], logd[0, 12], np.abs(logd[:, 12] - logd[0, 12]).max()))
    finally:
        if built:
            os.remove(os.path.join(REF, "Mixture"))
    np.savez_compressed(os.path.join(HERE, "md_meoh_216.npz"), com=com, quat=q, L=L, rc=RC,
                        taper=TAPER, dt=DT, seed=SEED, nsteps=STEPS,
                        nose=np.array([300.0, 7.5e5, 0.0]), **res)


if __name__ == "__main__":
    main()
