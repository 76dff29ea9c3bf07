"""Writes tests/golden/mc2_tip4p_216.npz by running the reference's own MC
program mc2 (mc2.F90, built unmodified by `make -f oracle/ref.mk mc2` into
oracle/_ref/mc2, which this script removes afterwards: about 280 MB).

Input: the TIP4P-216 configuration of tip4p_216.npz, its box and cutoff, no
grid and no bookkeeping (mc2 then uses the all-pairs list), one node.  For
each case (seed, temperature, sweeps) the fixture keeps what mc2 writes:
the energy after every sweep of 2N trials (@MC2O, 17 digits), the final
step sizes and rotation ratios (@DELT, 17 digits) and the final
configuration (@NX4A, 10 digits).

    python tests/golden/make_mc2.py
"""
import os
import subprocess
import tempfile

import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
from configs import mc2_files, parse_mc2_out  # noqa: E402
ROOT = os.path.dirname(os.path.dirname(HERE))
EXE = os.path.join(ROOT, "oracle", "_ref", "mc2")
CASES = [(4321, 300.0, 3), (97, 230.0, 4)]


def run_mc2(com, q, L, rc, taper, seed, temp, sweeps):
    cfg, deck = mc2_files(com, q, L, rc, taper, seed, temp, sweeps)
    with tempfile.TemporaryDirectory() as d:
        with open(os.path.join(d, "mc2.cfg"), "w") as f:
            f.write(cfg)
        with open(os.path.join(d, "node0.in"), "w") as f:
            f.write(deck)
        r = subprocess.run("ulimit -s unlimited; %s mc2.cfg" % EXE, shell=True, cwd=d,
                           capture_output=True, text=True, timeout=600)
        if r.returncode != 0:
            raise RuntimeError(r.stderr[-2000:])
        text = open(os.path.join(d, "node0.in.out")).read()
    return parse_mc2_out(text, len(com))


def main():
    if not os.path.exists(EXE):
        subprocess.run(["make", "-f", "oracle/ref.mk", "mc2"], cwd=ROOT, check=True,
                       capture_output=True)
    g = np.load(os.path.join(HERE, "tip4p_216.npz"))
    out = {"cases": np.array(CASES, np.float64)}
    try:
        for c, (seed, temp, sweeps) in enumerate(CASES):
            eps, delta, rtr, com, quat = run_mc2(g["com"], g["quat"], float(g["L"]),
                                                 float(g["rc"]), float(g["taper"]), seed,
                                                 temp, sweeps)
            assert len(eps) == sweeps
            out.update({"epsum_%d" % c: eps, "delta_%d" % c: delta, "rtratio_%d" % c: rtr,
                        "com_%d" % c: com, "quat_%d" % c: quat})
            print("case", c, "energies", eps)
    finally:
        # the binary and mc2's own module objects (DistribModule.o alone
        # holds 50 MB of static arrays) are not kept
        os.remove(EXE)
        for m in ("Tetrahedrality", "Set", "DistribModule", "Triangle", "q6", "Q6Local",
                  "mc2util"):
            for ext in (".o",):
                f = os.path.join(ROOT, "oracle", "_ref", m + ext)
                if os.path.exists(f):
                    os.remove(f)
        stamp = os.path.join(ROOT, "oracle", "_ref", "mc2.stamp")
        if os.path.exists(stamp):
            os.remove(stamp)
    np.savez_compressed(os.path.join(HERE, "mc2_tip4p_216.npz"), **out)


if __name__ == "__main__":
    main()
