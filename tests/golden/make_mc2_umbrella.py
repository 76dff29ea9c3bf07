"""Writes tests/golden/mc2_umbrella_tip4p_216.npz by running the reference's
MC program mc2 (mc2.F90 with its umbrella step, built unmodified by
`make -f oracle/ref.mk mc2` into oracle/_ref/mc2, removed afterwards) on
TIP4P-216 starts with umbrella potentials (@GEUM, mc2util.F90:174-191): the
lattice of tip4p_216.npz, and the equilibrated liquid of
md_tip4p_216_nve1000.npz for the Q6 umbrella (every molecule of the liquid
has bonds within 3 A; on the lattice none has, and q6 is 0/0).  The
reference histogram of the triangle order parameter is the @HIST block of
the reference's own sample input mc2.input.

For each case the fixture keeps mc2's @MC2O rows after every sweep of 2N
trials (loop, energy, 1/beta, then op, center, coeffi per umbrella, 17
digits), the final step sizes and rotation ratios (@DELT) and the final
configuration (@NX4A, 10 digits).

    python tests/golden/make_mc2_umbrella.py
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
from configs import mc2_files, parse_mc2_out, parse_mc2o_rows  # noqa: E402
ROOT = os.path.dirname(os.path.dirname(HERE))
EXE = os.path.join(ROOT, "oracle", "_ref", "mc2")
SAMPLE = "/root/reference/mc2.input"
# (start, seed, temperature, sweeps, umbrellas as (kind, center, coeffi))
CASES = [
    ("tip4p_216", 4321, 300.0, 3, [("TRIANGLE", 0.05, 1.0e4), ("EPOP____", -1.5e5, 1.0e-6)]),
    ("tip4p_216", 97, 230.0, 3, [("LOGTRIAN", -1.0, 3.0e3)]),
    ("md_tip4p_216_nve1000", 555, 300.0, 3, [("Q6LOCAL_", 0.2, 1.0e7)]),
    ("md_tip4p_216_nve1000", 8, 260.0, 2, [("Q6LOCAL_", 0.3, 3.0e6), ("TRIANGLE", 0.4, 2.0e3),
                                           ("EPOP____", -2.0e5, 1.0e-6)]),
]


def sample_refhist():
    lines = open(SAMPLE, encoding="latin-1").read().split("\n")
    k = lines.index("@HIST")
    assert lines[k + 1].split()[0] == "21"
    return np.array([float(v) for v in lines[k + 2:k + 23]])


def run_mc2(com, q, L, rc, taper, seed, temp, sweeps, umbrellas, refhist):
    cfg, deck = mc2_files(com, q, L, rc, taper, seed, temp, sweeps, umbrellas, refhist)
    with tempfile.TemporaryDirectory() as d:
        with open(os.path.join(d, "mc2.cfg"), "w") as f:
            f.write(cfg)
        with open(os.path.join(d, "node0.in"), "w") as f:
            f.write(deck)
        r = subprocess.run("ulimit -s unlimited; %s mc2.cfg" % EXE, shell=True, cwd=d,
                           capture_output=True, text=True, timeout=600)
        if r.returncode != 0:
            raise RuntimeError(r.stderr[-2000:])
        return open(os.path.join(d, "node0.in.out")).read()


def main():
    if not os.path.exists(EXE):
        subprocess.run(["make", "-f", "oracle/ref.mk", "mc2"], cwd=ROOT, check=True,
                       capture_output=True)
    refhist = sample_refhist()
    out = {"refhist": refhist, "ncases": np.array(len(CASES))}
    try:
        for c, (start, seed, temp, sweeps, umb) in enumerate(CASES):
            g = np.load(os.path.join(HERE, start + ".npz"))
            text = run_mc2(g["com"], g["quat"], float(g["L"]), float(g["rc"]),
                           float(g["taper"]), seed, temp, sweeps, umb, refhist)
            rows = parse_mc2o_rows(text)
            _, delta, rtr, com, quat = parse_mc2_out(text, len(g["com"]))
            assert rows.shape == (sweeps, 3 + 3 * len(umb))
            out.update({"start_%d" % c: np.array(start),
                        "case_%d" % c: np.array([seed, temp, sweeps], np.float64),
                        "kinds_%d" % c: np.array([u[0] for u in umb]),
                        "center_%d" % c: np.array([u[1] for u in umb], np.float64),
                        "coeffi_%d" % c: np.array([u[2] for u in umb], np.float64),
                        "mc2o_%d" % c: rows, "delta_%d" % c: delta, "rtratio_%d" % c: rtr,
                        "com_%d" % c: com, "quat_%d" % c: quat})
            print("case", c, rows[:, [1] + [3 + 3 * k for k in range(len(umb))]])
    finally:
        # the binary and mc2's own module objects are not kept
        os.remove(EXE)
        for m in ("Tetrahedrality", "Set", "DistribModule", "Triangle", "q6", "Q6Local",
                  "mc2util"):
            f = os.path.join(ROOT, "oracle", "_ref", m + ".o")
            if os.path.exists(f):
                os.remove(f)
        stamp = os.path.join(ROOT, "oracle", "_ref", "mc2.stamp")
        if os.path.exists(stamp):
            os.remove(stamp)
    np.savez_compressed(os.path.join(HERE, "mc2_umbrella_tip4p_216.npz"), **out)


if __name__ == "__main__":
    main()
