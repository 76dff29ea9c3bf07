"""Writes tests/golden/order_*.npz: the umbrella order parameters of mc2 as
the reference computes them (Q6Local.F90 sq6system_prepareall,
Triangle.F90 sTriangle_prepareall), through oracle/_ref/refop, our dumper
linked against the unmodified reference modules
(`make -f oracle/ref.mk oracle/_ref/refop`; removed afterwards with the
module objects it needs).

Inputs: liquid configurations (the equilibrated 216-molecule start of
md_tip4p_216_nve1000.npz, and 1024 molecules equilibrated here on the
oracle), their all-pairs list within 8.5 A from the oracle, and the
reference histogram of the triangle order parameter in the reference's
mc2.input (@HIST).  Outputs: q6 and the bond count of every molecule, the
sum and the order parameter; the triangle histogram and both forms of its
order parameter; and for eight seeded trial displacements of one molecule
(up to 0.3 A per axis) the order parameters that the reference's incremental
sq6system_Differentiate and sTriangle_Differentiate return.

    python tests/golden/make_order.py
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
from oracle import pyoracle as po  # noqa: E402

EXE = os.path.join(ROOT, "oracle", "_ref", "refop")
# @HIST of the reference's mc2.input: 21 bins of 0.1 from -1
REFHIST = [2.18, 20.292, 65.672, 138.316, 245.649, 374.39, 478.799, 515.596, 465.066, 355.355,
           226.949, 117.862, 46.259, 13.03, 3.083, 1.329, 1.811, 0.359, 0.003, 0.0, 0.0]


def run(com, L, rc):
    _, pi, pj, sh = po.pairlist(com, L, rc)
    n = len(com)
    rng = np.random.default_rng(n)
    targets = rng.choice(n, 8, replace=False).astype(np.int32)
    delta = rng.uniform(-0.3, 0.3, (8, 3))
    lines = ["%d %d" % (n, len(pi))]
    lines += ["%d %d %.17e %.17e %.17e" % (pi[k] + 1, pj[k] + 1, *sh[k]) for k in range(len(pi))]
    lines += ["%.17e %.17e %.17e" % tuple(c) for c in com]
    lines += ["%.17e" % v for v in REFHIST]
    lines += ["%d" % len(targets)]
    lines += ["%d %.17e %.17e %.17e" % (targets[c] + 1, *delta[c]) for c in range(len(targets))]
    r = subprocess.run("ulimit -s unlimited; " + EXE, shell=True, input="\n".join(lines) + "\n",
                       capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        raise RuntimeError(r.stderr[-2000:])
    t = r.stdout.split("\n")
    k = t.index("Q6")
    q6 = np.array([float(x.split()[0]) for x in t[k + 1:k + 1 + n]])
    nb = np.array([int(x.split()[1]) for x in t[k + 1:k + 1 + n]], np.int32)
    sumq6, op = map(float, t[k + 1 + n].split())
    out = dict(com=com, L=L, rc=rc, pi=pi, pj=pj, shift=sh, refhist=np.array(REFHIST), q6=q6,
               nbond=nb, sumq6=sumq6, op_q6=op)
    for tag in ("TRI", "LOGTRI"):
        k = t.index(tag)
        out["hist_" + tag.lower()] = np.array([float(x) for x in t[k + 1:k + 22]])
        out["op_" + tag.lower()] = float(t[k + 22])
    out["disp_target"], out["disp_delta"] = targets, delta
    for tag in ("DQ6", "DTRI"):
        out["disp_op_" + tag[1:].lower()] = np.array(
            [float(t[k + 1]) for k in range(len(t)) if t[k] == tag])
        assert len(out["disp_op_" + tag[1:].lower()]) == len(targets)
    return out


def liquid_1024():
    n, L = 1024, 31.289
    com, q = lattice_start(n, L, 20261202 + n)
    md = po.MDRun(po.TIP4P, com, q.T.copy(), L, 8.5, 0.5, 0.002, nose=(300.0, 7.5e5, 0.0))
    md.vreset(300.0, 4247)
    md.step(2000)
    return md.com.copy(), L


def main():
    built = not os.path.exists(EXE)
    if built:
        subprocess.run(["make", "-f", "oracle/ref.mk", "oracle/_ref/refop"], cwd=ROOT,
                       check=True, capture_output=True)
    try:
        g = np.load(os.path.join(HERE, "md_tip4p_216_nve1000.npz"))
        for name, (com, L) in [("order_tip4p_216", (g["com"], float(g["L"]))),
                               ("order_tip4p_1024", liquid_1024())]:
            out = run(np.ascontiguousarray(com), L, 8.5)
            np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
            print(name, "pairs", len(out["pi"]), "op_q6 %.6f op_tri %.6f nbond %d..%d"
                  % (out["op_q6"], out["op_tri"], out["nbond"].min(), out["nbond"].max()))
    finally:
        if built:
            os.remove(EXE)
            for m in ("Tetrahedrality", "Set", "DistribModule", "Triangle", "q6", "Q6Local",
                      "mc2util"):
                f = os.path.join(ROOT, "oracle", "_ref", m + ".o")
                if os.path.exists(f):
                    os.remove(f)
            stamp = os.path.join(ROOT, "oracle", "_ref", "mc2.stamp")
            if os.path.exists(stamp):
                os.remove(stamp)


if __name__ == "__main__":
    main()
