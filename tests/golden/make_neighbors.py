"""Writes tests/golden/neighbors_*.npz by running the reference's own
neighbour routines (Neighbor.F90: neighbors_all2, neighbor1_compress2,
neighbors_displacement) through oracle/_ref/refneigh, our dumper linked
against the unmodified reference modules (`make -f oracle/ref.mk`).

The fixtures are data: the input pair list and centres, the cases, and the
arrays the reference built.  Two inputs: the TIP4P-216 pair list of
tip4p_216.npz (9 A cutoff; sha256 digests of the arrays, in
neighbors_tip4p_216.json), and a 4x4x4 simple-cubic lattice with an
all-pairs minimum-image list, whose equal distances exercise the ties
(every array, in neighbors_lattice_64.npz).

    python tests/golden/make_neighbors.py
"""
import hashlib
import json
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
EXE = os.path.join(ROOT, "oracle", "_ref", "refneigh")
NEAR = 151
NONE, DIST, LABEL = 0, 1, 2

# maxnei order radius | cmaxnei corder cradius | target dx dy dz
# (radius <= 0: absent); the patterns of Q6Local.F90:123-129,
# Triangle.F90:81-86, traj2.F90:437-536 and mctest.F90:501-504, and more
CASES = [
    (0, LABEL, 6.0, 0, LABEL, 3.0, 5, 0.11, -0.07, 0.05),
    (0, LABEL, 6.0, 0, LABEL, 5.0, 17, -0.2, 0.3, 0.1),
    (0, DIST, 5.0, 4, DIST, 0.0, 3, 0.05, 0.02, -0.04),
    (10, DIST, 0.0, 4, DIST, 0.0, 40, 0.3, -0.25, 0.2),
    (4, DIST, 0.0, 0, LABEL, 3.5, 0, -0.6, 0.4, 0.5),
    (16, DIST, 0.0, 10, DIST, 4.0, 77, 0.01, 0.01, 0.01),
    (0, NONE, 3.0, 0, NONE, 0.0, 11, 0.4, 0.0, -0.3),
    (0, NONE, 0.0, 5, DIST, 5.5, 2, -0.15, -0.15, 0.15),
    (10, LABEL, 0.0, 3, LABEL, 0.0, 9, 0.2, 0.2, 0.2),
    (0, DIST, 0.0, 0, LABEL, 0.0, 30, 1.2, -0.9, 0.7),
]


def lattice64():
    m, a = 4, 3.1
    L = m * a
    g = np.arange(m) * a + 0.5 * a
    com = np.array([(x, y, z) for x in g for y in g for z in g])
    pi, pj, sh = [], [], []
    for i in range(len(com)):
        for j in range(i + 1, len(com)):
            d = com[i] - com[j]
            o = L * np.round(d / L)
            if np.sum((d - o) ** 2) < 6.5 ** 2:
                pi.append(i); pj.append(j); sh.append(o)
    return com, np.array(pi, np.int32), np.array(pj, np.int32), np.array(sh)


def run(com, pi, pj, sh, cases):
    n = len(com)
    lines = ["%d %d" % (n, len(pi))]
    lines += ["%d %d %.17e %.17e %.17e" % (pi[k] + 1, pj[k] + 1, *sh[k]) for k in range(len(pi))]
    lines += ["%.17e %.17e %.17e" % tuple(c) for c in com]
    lines += ["%d" % len(cases)]
    for c in cases:
        lines.append("%d %d %.17e %d %d %.17e %d %.17e %.17e %.17e"
                     % (c[0], c[1], c[2], c[3], c[4], c[5], c[6] + 1, c[7], c[8], c[9]))
    r = subprocess.run("ulimit -s unlimited; " + EXE, shell=True, input="\n".join(lines) + "\n",
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(r.stderr[-2000:])
    tok = r.stdout.split()
    p = 0

    def table(rows):
        nonlocal p
        nn = np.zeros(rows, np.int32)
        near = np.zeros((rows, NEAR), np.int32)
        vec = np.zeros((rows, NEAR, 4))
        for i in range(rows):
            nn[i] = int(tok[p]); p += 1
            for t in range(nn[i]):
                near[i, t] = int(tok[p]) - 1
                vec[i, t] = [float(v) for v in tok[p + 1:p + 5]]
                p += 5
        return nn, near, vec

    out = {}
    for ci in range(len(cases)):
        assert tok[p] == "ALL2"; p += 1
        out["all2_%d" % ci] = table(n)
        assert tok[p] == "COMP"; p += 1
        out["comp_%d" % ci] = table(n)
        assert tok[p] == "DISP"; p += 1
        m = int(tok[p]); p += 1
        idx = np.zeros(m, np.int32)
        nn = np.zeros(m, np.int32)
        near = np.zeros((m, NEAR), np.int32)
        vec = np.zeros((m, NEAR, 4))
        for r in range(m):
            idx[r] = int(tok[p]) - 1; p += 1
            nn[r] = int(tok[p]); p += 1
            for t in range(nn[r]):
                near[r, t] = int(tok[p]) - 1
                vec[r, t] = [float(v) for v in tok[p + 1:p + 5]]
                p += 5
        out["disp_%d" % ci] = (idx, nn, near, vec)
    assert p == len(tok)
    return out


def digest(nn, near, vec):
    """sha256 of a table's filled entries: nnear (<i4), then per array its
    labels (<i4) and (dx, dy, dz, dd) rows (<f8)."""
    h = hashlib.sha256(np.asarray(nn, "<i4").tobytes())
    for i in range(len(nn)):
        h.update(np.asarray(near[i, :nn[i]], "<i4").tobytes())
        h.update(np.asarray(vec[i, :nn[i]], "<f8").tobytes())
    return h.hexdigest()


def save_full(name, com, pi, pj, sh, cases):
    """Inputs and every array the reference built (small inputs)."""
    res = run(com, pi, pj, sh, cases)
    arrs = dict(coord=com, pi=pi, pj=pj, shift=sh, cases=np.array(cases, np.float64))
    for ci in range(len(cases)):
        for kind in ("all2", "comp"):
            nn, near, vec = res["%s_%d" % (kind, ci)]
            arrs["%s_nnear_%d" % (kind, ci)] = nn
            arrs["%s_near_%d" % (kind, ci)] = near
            arrs["%s_vec_%d" % (kind, ci)] = vec
        idx, nn, near, vec = res["disp_%d" % ci]
        arrs["disp_index_%d" % ci] = idx
        arrs["disp_nnear_%d" % ci] = nn
        arrs["disp_near_%d" % ci] = near
        arrs["disp_vec_%d" % ci] = vec
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **arrs)


def save_digests(name, src, cases):
    """Digests of the reference's arrays for the pair list of an existing
    fixture (the inputs stay in that fixture)."""
    g = np.load(os.path.join(HERE, src + ".npz"))
    res = run(g["com"], g["pi"], g["pj"], g["shift"], cases)
    out = {"source": src, "cases": [list(c) for c in cases], "digests": []}
    for ci in range(len(cases)):
        idx, nn, near, vec = res["disp_%d" % ci]
        out["digests"].append({
            "all2": digest(*res["all2_%d" % ci]), "comp": digest(*res["comp_%d" % ci]),
            "disp": digest(nn, near, vec), "disp_index": idx.tolist(),
            "all2_total": int(res["all2_%d" % ci][0].sum())})
    with open(os.path.join(HERE, name + ".json"), "w") as f:
        json.dump(out, f, indent=1)


def main():
    save_digests("neighbors_tip4p_216", "tip4p_216", CASES)
    com, pi, pj, sh = lattice64()
    lat = [c[:6] + (c[6] % 64,) + c[7:] for c in CASES]
    lat = [(c[0], c[1], min(c[2], 4.0) if c[2] > 0 else c[2], *c[3:]) for c in lat]
    save_full("neighbors_lattice_64", com, pi, pj, sh, lat)


if __name__ == "__main__":
    main()
