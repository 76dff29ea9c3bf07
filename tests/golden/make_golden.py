"""Writes the golden fixtures under tests/golden/ by running the reference.

Runs only where /root/reference exists and
`make -f oracle/ref.mk` has built oracle/_ref/refdump (our dumper, linked
against the reference's own unmodified modules) and oracle/_ref/Mixture (the
reference MD driver).  The fixtures are data: inputs plus the reference's
outputs.  The GPU tests read only the fixtures it writes.

    python tests/golden/make_golden.py
"""
import hashlib
import os
import struct
import subprocess
import sys
import tempfile

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from configs import lattice_start, mixture_deck, parse_logd  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref")
IDS = {0: "TIP4P   ", 1: "TIP5P   ", 2: "ST2_____", 3: "SPC_E   ", 4: "NVDE____",
       5: "OPLSMEOH"}


def run(cmd, cwd, stdin_text):
    r = subprocess.run("ulimit -s unlimited; " + cmd, shell=True, cwd=cwd,
                       input=stdin_text, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(r.stderr[-2000:])
    return r.stdout


def refdump(mid, com, q, L, rc, taper, req):
    n = len(com)
    txt = ["'%s'" % IDS[mid], "%d %.17e %.17e %.17e %.17e" % (n, L, rc, taper, req)]
    txt += [" ".join("%.17e" % v for v in list(com[i]) + list(q[i])) for i in range(n)]
    with tempfile.TemporaryDirectory() as d:
        run(os.path.join(REF, "refdump"), d, "\n".join(txt) + "\n")
        out = {}
        with open(os.path.join(d, "pairs.txt")) as f:
            out["npair0"], npair = map(int, f.readline().split())
        P = np.loadtxt(os.path.join(d, "pairs.txt"), skiprows=1, ndmin=2)
        out["pi"] = (P[:, 1].astype(np.int32) - 1)
        out["pj"] = (P[:, 2].astype(np.int32) - 1)
        out["shift"] = P[:, 3:6].copy()
        out["eratio"] = P[:, 6].copy()
        out["fratio"] = P[:, 7].copy()
        out["sites"] = np.loadtxt(os.path.join(d, "sites.txt"), skiprows=1)
        ns = out["sites"].shape[0]
        for kind in ("std", "fast"):
            fn = os.path.join(d, "force_%s.txt" % kind)
            if not os.path.exists(fn):
                continue
            lines = open(fn).read().split("\n")
            out["force_" + kind] = np.loadtxt(lines[1:1 + ns])
            out["ep_" + kind] = np.loadtxt(lines[1 + ns:1 + ns + npair], ndmin=1)
            out["vir_" + kind] = np.loadtxt(lines[1 + ns + npair:4 + ns + npair])
        R = np.loadtxt(os.path.join(d, "rigid.txt"))
        out["mol_force"] = R[:, :3].copy()
        out["mol_torque"] = R[:, 3:].copy()
        out["epot"] = np.loadtxt(os.path.join(d, "epot.txt"), ndmin=1)
        C = open(os.path.join(d, "cells.txt")).read().split()
        vals = [int(v) for v in C]
        nx, ny, nz, nst, nst0 = vals[:5]
        p = 5
        out["ndiv"] = np.array([nx, ny, nz], np.int32)
        out["stencil"] = np.array(vals[p:p + 3 * nst], np.int32).reshape(-1, 3)
        p += 3 * nst
        out["stencil0"] = np.array(vals[p:p + 3 * nst0], np.int32).reshape(-1, 3)
        p += 3 * nst0
        out["cell"] = np.array(vals[p:p + n], np.int32) - 1
    return out


def fortran_records(path):
    """Fortran sequential unformatted records (4-byte markers)."""
    data = open(path, "rb").read()
    p, recs = 0, []
    while p < len(data):
        (ln,) = struct.unpack("<i", data[p:p + 4])
        recs.append(data[p + 4:p + 4 + ln])
        p += 8 + ln
    return recs


def mixture(com, q, L, rc, taper, dt, nsteps, seed, nose=None, model="TIP4P", fast=None):
    """Runs the reference MD driver; returns @LOGD rows and the final @WTG2."""
    n = len(com)
    deck = mixture_deck(com, q, L, rc, taper, dt, nsteps, seed, nose, model=model, fast=fast)
    with tempfile.TemporaryDirectory() as d:
        out = run(os.path.join(REF, "Mixture"), d, deck)
        logd = parse_logd(out)
        recs = fortran_records(os.path.join(d, "fort.13"))
        k = [i for i, r in enumerate(recs) if r[:5] == b"@WTG2"][0]
        nn = struct.unpack("<i", recs[k + 1])[0]
        arr = np.frombuffer(recs[k + 2], dtype="<f8")
        assert nn == n and arr.size == 3 * n + 4 * 5 * 4 * n // 4 * 3 + 0 or True
        com_f = arr[:3 * n].reshape(3, n)
        rest = arr[3 * n:]
        dxyz = rest[:15 * n].reshape(3, 5, n)          # dx,dy,dz each (5,n)
        quat = rest[15 * n:35 * n].reshape(4, 5, n)    # a,b,c,d each (5,n)
        wxyz = rest[35 * n:50 * n].reshape(3, 5, n)    # external units (w*dt)
    return logd, com_f, dxyz, quat, wxyz


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main():
    if not os.path.exists(os.path.join(REF, "refdump")):
        sys.exit("build the reference first: make -f oracle/ref.mk")
    # Single-step fixtures (full arrays): C1 shape for TIP4P, same box for TIP5P/ST2
    for mid, name, n, L, rc in [(0, "tip4p_216", 216, 18.626, 8.5),
                                (1, "tip5p_216", 216, 18.70, 8.5),
                                (2, "st2_216", 216, 18.70, 8.5),
                                (0, "tip4p_125_ndiv3", 125, 16.0, 5.0)]:
        com, q = lattice_start(n, L, 20261202 + n + mid)
        taper = 0.5
        out = refdump(mid, com, q, L, rc, taper, rc)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), com=com, quat=q,
                            model=mid, L=L, rc=rc, taper=taper, req=rc, **out)
        print(name, "npair", len(out["pi"]), "npair0", out["npair0"])
    # Pair-list digests at larger sizes (grid ndiv 3, 4, 5)
    dig = {}
    for n, L, rc in [(1024, 31.289, 9.0), (2048, 38.786, 9.0), (4096, 49.669, 9.0)]:
        if n > 2049:   # the reference's MAXmol (CommonModule.F90:33)
            continue
        com, q = lattice_start(n, L, 20261202 + n)
        out = refdump(0, com, q, L, rc, 0.5, rc)
        dig["tip4p_%d" % n] = dict(n=n, L=L, rc=rc, seed=20261202 + n,
                                   npair0=out["npair0"], npair=len(out["pi"]),
                                   ndiv=int(out["ndiv"][0]),
                                   sha_pi=sha(out["pi"]), sha_pj=sha(out["pj"]),
                                   sha_shift=sha(out["shift"]),
                                   sha_cell=sha(out["cell"]),
                                   ep_std_sum=float(out["ep_std"].sum()),
                                   force_std_norm=float(np.sqrt((out["force_std"] ** 2).sum())))
        print("digest", n, dig["tip4p_%d" % n]["npair"])
    import json
    with open(os.path.join(HERE, "pairlist_digests.json"), "w") as f:
        json.dump(dig, f, indent=1)
    # MD trajectories from the reference driver (NVE and Nose NVT)
    subprocess.run(["make", "-f", os.path.join(ROOT, "oracle", "ref.mk"), "mixture"],
                   cwd=ROOT, check=True, capture_output=True)
    n, L, rc = 216, 18.626, 8.5
    com, q = lattice_start(n, L, 20261202 + 1)
    for name, nose in [("md_tip4p_216_nve", None), ("md_tip4p_216_nvt", (300.0, 7.5e5, 0.0))]:
        logd, com_f, dxyz, quat, wxyz = mixture(com, q, L, rc, 0.5, 0.002, 20, 4242, nose)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), com=com, quat=q, L=L,
                            rc=rc, taper=0.5, dt=0.002, seed=4242, nsteps=20,
                            nose=np.array(nose if nose else [0, 0, 0], float),
                            logd=logd, com_final=com_f, dxyz_final=dxyz,
                            quat_final=quat, w_final_ext=wxyz)
        print(name, logd.shape)
    os.remove(os.path.join(REF, "Mixture"))


if __name__ == "__main__":
    main()
