"""CPU baseline timing of the genuine reference beside the oracle.

Builds the reference MD driver Mixture from the sources under
/root/reference (unmodified; `make -f oracle/ref.mk mixture`) for the
216-molecule system, and for 4096 molecules from a scratch copy under /tmp
whose CommonModule selects the reference's own alternative MAXmol=4097 line
(CommonModule.F90:31, commented out in the distributed file) - nothing is
copied into the repository, and the built binaries go to oracle/_ref/.
Times, on this host's cores:
  * the reference Mixture, one process, and eight concurrent processes (the
    reference's replica-parallel mode: independent runs);
  * the oracle's FP64 step (the serial C restatement), one process;
for TIP4P 216 (rc 8.5, NVE) and TIP4P 4096 (rc 9.0, Nose NVT 300 K), both
from the seeded lattice start.  A rate is molecules x (steps_b - steps_a) /
(wall_b - wall_a) over two runs of different lengths, so deck parsing,
setup and output are excluded.  Prints one JSON object.

    python oracle/ref_timing.py [--quick]
"""
import concurrent.futures as cf
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)
REF = "/root/reference"
OUT = os.path.join(HERE, "_ref")
BIG = os.path.join(OUT, "big")

from tests.golden.configs import lattice_start, mixture_deck  # noqa: E402


def build_small():
    subprocess.run(["make", "-f", os.path.join(HERE, "ref.mk"), "mixture"], cwd=ROOT, check=True,
                   capture_output=True)
    return os.path.join(OUT, "Mixture")


def build_big():
    """Mixture with MAXmol = 4097, built from a scratch copy of the sources."""
    exe = os.path.join(BIG, "Mixture")
    if os.path.exists(exe):
        return exe
    os.makedirs(BIG, exist_ok=True)
    with tempfile.TemporaryDirectory(prefix="g6ref_") as d:
        for f in os.listdir(REF):
            if f.endswith((".F90", ".f", ".c", ".h")):
                shutil.copy(os.path.join(REF, f), d)
        cm = os.path.join(d, "CommonModule.F90")
        src = open(cm, encoding="latin-1").read()
        old, new = "integer,parameter :: MAXmol=2049", "!integer,parameter :: MAXmol=2049"
        assert old in src and "!integer,parameter :: MAXmol=4097" in src
        src = src.replace(old, new, 1).replace("!integer,parameter :: MAXmol=4097",
                                               "integer,parameter :: MAXmol=4097", 1)
        open(cm, "w", encoding="latin-1").write(src)
        r = subprocess.run(["make", "-f", os.path.join(HERE, "ref.mk"), "REF=" + d, "OUT=" + BIG,
                            "mixture"], cwd=ROOT, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(r.stderr[-3000:])
    return exe


def run_mixture(exe, deck):
    with tempfile.TemporaryDirectory() as d:
        t0 = time.perf_counter()
        r = subprocess.run("ulimit -s unlimited; " + exe + " > /dev/null", shell=True, cwd=d,
                           input=deck.encode(), capture_output=True)
        t1 = time.perf_counter()
    if r.returncode != 0:
        raise RuntimeError(r.stderr[-2000:].decode("latin-1"))
    return t1 - t0


CASES = {
    # name: (n, L, rc, Nose (T, q, dump) or None, steps a, steps b)
    "C1_tip4p_216_nve": (216, 18.626, 8.5, None, 200, 1200),
    "C2_tip4p_4096_nvt": (4096, 49.669, 9.0, (300.0, 7.5e5, 0.0), 20, 120),
}


def deck_of(name, nsteps):
    n, L, rc, nose, _, _ = CASES[name]
    com, q = lattice_start(n, L, 20261202 + (1 if n == 216 else 2))
    return mixture_deck(com, q, L, rc, 0.5, 0.002, nsteps, 4242, nose, fast=True, nlog=1000000)


def ref_rate(exe, name, procs=1):
    n, _, _, _, a, b = CASES[name]
    da, db = deck_of(name, a), deck_of(name, b)
    with cf.ThreadPoolExecutor(procs) as ex:
        ta = list(ex.map(lambda _: run_mixture(exe, da), range(procs)))
        tb = list(ex.map(lambda _: run_mixture(exe, db), range(procs)))
    return n * (b - a) * procs / (max(tb) - max(ta))


def oracle_rate(name):
    from oracle import pyoracle as O
    n, L, rc, nose, a, b = CASES[name]
    com, q = lattice_start(n, L, 20261202 + (1 if n == 216 else 2))
    md = O.MDRun(O.TIP4P, com, q.T.copy(), L, rc, 0.5, 0.002, fast=True, nose=nose)
    md.vreset(300.0, 4242)
    md.step(1)
    t0 = time.perf_counter()
    md.step(b - a)
    return n * (b - a) / (time.perf_counter() - t0)


def main():
    quick = "--quick" in sys.argv
    if quick:
        for k, v in CASES.items():
            CASES[k] = v[:4] + (v[4] // 4, v[5] // 4)
    small, big = build_small(), build_big()
    out = {"host": {"nproc": os.cpu_count(), "cpu": open("/proc/cpuinfo").read().split(
        "model name")[1].split("\n")[0].strip(": ")}}
    for name in CASES:
        exe = small if CASES[name][0] <= 2048 else big
        r1 = ref_rate(exe, name)
        r8 = ref_rate(exe, name, procs=8)
        o1 = oracle_rate(name)
        out[name] = {"reference_1proc": r1, "reference_8proc": r8, "oracle_1proc": o1,
                     "oracle_over_reference": o1 / r1}
        print(name, out[name], file=sys.stderr, flush=True)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
