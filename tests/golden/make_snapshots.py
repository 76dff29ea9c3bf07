"""Writes tests/golden/snapshots_tip4p_216.npz: the coordinate (@SNPO, every
5 steps) and coordinate + velocity (@SNPV, every 10 steps) snapshots that the
reference MD driver (Mixture) writes into its log over 20 Nose NVT steps of
216 TIP4P from the start of md_tip4p_216_nvt.npz (Mixture.F90:2338-2385,
RigidModule.F90:830-884).  Kept as data: per block the tag, step, box,
@FIXC and @ID08 text and the rows; plus the block order.  The Mixture
binary is built on demand by `make -f oracle/ref.mk mixture` and removed
afterwards.

    python tests/golden/make_snapshots.py
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, HERE)
from configs import mixture_deck, parse_logd, parse_snapshots  # noqa: E402
from make_golden import REF, run  # noqa: E402

SNPO, SNPV = 5, 10


def main():
    g = np.load(os.path.join(HERE, "md_tip4p_216_nvt.npz"))
    deck = mixture_deck(g["com"], g["quat"], float(g["L"]), float(g["rc"]), float(g["taper"]),
                        float(g["dt"]), int(g["nsteps"]), int(g["seed"]), tuple(g["nose"]),
                        snpo=SNPO, snpv=SNPV)
    built = not os.path.exists(os.path.join(REF, "Mixture"))
    if built:
        subprocess.run(["make", "-f", os.path.join(ROOT, "oracle", "ref.mk"), "mixture"],
                       cwd=ROOT, check=True, capture_output=True)
    try:
        with tempfile.TemporaryDirectory() as d:
            out = run(os.path.join(REF, "Mixture"), d, deck)
    finally:
        if built:
            os.remove(os.path.join(REF, "Mixture"))
    snaps = parse_snapshots(out)
    logd = parse_logd(out)
    assert np.array_equal(logd, g["logd"]), "the snapshots must not change the run"
    # the order of the output blocks: S @SNPO, V @SNPV, L @LOGD
    tags = [l.strip() for l in out.split("\n") if l.strip() in ("@ASTP", "@LOGD", "@NX4A", "@WTG3")]
    seq = []
    for t in tags:
        if t == "@LOGD":
            seq.append("L")
        elif t == "@NX4A":
            seq.append("S")
        elif t == "@WTG3":
            seq.append("V")
    arrays = {}
    for k, b in enumerate(snaps):
        arrays["b%d_rows" % k] = b["rows"]
    np.savez_compressed(os.path.join(HERE, "snapshots_tip4p_216.npz"), snpo=SNPO, snpv=SNPV,
                        sequence="".join(seq),
                        kind=np.array([b["kind"] for b in snaps]),
                        step=np.array([b["step"] for b in snaps]),
                        box=np.array([b["box"] for b in snaps]),
                        ncmp=np.array([b["ncmp"] for b in snaps]),
                        fixc=np.array([b["fixc"] or "" for b in snaps]),
                        ident=np.array([b["id"] for b in snaps]), **arrays)
    print("blocks", len(snaps), "sequence", "".join(seq))


if __name__ == "__main__":
    main()
