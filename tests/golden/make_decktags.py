"""Writes tests/golden/decktags_tip4p_216.npz: the @LOGD rows that the
reference MD driver (Mixture) writes for the 20-step Nose NVT run of
md_tip4p_216_nvt.npz when its deck uses the other forms of the tags
(@NOSE, @VOXN, @BOX3, @BOOK; tests.golden.configs.tag_decks).  The Mixture
binary is built on demand by `make -f oracle/ref.mk mixture` and removed
afterwards.

    python tests/golden/make_decktags.py
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, HERE)
from configs import parse_logd, tag_decks  # noqa: E402
from make_golden import REF, run  # noqa: E402


def main():
    g = np.load(os.path.join(HERE, "md_tip4p_216_nvt.npz"))
    built = not os.path.exists(os.path.join(REF, "Mixture"))
    if built:
        subprocess.run(["make", "-f", os.path.join(ROOT, "oracle", "ref.mk"), "mixture"],
                       cwd=ROOT, check=True, capture_output=True)
    out = {}
    try:
        for name, deck in tag_decks(g).items():
            with tempfile.TemporaryDirectory() as d:
                out[name] = parse_logd(run(os.path.join(REF, "Mixture"), d, deck))
            print(name, out[name].shape)
    finally:
        if built:
            os.remove(os.path.join(REF, "Mixture"))
    np.savez_compressed(os.path.join(HERE, "decktags_tip4p_216.npz"), **out)


if __name__ == "__main__":
    main()
