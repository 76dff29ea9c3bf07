"""CPU-side checks of the product library: it loads, exports every symbol
that include/g6gpu.h declares, and reports the absence of a device as an
error (no silent CPU fallback)."""
import ctypes
import os
import re

import numpy as np
import pytest

from genesis6_amd import g6

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "g6gpu.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\bint\s+(g6_\w+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 20
    L = g6.lib()
    for nm in names:
        assert hasattr(L, nm), nm


def test_model_site_counts():
    assert [g6.lib().g6_model_nsite(m) for m in (0, 1, 2)] == [5, 6, 6]


def test_lattice_start_is_uniform_and_deterministic():
    c1, q1 = g6.lattice_start(4096, 49.669, 20261204)
    c2, q2 = g6.lattice_start(4096, 49.669, 20261204)
    assert np.array_equal(c1, c2) and np.array_equal(q1, q2)
    assert np.allclose((q1 ** 2).sum(1), 1.0, atol=1e-14)
    assert c1.min() > 0 and c1.max() < 49.669
    # occupied lattice sites are distinct
    assert len({tuple(np.round(r, 6)) for r in c1}) == 4096


def test_no_device_is_an_error_not_a_fallback():
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a device is present")
    except ImportError:
        pass
    with pytest.raises(g6.G6Error):
        g6.System(g6.TIP4P, 216, 18.626, 8.5, 0.5, 0.002)
