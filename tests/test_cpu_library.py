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


def test_fortran_interface_binds_every_declared_symbol():
    """The ISO_C_BINDING module declares one interface per C entry point,
    under the C name."""
    src = open(os.path.join(ROOT, "genesis6_amd", "fortran", "g6gpu.f90")).read()
    bound = sorted(set(re.findall(r'bind\(C,\s*name="(g6_\w+)"\)', src)))
    assert bound == declared_functions()


def test_fortran_driver_reports_a_missing_device():
    """g6md is built and reads a deck; without a device it stops with the
    library's error code instead of running anything on the host."""
    import subprocess
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a device is present")
    except ImportError:
        pass
    exe = os.path.join(ROOT, "genesis6_amd", "lib", "g6md")
    if not os.path.exists(exe):
        pytest.skip("g6md not built (python -m genesis6_amd.build)")
    deck = "@BXLA\n18.626\n@ID08\nTIP4P\n@NX4A\n2\n0 0 0 1 0 0 0\n3 0 0 1 0 0 0\n"
    r = subprocess.run([exe], input=deck, capture_output=True, text=True, timeout=60)
    assert r.returncode != 0
    assert "g6_init_box3 failed with code -4" in r.stderr
