"""g6_eval_forces_noalloc, the entry point that allocates nothing, against
g6_eval_forces through the C ABI: the same forces, torques and thermo record
bit for bit, on two calls in a row (the second on a smaller system, so the
shared device array holds the first call's larger contents), for the rows
of both directions and the cell tiles."""
import ctypes as C
import os

import numpy as np
import pytest

from genesis6_amd import g6
from tests.golden.configs import lattice_start

pytestmark = pytest.mark.gpu


def call(fn, s):
    n = s.n
    F = np.zeros(3 * n)
    T = np.zeros(3 * n)
    th = np.zeros(g6.THERMO_LEN)
    rc = fn(s.h, F, T, th)
    return rc, F, T, th


@pytest.mark.parametrize("tiles", [0, 1])
def test_noalloc_equals_eval_forces(tiles):
    L = g6.lib()
    old = os.environ.get("G6_TILE_MIN")
    os.environ["G6_TILE_MIN"] = "0"
    os.environ["G6_TILES"] = str(tiles)
    try:
        cases = [(4096, 49.669), (1024, 31.289)]  # second call smaller
        systems = []
        for n, Lb in cases:
            print("noalloc tiles=%d n=%d" % (tiles, n), flush=True)
            s = g6.System(g6.TIP4P, n, Lb, 9.0, 0.5, 0.002, prec=g6.FP32, skin=1.0)
            com, q = lattice_start(n, Lb, 20261202 + n)
            s.upload(com, q)
            assert s.md_kernel()[0] == tiles
            systems.append(s)
        for s in systems:
            r1, F1, T1, th1 = call(L.g6_eval_forces_noalloc, s)
            r0, F0, T0, th0 = call(L.g6_eval_forces, s)
            assert r0 == 0 and r1 == 0
            assert np.array_equal(F0, F1) and np.array_equal(T0, T1) and np.array_equal(th0, th1)
            assert np.abs(F0).max() > 0 and th0[14] > 0
        # a repeat on the first system after the smaller one: still equal
        r2, F2, T2, th2 = call(L.g6_eval_forces_noalloc, systems[0])
        r0, F0, T0, th0 = call(L.g6_eval_forces, systems[0])
        assert r2 == 0 and np.array_equal(F0, F2) and np.array_equal(T0, T2)
        for s in systems:
            s.close()
    finally:
        os.environ.pop("G6_TILES", None)
        if old is None:
            os.environ.pop("G6_TILE_MIN", None)
        else:
            os.environ["G6_TILE_MIN"] = old


def test_noalloc_rejects_systems_beyond_its_array():
    n, Lb = 400000, 228.9   # 6 n doubles exceed the 16 MiB array
    s = g6.System(g6.SPCE, n, Lb, 9.0, 0.5, 0.002, prec=g6.FP32, skin=1.0)
    rc, *_ = call(g6.lib().g6_eval_forces_noalloc, s)
    assert rc == -3  # G6_ECAPACITY
    s.close()
