"""Restart files: the Fortran driver g6md writes Mixture's binary restart
(fort.13) record for record, and resumes from the reference's own restart
(fort.12) as the reference does (tests/golden/restart_tip4p_216.npz,
written by tests/golden/make_restart.py from the reference driver)."""
import os
import struct
import subprocess
import tempfile

import numpy as np
import pytest

from tests.golden.configs import parse_logd

pytestmark = pytest.mark.gpu
EXE = os.path.abspath(os.path.join(os.path.dirname(__file__), "..", "genesis6_amd", "lib",
                                   "g6md"))


def records(data):
    """Fortran sequential unformatted records (4-byte length markers)."""
    out, p = [], 0
    while p < len(data):
        (ln,) = struct.unpack("<i", data[p:p + 4])
        assert data[p + 4 + ln:p + 8 + ln] == data[p:p + 4]
        out.append(data[p + 4:p + 4 + ln])
        p += 8 + ln
    return out


def run_g6md(deck, fort12=None):
    with tempfile.TemporaryDirectory() as d:
        if fort12 is not None:
            with open(os.path.join(d, "fort.12"), "wb") as f:
                f.write(fort12)
        r = subprocess.run([EXE, "fp64"], input=deck, capture_output=True, text=True,
                           timeout=120, cwd=d)
        assert r.returncode == 0, r.stderr
        with open(os.path.join(d, "fort.13"), "rb") as f:
            return parse_logd(r.stdout), f.read()


def same_restart(got, ref, n):
    """Tags, record lengths and every setting bit for bit; the thermostat
    and the Gear state of @WTG2 within the trajectory tolerance."""
    g, r = records(got), records(ref)
    assert [len(x) for x in g] == [len(x) for x in r]
    tag = b""
    for a, b in zip(g, r):
        if len(b) == 5 and b[:1] == b"@":
            assert a == b                               # the tag itself
            tag = b
        elif tag == b"@NOS2":
            x, y = np.frombuffer(a, "<f8"), np.frombuffer(b, "<f8")
            assert np.allclose(x, y, rtol=1e-9, atol=1e-12), (x, y)
        elif tag in (b"@BOX3", b"@PCN5"):                  # a moving box: the barostat
            if tag == b"@PCN5":
                assert a[:4] == b[:4]                       # mode
                a, b = a[4:], b[4:]
            x, y = np.frombuffer(a, "<f8"), np.frombuffer(b, "<f8")
            assert np.allclose(x, y, rtol=1e-9, atol=1e-14), (tag, x, y)
        elif tag == b"@WTG2" and len(b) > 4:                # the state after n
            x, y = np.frombuffer(a, "<f8"), np.frombuffer(b, "<f8")
            assert np.abs(x - y).max() < 1e-9, np.abs(x - y).max()
        else:
            assert a == b, (tag, a, b)                  # settings, counts, n


def check_logd(got, ref):
    assert got.shape == ref.shape
    np.testing.assert_array_equal(got[:, 0], ref[:, 0])
    for col, tol in [(1, 1e-10), (4, 1e-9), (5, 1e-10), (12, 1e-10)]:
        err = np.abs(got[:, col] - ref[:, col]).max() / (np.abs(ref[:, col]).max() + 1e-300)
        assert err <= tol, (col, err)


def test_restart_file_written_like_the_reference(golden):
    g = golden("restart_tip4p_216")
    print("g6md run A (10 NVT steps)", flush=True)
    logd, f13 = run_g6md(str(g["deck_a"]))
    check_logd(logd, g["logd_a"])
    same_restart(f13, g["fort13_a"].tobytes(), 216)


def test_resume_from_the_reference_restart(golden):
    g = golden("restart_tip4p_216")
    print("g6md run B (resume from the reference's fort.13)", flush=True)
    logd, f13 = run_g6md(str(g["deck_b"]), fort12=g["fort13_a"].tobytes())
    check_logd(logd, g["logd_b"])
    same_restart(f13, g["fort13_b"].tobytes(), 216)
