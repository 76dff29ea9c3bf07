"""The Fortran replica-exchange MD driver g6rx (the replica branch of
Mixture with the MDReplica tags): four TIP4P replicas at 250-280 K from one
deck, exchanges every @INNR steps over an @RXPR chain.  Each node's log
(@MDRT temperatures, @LOGD rows) equals the same schedule run through the
Python binding (one ladder, one exchange group whose decisions are pinned to
the oracle elsewhere)."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from genesis6_amd import g6
from tests.golden.configs import lattice_start, mixture_deck, parse_logd

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "genesis6_amd", "lib", "g6rx")
N, L, RC, TAPER, DT = 128, 15.6446, 7.0, 0.5, 0.002
TEMPS = [250.0, 260.0, 270.0, 280.0]
STEPS, INNR, NLOG, NRXG, SEED = 40, 10, 5, 30, 777


def parse_node_log(text):
    lines = text.split("\n")
    mdrt = [float(lines[i + 1]) for i, l in enumerate(lines) if l.strip() == "@MDRT"]
    return mdrt, parse_logd(text)


def reference_schedule(com, q):
    """The driver's loop through the Python binding: chunks that stop at the
    log steps and before each exchange point."""
    systems = []
    for T in TEMPS:
        s = g6.System(g6.TIP4P, N, L, RC, TAPER, DT)
        s.set_fast(False)   # a deck without @FAST runs StdForce
        s.upload(com, q)
        s.set_nose(T, T * 2500.0)
        s.vreset(300.0, 11)
        systems.append(s)
    lad = g6.Ladder(systems)
    a, b = g6.rxpr_bonds(4, [(0, 1), (1, 2), (2, 3)])
    rx = g6.ReplicaExchange(4, (a, b), NRXG, SEED, systems=systems)
    logs = [[] for _ in TEMPS]
    mdrt = [[T] for T in TEMPS]
    done, intercept = 0, INNR
    while done < STEPS:
        chunk = min(STEPS - done, 1 + (-done) % NLOG)
        if intercept - 1 > done:
            chunk = min(chunk, intercept - 1 - done)
        lad.step(chunk)
        done += chunk
        th = lad.thermo()
        if (done - 1) % NLOG == 0:
            for k in range(4):
                logs[k].append((done - 1, th[k, 3], th[k, 13]))
        if done == intercept - 1 and intercept <= STEPS:
            intercept += INNR
            rx.exchange(th)
            for k, s in enumerate(systems):
                if s.nose()[0] != mdrt[k][-1]:
                    mdrt[k].append(s.nose()[0])
    for s in systems:
        s.close()
    return mdrt, logs


def test_g6rx_node_logs_follow_the_ladder_and_exchange_schedule():
    com, q = lattice_start(N, L, 31)
    deck = mixture_deck(com, q, L, RC, TAPER, DT, STEPS, 11, nlog=NLOG)
    deck = deck.replace("@ID08", "@MPIN\n4\nn0\nn1\nn2\nn3\n@SEED\n%d\n@INNR\n%d\n@NRXG\n%d\n"
                        "@RXPR\n4\n0 1\n1 2\n2 3\n-1 -1\n@ID08" % (SEED, INNR, NRXG), 1)
    print("g6rx 4 replicas", flush=True)
    with tempfile.TemporaryDirectory() as d:
        for k, T in enumerate(TEMPS):
            with open(os.path.join(d, "n%d" % k), "w") as f:
                f.write("@NOS2\n%r %r 0 0 0 0 0 0 0\n" % (T, T * 2500.0))
        r = subprocess.run([EXE, "fp64"], input=deck, capture_output=True, text=True,
                           timeout=300, cwd=d)
        assert r.returncode == 0, r.stderr
        got = [parse_node_log(open(os.path.join(d, "n%d.log" % k)).read()) for k in range(4)]
    mdrt, logs = reference_schedule(com, q)
    assert sum(len(m) for m in mdrt) > 4  # some exchange was accepted
    for k in range(4):
        assert got[k][0] == mdrt[k]
        rows = got[k][1]
        ref = np.array(logs[k])
        assert np.array_equal(rows[:, 0], ref[:, 0])
        # T and the conserved H as printed (17 significant digits)
        assert np.allclose(rows[:, 1], ref[:, 1], rtol=1e-15, atol=0)
        assert np.allclose(rows[:, 12], ref[:, 2], rtol=1e-15, atol=0)
