"""NPT Monte Carlo on the GPU against the oracle: the same proposals,
decisions and configurations bit for bit, the energy bookkeeping, the
surrogate's features, and delayed acceptance reduced to plain MC."""
import numpy as np
import pytest

from genesis6_amd import g6
from oracle import npt_oracle as NO
from oracle import pyoracle as O
from tests.golden.configs import lattice_start

pytestmark = pytest.mark.gpu

RC, TAPER, VOXA = 7.0, 0.5, 5.0


def box(n, rho):
    return (n * 18.015 / (6.0221e23 * rho) * 1e24) ** (1.0 / 3.0)


def system(n, rho, seed, steps=0, mid=g6.TIP4P):
    L = box(n, rho)
    s = g6.System(mid, n, L, RC, TAPER, 0.002, voxa=VOXA, prec=g6.FP64)
    com, q = lattice_start(n, L, seed)
    s.upload(com, q)
    if steps:
        s.set_nose(250.0, 250.0 * 2500.0)
        s.vreset(250.0, seed)
        s.step(steps)
    c, q5, _, _ = s.download()
    return s, c, q5[0].copy(), L


@pytest.mark.parametrize("mid,pvol", [(g6.TIP4P, 0.0), (g6.TIP4P, 0.3), (g6.SPCE, 0.3),
                                      (g6.NVDE, 0.3), (g6.ST2, 0.3)])
def test_plain_moves_match_oracle(mid, pvol):
    n = 128
    s, com, q, L = system(n, 1.0, 41, steps=40, mid=mid)
    print("npt model=%d n=%d L=%.4f pvol=%g" % (mid, n, L, pvol), flush=True)
    args = (250.0, 1.0, 1234567, 0.15, 8.0, 0.005, pvol, VOXA)
    s.npt_begin(*args)
    _, _, L0, U0 = s.npt_state()
    orc = O.NptRun(mid, com, q, L, RC, TAPER, VOXA, 250.0, 1.0, 1234567, 0.15, 8.0,
                   0.005, pvol, energy=U0)
    total = 0
    for rnd in range(3):
        st = s.npt_run(150)
        total += int(orc.run(150).sum())
        c, qq, Lg, Ug = s.npt_state()
        print("round", rnd, st, flush=True)
        assert st["accepted"] == total == orc.counts[1]
        assert st["proposals"] == orc.counts[0] and st["volume_proposals"] == orc.counts[2]
        assert st["volume_accepted"] == orc.counts[3]
        assert Lg == orc.L
        assert np.array_equal(c, orc.com) and np.array_equal(qq, orc.quat)
        assert abs(Ug - orc.U) < 1e-10 * abs(orc.U)
    assert st["accepted"] > 0
    # the incremental energy against a fresh PairInfo sum
    fresh = s.npt_energy()
    assert abs(Ug - fresh) < 1e-10 * abs(fresh)


def test_features_and_molecule_energies_match_oracle():
    n = 128
    s, com, q, L = system(n, 1.0, 43, steps=30)
    s.npt_begin(250.0, 1.0, 5, 0.15, 8.0, 0.005, 0.0, VOXA)
    x, e = s.npt_features()
    # site positions of the oracle's placement
    sites, _, _ = O.setsites(O.TIP4P, com, q)
    xs, ys, zs = (sites[:, k].reshape(n, 5) for k in range(3))
    Os = np.stack([xs[:, 0], ys[:, 0], zs[:, 0]], 1)
    Hs = np.stack([np.stack([xs[:, h], ys[:, h], zs[:, h]], 1) for h in (2, 3)], 1)
    xo = NO.features(Os, Hs, L)
    assert np.abs(x - xo).max() < 1e-4 * max(1.0, np.abs(xo).max())
    ep, _ = O.mc_pairinfo(O.TIP4P, com, q, L, RC, TAPER)
    eo = 0.5 * ep.sum(1)
    assert np.abs(e - eo).max() < 1e-10 * np.abs(eo).max()


def test_delayed_acceptance_with_a_zero_network_is_plain_mc():
    """With the network's output identically zero every block proposal
    passes the screen, so K = 1 rounds make exactly the plain chain's
    decisions on the same proposals (block moves only)."""
    n = 128
    s1, _, _, _ = system(n, 1.0, 47, steps=20)
    s2, _, _, _ = system(n, 1.0, 47, steps=20)
    for s in (s1, s2):
        s.npt_begin(250.0, 1.0, 99, 0.15, 8.0, 0.005, 0.0, VOXA)
    net = g6.Mlp(np.zeros(g6.MLP_NPARAM, np.float32))
    a = s1.npt_run(200)
    b = s2.npt_run(200, mlp=net, K=1)
    print(a, b, flush=True)
    # every proposal but the void ones passes a zero network and goes to the
    # exact decision; void proposals are rejected in both chains
    assert b["rounds"] == 200 and a["void"] == b["void"]
    assert b["surrogate_passes"] == b["exact_evaluations"] == 200 - b["void"]
    assert a["accepted"] == b["accepted"] and a["next_index"] == b["next_index"]
    for u, v in zip(s1.npt_state(), s2.npt_state()):
        assert np.array_equal(u, v)


def test_npt_state_is_freed_with_its_system():
    """A finalised system takes its NPT state with it: a new system (which
    may reuse the old one's host address) starts from its own configuration."""
    for n in (200, 100, 150):
        s, com, q, L = system(n, 1.0, 50 + n)
        s.npt_begin(250.0, 1.0, 5, 0.15, 8.0, 0.005, 0.2, VOXA)
        c, qq, Lg, U = s.npt_state()
        assert c.shape == (n, 3) and np.array_equal(c, com) and Lg == L
        _, fresh = O.mc_pairinfo(O.TIP4P, com, q, L, RC, TAPER)
        assert abs(U - fresh) < 1e-10 * abs(fresh)
        s.npt_run(50)
        s.close()


def c5_start(seed):
    """1024 TIP4P at 1.0 g/cm3 (L = 31.289 A), rc 9 A switched over 0.5 A,
    after 200 FP64 Nose steps at 250 K from the seeded lattice."""
    n, L = 1024, 31.289
    s = g6.System(g6.TIP4P, n, L, 9.0, TAPER, 0.002, voxa=9.0, prec=g6.FP64)
    com, q = lattice_start(n, L, seed)
    s.upload(com, q)
    s.set_nose(250.0, 250.0 * 2500.0)
    s.vreset(250.0, seed)
    s.step(200)
    c, q5, _, _ = s.download()
    return s, c, q5[0].copy(), L


def test_c5_plain_chain_matches_oracle():
    """The NPT workload at full size (1024 TIP4P, 250 K, 1 bar, rc 9 A, block
    cells of the 9 A Address_Assign grid, volume moves with probability
    1/28): 300 plain moves equal the oracle's chain bit for bit."""
    s, com, q, L = c5_start(61)
    args = (250.0, 1.0, 20261207, 0.15, 8.0, 0.005, 1.0 / 28.0, 9.0)
    s.npt_begin(*args)
    _, _, L0, U0 = s.npt_state()
    orc = O.NptRun(g6.TIP4P, com, q, L, 9.0, TAPER, 9.0, 250.0, 1.0, 20261207, 0.15, 8.0,
                   0.005, 1.0 / 28.0, energy=U0)
    for rnd in range(3):
        print("c5 plain round", rnd, flush=True)
        st = s.npt_run(100)
        orc.run(100)
        c, qq, Lg, Ug = s.npt_state()
        assert st["proposals"] == orc.counts[0] and st["accepted"] == orc.counts[1]
        assert st["volume_proposals"] == orc.counts[2]
        assert st["volume_accepted"] == orc.counts[3]
        assert Lg == orc.L
        assert np.array_equal(c, orc.com) and np.array_equal(qq, orc.quat)
        assert abs(Ug - orc.U) < 1e-10 * abs(orc.U)
    print("c5 plain", st, flush=True)
    assert st["accepted"] > 0 and st["volume_proposals"] > 0
    fresh = s.npt_energy()
    assert abs(Ug - fresh) < 1e-10 * abs(fresh)
    s.close()


def test_c5_zero_network_rounds_of_32_match_oracle():
    """Delayed acceptance at the workload's K = 32 with a network whose output
    is identically zero (block moves only, displaced <= 0.05 A and turned
    <= 2 deg so that some are accepted): every non-void proposal passes the
    screen, so round r decides the first non-void proposal of indices
    32 r .. 32 r + 31 on the exact energy.  The oracle's plain chain run on
    exactly those proposals makes the same decisions, bit for bit."""
    s, com, q, L = c5_start(62)
    args = (250.0, 1.0, 99, 0.05, 2.0, 0.005, 0.0, 9.0)
    s.npt_begin(*args)
    _, _, _, U0 = s.npt_state()
    net = g6.Mlp(np.zeros(g6.MLP_NPARAM, np.float32))
    orc = O.NptRun(g6.TIP4P, com, q, L, 9.0, TAPER, 9.0, 250.0, 1.0, 99, 0.05, 2.0, 0.005,
                   0.0, energy=U0)
    R = 100
    b = s.npt_run(R, mlp=net, K=32)
    print("c5 zero network", b, flush=True)
    assert b["rounds"] == R and b["next_index"] == 32 * R
    decided = 0
    for r in range(R):
        orc.index = 32 * r
        while orc.index < 32 * (r + 1):
            void = orc.counts[4]
            orc.run(1)
            if orc.counts[4] == void:
                decided += 1
                break
    assert b["surrogate_passes"] == b["exact_evaluations"] == decided
    assert b["accepted"] == orc.counts[1] and b["accepted"] > 0
    c, qq, Lg, Ug = s.npt_state()
    assert np.array_equal(c, orc.com) and np.array_equal(qq, orc.quat) and Lg == orc.L
    assert abs(Ug - orc.U) < 1e-10 * abs(orc.U)
    net.close()
    s.close()
