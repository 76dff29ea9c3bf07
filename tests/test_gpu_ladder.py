"""Replica ladders on the GPU: one launch per MD kernel for several systems
gives each system exactly what g6_step gives it alone; the thermostat
retarget of an exchange (Mixture.F90:1630-1653) matches the oracle; a full
exchange between ladder replicas makes the oracle's decisions."""
import numpy as np
import pytest

from genesis6_amd import g6
from oracle import pyoracle as O
from tests.golden.configs import lattice_start

pytestmark = pytest.mark.gpu

RC, TAPER, DT = 7.0, 0.5, 0.002


def box(n, rho):
    """Cubic box edge (A) for n waters at rho g/cm3 (M = 18.015 g/mol)."""
    return (n * 18.015 / (6.0221e23 * rho) * 1e24) ** (1.0 / 3.0)


def make(n, rho, seed, temp, prec, skin, nose=True, mid=g6.TIP4P):
    # methanol (M 32.04) at 0.35 times the water density of the case: its
    # lattice starts at liquid density put methyl groups in contact, and the
    # 2 fs steps blow such molecules apart (a parity test needs no liquid)
    L = box(n, rho * 0.35 * 18.015 / 32.04) if mid == g6.OPLSMEOH else box(n, rho)
    print("system model=%d n=%d L=%.4f T=%g prec=%d skin=%g nose=%d"
          % (mid, n, L, temp, prec, skin, nose), flush=True)
    s = g6.System(mid, n, L, RC, TAPER, DT, prec=prec, skin=skin)
    com, q = lattice_start(n, L, seed)
    s.upload(com, q)
    if nose:
        s.set_nose(temp, temp * 2500.0)
    s.vreset(temp, seed)
    return s


SPECS = [(128, 0.92, 11, 250.0, 0.5, True), (250, 1.0, 12, 280.0, 1.0, True),
         (300, 1.05, 13, 300.0, 0.0, False), (128, 1.0, 14, 320.0, 0.5, True)]


@pytest.mark.parametrize("prec,mid", [(g6.FP32, g6.TIP4P), (g6.FP64, g6.TIP4P),
                                      (g6.FP32, g6.NVDE), (g6.FP32, g6.SPCE),
                                      (g6.FP32, g6.OPLSMEOH)])
def test_ladder_steps_equal_single_system_steps_bitwise(prec, mid):
    lad_sys = [make(n, r, sd, T, prec, sk, ns, mid) for n, r, sd, T, sk, ns in SPECS]
    solo = [make(n, r, sd, T, prec, sk, ns, mid) for n, r, sd, T, sk, ns in SPECS]
    lad = g6.Ladder(lad_sys)
    for k in (7, 60, 1):
        print("steps", k, flush=True)
        lad.step(k)
        for s in solo:
            s.step(k)
        th = lad.thermo()
        for r, (a, b) in enumerate(zip(lad_sys, solo)):
            for x, y in zip(a.download(), b.download()):
                assert np.array_equal(x, y), (k, r)
            assert np.array_equal(th[r], b.thermo()), (k, r)
            assert np.array_equal(a.nose(), b.nose()), (k, r)
    # the systems stay usable on their own after ladder steps
    lad_sys[1].step(3)
    solo[1].step(3)
    for x, y in zip(lad_sys[1].download(), solo[1].download()):
        assert np.array_equal(x, y)


def test_set_temperature_matches_oracle():
    n, L = 128, box(128, 1.0)
    com, q = lattice_start(n, L, 21)
    s = g6.System(g6.TIP4P, n, L, RC, TAPER, DT, prec=g6.FP64)
    s.upload(com, q)
    s.set_nose(260.0, 260.0 * 2500.0)
    s.vreset(260.0, 21)
    md = O.MDRun(O.TIP4P, com, q.T.copy(), L, RC, TAPER, DT, fast=True,
                 nose=(260.0, 260.0 * 2500.0, 0.0))
    md.vreset(260.0, 21)
    s.step(4)
    md.step(4)
    _, _, dx0, w0 = s.download()
    z0 = s.nose()
    s.set_temperature(281.5)
    md.set_temperature(281.5)
    _, _, dx1, w1 = s.download()
    z1 = s.nose()
    ratio = np.sqrt(281.5 / z0[0])
    assert np.array_equal(dx1[0], dx0[0] * ratio)
    assert np.array_equal(w1[0], w0[0] * ratio)
    assert np.array_equal(dx1[1:], dx0[1:]) and np.array_equal(w1[1:], w0[1:])
    assert z1[0] == 281.5 and not z1[3:8].any()
    assert z1[1] == z0[1] and z1[2] == z0[2] and z1[8] == z0[8]
    # the retargeted run continues as the oracle's
    s.step(5)
    md.step(5)
    c, _, _, _ = s.download()
    assert np.abs(c - md.com).max() < 1e-10
    assert abs(s.thermo()[3] - md.s.temperature) < 1e-9 * md.s.temperature


def test_ladder_exchange_makes_the_oracle_decisions():
    temps = [250.0, 251.0, 252.0, 253.5]
    sys_ = [make(128, 1.0, 30 + k, T, g6.FP32, 0.5) for k, T in enumerate(temps)]
    lad = g6.Ladder(sys_)
    a, b = g6.rxpr_bonds(4, [(0, 1), (1, 2), (2, 3)])
    rx = g6.ReplicaExchange(4, (a, b), 30, 777, systems=sys_)
    orc = O.RxRun(4, a, b, 30, 777)
    total = 0
    for rnd in range(3):
        print("exchange round", rnd, flush=True)
        lad.step(20)
        th = lad.thermo()
        before = [s.download() for s in sys_]
        tnow = th[:, 29].copy()
        assert np.array_equal(tnow, np.array([s.nose()[0] for s in sys_]))
        beta_o, acc_o = orc.trials(th[:, 0], 1.0 / tnow)
        acc = rx.exchange(th)
        assert acc == acc_o
        total += acc
        assert list(rx.nodes()) == orc.nic
        for r, s in enumerate(sys_):
            z = s.nose()
            _, _, dx, w = s.download()
            if beta_o[r] != 1.0 / tnow[r]:
                tn = 1.0 / beta_o[r]
                assert z[0] == tn and not z[3:8].any()
                ratio = np.sqrt(tn / tnow[r])
                assert np.array_equal(dx[0], before[r][2][0] * ratio)
                assert np.array_equal(w[0], before[r][3][0] * ratio)
            else:
                assert z[0] == tnow[r]
                assert np.array_equal(dx, before[r][2]) and np.array_equal(w, before[r][3])
    assert total > 0  # the rule was exercised on accepted swaps


def test_block_assignment_changes_no_result():
    """The autotuner's block-assignment knob: the flat map gives the same
    states as the XCD map, bit for bit."""
    a = [make(n, r, sd, T, g6.FP32, sk, ns) for n, r, sd, T, sk, ns in SPECS]
    b = [make(n, r, sd, T, g6.FP32, sk, ns) for n, r, sd, T, sk, ns in SPECS]
    la, lb = g6.Ladder(a), g6.Ladder(b)
    lb.set_map(g6.Ladder.MAP_FLAT)
    la.step(25)
    lb.step(25)
    for x, y in zip(a, b):
        for u, v in zip(x.download(), y.download()):
            assert np.array_equal(u, v)
    assert np.array_equal(la.thermo(), lb.thermo())


def test_autotuner_returns_a_measured_candidate():
    from genesis6_amd.autotune import tune_ladder
    specs = SPECS[:2]
    best, table = tune_ladder(lambda skin: [make(n, r, sd, T, g6.FP32, skin, ns)
                                            for n, r, sd, T, _, ns in specs],
                              skins=(0.5, 1.0), steps=20, warmup=5)
    assert len(table) == 4 and best in table
    assert all(r["ms_per_step"] > 0 for r in table)


def test_c4_ladders_full_shape():
    """The 64-replica workload: 4 TIP4P ladders (N 256, 512, 1024, 2048 at
    0.92, 0.97, 1.00, 1.05 g/cm3) of 16 rungs 220-320 K (geometric), rc 9 A
    switched over 0.5 A, FP32 pairs, 1 A skin.  Two 500-step intervals of
    one ladder of all 64 equal 64 systems stepped alone, bit for bit, and
    the exchange between them (30 trials per ladder over the adjacent-rung
    chain) makes the oracle's decisions; the solo systems get the same
    thermostat retargets."""
    rc9 = 9.0
    specs = [(256, 0.92, 20.266), (512, 0.97, 25.088), (1024, 1.00, 31.289),
             (2048, 1.05, 38.786)]
    temps = [220.0 * (320.0 / 220.0) ** (k / 15.0) for k in range(16)]

    def one(n, Lb, T, seed):
        s = g6.System(g6.TIP4P, n, Lb, rc9, TAPER, DT, prec=g6.FP32, skin=1.0)
        com, q = lattice_start(n, Lb, seed)
        s.upload(com, q)
        s.set_nose(T, T * 2500.0)
        s.vreset(T, seed)
        return s

    lad_sys, solo = [], []
    for li, (n, rho, Lb) in enumerate(specs):
        assert abs(box(n, rho) - Lb) < 2e-3
        for k, T in enumerate(temps):
            seed = 20261206 + 100 * li + k
            lad_sys.append(one(n, Lb, T, seed))
            solo.append(one(n, Lb, T, seed))
    print("c4: 64 + 64 systems", flush=True)
    lad = g6.Ladder(lad_sys)
    groups, orcs = [], []
    for li in range(4):
        a, b = g6.rxpr_bonds(16, [(k, k + 1) for k in range(15)])
        groups.append(g6.ReplicaExchange(16, (a, b), 30, 4242 + li,
                                         systems=lad_sys[16 * li:16 * li + 16]))
        orcs.append(O.RxRun(16, a, b, 30, 4242 + li))
    accepted = 0
    for interval in range(2):
        print("c4 interval", interval, flush=True)
        lad.step(500)
        for s in solo:
            s.step(500)
        th = lad.thermo()
        for r, (x, y) in enumerate(zip(lad_sys, solo)):
            for u, v in zip(x.download(), y.download()):
                assert np.array_equal(u, v), (interval, r)
            assert np.array_equal(th[r], y.thermo()), (interval, r)
        for li in range(4):
            sl = slice(16 * li, 16 * li + 16)
            tnow = th[sl, 29].copy()
            beta_o, acc_o = orcs[li].trials(th[sl, 0], 1.0 / tnow)
            acc = groups[li].exchange(th[sl])
            assert acc == acc_o and list(groups[li].nodes()) == orcs[li].nic
            accepted += acc
            for r in range(16):
                if beta_o[r] != 1.0 / tnow[r]:
                    solo[16 * li + r].set_temperature(1.0 / beta_o[r])
                assert lad_sys[16 * li + r].nose()[0] == solo[16 * li + r].nose()[0]
    print("c4 accepted swaps", accepted, flush=True)
    assert accepted > 0
    lad.close()
    for g in groups:
        g.close()
    for s in lad_sys + solo:
        s.close()
