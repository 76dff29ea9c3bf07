"""The NPT Monte Carlo oracle: its proposal stream against the published
Philox4x32-10 vectors, the volume move against the closed-form isobaric
ideal-gas law, and the incremental energy against a fresh PairInfo sum."""
import numpy as np
from scipy.special import digamma

from oracle import pyoracle as O
from tests.golden.configs import lattice_start


def test_philox_known_answers():
    # Random123 known-answer vectors for philox4x32-10
    assert O.philox10([0, 0, 0, 0], [0, 0]) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert O.philox10([0xffffffff] * 4, [0xffffffff] * 2) == [
        0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert O.philox10([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344],
                      [0xa4093822, 0x299f31d0]) == [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


def test_volume_moves_sample_the_isobaric_ideal_gas():
    """Without interactions (rc below every COM distance) the ln V walk with
    -beta P dV + (N+1) dlnV samples p(V) ~ V^N exp(-beta P V): V is Gamma
    distributed and E[ln V] = digamma(N+1) + ln(T / P).  A dropped +1 in the
    Jacobian moves the mean by 0.5, a wrong PV sign makes V run away."""
    n, L = 2, 20.0
    com = np.array([[0.0, 0.0, 0.0], [10.0, 10.0, 10.0]])
    quat = np.zeros((4, n)); quat[0] = 1.0
    temp, press_bar = 300.0, 5000.0
    run = O.NptRun(O.TIP4P, com, quat, L, 0.01, 0.005, 5.0, temp, press_bar, 99, 0.1, 5.0,
                   0.4, 1.0, energy=0.0)
    press_k = press_bar * 1e5 * 1e-30 / 1.380066e-23
    lnv = []
    for _ in range(400):
        run.run(100)
        lnv.append(3.0 * np.log(run.L))
    lnv = np.array(lnv[40:])
    expect = digamma(n + 1) + np.log(temp / press_k)
    # 360 samples 100 moves apart; std of ln V is sqrt(trigamma(3)) = 0.63
    assert abs(lnv.mean() - expect) < 0.12, (lnv.mean(), expect)
    assert run.U == 0.0 and run.counts[3] > 0


def test_incremental_energy_equals_a_fresh_sum():
    n, L = 64, 12.4
    com, q = lattice_start(n, L, 5)
    run = O.NptRun(O.TIP4P, com, q.T.copy(), L, 5.5, 0.5, 4.0, 250.0, 1.0, 7, 0.15, 8.0,
                   0.005, 0.2)
    acc = run.run(120)
    _, fresh = O.mc_pairinfo(O.TIP4P, run.com, run.quat, run.L, 5.5, 0.5)
    assert acc.sum() > 0 and run.counts[2] > 0
    assert abs(run.U - fresh) < 1e-9 * abs(fresh)


def test_block_moves_alone_sample_the_ideal_gas():
    """Detailed balance and ergodicity of the block move.  Without
    interactions every molecule position is equally likely, so the fraction
    of molecule pairs sharing a cell of any fixed grid is 1/ncell.  A block
    move whose member may leave the cell has no reverse move (it picks a
    different member set) and drifts above it (0.141 for 1/8 here); a block
    grid that never shifts keeps the start's occupancy for ever.  The chain
    starts with every molecule in one cell."""
    n, L, voxa, nd = 16, 8.0, 4.0, 2
    rng = np.random.default_rng(3)
    com = rng.uniform(-1.5, 1.5, (n, 3))
    quat = np.zeros((4, n)); quat[0] = 1.0
    run = O.NptRun(O.TIP4P, com, quat, L, 0.01, 0.005, voxa, 300.0, 1.0, 7, 1.6, 5.0,
                   0.0, 0.0, energy=0.0)
    frac = []
    for _ in range(1500):
        run.run(50)
        c = run.com - L * np.round(run.com / L)
        j = np.round(c / (L / nd)).astype(int) % nd
        cnt = np.bincount((j[:, 2] * nd + j[:, 1]) * nd + j[:, 0], minlength=nd ** 3)
        frac.append((cnt * (cnt - 1)).sum() / (n * (n - 1)))
    frac = np.array(frac[200:])
    err = frac.reshape(13, -1).mean(1).std() / np.sqrt(13)  # batch means
    assert run.counts[1] > 0 and run.counts[4] > 0  # accepted and void block moves
    assert abs(frac.mean() - 1.0 / nd ** 3) < max(4 * err, 0.006), (frac.mean(), err)
