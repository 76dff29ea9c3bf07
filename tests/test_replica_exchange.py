"""Replica exchange on the host: the exchange graph order, the oracle's swap
rule against the closed-form stationary law, and the library's host rule
(g6_rx_trials, no device needed) against the oracle, decision for decision.

Reference: replica_exchange_trial (Replica.F90:7-73), the exchange loop of
Mixture.F90:1605-1622, graph_read / graph_perfect / graph2pair
(Graph.F90:34-52, 93-127)."""
import math

import numpy as np

from genesis6_amd import g6
from oracle import pyoracle as O

C = 10.0 * 0.120273  # I2J * J2K: internal energy units to K


def test_rxpr_chain_bonds_in_graph2pair_order():
    # graph_read links both directions per deck line; graph2pair walks node
    # 1..n and each node's neighbours in link order
    a, b = g6.rxpr_bonds(4, [(0, 1), (1, 2), (2, 3)])
    assert list(zip(a, b)) == [(1, 2), (2, 1), (2, 3), (3, 2), (3, 4), (4, 3)]
    a, b = g6.rxpr_bonds(3, None)  # graph_perfect
    assert list(zip(a, b)) == [(1, 2), (1, 3), (2, 1), (2, 3), (3, 1), (3, 2)]


def test_oracle_swap_rule_has_the_boltzmann_stationary_law():
    """Two replicas, energies fixed to the nodes: the chain over the two
    beta assignments spends the fraction w_B / (w_A + w_B) of its trials in
    the swapped one, w = exp(-(E0 b0 + E1 b1) I2J J2K).  A wrong sign or a
    dropped unit factor moves the fraction far outside the tolerance."""
    e = np.array([-1862.0, -1000.0])
    b0 = np.array([1 / 250.0, 1 / 300.0])
    a, b = g6.rxpr_bonds(2, [(0, 1)])
    rx = O.RxRun(2, a, b, 1, 4357)
    beta, inB = b0.copy(), 0
    ntr = 40000
    for _ in range(ntr):
        beta, _ = rx.trials(e, beta)
        inB += beta[0] != b0[0]
    wa = -(e[0] * b0[0] + e[1] * b0[1]) * C
    wb = -(e[0] * b0[1] + e[1] * b0[0]) * C
    x = math.exp(wb - wa)
    assert 0.2 < x < 0.8  # a case where the rule matters
    expect = x / (1 + x)
    assert abs(inB / ntr - expect) < 0.02, (inB / ntr, expect)


def test_oracle_trial_count_and_random_draws():
    """Every trial draws the bond; only trials with acceptance0 >=
    acceptance1 draw a second number."""
    a, b = g6.rxpr_bonds(4, [(0, 1), (1, 2), (2, 3)])
    e = np.array([-5000.0, -5100.0, -5200.0, -5300.0])
    beta = 1.0 / np.array([250.0, 260.0, 270.0, 280.0])
    rx = O.RxRun(4, a, b, 25, 99)
    rx.trials(e, beta)
    assert 25 <= rx.pos <= 50


def test_library_host_rule_matches_oracle():
    rng = np.random.default_rng(5)
    nrep = 16
    temps = 220.0 * (320.0 / 220.0) ** (np.arange(nrep) / 15.0)
    a, b = g6.rxpr_bonds(nrep, [(k, k + 1) for k in range(nrep - 1)])
    rx = g6.ReplicaExchange(nrep, (a, b), 30, 20261299)
    orc = O.RxRun(nrep, a, b, 30, 20261299)
    beta = 1.0 / temps
    for call in range(40):
        # energies near the ladder's mean potential energy at each temperature
        e = -9800.0 + 4.0 * temps + rng.normal(0, 6.0, nrep)
        bl, acc = rx.trials(e, beta)
        bo, acco = orc.trials(e, beta)
        assert acc == acco and np.array_equal(bl, bo), call
        assert list(rx.nodes()) == orc.nic
        beta = bl
    assert sorted(1.0 / beta) == sorted(1.0 / (1.0 / temps))


def test_oracle_umbrella_swap_rule_has_the_stationary_law():
    """Two MC nodes at one temperature, fixed energies and order parameters,
    different umbrellas (Replica.F90:30-49): the umbrellas travel between
    the nodes, and the chain spends the fraction w_B / (w_A + w_B) of its
    trials with them swapped, w = exp(-beta (E0 + U_0(op0) + E1 + U_1(op1))
    I2J J2K).  Evaluating the swapped state without exchanging the umbrellas
    gives one half instead."""
    e = np.array([-1000.0, -1000.5])
    beta = np.full(2, 1 / 250.0)
    op = np.array([[0.30], [0.55]])
    center0 = np.array([[0.3], [0.6]])
    coeffi = np.array([[800.0], [800.0]])
    a, b = g6.rxpr_bonds(2, [(0, 1)])
    rx = O.RxRun(2, a, b, 1, 77)
    center, k, inB = center0.copy(), coeffi.copy(), 0
    ntr = 40000
    for _ in range(ntr):
        beta, center, k, _ = rx.trials(e, beta, op, center, k)
        inB += center[0, 0] != center0[0, 0]

    def w(c):
        u = k[0, 0] * (op[:, 0] - c[:, 0]) ** 2
        return -(beta[0] * (e[0] + u[0]) + beta[1] * (e[1] + u[1])) * C
    x = math.exp(w(center0[::-1]) - w(center0))
    assert 0.1 < x < 0.6
    assert abs(inB / ntr - x / (1 + x)) < 0.02, (inB / ntr, x / (1 + x))


def test_library_umbrella_rule_matches_oracle():
    """g6_rx_trials_umbrella against the oracle, decision for decision:
    eight nodes, two umbrellas each, a temperature ladder."""
    rng = np.random.default_rng(9)
    nrep, nu = 8, 2
    temps = 240.0 + 10.0 * np.arange(nrep)
    a, b = g6.rxpr_bonds(nrep, [(k, k + 1) for k in range(nrep - 1)])
    rx = g6.ReplicaExchange(nrep, (a, b), 20, 4242)
    orc = O.RxRun(nrep, a, b, 20, 4242)
    beta = 1.0 / temps
    center = np.stack([np.linspace(0.2, 0.9, nrep), np.linspace(-3e5, -2.5e5, nrep)], 1)
    coeffi = np.stack([np.full(nrep, 500.0), np.full(nrep, 2e-6)], 1)
    cl, kl, co, ko = center.copy(), coeffi.copy(), center.copy(), coeffi.copy()
    bl = bo = beta
    total = 0
    for call in range(30):
        e = -2.8e5 + 60.0 * temps + rng.normal(0, 40.0, nrep)
        op = np.stack([cl[:, 0] + rng.normal(0, 0.05, nrep), e], 1)
        bl, cl, kl, acc = rx.trials_umbrella(e, bl, op, cl, kl)
        bo, co, ko, acco = orc.trials(e, bo, op, co, ko)
        assert acc == acco and np.array_equal(bl, bo), call
        assert np.array_equal(cl, co) and np.array_equal(kl, ko), call
        assert list(rx.nodes()) == orc.nic
        total += acc
    assert total > 0


def test_library_umbrella_rule_without_umbrellas_is_the_plain_rule():
    a, b = g6.rxpr_bonds(4, [(0, 1), (1, 2), (2, 3)])
    e = np.array([-5000.0, -5003.0, -5006.0, -5009.0])
    beta = 1.0 / np.array([250.0, 255.0, 260.0, 265.0])
    r1 = g6.ReplicaExchange(4, (a, b), 25, 31)
    r2 = g6.ReplicaExchange(4, (a, b), 25, 31)
    b1, acc1 = r1.trials(e, beta)
    b2, _, _, acc2 = r2.trials_umbrella(e, beta, np.zeros((4, 0)), np.zeros((4, 0)),
                                        np.zeros((4, 0)))
    assert acc1 == acc2 and np.array_equal(b1, b2)
