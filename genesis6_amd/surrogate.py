"""Data and training of the per-molecule energy surrogate (host side).

Configurations come from the MD path itself: Nose NVT runs of rigid TIP4P
at a set of temperatures, sampled every few steps.  For each sampled
configuration the device computes the 32 features of every molecule and its
exact FP64 energy e_i = (1/2) sum_j PairInfo(i, j); the network is then
trained on the GPU (g6_mlp_train: BF16 matrix-core forward and backward,
Adam) to map features to e_i.  The sum over molecules is the surrogate of
the configuration energy that screens NPT Monte Carlo moves.
"""
import time

import numpy as np

from . import g6


def md_dataset(n, L, temps, configs_per_run, steps_between, seed, rc=9.0, taper=0.5,
               equil=500, skin=1.0, scales=(1.0,), log=None):
    """(features (m, 32) float32, energies (m,) float64) of configurations
    from FP32-pair Nose NVT MD runs, one per temperature and box edge
    L * scale (the network then sees the density dependence that volume
    moves probe)."""
    runs = [(T, f) for T in temps for f in scales]
    m = len(runs) * configs_per_run * n
    X = np.empty((m, g6.MLP_NIN), np.float32)
    E = np.empty(m, np.float64)
    r = 0
    for k, (T, f) in enumerate(runs):
        t0 = time.perf_counter()
        Lb = L * f
        s = g6.System(g6.TIP4P, n, Lb, rc, taper, 0.002, prec=g6.FP32, skin=skin)
        com, q = g6.lattice_start(n, Lb, seed + k)
        s.upload(com, q)
        s.set_nose(T, T * 2500.0)
        s.vreset(T, seed + k)
        s.step(equil)
        for _ in range(configs_per_run):
            s.step(steps_between)
            s.npt_begin(T, 1.0, 1)          # loads the MD configuration
            x, e = s.npt_features()
            X[r:r + n] = x
            E[r:r + n] = e
            r += n
        s.close()
        if log:
            log("dataset: %g K, L %.3f, %d configurations, %.1f s"
                % (T, Lb, configs_per_run, time.perf_counter() - t0))
    return X, E


def normalisation(X, E):
    """Feature mean and 1/std, target shift and scale (the g6_mlp norm)."""
    mu = X.astype(np.float64).mean(0)
    sd = X.astype(np.float64).std(0)
    sd[sd < 1e-6] = 1e-6
    return np.concatenate([mu, 1.0 / sd, [E.mean(), E.std()]]).astype(np.float32)


def train(X, E, epochs=8, batch=4096, lr=2e-3, seed=11, init_seed=3, log=None):
    """A new network trained on (X, E); returns it and the per-epoch loss
    (mean squared error in units of the target's standard deviation)."""
    net = g6.Mlp(g6.mlp_init_params(init_seed), normalisation(X, E))
    t0 = time.perf_counter()
    losses = net.train(X, E.astype(np.float32), epochs, batch, lr, seed)
    if log:
        log("training: %d rows, %d epochs, %.1f s, loss %.4f -> %.4f"
            % (len(X), epochs, time.perf_counter() - t0, losses[0], losses[-1]))
    return net, losses
