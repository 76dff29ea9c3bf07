"""Synthetic starting configurations of the fixtures and tests (data only)."""
import numpy as np


def lattice_start(n, L, seed):
    """Simple-cubic lattice, m = ceil(n^(1/3)) per edge, sites chosen by a
    seeded permutation, small seeded jitter, Shoemake-uniform quaternions."""
    rng = np.random.default_rng(seed)
    m = int(np.ceil(round(n ** (1.0 / 3.0), 9)))
    while m ** 3 < n:
        m += 1
    g = (np.arange(m) + 0.5) * L / m
    pts = np.array(np.meshgrid(g, g, g, indexing="ij")).reshape(3, -1).T
    idx = np.sort(rng.permutation(len(pts))[:n])
    com = pts[idx] + rng.normal(0.0, 0.05, (n, 3))
    u = rng.random((n, 3))
    q = np.stack([np.sqrt(1 - u[:, 0]) * np.sin(2 * np.pi * u[:, 1]),
                  np.sqrt(1 - u[:, 0]) * np.cos(2 * np.pi * u[:, 1]),
                  np.sqrt(u[:, 0]) * np.sin(2 * np.pi * u[:, 2]),
                  np.sqrt(u[:, 0]) * np.cos(2 * np.pi * u[:, 2])], 1)
    return com, q
