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


def mixture_deck(com, q, L, rc, taper, dt, nsteps, seed, nose=None, model="TIP4P", nlog=1):
    """A Mixture input deck (the reference's tags and payload layout,
    Mixture.F90:600-1140): @DTPS, @MDLP, @NLOG, @RCOA, @BXLA, @VOXA, @VSET,
    optional @NOS2 (T, q, dump), @ID08, @NX4A with x y z a b c d per line."""
    n = len(com)
    dt, rc, taper, L = float(dt), float(rc), float(taper), float(L)
    deck = ["@DTPS", repr(dt), "@MDLP", str(int(nsteps)), "@NLOG", str(int(nlog)), "@RCOA",
            "%r %r" % (rc, taper), "@BXLA", repr(L), "@VOXA", "%r %r %r" % (rc, rc, rc),
            "@VSET", "300.0 %d" % int(seed)]
    if nose is not None:
        deck += ["@NOS2", "%r %r %r 0 0 0 0 0 0" % tuple(float(v) for v in nose)]
    deck += ["@ID08", model, "@NX4A", str(n)]
    deck += [" ".join("%.17e" % v for v in list(com[i]) + list(q[i])) for i in range(n)]
    return "\n".join(deck) + "\n"


def parse_logd(text):
    """The @LOGD rows of a Mixture-format log as an array (one row per line
    after each @LOGD tag)."""
    lines = text.split("\n")
    rows = [lines[i + 1] for i, l in enumerate(lines[:-1]) if l.strip() == "@LOGD"]
    return np.array([[float(v) for v in r.split()] for r in rows])
