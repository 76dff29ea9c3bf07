"""The surrogate network's CPU oracle: BF16 rounding against its definition
and the float64 gradient against central finite differences of the loss."""
import numpy as np

from genesis6_amd import g6
from oracle import mlp_oracle as M


def test_bf16_round_to_nearest_even():
    e = 2.0 ** -8
    vals = np.array([1.0, 1.0 + e, 1.0 + 3 * e, 1.0 + 1.5 * e, -2.0 - 2 * e, 3.0e-39],
                    np.float32)
    want = [1.0, 1.0, 1.0 + 4 * e, 1.0 + 2 * e, -2.0, 3.0e-39]
    got = M.bf16(vals)
    assert got[:5].tolist() == want[:5]
    # a subnormal keeps 7 mantissa bits of spacing 2^-133: within half of it
    assert abs(got[5] - want[5]) <= 2.0 ** -134


def _case(seed, m=40):
    rng = np.random.default_rng(seed)
    p = g6.mlp_init_params(seed).astype(np.float64)
    p[M.NHID * M.NIN:M.NHID * M.NIN + M.NHID] = rng.normal(0, 0.1, M.NHID)  # b1
    norm = np.concatenate([rng.normal(0, 1, M.NIN), rng.uniform(0.5, 2, M.NIN), [-3.0, 2.5]])
    x = rng.normal(0, 1.5, (m, M.NIN))
    t = rng.normal(-3.0, 2.0, m)
    return p, norm, x, t


def test_float64_gradient_matches_finite_differences():
    p, norm, x, t = _case(3)
    loss, g = M.loss_grad(p, norm, x, t, quant=False)
    rng = np.random.default_rng(9)
    # every bias and output weight, and a sample of each weight matrix
    idx = list(range(M.NHID * M.NIN, M.NHID * M.NIN + M.NHID))
    o2 = M.NHID * M.NIN + M.NHID
    idx += list(rng.choice(M.NHID * M.NIN, 60, replace=False))
    idx += list(o2 + rng.choice(M.NHID * M.NHID, 60, replace=False))
    idx += list(range(M.NPARAM - M.NHID - M.NHID - 1, M.NPARAM))
    h = 1e-6
    for k in idx:
        pp = p.copy(); pp[k] += h
        pm = p.copy(); pm[k] -= h
        fd = (M.loss_grad(pp, norm, x, t, quant=False)[0] -
              M.loss_grad(pm, norm, x, t, quant=False)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])


def test_quantised_oracle_stays_close_to_float64():
    p, norm, x, t = _case(4, 200)
    l0, g0 = M.loss_grad(p, norm, x, t, quant=False)
    l1, g1 = M.loss_grad(p, norm, x, t, quant=True)
    assert abs(l1 - l0) < 0.02 * l0
    assert np.linalg.norm(g1 - g0) < 0.05 * np.linalg.norm(g0)


def test_split_forward_resolves_small_input_changes():
    """The two-term BF16 activations keep 16 significant bits: the oracle's
    output change under a 0.1 % scaling of every feature follows the float64
    network's to a few per cent, where single BF16 operands (8 bits) lose
    it in rounding steps."""
    p, norm, x, _ = _case(5, 4000)
    x2 = (x * np.float32(1.001)).astype(np.float32)
    d64 = M.forward(p, norm, x2, quant=False) - M.forward(p, norm, x, quant=False)
    dq = M.forward(p, norm, x2) - M.forward(p, norm, x)
    assert np.linalg.norm(dq - d64) < 0.05 * np.linalg.norm(d64)
    # single BF16 operands for comparison
    W1, b1, W2, b2, w3, b3 = M.unpack(p)
    nm = np.asarray(norm, np.float64)

    def one_term(xx):
        xn = M.bf16((xx - nm[:M.NIN].astype(np.float32)) * nm[M.NIN:2 * M.NIN].astype(np.float32))
        h1 = M.bf16(M.silu(xn @ M.bf16(W1).T + b1))
        return M.silu(h1 @ M.bf16(W2).T + b2) @ w3 + b3

    d1 = one_term(x2) - one_term(x)
    assert np.linalg.norm(d1 * nm[-1] - d64) > 0.3 * np.linalg.norm(d64)
