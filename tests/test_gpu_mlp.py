"""The surrogate network's kernels (BF16 MFMA, FP32 accumulation) against
the CPU oracle: inference, the loss gradient, determinism, and training."""
import numpy as np
import pytest

from genesis6_amd import g6
from oracle import mlp_oracle as M

pytestmark = pytest.mark.gpu


def _case(seed, m):
    rng = np.random.default_rng(seed)
    p = g6.mlp_init_params(seed)
    p[M.NHID * M.NIN:M.NHID * M.NIN + M.NHID] = rng.normal(0, 0.1, M.NHID)
    norm = np.concatenate([rng.normal(0, 1, M.NIN), rng.uniform(0.5, 2, M.NIN),
                           [-3.0, 2.5]]).astype(np.float32)
    x = rng.normal(0, 1.5, (m, M.NIN)).astype(np.float32)
    t = rng.normal(-3.0, 2.0, m).astype(np.float32)
    return p, norm, x, t


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, float) - b) / np.linalg.norm(b))


@pytest.mark.parametrize("m", [16, 37, 1000, 5003])
def test_forward_matches_oracle(m):
    print("mlp forward m=%d" % m, flush=True)
    p, norm, x, _ = _case(m, m)
    net = g6.Mlp(p, norm)
    y = net.forward(x)
    yo = M.forward(p.astype(np.float64), norm, x)
    assert rel(y, yo) < 1e-4


@pytest.mark.parametrize("m", [16, 300, 4096, 20001])
def test_gradient_matches_oracle(m):
    print("mlp gradient m=%d" % m, flush=True)
    p, norm, x, t = _case(100 + m, m)
    net = g6.Mlp(p, norm)
    loss, g = net.loss_grad(x, t)
    lo, go = M.loss_grad(p.astype(np.float64), norm, x, t)
    assert abs(loss - lo) < 1e-4 * lo
    o = 0
    for name, size in [("W1", 2048), ("b1", 64), ("W2", 4096), ("b2", 64), ("w3", 64), ("b3", 1)]:
        assert rel(g[o:o + size], go[o:o + size]) < 3e-3, name
        o += size
    # the fixed-order reductions make the gradient reproducible bit for bit
    loss2, g2 = net.loss_grad(x, t)
    assert loss2 == loss and np.array_equal(g, g2)


def test_training_reduces_the_loss():
    rng = np.random.default_rng(7)
    x = rng.normal(0, 1, (16384, M.NIN)).astype(np.float32)
    t = (np.sin(x[:, 0]) * 2 + x[:, 1] * x[:, 2] - 0.5 * x[:, 3] ** 2).astype(np.float32)
    norm = np.concatenate([np.zeros(M.NIN), np.ones(M.NIN), [t.mean(), t.std()]]).astype(np.float32)
    net = g6.Mlp(g6.mlp_init_params(1), norm)
    l0, _ = net.loss_grad(x, t)
    losses = net.train(x, t, epochs=30, batch=256, lr=3e-3, seed=5)
    l1, _ = net.loss_grad(x, t)
    print("loss %.4f -> %.4f" % (l0, l1), flush=True)
    assert losses[-1] < losses[0] and l1 < 0.1 * l0


def test_forward_resolves_small_input_changes():
    """The kernel's two-term BF16 activations: the output change under a
    0.1 % scaling of every feature (a volume move's size) follows the float64
    network's to a few per cent."""
    p, norm, x, _ = _case(9, 4000)
    x2 = (x * np.float32(1.001)).astype(np.float32)
    net = g6.Mlp(p, norm)
    d = net.forward(x2).astype(np.float64) - net.forward(x)
    d64 = (M.forward(p.astype(np.float64), norm, x2, quant=False)
           - M.forward(p.astype(np.float64), norm, x, quant=False))
    assert rel(d, d64) < 0.05
