"""CPU oracle of the energy surrogate's network (numpy, float64).

Test infrastructure only: imported by tests/ and nothing in the product.
The network has no counterpart in the reference: this module restates the
definition the device kernels implement (an MLP 32 -> 64 -> 64 -> 1 with
SiLU; BF16 weights; forward activations as two BF16 terms hi = bf16(v),
lo = bf16(v - hi); the backward products on BF16 operands (the hi terms);
the loss mean((o - (t - y_shift)/y_scale)^2)), so that the kernels' forward
pass and gradient can be compared with it.  With quant=False every BF16
rounding is skipped and the arithmetic is plain float64: that variant is
pinned by central finite differences of its own loss.

Parameter layout: W1[64][32], b1[64], W2[64][64], b2[64], w3[64], b3.
Normalisation: mean[32], inv_std[32], y_shift, y_scale.
"""
import numpy as np

NIN, NHID = 32, 64
NPARAM = NHID * NIN + NHID + NHID * NHID + NHID + NHID + 1
NNORM = 2 * NIN + 2


def bf16(a):
    """Round float32 values to BF16 (nearest, ties to even), as float64."""
    u = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, np.float64)
    o = 0
    W1 = p[o:o + NHID * NIN].reshape(NHID, NIN); o += NHID * NIN
    b1 = p[o:o + NHID]; o += NHID
    W2 = p[o:o + NHID * NHID].reshape(NHID, NHID); o += NHID * NHID
    b2 = p[o:o + NHID]; o += NHID
    w3 = p[o:o + NHID]; o += NHID
    b3 = p[o]
    return W1, b1, W2, b2, w3, b3


def pack(W1, b1, W2, b2, w3, b3):
    return np.concatenate([np.ravel(W1), b1, np.ravel(W2), b2, w3, [b3]])


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def _q(a, quant):
    return bf16(a) if quant else np.asarray(a, np.float64)


def split(a):
    """(hi, lo) = (bf16(v), bf16(v - hi)) of float32 values v: the two-term
    BF16 form of a forward activation (the difference is exact in FP32)."""
    v = np.asarray(a, np.float32)
    hi = bf16(v)
    lo = bf16((v.astype(np.float64) - hi).astype(np.float32))
    return hi, lo


def forward(p, norm, x, quant=True, keep=False):
    """Energies y[m] = y_scale * o + y_shift of feature rows x[m][32]."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    norm = np.asarray(norm, np.float64)
    mean, istd, ysh, ysc = norm[:NIN], norm[NIN:2 * NIN], norm[2 * NIN], norm[2 * NIN + 1]
    x = np.asarray(x, np.float64)
    if quant:  # the kernel normalises in FP32
        xh, xl = split((x.astype(np.float32) - mean.astype(np.float32)) * istd.astype(np.float32))
        z1 = (xh + xl) @ bf16(W1).T + b1
        hh, hl = split(silu(z1))
        z2 = (hh + hl) @ bf16(W2).T + b2
        xn, h1 = xh, hh       # the backward's operands
    else:
        xn = (x - mean) * istd
        z1 = xn @ W1.T + b1
        h1 = silu(z1)
        z2 = h1 @ W2.T + b2
    o = silu(z2) @ w3 + b3
    y = ysc * o + ysh
    if keep:
        return y, dict(xn=xn, z1=z1, h1=h1, z2=z2, o=o)
    return y


def loss_grad(p, norm, x, t, quant=True):
    """(loss, grad[NPARAM]) of mean((o - (t - y_shift)/y_scale)^2)."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    norm = np.asarray(norm, np.float64)
    ysh, ysc = norm[2 * NIN], norm[2 * NIN + 1]
    _, c = forward(p, norm, x, quant, keep=True)
    m = len(c["o"])
    d = c["o"] - (np.asarray(t, np.float64) - ysh) / ysc
    loss = float(np.mean(d * d))
    g = 2.0 * d / m
    dz2 = g[:, None] * w3[None, :] * dsilu(c["z2"])
    gw3 = (g[:, None] * silu(c["z2"])).sum(0)
    gb2 = dz2.sum(0)
    gb3 = g.sum()
    dz2q = _q(dz2, quant)
    gW2 = dz2q.T @ c["h1"]
    dh1 = dz2q @ _q(W2, quant)
    dz1 = dh1 * dsilu(c["z1"])
    gb1 = dz1.sum(0)
    gW1 = _q(dz1, quant).T @ c["xn"]
    return loss, pack(gW1, gb1, gW2, gb2, gw3, gb3)
