"""CPU oracle of the surrogate's per-molecule features (numpy, float64).

Test infrastructure only.  The features have no counterpart in the
reference; this restates the definition the device kernel implements, for
molecule i with neighbours j whose O lies within Rc = 6 A of i's O
(minimum image d = O_j - O_i - L rint((O_j - O_i)/L)):
  0-7   sum_j exp(-4 (r - mu)^2) fc(r), mu = 2.4 + 0.5 k
  8-11  sum_j sum_h exp(-10 (|O_j' - H_ih| - mu)^2) fc(r), mu = 1.6, 1.9, 2.2, 2.6
  12-15 sum_j sum_h exp(-10 (|H_jh' - O_i| - mu)^2) fc(r)   (j' the image next to i)
  16-31 2^(1-z) sum_{j<k} (1 + l cos t)^z exp(-eta (r_j^2 + r_k^2)) fc_j fc_k,
        index 16 + 8 e + 4 l + z for eta = 0.02, 0.08; l = +1, -1; z = 1, 2, 4, 8
with fc(r) = (cos(pi r / Rc) + 1) / 2.

Pinned by tests/test_features_oracle.py against what the geometry fixes:
hand-computed dimer and trimer values for every block (radial, both H-bond
blocks, all sixteen angular indices at cos t = 0, -1, 1/2), invariance under
translation, rotation, relabelling and periodic images, and continuity to
zero at Rc.
"""
import numpy as np

RC = 6.0


def features(O, H, L):
    """O (n,3), H (n,2,3) site positions, box edge L -> (n,32)."""
    O = np.asarray(O, np.float64)
    H = np.asarray(H, np.float64)
    n = len(O)
    out = np.zeros((n, 32))
    mu_r = 2.4 + 0.5 * np.arange(8)
    mu_h = np.array([1.6, 1.9, 2.2, 2.6])
    zs = np.array([1.0, 2.0, 4.0, 8.0])
    for i in range(n):
        d = O - O[i]
        d -= L * np.rint(d / L)
        r2 = (d * d).sum(1)
        nb = np.nonzero((r2 < RC * RC) & (np.arange(n) != i))[0]
        d = d[nb]
        r = np.sqrt(r2[nb])
        fc = 0.5 * (np.cos(np.pi * r / RC) + 1.0)
        out[i, :8] = (np.exp(-4.0 * (r[:, None] - mu_r) ** 2) * fc[:, None]).sum(0)
        oj = O[i] + d                      # images of the neighbours' O next to i
        shift = oj - O[nb]
        for h in range(2):
            ra = np.linalg.norm(oj - H[i, h], axis=1)
            rb = np.linalg.norm(H[nb, h] + shift - O[i], axis=1)
            out[i, 8:12] += (np.exp(-10.0 * (ra[:, None] - mu_h) ** 2) * fc[:, None]).sum(0)
            out[i, 12:16] += (np.exp(-10.0 * (rb[:, None] - mu_h) ** 2) * fc[:, None]).sum(0)
        m = len(nb)
        for a in range(m):
            for b in range(a + 1, m):
                cs = d[a] @ d[b] / (r[a] * r[b])
                f = fc[a] * fc[b]
                e2 = r[a] ** 2 + r[b] ** 2
                for e, eta in enumerate((0.02, 0.08)):
                    ex = np.exp(-eta * e2) * f
                    for l, lam in enumerate((1.0, -1.0)):
                        p = 1.0 + lam * cs
                        out[i, 16 + 8 * e + 4 * l:20 + 8 * e + 4 * l] += p ** zs * ex
    out[:, 16:] *= np.tile(2.0 ** (1.0 - zs), 4)
    return out
