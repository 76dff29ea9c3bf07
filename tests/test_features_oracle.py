"""The surrogate's feature oracle (oracle/npt_oracle.py) pinned by what the
geometry fixes, not by restating its own loops: closed forms of dimers and
trimers placed so that every term has a hand value (fc(3 A) = 1/2,
r = mu exactly, cos t in {0, -1, 1/2}), the invariances of the definition
(translation, rotation, relabelling, periodic images of O and of H) and its
continuity to zero at the 6 A cutoff.  A wrong 2^(1-z) factor, eta, lambda
sign, feature index, cutoff function or H image fails one of them."""
import math

import numpy as np
import pytest

from oracle.npt_oracle import RC, features

BIG = 60.0  # box edge far beyond the cutoff: no periodic image within 6 A


def water(o, hdirs, d=0.9572):
    """O position and two H at distance d along the given unit directions."""
    o = np.asarray(o, float)
    h = np.array([o + d * np.asarray(u, float) / np.linalg.norm(u) for u in hdirs])
    return o, h


def sysof(mols):
    O = np.array([m[0] for m in mols])
    H = np.array([m[1] for m in mols])
    return O, H


def test_dimer_radial_closed_form():
    # two O 3 A apart: fc(3) = (cos(pi/2) + 1)/2 = 1/2 exactly
    mols = [water([0, 0, 0], [[0, 0, 1], [0, 1, 0]]), water([3, 0, 0], [[0, 0, 1], [0, 1, 0]])]
    f = features(*sysof(mols), BIG)
    for k in range(8):
        mu = 2.4 + 0.5 * k
        assert f[0, k] == pytest.approx(0.5 * math.exp(-4.0 * (3.0 - mu) ** 2), rel=1e-13, abs=1e-300)
    # a single neighbour has no angular pair
    assert np.all(f[:, 16:] == 0.0)
    # the dimer is symmetric in the radial block
    assert np.allclose(f[0, :8], f[1, :8], rtol=1e-14)


def test_dimer_at_a_radial_centre():
    # r = mu_0 = 2.4: the k = 0 Gaussian is 1, fc(2.4) = (cos(0.4 pi) + 1)/2
    mols = [water([0, 0, 0], [[0, 0, 1], [0, 1, 0]]), water([0, 0, 2.4], [[1, 0, 0], [0, 1, 0]])]
    f = features(*sysof(mols), BIG)
    fc = 0.5 * (math.cos(0.4 * math.pi) + 1.0)
    assert f[0, 0] == pytest.approx(fc, rel=1e-14)
    assert f[0, 1] == pytest.approx(fc * math.exp(-1.0), rel=1e-13)


def test_dimer_hydrogen_bond_terms():
    # i's first H points straight at O_j 3 A away: |O_j - H_i0| = 3 - 0.9572;
    # its second H points away (|O_j - H_i1| from the law of cosines).  j's H
    # point along +z and +y, so |H_j - O_i| = sqrt(9 + 0.9572^2) for both.
    d = 0.9572
    mols = [water([0, 0, 0], [[1, 0, 0], [-1, 0, 0]]), water([3, 0, 0], [[0, 0, 1], [0, 1, 0]])]
    f = features(*sysof(mols), BIG)
    fc = 0.5
    ra = [3.0 - d, 3.0 + d]
    rb = [math.sqrt(9.0 + d * d)] * 2
    for t, mu in enumerate((1.6, 1.9, 2.2, 2.6)):
        ea = sum(math.exp(-10.0 * (r - mu) ** 2) for r in ra) * fc
        eb = sum(math.exp(-10.0 * (r - mu) ** 2) for r in rb) * fc
        assert f[0, 8 + t] == pytest.approx(ea, rel=1e-12, abs=1e-300)
        assert f[0, 12 + t] == pytest.approx(eb, rel=1e-12, abs=1e-300)
    # from j's side the roles of the two blocks swap for this geometry's
    # distances: |O_i - H_j| are rb, |H_i - O_j| are ra
    for t, mu in enumerate((1.6, 1.9, 2.2, 2.6)):
        assert f[1, 8 + t] == pytest.approx(sum(math.exp(-10.0 * (r - mu) ** 2) for r in rb) * fc,
                                            rel=1e-12, abs=1e-300)
        assert f[1, 12 + t] == pytest.approx(sum(math.exp(-10.0 * (r - mu) ** 2) for r in ra) * fc,
                                             rel=1e-12, abs=1e-300)


@pytest.mark.parametrize("angle,cos_t", [(90.0, 0.0), (180.0, -1.0), (60.0, 0.5)])
def test_trimer_angular_closed_form(angle, cos_t):
    # neighbours at 3 A on the x axis and at the given angle in the xy plane:
    # each angular term is 2^(1-z) (1 + l cos t)^z exp(-eta (9 + 9)) fc(3)^2
    a = math.radians(angle)
    up = [[0, 0, 1], [0, 0, -1]]
    mols = [water([0, 0, 0], up), water([3, 0, 0], up),
            water([3 * math.cos(a), 3 * math.sin(a), 0], up)]
    f = features(*sysof(mols), BIG)
    for e, eta in enumerate((0.02, 0.08)):
        for l, lam in enumerate((1.0, -1.0)):
            for zi, z in enumerate((1, 2, 4, 8)):
                want = 2.0 ** (1 - z) * (1.0 + lam * cos_t) ** z * math.exp(-eta * 18.0) * 0.25
                got = f[0, 16 + 8 * e + 4 * l + zi]
                assert got == pytest.approx(want, rel=1e-12, abs=1e-15), (e, l, z)


def _cluster(seed, n=6):
    """A loose cluster of n waters within 5 A of each other, far from the box faces."""
    rng = np.random.default_rng(seed)
    mols = []
    while len(mols) < n:
        o = rng.uniform(-2.5, 2.5, 3) + BIG / 2
        if all(np.linalg.norm(o - m[0]) > 2.3 for m in mols):
            u = rng.normal(size=3)
            v = np.cross(u, rng.normal(size=3))
            mols.append(water(o, [u, v]))
    return sysof(mols)


def _rot(seed):
    q, _ = np.linalg.qr(np.random.default_rng(seed).normal(size=(3, 3)))
    return q * np.sign(np.linalg.det(q))  # a proper rotation


def test_rotation_translation_and_relabelling_invariance():
    O, H = _cluster(11)
    f = features(O, H, BIG)
    R = _rot(3)
    c = np.full(3, BIG / 2)
    Or = (O - c) @ R.T + c
    Hr = (H - c) @ R.T + c
    assert np.allclose(features(Or, Hr, BIG), f, rtol=1e-11, atol=1e-13)
    t = np.array([1.7, -3.3, 0.4])
    assert np.allclose(features(O + t, H + t, BIG), f, rtol=1e-12, atol=1e-13)
    perm = np.random.default_rng(5).permutation(len(O))
    assert np.allclose(features(O[perm], H[perm], BIG), f[perm], rtol=1e-12, atol=1e-13)
    # the H of a molecule are interchangeable
    assert np.allclose(features(O, H[:, ::-1], BIG), f, rtol=1e-12, atol=1e-13)


def test_periodic_images_of_o_and_h():
    """A cluster straddling the box corner equals the same cluster inside the
    box (H sites are given in their own O's frame, as the device builds them
    from one COM)."""
    L = 20.0
    O, H = _cluster(21)
    O = O - BIG / 2 + L / 2
    H = H - BIG / 2 + L / 2
    f = features(O, H, L)
    shift = np.array([L / 2, L / 2 - 0.5, L / 2 + 0.3])
    Os = O + shift
    Hs = H + shift
    wrap = np.floor(Os / L) * L
    Ow = Os - wrap
    Hw = Hs - wrap[:, None, :]
    assert np.allclose(features(Ow, Hw, L), f, rtol=1e-12, atol=1e-13)
    # one molecule (its O and H together) moved by box vectors: every term,
    # the H ones included, sees its image next to the others
    Om, Hm = Ow.copy(), Hw.copy()
    k = np.array([1.0, -1.0, 2.0]) * L
    Om[2] += k
    Hm[2] += k
    assert np.allclose(features(Om, Hm, L), f, rtol=1e-12, atol=1e-13)


def test_continuity_at_the_cutoff():
    up = [[0, 0, 1], [0, 1, 0]]
    vals = []
    for r in (RC - 1e-3, RC - 1e-6, RC, RC + 1e-6):
        mols = [water([0, 0, 0], up), water([r, 0, 0], up), water([0, r, 0], up)]
        vals.append(features(*sysof(mols), BIG)[0])
    # fc ~ (pi eps / Rc)^2 / 4 near the cutoff: 1e-3 below it the terms are O(1e-8)
    assert np.abs(vals[0]).max() < 1e-6
    assert np.abs(vals[1]).max() < 1e-12
    assert np.all(vals[2] == 0.0) and np.all(vals[3] == 0.0)


def test_a_planted_mistake_fails_the_closed_forms():
    """The trimer closed form catches a dropped 2^(1-z) normalisation."""
    up = [[0, 0, 1], [0, 0, -1]]
    mols = [water([0, 0, 0], up), water([3, 0, 0], up), water([0, 3, 0], up)]
    f = features(*sysof(mols), BIG)
    z = np.array([1, 2, 4, 8])
    unnormalised = f[0, 16:20] / 2.0 ** (1 - z)
    want = 2.0 ** (1 - z) * math.exp(-0.02 * 18.0) * 0.25
    assert np.allclose(f[0, 16:20], want, rtol=1e-12)
    assert not np.allclose(unnormalised, want, rtol=1e-3)
