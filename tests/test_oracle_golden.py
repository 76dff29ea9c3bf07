"""The oracle against the reference's own outputs (tests/golden/, written by
running the reference built from its sources) and against closed forms."""
import json
import os

import numpy as np
import pytest

from oracle import pyoracle as O
from tests.golden.configs import lattice_start

GOLD = os.path.join(os.path.dirname(__file__), "golden")
CASES = ["tip4p_216", "tip5p_216", "st2_216", "tip4p_125_ndiv3", "spce_216", "nvde_216",
         "meoh_216"]


@pytest.mark.parametrize("name", CASES)
def test_grid_and_pairlist_bit_exact(golden, name):
    g = golden(name)
    com, L, rc = g["com"], float(g["L"]), float(g["rc"])
    ndiv = O.lib().g6o_grid_ndiv(L, float(g["req"]))
    assert ndiv == g["ndiv"][0]
    st, st0 = O.stencil(ndiv)
    np.testing.assert_array_equal(st, g["stencil"].reshape(-1, 3))
    np.testing.assert_array_equal(st0, g["stencil0"].reshape(-1, 3))
    np.testing.assert_array_equal(O.address_assign(com, L, ndiv), g["cell"])
    np0, pi, pj, sh = O.pairlist(com, L, rc, float(g["req"]))
    assert np0 == int(g["npair0"])
    np.testing.assert_array_equal(pi, g["pi"])
    np.testing.assert_array_equal(pj, g["pj"])
    assert np.array_equal(sh, g["shift"])      # FP64 shifts bit-exact
    assert np.array_equal(np.signbit(sh), np.signbit(g["shift"]))


@pytest.mark.parametrize("name", CASES)
def test_cutoff_sites_forces(golden, name):
    g = golden(name)
    mid = int(g["model"])
    rc, taper = float(g["rc"]), float(g["taper"])
    er, fr = O.cutoff(g["pi"], g["pj"], g["com"], g["shift"], rc - taper, rc)
    assert np.array_equal(er, g["eratio"]) and np.array_equal(fr, g["fratio"])
    sites, tm, intra = O.setsites(mid, g["com"], g["quat"].T.copy())
    assert np.array_equal(sites, g["sites"])
    F, ep, vr = O.force(mid, "std", g["pi"], g["pj"], g["shift"], er, fr, sites)
    assert np.array_equal(F, g["force_std"])
    assert np.array_equal(ep, g["ep_std"])
    assert np.array_equal(vr, g["vir_std"])
    Fm, T = O.collect(mid, F, tm, intra)
    assert np.array_equal(Fm, g["mol_force"]) and np.array_equal(T, g["mol_torque"])
    assert np.array_equal(O.energy(mid, g["pi"], g["pj"], g["shift"], er, sites), g["epot"])
    if "force_fast" in g:
        F2, ep2, vr2 = O.force(mid, "fast", g["pi"], g["pj"], g["shift"], er, fr, sites)
        # TIP4P custom kernel: same operation order -> bit-exact; ST2: the oracle
        # loops the 16 Coulomb terms in a different order (<= 1e-14 relative).
        tol = 0.0 if mid == O.TIP4P else 1e-14
        scale = np.abs(g["force_fast"]).max()
        assert np.abs(F2 - g["force_fast"]).max() <= tol * scale
        assert np.abs(ep2 - g["ep_fast"]).max() <= tol * np.abs(g["ep_fast"]).max()


def test_pairlist_digests_larger_grids():
    dig = json.load(open(os.path.join(GOLD, "pairlist_digests.json")))
    import hashlib
    sha = lambda a: hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()
    for key, d in dig.items():
        com, q = lattice_start(d["n"], d["L"], d["seed"])
        np0, pi, pj, sh = O.pairlist(com, d["L"], d["rc"], d["rc"])
        assert (np0, len(pi)) == (d["npair0"], d["npair"]), key
        assert sha(pi) == d["sha_pi"] and sha(pj) == d["sha_pj"], key
        assert sha(sh) == d["sha_shift"], key
        assert sha(O.address_assign(com, d["L"], d["ndiv"])) == d["sha_cell"]


def md_trajectory(g, mid, nose, pre="", fast=False, tol=0.0):
    """The oracle's steps against a reference driver log and final state:
    bit for bit (tol 0), or within tol where the oracle sums a custom
    kernel's terms in another order."""
    md = O.MDRun(mid, g["com"], g["quat"].T.copy(), float(g["L"]), float(g["rc"]),
                 float(g["taper"]), float(g["dt"]), fast=fast, nose=nose)
    md.vreset(300.0, int(g["seed"]))
    close = lambda a, b: np.abs(np.asarray(a) - b).max() <= tol * max(1.0, np.abs(b).max())
    for s in range(int(g["nsteps"])):
        md.step()
        row = g[pre + "logd"][s]
        assert close(md.s.temperature, row[1])            # @LOGD column 2
        assert close(md.s.ep * 10 * 1e-3, row[4])         # column 5, kJ/mol
        assert abs(md.hamiltonian() - row[12]) <= max(1e-12, tol) * abs(row[12])
    assert close(md.com.T, g[pre + "com_final"])
    assert close(md.dx, g[pre + "dxyz_final"].transpose(1, 0, 2))
    assert close(md.quat, g[pre + "quat_final"].transpose(1, 0, 2))
    assert close(md.w * float(g["dt"]), g[pre + "w_final_ext"].transpose(1, 0, 2))


@pytest.mark.parametrize("fast", [False, True])
def test_st2_md_trajectory_with_and_without_fast(golden, fast):
    """ST2 through the reference driver without @FAST (StdForce, no
    Stillinger-Rahman switch; bit for bit) and with @FAST 1 (force_st2, whose
    16 Coulomb terms the oracle sums in another order: within 1e-11)."""
    g = golden("md_st2_216")
    md_trajectory(g, O.ST2, None, "fast_" if fast else "std_", fast=fast,
                  tol=1e-11 if fast else 0.0)


@pytest.mark.parametrize("name", ["md_tip4p_216_nve", "md_tip4p_216_nvt"])
def test_md_trajectory_matches_reference_driver(golden, name):
    g = golden(name)
    nose = None if g["nose"][1] == 0 else tuple(g["nose"])
    md_trajectory(g, O.TIP4P, nose)


@pytest.mark.parametrize("name,mid", [("md_spce_216", O.SPCE), ("md_nvde_216", O.NVDE),
                                      ("md_tip5p_216", O.TIP5P),
                                      ("md_meoh_216", O.OPLSMEOH)])
@pytest.mark.parametrize("ens", ["nve", "nvt"])
def test_std_model_md_trajectory_matches_reference_driver(golden, name, mid, ens):
    """TIP5P, SPC/E (the O site charged and LJ), NvdE (LJ on O and H,
    charges on H, M and the lone pairs) and OPLS methanol (LJ and charges on
    Me and O, a charge on H) through the reference driver's StdForce path:
    the oracle's step reproduces its log and final state bit for bit."""
    g = golden(name)
    md_trajectory(g, mid, tuple(g["nose"]) if ens == "nvt" else None, ens + "_")


def test_mt19937_known_answers():
    # std::mt19937 default seed 5489: first output, and the 10000th output
    # required by the C++ standard ([rand.predef]).
    s = O.mt_stream(5489, 10000)
    assert s[0] == 3499211612
    assert s[9999] == 4123659995


def test_cutoff_closed_form():
    rin, rout = 8.5, 9.0
    hh, dh = np.zeros(1), np.zeros(1)
    import ctypes as C
    def f(r):
        a, b = C.c_double(), C.c_double()
        O.lib().g6o_cutoff_func(rin, rout, r * r, C.byref(a), C.byref(b))
        return a.value, b.value
    assert f(8.0) == (1.0, 0.0) and f(9.0) == (0.0, 0.0)
    # S continuous at both ends, S' = -dh*r matches a central difference
    for r in np.linspace(8.51, 8.99, 17):
        e = 1e-6
        d = (f(r + e)[0] - f(r - e)[0]) / (2 * e)
        assert abs(d - (-f(r)[1] * r)) < 1e-7
    assert abs(f(rin + 1e-7)[0] - 1) < 1e-12 and abs(f(rout - 1e-7)[0]) < 1e-12


@pytest.mark.parametrize("mid", [O.TIP4P, O.TIP5P, O.ST2, O.SPCE, O.NVDE, O.OPLSMEOH])
def test_force_is_minus_gradient(mid):
    """F = -dE/dx by central differences on one molecule's COM; sum F = 0."""
    L, rc, taper = 14.0, 6.5, 1.0
    com, q = lattice_start(27, L, 7 + mid)
    def energy_and_force(c):
        _, pi, pj, sh = O.pairlist(c, L, rc, alltoall=True)
        er, fr = O.cutoff(pi, pj, c, sh, rc - taper, rc)
        sites, tm, intra = O.setsites(mid, c, q.T.copy())
        kind = "fast" if mid in (O.TIP4P, O.ST2) else "std"
        F, ep, _ = O.force(mid, kind, pi, pj, sh, er, fr, sites)
        Fm, _ = O.collect(mid, F, tm, intra)
        return ep.sum(), Fm
    E0, F0 = energy_and_force(com)
    assert np.abs(F0.sum(0)).max() < 1e-9 * np.abs(F0).max()
    h = 1e-6
    for i in (0, 13):
        for a in range(3):
            cp = com.copy(); cp[i, a] += h
            cm = com.copy(); cm[i, a] -= h
            num = -(energy_and_force(cp)[0] - energy_and_force(cm)[0]) / (2 * h)
            assert abs(num - F0[i, a]) < 1e-5 * max(1.0, abs(F0[i, a]))


def test_rotation_matrix_orthonormal():
    com, q = lattice_start(50, 20.0, 3)
    _, tm, _ = O.setsites(O.TIP4P, com, q.T.copy())
    for t in tm.reshape(-1, 3, 3):
        assert np.allclose(t @ t.T, np.eye(3), atol=1e-14)
        assert abs(np.linalg.det(t) - 1) < 1e-14
    _, tm, _ = O.setsites(O.TIP4P, com[:1], np.array([[1.0], [0.0], [0.0], [0.0]]))
    assert np.array_equal(tm.reshape(3, 3), np.eye(3))


def test_lj_coulomb_dimer_closed_form():
    """Two TIP4P molecules with parallel orientation: energy from textbook
    Coulomb + LJ sums over the site table equals the kernel's pair energy."""
    m = O.model(O.TIP4P)
    K = O.constants()
    com = np.array([[0.0, 0.0, 0.0], [3.0, 0.4, -0.2]])
    q = np.array([[1.0, 0.0, 0.0, 0.0], [1.0, 0.0, 0.0, 0.0]])
    sites, tm, intra = O.setsites(O.TIP4P, com, q.T.copy())
    _, pi, pj, sh = O.pairlist(com, 100.0, 9.0, alltoall=True)
    er, fr = O.cutoff(pi, pj, com, sh, 8.5, 9.0)
    _, ep, _ = O.force(O.TIP4P, "fast", pi, pj, sh, er, fr, sites)
    body = np.stack([m.bx[:5], m.by[:5], m.bz[:5]], 1)
    e = 0.0
    qs = [0.0, -1.04, 0.52, 0.52]
    C = 4.184 * 332.17752 * 100.0
    for a in range(1, 4):
        for b in range(1, 4):
            r = np.linalg.norm(body[a] - (body[b] + com[1]))
            e += qs[a] * qs[b] * C / r
    r = np.linalg.norm(body[0] - (body[0] + com[1]))
    e += K[3] / r ** 12 + K[4] / r ** 6
    assert abs(ep[0] - e) < 1e-12 * abs(e)


@pytest.mark.parametrize("name,mid", [("tip4p_216", 0), ("tip5p_216", 1), ("spce_216", 3),
                                      ("nvde_216", 4), ("meoh_216", 5)])
def test_pairinfo_matches_reference_pair_energies(golden, name, mid):
    """PairInfo (pairenergy over the all-pairs list within rc) holds, for
    every pair of the reference's own list, the reference's StdForce pair
    energy; pairs off the list are zero."""
    g = golden(name)
    L, rc, taper = float(g["L"]), float(g["rc"]), float(g["taper"])
    P, e = O.mc_pairinfo(mid, g["com"], g["quat"].T.copy(), L, rc, taper)
    pi, pj, ref = g["pi"], g["pj"], g["ep_std"]
    got = P[pi, pj]
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
    assert (P != 0).sum() == 2 * len(pi)
    assert np.array_equal(P, P.T)
    assert abs(e - ref.sum()) <= 1e-11 * np.abs(ref).sum()


def test_mc_energy_bookkeeping_and_quaternions():
    """After 1500 mc2 trials the running energy (sum of accepted dE) equals a
    fresh PairInfo of the final configuration, and every quaternion is still
    a unit quaternion (qadd of unit quaternions)."""
    from tests.golden.configs import lattice_start
    com, q = lattice_start(216, 18.626, 7)
    r = O.MCRun(0, com, q.T.copy(), 18.626, 8.5, 0.5, 250.0, 4242, 0.3, 1.0)
    nacc = r.trials(1500)
    assert 100 < nacc < 1500
    P, e = O.mc_pairinfo(0, r.com, r.quat, 18.626, 8.5, 0.5)
    assert abs(r.epsum[0] - e) <= 1e-12 * abs(e)
    assert np.abs(P - r.ep.reshape(216, 216)).max() <= 1e-9
    assert np.abs((r.quat ** 2).sum(0) - 1.0).max() < 1e-14
    assert r.delta.min() >= np.float32(0.05)


def test_nve_1000_steps_match_reference_driver(golden):
    """1000 NVE steps of 216 TIP4P from an equilibrated start: the oracle's
    conserved H equals the reference driver's at every step (the reference
    itself drifts upward by about 25 kJ/mol per ps here)."""
    g = golden("md_tip4p_216_nve1000")
    md = O.MDRun(O.TIP4P, g["com"], g["quat"].T.copy(), float(g["L"]), float(g["rc"]),
                 float(g["taper"]), float(g["dt"]))
    md.vreset(300.0, int(g["seed"]))
    H = []
    for _ in range(int(g["nsteps"])):
        md.step()
        H.append(md.hamiltonian())
    ref = g["logd"][:, 12]
    assert np.abs(np.array(H) - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("name,nose", [("npt", (300.0, 7.5e5, 0.0)), ("nph", None)])
def test_andersen_npt_matches_reference_driver(golden, name, nose):
    """50 steps with the Andersen barostat (@PCN4 1 1.0 1e-4), with and
    without the Nose thermostat: the oracle's temperature, pressure, volume
    and barostat energies equal the reference's @LOGD at every step."""
    g = golden("md_tip4p_216_nve1000")
    f = golden("md_tip4p_216_npt")
    md = O.MDRun(O.TIP4P, g["com"], g["quat"].T.copy(), float(g["L"]), float(g["rc"]),
                 float(g["taper"]), float(g["dt"]), nose=nose)
    md.set_andersen(1.0, 1e-4)
    md.vreset(300.0, 4246)
    ref = f["logd_" + name]
    for k in range(len(ref)):
        md.step()
        r = ref[k]
        assert md.s.temperature == r[1] and md.s.pressure == r[2] and md.s.volume == r[3], k
        assert md.s.andersen_ep * 1e-3 == r[10] and md.s.andersen_ek * 1e-3 == r[11], k
        assert abs(md.hamiltonian() - r[12]) <= 1e-12 * abs(r[12])
