"""HIP path against the oracle and the reference's golden fixtures (GPU).

Bars: integer/index work bit-exact (cells, pair list order and shifts); FP64
mode <= 1e-12 relative RMS force error; FP32-pair mode <= 1e-5 relative RMS
force error, both against the FP64 oracle.
"""
import json
import os
import tempfile

import numpy as np
import pytest

from genesis6_amd import g6
from oracle import pyoracle as O
from tests.golden.configs import lattice_start

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")
CASES = ["tip4p_216", "tip5p_216", "st2_216", "tip4p_125_ndiv3", "spce_216", "nvde_216",
         "meoh_216"]


def relrms(a, b):
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


def system_for(g, prec=g6.FP64, skin=0.0, dt=0.002):
    print("case model=%d n=%d L=%g prec=%d" % (int(g["model"]), len(g["com"]), float(g["L"]), prec),
          flush=True)
    return g6.System(int(g["model"]), len(g["com"]), float(g["L"]), float(g["rc"]),
                     float(g["taper"]), dt, voxa=float(g["req"]), prec=prec, skin=skin)


@pytest.mark.parametrize("name", CASES)
def test_cells_and_pairlist_bit_exact(golden, name):
    g = golden(name)
    s = system_for(g)
    cell, nd = s.cells(g["com"])
    assert nd == int(g["ndiv"][0])
    np.testing.assert_array_equal(cell, g["cell"])
    np0, pi, pj, sh = s.pairlist(g["com"])
    assert np0 == int(g["npair0"])
    np.testing.assert_array_equal(pi, g["pi"])
    np.testing.assert_array_equal(pj, g["pj"])
    assert np.array_equal(sh, g["shift"])
    assert np.array_equal(np.signbit(sh), np.signbit(g["shift"]))   # -0.0 kept


def test_pairlist_digests_and_full_size():
    """Reference digests at 1024/2048 molecules; oracle at 4096 (C2 size)."""
    import hashlib
    sha = lambda a: hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()
    dig = json.load(open(os.path.join(GOLD, "pairlist_digests.json")))
    for key, d in dig.items():
        com, _ = lattice_start(d["n"], d["L"], d["seed"])
        s = g6.System(g6.TIP4P, d["n"], d["L"], d["rc"], 0.5, 0.002)
        np0, pi, pj, sh = s.pairlist(com)
        assert (np0, len(pi)) == (d["npair0"], d["npair"]), key
        assert sha(pi) == d["sha_pi"] and sha(pj) == d["sha_pj"] and sha(sh) == d["sha_shift"]
    n, L, rc = 4096, 49.669, 9.0
    com, _ = lattice_start(n, L, 20261202 + 2)
    com = com + np.random.default_rng(5).normal(0, 0.3, com.shape) + 3.0 * L  # unwrapped
    s = g6.System(g6.TIP4P, n, L, rc, 0.5, 0.002)
    got = s.pairlist(com)
    ref = O.pairlist(com, L, rc, rc)
    assert got[0] == ref[0]
    np.testing.assert_array_equal(got[1], ref[1])
    np.testing.assert_array_equal(got[2], ref[2])
    assert np.array_equal(got[3], ref[3])


@pytest.mark.parametrize("name", CASES)
def test_sites_cutoff_collect(golden, name):
    g = golden(name)
    s = system_for(g)
    sites = s.site_positions(g["com"], g["quat"])
    assert np.abs(sites - g["sites"]).max() <= 1e-14 * np.abs(g["sites"]).max()
    er, fr = s.cutoff(g["pi"], g["pj"], g["com"], g["shift"])
    assert np.abs(er - g["eratio"]).max() <= 1e-14
    assert np.abs(fr - g["fratio"]).max() <= 1e-14 * max(1e-300, np.abs(g["fratio"]).max())
    F, T = s.collect(g["quat"], g["force_std"])
    assert relrms(F, g["mol_force"]) <= 1e-13
    assert relrms(T, g["mol_torque"]) <= 1e-13


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("prec,tol", [(g6.FP64, 1e-12), (g6.FP32, 1e-5)])
def test_force_seam(golden, name, prec, tol):
    """force() drop-in: site forces, pair energies, virial vs the reference."""
    g = golden(name)
    s = system_for(g, prec)
    ref = g["force_fast"] if "force_fast" in g else g["force_std"]
    ref_ep = g["ep_fast"] if "ep_fast" in g else g["ep_std"]
    ref_vr = g["vir_fast"] if "vir_fast" in g else g["vir_std"]
    F, ep, vr = s.force(g["pi"], g["pj"], g["shift"], g["eratio"], g["fratio"], g["sites"])
    assert relrms(F, ref) <= tol, relrms(F, ref)
    assert relrms(ep, ref_ep) <= tol
    assert relrms(vr, ref_vr) <= tol
    e = s.pair_energy(g["pi"], g["pj"], g["shift"], g["eratio"], g["sites"])
    assert relrms(e, ref_ep) <= tol


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("prec,tol", [(g6.FP64, 1e-12), (g6.FP32, 1e-5)])
def test_md_kernels_forces(golden, name, prec, tol):
    """The hot path (row kernel + fused collect) vs the reference's
    rigid_collectforce output on the same state."""
    g = golden(name)
    s = system_for(g, prec)
    s.upload(g["com"], g["quat"])
    F, T, th = s.eval_forces()
    Fr, Tr = g["mol_force"], g["mol_torque"]          # collect of StdForce
    if int(g["model"]) == O.ST2:
        # the MD path runs the switched custom ST2 kernel: collect its site
        # forces with the (bit-exact) oracle collect
        _, tm, intra = O.setsites(O.ST2, g["com"], g["quat"].T.copy())
        Fr, Tr = O.collect(O.ST2, g["force_fast"], tm, intra)
    assert relrms(F, Fr) <= tol, relrms(F, Fr)
    assert relrms(T, Tr) <= tol, relrms(T, Tr)
    ep_ref = (g["ep_fast"] if "ep_fast" in g else g["ep_std"]).sum()
    assert abs(th[0] - ep_ref) <= tol * np.abs(g["ep_std"]).sum()
    assert int(th[14]) == len(g["pi"])


@pytest.mark.parametrize("prec,tol", [(g6.FP64, 1e-12), (g6.FP32, 1e-5)])
def test_st2_without_fast_is_stdforce(golden, prec, tol):
    """g6_set_fast(0): ST2 through StdForce (no Stillinger-Rahman switch), as
    the reference driver runs it without @FAST: the force seam and the MD
    kernels against the reference's StdForce output."""
    g = golden("st2_216")
    s = system_for(g, prec)
    s.set_fast(False)
    F, ep, vr = s.force(g["pi"], g["pj"], g["shift"], g["eratio"], g["fratio"], g["sites"])
    assert relrms(F, g["force_std"]) <= tol, relrms(F, g["force_std"])
    assert relrms(ep, g["ep_std"]) <= tol and relrms(vr, g["vir_std"]) <= tol
    e = s.pair_energy(g["pi"], g["pj"], g["shift"], g["eratio"], g["sites"])
    assert relrms(e, g["ep_std"]) <= tol
    s.upload(g["com"], g["quat"])
    Fm, T, th = s.eval_forces()
    assert relrms(Fm, g["mol_force"]) <= tol and relrms(T, g["mol_torque"]) <= tol
    assert abs(th[0] - g["ep_std"].sum()) <= tol * np.abs(g["ep_std"]).sum()
    # and back: the custom kernel again
    s.set_fast(True)
    F2, _, _ = s.force(g["pi"], g["pj"], g["shift"], g["eratio"], g["fratio"], g["sites"])
    assert relrms(F2, g["force_fast"]) <= tol


@pytest.mark.parametrize("skin", [0.0, 1.0])
def test_md_kernels_forces_c2_size(skin):
    """4096 TIP4P at C2 density: FP64 and FP32 hot path vs the FP64 oracle."""
    n, L, rc, taper = 4096, 49.669, 9.0, 0.5
    com, q = lattice_start(n, L, 20261202 + 2)
    np0, pi, pj, sh = O.pairlist(com, L, rc, rc)
    er, fr = O.cutoff(pi, pj, com, sh, rc - taper, rc)
    sites, tm, intra = O.setsites(O.TIP4P, com, q.T.copy())
    Fs, ep, vr = O.force(O.TIP4P, "fast", pi, pj, sh, er, fr, sites)
    Fm, Tm = O.collect(O.TIP4P, Fs, tm, intra)
    for prec, tol in [(g6.FP64, 1e-12), (g6.FP32, 1e-5)]:
        s = g6.System(g6.TIP4P, n, L, rc, taper, 0.002, prec=prec, skin=skin)
        s.upload(com, q)
        F, T, th = s.eval_forces()
        assert relrms(F, Fm) <= tol and relrms(T, Tm) <= tol, (prec, relrms(F, Fm), relrms(T, Tm))
        assert int(th[14]) == len(pi)
        assert abs(th[0] - ep.sum()) <= tol * np.abs(ep).sum()
        assert relrms(th[4:13].reshape(3, 3), vr) <= tol


# (fixture, model, ensemble): the TIP4P pair of fixtures and the SPC/E and
# NvdE ones, whose arrays carry an "nve_" / "nvt_" prefix
TRAJ = [("md_tip4p_216_nve", g6.TIP4P, ""), ("md_tip4p_216_nvt", g6.TIP4P, ""),
        ("md_spce_216", g6.SPCE, "nve"), ("md_spce_216", g6.SPCE, "nvt"),
        ("md_nvde_216", g6.NVDE, "nve"), ("md_nvde_216", g6.NVDE, "nvt"),
        ("md_st2_216", g6.ST2, "std"), ("md_st2_216", g6.ST2, "fast"),
        ("md_tip5p_216", g6.TIP5P, "nve"), ("md_tip5p_216", g6.TIP5P, "nvt"),
        ("md_meoh_216", g6.OPLSMEOH, "nve"), ("md_meoh_216", g6.OPLSMEOH, "nvt")]
DECK_ID = {g6.TIP4P: "TIP4P", g6.SPCE: "SPC_E", g6.NVDE: "NVDE____", g6.ST2: "ST2_____",
           g6.TIP5P: "TIP5P", g6.OPLSMEOH: "OPLSMEOH"}


def traj_case(g, ens):
    """(prefix, Nose parameters or None, @FAST) of one trajectory fixture:
    the reference driver runs without @FAST unless the case is "fast"."""
    if ens in ("std", "fast"):
        return ens + "_", None, ens == "fast"
    if ens:
        return ens + "_", (tuple(g["nose"]) if ens == "nvt" else None), False
    return "", (tuple(g["nose"]) if g["nose"][1] != 0 else None), False


@pytest.mark.parametrize("name,mid,ens", TRAJ)
def test_md_trajectory_vs_reference_driver(golden, name, mid, ens):
    """20 FP64 steps against the reference Mixture run (the same MB start
    from MT19937): temperature and energy per step, final Gear state."""
    g = golden(name)
    pre, nose, fast = traj_case(g, ens)
    n = len(g["com"])
    print("trajectory", name, ens, flush=True)
    s = g6.System(mid, n, float(g["L"]), float(g["rc"]), float(g["taper"]), float(g["dt"]))
    s.set_fast(fast)
    s.upload(g["com"], g["quat"])
    if nose is not None:
        s.set_nose(*nose)
    s.vreset(300.0, int(g["seed"]))
    for k in range(int(g["nsteps"])):
        s.step(1)
        th = s.thermo()
        row = g[pre + "logd"][k]
        assert abs(th[3] - row[1]) <= 1e-10 * row[1], (k, th[3], row[1])
        assert abs(th[0] * 1e-2 - row[4]) <= 1e-9 * abs(row[4]) + 1e-9, (k, th[0] * 1e-2, row[4])
    com, quat, dx, w = s.download()
    assert np.abs(com.T - g[pre + "com_final"]).max() < 1e-9
    assert np.abs(quat - g[pre + "quat_final"].transpose(1, 0, 2)).max() < 1e-9
    assert np.abs(dx - g[pre + "dxyz_final"].transpose(1, 0, 2)).max() < 1e-12


@pytest.mark.parametrize("prec", [g6.FP64, g6.FP32])
def test_graph_replay_equals_eager_bitwise(prec):
    """hipGraph-captured steps give the same trajectory as eager launches,
    bit for bit: the MD path has no floating-point atomics."""
    n, L = 512, 24.83
    com, q = lattice_start(n, L, 11)
    a = g6.System(g6.TIP4P, n, L, 9.0, 0.5, 0.002, prec=prec, skin=1.0)
    b = g6.System(g6.TIP4P, n, L, 9.0, 0.5, 0.002, prec=prec, skin=1.0)
    for s in (a, b):
        s.upload(com, q)
        s.set_nose(300.0, 7.5e5)
        s.vreset(300.0, 7)
    a.step(100)           # two graph replays of 50 steps
    for _ in range(100):
        b.step(1)         # eager launches
    for x, y in zip(a.download(), b.download()):
        assert np.array_equal(x, y)
    assert np.array_equal(a.thermo(), b.thermo())
    assert np.array_equal(a.nose(), b.nose())


def test_skin_list_gives_the_reference_interacting_set():
    """A Verlet skin changes when rows are rebuilt, not which pairs interact:
    50 FP64 steps with skin 0 and skin 1.5 A agree to round-off."""
    n, L = 1024, 31.289
    com, q = lattice_start(n, L, 12)
    out = []
    for skin in (0.0, 1.5):
        s = g6.System(g6.TIP4P, n, L, 9.0, 0.5, 0.002, skin=skin)
        s.upload(com, q)
        s.vreset(300.0, 9)
        pairs = []
        for _ in range(50):
            s.step(1)
            pairs.append(s.thermo()[14])
        out.append((s.download(), pairs))
    assert out[0][1] == out[1][1]          # same interacting pairs every step
    assert np.abs(out[0][0][0] - out[1][0][0]).max() < 1e-9


def test_capacity_error_is_reported():
    n, L = 216, 18.626
    com, q = lattice_start(n, L, 3)
    s = g6.System(g6.TIP4P, n, L, 8.5, 0.5, 0.002)
    with pytest.raises(g6.G6Error):
        s.pairlist(com, cap=10)


@pytest.mark.parametrize("name,mid,ens", TRAJ)
def test_fortran_driver_reproduces_the_reference_log(golden, name, mid, ens):
    """The Fortran driver g6md reads the same Mixture deck as the reference
    and writes @LOGD rows that match the reference's own log: step number,
    temperature, pressure, volume, energies and the conserved quantity."""
    import subprocess
    from tests.golden.configs import mixture_deck, parse_logd
    g = golden(name)
    pre, nose, fast = traj_case(g, ens)
    deck = mixture_deck(g["com"], g["quat"], float(g["L"]), float(g["rc"]), float(g["taper"]),
                        float(g["dt"]), int(g["nsteps"]), int(g["seed"]), nose,
                        model=DECK_ID[mid], fast=True if fast else None)
    exe = os.path.join(os.path.dirname(GOLD), "..", "genesis6_amd", "lib", "g6md")
    print("g6md deck", name, flush=True)
    with tempfile.TemporaryDirectory() as d:   # g6md writes its restart file fort.13
        r = subprocess.run([os.path.abspath(exe), "fp64"], input=deck, capture_output=True,
                           text=True, timeout=120, cwd=d)
    assert r.returncode == 0, r.stderr
    got = parse_logd(r.stdout)
    ref = g[pre + "logd"]
    assert got.shape == ref.shape
    np.testing.assert_array_equal(got[:, 0], ref[:, 0])
    for col, tol in [(1, 1e-10), (2, 1e-6), (3, 1e-14), (4, 1e-9), (5, 1e-10), (8, 1e-8),
                     (9, 1e-8), (12, 1e-10)]:
        scale = np.abs(ref[:, col]).max() + 1e-300
        err = np.abs(got[:, col] - ref[:, col]).max() / scale
        assert err <= tol, (col, err)


def test_fortran_driver_snapshots_like_the_reference(golden):
    """@SNPO / @SNPV: the Fortran driver writes the coordinate and the
    coordinate + velocity snapshots of the reference's log at the same steps,
    in the same order among the @LOGD rows, with the same records and values
    (11 significant digits as printed; the trajectories agree far below)."""
    import subprocess
    from tests.golden.configs import mixture_deck, parse_snapshots
    g = golden("md_tip4p_216_nvt")
    ref = golden("snapshots_tip4p_216")
    deck = mixture_deck(g["com"], g["quat"], float(g["L"]), float(g["rc"]), float(g["taper"]),
                        float(g["dt"]), int(g["nsteps"]), int(g["seed"]), tuple(g["nose"]),
                        snpo=int(ref["snpo"]), snpv=int(ref["snpv"]))
    exe = os.path.join(os.path.dirname(GOLD), "..", "genesis6_amd", "lib", "g6md")
    print("g6md snapshots", flush=True)
    with tempfile.TemporaryDirectory() as d:
        r = subprocess.run([os.path.abspath(exe), "fp64"], input=deck, capture_output=True,
                           text=True, timeout=120, cwd=d)
    assert r.returncode == 0, r.stderr
    seq = "".join({"@LOGD": "L", "@NX4A": "S", "@WTG3": "V"}[l.strip()]
                  for l in r.stdout.split("\n") if l.strip() in ("@LOGD", "@NX4A", "@WTG3"))
    assert seq == str(ref["sequence"])
    snaps = parse_snapshots(r.stdout)
    assert len(snaps) == len(ref["step"])
    for k, b in enumerate(snaps):
        assert b["kind"] == ref["kind"][k] and b["step"] == ref["step"][k]
        assert b["ncmp"] == ref["ncmp"][k] and (b["fixc"] or "") == ref["fixc"][k]
        assert b["id"] == ref["ident"][k]
        assert np.array_equal(b["box"], ref["box"][k])
        rows = ref["b%d_rows" % k]
        assert b["rows"].shape == rows.shape
        err = np.abs(b["rows"] - rows) / (np.abs(rows).max(0) + 1e-300)
        assert err.max() <= 1e-9, (k, err.max(0))


@pytest.mark.parametrize("mid", [0, 1, 2, 3, 4, 5])
def test_mc_pairinfo_matches_oracle(mid):
    """PairInfo on the GPU equals the oracle's matrix bit for bit."""
    n, L = 216, 18.626
    com, q = lattice_start(n, L, 21 + mid)
    s = g6.System(mid, n, L, 8.5, 0.5, 0.002)
    s.upload(com, q)
    s.mc_begin(250.0, 99, 0.3, 1.0)
    P, e = s.mc_pairinfo()
    Pr, er = O.mc_pairinfo(mid, com, q.T.copy(), L, 8.5, 0.5)
    np.testing.assert_array_equal(P, Pr)
    assert abs(e - er) <= 1e-12 * abs(er)


@pytest.mark.parametrize("mid", [g6.TIP4P, g6.ST2, g6.SPCE, g6.NVDE, g6.OPLSMEOH])
def test_mc_trials_match_oracle(mid):
    """2000 mc2 trials on the GPU take the oracle's decisions: same accept
    count, the same final configuration bit for bit, the same PairInfo; the
    running energy agrees with a fresh recompute."""
    n, L = 512, 25.088
    com, q = lattice_start(n, L, 31)
    print("mc trials model", mid, flush=True)
    s = g6.System(mid, n, L, 9.0, 0.5, 0.002)
    s.upload(com, q)
    s.mc_begin(250.0, 4242, 0.3, 1.0)
    r = O.MCRun(mid, com, q.T.copy(), L, 9.0, 0.5, 250.0, 4242, 0.3, 1.0)
    for chunk in (700, 1300):
        acc = s.mc_trials(chunk)
        assert acc == r.trials(chunk), chunk
    c, qq, d, rt, e = s.mc_state()
    np.testing.assert_array_equal(c, r.com)
    np.testing.assert_array_equal(qq.T, r.quat)
    np.testing.assert_array_equal(d, r.delta)
    np.testing.assert_array_equal(rt, r.rtr)
    assert abs(e - r.epsum[0]) <= 1e-10 * abs(e)
    P, _ = s.mc_pairinfo()
    np.testing.assert_array_equal(P, r.ep.reshape(n, n))
    assert abs(s.mc_energy() - e) <= 1e-10 * abs(e)


@pytest.mark.parametrize("name", ["nose", "voxn", "box3", "book"])
def test_fortran_driver_reads_the_other_tag_forms(golden, name):
    """g6md on the decks that use @NOSE (the older thermostat tag: no zetasum,
    the mass in the older unit), @VOXN and @BOX3 writes the @LOGD rows that
    the reference driver writes for the same decks.  With @BOOK 10 1.0 the
    reference evaluates a list booked every 10 steps, and its run departs
    from its own unbooked run (T by 2e-5 relative within 20 steps); g6md
    takes the margin as its skin and keeps the exact interacting set, so its
    @BOOK run is the reference's unbooked run."""
    import subprocess
    from tests.golden.configs import parse_logd, tag_decks
    g = golden("md_tip4p_216_nvt")
    ref = g["logd"] if name == "book" else golden("decktags_tip4p_216")[name]
    deck = tag_decks(g)[name]
    exe = os.path.join(os.path.dirname(GOLD), "..", "genesis6_amd", "lib", "g6md")
    print("g6md tag deck", name, flush=True)
    with tempfile.TemporaryDirectory() as d:
        r = subprocess.run([os.path.abspath(exe), "fp64"], input=deck, capture_output=True,
                           text=True, timeout=120, cwd=d)
    assert r.returncode == 0, r.stderr
    got = parse_logd(r.stdout)
    assert got.shape == ref.shape
    np.testing.assert_array_equal(got[:, 0], ref[:, 0])
    for col, tol in [(1, 1e-10), (2, 1e-6), (3, 1e-14), (4, 1e-9), (5, 1e-10), (8, 1e-8),
                     (9, 1e-8), (12, 1e-10)]:
        scale = np.abs(ref[:, col]).max() + 1e-300
        err = np.abs(got[:, col] - ref[:, col]).max() / scale
        assert err <= tol, (col, err)
