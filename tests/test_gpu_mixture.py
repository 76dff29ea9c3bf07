"""Two-component MD on the GPU (g6_mixture_*): 448 TIP4P water and 64 OPLS
methanol in a 29 A box against the reference MD driver's run of the same
deck (tests/golden/md_mixture_tip4p_meoh.npz, written by
tests/golden/make_mixture_md.py): @VSET 300 K drawn over the water, then
the methanol, from one random stream; one Nose thermostat over both; the
pairs between the components by StdForce with Lorentz-Berthelot site
parameters.  Every @LOGD column of the 30 steps is compared: T, P, V, the
pair energy of each pair set (water-water, methanol-methanol,
water-methanol), the kinetic energy of each component, the per-molecule
sums, the thermostat energies and H.  The Fortran driver g6md runs the
reference's deck itself and continues it from its own restart."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from genesis6_amd import g6
from tests.golden.configs import parse_logd

pytestmark = pytest.mark.gpu
I2J, X2KX = 10.0, 1e-3
COLS = [(c, 1e-12) for c in (1, 2, 4, 5, 6, 7, 8, 9, 10, 11, 12, 15, 17)] + [(3, 1e-15)]


def logd_row(t, p, nmol):
    ep, ek = t[0], t[1] + t[2]
    return [t[3], t[25], t[26], p[0] * I2J * X2KX, p[1] * I2J * X2KX, p[2] * I2J * X2KX,
            p[3] * I2J * X2KX, p[4] * I2J * X2KX, ep * I2J / nmol * X2KX,
            ek * I2J / nmol * X2KX, t[27] * X2KX, t[28] * X2KX, t[30] * X2KX, t[31] * X2KX,
            t[13], 0.0, t[13], 0.0]


def mixture(f, fast=False):
    L, rc, taper, dt = float(f["L"]), float(f["rc"]), float(f["taper"]), float(f["dt"])
    w, m = f["water"], f["meoh"]
    sw = g6.System(g6.TIP4P, len(w), L, rc, taper, dt, prec=g6.FP64)
    sm = g6.System(g6.OPLSMEOH, len(m), L, rc, taper, dt, prec=g6.FP64)
    for s, x in ((sw, w), (sm, m)):
        s.set_fast(fast)
        s.upload(x[:, :3], x[:, 3:])
    sw.set_nose(300.0, 7.5e5)
    mx = g6.Mixture(sw, sm)
    mx.vreset(300.0, int(f["seed"]))
    return sw, sm, mx


def test_mixture_log_matches_reference_driver(golden):
    f = golden("md_mixture_tip4p_meoh")
    ref = f["logd"]
    print("mixture 448 TIP4P + 64 OPLSMEOH, %d steps" % len(ref), flush=True)
    sw, sm, mx = mixture(f)
    nmol = sw.n + sm.n
    got = []
    for k in range(len(ref)):
        mx.step(1)
        t, p = mx.thermo()
        got.append([k] + logd_row(t, p, nmol))
    got = np.array(got)
    assert got.shape == ref.shape
    np.testing.assert_array_equal(got[:, 0], ref[:, 0])
    # T, P, V, ep of the three sets, ek of both, per molecule, Nose, H: each
    # within 1e-12 of its column's largest magnitude (the differences are
    # summation order; measured at about 1e-14 and below)
    for col, tol in COLS:
        err = np.abs(got[:, col] - ref[:, col]).max() / (np.abs(ref[:, col]).max() + 1e-300)
        print("  column %d: %.2e" % (col, err))
        assert err <= tol, (col, err)
    for c in (13, 14, 16, 18):
        assert np.all(ref[:, c] == 0.0)
    mx.close()
    sw.close()
    sm.close()


def test_mixture_steps_in_chunks_as_one_by_one(golden):
    """k steps in one call end where k calls of one step end, bit for bit
    (the state between calls is the corrected state), and the methanol keeps
    its own thermostat buffer untouched."""
    f = golden("md_mixture_tip4p_meoh")
    ends = []
    for chunks in ([1] * 6, [4, 2]):
        sw, sm, mx = mixture(f)
        nose_m = sm.nose()
        for c in chunks:
            mx.step(c)
        t, p = mx.thermo()
        ends.append((t.copy(), p.copy(), sw.download()[0], sm.download()[0]))
        assert np.array_equal(sm.nose(), nose_m)
        mx.close()
        sw.close()
        sm.close()
    for a, b in zip(ends[0], ends[1]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("skin", ["0", "1.0"])
def test_g6md_fp32_mixture_follows_the_reference(golden, skin):
    """FP32 pair arithmetic (the pairs between components in FP64 from the
    FP32 site offsets of the records), with and without a pair-list skin:
    the temperature, the three pair energies and H follow the reference's
    FP64 run over the 30 steps."""
    f = golden("md_mixture_tip4p_meoh")
    ref = f["logd"]
    print("g6md mixture deck fp32 skin", skin, flush=True)
    with tempfile.TemporaryDirectory() as d:
        got = run_g6md(str(f["deck"]), d, "fp32", skin)
    assert got.shape == ref.shape
    # T and H against their own size; each pair set's energy against the
    # largest pair energy (the methanol-methanol set is a small sum of large
    # terms of both signs, so its own size is no scale for FP32 rounding)
    ep_scale = np.abs(ref[:, 4:7]).max()
    for col in (1, 4, 5, 6, 15):
        scale = ep_scale if 4 <= col <= 6 else np.abs(ref[:, col]).max()
        err = np.abs(got[:, col] - ref[:, col]).max() / scale
        print("  column %d: %.2e" % (col, err))
        assert err <= 1e-5, (col, err)


def test_mixture_refuses_mismatched_components(golden):
    f = golden("md_mixture_tip4p_meoh")
    L, rc, taper, dt = float(f["L"]), float(f["rc"]), float(f["taper"]), float(f["dt"])
    a = g6.System(g6.TIP4P, 448, L, rc, taper, dt, prec=g6.FP64)
    for kw in ({"prec": g6.FP32}, {"L": L + 1.0}, {"rc": rc - 1.0}, {"dt": 0.001}):
        # precision, box, cutoff, time step: each must be the first's
        args = {"L": L, "rc": rc, "dt": dt, "prec": g6.FP64}
        args.update(kw)
        b = g6.System(g6.OPLSMEOH, 64, args["L"], args["rc"], taper, args["dt"],
                      prec=args["prec"])
        with pytest.raises(g6.G6Error):
            g6.Mixture(a, b)
        b.close()
    with pytest.raises(g6.G6Error):
        g6.Mixture(a, a)
    a.close()


EXE = os.path.abspath(os.path.join(os.path.dirname(__file__), "..", "genesis6_amd", "lib",
                                   "g6md"))


def run_g6md(deck, d, *args):
    r = subprocess.run([EXE] + list(args), input=deck, capture_output=True, text=True,
                       timeout=120, cwd=d)
    assert r.returncode == 0, r.stderr
    return parse_logd(r.stdout)


def close_to(got, ref):
    assert got.shape == ref.shape
    np.testing.assert_array_equal(got[:, 0], ref[:, 0])
    for col, tol in COLS:
        err = np.abs(got[:, col] - ref[:, col]).max() / (np.abs(ref[:, col]).max() + 1e-300)
        assert err <= tol, (col, err)


def test_g6md_runs_the_two_component_deck(golden):
    """The reference's own deck (two @ID08/@NX4A blocks) through the Fortran
    driver: its @LOGD lines carry the three pair energies and two kinetic
    energies as the reference's do."""
    f = golden("md_mixture_tip4p_meoh")
    print("g6md mixture deck", flush=True)
    with tempfile.TemporaryDirectory() as d:
        close_to(run_g6md(str(f["deck"]), d, "fp64"), f["logd"])


def test_g6md_mixture_restart_continues_the_run(golden):
    """20 steps, the fort.13 restart (@NCMP 2, a @TIE4 per component, then
    @FIXC @ID08 @WTG2 per component) read back as fort.12, 10 more steps:
    the last 10 log lines are the reference's single run's."""
    f = golden("md_mixture_tip4p_meoh")
    deck = str(f["deck"]).replace("@MDLP\n30\n", "@MDLP\n20\n", 1)
    assert "@MDLP\n20\n" in deck
    print("g6md mixture restart", flush=True)
    with tempfile.TemporaryDirectory() as d:
        first = run_g6md(deck, d, "fp64")
        close_to(first, f["logd"][:20])
        os.rename(os.path.join(d, "fort.13"), os.path.join(d, "fort.12"))
        rest = run_g6md("@MDLP\n10\n", d, "fp64")
    close_to(rest, f["logd"][20:])


def test_g6md_mixture_with_fast_kernels(golden):
    """With @FAST 1 the reference runs its custom TIP4P, methanol and
    TIP4P x methanol kernels (Interaction2.F90:19-59).  Its two methanol
    kernels assign each pair's virial to vrmat instead of adding it
    (OPLSMeOHModule.F90:500-511, 1655-1666), so the reference's pressure then
    holds the last pair's virial of those two sets alone; every other column
    (energies, temperature, H: the trajectory) equals the StdForce run's.
    g6md with @FAST 1 gives the reference's @FAST columns except P, and the
    pressure of the full virial, the StdForce run's P."""
    f = golden("md_mixture_tip4p_meoh")
    ref, ref_fast = f["logd"], f["logd_fast"]
    print("g6md mixture deck with @FAST 1", flush=True)
    with tempfile.TemporaryDirectory() as d:
        got = run_g6md(str(f["deck_fast"]), d, "fp64")
    close_to(got, ref)
    pfast = np.abs(ref_fast[:, 2] - ref[:, 2]).max() / np.abs(ref[:, 2]).max()
    assert pfast > 0.5, pfast            # the reference's @FAST pressure is another number
    other = [c for c, _ in COLS if c != 2]
    err = np.abs(got[:, other] - ref_fast[:, other]).max(0) / np.abs(ref_fast[:, other]).max(0)
    assert err.max() <= 1e-12, err


def test_mixture_refuses_calls_after_a_component_is_freed(golden):
    f = golden("md_mixture_tip4p_meoh")
    sw, sm, mx = mixture(f)
    mx.step(2)
    sm.close()
    for call in (lambda: mx.step(1), mx.thermo, lambda: mx.vreset(300.0, 1)):
        with pytest.raises(g6.G6Error):
            call()
    mx.close()
    sw.step(1)          # the first component steps on its own again
    sw.close()


@pytest.mark.parametrize("lpm", ["16", "64", "256"])
def test_every_cross_kernel_form_follows_the_reference(golden, monkeypatch, lpm):
    """The cross kernel's three forms (a 16-lane row, a wave, or a block of
    four waves per molecule; chosen by partner count and size, set here by
    G6_MIX_LPM) each give the reference's log."""
    f = golden("md_mixture_tip4p_meoh")
    ref = f["logd"]
    monkeypatch.setenv("G6_MIX_LPM", lpm)
    sw, sm, mx = mixture(f)
    monkeypatch.delenv("G6_MIX_LPM")
    print("mixture, cross kernel form", lpm, flush=True)
    got = []
    for k in range(10):
        mx.step(1)
        t, p = mx.thermo()
        got.append([k] + logd_row(t, p, sw.n + sm.n))
    close_to(np.array(got), ref[:10])
    mx.close()
    sw.close()
    sm.close()


def lattice_mixture(nside=16, a=3.625, seed=7):
    """The benchmark's mixture start: water on a simple cubic lattice, OPLS
    methanol on the sites with even indices on all three axes."""
    rng = np.random.default_rng(seed)
    water, meoh = [], []
    for i in range(nside):
        for j in range(nside):
            for k in range(nside):
                q = rng.standard_normal(4)
                row = [(i + 0.5) * a, (j + 0.5) * a, (k + 0.5) * a] + list(q / np.linalg.norm(q))
                (meoh if i % 2 == 0 and j % 2 == 0 and k % 2 == 0 else water).append(row)
    return np.array(water), np.array(meoh), nside * a


def test_cross_forms_agree_at_4096_molecules(monkeypatch):
    """3584 water + 512 methanol (58 A box): the cross kernel's three forms
    give the same pair energies and thermo record after 20 steps within
    1e-12, and FP32 pairs follow FP64 (T within 1e-5)."""
    water, meoh, L = lattice_mixture()
    out = {}
    for name, prec, lpm in (("16", g6.FP64, "16"), ("64", g6.FP64, "64"),
                            ("256", g6.FP64, "256"), ("fp32", g6.FP32, "")):
        print("mixture 4096, form", name, flush=True)
        sw = g6.System(g6.TIP4P, len(water), L, 8.5, 0.5, 0.002, prec=prec, skin=1.0)
        sm = g6.System(g6.OPLSMEOH, len(meoh), L, 8.5, 0.5, 0.002, prec=prec, skin=1.0)
        sw.upload(water[:, :3], water[:, 3:])
        sm.upload(meoh[:, :3], meoh[:, 3:])
        sw.set_nose(300.0, 7.5e5)
        if lpm:
            monkeypatch.setenv("G6_MIX_LPM", lpm)
        mx = g6.Mixture(sw, sm)
        monkeypatch.delenv("G6_MIX_LPM", raising=False)
        mx.vreset(300.0, 4246)
        mx.step(20)
        out[name] = mx.thermo()
        mx.close()
        sw.close()
        sm.close()
    t0, p0 = out["64"]
    for name in ("16", "256"):
        t, p = out[name]
        assert np.abs(p[:3] - p0[:3]).max() <= 1e-12 * np.abs(p0[:3]).max(), (name, p, p0)
        assert abs(t[3] - t0[3]) <= 1e-12 * t0[3] and abs(t[13] - t0[13]) <= 1e-12 * abs(t0[13])
    t, _ = out["fp32"]
    assert abs(t[3] / t0[3] - 1.0) < 1e-5, (t[3], t0[3])
