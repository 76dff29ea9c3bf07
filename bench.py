"""Benchmark: TIP4P-4096 rigid-water MD step on one MI355X.

Workload (the headline configuration): 4096 TIP4P molecules at 1.0 g/cm3
(L = 49.669 A), Nose NVT at 300 K, cutoff 9.0 A switched over 0.5 A, dt 2 fs,
FP32 pair arithmetic with FP64 accumulation; synthetic seeded lattice start
with uniform random orientations and Maxwell-Boltzmann velocities, brought to
a 300 K liquid by an untimed equilibration (velocity redraws at 300 K, then
Nose NVT alone).  The FP64 GPU path starts from the same equilibrated state
and is timed with the same step counts and the same launch mode (every
g6_step call replays captured hipGraphs).  One "step" is one full MD step
(predict, cell grid + pair rows, forces, collect, correct, thermostat), all
resident in HBM.  Neighbour rows carry a 1.8 A Verlet skin by default and are
rebuilt when a molecule has moved half of it; the FP64 rc test on every
listed pair keeps the interacting set the reference's (rows rebuilt every
step) exactly.

Prints ONE JSON line (rank 0, N=1).  --only c2|c3|c4|c5|models|mix runs one workload
(for profiling); --eager launches every kernel without graphs (the mode
kernel traces are taken in).
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N, L, RC, TAPER, DT, TEMP = 4096, 49.669, 9.0, 0.5, 0.002, 300.0
NOSE_Q = 7.5e5          # @NOS2 mass: T0*tau^2/dt^2 with tau = 0.1 ps
SEED = 20261204
# algorithmic FLOP per interacting molecule pair (FMA = 2, rsqrt/rcp = 1):
# Coulomb site pair 21, LJ site pair 28, 40 per molecule pair
FLOP_PER_PAIR = {"tip4p": 9 * 21 + 28 + 40, "tip5p": 16 * 21 + 28 + 40}
FP32_PEAK = 157.3e12    # MI355X vector FP32 (vendor, packed)
FP32_PEAK_NONPACKED = 78.6e12  # without packed FP32 instructions (not used by these kernels)


def log(msg):
    print("bench: " + msg, file=sys.stderr, flush=True)


def new_system(mid, n, Lb, prec, skin, seed, temp, q_nos2):
    from genesis6_amd import g6
    com, q = g6.lattice_start(n, Lb, seed)
    s = g6.System(mid, n, Lb, RC, TAPER, DT, prec=prec, skin=skin)
    s.upload(com, q)
    s.set_nose(temp, q_nos2)
    s.vreset(temp, seed)
    return s


def equilibrate(s, temp, seed, redraws, every, settle):
    """Untimed: `redraws` rounds of a fresh Maxwell-Boltzmann draw at temp
    followed by `every` Nose NVT steps (they carry the lattice's excess
    potential energy off), then `settle` steps of Nose NVT alone."""
    for r in range(redraws):
        s.vreset(temp, seed + 1000 + r)
        s.step(every)
    s.step(settle)
    return s.thermo().copy()


def state_of(s):
    com, q, dx, w = s.download()
    return dict(com=com, quat=q, dx=dx, w=w, nose=s.nose())


def load_state(s, st):
    s.upload(st["com"], None, st["dx"], st["w"], quat=st["quat"])
    z = st["nose"]
    s.set_nose(z[0], z[1] / s.dt, z[2], tuple(z[3:8]), z[8])


def timed_leg(s, steps, warmup):
    """A step(steps) call to capture its graphs (untimed), `warmup` untimed
    steps, then exactly `steps` steps between two stream synchronizations."""
    s.step(steps)
    s.step(warmup)
    s.thermo()                       # synchronizes the stream
    t0 = time.perf_counter()
    s.step(steps)
    th = s.thermo()                  # synchronizes the stream
    return time.perf_counter() - t0, th.copy()


def roofline(flop_per_pair, pairs, kernel_ms, kernel, traffic, **extra):
    achieved = flop_per_pair * pairs / (kernel_ms * 1e-3)
    out = {"bound": "valu", "achieved": achieved / 1e12, "peak": FP32_PEAK / 1e12,
           "unit": "TFLOP/s", "frac": achieved / FP32_PEAK, "traffic": traffic,
           "kernel": kernel, "kernel_ms": kernel_ms, "flop_per_pair": flop_per_pair,
           "pairs_per_launch": float(pairs),
           "frac_of_nonpacked_peak": achieved / FP32_PEAK_NONPACKED}
    out.update(extra)
    return out


def traffic_of(name):
    """HBM bytes per force-kernel launch from a committed rocprofv3 --pmc
    summary (FETCH_SIZE + WRITE_SIZE, copy-kernel calibrated), if present."""
    try:
        return json.load(open(os.path.join(ROOT, "profiles", name)))["bytes_per_launch"]
    except (OSError, ValueError, KeyError):
        return None


def c2_run(a):
    """The headline: FP32 and FP64 legs on the equilibrated 300 K state."""
    from genesis6_amd import g6
    log("c2: TIP4P-%d, equilibration" % N)
    s = new_system(g6.TIP4P, N, L, g6.FP32, a.skin, SEED, TEMP, NOSE_Q)
    t0 = time.perf_counter()
    th_eq = equilibrate(s, TEMP, SEED, 10, 2000, 10000)
    log("c2: equilibrated in %.1f s: T %.1f K, U %.1f, %d pairs"
        % (time.perf_counter() - t0, th_eq[3], th_eq[0], th_eq[14]))
    st = state_of(s)
    legs = {}
    for name, prec in (("fp32", g6.FP32), ("fp64", g6.FP64)):
        if name == "fp64" and a.no_fp64:
            continue
        if name == "fp64":
            s = g6.System(g6.TIP4P, N, L, RC, TAPER, DT, prec=g6.FP64, skin=a.skin)
        load_state(s, st)
        el, th = timed_leg(s, a.steps, a.warmup)
        leg = dict(value=N * a.steps / el, steps=a.steps, warmup=a.warmup,
                   ms_per_step=el / a.steps * 1e3, temperature_K=float(th[3]),
                   interacting_pairs=float(th[14]))
        # a long run of the same leg (10,000 steps by default)
        if a.long_steps > 0:
            load_state(s, st)
            el_l, _ = timed_leg(s, a.long_steps, a.warmup)
            leg["long_run"] = dict(steps=a.long_steps, value=N * a.long_steps / el_l,
                                   ms_per_step=el_l / a.long_steps * 1e3)
        # the force kernel: back-to-back launches on the equilibrated state
        # (events on the library stream) and its share of eager steps
        load_state(s, st)
        s.step(1)
        leg["force_kernel_ms"] = s.time_force(200)
        t = s.step_timed(50)
        leg["eager_step"] = dict(force_kernel_ms=t[0], per_molecule_kernels_ms=t[2],
                                 step_ms=t[3], pairs=t[4])
        legs[name] = leg
        log("c2 %s: %.4g molecule-steps/s (%.4f ms/step), force kernel %.2f us, T %.1f K"
            % (name, leg["value"], leg["ms_per_step"], leg["force_kernel_ms"] * 1e3,
               leg["temperature_K"]))
        s.close()
    f32 = legs["fp32"]
    rl = roofline(FLOP_PER_PAIR["tip4p"], f32["eager_step"]["pairs"], f32["force_kernel_ms"],
                  "k_md_main<TIP4P,float>", traffic_of("md_main_tip4p4096_traffic.json"),
                  kernel_in_eager_step_ms=f32["eager_step"]["force_kernel_ms"])
    return legs, rl, st, th_eq


def tip5p_run(a):
    """65,536 TIP5P molecules at 1.0 g/cm3 (L = 125.157 A), Nose NVT 275 K,
    rc 9.0 A switched over 0.5 A, FP32 pairs (the cell-tile force kernel),
    timed from an equilibrated start."""
    from genesis6_amd import g6
    n, Lb, T = 65536, 125.157, 275.0
    s = new_system(g6.TIP5P, n, Lb, g6.FP32, a.tip5p_skin, 20261202 + 3, T, T * 2500.0)
    equilibrate(s, T, 20261202 + 3, 4, 250, 500)
    el, th = timed_leg(s, a.tip5p_steps, 50)
    kern = s.md_kernel()
    s.step(1)
    kms = s.time_force(50)
    pairs = float(s.thermo()[14])
    s.close()
    name = "k_md_tile<TIP5P>" if kern[0] else "k_md_main<TIP5P,float>"
    out = dict(value=n * a.tip5p_steps / el, unit="molecule-steps/s", molecules=n,
               steps=a.tip5p_steps, ms_per_step=el / a.tip5p_steps * 1e3,
               pair_list_skin_A=a.tip5p_skin,
               pairs_per_step=float(th[14]), temperature_K=float(th[3]),
               force_kernel={"tiles": bool(kern[0]), "cells_per_axis": kern[1],
                             "workgroups_per_tile": kern[2], "waves": kern[3],
                             "slots": kern[4], "left_tiles_on_code": kern[5]})
    out["roofline"] = roofline(FLOP_PER_PAIR["tip5p"], pairs, kms, name,
                               traffic_of("md_tile_tip5p65536_traffic.json"))
    return out


def cpu_child(path, budget):
    """One process of the all-cores CPU baseline: the FP64 oracle from the
    saved state for `budget` seconds; prints molecule-steps/s."""
    from oracle import pyoracle as O
    z = np.load(path)
    md = O.MDRun(O.TIP4P, z["com"], z["quat"][0], L, RC, TAPER, DT, fast=True,
                 nose=(TEMP, NOSE_Q, 0.0))
    md.quat[:] = z["quat"]
    md.dx[:] = z["dx"]
    md.w[:] = z["w"]
    for i in range(9):
        md.s.nose[i] = z["nose"][i]
    md.step(1)
    t0 = time.perf_counter()
    k = 0
    while True:
        md.step(1)
        k += 1
        el = time.perf_counter() - t0
        if el >= budget:
            break
    return N * k / el, k, el


def cpu_baseline(st, budget, procs):
    """The FP64 oracle (serial C restatement of the reference step, custom
    TIP4P kernel) on the equilibrated C2 state: one process in this one,
    then `procs` concurrent child processes (independent replicas, the
    reference's replica-parallel mode)."""
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "state.npz")
        np.savez(path, com=st["com"], quat=st["quat"], dx=st["dx"], w=st["w"], nose=st["nose"])
        v1, k1, e1 = cpu_child(path, budget)
        code = ("import sys; sys.path.insert(0, %r); import bench; "
                "print(bench.cpu_child(%r, %r)[0])" % (ROOT, path, budget))
        ps = [subprocess.Popen([sys.executable, "-c", code], stdout=subprocess.PIPE, text=True)
              for _ in range(procs)]
        rates = []
        for p in ps:
            out, _ = p.communicate()
            if p.returncode != 0:
                raise RuntimeError("CPU baseline child failed")
            rates.append(float(out.strip().splitlines()[-1]))
    return v1, k1, e1, sum(rates)


def reference_calibration():
    """The oracle / flang-built reference rate ratio measured where the
    reference sources are (committed by oracle/ref_timing.py)."""
    try:
        return json.load(open(os.path.join(ROOT, "profiles", "cpu_reference_timing.json")))
    except (OSError, ValueError):
        return None


# Replica ladders: 4 ladders x 16 rungs, T_k = 220 (320/220)^(k/15) K; per
# ladder (N, rho, L): TIP4P, rc 9.0 A switched over 0.5 A, FP32 pairs, Nose
# Q = T tau^2/dt^2 (tau 0.1 ps); one exchange of 30 trials per ladder every
# 500 steps over the adjacent-rung chain.
LADDERS = [(256, 0.92, 20.266), (512, 0.97, 25.088), (1024, 1.00, 31.289),
           (2048, 1.05, 38.786)]
RUNGS, RX_EVERY, RX_TRIALS = 16, 500, 30


TEMPS = [220.0 * (320.0 / 220.0) ** (k / 15.0) for k in range(RUNGS)]


def ladder_systems(skin):
    """The 64 replicas (4 ladders x 16 rungs) from their seeded starts."""
    from genesis6_amd import g6
    systems = []
    for li, (n, rho, Lb) in enumerate(LADDERS):
        for k, T in enumerate(TEMPS):
            seed = 20261202 + 4 + 100 * li + k
            s = g6.System(g6.TIP4P, n, Lb, RC, TAPER, DT, prec=g6.FP32, skin=skin)
            com, q = g6.lattice_start(n, Lb, seed)
            s.upload(com, q)
            s.set_nose(T, T * 2500.0)
            s.vreset(T, seed)
            systems.append(s)
    return systems


def ladder_run(intervals, skin, tune=False, log=None):
    """All 64 replicas stepped by one ladder (one launch per MD kernel for
    every replica, one graph per 500-step interval) with the reference's
    replica exchange between intervals.  With tune, the autotuner first
    picks the skin and block assignment.  Returns molecule-steps/s over the
    timed intervals, exchanges included."""
    from genesis6_amd import g6
    from genesis6_amd.autotune import tune_ladder
    best, table = None, None
    mode = g6.Ladder.MAP_FLAT
    if tune:
        best, table = tune_ladder(ladder_systems, log=log)
        skin, mode = best["skin"], best["map_mode"]
    systems = ladder_systems(skin)
    groups = []
    for li in range(len(LADDERS)):
        bonds = g6.rxpr_bonds(RUNGS, [(k, k + 1) for k in range(RUNGS - 1)])
        groups.append(g6.ReplicaExchange(RUNGS, bonds, RX_TRIALS, 20261202 + 4 + 100 * li + 99,
                                         systems=systems[li * RUNGS:(li + 1) * RUNGS]))
    lad = g6.Ladder(systems)
    lad.set_map(mode)
    nmol = sum(s.n for s in systems)

    def interval():
        lad.step(RX_EVERY)
        th = lad.thermo()
        acc = 0
        for li, grp in enumerate(groups):
            acc += grp.exchange(th[li * RUNGS:(li + 1) * RUNGS])
        return acc, th

    interval()                                   # warm-up: graph capture
    t0 = time.perf_counter()
    acc = 0
    for _ in range(intervals):
        a, th = interval()
        acc += a
    el = time.perf_counter() - t0
    steps = intervals * RX_EVERY
    kms = lad.time_force(100)
    out = dict(value=nmol * steps / el, unit="molecule-steps/s", replicas=len(systems),
               molecules=nmol, steps=steps, ms_per_step=el / steps * 1e3,
               skin_A=skin, block_map="xcd" if mode == g6.Ladder.MAP_XCD else "flat",
               autotune=table,
               exchange_trials=intervals * RX_TRIALS * len(groups), accepted=acc,
               pairs_per_step=float(th[:, 14].sum()),
               mean_temperature_K=float(th[:, 3].mean()))
    out["roofline"] = roofline(FLOP_PER_PAIR["tip4p"], th[:, 14].sum(), kms,
                               "k_md_main_lad<TIP4P,float>",
                               traffic_of("md_main_lad_tip4p64x960_traffic.json"))
    lad.close()
    for g in groups:
        g.close()
    for s in systems:
        s.close()
    return out


def models_run(steps, skin):
    """The other rigid models on the headline workload (4096 molecules at
    1.0 g/cm3, L = 49.669 A, Nose NVT 300 K, rc 9.0 A switched over 0.5 A,
    FP32 pairs): SPC/E, NvdE and ST2 (its custom kernel), molecule-steps/s,
    each timed from an equilibrated start."""
    from genesis6_amd import g6
    n, Lb, T = 4096, 49.669, 300.0
    out = {}
    for name, mid in (("spce", g6.SPCE), ("nvde", g6.NVDE), ("st2", g6.ST2)):
        s = new_system(mid, n, Lb, g6.FP32, skin, 20261204, T, T * 2500.0)
        equilibrate(s, T, 20261204, 5, 1000, 2000)
        el, th = timed_leg(s, steps, 50)
        s.close()
        out[name] = dict(value=n * steps / el, unit="molecule-steps/s", steps=steps,
                         ms_per_step=el / steps * 1e3, pairs_per_step=float(th[14]),
                         temperature_K=float(th[3]))
    return out


def mixture_run(steps, skin):
    """A two-component mixture (the reference's @NCMP 2 run): 3584 TIP4P
    water and 512 OPLS methanol on a 16 x 16 x 16 simple cubic lattice of
    spacing 3.625 A (L = 58.0 A; methanol on the sites with even indices on
    all three axes), uniform random orientations (seed 7), @VSET 300 K over
    both from one stream, one Nose thermostat at 300 K, rc 8.5 A switched
    over 0.5 A, dt 2 fs.  200 untimed steps from the lattice, then `steps`
    timed steps, for FP32 and FP64 pairs; molecule-steps/s."""
    from genesis6_amd import g6
    nside, a = 16, 3.625
    L = nside * a
    rng = np.random.default_rng(7)
    water, meoh = [], []
    for i in range(nside):
        for j in range(nside):
            for k in range(nside):
                q = rng.standard_normal(4)
                row = [(i + 0.5) * a, (j + 0.5) * a, (k + 0.5) * a] + list(q / np.linalg.norm(q))
                (meoh if i % 2 == 0 and j % 2 == 0 and k % 2 == 0 else water).append(row)
    water, meoh = np.array(water), np.array(meoh)
    out = {"water": len(water), "methanol": len(meoh), "L_A": L, "pair_list_skin_A": skin}
    for name, prec in (("fp32", g6.FP32), ("fp64", g6.FP64)):
        sw = g6.System(g6.TIP4P, len(water), L, 8.5, 0.5, 0.002, prec=prec, skin=skin)
        sm = g6.System(g6.OPLSMEOH, len(meoh), L, 8.5, 0.5, 0.002, prec=prec, skin=skin)
        sw.upload(water[:, :3], water[:, 3:])
        sm.upload(meoh[:, :3], meoh[:, 3:])
        sw.set_nose(300.0, 7.5e5)
        mx = g6.Mixture(sw, sm)
        mx.vreset(300.0, 4246)
        mx.step(200)
        mx.thermo()
        t0 = time.perf_counter()
        mx.step(steps)
        th, parts = mx.thermo()
        el = time.perf_counter() - t0
        n = len(water) + len(meoh)
        out[name] = dict(value=n * steps / el, unit="molecule-steps/s", steps=steps,
                         ms_per_step=el / steps * 1e3, temperature_K=float(th[3]),
                         pairs_per_step=float(th[14]))
        mx.close()
        sw.close()
        sm.close()
    return out


def by_move_type(plain, da):
    """Stage-1 (surrogate) passes and stage-2 (exact) acceptances of the
    delayed-acceptance chain by move type, beside the plain chain's
    acceptance of each type."""
    pv, pb = plain["volume_proposals"], plain["proposals"] - plain["volume_proposals"]
    dv, db = da["volume_proposals"], da["proposals"] - da["volume_proposals"]
    s1v = da["volume_surrogate_passes"]
    s1b = da["surrogate_passes"] - s1v
    av, ab = da["volume_accepted"], da["accepted"] - da["volume_accepted"]
    frac = lambda a, b: a / b if b else None
    return {
        "plain": {"volume_acceptance": frac(plain["volume_accepted"], pv),
                  "block_acceptance": frac(plain["accepted"] - plain["volume_accepted"], pb)},
        "delayed_acceptance": {
            "volume": {"proposals": dv, "stage1_passes": s1v, "stage2_accepts": av,
                       "stage1_pass_rate": frac(s1v, dv), "stage2_accept_rate": frac(av, s1v),
                       "acceptance": frac(av, dv)},
            "block": {"proposals": db, "stage1_passes": s1b, "stage2_accepts": ab,
                      "stage1_pass_rate": frac(s1b, db), "stage2_accept_rate": frac(ab, s1b),
                      "acceptance": frac(ab, db)}}}


def npt_ensemble(net, moves, rounds, K, p_volume, chains, log):
    """<V> and <U> of the plain and delayed-acceptance chains: `chains`
    independent pairs of chains, each pair from its own MD start (3000 Nose
    steps at 250 K from its own lattice seed) and its own proposal stream,
    both chains of a pair from the same 2000-move plain equilibration; the
    box volume and energy are read 20 times along each chain.  Errors are the
    standard errors over the independent chains."""
    from genesis6_amd import g6
    n, L, T, P = 1024, 31.289, 250.0, 1.0
    means = {"plain": [], "delayed_acceptance": []}
    for c in range(chains):
        s = g6.System(g6.TIP4P, n, L, RC, TAPER, DT, prec=g6.FP32, skin=1.0)
        com, q = g6.lattice_start(n, L, 20261300 + c)
        s.upload(com, q)
        s.set_nose(T, T * 2500.0)
        s.vreset(T, 20261300 + c)
        s.step(3000)
        for kind, count in (("plain", moves), ("delayed_acceptance", rounds)):
            s.npt_begin(T, P, 20261400 + c, 0.15, 8.0, 0.005, p_volume, 9.0)
            s.npt_run(2000)
            vs, us = [], []
            for _ in range(20):
                if kind == "plain":
                    s.npt_run(count // 20)
                else:
                    s.npt_run(count // 20, mlp=net, K=K)
                _, _, Lc, Uc = s.npt_state()
                vs.append(Lc ** 3)
                us.append(Uc)
            means[kind].append((float(np.mean(vs)), float(np.mean(us))))
        s.close()
    out = {"chains": chains, "unit_volume": "A^3", "unit_energy": "internal (10 J/mol)"}
    for kind, m in means.items():
        a = np.array(m)
        se = a.std(0, ddof=1) / np.sqrt(len(a)) if len(a) > 1 else np.zeros(2)
        out[kind] = {"V": float(a[:, 0].mean()), "V_err": float(se[0]),
                     "U": float(a[:, 1].mean()), "U_err": float(se[1]),
                     "per_chain": m}
    p, d = out["plain"], out["delayed_acceptance"]
    for q in ("V", "U"):
        err = np.hypot(p[q + "_err"], d[q + "_err"])
        out["z_" + q] = float((d[q] - p[q]) / err) if err > 0 else None
    log("npt ensemble: <V> plain %.0f +- %.0f, DA %.0f +- %.0f; <U> plain %.0f +- %.0f, "
        "DA %.0f +- %.0f" % (p["V"], p["V_err"], d["V"], d["V_err"], p["U"], p["U_err"],
                             d["U"], d["U_err"]))
    return out


def npt_run(configs, epochs, moves, rounds, K, p_volume, log, chains=4):
    """1024 TIP4P molecules, NPT Monte Carlo at 250 K and 1 bar (L = 31.289 A
    at the start, rc 9.0 A switched over 0.5 A, FP64 PairInfo energies):
    volume moves ln V +- 0.5%, block moves of every molecule of one
    Address_Assign cell (ndiv 3) displaced <= 0.15 A and turned <= 8 deg.
    Plain MC decides every proposal on the exact energy; delayed acceptance
    screens rounds of K proposals with the surrogate, trained here on the
    GPU on `configs` MD configurations at 240-300 K and five densities.
    Both chains start from the same equilibrated state.  Returns the
    exact-accepted moves per second of both."""
    from genesis6_amd import g6, surrogate as S
    n, L, T, P = 1024, 31.289, 250.0, 1.0
    temps = [240.0, 250.0, 260.0, 270.0, 280.0, 290.0, 300.0]
    scales = (0.97, 0.985, 1.0, 1.015, 1.03)
    per = max(1, -(-configs // (len(temps) * len(scales))))
    t0 = time.perf_counter()
    X, E = S.md_dataset(n, L, temps, per, 20, 20261207, scales=scales)
    t_data = time.perf_counter() - t0
    t0 = time.perf_counter()
    net, losses = S.train(X, E, epochs=epochs, batch=4096, lr=2e-3)
    t_train = time.perf_counter() - t0
    rows = len(X)
    del X, E
    log("npt: %d configurations (%.1f s), trained %d epochs (%.1f s), loss %.3f -> %.3f"
        % (rows // n, t_data, epochs, t_train, losses[0], losses[-1]))
    s = g6.System(g6.TIP4P, n, L, RC, TAPER, DT, prec=g6.FP32, skin=1.0)
    com, q = g6.lattice_start(n, L, 20261202 + 5)
    s.upload(com, q)
    s.set_nose(T, T * 2500.0)
    s.vreset(T, 20261202 + 5)
    s.step(3000)

    def start():
        s.npt_begin(T, P, 20261207, 0.15, 8.0, 0.005, p_volume, 9.0)
        s.npt_run(2000)  # the same equilibration for both chains
        return s.npt_run(0)

    a = start()
    t0 = time.perf_counter()
    b = s.npt_run(moves)
    el_p = time.perf_counter() - t0
    plain = {k: b[k] - a[k] for k in b}
    a = start()
    s.npt_run(2, mlp=net, K=K)   # graph capture
    a = s.npt_run(0)
    t0 = time.perf_counter()
    b = s.npt_run(rounds, mlp=net, K=K)
    el_d = time.perf_counter() - t0
    da = {k: b[k] - a[k] for k in b}
    _, _, Lf, Uf = s.npt_state()
    s.close()
    ens = npt_ensemble(net, moves, rounds, K, p_volume, chains, log)
    net.close()
    vp = plain["accepted"] / el_p
    vd = da["accepted"] / el_d
    return {"metric": "exact-accepted MC moves/s", "unit": "moves/s",
            "plain": vp, "delayed_acceptance": vd, "speedup": vd / vp,
            "K": K, "p_volume": p_volume,
            "plain_counts": dict(plain, seconds=el_p),
            "delayed_acceptance_counts": dict(da, seconds=el_d),
            "by_move_type": by_move_type(plain, da),
            "ensemble": ens,
            "training": {"configurations": rows // n, "rows": rows, "epochs": epochs,
                         "loss_first_epoch": float(losses[0]), "loss_last_epoch": float(losses[-1]),
                         "data_seconds": t_data, "train_seconds": t_train},
            "final_box_A": Lf}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--long-steps", type=int, default=10000,
                    help="a second timing of each TIP4P-4096 leg over this many steps "
                         "(0: skip)")
    ap.add_argument("--skin", type=float, default=1.8,
                    help="Verlet skin in A (0: rows rebuilt every step, as the reference); "
                         "1.8 A measured fastest for TIP4P-4096 among 0.7-3.0 A")
    ap.add_argument("--only", choices=["c2", "c3", "c4", "c5", "models", "mix"], default=None,
                    help="run one workload only (profiling)")
    ap.add_argument("--eager", action="store_true",
                    help="launch every kernel eagerly (no hipGraphs), as kernel traces are taken")
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="seconds of each CPU-baseline process")
    ap.add_argument("--cpu-procs", type=int, default=0,
                    help="concurrent CPU-baseline processes (0: the usable cores, at most 16)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--ladder-intervals", type=int, default=4,
                    help="timed 500-step intervals of the 64-replica ladder (0: skip)")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the ladder autotuner (skin and block assignment)")
    ap.add_argument("--tip5p-steps", type=int, default=500,
                    help="timed steps of the TIP5P-65536 system (0: skip)")
    ap.add_argument("--tip5p-skin", type=float, default=1.0,
                    help="Verlet skin of the TIP5P-65536 system in A (1.0-1.4 A measured "
                         "fastest for its cell tiles among 0.6-1.8 A)")
    ap.add_argument("--model-steps", type=int, default=2000,
                    help="timed steps of SPC/E, NvdE and ST2 at 4096 molecules (0: skip)")
    ap.add_argument("--mix-steps", type=int, default=2000,
                    help="timed steps of the 3584 water + 512 methanol mixture (0: skip)")
    ap.add_argument("--npt-configs", type=int, default=20000,
                    help="MD configurations the surrogate is trained on (0: skip NPT MC)")
    ap.add_argument("--npt-epochs", type=int, default=10)
    ap.add_argument("--npt-moves", type=int, default=20000, help="timed plain MC moves")
    ap.add_argument("--npt-rounds", type=int, default=2000,
                    help="timed delayed-acceptance rounds")
    ap.add_argument("--npt-k", type=int, default=32, help="proposals per round")
    ap.add_argument("--npt-pvol", type=float, default=1.0 / 28.0,
                    help="volume-move probability (1/28: one per sweep of the 27 cell blocks)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run five device-to-device copies of this many MiB and exit "
                         "(a plain copy kernel of known size, which calibrates the "
                         "FETCH_SIZE / WRITE_SIZE counters of rocprofv3 --pmc)")
    a = ap.parse_args()
    if a.eager:
        os.environ["G6_GRAPH_STEPS"] = "0"
    if a.copy_probe_mib > 0:
        import torch
        x = torch.ones(a.copy_probe_mib << 18, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        log("copy probe: 5 x %d MiB device-to-device" % a.copy_probe_mib)
        return
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the system is not sharded across GPUs "
              "(the single-system step runs on one device); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    only = a.only
    if only not in (None, "c2"):
        out = {"only": only}
        if only == "c3":
            out["tip5p_65536"] = tip5p_run(a)
        elif only == "c4":
            out["replica_ladders"] = ladder_run(a.ladder_intervals or 1, a.skin,
                                                tune=not a.no_tune, log=log)
        elif only == "c5":
            out["npt_mc"] = npt_run(a.npt_configs, a.npt_epochs, a.npt_moves, a.npt_rounds,
                                    a.npt_k, a.npt_pvol, log)
        elif only == "mix":
            out["mixture_4096"] = mixture_run(a.mix_steps, a.skin)
        else:
            out["other_models_4096"] = models_run(a.model_steps, a.skin)
        print(json.dumps(out))
        return
    legs, rl, st, th_eq = c2_run(a)
    f32 = legs["fp32"]
    fp64 = legs.get("fp64")
    ladder = tip5p = models = npt = mix = None
    if only is None and a.ladder_intervals > 0:
        ladder = ladder_run(a.ladder_intervals, a.skin, tune=not a.no_tune, log=log)
        log("ladder %.4g molecule-steps/s (%d replicas, %.3f ms/step), force kernel %.1f%% of "
            "FP32 peak" % (ladder["value"], ladder["replicas"], ladder["ms_per_step"],
                           100 * ladder["roofline"]["frac"]))
    if only is None and a.tip5p_steps > 0:
        tip5p = tip5p_run(a)
        log("TIP5P-65536 %.4g molecule-steps/s, force kernel %.1f%% of FP32 peak"
            % (tip5p["value"], 100 * tip5p["roofline"]["frac"]))
    if only is None and a.model_steps > 0:
        models = models_run(a.model_steps, a.skin)
        log("4096 molecules, " + ", ".join(
            "%s %.4g" % (k, v["value"]) for k, v in models.items()) + " molecule-steps/s")
    if only is None and a.mix_steps > 0:
        mix = mixture_run(a.mix_steps, a.skin)
        log("mixture 3584 water + 512 methanol %.4g (FP32) / %.4g (FP64) molecule-steps/s"
            % (mix["fp32"]["value"], mix["fp64"]["value"]))
    if only is None and a.npt_configs > 0:
        npt = npt_run(a.npt_configs, a.npt_epochs, a.npt_moves, a.npt_rounds, a.npt_k,
                      a.npt_pvol, log)
        log("NPT MC exact-accepted moves/s plain %.0f, delayed acceptance %.0f (%.2fx)"
            % (npt["plain"], npt["delayed_acceptance"], npt["speedup"]))
    cpu = None
    if not a.no_cpu:
        usable = len(os.sched_getaffinity(0))
        procs = a.cpu_procs or min(16, usable)
        v1, k1, e1, vall = cpu_baseline(st, a.cpu_seconds, procs)
        cal = reference_calibration()
        ratio = None
        if cal and "C2_tip4p_4096_nvt" in cal:
            ratio = cal["C2_tip4p_4096_nvt"]["oracle_over_reference"]
        cpu = {"value": vall, "unit": "molecule-steps/s", "cores": procs, "kind": "port",
               "sample": "TIP4P-4096 Nose NVT from the equilibrated 300 K state, FP64 oracle "
                         "(serial C restatement of the reference step, custom TIP4P kernel): "
                         "%d concurrent processes (independent replicas) for %.0f s each; "
                         "one process alone: %.4g molecule-steps/s (%d steps in %.1f s)"
                         % (procs, a.cpu_seconds, v1, k1, e1),
               "one_core": v1, "nproc": os.cpu_count(), "usable_cores": usable,
               "oracle_over_reference": ratio,
               "reference_equivalent_one_core": (v1 / ratio) if ratio else None,
               "reference_timing_here": cal}
        log("CPU oracle %.4g molecule-steps/s on %d processes, %.4g on one"
            % (vall, procs, v1))
    out = {
        "metric": "TIP4P-4096 molecule-steps/s",
        "value": f32["value"],
        "unit": "molecule-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": f32["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic: seeded lattice start, uniform random orientations, "
                "Maxwell-Boltzmann velocities (MT19937 seed %d), equilibrated untimed to a "
                "300 K liquid (10 x 2000 steps with velocity redraws, 10000 Nose steps)" % SEED,
        "config": {"workload": "TIP4P rigid water, 4096 molecules, 1.0 g/cm3, L=49.669 A, "
                               "Nose NVT 300 K, rc 9.0 A switched over 0.5 A, dt 2 fs, "
                               "FP32 pair arithmetic / FP64 accumulation",
                   "n_molecules": N, "pair_list_skin_A": a.skin,
                   "launch": "eager" if a.eager else "hipGraph replay",
                   "equilibrated_temperature_K": float(th_eq[3]),
                   "final_temperature_K": f32["temperature_K"],
                   "interacting_pairs_per_step": f32["interacting_pairs"]},
        "fp32_gpu": f32,
        "fp64_gpu": fp64,
        "speedup_vs_fp64_gpu": (f32["value"] / fp64["value"]) if fp64 else None,
        "roofline": rl,
        "cpu_baseline": cpu,
        "npt_mc": npt,
        "replica_ladders": ladder,
        "tip5p_65536": tip5p,
        "other_models_4096": models,
        "mixture_4096": mix,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
