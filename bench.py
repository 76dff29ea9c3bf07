"""Benchmark: TIP4P-4096 rigid-water MD step on one MI355X.

Workload (the headline configuration): 4096 TIP4P molecules at 1.0 g/cm3
(L = 49.669 A), Nose NVT at 300 K, cutoff 9.0 A switched over 0.5 A, dt 2 fs,
FP32 pair arithmetic with FP64 accumulation; synthetic seeded lattice start
with uniform random orientations and Maxwell-Boltzmann velocities.  One
"step" is one full MD step (predict, cell grid + pair list, forces,
collect, correct, thermostat), all resident in HBM and replayed as hipGraphs.
Neighbour rows carry a 1.8 A Verlet skin by default (the fastest of 0.7-3.0 A
measured on this workload) and are rebuilt when a molecule has moved half of
it; the FP64 rc test on every listed pair keeps
the interacting set the reference's (rows rebuilt every step) exactly.

Prints ONE JSON line (rank 0, N=1).  The FP64 GPU path on the same workload
is timed too and reported beside it.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N, L, RC, TAPER, DT, TEMP = 4096, 49.669, 9.0, 0.5, 0.002, 300.0
NOSE_Q = 7.5e5          # @NOS2 mass: T0*tau^2/dt^2 with tau = 0.1 ps
SEED = 20261204
FLOP_PER_PAIR = 257     # TIP4P: 9 Coulomb x 21 + LJ 28 + 40 per molecule pair
FP32_PEAK = 157.3e12    # MI355X vector FP32 (vendor)
HBM_PEAK = 8.0e12


def make_system(prec, skin):
    from genesis6_amd import g6
    com, q = g6.lattice_start(N, L, SEED)
    s = g6.System(g6.TIP4P, N, L, RC, TAPER, DT, prec=prec, skin=skin)
    s.upload(com, q)
    s.set_nose(TEMP, NOSE_Q)
    s.vreset(TEMP, SEED)
    return s, com, q


def timed_run(s, steps, warmup):
    s.step(warmup)
    s.thermo()                       # synchronizes the stream
    t0 = time.perf_counter()
    s.step(steps)
    th = s.thermo()                  # synchronizes the stream
    t1 = time.perf_counter()
    return t1 - t0, th


def cpu_baseline(com, q, budget_s):
    """The FP64 oracle (serial C restatement of the reference step) on the
    same start, on one host core, for a bounded number of steps."""
    from oracle import pyoracle as O
    md = O.MDRun(O.TIP4P, com, q.T.copy(), L, RC, TAPER, DT, fast=True,
                 nose=(TEMP, NOSE_Q, 0.0))
    md.vreset(TEMP, SEED)
    md.step(1)
    t0 = time.perf_counter()
    k = 0
    while True:
        md.step(1)
        k += 1
        el = time.perf_counter() - t0
        if el >= budget_s:
            break
    return N * k / el, k, el


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
    out = dict(value=nmol * steps / el, unit="molecule-steps/s", replicas=len(systems),
               molecules=nmol, steps=steps, ms_per_step=el / steps * 1e3,
               skin_A=skin, block_map="xcd" if mode == g6.Ladder.MAP_XCD else "flat",
               autotune=table,
               exchange_trials=intervals * RX_TRIALS * len(groups), accepted=acc,
               pairs_per_step=float(th[:, 14].sum()),
               mean_temperature_K=float(th[:, 3].mean()))
    lad.close()
    for g in groups:
        g.close()
    for s in systems:
        s.close()
    return out


def tip5p_run(steps, skin):
    """65,536 TIP5P molecules at 1.0 g/cm3 (L = 125.157 A), Nose NVT 275 K,
    rc 9.0 A switched over 0.5 A, FP32 pairs."""
    from genesis6_amd import g6
    n, Lb, T = 65536, 125.157, 275.0
    s = g6.System(g6.TIP5P, n, Lb, RC, TAPER, DT, prec=g6.FP32, skin=skin)
    com, q = g6.lattice_start(n, Lb, 20261202 + 3)
    s.upload(com, q)
    s.set_nose(T, T * 2500.0)
    s.vreset(T, 20261202 + 3)
    el, th = timed_run(s, steps, 50)
    s.close()
    return dict(value=n * steps / el, unit="molecule-steps/s", molecules=n, steps=steps,
                ms_per_step=el / steps * 1e3, pairs_per_step=float(th[14]),
                temperature_K=float(th[3]))


def models_run(steps, skin):
    """The other rigid models on the headline workload (4096 molecules at
    1.0 g/cm3, L = 49.669 A, Nose NVT 300 K, rc 9.0 A switched over 0.5 A,
    FP32 pairs): SPC/E, NvdE and ST2 (its custom kernel), molecule-steps/s."""
    from genesis6_amd import g6
    n, Lb, T = 4096, 49.669, 300.0
    out = {}
    for name, mid in (("spce", g6.SPCE), ("nvde", g6.NVDE), ("st2", g6.ST2)):
        s = g6.System(mid, n, Lb, RC, TAPER, DT, prec=g6.FP32, skin=skin)
        com, q = g6.lattice_start(n, Lb, 20261204)
        s.upload(com, q)
        s.set_nose(T, T * 2500.0)
        s.vreset(T, 20261204)
        el, th = timed_run(s, steps, 50)
        s.close()
        out[name] = dict(value=n * steps / el, unit="molecule-steps/s", steps=steps,
                         ms_per_step=el / steps * 1e3, pairs_per_step=float(th[14]),
                         temperature_K=float(th[3]))
    return out


def npt_run(configs, epochs, moves, rounds, K, p_volume, log):
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
    net.close()
    vp = plain["accepted"] / el_p
    vd = da["accepted"] / el_d
    return {"metric": "exact-accepted MC moves/s", "unit": "moves/s",
            "plain": vp, "delayed_acceptance": vd, "speedup": vd / vp,
            "K": K, "p_volume": p_volume,
            "plain_counts": dict(plain, seconds=el_p),
            "delayed_acceptance_counts": dict(da, seconds=el_d),
            "training": {"configurations": rows // n, "rows": rows, "epochs": epochs,
                         "loss_first_epoch": float(losses[0]), "loss_last_epoch": float(losses[-1]),
                         "data_seconds": t_data, "train_seconds": t_train},
            "final_box_A": Lf}


def traffic_from_profile():
    """HBM bytes per force-kernel launch from the committed rocprofv3 --pmc
    summary (FETCH_SIZE + WRITE_SIZE, copy-kernel calibrated), if present."""
    p = os.path.join(ROOT, "profiles", "md_main_tip4p4096_traffic.json")
    try:
        return json.load(open(p))["bytes_per_launch"]
    except Exception:
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--skin", type=float, default=1.8,
                    help="Verlet skin in A (0: rows rebuilt every step, as the reference); "
                         "1.8 A measured fastest for TIP4P-4096 among 0.7-3.0 A")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--ladder-intervals", type=int, default=4,
                    help="timed 500-step intervals of the 64-replica ladder (0: skip)")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the ladder autotuner (skin and block assignment)")
    ap.add_argument("--tip5p-steps", type=int, default=500,
                    help="timed steps of the TIP5P-65536 system (0: skip)")
    ap.add_argument("--model-steps", type=int, default=2000,
                    help="timed steps of SPC/E, NvdE and ST2 at 4096 molecules (0: skip)")
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
    if a.copy_probe_mib > 0:
        import torch
        x = torch.ones(a.copy_probe_mib << 18, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        print("copy probe: 5 x %d MiB device-to-device" % a.copy_probe_mib, file=sys.stderr)
        return
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the system is not sharded across GPUs "
              "(the single-system step runs on one device); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    from genesis6_amd import g6
    print("bench: TIP4P-%d FP32-pair, %d warmup + %d steps" % (N, a.warmup, a.steps),
          file=sys.stderr, flush=True)
    s, com, q = make_system(g6.FP32, a.skin)
    el, th = timed_run(s, a.steps, a.warmup)
    value = N * a.steps / el
    # per-kernel timing with HIP events on the library stream: each phase of
    # eager steps bracketed (the force kernel with its dispatch: what the
    # roofline uses), and the force kernel alone, back to back in one graph
    # on warm caches (a lower bound: in a step its inputs come fresh from the
    # Gear kernel)
    nt = min(200, max(20, a.steps // 50))
    t = s.step_timed(nt)
    npair = t[4]
    s.step(10)
    t_b2b_ms = s.time_force(200)
    t_force_ms = t[0]
    t_force_s = t_force_ms * 1e-3
    achieved = npair * FLOP_PER_PAIR / t_force_s
    print("bench: FP32 %.4g molecule-steps/s, T=%.1f K, force kernel %.1f us "
          "(%.1f us back to back), npair %.0f"
          % (value, th[3], t_force_ms * 1e3, t_b2b_ms * 1e3, npair), file=sys.stderr,
          flush=True)
    s.close()
    fp64 = None
    if not a.no_fp64:
        s64, _, _ = make_system(g6.FP64, a.skin)
        k64 = max(100, a.steps // 5)
        el64, _ = timed_run(s64, k64, a.warmup)
        s64.step(10)
        fp64 = dict(value=N * k64 / el64, steps=k64, force_kernel_us=s64.time_force(100) * 1e3)
        s64.close()
        print("bench: FP64 %.4g molecule-steps/s" % fp64["value"], file=sys.stderr, flush=True)
    ladder = None
    if a.ladder_intervals > 0:
        ladder = ladder_run(a.ladder_intervals, a.skin, tune=not a.no_tune,
                            log=lambda m: print("bench: " + m, file=sys.stderr, flush=True))
        print("bench: ladder %.4g molecule-steps/s (%d replicas, %.3f ms/step)"
              % (ladder["value"], ladder["replicas"], ladder["ms_per_step"]),
              file=sys.stderr, flush=True)
    tip5p = None
    if a.tip5p_steps > 0:
        tip5p = tip5p_run(a.tip5p_steps, a.skin)
        print("bench: TIP5P-65536 %.4g molecule-steps/s" % tip5p["value"], file=sys.stderr,
              flush=True)
    models = None
    if a.model_steps > 0:
        models = models_run(a.model_steps, a.skin)
        print("bench: 4096 molecules, " + ", ".join(
            "%s %.4g" % (k, v["value"]) for k, v in models.items()) + " molecule-steps/s",
            file=sys.stderr, flush=True)
    npt = None
    if a.npt_configs > 0:
        npt = npt_run(a.npt_configs, a.npt_epochs, a.npt_moves, a.npt_rounds, a.npt_k,
                      a.npt_pvol, lambda m: print("bench: " + m, file=sys.stderr, flush=True))
        print("bench: NPT MC exact-accepted moves/s plain %.0f, delayed acceptance %.0f (%.2fx)"
              % (npt["plain"], npt["delayed_acceptance"], npt["speedup"]), file=sys.stderr,
              flush=True)
    cpu = None
    if not a.no_cpu:
        v, k, e = cpu_baseline(com, q, a.cpu_seconds)
        cpu = {"value": v, "unit": "molecule-steps/s", "cores": 1, "kind": "port",
               "sample": "TIP4P-4096 Nose NVT, %d MD steps (%.1f s) from the same start, "
                         "FP64 oracle (serial C restatement of the reference step, "
                         "custom TIP4P kernel)" % (k, e)}
    out = {
        "metric": "TIP4P-4096 molecule-steps/s",
        "value": value,
        "unit": "molecule-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": el / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic: seeded lattice start, uniform random orientations, "
                "Maxwell-Boltzmann velocities (MT19937 seed %d)" % SEED,
        "config": {"workload": "TIP4P rigid water, 4096 molecules, 1.0 g/cm3, L=49.669 A, "
                               "Nose NVT 300 K, rc 9.0 A switched over 0.5 A, dt 2 fs, "
                               "FP32 pair arithmetic / FP64 accumulation",
                   "n_molecules": N, "pair_list_skin_A": a.skin,
                   "final_temperature_K": float(th[3]),
                   "interacting_pairs_per_step": float(npair)},
        "fp64_gpu": fp64,
        "speedup_vs_fp64_gpu": (value / fp64["value"]) if fp64 else None,
        "roofline": {"bound": "valu", "achieved": achieved / 1e12, "peak": FP32_PEAK / 1e12,
                     "unit": "TFLOP/s", "frac": achieved / FP32_PEAK,
                     "traffic": traffic_from_profile(),
                     "kernel": "k_md_main<TIP4P,float>",
                     "kernel_ms": t_force_ms, "kernel_back_to_back_ms": t_b2b_ms,
                     "flop_per_pair": FLOP_PER_PAIR,
                     "pairs_per_launch": float(npair),
                     "per_molecule_kernels_ms": t[2],
                     "eager_step_ms": t[3]},
        "cpu_baseline": cpu,
        "npt_mc": npt,
        "replica_ladders": ladder,
        "tip5p_65536": tip5p,
        "other_models_4096": models,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
