"""Autotuner of the replica-ladder launch shapes.

A ladder's step time depends on two shapes that no formula settles well
across system sizes and densities:
  * the Verlet skin, i.e. the cell tile of the MD grid (cells of edge
    >= (rc + skin)/2) and the row length against the rebuild rate;
  * the block assignment of the ladder launches: every block of a system on
    one XCD (records, rows and cell lists stay in one L2) or the systems'
    blocks one after another (every system spread over all XCDs).
The tuner builds the ladder for each candidate skin from the same start,
times each assignment on the real workload, and returns the fastest.  The
assignment changes no result; the skin changes only when rows are rebuilt
(the interacting set is the reference's either way).
"""
import time

from . import g6


def tune_ladder(make_systems, skins=(0.6, 1.0, 1.4), maps=(0, 1), steps=500, warmup=100,
                log=None):
    """make_systems(skin) -> list of g6.System.  Returns (best, table): best
    is the fastest {skin, map, ms_per_step}; table holds every candidate."""
    table = []
    for skin in skins:
        systems = make_systems(skin)
        lad = g6.Ladder(systems)
        for mode in maps:
            lad.set_map(mode)
            lad.step(warmup)
            t0 = time.perf_counter()
            lad.step(steps)
            el = time.perf_counter() - t0
            row = {"skin": skin, "map": "xcd" if mode == g6.Ladder.MAP_XCD else "flat",
                   "map_mode": mode, "ms_per_step": el / steps * 1e3}
            table.append(row)
            if log:
                log("autotune: skin %.2f, %s map: %.4f ms/step" % (skin, row["map"],
                                                                   row["ms_per_step"]))
        lad.close()
        for s in systems:
            s.close()
    best = min(table, key=lambda r: r["ms_per_step"])
    return best, table
