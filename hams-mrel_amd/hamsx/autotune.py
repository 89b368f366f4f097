"""Autotuner of the per-size LU options (hamsx_set_option) for the serving
loop, and the lookup of its table.

For each panel-count bucket, depth class (deep / finite) and precision the
tuner builds a seeded request mesh of that size, keeps `in_flight` solves in
flight on as many contexts (the serving loop's pattern: one stream each,
submitted solves replayed from their recorded HIP graphs), and times the
frequency-solve throughput under every combination of
  * graphs          -- replay each context's recorded HIP graph of the solve
                       (one CPU call per solve instead of ~1,000 launches) or
                       launch eagerly (0, 1),
  * lu_select_rows  -- rows per pivot-selection block (256, 512),
  * lu_lookahead    -- the trailing update beyond the next panel on a second
                       stream (0, 1),
  * lu_outer_min_n  -- outer blocks of four panels from this order on (never,
                       or from 4,096 on, for buckets of at least 4,096 panels).
The fastest combination of each bucket is written to hamsx/tuning.json with
every trial's solves/s, and Server applies the nearest bucket's choice to
each request (options_for).  None of these options changes a result beyond
the LU's rounding: look-ahead is bit-identical, and selection-block size and
outer blocks give equally valid pivoted factorisations.

Run on the GPU:  python -m hamsx.autotune [--out PATH] [--quick]
(with eight hardware queues, GPU_MAX_HW_QUEUES, as bench.py runs: four
contexts of two streams each)

This has no counterpart in the reference, whose ZGETRF (MKL) has no such
choices (src/AssbMatxMulti.f90:110-114).
"""
from __future__ import annotations

import argparse
import collections
import itertools
import json
import math
import os
import time

import numpy as np

TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")
BUCKETS = (1000, 2000, 4000, 8000, 12000)
FINITE_DEPTH = 50.0


def depth_class(depth: float) -> str:
    return "deep" if depth < 0 else "finite"


def load_table(path: str = TABLE) -> dict:
    """The tuner's table, or an empty one when none has been written."""
    try:
        with open(path) as f:
            return json.load(f)
    except OSError:
        return {}


def options_for(table: dict, panels: int, depth: float, precision: str) -> dict:
    """LU options of the bucket nearest to `panels` (in log scale) with the
    same depth class and precision; {} when the table has none."""
    cands = [e for e in table.get("entries", []) if e["depth"] == depth_class(depth) and e["precision"] == precision]
    if not cands:
        return {}
    best = min(cands, key=lambda e: abs(math.log(panels / e["panels"])))
    return dict(best["best"])


def _mesh(panels: int, seed: int):
    from .configs import random_request
    from .mesh import combine
    case = random_request(np.random.default_rng(seed), total=panels)
    return combine(case.bodies, case.xr)


def measure(mesh, depth: float, precision: str, opts: dict, in_flight: int, solves: int, solvers=None):
    """Frequency solves per second of `solves` timed solves, `in_flight` at a
    time, after one untimed round (first solves record their graphs)."""
    from .solver import FrequencySolver
    own = solvers is None
    if own:
        solvers = [FrequencySolver(mesh, precision=precision) for _ in range(in_flight)]
    try:
        for s in solvers:
            s.set_options(**opts)
            s.set_waves(np.deg2rad([0.0, 90.0]))
        ws = np.linspace(0.3, 2.5, 2 * in_flight + solves)
        k = 0
        for _ in range(2):  # eager solve, then the recorded one
            for s in solvers:
                s.submit(float(ws[k]), depth)
                k += 1
            for s in solvers:
                s.collect()
        t0 = time.perf_counter()
        busy = collections.deque()
        for s in solvers[:solves]:
            s.submit(float(ws[k]), depth)
            k += 1
            busy.append(s)
        submitted, done = len(busy), 0
        while done < solves:
            s = busy.popleft()
            s.collect()
            done += 1
            if submitted < solves:
                s.submit(float(ws[k % len(ws)]), depth)
                k += 1
                submitted += 1
                busy.append(s)
        return solves / (time.perf_counter() - t0)
    finally:
        if own:
            for s in solvers:
                s.close()


def tune(buckets=BUCKETS, precisions=("fp64", "mixed"), depths=(-1.0, FINITE_DEPTH), in_flight=4, solves=8,
         seed=2025, log=print) -> dict:
    from .solver import FrequencySolver
    entries = []
    for panels, depth, prec in itertools.product(buckets, depths, precisions):
        mesh = _mesh(panels, seed)
        grid = {"graphs": (0, 1), "lu_select_rows": (256, 512), "lu_lookahead": (0, 1),
                "lu_outer_min_n": (0,) if mesh.nelem < 4096 else (0, 4096)}
        solvers = [FrequencySolver(mesh, precision=prec) for _ in range(in_flight)]
        trials = []
        try:
            for combo in itertools.product(*grid.values()):
                opts = dict(zip(grid.keys(), combo))
                if opts["graphs"] and opts["lu_lookahead"]:
                    continue  # recordings use the one-stream LU schedule
                sps = measure(mesh, depth, prec, opts, in_flight, solves, solvers)
                trials.append({"options": opts, "solves_per_s": sps})
                log(f"[autotune] N={mesh.nelem} {depth_class(depth)} {prec} {opts}: {sps:.2f} solves/s")
        finally:
            for s in solvers:
                s.close()
        best = max(trials, key=lambda t: t["solves_per_s"])
        entries.append({"panels": mesh.nelem, "bucket": panels, "depth": depth_class(depth), "precision": prec,
                        "best": best["options"], "solves_per_s": best["solves_per_s"], "trials": trials})
    return {"in_flight": in_flight, "timed_solves": solves, "seed": seed, "finite_depth_m": FINITE_DEPTH,
            "entries": entries}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=TABLE)
    ap.add_argument("--quick", action="store_true", help="two buckets only (a smoke run of the tuner)")
    args = ap.parse_args(argv)
    # two HIP streams per context (its own and the LU look-ahead's): as many
    # hardware queues, raised before the runtime starts, as bench.py and the
    # serving bench run with
    if int(os.environ.get("GPU_MAX_HW_QUEUES", "4") or 4) < 8:
        os.environ["GPU_MAX_HW_QUEUES"] = "8"
    buckets = (1000, 4000) if args.quick else BUCKETS
    t0 = time.perf_counter()
    table = tune(buckets=buckets, log=lambda s: print(s, flush=True))
    table["tuning_seconds"] = time.perf_counter() - t0
    with open(args.out, "w") as f:
        json.dump(table, f, indent=1)
        f.write("\n")
    print(f"[autotune] wrote {args.out} in {table['tuning_seconds']:.0f} s", flush=True)


if __name__ == "__main__":
    main()
