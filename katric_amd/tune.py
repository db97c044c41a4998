"""Autotuner for the intersection knobs (kt_params).

The knobs -- wave/workgroup bin thresholds, how a team walks its T lists
(streamed or grouped), from which side each edge is intersected (min_side),
queue order (degree or vertex), persistent workgroups per CU -- move work between
kernels and code paths and never change a count.  Their best setting depends
on the degree profile of the input, so the tuner

  * reduces an input to a *graph class*: log2 buckets of the vertex count,
    of the average degree and of the largest degree (for a batch: of the
    graph count, the average degree and the largest degree);
  * times every candidate setting on the device (HIP events inside the
    library: binning + intersection + everything after orientation), keeping
    the fastest by median over a few repetitions;
  * stores the winner per class in a JSON table.

The table shipped in the package (tuned_params.json) was measured on MI355X;
a user table (KATRIC_TUNE_TABLE, default ~/.cache/katric_amd/tuned_params.json)
takes precedence and receives new results.  `apply(ctx, key)` installs the
tuned knobs of a class on a context, falling back to the library defaults.
"""
from __future__ import annotations

import itertools
import json
import math
import os
import time
from typing import Callable, Dict, Iterable, List, Optional

import numpy as np

PACKAGE_TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned_params.json")
DEFAULTS = {"wave_max_deg": 128, "group_max_deg": 2048, "grid_per_cu": 4, "strategy": 0, "order": 0, "min_side": 1}


def user_table_path() -> str:
    return os.environ.get("KATRIC_TUNE_TABLE",
                          os.path.join(os.path.expanduser("~"), ".cache", "katric_amd", "tuned_params.json"))


def _lg(x: float) -> int:
    return int(math.ceil(math.log2(max(x, 1.0))))


def graph_key(n: int, offsets: np.ndarray) -> str:
    """Class of a symmetric CSR: plain/n<lg n>/d<lg avg deg>/D<lg max deg>."""
    deg = np.diff(np.asarray(offsets, dtype=np.int64))
    avg = float(deg.mean()) if n else 0.0
    mx = int(deg.max()) if n else 0
    return f"plain/n{_lg(n)}/d{_lg(avg)}/D{_lg(mx)}"


def batch_key(num_graphs: int, offsets: np.ndarray) -> str:
    """Class of a disjoint-union batch: batch/g<lg graphs>/d<lg avg deg>/D<lg max deg>."""
    deg = np.diff(np.asarray(offsets, dtype=np.int64))
    avg = float(deg.mean()) if len(deg) else 0.0
    mx = int(deg.max()) if len(deg) else 0
    return f"batch/g{_lg(num_graphs)}/d{_lg(avg)}/D{_lg(mx)}"


STAGES = [
    # stage 1: how lists are walked, from which side of each edge, and in
    # which order tails are taken
    {"strategy": [0, 1, 2, 3, 4, 5], "order": [0, 1], "min_side": [1, 0]},
    # stage 2: the bin thresholds and the grid, under the best stage-1 choice
    {"wave_max_deg": [32, 64, 128, 256], "group_max_deg": [1024, 2048], "grid_per_cu": [2, 3, 4]},
]


def candidates(stage: dict, base: Dict[str, int]) -> List[Dict[str, int]]:
    """Every combination of one stage's knobs on top of `base` (combinations
    the library rejects -- a workgroup bin below the wave bin -- are skipped)."""
    keys = sorted(stage)
    out = []
    for vals in itertools.product(*(stage[k] for k in keys)):
        p = dict(base)
        p.update(zip(keys, vals))
        if p["group_max_deg"] >= p["wave_max_deg"]:
            out.append(p)
    return out


def load_table(path: Optional[str] = None) -> Dict[str, dict]:
    table: Dict[str, dict] = {}
    for p in ([PACKAGE_TABLE, user_table_path()] if path is None else [path]):
        if os.path.exists(p):
            with open(p) as f:
                table.update(json.load(f).get("classes", {}))
    return table


def save_result(key: str, entry: dict, path: Optional[str] = None) -> str:
    p = path or user_table_path()
    os.makedirs(os.path.dirname(p) or ".", exist_ok=True)
    data = {"device": "MI355X (gfx950)", "classes": {}}
    if os.path.exists(p):
        with open(p) as f:
            data = json.load(f)
    data.setdefault("classes", {})[key] = entry
    tmp = p + ".tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, p)
    return p


def lookup(key: str, table: Optional[Dict[str, dict]] = None) -> Dict[str, int]:
    """Tuned knobs of a class, or the nearest tuned class of the same kind
    (smallest distance over the log2 buckets), or the library defaults."""
    table = load_table() if table is None else table
    if key in table:
        return dict(table[key]["params"])
    kind = key.split("/")[0]

    def coords(k):
        return [int(part[1:]) for part in k.split("/")[1:]]

    best, dist = None, None
    for k in table:
        if k.split("/")[0] != kind:
            continue
        d = sum(abs(a - b) for a, b in zip(coords(k), coords(key)))
        if dist is None or d < dist:
            best, dist = k, d
    return dict(table[best]["params"]) if best is not None else dict(DEFAULTS)


def apply(ctx, key: str, table: Optional[Dict[str, dict]] = None) -> Dict[str, int]:
    p = lookup(key, table)
    ctx.set_params(**p)
    return p


def tune(ctx, run: Callable[[], float], reps: int = 3, key: Optional[str] = None, save: bool = False,
         path: Optional[str] = None, log=None, stages=None) -> dict:
    """Times `run` (one counting call on `ctx` returning its device
    milliseconds) stage by stage: each stage tries every combination of its
    knobs on top of the best setting so far and keeps the fastest (median of
    `reps` runs after one warm-up run).  Returns {"params", "ms",
    "default_ms", "trials"} and installs the winner on `ctx`; with `save`,
    stores it under `key` in the user table (or `path`)."""
    saved = ctx.params()
    best = dict(DEFAULTS)
    trials = 0

    def measure(p):
        ctx.set_params(**p)
        run()  # warm-up (and, for batches, the capture of this setting)
        return float(np.median([run() for _ in range(reps)]))

    try:
        default_ms = best_ms = measure(best)
        for stage in (stages or STAGES):
            for p in candidates(stage, best):
                ms = measure(p)
                trials += 1
                if log:
                    log(f"[tune] {p} {ms:.3f} ms")
                if ms < best_ms:
                    best, best_ms = p, ms
    finally:
        ctx.set_params(**saved)
    entry = {"params": best, "ms": best_ms, "default_ms": default_ms, "trials": trials,
             "measured": time.strftime("%Y-%m-%d")}
    ctx.set_params(**best)
    if save and key:
        save_result(key, entry, path)
    return entry


def tune_graph(ctx, g, reps: int = 3, save: bool = False, path: Optional[str] = None, log=None) -> dict:
    """Tunes kt_count on a graph resident in HBM (the graph is uploaded once)."""
    dg = ctx.upload(g)
    pv = ctx.alloc(max(g.n, 1) * 8)

    def run():
        ctx.count_device(dg, per_vertex_out=pv)
        return ctx.timing()["count_ms"]

    key = graph_key(g.n, g.offsets)
    e = tune(ctx, run, reps, key, save, path, log)
    e["key"] = key
    return e


def tune_batch(ctx, b, reps: int = 3, save: bool = False, path: Optional[str] = None, log=None) -> dict:
    """Tunes kt_batch_count on a batch resident in HBM; each setting is
    captured once and then replayed."""
    from . import batch as KB

    db = KB.upload_batch(ctx, b)
    pv = ctx.alloc(max(b.n_total, 1) * 8)

    def run():
        KB.count_batch(ctx, db, per_vertex_out=pv)
        return ctx.timing()["count_ms"]  # the replayed device pipeline

    key = batch_key(b.num_graphs, b.offsets)
    e = tune(ctx, run, reps, key, save, path, log)
    e["key"] = key
    return e


def main(argv=None) -> None:
    """Tunes the standard workloads on device 0 and writes the table:
    python -m katric_amd.tune [--out PATH] [--workloads rmat22,batch,serve]"""
    import argparse
    import sys

    from . import Context, gen_gnm, gen_rmat
    from . import batch as KB

    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=PACKAGE_TABLE)
    ap.add_argument("--workloads", default="rmat22,gnm,batch,serve")
    ap.add_argument("--reps", type=int, default=3)
    args = ap.parse_args(argv)

    def log(m):
        print(m, file=sys.stderr, flush=True)

    ctx = Context(0)
    for wl in args.workloads.split(","):
        if wl == "rmat22":  # RMAT scale 22, edge factor 16 (the plain drop-in workload)
            e = tune_graph(ctx, gen_rmat(22, 16, seed=1), args.reps, True, args.out, log)
        elif wl == "gnm":  # G(n,m), n = 2^16, 2^20 pairs
            e = tune_graph(ctx, gen_gnm(1 << 16, 1 << 20, 1), args.reps, True, args.out, log)
        elif wl == "batch":  # 8192 mixed small graphs (the batched workload)
            e = tune_batch(ctx, KB.gen_batch(8192, seed=1), args.reps, True, args.out, log)
        elif wl == "serve":  # a device batch of served requests: 64 graphs, n in [2^10, 2^20]
            b = KB.gen_batch(64, seed=2, n_min=1 << 10, n_max=1 << 20, deg_min=4.0, deg_max=32.0)
            e = tune_batch(ctx, b, args.reps, True, args.out, log)
        else:
            raise SystemExit(f"unknown workload {wl}")
        print(json.dumps({"workload": wl, **e}), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
