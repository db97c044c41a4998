"""Reference-shaped statistics: the JSON document KaTric writes with
--json-output (apps/katric.cpp:278-297), filled from this library's counters
and device timings, so that tooling written for the reference's output can
read runs of this one.

Layout (field names and nesting as in include/cetric/statistics.h):
  {"stats": [ {"local_statistics": [LocalStatistics per rank],
               "global_wall_time", "triangles", "counted_triangles"} ],
   "io_time" (optional), "config": {...}}
LocalStatistics (statistics.h:193-249): timings in seconds, the work
counters wedge_checks / intersection_size_local / intersection_size_global,
local_triangles / type3_triangles, wedges, plus the message and load-balancing
blocks, which are zero here (one GPU exchanges no messages).
"""
from __future__ import annotations

import json
from typing import Dict, List, Optional

import numpy as np


def _messages() -> Dict[str, int]:
    # MessageStatistics (statistics.h:38-72)
    return {"sent_messages": 0, "received_messages": 0, "send_volume": 0, "receive_volume": 0}


def _preprocessing(orientation_s: float = 0.0, sorting_s: float = 0.0) -> dict:
    # PreprocessingStatistics (statistics.h:76-100)
    return {"degree_exchange_time": 0.0, "orientation_time": orientation_s, "sorting_time": sorting_s,
            "message_statistics": _messages()}


def _load_balancing() -> dict:
    # LoadBalancingStatistics (statistics.h:102-126)
    return {"cost_function_evaluation_time": 0.0, "cost_function_communication_time": 0.0,
            "computation_time": 0.0, "redistribution_time": 0.0, "phase_time": 0.0,
            "message_statistics": _messages()}


def wedges(offsets: np.ndarray) -> int:
    """sum over v of deg(v) * (deg(v) - 1) / 2, as shmetric reports it (apps/shmetric.cpp:121-129)."""
    d = np.diff(np.asarray(offsets, dtype=np.int64))
    return int((d * (d - 1) // 2).sum())


def local_statistics(stats: dict, timing: dict, rank: int = 0, wedge_count: int = 0) -> dict:
    """One rank's LocalStatistics from Context.stats() / Context.timing() of
    the call that counted it (kt_count or kt_count_shard)."""
    ms = 1e-3
    local_phase = timing.get("intersect_ms", 0.0) * ms
    count_phase = timing.get("count_ms", 0.0) * ms
    total = timing.get("total_ms", 0.0) * ms
    return {
        "rank": rank,
        "io_time": 0.0,
        "ghost_rank_gather": 0.0,
        "preprocessing_local_phase": _preprocessing(orientation_s=timing.get("prep_ms", 0.0) * ms),
        "preprocessing_global_phase": _preprocessing(),
        "primary_load_balancing": _load_balancing(),
        "secondary_load_balancing": _load_balancing(),
        "local_phase_time": local_phase,
        "contraction_time": 0.0,
        "global_phase_time": max(0.0, count_phase - local_phase) if stats.get("type3_triangles") else 0.0,
        "reduce_time": 0.0,
        "preprocessing_time": timing.get("prep_ms", 0.0) * ms,
        "preprocessing_global_time": 0.0,
        "message_statistics": _messages(),
        "global_phase_threshold": 0,
        "wedges": int(wedge_count),
        "local_wall_time": total,
        "local_time": total,
        "skipped_nodes": int(stats.get("skipped_nodes", 0)),
        "wedge_checks": int(stats.get("wedge_checks", 0)),
        "intersection_size_local": int(stats.get("intersection_size_local", 0)),
        "intersection_size_global": int(stats.get("intersection_size_global", 0)),
        "nodes_parallel2d": 0,
        "local_triangles": int(stats.get("local_triangles", 0)),
        "type3_triangles": int(stats.get("type3_triangles", 0)),
    }


def statistics(locals_: List[dict]) -> dict:
    """Statistics after the reduction to rank 0 (statistics.h:300-358)."""
    tri = sum(s["local_triangles"] + s["type3_triangles"] for s in locals_)
    return {"local_statistics": locals_,
            "global_wall_time": max((s["local_wall_time"] for s in locals_), default=0.0),
            "triangles": tri, "counted_triangles": tri}


def document(runs: List[dict], config: Optional[dict] = None, io_time: Optional[float] = None) -> dict:
    """The --json-output document: {"stats": [...], "io_time"?, "config": {...}}."""
    doc = {"stats": runs}
    if io_time is not None:
        doc["io_time"] = io_time
    doc["config"] = dict(config or {})
    return doc


def run_shmem_json(ctx, g, config: Optional[dict] = None, iterations: int = 1) -> str:
    """Counts `g` `iterations` times (like shmetric's repetitions) and returns
    the reference-shaped JSON text."""
    runs = []
    w = wedges(g.offsets)
    for _ in range(iterations):
        ctx.count(g, per_vertex=False)
        runs.append(statistics([local_statistics(ctx.stats(), ctx.timing(), 0, w)]))
    cfg = {"algorithm": "cetric", "num_threads": 1, "intersection_method": "device"}
    cfg.update(config or {})
    return json.dumps(document(runs, cfg), indent=2)
