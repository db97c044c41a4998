"""Benchmark of the triangle-counting core on one MI355X.

Default (--mode plain): exact triangle counting of RMAT scale 22 (edge factor
16, a=0.57 b=0.19 c=0.19, seeded, permuted, symmetrised).  One step = one
kt_count call on the symmetric CSR already resident in HBM: degree
orientation + oriented-CSR construction + degree-binned intersection,
producing the uint64 total and the uint64 per-vertex counts in HBM.

--mode batch: 8192 mixed G(n,m)/R-MAT graphs (n log-uniform in [64, 65536],
average degree U[4, 64]) counted by the hipGraph-replayed batch pipeline,
against one kt_count launch per graph on the same GPU.

--mode serve: a Poisson stream of requests through the continuous-batching
server (kt_serve_*), against one kt_count call per request.

--mode shard: rank 0 of RMAT scale 26 (edge factor 16) partitioned into 16
contiguous vertex ranges; one step = one kt_count_shard call (CeTriC local
phase over owned + ghost tails, then the contracted phase over the received
N+cut messages) on inputs resident in HBM, checked bit-exactly (local, type-3
and per-vertex counts) against the CPU rank simulation of the oracle.

--mode copy: a plain device-to-device copy of --copy-mib MiB (torch only, the
library is not loaded); the HBM-traffic counters of this kernel calibrate the
FETCH_SIZE / WRITE_SIZE readings of the counting kernels.

Prints one JSON line (see the "metric" field).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X published HBM3E bandwidth

# HBM bytes per k_tc_main launch, measured with two rocprofv3 --pmc runs of
# this workload under its autotuned knobs (FETCH_SIZE, then WRITE_SIZE, KiB
# per dispatch).  A plain 1 GiB device copy reads FETCH_SIZE = 524,308 KiB and
# WRITE_SIZE = 1,048,576 KiB on MI355X, so FETCH_SIZE is scaled by 2.000 and
# WRITE_SIZE by 1.000.  RMAT-22 ef16 seed 1 (min-side walk):
# 34,703,157 KiB x 2.000 + 3,177,102 KiB = 74.33 GB.
# Shard (RMAT-26 ef16 seed 1, 16 ranks, rank 0, CeTriC), the count phase's
# intersection launches (k_tc_main + k_tc_big, local and contracted, binning
# and k_tc_huge): FETCH_SIZE 292,888,269 KiB x 2.000 + WRITE_SIZE 15,960,809 KiB = 616.18 GB.
MEASURED_TRAFFIC = {("plain", 22, 16, 1): 74.33e9, ("shard-cetric", 26, 16, 1, 16, 0): 616.18e9}



def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--mode", choices=["plain", "unsorted", "batch", "serve", "shard", "copy"], default="plain")
    ap.add_argument("--requests", type=int, default=20000, help="serve mode: requests in the Poisson stream")
    ap.add_argument("--rate", type=float, default=0.0, help="serve mode: arrivals/s (0: 70%% of saturation)")
    ap.add_argument("--small-mix", action="store_true",
                    help="serve mode: the config-4 graph mix (n in [64, 65536]) instead of config 5")
    ap.add_argument("--no-extra", action="store_true",
                    help="plain mode: only the RMAT-22 line (no unsorted, batch and serve legs)")
    ap.add_argument("--max-batch-edges", type=int, default=0,
                    help="serve mode: edges per device batch (0: library default)")
    ap.add_argument("--copy-mib", type=int, default=1024, help="copy mode: bytes copied per launch, MiB")
    ap.add_argument("--ranks", type=int, default=16, help="shard mode: number of contiguous vertex ranges")
    ap.add_argument("--rank", type=int, default=0, help="shard mode: the rank counted")
    ap.add_argument("--ditric", action="store_true", help="shard mode: DiTriC (patric) attribution instead of CeTriC")
    ap.add_argument("--all-ranks", action="store_true",
                    help="shard mode: also count every rank's shard and check the partition identities")
    ap.add_argument("--graphs", type=int, default=8192, help="batch mode: number of graphs")
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--cpu-runs", type=int, default=3, help="plain mode: CPU baseline runs (minimum reported)")
    ap.add_argument("--params", default="",
                    help="intersection knobs, e.g. strategy=1,wave_max_deg=64 (kt_params fields)")
    ap.add_argument("--no-cpu-baseline", action="store_true",
                    help="omit the CPU reference timing (profiling runs)")
    ap.add_argument("--no-check", action="store_true", help="omit the total cross-check against the CPU run")
    return ap.parse_args()


def apply_params(ctx, args, key=None):
    """Installs the intersection knobs: --params if given, else the autotuned
    setting of the workload's class (katric_amd/tuned_params.json, nearest
    class); returns the knobs in effect."""
    if args.params:
        kv = dict(item.split("=") for item in args.params.split(","))
        ctx.set_params(**{k.strip(): int(v) for k, v in kv.items()})
    elif key is not None:
        from katric_amd import tune as T
        T.apply(ctx, key)
    return ctx.params()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_info():
    """The host the CPU baseline ran on: CPU model and logical CPU count."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count()}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only: the distributed path is represented by "
                                   "one rank's shard on one device (kt_count_shard); run with --gpus 1"}))
        sys.exit(2)
    if args.mode == "batch":
        return run_batch(args)
    if args.mode == "serve":
        return run_serve(args)
    if args.mode == "copy":
        return run_copy(args)
    if args.mode == "unsorted":
        return run_unsorted(args)
    if args.mode == "shard":
        if args.scale == 22 and "--scale" not in sys.argv:
            args.scale = 26
        return run_shard(args)
    return run_plain(args)


def run_shard(args):
    import katric_amd as K
    from katric_amd import shard as S

    t0 = time.time()
    g = K.gen_rmat(args.scale, args.edge_factor, seed=args.seed)
    n, m = g.n, g.m
    log(f"[bench] RMAT-{args.scale} ef{args.edge_factor}: n={n} m={m} gen {time.time() - t0:.1f}s")
    algo = S.DITRIC if args.ditric else S.CETRIC
    t0 = time.time()
    s = S.make_shard(g, args.ranks, args.rank, algo)
    mtot = len(s.msg_heads)
    log(f"[bench] shard {args.rank}/{args.ranks}: owned {s.n_local}, local heads {len(s.heads)}, "
        f"ghosts {len(s.ghost_ids)}, messages {len(s.msg_src)} ({mtot} ids), make {time.time() - t0:.1f}s")
    from katric_amd import tune as T
    ctx = K.Context(0)
    knobs = apply_params(ctx, args, T.graph_key(g.n, g.offsets))
    ds = S.upload_shard(ctx, s)
    X = s.n_local + len(s.ghost_ids)
    d_pv = ctx.alloc(max(X, 1) * 8)
    ctx.synchronize()
    res = None
    for _ in range(args.warmup):
        res = S.count_shard_device(ctx, ds, per_vertex_out=d_pv)
    st = ctx.stats()
    step_ms, cnt_ms = [], []
    ctx.synchronize()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        r2 = S.count_shard_device(ctx, ds, per_vertex_out=d_pv)
        assert r2 == res, "non-deterministic shard counts"
        cnt_ms.append(ctx.timing()["count_ms"])
    ctx.synchronize()
    ms_per_step = (time.perf_counter() - t_start) * 1e3 / args.steps
    tm = ctx.timing()
    pv = d_pv.download(np.uint64, X)
    loc, t3 = res
    Mk, W = st["oriented_edges"], st["intersection_size_local"] + st["intersection_size_global"]
    nl, ng, nmsg = s.n_local, len(s.ghost_ids), len(s.msg_src)
    # algorithmic bytes (unit = one oriented shard edge or one received cut id):
    # prep: read local CSR, ghost table and messages, write the oriented combined CSR;
    # count: both lists per intersected pair (4W), heads once, offsets and per-vertex counts
    b_pre = 8 * (nl + 1) + 4 * len(s.heads) + 12 * ng + 12 * nmsg + 4 * mtot + 8 * (X + 1) + 4 * Mk
    b_count = 4 * W + 4 * (Mk + mtot) + 8 * (X + 1) + 8 * X
    c_ms = float(np.mean(cnt_ms))
    # design bytes of the two intersection launches: every streamed T-list id
    # (4 B), every S entry once (4 B per local edge and received id), one walk
    # entry per local edge and received id (4 B id + 16 B T bounds), tail
    # offsets and per-vertex counts (16 B per tail)
    b_design = 4 * st["streamed_ids"] + 24 * (Mk + mtot) + 16 * X
    achieved = b_design / (c_ms * 1e-3) / 1e9
    units = Mk + mtot
    result = {
        "metric": f"shard edges/s (exact per-rank triangle counts, {'DiTriC' if args.ditric else 'CeTriC'} "
                  f"rank {args.rank} of {args.ranks}, RMAT-{args.scale} ef{args.edge_factor})",
        "value": units / (ms_per_step * 1e-3),
        "unit": "edges/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u32 ids / u64 counts",
        "data": "synthetic (seeded RMAT generator, permuted ids); shard inputs built on the host from the full graph",
        "config": {
            "params": knobs,
            "workload": f"RMAT-{args.scale} ef{args.edge_factor} seed={args.seed}, {args.ranks} contiguous ranges, "
                        f"rank {args.rank}: local CSR + ghost degrees + N+cut messages resident in HBM -> "
                        f"local / type-3 totals + per-vertex uint64 in HBM",
            "n": n, "m": m, "owned": nl, "ghosts": ng, "messages": nmsg, "message_ids": mtot,
            "oriented_shard_edges": Mk, "W": W, "wedge_checks": st["wedge_checks"], "edges_counted": units,
            "local_triangles": loc, "type3_triangles": t3,
            "device_ms": tm["total_ms"], "prep_ms": tm["prep_ms"], "count_ms": tm["count_ms"],
            "pipeline_algorithmic_GBs": (b_pre + b_count) / (ms_per_step * 1e-3) / 1e9,
        },
        "roofline": {
            "kernel": "k_tc_main (local + contracted phase)", "bound": "hbm", "achieved": achieved,
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
            "traffic": None if (args.params or args.ditric) else MEASURED_TRAFFIC.get(
                ("shard-cetric", args.scale, args.edge_factor, args.seed, args.ranks, args.rank)),
            "kernel_ms": c_ms,
            "bytes_per_launch": b_design,
            "bytes": "4 * streamed_ids + 24 * (oriented_shard_edges + message_ids) + 16 * tails",
            "algorithmic_effective_GBs": b_count / (c_ms * 1e-3) / 1e9,
            "algorithmic_bytes": "4 * W + 4 * (oriented_shard_edges + message_ids) + 8 * (tails + 1) + 8 * tails",
        },
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        thr = args.cpu_threads
        ts = time.perf_counter()
        ol, og, opv, _ = O.rank_sim(g, args.ranks, args.rank, O.DITRIC if args.ditric else O.CETRIC,
                                    threads=thr, per_vertex=True)
        cpu_s = time.perf_counter() - ts
        if not args.no_check:
            assert (ol, og) == (loc, t3), f"GPU ({loc}, {t3}) != CPU ({ol}, {og})"
            ids = s.vertex_ids()
            assert np.array_equal(pv, opv[ids]), "per-vertex shard counts differ from the CPU rank simulation"
            assert int(opv.sum(dtype=np.uint64)) == int(pv.sum(dtype=np.uint64)), "CPU counts outside the shard"
            result["config"]["checked"] = "local, type-3 and per-vertex counts equal the CPU rank simulation"
        result["cpu_baseline"] = {
            "value": units / cpu_s, "unit": "edges/s", "cores": thr, "kind": "port", **host_info(),
            "sample": f"full rank {args.rank} simulation (whole-graph degree orientation, ghost lists, message "
                      f"lists, local + contracted phase, {thr} OpenMP threads), {cpu_s:.2f} s",
        }
    if args.all_ranks:
        result["config"]["all_ranks"] = shard_sweep(args, ctx, g, algo, ds, d_pv, res)
    print(json.dumps(result), flush=True)


def shard_sweep(args, ctx, g, algo, ds, d_pv, rank_result):
    """Counts every rank on the GPU from one resident copy of the whole graph
    (kt_count_shard_graph builds each rank's inputs on the device) and checks
    the size-independent identities of the partition against the one-graph
    count: sum over ranks of local + type-3 triangles = total, and the
    per-vertex counts accumulated over all ranks = the per-vertex counts.  The
    timed rank's device-produced counts must equal its host-produced ones."""
    from katric_amd import shard as S

    for b in list(ds.bufs.values()) + [d_pv]:  # the timed rank's device inputs
        b.free()
    total, pv = ctx.count(g)
    dg = ctx.upload(g)
    d_acc = ctx.alloc(g.n * 8)
    d_acc.upload(np.zeros(g.n, np.uint64))
    per_rank, rank_ms, t_all = [], [], 0
    ctx.synchronize()
    t_sweep = time.perf_counter()
    for r in range(args.ranks):
        t0 = time.perf_counter()
        loc, t3, _ = S.count_shard_graph(ctx, dg, args.ranks, r, algo, per_vertex_out=d_acc, accumulate=True)
        tm = ctx.timing()
        per_rank.append([loc, t3])
        rank_ms.append(round(tm["total_ms"], 3))
        t_all += loc + t3
        log(f"[bench] rank {r}: local {loc} type-3 {t3} device {tm['total_ms']:.1f} ms "
            f"(producer + prep {tm['prep_ms']:.1f}, count {tm['count_ms']:.1f}), wall {(time.perf_counter() - t0) * 1e3:.0f} ms")
        if r == args.rank:
            assert (loc, t3) == rank_result, f"device producer {loc, t3} != host producer {rank_result}"
    sweep_ms = (time.perf_counter() - t_sweep) * 1e3
    acc = d_acc.download(np.uint64, g.n)
    assert t_all == total, f"sum over ranks {t_all} != total {total}"
    assert np.array_equal(acc, pv), "per-vertex counts summed over ranks differ from the one-graph counts"
    d_acc.free()
    dg.free()
    return {"per_rank_local_type3": per_rank, "per_rank_device_ms": rank_ms, "sweep_wall_ms": sweep_ms,
            "total": total, "producer": "device (kt_count_shard_graph from the resident full graph)",
            "checked": "sum over ranks of local + type-3 = one-graph total; per-vertex counts accumulated over ranks "
                       "= one-graph per-vertex counts; the timed rank equals its host-produced shard"}


def run_copy(args):
    import torch

    nbytes = args.copy_mib << 20
    src = torch.ones(nbytes // 4, dtype=torch.int32, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(args.warmup):
        dst.copy_(src)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(args.steps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    assert bool((dst == 1).all())
    print(json.dumps({"metric": "copy GB/s (read + write)", "value": 2 * nbytes / (ms * 1e-3) / 1e9, "unit": "GB/s",
                      "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
                      "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int32",
                      "data": "synthetic", "config": {"workload": f"device copy of {args.copy_mib} MiB",
                                                      "bytes_read": nbytes, "bytes_written": nbytes}}), flush=True)


def _percentiles(lat):
    lat = np.asarray(lat, dtype=np.float64)
    return float(np.percentile(lat, 50)), float(np.percentile(lat, 99))


def _poisson_times(n, rate, rng):
    return np.cumsum(rng.exponential(1.0 / rate, size=n))


def _sleep_until(t_abs):
    while True:
        d = t_abs - time.perf_counter()
        if d <= 0:
            return
        time.sleep(d - 0.0002 if d > 0.0005 else 0)


# Request mixes of the served stream: config 5 (n log-uniform in
# [2^10, 2^20], average degree U[4, 32]) and the small-graph mix of config 4
# (n in [64, 65536], average degree U[4, 64]), where batching pays most.
SERVE_MIXES = {
    "config5": dict(n_min=1 << 10, n_max=1 << 20, deg_min=4.0, deg_max=32.0, pool=256, baseline=1500,
                    saturation=4000),
    "small": dict(n_min=64, n_max=65536, deg_min=4.0, deg_max=64.0, pool=2048, baseline=4000, saturation=20000),
}


def serve_leg(args, ctx, mix, requests):
    """A Poisson stream of `requests` count requests of the given mix through
    the continuous-batching server (kt_serve_*), against one kt_count call per
    request on the same GPU.  Requests draw from a seeded pool of distinct
    graphs in page-locked memory; every request uploads its CSR again (the
    server keeps no results), so reuse of a pool graph saves no work."""
    import threading

    import katric_amd as K
    from katric_amd import batch as KB
    from katric_amd import serve as KS
    from katric_amd import tune as T

    mx = SERVE_MIXES[mix]
    t0 = time.time()
    pool_b = KB.gen_batch(mx["pool"], seed=args.seed + (0 if mix == "config5" else 17), n_min=mx["n_min"],
                          n_max=mx["n_max"], deg_min=mx["deg_min"], deg_max=mx["deg_max"])
    pool, owners = [], []
    for g in range(pool_b.num_graphs):
        view, own = KS.pinned_graph(pool_b.graph(g))
        pool.append(view)
        owners += own
    del pool_b
    pool_edges = np.array([g.m for g in pool], np.int64)
    log(f"[bench] serve {mix}: pool {len(pool)} graphs, mean m {pool_edges.mean():.0f}, gen {time.time() - t0:.1f}s")
    rng = np.random.default_rng(args.seed + 3)
    R = requests
    which = rng.integers(0, len(pool), size=R)
    sample = [pool[i] for i in range(min(64, len(pool)))]
    sample_off = np.concatenate([[0]] + [np.diff(gr.offsets.astype(np.int64)) for gr in sample]).cumsum()
    knobs = apply_params(ctx, args, T.batch_key(len(sample), sample_off))
    # reference totals: one kt_count per pool graph (this also warms the
    # one-graph path); a sample of them against the oracle
    want = np.array([ctx.count(g, per_vertex=False)[0] for g in pool], np.uint64)
    checked = 0
    if not args.no_check:
        from oracle import oracle as O
        for i in range(0, len(pool), max(1, len(pool) // 16)):
            assert O.run_shmem(pool[i], threads=args.cpu_threads, per_vertex=False)[0] == want[i]
            checked += 1

    # the load generator calls the C ABI directly with the pool graphs' raw
    # pointers, so that its own per-request cost (a few microseconds) stays
    # far below the server's
    import ctypes as C
    from katric_amd._ffi import lib as kt_lib
    L = kt_lib()
    raw = [(g.n, g.offsets.ctypes.data, g.heads.ctypes.data if g.nnz else None) for g in pool]
    submit_fn, wait_fn = L.kt_serve_submit, L.kt_serve_wait

    def run_stream(srv, n, arrivals):
        h = ctx.handle
        tickets = [0] * n

        def submitter():
            t = C.c_uint64(0)
            tref = C.byref(t)
            base = time.perf_counter()
            for i in range(n):
                if arrivals is not None:
                    _sleep_until(base + arrivals[i])
                gn, po, ph = raw[which[i]]
                rc = submit_fn(h, gn, po, ph, None, tref)
                assert rc == 0, rc
                tickets[i] = t.value

        th = threading.Thread(target=submitter)
        r = KS.KtServeResult()
        rref = C.byref(r)
        status = np.empty(n, np.int64)
        totals = np.empty(n, np.uint64)
        lat = np.empty(n)
        ts = time.perf_counter()
        th.start()
        for i in range(n):
            while tickets[i] == 0:
                time.sleep(0.0001)
            rc = wait_fn(h, tickets[i], -1.0, rref)
            assert rc == 0, rc
            status[i], totals[i], lat[i] = r.status, r.total, r.latency_ms
        th.join()
        wall = time.perf_counter() - ts
        assert (status == 0).all() and np.array_equal(totals, want[which[:n]]), \
            "served count differs from the one-graph count"
        return lat, wall

    srv = KS.Server(ctx, max_batch_edges=args.max_batch_edges)
    # saturation: every request submitted at once
    n_sat = min(R, mx["saturation"])
    run_stream(srv, min(n_sat, 500), None)  # warm-up (captures, arena growth)
    _, wall = run_stream(srv, n_sat, None)
    sat_rate = n_sat / wall
    rate = args.rate if args.rate > 0 else 0.7 * sat_rate
    arrivals = _poisson_times(R, rate, np.random.default_rng(args.seed + 1))
    lat, wall = run_stream(srv, R, arrivals)
    stats = srv.stats()
    srv.stop()
    p50, p99 = _percentiles(lat)
    edges = int(pool_edges[which[:R]].sum())
    # one kt_count per request, same GPU: saturated, then on the same arrivals
    nb = min(R, mx["baseline"])
    ts = time.perf_counter()
    for i in range(nb):
        ctx.count(pool[which[i]], per_vertex=False)
    b_sat = nb / (time.perf_counter() - ts)
    b_lat = []
    base = time.perf_counter()
    for i in range(nb):
        _sleep_until(base + arrivals[i])
        tot, _ = ctx.count(pool[which[i]], per_vertex=False)
        assert tot == want[which[i]]
        b_lat.append((time.perf_counter() - (base + arrivals[i])) * 1e3)
    b50, b99 = _percentiles(b_lat)
    for o in owners:
        o.free()
    return {
        "mix": mix,
        "graphs_per_s": R / wall, "edges_per_s": edges / wall, "requests": R, "edges": edges,
        "arrival_rate_per_s": rate, "latency_p50_ms": p50, "latency_p99_ms": p99,
        "saturation_graphs_per_s": sat_rate,
        "one_graph_per_launch_graphs_per_s": b_sat,
        "speedup_saturation_vs_one_graph_per_launch": sat_rate / b_sat,
        "one_graph_per_launch_latency_p50_ms_same_arrivals": b50,
        "one_graph_per_launch_latency_p99_ms_same_arrivals": b99,
        "one_graph_per_launch_requests": nb,
        "batches": stats["batches"], "mean_batch_graphs": stats["graphs"] / max(stats["batches"], 1),
        "pool_graphs": len(pool), "pool_mean_edges": float(pool_edges.mean()),
        "params": knobs,
        "checked": f"every served total equals the one-graph count; {checked} pool graphs equal the oracle",
        "pcie": "CSR uploaded per request from page-locked memory (PCIe inclusive)",
    }


def run_serve(args):
    import katric_amd as K

    ctx = K.Context(0)
    mix = "small" if args.small_mix else "config5"
    leg = serve_leg(args, ctx, mix, args.requests)
    result = {
        "metric": f"graphs/s served (exact triangle counts, Poisson stream, continuous batching, {mix} mix)",
        "value": leg["graphs_per_s"], "unit": "graphs/s", "n_gpus": 1, "steps": leg["requests"], "warmup": 0,
        "ms_per_step": 1e3 / leg["graphs_per_s"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u32 ids / u64 counts", "data": "synthetic (seeded pool of G(n,m)/RMAT graphs)",
        "config": {"workload": f"served stream, {mix} mix", **leg}, "cpu_baseline": None,
    }
    print(json.dumps(result), flush=True)


def batch_leg(args, ctx, graphs=8192):
    """Config 4: `graphs` mixed graphs counted by the hipGraph-replayed batch
    pipeline, against one kt_count launch per graph (CSR in HBM) on the same
    GPU; a sample of the per-graph totals against the oracle."""
    import katric_amd as K
    from katric_amd import batch as KB
    from katric_amd import tune as T

    t0 = time.time()
    b = KB.gen_batch(graphs, seed=args.seed)
    G, N, E = b.num_graphs, b.n_total, b.nnz // 2
    log(f"[bench] batch: {G} graphs, {N} vertices, {E} edges, gen {time.time() - t0:.1f}s")
    knobs = apply_params(ctx, args, T.batch_key(G, b.offsets))
    db = KB.upload_batch(ctx, b)
    pv = ctx.alloc(max(N, 1) * 8)
    tot = None
    for _ in range(max(args.warmup, 1)):  # the first call captures the pipeline
        tot = KB.count_batch(ctx, db, per_vertex_out=pv)
    ctx.synchronize()
    steps = max(args.steps // 2, 3)
    ts = time.perf_counter()
    step_ms = []
    for _ in range(steps):
        t1 = time.perf_counter()
        t2 = KB.count_batch(ctx, db, per_vertex_out=pv)  # returns after the totals are on the host
        step_ms.append((time.perf_counter() - t1) * 1e3)
        assert np.array_equal(t2, tot), "batch totals differ between replays"
    ctx.synchronize()
    batch_s = (time.perf_counter() - ts) / steps
    replays = KB.replays(ctx)
    # one kt_count launch per graph on per-graph CSRs (local ids) in HBM
    vb, off, heads = b.vertex_base.astype(np.int64), b.offsets, b.heads
    loc_off = np.empty(N + G, np.uint64)
    loc_heads = np.empty(max(len(heads), 1), np.uint32)
    for g in range(G):
        a, z = vb[g], vb[g + 1]
        e0, e1 = int(off[a]), int(off[z])
        loc_off[a + g:z + g + 1] = off[a:z + 1] - np.uint64(e0)
        loc_heads[e0:e1] = heads[e0:e1] - np.uint32(a)
    d_off = ctx.alloc(loc_off.nbytes)
    d_off.upload(loc_off)
    d_hd = ctx.alloc(loc_heads.nbytes)
    d_hd.upload(loc_heads)
    del loc_off, loc_heads
    single = np.zeros(G, np.uint64)
    for g in range(min(G, 16)):  # warm the per-graph path (same knobs)
        single[g] = ctx.count_device(int(vb[g + 1] - vb[g]), d_off.ptr + 8 * int(vb[g] + g),
                                     d_hd.ptr + 4 * int(off[vb[g]]), pv.ptr + 8 * int(vb[g]))
    ctx.synchronize()
    ts = time.perf_counter()
    done_ms = np.empty(G)
    for g in range(G):
        single[g] = ctx.count_device(int(vb[g + 1] - vb[g]), d_off.ptr + 8 * int(vb[g] + g),
                                     d_hd.ptr + 4 * int(off[vb[g]]), pv.ptr + 8 * int(vb[g]))
        done_ms[g] = (time.perf_counter() - ts) * 1e3
    ctx.synchronize()
    single_s = time.perf_counter() - ts
    b50, b99 = _percentiles(step_ms)
    assert np.array_equal(single, tot), "batched totals differ from one-graph-per-launch totals"
    checked = 0
    if not args.no_check:
        from oracle import oracle as O
        for g in range(0, G, max(1, G // 64)):
            assert O.run_shmem(b.graph(g), threads=args.cpu_threads, per_vertex=False)[0] == tot[g]
            checked += 1
    for buf in (d_off, d_hd, pv, db.offsets, db.heads):
        buf.free()
    return {
        "graphs_per_s": G / batch_s, "edges_per_s": E / batch_s, "ms_per_batch": batch_s * 1e3,
        "graphs": G, "vertices": N, "edges": E, "triangles": int(tot.sum()),
        "one_graph_per_launch_graphs_per_s": G / single_s, "one_graph_per_launch_s": single_s,
        "speedup_vs_one_graph_per_launch": single_s / batch_s,
        # all graphs of the batch submitted at once: a graph's latency is its
        # batch's wall time (batched) or the time until its own launch
        # returned (one graph per launch, in submission order)
        "latency_p50_ms": b50, "latency_p99_ms": b99,
        "one_graph_per_launch_latency_p50_ms": float(np.percentile(done_ms, 50)),
        "one_graph_per_launch_latency_p99_ms": float(np.percentile(done_ms, 99)),
        "graph_replays": replays, "params": knobs,
        "workload": f"{G} graphs, n log-uniform [64, 65536], avg degree U[4,64], G(n,m)/RMAT 50/50, "
                    f"seed={args.seed}; disjoint-union CSR resident in HBM -> per-graph uint64 totals",
        "checked": f"batched totals = one-graph-per-launch totals (all {G}); {checked} graphs equal the oracle",
    }


def run_batch(args):
    import katric_amd as K

    ctx = K.Context(0)
    leg = batch_leg(args, ctx, args.graphs)
    result = {
        "metric": f"graphs/s (exact triangle counts, {leg['graphs']} mixed G(n,m)/RMAT graphs, batched)",
        "value": leg["graphs_per_s"], "unit": "graphs/s", "n_gpus": 1, "steps": max(args.steps // 2, 3),
        "warmup": args.warmup, "ms_per_step": leg["ms_per_batch"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u32 ids / u64 counts", "data": "synthetic (seeded mixed batch generator)",
        "config": leg, "cpu_baseline": None,
    }
    print(json.dumps(result), flush=True)


def design_bytes(streamed, m, n, entries):
    """HBM bytes the hash-probe intersection must move per launch: every
    streamed T-list id once (4 B; the library's streamed_ids, close to the
    reference's wedge_checks when every oriented edge is intersected from its
    first vertex and far smaller with min_side), every S(x) once when its
    table is built (4m), per walk entry the neighbour id and the bounds of
    its T list (4 + 16 B; entries = 2m for the min-side walk over whole input
    rows, m otherwise), and the tail offsets and per-vertex counts (16n)."""
    return 4 * streamed + 4 * m + 20 * entries + 16 * n


def unsorted_leg(args, ctx, g, total):
    """The same graph with every adjacency list in descending order, resident
    in HBM: orientation finds the lists unsorted and sorts the oriented lists
    on the device (each list in place by a lane, a wave or a workgroup; a
    list above 4096 ids would send the graph to the LSD radix sort) -- the
    path the reference always pays (sort_neighborhoods,
    distributed_graph.h:577-593)."""
    deg = np.diff(g.offsets.astype(np.int64))
    rows = np.repeat(np.arange(g.n, dtype=np.int64), deg)
    pos = g.offsets[rows].astype(np.int64) + g.offsets[rows + 1].astype(np.int64) - 1 - np.arange(g.nnz)
    heads = np.empty_like(g.heads)
    heads[pos] = g.heads
    del rows, pos, deg
    import katric_amd as K
    dg = ctx.upload(K.Graph(g.n, g.offsets, heads))
    del heads
    d_pv = ctx.alloc(g.n * 8)
    for _ in range(2):
        assert ctx.count_device(dg, per_vertex_out=d_pv) == total
    ctx.synchronize()
    steps = max(args.steps // 4, 3)
    prep, dev = [], []
    ts = time.perf_counter()
    for _ in range(steps):
        assert ctx.count_device(dg, per_vertex_out=d_pv) == total, "unsorted input counts differ"
        tm = ctx.timing()
        prep.append(tm["prep_ms"])
        dev.append(tm["total_ms"])
    ctx.synchronize()
    ms = (time.perf_counter() - ts) * 1e3 / steps
    dg.free()
    d_pv.free()
    return {"ms_per_step": ms, "edges_per_s": g.m / (ms * 1e-3), "prep_ms": float(np.mean(prep)),
            "device_ms": float(np.mean(dev)), "steps": steps,
            "workload": f"the same RMAT-{args.scale} CSR with every list in descending order, resident in HBM; "
                        "orientation + segmented device sort of the oriented lists + intersection",
            "checked": "total equals the sorted input's"}


def run_unsorted(args):
    """RMAT-22 with every adjacency list in descending order (the unsorted leg
    of the default run, on its own)."""
    import katric_amd as K
    from katric_amd import tune as T

    g = K.gen_rmat(args.scale, args.edge_factor, seed=args.seed)
    ctx = K.Context(0)
    knobs = apply_params(ctx, args, T.graph_key(g.n, g.offsets))
    total = ctx.count(g, per_vertex=False)[0]
    leg = unsorted_leg(args, ctx, g, total)
    print(json.dumps({
        "metric": f"edges/s (exact triangle count, RMAT-{args.scale} ef{args.edge_factor}, unsorted lists)",
        "value": leg["edges_per_s"], "unit": "edges/s", "n_gpus": 1, "steps": leg["steps"], "warmup": 2,
        "ms_per_step": leg["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u32 ids / u64 counts", "data": "synthetic (seeded RMAT generator, permuted ids, lists reversed)",
        "config": {"params": knobs, **leg}, "cpu_baseline": None}), flush=True)


def run_plain(args):
    import katric_amd as K

    t0 = time.time()
    g = K.gen_rmat(args.scale, args.edge_factor, seed=args.seed)
    n, m, nnz = g.n, g.m, g.nnz
    log(f"[bench] RMAT-{args.scale} ef{args.edge_factor}: n={n} m={m} gen {time.time() - t0:.1f}s")
    checksum = g.checksum()

    from katric_amd import tune as T
    ctx = K.Context(0)
    knobs = apply_params(ctx, args, T.graph_key(g.n, g.offsets))
    dg = ctx.upload(g)          # the symmetric CSR resident in HBM
    d_pv = ctx.alloc(n * 8)     # per-vertex uint64 counts, in HBM
    ctx.synchronize()
    total = None
    for _ in range(args.warmup):
        total = ctx.count_device(dg, per_vertex_out=d_pv)
    ctx.synchronize()
    stats = ctx.stats()
    W = stats["intersection_size_local"]
    WC = stats["wedge_checks"]
    streamed = stats["streamed_ids"]
    entries = nnz if ctx.params()["min_side"] else m
    kern_ms = []
    ctx.synchronize()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        tt = ctx.count_device(dg, per_vertex_out=d_pv)
        assert tt == total, "non-deterministic total"
        kern_ms.append(ctx.timing()["intersect_ms"])
    ctx.synchronize()
    t_end = time.perf_counter()
    ms_per_step = (t_end - t_start) * 1e3 / args.steps
    tm = ctx.timing()
    pv_sum = int(d_pv.download(np.uint64, n).sum(dtype=np.uint64))
    assert pv_sum == 3 * total, "per-vertex counts inconsistent with the total"
    dg.free()
    d_pv.free()

    # roofline of k_tc_main on the bytes this design must move (unit = one
    # oriented edge: 4 B per streamed T-list id -- streamed_ids in all -- plus
    # 4 B of S(x) per kept slot, 20 B per walk entry and 16 B per vertex); the
    # reference's intersection_size_local figure, which charges both lists
    # per edge (B_count = 4W + 4m + 8(n+1) + 8n), is reported beside it
    b_design = design_bytes(streamed, m, n, entries)
    b_count = 4 * W + 4 * m + 8 * (n + 1) + 8 * n
    b_pre = 16 * (n + 1) + 12 * m
    k_ms = float(np.mean(kern_ms))
    achieved = b_design / (k_ms * 1e-3) / 1e9
    pipeline_gbs = (b_pre + b_count) / (ms_per_step * 1e-3) / 1e9

    result = {
        "metric": "edges/s (exact triangle count, RMAT-22 ef16, plain drop-in)",
        "value": m / (ms_per_step * 1e-3),
        "unit": "edges/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32 ids / u64 counts",
        "data": "synthetic (seeded RMAT generator, permuted ids)",
        "config": {
            "params": knobs,
            "workload": f"RMAT-{args.scale} ef{args.edge_factor} a=0.57 b=0.19 c=0.19 seed={args.seed}, "
                        f"symmetric CSR resident in HBM -> total + per-vertex uint64 in HBM",
            "n": n, "m": m, "triangles": total, "W": W, "wedge_checks": WC, "streamed_ids": streamed,
            "walk_entries": entries, "csr_checksum": f"{checksum:016x}",
            "device_ms": tm["total_ms"], "prep_ms": tm["prep_ms"], "count_ms": tm["count_ms"],
            "pipeline_reference_bytes_GBs": pipeline_gbs,
        },
        "roofline": {
            "kernel": "k_tc_main (intersection)",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            # measured under the autotuned knobs; other knobs move other bytes
            "traffic": None if args.params else MEASURED_TRAFFIC.get(("plain", args.scale, args.edge_factor, args.seed)),
            "kernel_ms": k_ms,
            "bytes_per_launch": b_design,
            "bytes": "4 * streamed_ids + 4 * m + 20 * walk_entries + 16 * n",
            "algorithmic_effective_GBs": b_count / (k_ms * 1e-3) / 1e9,
            "algorithmic_bytes": "4 * W + 4 * m + 8 * (n + 1) + 8 * n, W = intersection_size_local",
        },
        "cpu_baseline": None,
    }

    if not args.no_cpu_baseline:
        from oracle import oracle as O
        thr = args.cpu_threads
        runs = []
        for _ in range(args.cpu_runs):  # shmetric reports the minimum over its iterations
            ts = time.perf_counter()
            ct, _, _ = O.run_shmem(g, threads=thr, per_vertex=False)
            runs.append(time.perf_counter() - ts)
        cpu_s = min(runs)
        if not args.no_check:
            assert ct == total, f"GPU total {total} != CPU total {ct}"
        result["cpu_baseline"] = {
            "value": m / cpu_s,
            "unit": "edges/s",
            "cores": thr,
            "kind": "port",
            **host_info(),
            "sample": f"full RMAT-{args.scale} workload, minimum of {len(runs)} runs (orient + sort + "
                      f"merge-intersection local phase, {thr} OpenMP threads), {cpu_s:.2f} s",
        }
        # single-rank katric (CeTriC at P = 1 on one thread: with one rank
        # there are no ghosts and no messages, so it is the one-thread local
        # phase) on the same RMAT-22 graph: preprocessing in full, the local
        # phase over every 16th tail (ids are randomly permuted, so the sample
        # is representative); the full-workload time is prep + 16 x sample
        stride = 16
        ts = time.perf_counter()
        _, _, st1 = O.run_shmem_sample(g, stride=stride, per_vertex=False)
        one_s = time.perf_counter() - ts
        est_s = st1["prep_s"] + stride * st1["local_s"]
        result["cpu_baseline_single_rank"] = {
            "value": m / est_s, "unit": "edges/s", "cores": 1, "kind": "port", **host_info(),
            "sample": f"RMAT-{args.scale} (the same graph): orientation + sort of all {m} edges "
                      f"({st1['prep_s']:.2f} s) and the local phase over every {stride}th tail "
                      f"({st1['local_s']:.2f} s, {st1['sample_edges']} of {m} oriented edges), one thread; "
                      f"value = m / (prep + {stride} x local) = m / {est_s:.1f} s; measured {one_s:.1f} s",
        }
    if not args.no_extra:
        log("[bench] unsorted leg")
        result["unsorted"] = unsorted_leg(args, ctx, g, total)
        del g
        log("[bench] batch leg")
        result["batch"] = batch_leg(args, ctx)
        log("[bench] serve legs")
        result["serve"] = serve_leg(args, ctx, "config5", args.requests)
        result["serve_small"] = serve_leg(args, ctx, "small", args.requests)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
