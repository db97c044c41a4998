"""Every benchmark configuration at its stated size, on the GPU, against the
oracle: totals, per-vertex counts, per-rank counts and work counters.

  1. G(n,m), n = 2^16, 2^20 endpoint pairs: shmetric and single-rank katric
  2. R-MAT scale 22, edge factor 16, permuted: sorted and unsorted lists
  3. R-MAT scale 26 in 16 ranges, rank 0: CeTriC and DiTriC, host- and
     device-produced shard inputs
  4. 8192 mixed G(n,m) / R-MAT graphs in one batch
  5. a served Poisson stream of config-5 graphs, and one of config-4 graphs

Each graph is generated once; the oracle runs on the host's cores (16 at
most).  Reference: shared_memory_cetric.h:105-170, cetric.h:160-316 and
665-1002 (paths relative to the reference root).
"""
import os
import threading
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import katric_amd as K
from katric_amd import batch as KB
from katric_amd import serve as KS
from katric_amd import shard as S
from oracle import oracle as O

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(900)]

THREADS = max(1, min(16, os.cpu_count() or 1))


def log(msg):
    print(f"[configs] {msg}", flush=True)


def reversed_rows(g):
    """The same graph with every adjacency list in descending order."""
    deg = np.diff(g.offsets.astype(np.int64))
    rows = np.repeat(np.arange(g.n, dtype=np.int64), deg)
    pos = g.offsets[rows].astype(np.int64) + g.offsets[rows + 1].astype(np.int64) - 1 - np.arange(g.nnz)
    heads = np.empty_like(g.heads)
    heads[pos] = g.heads
    del rows, pos
    return K.Graph(g.n, g.offsets, heads)


def oracle_graphs(graphs, per_vertex=True):
    """(total, per-vertex) of each graph, one oracle thread per graph."""
    with ThreadPoolExecutor(THREADS) as ex:
        return list(ex.map(lambda g: O.run_shmem(g, threads=1, per_vertex=per_vertex)[:2], graphs))


# ---------------------------------------------------------------------------
# 1: G(n,m) 2^16 / 2^20
# ---------------------------------------------------------------------------
def test_config1_gnm(ctx):
    g = K.gen_gnm(1 << 16, 1 << 20, seed=1)
    log(f"config 1: n={g.n} m={g.m}")
    total, pv = ctx.count(g)
    st = ctx.stats()
    ot, opv, ost = O.run_shmem(g, threads=THREADS)
    assert total == ot
    np.testing.assert_array_equal(pv, opv)
    assert st["wedge_checks"] == ost["wedge_checks"]
    assert st["intersection_size_local"] == ost["intersection_size_local"]
    # single-rank katric: CeTriC at P = 1 (one thread), and the device shard path
    ol, og, opv1, _ = O.rank_sim(g, 1, 0, O.CETRIC, threads=1)
    assert (ol, og) == (total, 0)
    np.testing.assert_array_equal(opv1, opv)
    loc, t3, spv = S.count_shard_graph(ctx, g, 1, 0, S.CETRIC)
    assert (loc, t3) == (ol, og)
    np.testing.assert_array_equal(spv, opv)


# ---------------------------------------------------------------------------
# 2: R-MAT 22
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def rmat22():
    t = time.time()
    g = K.gen_rmat(22, 16, seed=1)
    log(f"config 2: RMAT-22 n={g.n} m={g.m} gen {time.time() - t:.1f}s")
    return g


@pytest.fixture(scope="module")
def rmat22_oracle(rmat22):
    t = time.time()
    r = O.run_shmem(rmat22, threads=THREADS)
    log(f"config 2: oracle {time.time() - t:.1f}s total={r[0]}")
    return r


def test_config2_rmat22(ctx, rmat22, rmat22_oracle):
    g = rmat22
    dg = ctx.upload(g)
    d_pv = ctx.alloc(g.n * 8)
    total = ctx.count_device(dg, per_vertex_out=d_pv)
    st = ctx.stats()
    pv = d_pv.download(np.uint64, g.n)
    dg.free()
    d_pv.free()
    ot, opv, ost = rmat22_oracle
    assert total == ot
    np.testing.assert_array_equal(pv, opv)
    for k in ("wedge_checks", "intersection_size_local"):
        assert st[k] == ost[k], k
    assert st["oriented_edges"] == g.m


def test_config2_rmat22_unsorted(ctx, rmat22, rmat22_oracle):
    # every list in descending order: the oriented lists go through the
    # device radix sort
    g = reversed_rows(rmat22)
    total, pv = ctx.count(g)
    tm = ctx.timing()
    ot, opv, _ = rmat22_oracle
    assert total == ot
    np.testing.assert_array_equal(pv, opv)
    assert tm["prep_ms"] > 0 and tm["total_ms"] > 0
    log(f"config 2 unsorted: prep {tm['prep_ms']:.2f} ms, total {tm['total_ms']:.2f} ms")


# ---------------------------------------------------------------------------
# 3: R-MAT 26 in 16 ranges, rank 0
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def rmat26():
    t = time.time()
    g = K.gen_rmat(26, 16, seed=1)
    log(f"config 3: RMAT-26 n={g.n} m={g.m} gen {time.time() - t:.1f}s")
    return g


@pytest.mark.parametrize("algo", [S.CETRIC, S.DITRIC], ids=["cetric", "ditric"])
def test_config3_rmat26_rank0(ctx, rmat26, algo):
    g = rmat26
    P, r = 16, 0
    t = time.time()
    s = S.make_shard(g, P, r, algo)
    loc, t3, pv = S.count_shard(ctx, s)
    st = ctx.stats()
    log(f"config 3 algo {algo}: local {loc} type3 {t3} (shard + count {time.time() - t:.1f}s)")
    t = time.time()
    ol, og, opv, ost = O.rank_sim(g, P, r, algo, threads=THREADS)
    log(f"config 3 algo {algo}: oracle {time.time() - t:.1f}s")
    assert (loc, t3) == (ol, og)
    ids = s.vertex_ids()
    np.testing.assert_array_equal(pv, opv[ids])
    assert int(opv.sum(dtype=np.uint64)) == int(pv.sum(dtype=np.uint64))  # nothing outside the shard
    for k in ("wedge_checks", "intersection_size_local", "intersection_size_global"):
        assert st[k] == ost[k], k
    del s
    # the same rank with its inputs built on the device from the whole graph
    dg = ctx.upload(g)
    d_pv = ctx.alloc(g.n * 8)
    loc2, t32, _ = S.count_shard_graph(ctx, dg, P, r, algo, per_vertex_out=d_pv)
    assert (loc2, t32) == (ol, og)
    np.testing.assert_array_equal(d_pv.download(np.uint64, g.n), opv)
    dg.free()
    d_pv.free()


# ---------------------------------------------------------------------------
# 4: 8192 graphs in one batch
# ---------------------------------------------------------------------------
def test_config4_batch8192(ctx):
    t = time.time()
    b = KB.gen_batch(8192, seed=1)
    log(f"config 4: {b.num_graphs} graphs, {b.n_total} vertices, {b.nnz // 2} edges, gen {time.time() - t:.1f}s")
    db = KB.upload_batch(ctx, b)
    d_pv = ctx.alloc(b.n_total * 8)
    tot = KB.count_batch(ctx, db, per_vertex_out=d_pv)       # captures the pipeline
    tot2 = KB.count_batch(ctx, db, per_vertex_out=d_pv)      # replays it
    np.testing.assert_array_equal(tot, tot2)
    pv = d_pv.download(np.uint64, b.n_total)
    db.offsets.free()
    db.heads.free()
    d_pv.free()
    t = time.time()
    res = oracle_graphs([b.graph(i) for i in range(b.num_graphs)])
    log(f"config 4: oracle {time.time() - t:.1f}s")
    np.testing.assert_array_equal(tot, np.array([r[0] for r in res], np.uint64))
    np.testing.assert_array_equal(pv, np.concatenate([r[1] for r in res]))


# ---------------------------------------------------------------------------
# 5: served streams
# ---------------------------------------------------------------------------
def served_stream(ctx, pool, which, rate, seed):
    """Submits pool[which[i]] at Poisson arrival times (per-vertex outputs
    requested); returns the results and per-vertex arrays in order."""
    srv = KS.Server(ctx)
    R = len(which)
    arrivals = np.cumsum(np.random.default_rng(seed).exponential(1.0 / rate, size=R))
    pvs = [np.zeros(pool[w].n, np.uint64) for w in which]
    tickets = [0] * R

    def submit():
        base = time.perf_counter()
        for i in range(R):
            d = base + arrivals[i] - time.perf_counter()
            if d > 0:
                time.sleep(d)
            tickets[i] = srv.submit(pool[which[i]], per_vertex=pvs[i])

    th = threading.Thread(target=submit)
    th.start()
    res = []
    for i in range(R):
        while tickets[i] == 0:
            time.sleep(0.0005)
        res.append(srv.wait(tickets[i]))
    th.join()
    stats = srv.stats()
    srv.stop()
    return res, pvs, stats


def check_stream(res, pvs, which, want):
    bad = [i for i, r in enumerate(res) if r.status != 0 or r.total != want[which[i]][0]]
    assert not bad, f"{len(bad)} served results differ from the oracle (first: request {bad[0]})"
    for i, p in enumerate(pvs):
        np.testing.assert_array_equal(p, want[which[i]][1])


@pytest.mark.parametrize("mix", ["config5", "config4-small"])
def test_config5_served_stream(ctx, mix):
    if mix == "config5":
        P, R, rate = 192, 2000, 1500.0
        pool_b = KB.gen_batch(P, seed=5, n_min=1 << 10, n_max=1 << 20, deg_min=4.0, deg_max=32.0)
    else:
        P, R, rate = 512, 4000, 20000.0
        pool_b = KB.gen_batch(P, seed=6)
    pool = [pool_b.graph(i) for i in range(P)]
    del pool_b
    t = time.time()
    want = oracle_graphs(pool)
    log(f"{mix}: pool {P} graphs, {sum(g.m for g in pool)} edges, oracle {time.time() - t:.1f}s")
    which = np.random.default_rng(7).integers(0, P, size=R)
    t = time.time()
    res, pvs, stats = served_stream(ctx, pool, which, rate, seed=8)
    lat = np.array([r.latency_ms for r in res])
    log(f"{mix}: {R} requests in {time.time() - t:.1f}s, {stats['batches']} batches, "
        f"p50 {np.percentile(lat, 50):.1f} ms p99 {np.percentile(lat, 99):.1f} ms")
    check_stream(res, pvs, which, want)
    # alone: the same graphs one call each equal the oracle too
    for i in range(0, P, max(1, P // 16)):
        t1, pv1 = ctx.count(pool[i])
        assert t1 == want[i][0]
        np.testing.assert_array_equal(pv1, want[i][1])
