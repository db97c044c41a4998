"""Parity of the batched, hipGraph-replayed path (kt_batch_count) with the
oracle, graph by graph, and with the one-graph path."""
import numpy as np
import pytest

import katric_amd as K
from katric_amd import batch as KB
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def oracle_totals(b):
    tot, pvs = [], []
    for g in range(b.num_graphs):
        t, pv, _ = O.run_shmem(b.graph(g))
        tot.append(t)
        pvs.append(pv)
    return np.array(tot, np.uint64), np.concatenate(pvs) if pvs else np.zeros(0, np.uint64)


def test_batch_small_mixed(ctx):
    b = KB.gen_batch(48, seed=3, n_min=64, n_max=3000)
    tot, pv = KB.count_batch(ctx, b, per_vertex=True)
    otot, opv = oracle_totals(b)
    np.testing.assert_array_equal(tot, otot)
    np.testing.assert_array_equal(pv, opv)


def test_batch_graph_replays_match_eager(ctx):
    b = KB.gen_batch(64, seed=5, n_min=64, n_max=4096)
    db = KB.upload_batch(ctx, b)
    otot, _ = oracle_totals(b)
    r0 = KB.replays(ctx)
    for i in range(4):
        tot = KB.count_batch(ctx, db)
        np.testing.assert_array_equal(tot, otot)
        # eager work between replays
        g = b.graph(i)
        assert ctx.count(g)[0] == otot[i]
    assert KB.replays(ctx) - r0 >= 4
    db.offsets.free()
    db.heads.free()


def test_batch_equals_single_graph_calls(ctx):
    graphs = [K.gen_rmat(s, 8, seed=s) for s in (6, 9, 11)] + [K.read_metis(__import__("tests.conftest", fromlist=["golden"]).golden("hiv.metis"))]
    b = KB.from_graphs(graphs)
    tot = KB.count_batch(ctx, b)
    assert tot.tolist() == [ctx.count(g)[0] for g in graphs]
    assert tot[-1] == 45


def test_batch_unsorted_and_empty_graphs(ctx):
    rng = np.random.default_rng(1)
    graphs = [K.gen_gnm(500, 4000, seed=2), K.Graph(7, np.zeros(8, np.uint64), np.zeros(0, np.uint32)),
              K.gen_rmat(10, 16, seed=4)]
    b = KB.from_graphs(graphs)
    heads = b.heads.copy()
    for v in range(b.n_total):
        a, e = int(b.offsets[v]), int(b.offsets[v + 1])
        heads[a:e] = rng.permutation(heads[a:e])
    b2 = KB.Batch(b.vertex_base, b.offsets, heads)
    t1 = KB.count_batch(ctx, b)
    t2 = KB.count_batch(ctx, b2)
    np.testing.assert_array_equal(t1, t2)
    assert t1[1] == 0


def test_batch_unsorted_long_list(ctx):
    # one graph's longest oriented list (4098 ids) is beyond the in-place
    # sort: the batch is redone with the radix sort; K_n minus a perfect
    # matching has C(n,3) - (n/2)(n-2) triangles
    n = 4100
    iu, ju = np.triu_indices(n, 1)
    keep = ~((iu == 0) & (ju == 2)) & ~((iu == 1) & (ju == 3)) & ~((iu >= 4) & (iu % 2 == 0) & (ju == iu + 1))
    big = K.Graph.from_edges(n, np.stack([iu[keep], ju[keep]], 1))
    graphs = [K.gen_gnm(300, 2000, seed=5), big, K.gen_rmat(9, 16, seed=6)]
    b = KB.from_graphs(graphs)
    rng = np.random.default_rng(3)
    heads = b.heads.copy()
    for v in range(b.n_total):
        a, e = int(b.offsets[v]), int(b.offsets[v + 1])
        heads[a:e] = rng.permutation(heads[a:e])
    tot = KB.count_batch(ctx, KB.Batch(b.vertex_base, b.offsets, heads))
    assert tot[1] == n * (n - 1) * (n - 2) // 6 - (n // 2) * (n - 2)
    assert tot[0] == ctx.count(graphs[0])[0] and tot[2] == ctx.count(graphs[2])[0]


def test_batch_totals_across_chunk_boundaries(ctx):
    # per-graph totals are summed in 512-vertex chunks: graphs with no vertices,
    # graphs that end exactly on a chunk boundary and graphs spanning many chunks
    empty = K.Graph(0, np.zeros(1, np.uint64), np.zeros(0, np.uint32))
    graphs = [empty, K.gen_rmat(15, 16, seed=7), empty, K.gen_gnm(512, 6000, seed=8), K.gen_gnm(511, 6000, seed=9),
              empty, K.gen_rmat(9, 16, seed=10), K.gen_gnm(1025, 9000, seed=11), empty]
    b = KB.from_graphs(graphs)
    tot, pv = KB.count_batch(ctx, b, per_vertex=True)
    want = [ctx.count(g)[0] if g.n else 0 for g in graphs]
    assert tot.tolist() == want
    assert int(pv.sum()) == 3 * sum(want)


def test_batch_invalid(ctx):
    g = K.Graph(3, np.array([0, 1, 1, 2], np.uint64), np.array([1, 1], np.uint32))  # asymmetric
    b = KB.from_graphs([K.gen_gnm(100, 500, seed=1), g])
    with pytest.raises(K.KatricError):
        KB.count_batch(ctx, b)
    # context still usable
    assert KB.count_batch(ctx, KB.from_graphs([K.gen_gnm(100, 500, seed=1)]))[0] == ctx.count(K.gen_gnm(100, 500, seed=1))[0]
