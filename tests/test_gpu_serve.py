"""Serving loop: every result equals the one-graph count, alone and in the mix;
admission by budget, isolation of malformed requests, cancellation."""
import time

import numpy as np
import pytest

import katric_amd as K
from katric_amd import serve as KS

pytestmark = pytest.mark.gpu


def mix(count, seed):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        if rng.random() < 0.5:
            out.append(K.gen_rmat(int(rng.integers(6, 14)), int(rng.integers(4, 17)), seed=int(rng.integers(1 << 30))))
        else:
            n = int(rng.integers(50, 20000))
            out.append(K.gen_gnm(n, int(n * rng.uniform(2, 16)), seed=int(rng.integers(1 << 30))))
    return out


def test_serve_results_equal_alone(ctx):
    graphs = mix(40, 1)
    alone = [ctx.count(g) for g in graphs]
    with K.Context(0) as sctx:
        srv = KS.Server(sctx)
        pvs = [np.zeros(g.n, np.uint64) for g in graphs]
        tickets = [srv.submit(g, pv) for g, pv in zip(graphs, pvs)]
        results = [srv.wait(t) for t in tickets]
        st = srv.stats()
        srv.stop()
    for (t, pv), r, p in zip(alone, results, pvs):
        assert r.status == 0 and r.total == t
        np.testing.assert_array_equal(p, pv)
    assert st["graphs"] == 40
    assert max(r.batch_graphs for r in results) > 1, "requests were batched"


def test_serve_arrivals_during_batches(ctx):
    graphs = mix(60, 2)
    alone = [ctx.count(g, per_vertex=False)[0] for g in graphs]
    with K.Context(0) as sctx:
        srv = KS.Server(sctx)
        tickets = []
        for i, g in enumerate(graphs):
            tickets.append(srv.submit(g))
            if i % 7 == 0:
                time.sleep(0.002)
        res = [srv.wait(t) for t in tickets]
        srv.stop()
    assert [r.total for r in res] == alone
    assert len({r.batch_id for r in res}) > 1


def test_serve_isolates_malformed(ctx):
    graphs = mix(10, 3)
    bad = K.Graph(3, np.array([0, 1, 1, 2], np.uint64), np.array([1, 1], np.uint32))  # asymmetric
    out_of_range = K.Graph(2, np.array([0, 1, 2], np.uint64), np.array([1, 7], np.uint32))
    alone = [ctx.count(g, per_vertex=False)[0] for g in graphs]
    with K.Context(0) as sctx:
        srv = KS.Server(sctx)
        tk = [srv.submit(g) for g in graphs[:5]] + [srv.submit(bad), srv.submit(out_of_range)] + \
             [srv.submit(g) for g in graphs[5:]]
        res = [srv.wait(t) for t in tk]
        srv.stop()
    good = res[:5] + res[7:]
    assert [r.total for r in good] == alone and all(r.status == 0 for r in good)
    assert res[5].status == K.KT_EINVAL and res[6].status == K.KT_EINVAL


def test_serve_budget_admission(ctx):
    graphs = mix(20, 4)
    big = K.gen_rmat(16, 16, seed=1)
    alone = [ctx.count(g, per_vertex=False)[0] for g in graphs]
    with K.Context(0) as sctx:
        srv = KS.Server(sctx, hbm_budget_bytes=24 << 20)  # 12 MiB per batch slot; RMAT-16 needs ~17 MiB
        tk = [srv.submit(g) for g in graphs]
        tbig = srv.submit(big)
        res = [srv.wait(t) for t in tk]
        rbig = srv.wait(tbig)
        srv.stop()
    assert [r.total for r in res] == alone
    assert rbig.status == K.KT_ENOMEM
    assert max(r.batch_graphs for r in res) < 20  # the budget split the stream into several batches


def test_serve_cancel_and_unsorted(ctx):
    g = K.gen_rmat(11, 16, seed=9)
    rng = np.random.default_rng(0)
    heads = g.heads.copy()
    for v in range(g.n):
        a, b = int(g.offsets[v]), int(g.offsets[v + 1])
        heads[a:b] = rng.permutation(heads[a:b])
    ug = K.Graph(g.n, g.offsets, heads)
    want = ctx.count(g, per_vertex=False)[0]
    with K.Context(0) as sctx:
        srv = KS.Server(sctx)
        tk = [srv.submit(ug) for _ in range(30)]
        srv.cancel(tk[-1])
        res = [srv.wait(t) for t in tk]
        srv.stop()
    assert all(r.total == want and r.status == 0 for r in res[:-1])
    assert res[-1].status in (0, K.KT_ECANCELED)


def test_serve_page_locked_and_pageable_mix(ctx):
    # page-locked requests are read by the device across PCIe (one gather
    # launch per batch); pageable ones are copied in first -- results agree
    graphs = mix(30, 5)
    alone = [ctx.count(g) for g in graphs]
    with K.Context(0) as sctx:
        srv = KS.Server(sctx)
        keep, tickets, pvs = [], [], []
        for i, g in enumerate(graphs):
            if i % 3:
                view, own = KS.pinned_graph(g)
                keep += own
                g = view
            pv = np.zeros(g.n, np.uint64)
            pvs.append(pv)
            tickets.append(srv.submit(g, pv))
        res = [srv.wait(t) for t in tickets]
        srv.stop()
    for (t, pv), r, p in zip(alone, res, pvs):
        assert r.status == 0 and r.total == t
        np.testing.assert_array_equal(p, pv)


def test_serve_page_locked_malformed_isolated(ctx):
    graphs = mix(6, 6)
    out_of_range = K.Graph(2, np.array([0, 1, 2], np.uint64), np.array([1, 7], np.uint32))
    alone = [ctx.count(g, per_vertex=False)[0] for g in graphs]
    keep = []
    with K.Context(0) as sctx:
        srv = KS.Server(sctx)
        tk = []
        for g in graphs[:3] + [out_of_range] + graphs[3:]:
            view, own = KS.pinned_graph(g)
            keep += own
            tk.append(srv.submit(view))
        res = [srv.wait(t) for t in tk]
        srv.stop()
    good = res[:3] + res[4:]
    assert [r.total for r in good] == alone
    assert res[3].status == K.KT_EINVAL


def test_serve_unsorted_under_tight_budget(ctx):
    # unsorted requests: a batch that finds unsorted lists is re-launched with
    # them sorted in place (no scratch beyond the first launch's), under
    # budgets from a few MiB up: every result must equal the oracle
    from oracle import oracle as O

    rng = np.random.default_rng(4)
    graphs = []
    for s in range(8):
        g = K.gen_gnm(2000, 16000, seed=40 + s)
        heads = g.heads.copy()
        for v in range(g.n):
            a, b = int(g.offsets[v]), int(g.offsets[v + 1])
            heads[a:b] = rng.permutation(heads[a:b])
        graphs.append(K.Graph(g.n, g.offsets, heads))
    want = [O.run_shmem(g)[:2] for g in graphs]
    isolated = 0
    for mib in (3, 4, 5, 6, 8, 12, 64):
        with K.Context(0) as sctx:
            srv = KS.Server(sctx, hbm_budget_bytes=mib << 20)
            pvs = [np.zeros(g.n, np.uint64) for g in graphs]
            tickets = [srv.submit(g, pv) for g, pv in zip(graphs, pvs)]
            res = [srv.wait(t) for t in tickets]
            isolated += srv.stats()["isolation_reruns"]
            srv.stop()
        for (t, opv), r, pv in zip(want, res, pvs):
            assert r.status == 0 and r.total == t, (mib, r.status)
            np.testing.assert_array_equal(pv, opv)


def _clique_minus_matching_shuffled(n, seed):
    """K_n minus the matching (0,2), (1,3), (4,5), ...: all degrees n - 2,
    vertex 0's oriented list n - 2 ids; rows shuffled.  Triangles:
    C(n,3) - (n/2)(n-2)."""
    iu, ju = np.triu_indices(n, 1)
    keep = ~((iu == 0) & (ju == 2)) & ~((iu == 1) & (ju == 3)) & ~((iu >= 4) & (iu % 2 == 0) & (ju == iu + 1))
    g = K.Graph.from_edges(n, np.stack([iu[keep], ju[keep]], 1))
    rng = np.random.default_rng(seed)
    heads = g.heads.copy()
    for v in range(g.n):
        a, b = int(g.offsets[v]), int(g.offsets[v + 1])
        heads[a:b] = rng.permutation(heads[a:b])
    return K.Graph(g.n, g.offsets, heads)


def test_serve_unsorted_long_lists(ctx):
    # a request whose longest oriented list (4098 ids) is beyond the in-place
    # sort: its batch is re-launched with the radix sort; small requests
    # queued around it are answered as usual
    from oracle import oracle as O

    n = 4100
    big = _clique_minus_matching_shuffled(n, 7)
    small = [K.gen_gnm(300, 2000, seed=60 + s) for s in range(3)]
    with K.Context(0) as sctx:
        srv = KS.Server(sctx)
        tickets = [srv.submit(g) for g in [small[0], big] + small[1:]]
        res = [srv.wait(t) for t in tickets]
        srv.stop()
    assert res[1].status == 0 and res[1].total == n * (n - 1) * (n - 2) // 6 - (n // 2) * (n - 2)
    for r, g in zip([res[0]] + res[2:], small):
        assert r.status == 0 and r.total == O.run_shmem(g)[0]


def test_serve_large_requests_run_alone(ctx):
    # requests above large_request_edges run in batches of their own (one at
    # a time); the small ones around them may go ahead; every result equals
    # the one-graph count
    graphs = mix(50, 9)
    alone = [ctx.count(g) for g in graphs]
    big = 30000
    with K.Context(0) as sctx:
        srv = KS.Server(sctx, large_request_edges=big)
        pvs = [np.zeros(g.n, np.uint64) for g in graphs]
        tickets = [srv.submit(g, pv) for g, pv in zip(graphs, pvs)]
        res = [srv.wait(t) for t in tickets]
        srv.stop()
    n_large = 0
    for g, (t, pv), r, p in zip(graphs, alone, res, pvs):
        assert r.status == 0 and r.total == t
        np.testing.assert_array_equal(p, pv)
        if g.m > big:
            n_large += 1
            assert r.batch_graphs == 1
    assert 0 < n_large < len(graphs)
    assert max(r.batch_graphs for r in res) > 1
