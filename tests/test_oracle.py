"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the checker.

Every function of the oracle is checked here against something it does not
compute itself: the reference's own test vectors and known answers, closed
forms, brute force, networkx golden vectors, or an independent Python
enumeration.
"""
import itertools
import json
import math

import numpy as np
import pytest

import katric_amd as K
from oracle import oracle as O
from tests.conftest import KNOWN_TOTALS, golden

# tests/intersection_test.cpp:41-52 of the reference (hand vectors)
HAND_VECTORS = [
    ([0, 1, 3, 6, 9, 478], [1, 6, 7, 8, 478], [1, 6, 478]),
    ([], [1, 6, 7, 8, 478], []),
    ([], [], []),
    ([1], [], []),
    ([1], [1], [1]),
    ([1], [42], []),
    ([13, 46, 54, 58, 64], [6, 21, 43, 47, 75], []),
    ([1, 2, 5, 9, 10], [1, 2, 4, 5, 6, 7, 8], [1, 2, 5]),
    ([2, 3, 4, 7, 10], [3, 5, 6], [3]),
    ([1, 2, 3, 4, 5], [6, 7, 8, 9, 10, 11, 12], []),
]


@pytest.mark.parametrize("case", range(len(HAND_VECTORS)))
@pytest.mark.parametrize("swap", [False, True])
def test_intersection_hand_vectors(case, swap):
    a, b, exp = HAND_VECTORS[case]
    if swap:
        a, b = b, a
    assert O.intersection(a, b, method="merge_only") == exp
    assert O.intersection(a, b, method="binary_search_only") == exp
    # binary_intersection emits in arbitrary order (compared unordered there)
    assert sorted(O.intersection(a, b, method="binary_only")) == exp
    assert sorted(O.intersection(a, b, method="binary_only", cutoff=1)) == exp
    for m in (O.MERGE, O.BINARY_SEARCH, O.HYBRID):
        assert O.intersection(a, b, method=m) == exp


def test_intersection_random():
    # tests/intersection_test.cpp:101-173: 100 random size triples <= 1000,
    # values in [1, max(10, 100)], against std::set_intersection
    rng = np.random.default_rng(12345)
    for _ in range(100):
        s1, s2 = rng.integers(0, 1001, size=2)
        v1 = sorted(set(rng.integers(1, 101, size=s1).tolist()))
        v2 = sorted(set(rng.integers(1, 101, size=s2).tolist()))
        for v2_empty, swap in itertools.product([False, True], [False, True]):
            a, b = v1, ([] if v2_empty else v2)
            if swap:
                a, b = b, a
            exp = sorted(set(a) & set(b))
            assert O.intersection(a, b, method="merge_only") == exp
            assert O.intersection(a, b, method="binary_search_only") == exp
            assert sorted(O.intersection(a, b, method="binary_only", cutoff=8)) == exp
            assert O.intersection(a, b, method=O.HYBRID) == exp


def test_log2_closed_form():
    assert O.log2(1) == 0
    for k in range(1, 63):
        assert O.log2(1 << k) == k
        assert O.log2((1 << k) - 1) == k - 1
        assert O.log2((1 << k) + 1) == k


@pytest.mark.parametrize("name", sorted(KNOWN_TOTALS))
def test_known_answers(name):
    # examples/examples.yaml of the reference
    g = K.read_metis(golden(name))
    for threads in (1, 4):
        t, pv, st = O.run_shmem(g, threads=threads)
        assert t == KNOWN_TOTALS[name]
        assert int(pv.sum()) == 3 * t
    bt, bpv = O.brute_force(g)
    assert bt == KNOWN_TOTALS[name]
    np.testing.assert_array_equal(O.run_shmem(g)[1], bpv)


def test_binary_fixture_matches_metis():
    # examples/metis-sample.first_out/.head hold the same graph as the .metis file
    a = K.read_metis(golden("metis-sample.metis"))
    b = K.read_graphio_binary(golden("metis-sample.first_out"), golden("metis-sample.head"))
    np.testing.assert_array_equal(a.offsets, b.offsets)
    np.testing.assert_array_equal(a.heads, b.heads)
    assert a.offsets.tolist() == [0, 3, 6, 10, 14, 17, 20, 22]


def _orientation_invariants(g, ooff, oheads):
    deg = np.diff(g.offsets).astype(np.int64)
    m = g.nnz // 2
    assert int(ooff[-1]) == m
    seen = set()
    for v in range(g.n):
        out = oheads[ooff[v]:ooff[v + 1]].tolist()
        assert out == sorted(set(out)), "out-lists ascending, no duplicates"
        nb = set(g.heads[g.offsets[v]:g.offsets[v + 1]].tolist())
        for u in out:
            assert u in nb
            assert (deg[v], v) < (deg[u], u)
            seen.add((min(u, v), max(u, v)))
        assert len(out) <= math.sqrt(2 * m) + 1
        # out-neighbours plus in-neighbours give N(v) (orientation_test.cpp:81-133)
        ins = {u for u in nb if (deg[u], u) < (deg[v], v)}
        assert set(out) | ins == nb and not (set(out) & ins)
    assert len(seen) == m, "every undirected edge oriented exactly once"


@pytest.mark.parametrize("name", ["metis-sample.metis", "triangle.metis", "hiv.metis"])
def test_orientation_invariants_examples(name):
    g = K.read_metis(golden(name))
    _orientation_invariants(g, *O.orient(g))


def test_orientation_invariants_random():
    g = K.gen_rmat(9, 8, seed=4)
    _orientation_invariants(g, *O.orient(g))


def _golden_cases():
    with open(golden("pervertex.json")) as f:
        return json.load(f)["cases"]


@pytest.mark.parametrize("case", _golden_cases(), ids=lambda c: c["name"])
def test_per_vertex_networkx_golden(case):
    g = K.Graph.from_edges(case["n"], case["edges"])
    for method in (O.MERGE, O.BINARY_SEARCH, O.BINARY, O.HYBRID):
        t, pv, _ = O.run_shmem(g, method=method)
        assert t == case["total"]
        assert pv.tolist() == case["per_vertex"]


def test_closed_forms():
    for n in (3, 4, 7, 20):
        g = K.Graph.from_edges(n, [(i, j) for i in range(n) for j in range(i + 1, n)])
        t, pv, _ = O.run_shmem(g)
        assert t == math.comb(n, 3)
        assert (pv == math.comb(n - 1, 2)).all()
        assert O.brute_force(g)[0] == math.comb(n, 3)
    star = K.Graph.from_edges(50, [(0, i) for i in range(1, 50)])
    assert O.run_shmem(star)[0] == 0 and O.brute_force(star)[0] == 0
    k = 30  # wheel: hub 0 + rim cycle 1..k: k triangles, hub in all, rim vertex in 2
    wheel = K.Graph.from_edges(k + 1, [(0, i) for i in range(1, k + 1)] + [(i, i % k + 1) for i in range(1, k + 1)])
    t, pv, _ = O.run_shmem(wheel)
    assert t == k and pv[0] == k and (pv[1:] == 2).all()
    empty = K.Graph(4, np.zeros(5, np.uint64), np.zeros(0, np.uint32))
    assert O.run_shmem(empty)[0] == 0


def test_relabelling_invariance():
    g = K.gen_gnm(300, 3000, seed=9)
    t, pv, _ = O.run_shmem(g)
    perm = np.random.default_rng(3).permutation(g.n)
    tails = np.repeat(np.arange(g.n), np.diff(g.offsets).astype(np.int64))
    e = np.stack([perm[tails], perm[g.heads]], 1)
    h = K.Graph.from_edges(g.n, e[e[:, 0] < e[:, 1]])
    t2, pv2, _ = O.run_shmem(h)
    assert t2 == t
    np.testing.assert_array_equal(pv2[perm], pv)


def test_parallel_equals_sequential():
    g = K.gen_rmat(12, 16, seed=2)
    a = O.run_shmem(g, threads=1)
    b = O.run_shmem(g, threads=8)
    assert a[0] == b[0]
    np.testing.assert_array_equal(a[1], b[1])
    assert a[2]["intersection_size_local"] == b[2]["intersection_size_local"]
    # W = sum over oriented edges of |N+(v)| + |N+(u)|, recomputed in numpy
    ooff, oh = O.orient(g)
    dp = np.diff(ooff).astype(np.int64)
    tails = np.repeat(np.arange(g.n), dp)
    assert a[2]["intersection_size_local"] == int((dp[tails] + dp[oh]).sum())
    assert a[2]["wedge_checks"] == int(dp[oh].sum())


def _python_rank_attribution(g, bounds, P):
    """Independent enumeration: every triangle by brute force over edges,
    ordered by (degree, id), attributed by owner pattern."""
    deg = np.diff(g.offsets).astype(np.int64)
    adj = [set(g.heads[g.offsets[v]:g.offsets[v + 1]].tolist()) for v in range(g.n)]

    def owner(x):
        return int(np.searchsorted(bounds, x, side="right") - 1)

    cet_local, cet_t3 = np.zeros(P, np.int64), np.zeros(P, np.int64)
    dit_local, dit_glob = np.zeros(P, np.int64), np.zeros(P, np.int64)
    pv_cet = np.zeros((P, g.n), np.int64)
    for a in range(g.n):
        for b in adj[a]:
            if b <= a:
                continue
            for c in adj[a] & adj[b]:
                if c <= b:
                    continue
                lo, mid, hi = sorted((a, b, c), key=lambda x: (deg[x], x))
                ro = [owner(lo), owner(mid), owner(hi)]
                if len(set(ro)) < 3:
                    r = max(set(ro), key=ro.count)
                    cet_local[r] += 1
                else:
                    r = ro[1]
                    cet_t3[r] += 1
                pv_cet[r, [a, b, c]] += 1
                if ro[0] == ro[1]:
                    dit_local[ro[1]] += 1
                else:
                    dit_glob[ro[1]] += 1
    return cet_local, cet_t3, dit_local, dit_glob, pv_cet


@pytest.mark.parametrize("P", [1, 2, 3, 4, 16])
def test_rank_attribution(P):
    g = K.gen_rmat(9, 16, seed=P)
    bounds = O.even_bounds(g.n, P)
    cl, c3, dl, dg, pvc = _python_rank_attribution(g, bounds, P)
    total = O.run_shmem(g)[0]
    assert cl.sum() + c3.sum() == total
    cf_l, cf_3 = O.rank_closed_form(g, P, O.CETRIC)
    np.testing.assert_array_equal(cf_l, cl)
    np.testing.assert_array_equal(cf_3, c3)
    df_l, df_g = O.rank_closed_form(g, P, O.DITRIC)
    np.testing.assert_array_equal(df_l, dl)
    np.testing.assert_array_equal(df_g, dg)
    pv_sum = np.zeros(g.n, np.int64)
    for r in range(P):
        loc, t3, pv, st = O.rank_sim(g, P, r, O.CETRIC)
        assert (loc, t3) == (cl[r], c3[r])
        np.testing.assert_array_equal(pv, pvc[r])
        pv_sum += pv.astype(np.int64)
        dloc, dglo, _, _ = O.rank_sim(g, P, r, O.DITRIC)
        assert (dloc, dglo) == (dl[r], dg[r])
        if P > 1:
            loc2, t32, pv2, _ = O.rank_sim(g, P, r, O.CETRIC, threads=4)
            assert (loc2, t32) == (loc, t3)
            np.testing.assert_array_equal(pv2, pv)
    np.testing.assert_array_equal(pv_sum, O.run_shmem(g)[1].astype(np.int64))


def test_rank_attribution_uneven_ranges():
    g = K.gen_gnm(400, 4000, seed=5)
    bounds = np.array([0, 10, 11, 200, 390, 400], np.uint64)
    P = 5
    cl, c3, dl, dg, _ = _python_rank_attribution(g, bounds, P)
    for r in range(P):
        assert O.rank_sim(g, P, r, O.CETRIC, bounds=bounds)[:2] == (cl[r], c3[r])
        assert O.rank_sim(g, P, r, O.DITRIC, bounds=bounds)[:2] == (dl[r], dg[r])


def rank_work_counters(g, bounds, r, algorithm):
    """(wedge_checks, intersection_size_local, intersection_size_global) of
    rank r as sums over oriented edges (numpy), independent of the oracle's
    message loops.  Local phase: per tail v and out-neighbour u, |list(v)| +
    |list(u)| intersected ids and |list(u)| wedge checks.  Message phase: per
    edge a -> u into the rank from a remote a, |message(a)| + |list(u)| and
    |list(u)|, where CeTriC's lists and messages are the cut lists (remote
    out-neighbours) and receivers without one are compacted away."""
    ooff, oh = O.orient(g)
    dp = np.diff(ooff).astype(np.int64)
    tails = np.repeat(np.arange(g.n), dp)
    owner = np.searchsorted(bounds.astype(np.int64), np.arange(g.n), side="right") - 1
    mine = owner == r
    cut = np.bincount(tails, weights=owner[oh] != owner[tails], minlength=g.n).astype(np.int64)
    into = ~mine[tails] & mine[oh]  # a -> u, a remote, u owned
    if algorithm == O.DITRIC:
        loc = mine[tails] & mine[oh]
        w = dp[oh][loc].sum() + dp[oh][into].sum()
        return int(w), int((dp[tails] + dp[oh])[loc].sum()), int((dp[tails] + dp[oh])[into].sum())
    # CeTriC: a remote tail's list holds its owned out-neighbours; an owned
    # tail sees a ghost head's owned out-neighbours only
    own_out = np.bincount(tails, weights=mine[oh], minlength=g.n).astype(np.int64)
    dlist = np.where(mine, dp, own_out)
    loc = mine[tails] | (into & (own_out[tails] > 0))
    wl = dlist[oh][loc]
    isz_l = (dlist[tails][loc] + wl).sum()
    recv = into & (cut[oh] > 0)
    return (int(wl.sum() + cut[oh][recv].sum()), int(isz_l),
            int((cut[tails] + cut[oh])[recv].sum()))


@pytest.mark.parametrize("algorithm", [O.CETRIC, O.DITRIC])
@pytest.mark.parametrize("P", [1, 3, 8])
def test_rank_work_counters(P, algorithm):
    g = K.gen_rmat(10, 16, seed=20 + P)
    bounds = O.even_bounds(g.n, P)
    for r in range(P):
        st = O.rank_sim(g, P, r, algorithm, per_vertex=False)[3]
        got = (st["wedge_checks"], st["intersection_size_local"], st["intersection_size_global"])
        assert got == rank_work_counters(g, bounds, r, algorithm), (P, r)


def test_planted_mistake_is_caught():
    """A plausible attribution mistake (crediting type-3 triangles to the
    owner of the lowest instead of the middle vertex) disagrees with the
    simulation: the pins above can tell the two apart."""
    g = K.gen_rmat(10, 16, seed=8)
    P = 8
    bounds = O.even_bounds(g.n, P)
    deg = np.diff(g.offsets).astype(np.int64)
    ooff, oh = O.orient(g)
    owner = lambda x: int(np.searchsorted(bounds, x, side="right") - 1)  # noqa: E731
    wrong = np.zeros(P, np.int64)
    for a in range(g.n):
        na = set(oh[ooff[a]:ooff[a + 1]].tolist())
        for b in na:
            for c in na & set(oh[ooff[b]:ooff[b + 1]].tolist()):
                ro = (owner(a), owner(b), owner(c))
                if len(set(ro)) == 3:
                    wrong[ro[0]] += 1
    sim = np.array([O.rank_sim(g, P, r, O.CETRIC)[1] for r in range(P)])
    assert sim.sum() == wrong.sum()
    assert not np.array_equal(sim, wrong)


@pytest.mark.parametrize("seed", [3, 4])
def test_pseudo2core_counters(seed):
    """Config::pseudo2core (cetric_edge_iterator.h:239-242): tails with
    d+(v) < 2 are skipped and counted; counts are unchanged; the work counters
    lose exactly the skipped tails' terms, recomputed here in numpy."""
    g = K.gen_rmat(11, 8, seed=seed)
    t0, pv0, s0 = O.run_shmem(g, threads=3)
    t1, pv1, s1 = O.run_shmem(g, threads=3, pseudo2core=True)
    assert t0 == t1
    np.testing.assert_array_equal(pv0, pv1)
    ooff, oh = O.orient(g)
    dp = np.diff(ooff).astype(np.int64)
    tails = np.repeat(np.arange(g.n), dp)
    keep = dp[tails] >= 2
    assert s1["skipped_nodes"] == int(np.sum(dp < 2))
    assert s1["wedge_checks"] == int(dp[oh][keep].sum())
    assert s1["intersection_size_local"] == int((dp[tails] + dp[oh])[keep].sum())
    assert s0["skipped_nodes"] == 0
    assert s0["wedge_checks"] == int(dp[oh].sum())


def test_single_rank_katric_is_one_thread_shmetric():
    """CeTriC at P = 1: no ghosts, no messages -- the same emits and work
    counters as shmetric's one-thread local phase."""
    g = K.gen_rmat(11, 16, seed=6)
    loc, t3, pv, st = O.rank_sim(g, 1, 0, O.CETRIC)
    t, pv1, st1 = O.run_shmem(g, threads=1)
    assert (loc, t3) == (t, 0)
    np.testing.assert_array_equal(pv, pv1)
    for k in ("wedge_checks", "intersection_size_local"):
        assert st[k] == st1[k]


def test_sampled_timing_helper():
    g = K.gen_rmat(10, 16, seed=2)
    t, _, _ = O.run_shmem(g, per_vertex=False)
    tri, _, st = O.run_shmem_sample(g, stride=1)
    assert tri == t and st["sample_edges"] == g.m
    tri4, _, st4 = O.run_shmem_sample(g, stride=4)
    ooff, _ = O.orient(g)
    assert st4["sample_edges"] == int(np.diff(ooff.astype(np.int64))[::4].sum())
    assert tri4 <= t
