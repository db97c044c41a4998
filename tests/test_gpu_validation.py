"""Input validation of every counting entry point (plain, listing, orient,
batch, serve, device shard producer), the Config knobs of the shared-memory
path (pseudo2core, binary-search probes) and the device intersection of
caller-given lists, against the oracle and closed forms."""
import numpy as np
import pytest

import katric_amd as K
from katric_amd import batch as KB
from katric_amd import serve as KS
from katric_amd import shard as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu

# 0 -> 1 and 1 -> 2 only: half of the slots are kept under the (degree, id)
# order (0 -> 1 is kept, 1 -> 2 is not), so a kept-count check alone accepts it
COUNTER = K.Graph(4, np.array([0, 1, 2, 2, 2], np.uint64), np.array([1, 2], np.uint32))


def kept_slots(g):
    """Slots v -> u with (deg v, v) < (deg u, u), counted in numpy."""
    deg = np.diff(g.offsets.astype(np.int64))
    tails = np.repeat(np.arange(g.n), deg)
    u = g.heads.astype(np.int64)
    return int(np.sum((deg[tails] < deg[u]) | ((deg[tails] == deg[u]) & (tails < u))))


def planted_asymmetric(seed=3):
    """RMAT-12 with one slot a -> b replaced by a -> c (c not adjacent to a):
    the pairs {a, b} and {a, c} then have one slot each, and a, b, c are
    chosen so that the kept count stays exactly nnz / 2."""
    g = K.gen_rmat(12, 16, seed=seed)
    rng = np.random.default_rng(seed)
    deg = np.diff(g.offsets.astype(np.int64))

    def before(x, y):
        return deg[x] < deg[y] or (deg[x] == deg[y] and x < y)

    for _ in range(100000):
        a = int(rng.integers(g.n))
        if deg[a] < 2:
            continue
        row = g.heads[g.offsets[a]:g.offsets[a + 1]]
        b = int(row[rng.integers(len(row))])
        c = int(rng.integers(g.n))
        if c == a or c in set(row.tolist()):
            continue
        # originally exactly one of a -> b, b -> a is kept; afterwards a -> c
        # and b -> a remain: keep the count when exactly one of them is kept
        if before(a, c) == before(b, a):
            continue
        heads = g.heads.copy()
        new = np.sort(np.array([x for x in row.tolist() if x != b] + [c], np.uint32))
        heads[g.offsets[a]:g.offsets[a + 1]] = new
        h = K.Graph(g.n, g.offsets, heads)
        if 2 * kept_slots(h) == h.nnz:
            return h
    raise AssertionError("no planted case found")


def test_counterexample_kept_count_balanced():
    assert 2 * kept_slots(COUNTER) == COUNTER.nnz


def test_symmetry_exhaustive_4_vertices(ctx):
    # every CSR on 4 vertices whose rows are sorted subsets of the other
    # vertices (8^4 = 4096 inputs): kt_count accepts exactly the symmetric
    # ones -- out-lists and in-lists agree, the reference's orientation
    # contract (tests/orientation_test.cpp:81-133) -- and counts them right
    import itertools
    n = 4
    others = [[u for u in range(n) if u != v] for v in range(n)]
    rows_of = [[[o[i] for i in range(3) if mask >> i & 1] for mask in range(8)] for o in others]
    accepted = 0
    for pick in itertools.product(range(8), repeat=n):
        rows = [rows_of[v][pick[v]] for v in range(n)]
        off = np.zeros(n + 1, np.uint64)
        off[1:] = np.cumsum([len(r) for r in rows])
        heads = np.array([u for r in rows for u in r], np.uint32)
        g = K.Graph(n, off, heads)
        sym = all((v in rows[u]) for v in range(n) for u in rows[v])
        if sym:
            t = ctx.count(g, per_vertex=False)[0]
            assert t == sum(1 for a, b, c in itertools.combinations(range(n), 3)
                            if b in rows[a] and c in rows[a] and c in rows[b]), rows
            accepted += 1
        else:
            with pytest.raises(K.KatricError) as e:
                ctx.count(g, per_vertex=False)
            assert e.value.status == K.KT_EINVAL, rows
    assert accepted == 2 ** 6  # one symmetric CSR per edge subset of K_4


@pytest.mark.parametrize("case", ["n4", "rmat12-planted"])
def test_asymmetric_plain_paths(ctx, case):
    g = COUNTER if case == "n4" else planted_asymmetric()
    for call in (lambda: ctx.count(g), lambda: ctx.orient(g), lambda: ctx.list_triangles(g)):
        with pytest.raises(K.KatricError) as e:
            call()
        assert e.value.status == K.KT_EINVAL
    # the device copy of the input is checked the same way
    dg = ctx.upload(g)
    with pytest.raises(K.KatricError) as e:
        ctx.count_device(dg)
    assert e.value.status == K.KT_EINVAL
    dg.free()
    assert ctx.count(K.Graph.from_edges(3, [(0, 1), (1, 2), (0, 2)]))[0] == 1


def test_asymmetric_unsorted_rows(ctx):
    g = planted_asymmetric(seed=5)
    rng = np.random.default_rng(1)
    heads = g.heads.copy()
    for v in range(g.n):
        a, b = int(g.offsets[v]), int(g.offsets[v + 1])
        heads[a:b] = rng.permutation(heads[a:b])
    with pytest.raises(K.KatricError) as e:
        ctx.count(K.Graph(g.n, g.offsets, heads))
    assert e.value.status == K.KT_EINVAL


def test_asymmetric_in_batch(ctx):
    good = [K.gen_gnm(300, 2000, seed=4), K.read_metis(__import__("tests.conftest", fromlist=["golden"]).golden(
        "hiv.metis"))]
    for bad in (COUNTER, planted_asymmetric()):
        b = KB.from_graphs([good[0], bad, good[1]])
        with pytest.raises(K.KatricError) as e:
            KB.count_batch(ctx, b)
        assert e.value.status == K.KT_EINVAL
    # one graph with a surplus of kept slots, another with a shortfall:
    # together the kept count is balanced (a per-union count check passes)
    # surplus: 1 <-> 2 plus 0 -> 1, kept because (1, 0) < (1, 1); short: 1 -> 0 alone, not kept
    surplus = K.Graph(3, np.array([0, 1, 2, 3], np.uint64), np.array([1, 2, 1], np.uint32))
    short = K.Graph(2, np.array([0, 0, 1], np.uint64), np.array([0], np.uint32))
    b = KB.from_graphs([surplus, short])
    assert 2 * sum(kept_slots(x) for x in (surplus, short)) == b.nnz
    with pytest.raises(K.KatricError) as e:
        KB.count_batch(ctx, b)
    assert e.value.status == K.KT_EINVAL
    tot = KB.count_batch(ctx, KB.from_graphs(good))
    assert tot[1] == 45


def test_cross_graph_edges_in_batch(ctx):
    # three triangles each split 2 + 1 over two graphs of 6 vertices: the union
    # is symmetric and every graph's per-vertex sum is a multiple of 3, but
    # heads leave their row's graph
    tri = [(0, 1, 6), (2, 3, 7), (4, 5, 8)]
    edges = []
    for a, b, c in tri:
        edges += [(a, b), (b, c), (a, c)]
    union = K.Graph.from_edges(12, edges)
    b = KB.Batch(np.array([0, 6, 12], np.uint64), union.offsets, union.heads)
    with pytest.raises(K.KatricError) as e:
        KB.count_batch(ctx, b)
    assert e.value.status == K.KT_EINVAL
    # the same union as one graph is valid
    assert ctx.count(union)[0] == 3


def test_asymmetric_served_request_is_isolated(ctx):
    srv = KS.Server(ctx)
    goods = [K.gen_gnm(500, 3000, seed=s) for s in range(6)]
    want = [ctx.count(g)[0] for g in goods]
    bad = planted_asymmetric()
    tickets = [srv.submit(g) for g in goods[:3]] + [srv.submit(bad), srv.submit(COUNTER)] + \
              [srv.submit(g) for g in goods[3:]]
    res = [srv.wait(t) for t in tickets]
    srv.stop()
    st = [r.status for r in res]
    assert st[3] == K.KT_EINVAL and st[4] == K.KT_EINVAL
    ok = [r for i, r in enumerate(res) if i not in (3, 4)]
    assert all(r.status == 0 for r in ok)
    assert [r.total for r in ok] == want


def test_asymmetric_shard_graph(ctx):
    g = planted_asymmetric()
    bounds = O.even_bounds(g.n, 4)
    for algo in (S.CETRIC, S.DITRIC):
        with pytest.raises(K.KatricError) as e:
            S.count_shard_graph(ctx, g, 4, 0, algo)
        assert e.value.status == K.KT_EINVAL
    good = K.gen_rmat(12, 16, seed=3)
    loc, t3, _ = S.count_shard_graph(ctx, good, 4, 1, S.CETRIC)
    ol, og, _, _ = O.rank_sim(good, 4, 1, O.CETRIC, bounds=bounds)
    assert (loc, t3) == (ol, og)


# ---------------------------------------------------------------------------
# Config knobs of the shared-memory path
# ---------------------------------------------------------------------------
@pytest.fixture
def knobs(ctx):
    saved = ctx.params()
    yield ctx
    ctx.set_params(**saved)


@pytest.mark.parametrize("seed", [2, 9])
def test_pseudo2core_matches_oracle(knobs, seed):
    g = K.gen_rmat(13, 8, seed=seed)
    knobs.set_params(pseudo2core=1)
    total, pv = knobs.count(g)
    st = knobs.stats()
    ot, opv, ost = O.run_shmem(g, threads=4, pseudo2core=True)
    assert total == ot
    np.testing.assert_array_equal(pv, opv)
    for k in ("wedge_checks", "intersection_size_local", "skipped_nodes"):
        assert st[k] == ost[k], k
    # closed form: skipped = #{v : d+(v) < 2}
    ooff, _ = O.orient(g)
    assert st["skipped_nodes"] == int(np.sum(np.diff(ooff.astype(np.int64)) < 2))
    # the counters without pseudo2core differ by the skipped tails' work only
    knobs.set_params(pseudo2core=0)
    knobs.count(g)
    st0 = knobs.stats()
    assert st0["skipped_nodes"] == 0 and st0["wedge_checks"] >= st["wedge_checks"]


def test_pseudo2core_batch(knobs):
    b = KB.gen_batch(24, seed=7, n_min=64, n_max=2048)
    knobs.set_params(pseudo2core=1)
    tot = KB.count_batch(knobs, b)
    skipped = knobs.stats()["skipped_nodes"]
    want, wskip = [], 0
    for i in range(b.num_graphs):
        t, _, s = O.run_shmem(b.graph(i), pseudo2core=True)
        want.append(t)
        wskip += s["skipped_nodes"]
    assert tot.tolist() == want
    assert skipped == wskip


def test_pseudo2core_rejected_on_shards(knobs):
    g = K.gen_rmat(10, 16, seed=1)
    knobs.set_params(pseudo2core=1)
    with pytest.raises(K.KatricError) as e:
        S.count_shard_graph(knobs, g, 2, 0, S.CETRIC)
    assert e.value.status == K.KT_EINVAL


@pytest.mark.parametrize("thr", [1, 16, 300])
def test_binary_search_probes(knobs, thr):
    # every tail above thr probes by binary search in global memory
    g = K.gen_rmat(13, 16, seed=4)
    knobs.set_params(bsearch_min_deg=thr)
    total, pv = knobs.count(g)
    ot, opv, ost = O.run_shmem(g, threads=4)
    assert total == ot
    np.testing.assert_array_equal(pv, opv)
    st = knobs.stats()
    assert st["wedge_checks"] == ost["wedge_checks"]
    assert st["intersection_size_local"] == ost["intersection_size_local"]


def test_params_validation(knobs):
    for bad in (dict(bsearch_min_deg=0), dict(bsearch_min_deg=8193), dict(pseudo2core=2)):
        with pytest.raises(K.KatricError):
            knobs.set_params(**bad)


# ---------------------------------------------------------------------------
# kt_intersect: caller-given lists
# ---------------------------------------------------------------------------
HAND = [  # tests/intersection_test.cpp:41-52 of the reference
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


def test_intersect_hand_vectors(ctx):
    pairs = [(a, b) for a, b, _ in HAND] + [(b, a) for a, b, _ in HAND]
    counts, matches = ctx.intersect(pairs)
    want = [w for _, _, w in HAND] * 2
    assert counts.tolist() == [len(w) for w in want]
    assert [m.tolist() for m in matches] == want


def test_intersect_random_vs_oracle(ctx):
    # tests/intersection_test.cpp:101-173: random sorted sets of sizes <= 1000
    # against std::set_intersection (the oracle's merge), plus long lists
    rng = np.random.default_rng(11)
    pairs = []
    for i in range(300):
        na, nb = rng.integers(0, 1000, size=2)
        if i % 50 == 0:
            nb = 200000
        u = int(rng.integers(max(na, nb, 1), 4 * max(na, nb, 1) + 2))
        pairs.append((np.sort(rng.choice(u, size=na, replace=False)), np.sort(rng.choice(u, size=nb, replace=False))))
    counts, matches = ctx.intersect(pairs)
    for (a, b), c, m in zip(pairs, counts, matches):
        want = O.intersection(a, b, method="merge_only")
        assert int(c) == len(want)
        assert m.tolist() == want
    assert ctx.intersect(pairs, with_matches=False)[0].tolist() == counts.tolist()


def test_intersect_rejects_unsorted(ctx):
    for a, b in (([3, 2], [2]), ([1, 1], [1]), ([1, 2], [5, 4, 6])):
        with pytest.raises(K.KatricError) as e:
            ctx.intersect([(a, b)])
        assert e.value.status == K.KT_EINVAL
