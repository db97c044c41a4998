"""Every intersection strategy and bin threshold gives the oracle's counts.

The knobs (kt_params: wave/workgroup bin thresholds, streamed or grouped list
walk, grid size) only move work between kernels and code paths; totals,
per-vertex counts and the reference's work counters must not change.
"""
import itertools

import numpy as np
import pytest

import katric_amd as K
from katric_amd import batch as KB
from katric_amd import shard as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu

STRATEGIES = [K.STRATEGY_STREAMED, K.STRATEGY_GROUPED16, K.STRATEGY_GROUPED8, K.STRATEGY_GROUPED32,
              K.STRATEGY_GROUPED4, K.STRATEGY_GROUPED2]
# (wave_max_deg, group_max_deg, grid_per_cu)
BINS = [(128, 2048, 4), (8, 256, 1), (256, 512, 2), (2, 2048, 3)]


ORDERS = [K.ORDER_DEGREE, K.ORDER_VERTEX]


@pytest.fixture
def knobs(ctx):
    saved = ctx.params()

    def apply(strategy, wave, group, grid, order=K.ORDER_DEGREE):
        ctx.set_params(wave_max_deg=wave, group_max_deg=group, grid_per_cu=grid, strategy=strategy, order=order)

    yield apply
    ctx.set_params(**saved)


def _graphs():
    yield "rmat14", K.gen_rmat(14, 16, seed=21)
    yield "gnm", K.gen_gnm(5000, 60000, 22)
    # one hub joined to a clique: many short T lists (the hub's neighbours)
    # next to long ones
    n = 400
    iu = np.triu_indices(60, 1)
    clique = np.stack(iu, 1) + 1
    hub = np.stack([np.zeros(n - 1, np.int64), np.arange(1, n)], 1)
    ring = np.stack([np.arange(61, n - 1), np.arange(62, n)], 1)
    yield "hub", K.Graph.from_edges(n, np.concatenate([clique, hub, ring]))


GRAPHS = list(_graphs())


@pytest.mark.parametrize("strategy,bins,order", list(itertools.product(STRATEGIES, BINS, ORDERS)))
def test_plain_counts_every_strategy(ctx, knobs, strategy, bins, order):
    knobs(strategy, *bins, order=order)
    for name, g in GRAPHS:
        total, pv = ctx.count(g)
        ot, opv, ost = O.run_shmem(g, threads=4)
        assert total == ot, name
        np.testing.assert_array_equal(pv, opv, err_msg=name)
        st = ctx.stats()
        assert st["intersection_size_local"] == ost["intersection_size_local"], name
        assert st["wedge_checks"] == ost["wedge_checks"], name


@pytest.mark.parametrize("n", [300, 1200])
@pytest.mark.parametrize("strategy", STRATEGIES[1:])
def test_grouped_complete_graph(ctx, knobs, n, strategy):
    # K_n: lists of every length up to n - 1, several blocks per list
    knobs(strategy, 128, 2048, 4)
    iu = np.triu_indices(n, 1)
    g = K.Graph.from_edges(n, np.stack(iu, 1))
    t, pv = ctx.count(g)
    assert t == n * (n - 1) * (n - 2) // 6
    assert (pv == (n - 1) * (n - 2) // 2).all()


@pytest.mark.parametrize("order", ORDERS)
@pytest.mark.parametrize("strategy", STRATEGIES[1:])
def test_grouped_batch(ctx, knobs, order, strategy):
    knobs(strategy, 64, 1024, 4, order)
    b = KB.gen_batch(40, seed=9, n_min=64, n_max=4000)
    tot, pv = KB.count_batch(ctx, b, per_vertex=True)
    for g in range(b.num_graphs):
        t, opv, _ = O.run_shmem(b.graph(g))
        assert tot[g] == t
        a, z = int(b.vertex_base[g]), int(b.vertex_base[g + 1])
        np.testing.assert_array_equal(pv[a:z], opv)


@pytest.mark.parametrize("algo", [S.CETRIC, S.DITRIC])
@pytest.mark.parametrize("order", ORDERS)
def test_grouped_shard(ctx, knobs, algo, order):
    knobs(K.STRATEGY_GROUPED8, 32, 2048, 4, order)
    g = K.gen_rmat(12, 16, seed=5)
    P = 4
    total = 0
    for r in range(P):
        sh = S.make_shard(g, P, r, algo)
        loc, t3, pv = S.count_shard(ctx, sh)
        oloc, ot3, opv, _ = O.rank_sim(g, P, r, algo, threads=4)
        assert (loc, t3) == (oloc, ot3)
        full = np.zeros(g.n, np.uint64)
        full[sh.vertex_ids()] = pv
        np.testing.assert_array_equal(full, opv)
        total += loc + t3
    assert total == O.run_shmem(g)[0]


def test_set_params_rejects_bad_knobs(ctx):
    with pytest.raises(K.KatricError):
        ctx.set_params(strategy=7)
    with pytest.raises(K.KatricError):
        ctx.set_params(order=2)
    with pytest.raises(K.KatricError):
        ctx.set_params(min_side=2)


def _expected_streamed(g, min_side=True, hash_max=8192):
    """Streamed T-list ids, from the oracle's oriented CSR.  Forward walk:
    every oriented edge (x, u) streams |N+(u)| from x, except where x has a
    single out-neighbour (no triangle can close there; the reference still
    charges those lists to wedge_checks).  Min-side walk: |N+(u)| from x
    when |N+(u)| <= |N+(x)| or u's set is beyond the LDS tables, else
    |N+(x)| from u."""
    ooff, oh = O.orient(g)
    d = np.diff(ooff.astype(np.int64))
    x = np.repeat(np.arange(g.n), d)
    dx, du = d[x], d[oh]
    fwd = np.where(dx >= 2, du, 0)
    if not min_side:
        return int(fwd.sum())
    at_x = (du <= dx) | (du > hash_max)
    return int(np.where(at_x, fwd, dx).sum())


@pytest.mark.parametrize("bins", BINS)
def test_min_side_walk(ctx, knobs, bins):
    # min_side moves every edge's intersection to the side with the shorter
    # T list: totals, per-vertex counts and the reference's work counters
    # stay those of the forward walk (and the oracle's); only streamed_ids
    # changes, to the closed form above
    for strategy in (K.STRATEGY_STREAMED, K.STRATEGY_GROUPED8):
        knobs(strategy, *bins)
        for name, g in GRAPHS:
            ctx.set_params(min_side=0)
            t0, pv0 = ctx.count(g)
            s0 = ctx.stats()
            ctx.set_params(min_side=1)
            t1, pv1 = ctx.count(g)
            s1 = ctx.stats()
            ot, opv, ost = O.run_shmem(g, threads=4)
            assert t0 == t1 == ot, name
            np.testing.assert_array_equal(pv0, opv)
            np.testing.assert_array_equal(pv1, opv)
            for k in ("wedge_checks", "intersection_size_local"):
                assert s0[k] == s1[k] == ost[k], (name, k)
            assert s0["streamed_ids"] == _expected_streamed(g, min_side=False), name
            assert s1["streamed_ids"] == _expected_streamed(g), name


def test_min_side_batch(ctx):
    b = KB.gen_batch(48, seed=9, n_min=64, n_max=4000)
    ctx.set_params(min_side=0)
    t0, pv0 = KB.count_batch(ctx, b, per_vertex=True)
    ctx.set_params(min_side=1)
    t1, pv1 = KB.count_batch(ctx, b, per_vertex=True)
    np.testing.assert_array_equal(t0, t1)
    np.testing.assert_array_equal(pv0, pv1)
    for g in range(0, b.num_graphs, 7):
        t, opv, _ = O.run_shmem(b.graph(g))
        assert t1[g] == t
        a, z = int(b.vertex_base[g]), int(b.vertex_base[g + 1])
        np.testing.assert_array_equal(pv1[a:z], opv)
