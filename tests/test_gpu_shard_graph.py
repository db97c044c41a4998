"""Device-side shard producer (kt_count_shard_graph): every rank's counts,
built on the device from the whole graph, equal the oracle's P-rank
simulation and the host producer + kt_count_shard, for CeTriC and DiTriC."""
import numpy as np
import pytest

import katric_amd as K
from katric_amd import shard as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _shuffled_rows(g, seed):
    """The same graph with every adjacency list in random order."""
    rng = np.random.default_rng(seed)
    hd = g.heads.copy()
    for v in range(g.n):
        a, b = int(g.offsets[v]), int(g.offsets[v + 1])
        rng.shuffle(hd[a:b])
    return K.Graph(g.n, g.offsets.copy(), hd)


def check_all_ranks(ctx, g, P, algo, bounds=None, device=False):
    src = ctx.upload(g) if device else g
    total = 0
    acc = np.zeros(g.n, np.uint64)
    for r in range(P):
        loc, t3, pv = S.count_shard_graph(ctx, src, P, r, algo, bounds=bounds)
        st = ctx.stats()
        oloc, ot3, opv, ost = O.rank_sim(g, P, r, algo, bounds=bounds, threads=4)
        assert (loc, t3) == (oloc, ot3), (P, r, algo)
        np.testing.assert_array_equal(pv, opv, err_msg=f"P={P} r={r}")
        # the reference's work counters of the rank, as the host producer's shard gives them
        S.count_shard(ctx, S.make_shard(g, P, r, algo, bounds=bounds), per_vertex=False)
        hst = ctx.stats()
        for k in ("wedge_checks", "intersection_size_local", "intersection_size_global", "oriented_edges",
                  "vertices"):
            assert st[k] == hst[k], (k, P, r)
        for k in ("wedge_checks", "intersection_size_local", "intersection_size_global"):
            assert st[k] == ost[k], (k, P, r, algo)
        total += loc + t3
        acc += pv
    t, pv_all = ctx.count(g)
    assert total == t
    np.testing.assert_array_equal(acc, pv_all)
    if device:
        src.free()


@pytest.mark.parametrize("algo", [S.CETRIC, S.DITRIC])
@pytest.mark.parametrize("P", [1, 3, 16])
def test_rmat_all_ranks(ctx, algo, P):
    check_all_ranks(ctx, K.gen_rmat(12, 16, seed=P + 7), P, algo, device=(P != 3))


@pytest.mark.parametrize("algo", [S.CETRIC, S.DITRIC])
def test_uneven_and_empty_ranges(ctx, algo):
    g = K.gen_gnm(3000, 40000, seed=4)
    bounds = np.array([0, 0, 5, 6, 1000, 2990, 2990, 3000], np.uint64)
    check_all_ranks(ctx, g, 7, algo, bounds=bounds, device=True)


@pytest.mark.parametrize("algo", [S.CETRIC, S.DITRIC])
def test_unsorted_rows(ctx, algo):
    # rows in random order: message lists are put in order on the device
    g = K.gen_rmat(11, 16, seed=3)
    gs = _shuffled_rows(g, 5)
    for r in range(4):
        a = S.count_shard_graph(ctx, gs, 4, r, algo)
        b = S.count_shard_graph(ctx, g, 4, r, algo)
        assert a[:2] == b[:2]
        np.testing.assert_array_equal(a[2], b[2])
        oloc, ot3, opv, _ = O.rank_sim(g, 4, r, algo, threads=4)
        assert a[:2] == (oloc, ot3)


def test_reference_examples(ctx):
    from tests.conftest import golden
    for name, tri in (("hiv.metis", 45), ("metis-sample.metis", 4), ("triangle.metis", 1)):
        g = K.read_metis(golden(name))
        for P in (1, 2, 3):
            for algo in (S.CETRIC, S.DITRIC):
                got = sum(sum(S.count_shard_graph(ctx, g, P, r, algo, per_vertex=False)[:2]) for r in range(P))
                assert got == tri, (name, P, algo)


def test_accumulate_device_output(ctx):
    # a sweep over all ranks into one device buffer gives kt_count's per-vertex counts
    g = K.gen_rmat(13, 16, seed=11)
    dg = ctx.upload(g)
    out = ctx.alloc(g.n * 8)
    out.upload(np.full(g.n, 7, np.uint64))
    total = 0
    for r in range(16):
        loc, t3, _ = S.count_shard_graph(ctx, dg, 16, r, S.CETRIC, per_vertex_out=out, accumulate=True)
        total += loc + t3
    t, pv = ctx.count(g)
    assert total == t
    np.testing.assert_array_equal(out.download(np.uint64, g.n), pv + 7)
    # without accumulate the output is overwritten (zero outside the rank's view)
    loc, t3, _ = S.count_shard_graph(ctx, dg, 16, 5, S.DITRIC, per_vertex_out=out)
    _, _, opv, _ = O.rank_sim(g, 16, 5, S.DITRIC, threads=4)
    np.testing.assert_array_equal(out.download(np.uint64, g.n), opv)
    out.free()
    dg.free()


def test_accumulate_host_output(ctx):
    g = K.gen_gnm(2000, 30000, seed=12)
    acc = np.ones(g.n, np.uint64)
    for r in range(3):
        S.count_shard_graph(ctx, g, 3, r, S.DITRIC, per_vertex=acc, accumulate=True)
    np.testing.assert_array_equal(acc, ctx.count(g)[1] + 1)


def test_device_matches_host_producer_rmat16(ctx):
    g = K.gen_rmat(16, 16, seed=1)
    dg = ctx.upload(g)
    for algo in (S.CETRIC, S.DITRIC):
        for r in (0, 7):
            sh = S.make_shard(g, 16, r, algo)
            loc, t3, pv = S.count_shard(ctx, sh)
            full = np.zeros(g.n, np.uint64)
            full[sh.vertex_ids()] = pv
            dl, dt3, dpv = S.count_shard_graph(ctx, dg, 16, r, algo)
            assert (dl, dt3) == (loc, t3)
            np.testing.assert_array_equal(dpv, full)
    dg.free()


def test_rejects_malformed_input(ctx):
    g = K.gen_rmat(10, 8, seed=2)
    bad = K.Graph(g.n, g.offsets.copy(), g.heads.copy())
    bad.heads[-1] = g.n + 5  # a head outside [0, n) in the last rank's slice
    with pytest.raises(K.KatricError) as e:
        S.count_shard_graph(ctx, bad, 4, 3)
    assert e.value.status == K.KT_EINVAL
    with pytest.raises(K.KatricError):
        S.count_shard_graph(ctx, g, 4, 4)  # rank out of range
    with pytest.raises(K.KatricError):
        S.count_shard_graph(ctx, g, 2, 0, bounds=np.array([0, 10, g.n - 1], np.uint64))  # bounds miss a vertex
    off = g.offsets.copy()
    off[6] = off[7] + 1  # not monotone
    with pytest.raises(K.KatricError):
        S.count_shard_graph(ctx, K.Graph(g.n, off, g.heads.copy()), 4, 0)
    # the context still counts after the errors
    assert S.count_shard_graph(ctx, g, 4, 0, per_vertex=False)[:2] == O.rank_sim(g, 4, 0)[:2]


@pytest.mark.parametrize("algo", [S.CETRIC, S.DITRIC])
def test_dense_messages(ctx, algo):
    # average degree 800 over 8 ranks: most contracted-phase messages carry
    # 256-600 ghost entries, so workgroup teams (receivers shared a few per
    # grab) count them, not only single waves
    check_all_ranks(ctx, K.gen_gnm(4000, 1600000, seed=11), 8, algo, device=True)
