"""Parity of the HIP shard path (kt_count_shard) with the oracle's P-rank
simulation, per rank, for CeTriC (cetric-x) and DiTriC (patric)."""
import numpy as np
import pytest

import katric_amd as K
from katric_amd import shard as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def check_rank(ctx, g, P, r, algo, bounds=None):
    sh = S.make_shard(g, P, r, algo, bounds=bounds)
    loc, t3, pv = S.count_shard(ctx, sh)
    oloc, ot3, opv, _ = O.rank_sim(g, P, r, algo, bounds=bounds, threads=4)
    assert (loc, t3) == (oloc, ot3), (P, r, algo)
    full = np.zeros(g.n, np.uint64)
    full[sh.vertex_ids()] = pv
    np.testing.assert_array_equal(full, opv)
    return loc + t3


@pytest.mark.parametrize("algo", [S.CETRIC, S.DITRIC])
@pytest.mark.parametrize("P", [1, 2, 4, 16])
def test_shard_rmat_all_ranks(ctx, algo, P):
    g = K.gen_rmat(12, 16, seed=P)
    total = sum(check_rank(ctx, g, P, r, algo) for r in range(P))
    assert total == O.run_shmem(g)[0]


@pytest.mark.parametrize("algo", [S.CETRIC, S.DITRIC])
def test_shard_uneven_ranges(ctx, algo):
    g = K.gen_gnm(3000, 40000, seed=4)
    bounds = np.array([0, 5, 6, 1000, 2990, 3000], np.uint64)
    total = sum(check_rank(ctx, g, 5, r, algo, bounds=bounds) for r in range(5))
    assert total == O.run_shmem(g)[0]


def test_shard_reference_examples(ctx):
    for name, tri in (("hiv.metis", 45), ("metis-sample.metis", 4)):
        from tests.conftest import golden
        g = K.read_metis(golden(name))
        for P in (2, 3):
            for algo in (S.CETRIC, S.DITRIC):
                assert sum(check_rank(ctx, g, P, r, algo) for r in range(P)) == tri


def test_shard_rmat16_rank0(ctx):
    g = K.gen_rmat(16, 16, seed=1)
    for algo in (S.CETRIC, S.DITRIC):
        check_rank(ctx, g, 16, 0, algo)


def test_shard_bad_ghost_table(ctx):
    g = K.gen_rmat(10, 8, seed=2)
    sh = S.make_shard(g, 4, 1)
    sh.ghost_ids = sh.ghost_ids[1:].copy()
    sh.ghost_degree = sh.ghost_degree[1:].copy()
    with pytest.raises(K.KatricError) as e:
        S.count_shard(ctx, sh)
    assert e.value.status == K.KT_EINVAL


def test_shard_bad_message(ctx):
    # a message list out of order is rejected, and the context keeps working
    g = K.gen_rmat(10, 8, seed=2)
    sh = S.make_shard(g, 4, 1)
    lens = np.diff(sh.msg_offsets.astype(np.int64))
    i = int(np.nonzero(lens >= 2)[0][0])
    a, b = int(sh.msg_offsets[i]), int(sh.msg_offsets[i + 1])
    sh.msg_heads[a:b] = sh.msg_heads[a:b][::-1].copy()
    with pytest.raises(K.KatricError) as e:
        S.count_shard(ctx, sh)
    assert e.value.status == K.KT_EINVAL
    check_rank(ctx, g, 4, 1, S.CETRIC)
