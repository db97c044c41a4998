"""Tails above the workgroup bin: the 16-wave teams with a 16384-slot table
(k_tc_big, |S| <= 8192) and the binary-search teams beyond (k_tc_huge).

Complete graphs have closed forms (C(n,3) triangles, C(n-1,2) per vertex) and
put every tail x at |N+(x)| = n-1-x, so K_3000 reaches the big bin and K_8300
the huge one.  Lowering the workgroup bin sends ordinary tails to the big
kernel, where the oracle checks them."""
import numpy as np
import pytest

import katric_amd as K
from katric_amd import batch as KB
from katric_amd import shard as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def complete_graph(n):
    off = np.arange(n + 1, dtype=np.uint64) * np.uint64(n - 1)
    ids = np.arange(n, dtype=np.uint32)
    heads = np.empty(n * (n - 1), np.uint32)
    for v in range(n):
        row = heads[v * (n - 1):(v + 1) * (n - 1)]
        row[:v] = ids[:v]
        row[v:] = ids[v + 1:]
    return K.Graph(n, off, heads)


@pytest.fixture
def knobs(ctx):
    saved = ctx.params()
    yield ctx.set_params
    ctx.set_params(**saved)


@pytest.mark.parametrize("n", [3000, 8300])
def test_complete_graph_big_and_huge(ctx, n):
    g = complete_graph(n)
    t, pv = ctx.count(g)
    assert t == n * (n - 1) * (n - 2) // 6
    assert (pv == (n - 1) * (n - 2) // 2).all()


def test_big_bin_against_oracle(ctx, knobs):
    knobs(wave_max_deg=32, group_max_deg=64)
    for g in (K.gen_rmat(14, 16, seed=31), K.gen_gnm(6000, 120000, 32)):
        t, pv = ctx.count(g)
        ot, opv, ost = O.run_shmem(g, threads=4)
        assert t == ot
        np.testing.assert_array_equal(pv, opv)
        st = ctx.stats()
        assert st["wedge_checks"] == ost["wedge_checks"]
        assert st["intersection_size_local"] == ost["intersection_size_local"]


@pytest.mark.parametrize("algo", [S.CETRIC, S.DITRIC])
def test_big_bin_shard(ctx, knobs, algo):
    knobs(wave_max_deg=16, group_max_deg=32)
    g = K.gen_rmat(12, 16, seed=33)
    total = 0
    for r in range(4):
        loc, t3, pv = S.count_shard_graph(ctx, g, 4, r, algo)
        oloc, ot3, opv, _ = O.rank_sim(g, 4, r, algo, threads=4)
        assert (loc, t3) == (oloc, ot3)
        np.testing.assert_array_equal(pv, opv)
        total += loc + t3
    assert total == O.run_shmem(g)[0]


def test_big_bin_batch(ctx, knobs):
    knobs(wave_max_deg=16, group_max_deg=32)
    b = KB.gen_batch(24, seed=34, n_min=256, n_max=6000)
    tot, pv = KB.count_batch(ctx, b, per_vertex=True)
    for g in range(b.num_graphs):
        t, opv, _ = O.run_shmem(b.graph(g))
        assert tot[g] == t
        a, z = int(b.vertex_base[g]), int(b.vertex_base[g + 1])
        np.testing.assert_array_equal(pv[a:z], opv)
