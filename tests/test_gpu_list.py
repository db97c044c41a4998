"""Triangle listing (kt_list_triangles): every triangle once, as the
reference's emit callback receives it -- (v, u, w) with v < u < w in the
(degree, id) order -- checked against an enumeration on the CPU."""
import numpy as np
import pytest

import katric_amd as K
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def enumerate_triangles(g):
    """Sorted (v, u, w) triples by the (degree, id) orientation, plain Python."""
    deg = np.diff(g.offsets.astype(np.int64))
    rank = lambda a: (int(deg[a]), int(a))  # noqa: E731
    out_nb = []
    for v in range(g.n):
        row = g.heads[int(g.offsets[v]):int(g.offsets[v + 1])]
        out_nb.append({int(u) for u in row if rank(u) > rank(v)})
    tri = set()
    for v in range(g.n):
        for u in out_nb[v]:
            for w in out_nb[v] & out_nb[u]:
                tri.add((v, u, w))
    return tri


def check_listing(ctx, g):
    tri = ctx.list_triangles(g)
    total, _ = ctx.count(g, per_vertex=False)
    assert len(tri) == total
    got = {tuple(int(x) for x in t) for t in tri}
    assert len(got) == len(tri), "a triangle was listed twice"
    assert got == enumerate_triangles(g)
    return tri


def _hub_graph():
    n = 400
    iu = np.triu_indices(60, 1)
    clique = np.stack(iu, 1) + 1
    hub = np.stack([np.zeros(n - 1, np.int64), np.arange(1, n)], 1)
    ring = np.stack([np.arange(61, n - 1), np.arange(62, n)], 1)
    return K.Graph.from_edges(n, np.concatenate([clique, hub, ring]))


def test_list_reference_examples(ctx):
    from tests.conftest import golden
    for name, tri in (("triangle.metis", 1), ("metis-sample.metis", 4), ("hiv.metis", 45)):
        assert len(check_listing(ctx, K.read_metis(golden(name)))) == tri


@pytest.mark.parametrize("name", ["rmat11", "gnm", "hub", "k40"])
def test_list_matches_enumeration(ctx, name):
    g = {"rmat11": lambda: K.gen_rmat(11, 16, seed=41),
         "gnm": lambda: K.gen_gnm(2000, 20000, 42),
         "hub": _hub_graph,
         "k40": lambda: K.Graph.from_edges(40, np.stack(np.triu_indices(40, 1), 1))}[name]()
    check_listing(ctx, g)


def test_list_every_bin(ctx):
    # small bins push most tails to the workgroup and 16-wave teams
    saved = ctx.params()
    try:
        for wave, group in ((16, 32), (2, 4), (128, 2048)):
            ctx.set_params(wave_max_deg=wave, group_max_deg=group, strategy=K.STRATEGY_GROUPED8)
            check_listing(ctx, K.gen_rmat(12, 16, seed=43))
    finally:
        ctx.set_params(**saved)


def test_list_per_vertex_consistency(ctx):
    # every vertex appears in its triangles exactly t[x] times
    g = K.gen_rmat(14, 16, seed=44)
    tri = ctx.list_triangles(g)
    _, pv = ctx.count(g)
    np.testing.assert_array_equal(np.bincount(tri.ravel(), minlength=g.n).astype(np.uint64), pv)
    _, opv, _ = O.run_shmem(g, threads=4)
    np.testing.assert_array_equal(pv, opv)


def test_list_capacity_too_small(ctx):
    g = K.gen_rmat(10, 16, seed=45)
    total, _ = ctx.count(g, per_vertex=False)
    with pytest.raises(K.KatricError) as e:
        ctx.list_triangles(g, capacity=total - 1)
    assert e.value.status == K.KT_ERANGE
    assert len(ctx.list_triangles(g, capacity=total + 10)) == total
    assert len(ctx.list_triangles(K.Graph.from_edges(5, [(0, 1), (1, 2), (3, 4)]))) == 0
