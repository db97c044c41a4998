"""Parity of the HIP plain path (kt_count / kt_orient) with the CPU oracle."""
import numpy as np
import pytest

import katric_amd as K
from oracle import oracle as O
from tests.conftest import KNOWN_TOTALS, golden

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", sorted(KNOWN_TOTALS))
def test_reference_examples(ctx, name):
    g = K.read_metis(golden(name))
    total, pv = ctx.count(g)
    assert total == KNOWN_TOTALS[name]
    ot, opv, _ = O.run_shmem(g)
    assert total == ot
    np.testing.assert_array_equal(pv, opv)


def test_graphio_binary_fixture(ctx):
    g = K.read_graphio_binary(golden("hiv.first_out"), golden("hiv.head"))
    assert ctx.count(g)[0] == 45


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_gnm_small(ctx, seed):
    g = K.gen_gnm(2000, 30000, seed)
    total, pv = ctx.count(g)
    ot, opv, _ = O.run_shmem(g)
    assert total == ot
    np.testing.assert_array_equal(pv, opv)
    assert int(pv.sum()) == 3 * total


@pytest.mark.parametrize("scale", [10, 14])
def test_rmat_small(ctx, scale):
    g = K.gen_rmat(scale, 16, seed=scale)
    total, pv = ctx.count(g)
    ot, opv, ost = O.run_shmem(g, threads=4)
    assert total == ot
    np.testing.assert_array_equal(pv, opv)
    st = ctx.stats()
    assert st["intersection_size_local"] == ost["intersection_size_local"]
    assert st["wedge_checks"] == ost["wedge_checks"]


def test_orient_matches_oracle(ctx):
    g = K.gen_rmat(12, 16, seed=7)
    og = ctx.orient(g)
    ooff, oheads = O.orient(g)
    np.testing.assert_array_equal(og.offsets, ooff)
    np.testing.assert_array_equal(og.heads, oheads)


def test_orient_unsorted_many_tiles(ctx):
    # shuffled rows of a graph with ~2 x 10^5 oriented edges: the radix sort
    # runs over dozens of 4096-key tiles (the last one partial) and must give
    # exactly the oracle's ascending oriented lists
    g = K.gen_rmat(13, 16, seed=7)
    rng = np.random.default_rng(8)
    heads = g.heads.copy()
    for v in range(g.n):
        a, b = int(g.offsets[v]), int(g.offsets[v + 1])
        heads[a:b] = rng.permutation(heads[a:b])
    og = ctx.orient(K.Graph(g.n, g.offsets, heads))
    ooff, oheads = O.orient(g)
    np.testing.assert_array_equal(og.offsets, ooff)
    np.testing.assert_array_equal(og.heads, oheads)


def test_unsorted_input(ctx):
    # metis-sample's lists are not ascending in the file
    g = K.read_metis(golden("metis-sample.metis"))
    og = ctx.orient(g)
    ooff, oheads = O.orient(g)
    np.testing.assert_array_equal(og.offsets, ooff)
    np.testing.assert_array_equal(og.heads, oheads)
    rng = np.random.default_rng(5)
    g2 = K.gen_rmat(11, 16, seed=3)
    heads = g2.heads.copy()
    for v in range(g2.n):
        a, b = int(g2.offsets[v]), int(g2.offsets[v + 1])
        heads[a:b] = rng.permutation(heads[a:b])
    g3 = K.Graph(g2.n, g2.offsets, heads)
    t3, pv3 = ctx.count(g3)
    t2, pv2 = ctx.count(g2)
    assert t3 == t2
    np.testing.assert_array_equal(pv3, pv2)


def _clique_minus_matching(n, dup=False):
    """K_n minus the perfect matching (0,2), (1,3), (4,5), (6,7), ...: every
    vertex has degree n - 2, so the (degree, id) orientation sends every edge
    to the larger id and vertex 0's oriented list holds n - 2 ids.  With dup,
    edges (0,4) and (1,5) are traded for (4,5) and a second copy of (0,1):
    degrees stay n - 2 and vertex 0's oriented list (n - 2 ids) holds 1 twice."""
    assert n % 2 == 0 and n >= 8
    iu, ju = np.triu_indices(n, 1)
    keep = ~((iu == 0) & (ju == 2)) & ~((iu == 1) & (ju == 3)) & ~((iu >= 4) & (iu % 2 == 0) & (ju == iu + 1))
    e = np.stack([iu[keep], ju[keep]], 1)
    if dup:
        drop = ((e[:, 0] == 0) & (e[:, 1] == 4)) | ((e[:, 0] == 1) & (e[:, 1] == 5))
        e = np.concatenate([e[~drop], [[4, 5], [0, 1]]])
    src = np.concatenate([e[:, 0], e[:, 1]])
    dst = np.concatenate([e[:, 1], e[:, 0]])
    order = np.lexsort((dst, src))
    off = np.zeros(n + 1, np.uint64)
    np.add.at(off, src + 1, 1)
    return K.Graph(n, np.cumsum(off).astype(np.uint64), dst[order].astype(np.uint32))


def _shuffle_rows(g, seed):
    rng = np.random.default_rng(seed)
    heads = g.heads.copy()
    for v in range(g.n):
        a, b = int(g.offsets[v]), int(g.offsets[v + 1])
        heads[a:b] = rng.permutation(heads[a:b])
    return K.Graph(g.n, g.offsets, heads)


@pytest.mark.parametrize("n", [10, 18, 20, 68, 300, 4098, 4100])
def test_unsorted_list_lengths(ctx, n):
    # shuffled rows whose oriented lists run through every length from 0 to
    # n - 2: lists of up to 4096 ids are sorted per list on the device (by a
    # lane up to 16, a wave up to 64, a workgroup up to 4096), a longer one
    # sends the whole graph to the radix sort (n = 4100)
    g = _clique_minus_matching(n)
    gs = _shuffle_rows(g, n)
    og = ctx.orient(gs)
    # equal degrees: every edge points to the larger id
    rows = np.repeat(np.arange(n), np.diff(g.offsets.astype(np.int64)))
    up = g.heads > rows
    ooff = np.zeros(n + 1, np.uint64)
    ooff[1:] = np.cumsum(np.bincount(rows[up], minlength=n))
    np.testing.assert_array_equal(og.offsets, ooff)
    np.testing.assert_array_equal(og.heads, g.heads[up])
    # K_n minus a perfect matching: each of the n/2 missing edges was in n - 2
    # triangles of K_n, and no triangle holds two of them
    t, pv = ctx.count(gs)
    assert t == n * (n - 1) * (n - 2) // 6 - (n // 2) * (n - 2)
    assert int(pv.sum()) == 3 * t
    if n <= 300:
        ot, opv, _ = O.run_shmem(g)
        assert t == ot
        np.testing.assert_array_equal(pv, opv)
    # a duplicate edge in vertex 0's list (n - 2 ids) is found by the sort
    gd = _shuffle_rows(_clique_minus_matching(n, dup=True), n + 1)
    with pytest.raises(K.KatricError) as e:
        ctx.count(gd)
    assert e.value.status == K.KT_EINVAL
    assert ctx.count(gs)[0] == t


def test_empty_and_tiny(ctx):
    assert ctx.count(K.Graph(0, np.zeros(1, np.uint64), np.zeros(0, np.uint32)))[0] == 0
    g = K.Graph(5, np.zeros(6, np.uint64), np.zeros(0, np.uint32))
    t, pv = ctx.count(g)
    assert t == 0 and not pv.any()
    g = K.Graph.from_edges(3, [(0, 1), (1, 2), (0, 2)])
    t, pv = ctx.count(g)
    assert t == 1 and pv.tolist() == [1, 1, 1]


@pytest.mark.parametrize("n", [300, 1200])
def test_complete_graph(ctx, n):
    # K_n: C(n,3) triangles, C(n-1,2) per vertex; n = 1200 puts tails in the
    # workgroup bin (512 < d+ <= 2048)
    iu = np.triu_indices(n, 1)
    g = K.Graph.from_edges(n, np.stack(iu, 1))
    t, pv = ctx.count(g)
    assert t == n * (n - 1) * (n - 2) // 6
    assert (pv == (n - 1) * (n - 2) // 2).all()


def test_huge_tail_bin(ctx):
    # K_{a,a} plus a perfect matching on the right side: every left vertex has
    # d+ = a > 2048 (binary-search bin); triangles = a * a/2, per vertex
    # a/2 (left) and a (right).
    a = 2100
    left = np.arange(a)
    right = np.arange(a, 2 * a)
    bip = np.stack(np.meshgrid(left, right, indexing="ij"), -1).reshape(-1, 2)
    match = np.stack([right[0::2], right[1::2]], 1)
    g = K.Graph.from_edges(2 * a, np.concatenate([bip, match]))
    t, pv = ctx.count(g)
    assert t == a * (a // 2)
    assert (pv[:a] == a // 2).all() and (pv[a:] == a).all()


def test_invalid_inputs(ctx):
    bad = [
        K.Graph(2, np.array([0, 1, 2], np.uint64), np.array([1, 5], np.uint32)),   # head out of range
        K.Graph(2, np.array([0, 2, 3], np.uint64), np.array([0, 1, 0], np.uint32)),  # self loop
        K.Graph(3, np.array([0, 2, 3, 4], np.uint64), np.array([1, 1, 0, 0], np.uint32)),  # duplicate
        K.Graph(3, np.array([0, 1, 1, 2], np.uint64), np.array([1, 1], np.uint32)),  # asymmetric
    ]
    for g in bad:
        with pytest.raises(K.KatricError) as e:
            ctx.count(g)
        assert e.value.status == K.KT_EINVAL
    # the context stays usable
    assert ctx.count(K.Graph.from_edges(3, [(0, 1), (1, 2), (0, 2)]))[0] == 1


def test_device_pointers(ctx):
    g = K.gen_rmat(12, 16, seed=11)
    dg = ctx.upload(g)
    pv = ctx.alloc(g.n * 8)
    t = ctx.count_device(dg, per_vertex_out=pv)
    ot, opv, _ = O.run_shmem(g)
    assert t == ot
    np.testing.assert_array_equal(pv.download(np.uint64, g.n), opv)
    dg.free()
    pv.free()


def test_torch_tensor_interop():
    # torch ships its own HIP runtime: importing torch before the library makes
    # both share it; device tensors are then valid kt_count inputs.
    import subprocess
    import sys
    code = (
        "import torch, numpy as np, katric_amd as K\n"
        "g = K.gen_rmat(11, 16, seed=2)\n"
        "off = torch.from_numpy(g.offsets.view(np.int64)).cuda()\n"
        "hd = torch.from_numpy(g.heads.view(np.int32)).cuda()\n"
        "pv = torch.zeros(g.n, dtype=torch.int64, device='cuda')\n"
        "torch.cuda.synchronize()\n"
        "ctx = K.Context(0)\n"
        "t = ctx.count_device(g.n, off, hd, pv)\n"
        "t2, pv2 = ctx.count(g)\n"
        "assert t == t2, (t, t2)\n"
        "assert np.array_equal(pv.cpu().numpy().view(np.uint64), pv2)\n"
        "print('ok', t)\n"
    )
    res = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300,
                         cwd=__import__("os").path.dirname(__import__("os").path.dirname(__file__)))
    assert res.returncode == 0, res.stdout + res.stderr
    assert res.stdout.startswith("ok")


def test_deterministic_repeat(ctx):
    g = K.gen_rmat(15, 16, seed=5)
    a = ctx.count(g)
    b = ctx.count(g)
    assert a[0] == b[0]
    np.testing.assert_array_equal(a[1], b[1])
