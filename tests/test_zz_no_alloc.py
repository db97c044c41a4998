"""kt_count_noalloc (the entry point that allocates nothing) against
kt_count, through the C ABI: two calls in a row, the second on a smaller
graph, per-vertex counts included; a graph beyond the fixed workspace is
KT_ENOMEM and leaves the entry point usable."""
import ctypes as C

import numpy as np
import pytest

import katric_amd as K
from katric_amd._ffi import lib

pytestmark = pytest.mark.gpu

SCRATCH = 64 << 20  # KT_NOALLOC_SCRATCH_BYTES


def noalloc(ctx, dg, d_pv):
    total = C.c_uint64(0)
    rc = lib().kt_count_noalloc(ctx.handle, dg.n, dg.offsets.ptr, dg.heads.ptr, C.byref(total),
                                d_pv.ptr if d_pv is not None else None)
    return rc, int(total.value)


def test_noalloc_matches_kt_count():
    with K.Context(0) as ctx:
        big = K.gen_rmat(16, 16, seed=3)      # ~1.9 M slots: inside the fixed workspace
        small = K.gen_gnm(3000, 20000, seed=4)
        for g in (big, small):                # the second call is on a smaller input
            want_t, want_pv = ctx.count(g)
            dg = ctx.upload(g)
            d_pv = ctx.alloc(g.n * 8)
            before = ctx.device_bytes()
            rc, t = noalloc(ctx, dg, d_pv)
            assert rc == 0, ctx.check(rc, "kt_count_noalloc")
            assert ctx.device_bytes() == before   # nothing allocated through the context
            assert t == want_t
            np.testing.assert_array_equal(d_pv.download(np.uint64, g.n), want_pv)
            rc, t2 = noalloc(ctx, dg, None)       # without per-vertex counts
            assert rc == 0 and t2 == want_t
            dg.free()
            d_pv.free()


def test_noalloc_unsorted_and_errors():
    with K.Context(0) as ctx:
        g = K.gen_rmat(12, 16, seed=5)
        rng = np.random.default_rng(2)
        heads = g.heads.copy()
        for v in range(g.n):
            a, b = int(g.offsets[v]), int(g.offsets[v + 1])
            heads[a:b] = rng.permutation(heads[a:b])
        want_t, want_pv = ctx.count(g)
        dg = ctx.upload(K.Graph(g.n, g.offsets, heads))
        d_pv = ctx.alloc(g.n * 8)
        rc, t = noalloc(ctx, dg, d_pv)
        assert rc == 0 and t == want_t
        np.testing.assert_array_equal(d_pv.download(np.uint64, g.n), want_pv)
        dg.free()
        d_pv.free()
        # beyond the fixed workspace: KT_ENOMEM, nothing run
        huge = K.gen_rmat(21, 16, seed=1)     # ~6.4e7 slots, > 64 MiB of workspace
        dg = ctx.upload(huge)
        rc, _ = noalloc(ctx, dg, None)
        assert rc == K.KT_ENOMEM
        assert b"workspace" in lib().kt_last_error(ctx.handle)
        dg.free()
        # asymmetric input is still rejected, and the entry stays usable
        bad = K.Graph(4, np.array([0, 1, 2, 2, 2], np.uint64), np.array([1, 2], np.uint32))
        dg = ctx.upload(bad)
        assert noalloc(ctx, dg, None)[0] == K.KT_EINVAL
        dg.free()
        ok = K.gen_gnm(500, 3000, seed=1)
        dg = ctx.upload(ok)
        rc, t = noalloc(ctx, dg, None)
        assert rc == 0 and t == ctx.count(ok)[0]
        dg.free()
