"""One rank's shard of the distributed algorithms (include/katric_shard.h).

``make_shard`` produces what the reference's MPI exchange delivers to one
rank (local CSR slice, ghost degrees, received messages); ``count_shard``
runs that rank's local and contracted / distributed phase on the GPU.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import _ffi
from ._ffi import lib

CETRIC = 0  # configs/cetric.toml: "cetric-x"
DITRIC = 1  # configs/ditric.toml: "patric"

u32, u64, vp = C.c_uint32, C.c_uint64, C.c_void_p


class KtShardIn(C.Structure):
    _fields_ = [
        ("algorithm", u32), ("rank", u32), ("num_ranks", u32), ("flags", u32),
        ("range_bounds", vp), ("offsets", vp), ("heads", vp),
        ("n_ghost", u32), ("ghost_ids", vp), ("ghost_degree", vp),
        ("n_msg", u32), ("msg_src", vp), ("msg_offsets", vp), ("msg_heads", vp),
    ]


class KtShardOut(C.Structure):
    _fields_ = [("local_triangles", u64), ("type3_triangles", u64), ("per_vertex", vp)]


class KtShardHost(C.Structure):
    _fields_ = [
        ("rank", u32), ("num_ranks", u32), ("algorithm", u32), ("n_local", u32),
        ("offsets", C.POINTER(u64)), ("heads", C.POINTER(u32)),
        ("n_ghost", u32), ("ghost_ids", C.POINTER(u32)), ("ghost_degree", C.POINTER(u64)),
        ("n_msg", u32), ("msg_src", C.POINTER(u32)), ("msg_offsets", C.POINTER(u64)), ("msg_heads", C.POINTER(u32)),
    ]


_ffi.register("kt_count_shard", C.c_int, [vp, C.POINTER(KtShardIn), C.POINTER(KtShardOut)])
_ffi.register("kt_shard_make", C.c_int, [u32, vp, vp, u32, vp, u32, u32, C.POINTER(KtShardHost)])
_ffi.register("kt_shard_host_free", None, [C.POINTER(KtShardHost)])


def even_bounds(n: int, P: int) -> np.ndarray:
    """P contiguous ranges: rank r owns [r*n//P, (r+1)*n//P)."""
    return np.array([r * n // P for r in range(P + 1)], dtype=np.uint64)


@dataclass
class Shard:
    algorithm: int
    rank: int
    bounds: np.ndarray
    offsets: np.ndarray
    heads: np.ndarray
    ghost_ids: np.ndarray
    ghost_degree: np.ndarray
    msg_src: np.ndarray
    msg_offsets: np.ndarray
    msg_heads: np.ndarray

    @property
    def num_ranks(self) -> int:
        return len(self.bounds) - 1

    @property
    def n_local(self) -> int:
        return int(self.bounds[self.rank + 1] - self.bounds[self.rank])

    def vertex_ids(self) -> np.ndarray:
        """Global ids of the per-vertex output entries (owned, then ghosts)."""
        lo = int(self.bounds[self.rank])
        return np.concatenate([np.arange(lo, lo + self.n_local, dtype=np.int64),
                               self.ghost_ids.astype(np.int64)])


def _arr(p, count, dtype):
    if count == 0:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(p, shape=(count,)).copy()


def make_shard(g, num_ranks: int, rank: int, algorithm: int = CETRIC, bounds=None) -> Shard:
    """Shard inputs of `rank` for a symmetric CSR partitioned into contiguous ranges."""
    b = np.ascontiguousarray(bounds if bounds is not None else even_bounds(g.n, num_ranks), dtype=np.uint64)
    h = KtShardHost()
    rc = lib().kt_shard_make(g.n, g.offsets.ctypes.data, g.heads.ctypes.data if g.nnz else None, num_ranks,
                             b.ctypes.data, rank, algorithm, C.byref(h))
    if rc != 0:
        raise ValueError(f"kt_shard_make failed ({rc})")
    try:
        nl = h.n_local
        off = _arr(h.offsets, nl + 1, np.uint64)
        return Shard(
            algorithm=algorithm, rank=rank, bounds=b,
            offsets=off, heads=_arr(h.heads, int(off[-1]), np.uint32),
            ghost_ids=_arr(h.ghost_ids, h.n_ghost, np.uint32), ghost_degree=_arr(h.ghost_degree, h.n_ghost, np.uint64),
            msg_src=_arr(h.msg_src, h.n_msg, np.uint32),
            msg_offsets=_arr(h.msg_offsets, h.n_msg + 1, np.uint64) if h.n_msg else np.zeros(1, np.uint64),
            msg_heads=_arr(h.msg_heads, int(h.msg_offsets[h.n_msg]) if h.n_msg else 0, np.uint32),
        )
    finally:
        lib().kt_shard_host_free(C.byref(h))


def _p(a: np.ndarray) -> Optional[int]:
    return a.ctypes.data if a.size else None


def count_shard(ctx, s: Shard, per_vertex: bool = True):
    """(local_triangles, type3_triangles, per_vertex over s.vertex_ids() or None)."""
    si = KtShardIn(
        algorithm=s.algorithm, rank=s.rank, num_ranks=s.num_ranks, flags=0,
        range_bounds=s.bounds.ctypes.data, offsets=s.offsets.ctypes.data, heads=_p(s.heads),
        n_ghost=len(s.ghost_ids), ghost_ids=_p(s.ghost_ids), ghost_degree=_p(s.ghost_degree),
        n_msg=len(s.msg_src), msg_src=_p(s.msg_src), msg_offsets=s.msg_offsets.ctypes.data,
        msg_heads=_p(s.msg_heads),
    )
    pv = np.zeros(s.n_local + len(s.ghost_ids), dtype=np.uint64) if per_vertex else None
    so = KtShardOut(0, 0, pv.ctypes.data if pv is not None and pv.size else None)
    ctx.check(lib().kt_count_shard(ctx.handle, C.byref(si), C.byref(so)), "kt_count_shard")
    return int(so.local_triangles), int(so.type3_triangles), pv


@dataclass
class DeviceShard:
    """A shard's input arrays resident in device memory (range bounds stay on the host)."""
    shard: Shard
    bufs: dict


def upload_shard(ctx, s: Shard) -> DeviceShard:
    bufs = {}
    for name in ("offsets", "heads", "ghost_ids", "ghost_degree", "msg_src", "msg_offsets", "msg_heads"):
        a = getattr(s, name)
        b = ctx.alloc(max(a.nbytes, 8))
        if a.size:
            b.upload(a)
        bufs[name] = b
    return DeviceShard(s, bufs)


def count_shard_device(ctx, d: DeviceShard, per_vertex_out=None):
    """kt_count_shard on device-resident inputs (KT_DEVICE_INPUT); per-vertex
    counts go to the device buffer `per_vertex_out` when given.  Returns
    (local_triangles, type3_triangles)."""
    from . import KT_DEVICE_INPUT, KT_DEVICE_OUTPUT

    s, b = d.shard, d.bufs
    flags = KT_DEVICE_INPUT | (KT_DEVICE_OUTPUT if per_vertex_out is not None else 0)
    si = KtShardIn(
        algorithm=s.algorithm, rank=s.rank, num_ranks=s.num_ranks, flags=flags,
        range_bounds=s.bounds.ctypes.data, offsets=b["offsets"].ptr, heads=b["heads"].ptr,
        n_ghost=len(s.ghost_ids), ghost_ids=b["ghost_ids"].ptr, ghost_degree=b["ghost_degree"].ptr,
        n_msg=len(s.msg_src), msg_src=b["msg_src"].ptr, msg_offsets=b["msg_offsets"].ptr,
        msg_heads=b["msg_heads"].ptr,
    )
    pv = per_vertex_out.ptr if per_vertex_out is not None else None
    so = KtShardOut(0, 0, pv)
    ctx.check(lib().kt_count_shard(ctx.handle, C.byref(si), C.byref(so)), "kt_count_shard")
    return int(so.local_triangles), int(so.type3_triangles)


ACCUMULATE = 0x10  # KT_SHARD_ACCUMULATE: per-vertex counts are added to the output

_ffi.register("kt_count_shard_graph", C.c_int, [vp, u32, vp, vp, u32, u32, u32, vp, u32, C.POINTER(KtShardOut)])


def count_shard_graph(ctx, g, num_ranks: int, rank: int, algorithm: int = CETRIC, bounds=None,
                      per_vertex: bool = True, per_vertex_out=None, accumulate: bool = False):
    """One rank's counts with its shard built on the device from the whole
    graph `g` (a host Graph, or a DeviceGraph resident in HBM) --
    kt_count_shard_graph.  Per-vertex counts are indexed by global id: into
    the device buffer `per_vertex_out` (n entries) when given, else returned
    as a host array when `per_vertex`; with `accumulate` they are added to the
    output's contents (a host array passed as `per_vertex`, or the device
    buffer).  Returns (local_triangles, type3_triangles, per_vertex or None)."""
    from . import KT_DEVICE_INPUT, KT_DEVICE_OUTPUT, DeviceGraph

    b = np.ascontiguousarray(bounds if bounds is not None else even_bounds(g.n, num_ranks), dtype=np.uint64)
    flags = ACCUMULATE if accumulate else 0
    if isinstance(g, DeviceGraph):
        flags |= KT_DEVICE_INPUT
        off, hd = g.offsets.ptr, g.heads.ptr
    else:
        off, hd = g.offsets.ctypes.data, _p(g.heads)
    pv = None
    if per_vertex_out is not None:
        flags |= KT_DEVICE_OUTPUT
        pv_ptr = per_vertex_out.ptr
    elif per_vertex is not False and per_vertex is not None:
        if isinstance(per_vertex, np.ndarray):
            assert per_vertex.dtype == np.uint64 and per_vertex.size == g.n and per_vertex.flags.c_contiguous
            pv = per_vertex
        else:
            pv = np.zeros(g.n, dtype=np.uint64)
        pv_ptr = pv.ctypes.data if pv.size else None
    else:
        pv_ptr = None
    so = KtShardOut(0, 0, pv_ptr)
    ctx.check(lib().kt_count_shard_graph(ctx.handle, g.n, off, hd, flags, algorithm, num_ranks, b.ctypes.data, rank,
                                         C.byref(so)), "kt_count_shard_graph")
    return int(so.local_triangles), int(so.type3_triangles), pv
