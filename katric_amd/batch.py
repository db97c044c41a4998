"""Batched counting of many independent graphs (kt_batch_count).

A batch is one disjoint-union CSR: graph g owns the global ids
[vertex_base[g], vertex_base[g+1]).  ``gen_batch`` draws the mixed G(n,m) /
R-MAT batch, ``from_graphs`` packs given graphs, ``count_batch`` runs the
hipGraph-captured batch pipeline.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import List, Optional

import numpy as np

from . import _ffi, Graph, DeviceBuffer, KT_DEVICE_INPUT, KT_DEVICE_OUTPUT
from ._ffi import lib

u32, u64, vp = C.c_uint32, C.c_uint64, C.c_void_p


class KtHostBatch(C.Structure):
    _fields_ = [("num_graphs", u32), ("n_total", u64), ("nnz_total", u64), ("vertex_base", C.POINTER(u64)),
                ("offsets", C.POINTER(u64)), ("heads", C.POINTER(u32)), ("kind", C.POINTER(C.c_uint8))]


_ffi.register("kt_gen_batch", C.c_int, [u32, u64, u32, u32, C.c_double, C.c_double, C.c_double, C.c_double,
                                         C.c_double, C.c_double, C.POINTER(KtHostBatch)])
_ffi.register("kt_host_batch_free", None, [C.POINTER(KtHostBatch)])
_ffi.register("kt_batch_count", C.c_int, [vp, u32, vp, vp, vp, u32, vp, vp])
_ffi.register("kt_batch_replays", u64, [vp])


@dataclass
class Batch:
    vertex_base: np.ndarray  # num_graphs + 1 (uint64)
    offsets: np.ndarray      # n_total + 1 (uint64, global)
    heads: np.ndarray        # nnz_total (uint32, global ids)
    kind: Optional[np.ndarray] = None

    @property
    def num_graphs(self) -> int:
        return len(self.vertex_base) - 1

    @property
    def n_total(self) -> int:
        return int(self.vertex_base[-1])

    @property
    def nnz(self) -> int:
        return int(self.offsets[-1])

    def graph(self, g: int) -> Graph:
        """Graph g with local ids."""
        a, b = int(self.vertex_base[g]), int(self.vertex_base[g + 1])
        e0, e1 = int(self.offsets[a]), int(self.offsets[b])
        return Graph(b - a, self.offsets[a:b + 1] - np.uint64(e0), (self.heads[e0:e1] - np.uint32(a)))

    def edges_per_graph(self) -> np.ndarray:
        return (self.offsets[self.vertex_base[1:]] - self.offsets[self.vertex_base[:-1]]) // 2


def gen_batch(num_graphs: int, seed: int = 1, n_min: int = 64, n_max: int = 65536, deg_min: float = 4.0,
              deg_max: float = 64.0, rmat_fraction: float = 0.5, a=0.57, b=0.19, c=0.19) -> Batch:
    hb = KtHostBatch()
    rc = lib().kt_gen_batch(num_graphs, seed, n_min, n_max, deg_min, deg_max, rmat_fraction, a, b, c, C.byref(hb))
    if rc != 0:
        raise ValueError(f"kt_gen_batch failed ({rc})")
    try:
        G, N, E = hb.num_graphs, hb.n_total, hb.nnz_total
        vb = np.ctypeslib.as_array(hb.vertex_base, shape=(G + 1,)).copy()
        off = np.ctypeslib.as_array(hb.offsets, shape=(N + 1,)).copy()
        heads = np.ctypeslib.as_array(hb.heads, shape=(max(E, 1),))[:E].copy()
        kind = np.ctypeslib.as_array(hb.kind, shape=(G,)).copy()
        return Batch(vb, off, heads, kind)
    finally:
        lib().kt_host_batch_free(C.byref(hb))


def from_graphs(graphs: List[Graph]) -> Batch:
    vb = np.zeros(len(graphs) + 1, np.uint64)
    vb[1:] = np.cumsum([g.n for g in graphs])
    eb = np.zeros(len(graphs) + 1, np.uint64)
    eb[1:] = np.cumsum([g.nnz for g in graphs])
    off = np.concatenate([g.offsets[:-1] + eb[i] for i, g in enumerate(graphs)] + [eb[-1:]]).astype(np.uint64)
    heads = np.concatenate([g.heads.astype(np.uint64) + vb[i] for i, g in enumerate(graphs)] +
                           [np.zeros(0, np.uint64)]).astype(np.uint32)
    return Batch(vb, off, heads)


@dataclass
class DeviceBatch:
    batch: Batch
    offsets: DeviceBuffer
    heads: DeviceBuffer


def upload_batch(ctx, b: Batch) -> DeviceBatch:
    off = DeviceBuffer(ctx, b.offsets.nbytes)
    off.upload(b.offsets)
    hd = DeviceBuffer(ctx, max(b.heads.nbytes, 4))
    if b.heads.size:
        hd.upload(b.heads)
    return DeviceBatch(b, off, hd)


def count_batch(ctx, b, per_vertex: bool = False, per_vertex_out: Optional[DeviceBuffer] = None):
    """Per-graph totals (uint64[num_graphs]) and, with per_vertex=True, the
    per-vertex counts over the union.  `b` is a Batch (host) or DeviceBatch."""
    host = b.batch if isinstance(b, DeviceBatch) else b
    vb = np.ascontiguousarray(host.vertex_base, dtype=np.uint64)
    totals = np.zeros(host.num_graphs, np.uint64)
    pv = np.zeros(host.n_total, np.uint64) if per_vertex and per_vertex_out is None else None
    flags = 0
    if isinstance(b, DeviceBatch):
        flags |= KT_DEVICE_INPUT
        off_p, hd_p = b.offsets.ptr, b.heads.ptr
    else:
        off_p, hd_p = host.offsets.ctypes.data, (host.heads.ctypes.data if host.heads.size else None)
    pv_p = None
    if per_vertex_out is not None:
        flags |= KT_DEVICE_OUTPUT
        pv_p = per_vertex_out.ptr
    elif pv is not None:
        pv_p = pv.ctypes.data
    rc = lib().kt_batch_count(ctx.handle, host.num_graphs, vb.ctypes.data, off_p, hd_p, flags, totals.ctypes.data,
                              pv_p)
    ctx.check(rc, "kt_batch_count")
    return (totals, pv) if per_vertex else totals


def replays(ctx) -> int:
    return int(lib().kt_batch_replays(ctx.handle))
