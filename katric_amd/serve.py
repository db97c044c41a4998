"""Serving loop with continuous batching (include/katric_serve.h)."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Dict, Optional

import numpy as np

from . import _ffi, Graph, KatricError
from ._ffi import lib

u32, u64, vp = C.c_uint32, C.c_uint64, C.c_void_p


class KtServeConfig(C.Structure):
    _fields_ = [("hbm_budget_bytes", u64), ("max_batch_graphs", u32), ("large_request_edges", u32),
                ("max_batch_edges", u64)]


class KtServeResult(C.Structure):
    _fields_ = [("ticket", u64), ("status", C.c_int32), ("batch_graphs", u32), ("total", u64), ("batch_id", u64),
                ("queue_ms", C.c_double), ("latency_ms", C.c_double)]


class KtServeStats(C.Structure):
    _fields_ = [(f, u64) for f in ("submitted", "batches", "graphs", "edges", "isolation_reruns", "pending")]


_ffi.register("kt_serve_start", C.c_int, [vp, C.POINTER(KtServeConfig)])
_ffi.register("kt_serve_stop", C.c_int, [vp])
_ffi.register("kt_serve_submit", C.c_int, [vp, u32, vp, vp, vp, C.POINTER(u64)])
_ffi.register("kt_serve_poll", C.c_int, [vp, u64, C.POINTER(KtServeResult)])
_ffi.register("kt_serve_wait", C.c_int, [vp, u64, C.c_double, C.POINTER(KtServeResult)])
_ffi.register("kt_serve_cancel", C.c_int, [vp, u64])
_ffi.register("kt_serve_get_stats", C.c_int, [vp, C.POINTER(KtServeStats)])
_ffi.register("kt_host_alloc", C.c_int, [u64, C.POINTER(vp)])
_ffi.register("kt_host_free", C.c_int, [vp])

KT_ENOTREADY = -5


@dataclass
class Result:
    ticket: int
    status: int
    total: int
    batch_id: int
    batch_graphs: int
    queue_ms: float
    latency_ms: float


class PinnedArray:
    """numpy view of page-locked host memory (kt_host_alloc)."""

    def __init__(self, count: int, dtype):
        self.dtype = np.dtype(dtype)
        self.count = int(count)
        p = C.c_void_p()
        rc = lib().kt_host_alloc(max(self.count * self.dtype.itemsize, 1), C.byref(p))
        if rc != 0:
            raise KatricError(rc, "kt_host_alloc")
        self.ptr = p.value
        buf = (C.c_char * max(self.count * self.dtype.itemsize, 1)).from_address(self.ptr)
        self.array = np.frombuffer(buf, dtype=self.dtype, count=self.count)

    def free(self):
        if self.ptr:
            lib().kt_host_free(self.ptr)
            self.ptr = None
            self.array = None


def pinned_graph(g: Graph):
    """Copy of g in page-locked memory: (Graph view, [PinnedArray owners])."""
    po = PinnedArray(g.n + 1, np.uint64)
    po.array[:] = g.offsets
    ph = PinnedArray(max(g.nnz, 1), np.uint32)
    ph.array[: g.nnz] = g.heads
    view = Graph.__new__(Graph)
    view.n, view.offsets, view.heads = g.n, po.array, ph.array[: g.nnz]
    return view, [po, ph]


class Server:
    """Continuous-batching server on a Context's device."""

    def __init__(self, ctx, hbm_budget_bytes: int = 0, max_batch_graphs: int = 0, max_batch_edges: int = 0,
                 large_request_edges: int = 0):
        self.ctx = ctx
        cfg = KtServeConfig(hbm_budget_bytes, max_batch_graphs, large_request_edges, max_batch_edges)
        ctx.check(lib().kt_serve_start(ctx.handle, C.byref(cfg)), "kt_serve_start")
        self._keep: Dict[int, tuple] = {}

    def submit(self, g: Graph, per_vertex: Optional[np.ndarray] = None) -> int:
        t = C.c_uint64(0)
        rc = lib().kt_serve_submit(self.ctx.handle, g.n, g.offsets.ctypes.data,
                                   g.heads.ctypes.data if g.nnz else None,
                                   per_vertex.ctypes.data if per_vertex is not None else None, C.byref(t))
        if rc != 0:
            raise KatricError(rc, "kt_serve_submit")
        self._keep[t.value] = (g, per_vertex)  # buffers stay alive until the result is collected
        return t.value

    def _result(self, r: KtServeResult) -> Result:
        self._keep.pop(r.ticket, None)
        return Result(r.ticket, r.status, r.total, r.batch_id, r.batch_graphs, r.queue_ms, r.latency_ms)

    def poll(self, ticket: int) -> Optional[Result]:
        r = KtServeResult()
        rc = lib().kt_serve_poll(self.ctx.handle, ticket, C.byref(r))
        if rc == KT_ENOTREADY:
            return None
        if rc != 0:
            raise KatricError(rc, "kt_serve_poll")
        return self._result(r)

    def wait(self, ticket: int, timeout_ms: float = -1.0) -> Optional[Result]:
        r = KtServeResult()
        rc = lib().kt_serve_wait(self.ctx.handle, ticket, timeout_ms, C.byref(r))
        if rc == KT_ENOTREADY:
            return None
        if rc != 0:
            raise KatricError(rc, "kt_serve_wait")
        return self._result(r)

    def cancel(self, ticket: int) -> int:
        return int(lib().kt_serve_cancel(self.ctx.handle, ticket))

    def stats(self) -> dict:
        s = KtServeStats()
        self.ctx.check(lib().kt_serve_get_stats(self.ctx.handle, C.byref(s)), "kt_serve_get_stats")
        return {f: getattr(s, f) for f, _ in KtServeStats._fields_}

    def stop(self):
        self.ctx.check(lib().kt_serve_stop(self.ctx.handle), "kt_serve_stop")
