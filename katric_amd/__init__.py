"""katric_amd: MI355X-native per-rank triangle-counting core.

Python host layer over the C ABI (include/katric.h).  The device work runs in
hand-written HIP kernels (katric_amd/csrc); this module only moves buffers
and mirrors the reference's entry points:

  * ``run_shmem(G)``      -- cetric::run_shmem
                             (include/cetric/counters/shared_memory_cetric.h:105)
  * ``Context.count``     -- the same, with per-vertex counts
  * ``Context.orient``    -- preprocessing (DistributedGraph::orient +
                             sort_neighborhoods, distributed_graph.h:537-593)

Graphs are ``Graph`` objects: a symmetric CSR with uint64 offsets and uint32
heads held in numpy arrays (host) or torch tensors (device).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional, Tuple

import numpy as np

from . import _ffi
from ._ffi import KtHostGraph, KtParams, KtStats, KtTiming, lib

__all__ = [
    "KatricError",
    "Graph",
    "DeviceBuffer",
    "DeviceGraph",
    "Context",
    "run_shmem",
    "gen_gnm",
    "gen_rmat",
    "read_metis",
    "write_metis",
    "read_graphio_binary",
    "write_graphio_binary",
    "KT_DEVICE_INPUT",
    "KT_DEVICE_OUTPUT",
]

STRATEGY_STREAMED = 0  # KT_STRATEGY_STREAMED
STRATEGY_GROUPED16 = 1  # KT_STRATEGY_GROUPED16
STRATEGY_GROUPED8 = 2  # KT_STRATEGY_GROUPED8
STRATEGY_GROUPED32 = 3  # KT_STRATEGY_GROUPED32
STRATEGY_GROUPED4 = 4  # KT_STRATEGY_GROUPED4
STRATEGY_GROUPED2 = 5  # KT_STRATEGY_GROUPED2
ORDER_DEGREE = 0  # KT_ORDER_DEGREE
ORDER_VERTEX = 1  # KT_ORDER_VERTEX
KT_DEVICE_INPUT = 0x1
KT_DEVICE_OUTPUT = 0x2

KT_OK = 0
KT_EINVAL = -1
KT_ENOMEM = -2
KT_EHIP = -3
KT_ENODEV = -4
KT_ENOTREADY = -5
KT_ECANCELED = -6
KT_ENOTFOUND = -7
KT_ESTATE = -8
KT_ERANGE = -9


class KatricError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"katric status {status}: {message}")
        self.status = status


def _ptr(a) -> Optional[int]:
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        assert a.flags["C_CONTIGUOUS"]
        return a.ctypes.data
    # torch tensor (device or host)
    return a.data_ptr()


@dataclass
class Graph:
    """Symmetric CSR: offsets (n+1, uint64), heads (offsets[n], uint32)."""

    n: int
    offsets: np.ndarray
    heads: np.ndarray

    def __post_init__(self):
        self.offsets = np.ascontiguousarray(self.offsets, dtype=np.uint64)
        self.heads = np.ascontiguousarray(self.heads, dtype=np.uint32)
        if self.offsets.shape != (self.n + 1,):
            raise ValueError("offsets must have n+1 entries")

    @property
    def nnz(self) -> int:
        return int(self.offsets[-1]) if self.n else 0

    @property
    def m(self) -> int:
        return self.nnz // 2

    def degrees(self) -> np.ndarray:
        return np.diff(self.offsets).astype(np.uint64)

    def checksum(self) -> int:
        return int(lib().kt_csr_checksum(self.n, _ptr(self.offsets), _ptr(self.heads)))

    @staticmethod
    def from_edges(n: int, edges) -> "Graph":
        """Canonical symmetric CSR from an iterable of undirected (u, v)."""
        e = np.asarray(list(edges) if not isinstance(edges, np.ndarray) else edges, dtype=np.int64).reshape(-1, 2)
        e = e[e[:, 0] != e[:, 1]]
        both = np.concatenate([e, e[:, ::-1]]) if len(e) else e
        keys = np.unique(both[:, 0].astype(np.uint64) << np.uint64(32) | both[:, 1].astype(np.uint64))
        tails = (keys >> np.uint64(32)).astype(np.int64)
        heads = (keys & np.uint64(0xFFFFFFFF)).astype(np.uint32)
        off = np.zeros(n + 1, dtype=np.uint64)
        np.add.at(off, tails + 1, 1)
        return Graph(n, np.cumsum(off, dtype=np.uint64), heads)


def _take_host_graph(hg: KtHostGraph) -> Graph:
    try:
        n = hg.n
        off = np.ctypeslib.as_array(hg.offsets, shape=(n + 1,)).copy()
        heads = np.ctypeslib.as_array(hg.heads, shape=(max(hg.nnz, 1),))[: hg.nnz].copy()
        return Graph(n, off, heads)
    finally:
        lib().kt_host_graph_free(C.byref(hg))


def _gio(rc: int, what: str):
    if rc != 0:
        raise KatricError(rc, what)


def gen_gnm(n: int, samples: int, seed: int = 1) -> Graph:
    """G(n, m) with `samples` uniform endpoint pairs (self-loops/duplicates removed)."""
    hg = KtHostGraph()
    _gio(lib().kt_gen_gnm(n, samples, seed, C.byref(hg)), "kt_gen_gnm")
    return _take_host_graph(hg)


def gen_rmat(scale: int, edge_factor: int = 16, a=0.57, b=0.19, c=0.19, seed: int = 1, n: int = 0,
             samples: Optional[int] = None, permute: bool = True) -> Graph:
    """R-MAT graph: edge_factor * 2^scale samples unless `samples` is given."""
    hg = KtHostGraph()
    s = samples if samples is not None else edge_factor * (1 << scale)
    _gio(lib().kt_gen_rmat(scale, s, a, b, c, seed, n, 1 if permute else 0, C.byref(hg)), "kt_gen_rmat")
    return _take_host_graph(hg)


def read_metis(path: str) -> Graph:
    hg = KtHostGraph()
    _gio(lib().kt_read_metis(path.encode(), C.byref(hg)), f"kt_read_metis({path})")
    return _take_host_graph(hg)


def write_metis(path: str, g: Graph):
    _gio(lib().kt_write_metis(path.encode(), g.n, _ptr(g.offsets), _ptr(g.heads)), "kt_write_metis")


def read_graphio_binary(first_out: str, head: str) -> Graph:
    hg = KtHostGraph()
    _gio(lib().kt_read_graphio_binary(first_out.encode(), head.encode(), C.byref(hg)), "kt_read_graphio_binary")
    return _take_host_graph(hg)


def write_graphio_binary(first_out: str, head: str, g: Graph):
    _gio(lib().kt_write_graphio_binary(first_out.encode(), head.encode(), g.n, _ptr(g.offsets), _ptr(g.heads)),
         "kt_write_graphio_binary")


class DeviceBuffer:
    """Device allocation made through kt_buffer_alloc (freed with the context)."""

    def __init__(self, ctx: "Context", nbytes: int):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        ctx.check(lib().kt_buffer_alloc(ctx.handle, self.nbytes, C.byref(p)), "kt_buffer_alloc")
        self.ptr = int(p.value)

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        self.ctx.check(lib().kt_buffer_upload(self.ctx.handle, self.ptr, a.ctypes.data, a.nbytes), "upload")

    def download(self, dtype, count: int) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        assert out.nbytes <= self.nbytes
        self.ctx.check(lib().kt_buffer_download(self.ctx.handle, out.ctypes.data, self.ptr, out.nbytes),
                       "download")
        return out

    def free(self):
        if self.ptr and self.ctx.handle:
            self.ctx.check(lib().kt_buffer_free(self.ctx.handle, self.ptr), "kt_buffer_free")
        self.ptr = 0


@dataclass
class DeviceGraph:
    """A symmetric CSR resident in HBM."""

    n: int
    nnz: int
    offsets: DeviceBuffer
    heads: DeviceBuffer

    @property
    def m(self) -> int:
        return self.nnz // 2

    def free(self):
        self.offsets.free()
        self.heads.free()


class Context:
    """One kt_ctx: a device, a stream and a device workspace."""

    def __init__(self, device: int = 0, hbm_budget_bytes: int = 0):
        self._h = C.c_void_p()
        rc = lib().kt_create(C.byref(self._h), device, hbm_budget_bytes)
        if rc != 0:
            raise KatricError(rc, lib().kt_strerror(rc).decode())

    def close(self):
        if self._h:
            lib().kt_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def handle(self):
        return self._h

    def check(self, rc: int, what: str):
        if rc != 0:
            msg = lib().kt_last_error(self._h).decode() or lib().kt_strerror(rc).decode()
            raise KatricError(rc, f"{what}: {msg}")

    # -- knobs / telemetry ------------------------------------------------
    def params(self) -> dict:
        p = KtParams()
        self.check(lib().kt_get_params(self._h, C.byref(p)), "kt_get_params")
        return {f: getattr(p, f) for f, _ in KtParams._fields_}

    def set_params(self, wave_max_deg=None, group_max_deg=None, grid_per_cu=None, strategy=None, order=None,
                   bsearch_min_deg=None, pseudo2core=None, min_side=None, **extra):
        """Intersection knobs (see kt_params); unknown keys are rejected."""
        if extra:
            raise TypeError(f"unknown parameters: {sorted(extra)}")
        cur = self.params()
        new = dict(wave_max_deg=wave_max_deg, group_max_deg=group_max_deg, grid_per_cu=grid_per_cu,
                   strategy=strategy, order=order, bsearch_min_deg=bsearch_min_deg, pseudo2core=pseudo2core,
                   min_side=min_side)
        cur.update({k: int(v) for k, v in new.items() if v is not None})
        p = KtParams(**cur)
        self.check(lib().kt_set_params(self._h, C.byref(p)), "kt_set_params")

    def stats(self) -> dict:
        s = KtStats()
        self.check(lib().kt_last_stats(self._h, C.byref(s)), "kt_last_stats")
        return {f: getattr(s, f) for f, _ in KtStats._fields_}

    def timing(self) -> dict:
        t = KtTiming()
        self.check(lib().kt_last_timing(self._h, C.byref(t)), "kt_last_timing")
        return {f: getattr(t, f) for f, _ in KtTiming._fields_}

    def device_bytes(self) -> int:
        return int(lib().kt_device_bytes(self._h))

    def synchronize(self):
        self.check(lib().kt_synchronize(self._h), "kt_synchronize")

    def alloc(self, nbytes: int) -> DeviceBuffer:
        return DeviceBuffer(self, nbytes)

    def upload(self, g: Graph) -> DeviceGraph:
        off = DeviceBuffer(self, (g.n + 1) * 8)
        off.upload(g.offsets)
        hd = DeviceBuffer(self, max(g.nnz, 1) * 4)
        if g.nnz:
            hd.upload(g.heads)
        return DeviceGraph(g.n, g.nnz, off, hd)

    # -- counting -----------------------------------------------------------
    def count(self, g: Graph, per_vertex: bool = True) -> Tuple[int, Optional[np.ndarray]]:
        """Total and per-vertex triangle counts of a host CSR."""
        total = C.c_uint64(0)
        pv = np.zeros(g.n, dtype=np.uint64) if per_vertex else None
        rc = lib().kt_count(self._h, g.n, _ptr(g.offsets), _ptr(g.heads), 0, C.byref(total), _ptr(pv))
        self.check(rc, "kt_count")
        return int(total.value), pv

    def count_device(self, g, offsets=None, heads=None, per_vertex_out=None) -> int:
        """Counts a CSR resident in device memory: `g` is a DeviceGraph, or
        the vertex count with `offsets` / `heads` given as device pointers
        (ints, DeviceBuffers or torch tensors).  Per-vertex counts are written
        to `per_vertex_out` (device) when given."""
        total = C.c_uint64(0)
        flags = KT_DEVICE_INPUT | (KT_DEVICE_OUTPUT if per_vertex_out is not None else 0)

        def dptr(x):
            if x is None or isinstance(x, int):
                return x
            if isinstance(x, DeviceBuffer):
                return x.ptr
            return x.data_ptr()

        if isinstance(g, DeviceGraph):
            n, off_p, hd_p = g.n, g.offsets.ptr, g.heads.ptr
        else:
            n, off_p, hd_p = int(g), dptr(offsets), dptr(heads)
        pv_p = dptr(per_vertex_out)
        rc = lib().kt_count(self._h, n, off_p, hd_p, flags, C.byref(total), pv_p)
        self.check(rc, "kt_count")
        return int(total.value)

    def list_triangles(self, g: Graph, capacity: Optional[int] = None) -> np.ndarray:
        """Every triangle of a host CSR once, as a (T, 3) uint32 array of
        (v, u, w): v the tail (lowest in the (degree, id) order), u in N+(v),
        w in N+(v) and N+(u) -- the triples the reference's emit callback
        receives.  Without `capacity` a first call sizes the buffer."""
        total = C.c_uint64(0)
        if capacity is None:
            rc = lib().kt_list_triangles(self._h, g.n, _ptr(g.offsets), _ptr(g.heads), 0, None, 0, C.byref(total))
            if rc not in (0, KT_ERANGE):
                self.check(rc, "kt_list_triangles")
            capacity = int(total.value)
        out = np.zeros((max(capacity, 1), 3), dtype=np.uint32)
        rc = lib().kt_list_triangles(self._h, g.n, _ptr(g.offsets), _ptr(g.heads), 0, out.ctypes.data, capacity,
                                     C.byref(total))
        self.check(rc, "kt_list_triangles")
        return out[: int(total.value)]

    def intersect(self, pairs, with_matches: bool = True):
        """Intersections of (a, b) pairs of strictly increasing id lists on the
        device (kt_intersect).  Returns (counts, matches): counts[i] =
        |a_i cap b_i|, matches a list of ascending uint32 arrays (None without
        with_matches)."""
        pairs = list(pairs)
        P = len(pairs)
        if P == 0:
            return np.zeros(0, np.uint64), ([] if with_matches else None)
        la = [np.ascontiguousarray(a, dtype=np.uint32) for a, _ in pairs]
        lb = [np.ascontiguousarray(b, dtype=np.uint32) for _, b in pairs]
        aoff = np.zeros(P + 1, np.uint64)
        boff = np.zeros(P + 1, np.uint64)
        aoff[1:] = np.cumsum([len(x) for x in la])
        boff[1:] = np.cumsum([len(x) for x in lb])
        a = np.concatenate(la) if aoff[-1] else np.zeros(1, np.uint32)
        b = np.concatenate(lb) if boff[-1] else np.zeros(1, np.uint32)
        counts = np.zeros(P, np.uint64)
        total = C.c_uint64(0)
        cap = int(min(aoff[-1], boff[-1]))
        out = np.zeros(max(cap, 1), np.uint32) if with_matches else None
        rc = lib().kt_intersect(self._h, P, _ptr(aoff), _ptr(a), _ptr(boff), _ptr(b), 0, _ptr(counts), _ptr(out),
                                cap, C.byref(total))
        self.check(rc, "kt_intersect")
        if not with_matches:
            return counts, None
        pos = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
        return counts, [out[pos[i]:pos[i + 1]].copy() for i in range(P)]

    def orient(self, g: Graph) -> Graph:
        """Oriented (out-only) CSR under the (degree, id) order, ascending lists."""
        off = np.zeros(g.n + 1, dtype=np.uint64)
        heads = np.zeros(max(g.nnz // 2, 1), dtype=np.uint32)
        rc = lib().kt_orient(self._h, g.n, _ptr(g.offsets), _ptr(g.heads), 0, _ptr(off), _ptr(heads))
        self.check(rc, "kt_orient")
        return Graph(g.n, off, heads[: g.nnz // 2])


_default_ctx: Optional[Context] = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


def run_shmem(g: Graph, ctx: Optional[Context] = None) -> int:
    """Triangle count of a symmetric CSR; mirrors cetric::run_shmem
    (include/cetric/counters/shared_memory_cetric.h:105-170), which returns the
    number of emitted triangles."""
    total, _ = (ctx or default_context()).count(g, per_vertex=False)
    return total


# sub-modules register the signatures of their entry points on import
from . import batch, serve, shard  # noqa: E402,F401
