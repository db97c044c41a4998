"""ctypes wrapper of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker or the timed CPU baseline; the
product package never imports it.  See oracle.c for what each function
restates from the reference and how it is pinned.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

MERGE, BINARY_SEARCH, BINARY, HYBRID = 0, 1, 2, 3
CETRIC, DITRIC = 0, 1


class OrStats(C.Structure):
    _fields_ = [(f, C.c_uint64) for f in ("wedge_checks", "intersection_size_local", "intersection_size_global",
                                           "local_triangles", "type3_triangles", "skipped_nodes")]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        h = C.CDLL(LIB_PATH)
        vp, u32, u64, i32, sz = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int, C.c_size_t
        sig = {
            "or_log2": (u64, [u64]),
            "or_merge_intersection": (sz, [vp, sz, vp, sz, vp]),
            "or_binary_search_intersection": (sz, [vp, sz, vp, sz, vp]),
            "or_binary_intersection": (sz, [vp, sz, vp, sz, vp, sz]),
            "or_intersection": (sz, [vp, sz, vp, sz, vp, i32, sz, C.c_double]),
            "or_orient": (i32, [u32, vp, vp, vp, vp]),
            "or_run_shmem": (i32, [u32, vp, vp, i32, i32, C.POINTER(u64), vp, C.POINTER(OrStats)]),
            "or_run_shmem_ex": (i32, [u32, vp, vp, i32, i32, i32, C.POINTER(u64), vp, C.POINTER(OrStats)]),
            "or_brute_force": (i32, [u32, vp, vp, C.POINTER(u64), vp]),
            "or_run_shmem_sampled": (i32, [u32, vp, vp, u32, C.POINTER(u64), C.POINTER(u64),
                                           C.POINTER(C.c_double), C.POINTER(C.c_double)]),
            "or_owner": (i32, [u64, vp, i32]),
            "or_rank_sim": (i32, [u32, vp, vp, i32, vp, i32, i32, i32, C.POINTER(u64), C.POINTER(u64), vp,
                                  C.POINTER(OrStats)]),
            "or_rank_closed_form": (i32, [u32, vp, vp, i32, vp, i32, vp, vp]),
        }
        for k, (r, a) in sig.items():
            f = getattr(h, k)
            f.restype = r
            f.argtypes = a
        _lib = h
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


def _u32(x):
    return np.ascontiguousarray(x, dtype=np.uint32)


def _u64(x):
    return np.ascontiguousarray(x, dtype=np.uint64)


def intersection(a, b, method=MERGE, cutoff=1000, hybrid_scale=1.0):
    a, b = _u32(a), _u32(b)
    out = np.zeros(max(1, min(len(a), len(b))), dtype=np.uint32)
    pa, pb = _p(a) if len(a) else None, _p(b) if len(b) else None
    if method == "merge_only":
        k = lib().or_merge_intersection(pa, len(a), pb, len(b), _p(out))
    elif method == "binary_search_only":
        k = lib().or_binary_search_intersection(pa, len(a), pb, len(b), _p(out))
    elif method == "binary_only":
        k = lib().or_binary_intersection(pa, len(a), pb, len(b), _p(out), cutoff)
    else:
        k = lib().or_intersection(pa, len(a), pb, len(b), _p(out), method, cutoff, hybrid_scale)
    return out[:k].tolist()


def log2(x):
    return int(lib().or_log2(x))


def orient(g):
    off = np.zeros(g.n + 1, dtype=np.uint64)
    heads = np.zeros(max(g.nnz // 2, 1), dtype=np.uint32)
    rc = lib().or_orient(g.n, _p(_u64(g.offsets)), _p(_u32(g.heads)) if g.nnz else None, _p(off), _p(heads))
    if rc != 0:
        raise ValueError("not a symmetric CSR")
    return off, heads[: g.nnz // 2]


def run_shmem(g, method=MERGE, threads=1, per_vertex=True, pseudo2core=False):
    """(total, per_vertex or None, stats dict)."""
    tot = C.c_uint64(0)
    pv = np.zeros(g.n, dtype=np.uint64) if per_vertex else None
    st = OrStats()
    off, hd = _u64(g.offsets), _u32(g.heads)
    lib().or_run_shmem_ex(g.n, _p(off), _p(hd) if g.nnz else None, method, threads, 1 if pseudo2core else 0,
                          C.byref(tot), _p(pv), C.byref(st))
    return int(tot.value), pv, {f: getattr(st, f) for f, _ in OrStats._fields_}


def run_shmem_sample(g, stride=16, per_vertex=False):
    """One-thread timing sample: (sample triangles, None, {prep_s, local_s,
    sample_edges}) -- preprocessing of the whole graph, local phase over the
    tails v % stride == 0."""
    tri, edges = C.c_uint64(0), C.c_uint64(0)
    ps, ls = C.c_double(0), C.c_double(0)
    off, hd = _u64(g.offsets), _u32(g.heads)
    rc = lib().or_run_shmem_sampled(g.n, _p(off), _p(hd) if g.nnz else None, stride, C.byref(tri), C.byref(edges),
                                    C.byref(ps), C.byref(ls))
    if rc != 0:
        raise ValueError("bad stride")
    return int(tri.value), None, {"prep_s": ps.value, "local_s": ls.value, "sample_edges": int(edges.value)}


def brute_force(g):
    tot = C.c_uint64(0)
    pv = np.zeros(g.n, dtype=np.uint64)
    off, hd = _u64(g.offsets), _u32(g.heads)
    rc = lib().or_brute_force(g.n, _p(off), _p(hd) if g.nnz else None, C.byref(tot), _p(pv))
    if rc != 0:
        raise ValueError("graph too large for brute force")
    return int(tot.value), pv


def even_bounds(n, P):
    """Contiguous ranges: rank r owns [r*n//P, (r+1)*n//P)."""
    return np.array([r * n // P for r in range(P + 1)], dtype=np.uint64)


def owner(x, bounds):
    b = _u64(bounds)
    return int(lib().or_owner(x, _p(b), len(b) - 1))


def rank_sim(g, P, rank, algorithm=CETRIC, bounds=None, threads=1, per_vertex=True):
    """(local, global/type3, per_vertex[n] or None, stats) of one rank."""
    b = _u64(bounds if bounds is not None else even_bounds(g.n, P))
    loc, glo = C.c_uint64(0), C.c_uint64(0)
    pv = np.zeros(g.n, dtype=np.uint64) if per_vertex else None
    st = OrStats()
    off, hd = _u64(g.offsets), _u32(g.heads)
    rc = lib().or_rank_sim(g.n, _p(off), _p(hd) if g.nnz else None, P, _p(b), rank, algorithm, threads,
                           C.byref(loc), C.byref(glo), _p(pv), C.byref(st))
    if rc != 0:
        raise ValueError("bad rank simulation arguments")
    return int(loc.value), int(glo.value), pv, {f: getattr(st, f) for f, _ in OrStats._fields_}


def rank_closed_form(g, P, algorithm=CETRIC, bounds=None):
    b = _u64(bounds if bounds is not None else even_bounds(g.n, P))
    loc = np.zeros(P, dtype=np.uint64)
    t3 = np.zeros(P, dtype=np.uint64)
    off, hd = _u64(g.offsets), _u32(g.heads)
    rc = lib().or_rank_closed_form(g.n, _p(off), _p(hd) if g.nnz else None, P, _p(b), algorithm, _p(loc), _p(t3))
    if rc != 0:
        raise ValueError("not a symmetric CSR")
    return loc, t3
