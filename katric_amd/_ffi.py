"""ctypes declarations of the katric_amd C ABI (include/katric.h,
include/katric_graphio.h).

Loading fails loudly when the built library is missing: there is no CPU
fallback for any entry point.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libkatric_amd.so")

u32 = C.c_uint32
u64 = C.c_uint64
i32 = C.c_int
p_u64 = C.POINTER(C.c_uint64)
p_u32 = C.POINTER(C.c_uint32)
vp = C.c_void_p


class KtStats(C.Structure):
    _fields_ = [
        ("wedge_checks", u64),
        ("intersection_size_local", u64),
        ("intersection_size_global", u64),
        ("local_triangles", u64),
        ("type3_triangles", u64),
        ("oriented_edges", u64),
        ("vertices", u64),
        ("input_slots", u64),
        ("skipped_nodes", u64),
        ("streamed_ids", u64),
    ]


class KtTiming(C.Structure):
    _fields_ = [
        ("total_ms", C.c_float),
        ("h2d_ms", C.c_float),
        ("prep_ms", C.c_float),
        ("count_ms", C.c_float),
        ("intersect_ms", C.c_float),
        ("d2h_ms", C.c_float),
    ]


class KtParams(C.Structure):
    _fields_ = [
        ("wave_max_deg", u32),
        ("group_max_deg", u32),
        ("grid_per_cu", u32),
        ("strategy", u32),
        ("order", u32),
        ("bsearch_min_deg", u32),
        ("pseudo2core", u32),
        ("min_side", u32),
    ]


class KtHostGraph(C.Structure):
    _fields_ = [("n", u32), ("nnz", u64), ("offsets", p_u64), ("heads", p_u32)]


# name -> (restype, argtypes); every symbol declared in include/*.h
SIGNATURES = {
    # katric.h
    "kt_create": (i32, [C.POINTER(vp), i32, u64]),
    "kt_destroy": (None, [vp]),
    "kt_strerror": (C.c_char_p, [i32]),
    "kt_last_error": (C.c_char_p, [vp]),
    "kt_set_params": (i32, [vp, C.POINTER(KtParams)]),
    "kt_get_params": (i32, [vp, C.POINTER(KtParams)]),
    "kt_last_stats": (i32, [vp, C.POINTER(KtStats)]),
    "kt_last_timing": (i32, [vp, C.POINTER(KtTiming)]),
    "kt_device_bytes": (u64, [vp]),
    "kt_buffer_alloc": (i32, [vp, u64, C.POINTER(vp)]),
    "kt_buffer_free": (i32, [vp, vp]),
    "kt_buffer_upload": (i32, [vp, vp, vp, u64]),
    "kt_buffer_download": (i32, [vp, vp, vp, u64]),
    "kt_synchronize": (i32, [vp]),
    "kt_count": (i32, [vp, u32, vp, vp, u32, p_u64, vp]),
    "kt_list_triangles": (i32, [vp, u32, vp, vp, u32, vp, u64, p_u64]),
    "kt_orient": (i32, [vp, u32, vp, vp, u32, vp, vp]),
    "kt_intersect": (i32, [vp, u64, vp, vp, vp, vp, u32, vp, vp, u64, p_u64]),
    "kt_count_noalloc": (i32, [vp, u32, vp, vp, p_u64, vp]),
    # katric_graphio.h
    "kt_host_graph_free": (None, [C.POINTER(KtHostGraph)]),
    "kt_gen_gnm": (i32, [u32, u64, u64, C.POINTER(KtHostGraph)]),
    "kt_gen_rmat": (i32, [u32, u64, C.c_double, C.c_double, C.c_double, u64, u32, i32, C.POINTER(KtHostGraph)]),
    "kt_read_metis": (i32, [C.c_char_p, C.POINTER(KtHostGraph)]),
    "kt_write_metis": (i32, [C.c_char_p, u32, vp, vp]),
    "kt_read_graphio_binary": (i32, [C.c_char_p, C.c_char_p, C.POINTER(KtHostGraph)]),
    "kt_write_graphio_binary": (i32, [C.c_char_p, C.c_char_p, u32, vp, vp]),
    "kt_csr_checksum": (u64, [u32, vp, vp]),
}

_lib = None


def lib():
    """The loaded library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"katric_amd native library not built ({LIB_PATH}); run `python -m katric_amd.build`"
            )
        h = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(h, name)
            f.restype = res
            f.argtypes = args
        _lib = h
    return _lib


def register(name, restype, argtypes):
    """Adds a signature (used by sub-modules for their own entry points)."""
    SIGNATURES[name] = (restype, argtypes)
    if _lib is not None:
        f = getattr(_lib, name)
        f.restype = restype
        f.argtypes = argtypes
