"""ctypes binding of libpoyhip.so (the C ABI in include/poyhip.h).

The library is built in-tree by ``poy5_amd/csrc/Makefile`` (see
``__graft_entry__.build``).  Loading fails loudly when it is missing: there is
no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpoyhip.so")

POYHIP_OK = 0
ERRORS = {
    -1: "POYHIP_EINVAL", -2: "POYHIP_ENOMEM", -3: "POYHIP_EDEVICE",
    -4: "POYHIP_ESTATE", -5: "POYHIP_ECAP", -6: "POYHIP_EBUSY",
}

OP_ALIGN_AFFINE = 1
OP_COST_AFFINE = 2
OP_ALIGN_2D = 3
OP_COST_2D = 4
OP_ALIGN_3D = 5
OP_COST_3D = 6
OP_NEWKK_ALIGN = 7
OP_NEWKK_COST = 8

POST_MEDIAN_2 = 1
POST_MEDIAN_2_WITH_GAPS = 2
POST_CLOSEST = 4
POST_CLOSEST_EQ = 5


class PoyhipCM(C.Structure):
    _fields_ = [
        ("lcm", C.c_int), ("gap", C.c_int), ("a_sz", C.c_int),
        ("combinations", C.c_int), ("cost_model", C.c_int), ("gap_open", C.c_int),
        ("cost", C.c_int * 2048), ("worst", C.c_int * 2048),
        ("prepend", C.c_int * 64), ("tail", C.c_int * 64),
        ("median", C.c_uint8 * 2048),
    ]


class PoyhipCM3(C.Structure):
    _fields_ = [("cost", C.c_int * 32768), ("median", C.c_uint8 * 32768)]


class PoyhipServerOpts(C.Structure):
    _fields_ = [("n_slots", C.c_int), ("max_batch_pairs", C.c_int), ("max_len", C.c_int),
                ("slot_dm_bytes", C.c_size_t), ("pool_bytes", C.c_size_t)]


REQ_KEEP_MEDIAN = 1


class PoyhipError(RuntimeError):
    def __init__(self, code, msg=""):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


_lock = threading.Lock()
_lib = None

# name -> (restype, argtypes); every symbol here is declared in poyhip.h or
# poyhip_tools.h (benchmark generators and test hooks)
_u8p = C.POINTER(C.c_uint8)
_ip = C.POINTER(C.c_int)
_i64p = C.POINTER(C.c_int64)
_vp = C.c_void_p
_i32p = C.POINTER(C.c_int32)
SIGNATURES = {
    "poyhip_create": (C.c_int, [C.c_int, C.POINTER(_vp)]),
    "poyhip_destroy": (C.c_int, [_vp]),
    "poyhip_last_error": (C.c_char_p, [_vp]),
    "poyhip_set_cm": (C.c_int, [_vp, C.POINTER(PoyhipCM)]),
    "poyhip_flush": (C.c_int, [_vp]),
    "poyhip_cost_2": (C.c_int, [_vp, _u8p, C.c_int, _u8p, C.c_int, C.c_int, _ip]),
    "poyhip_backtrace_2d": (C.c_int, [_vp, _u8p, C.c_int, _u8p, C.c_int, _u8p, _u8p,
                                      C.c_int, _ip, C.c_int]),
    "poyhip_limit_2": (C.c_int, [_vp, _u8p, C.c_int, _u8p, C.c_int, C.c_int, C.c_int, C.c_int,
                                 C.c_int, C.c_int, C.c_int, _ip]),
    "poyhip_backtrace_2d_limit": (C.c_int, [_vp, _u8p, C.c_int, _u8p, C.c_int, _u8p, _u8p,
                                            C.c_int, _ip, C.c_int, C.c_int, C.c_int, C.c_int,
                                            C.c_int]),
    "poyhip_align_2d": (C.c_int, [_vp, _u8p, C.c_int, _u8p, C.c_int, C.c_int, C.c_int,
                                  _u8p, _u8p, C.c_int, _ip, _ip]),
    "poyhip_cost_affine_3": (C.c_int, [_vp, _u8p, C.c_int, _u8p, C.c_int, _ip]),
    "poyhip_union_merge": (C.c_int, [_vp, _u8p, _u8p, _u8p, C.c_int, _u8p, C.c_int, C.c_int,
                                     _i32p, _u8p, C.c_int, C.c_int, _i32p, _u8p, C.c_int,
                                     _i32p, _i32p, _i32p, _ip]),
    "poyhip_union_merge_batch": (C.c_int, [_vp, C.c_int, _vp, _vp, C.c_int64, _vp, C.c_int64,
                                           _vp, _vp, C.c_int]),
    "poyhip_newkk_cost_2_affine": (C.c_int, [_vp, _u8p, C.c_int, _u8p, C.c_int, C.c_int, _ip]),
    "poyhip_newkk_align_2_affine": (C.c_int, [_vp, _u8p, C.c_int, _u8p, C.c_int, C.c_int, _ip,
                                              _u8p, _u8p, C.c_int, _ip]),
    "poyhip_align_affine_3": (C.c_int, [_vp, _u8p, C.c_int, _u8p, C.c_int, C.c_int,
                                        _u8p, _ip, _u8p, _ip, _u8p, _u8p, _ip, C.c_int, _ip]),
    "poyhip_align_affine_3_noalloc": (C.c_int, [_vp, _u8p, C.c_int, _u8p, C.c_int, C.c_int,
                                                _u8p, _ip, _u8p, _ip, _u8p, _u8p, _ip, C.c_int,
                                                _ip]),
    "poyhip_ancestor_2": (C.c_int, [_vp, _u8p, _u8p, C.c_int, _u8p, C.c_int, _ip]),
    "poyhip_median_2": (C.c_int, [_vp, _u8p, _u8p, C.c_int, _u8p, C.c_int, _ip]),
    "poyhip_median_2_with_gaps": (C.c_int, [_vp, _u8p, _u8p, C.c_int, _u8p, C.c_int, _ip]),
    "poyhip_worst_2": (C.c_int, [_vp, _u8p, _u8p, C.c_int, _ip]),
    "poyhip_closest_2": (C.c_int, [_vp, _u8p, _u8p, C.c_int, C.c_int, _u8p, C.c_int, _ip]),
    "poyhip_pairs_post_batch": (C.c_int, [_vp, C.c_int, C.c_int, _vp, _vp, C.c_int64, _vp, _vp,
                                          C.c_int]),
    "poyhip_set_cm3": (C.c_int, [_vp, C.POINTER(PoyhipCM3)]),
    "poyhip_cost_3": (C.c_int, [_vp, _u8p, C.c_int, _u8p, C.c_int, _u8p, C.c_int, C.c_int, _ip]),
    "poyhip_backtrack_3d": (C.c_int, [_vp, _u8p, C.c_int, _u8p, C.c_int, _u8p, C.c_int, _u8p,
                                      _u8p, _u8p, C.c_int, _ip]),
    "poyhip_align_3d": (C.c_int, [_vp, _u8p, C.c_int, _u8p, C.c_int, _u8p, C.c_int, C.c_int,
                                  _u8p, _u8p, _u8p, C.c_int, _ip, _ip]),
    "poyhip_median_3": (C.c_int, [_vp, _u8p, _u8p, _u8p, C.c_int, _u8p, C.c_int, _ip]),
    "poyhip_batch3_create": (C.c_int, [_vp, C.c_int, C.c_int, _ip, _ip, _ip, C.c_size_t,
                                       C.POINTER(_vp)]),
    "poyhip_batch3_destroy": (C.c_int, [_vp]),
    "poyhip_batch3_out_bytes": (C.c_size_t, [_vp]),
    "poyhip_batch3_out_offset": (C.c_int64, [_vp, C.c_int]),
    "poyhip_batch3_out_cap": (C.c_int, [_vp, C.c_int]),
    "poyhip_batch3_run": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, C.c_int]),
    "poyhip_batch3_chunks": (C.c_int, [_vp]),
    "poyhip_batch3_last_fill_ms": (C.c_float, [_vp]),
    "poyhip_server_create": (C.c_int, [_vp, C.POINTER(PoyhipServerOpts), C.POINTER(_vp)]),
    "poyhip_server_destroy": (C.c_int, [_vp]),
    "poyhip_server_put_seq": (C.c_int, [_vp, _u8p, C.c_int, _i64p]),
    "poyhip_server_get_seq": (C.c_int, [_vp, C.c_int64, _u8p, C.c_int, _ip]),
    "poyhip_server_submit": (C.c_int, [_vp, C.c_int, C.c_int64, C.c_int64, C.c_int, C.c_int,
                                       _i64p, _i64p]),
    "poyhip_server_wait": (C.c_int, [_vp, C.c_int64, C.c_int]),
    "poyhip_server_result": (C.c_int, [_vp, C.c_int64, _ip, _u8p, _u8p, C.c_int, _ip,
                                       C.POINTER(C.c_double), C.c_int]),
    "poyhip_server_fail_next_launch": (C.c_int, [_vp]),
    "poyhip_server_stats": (C.c_int, [_vp, _i64p, _i64p, C.POINTER(C.c_double), _ip, _ip]),
    "poyhip_server_replay": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_double), _ip, _i64p, _i64p,
                                       _ip, _ip, _ip, C.POINTER(C.c_double), _i64p,
                                       C.POINTER(C.c_double)]),
    "poyhip_gen_triples": (C.c_int, [C.c_int, C.c_int, C.c_uint64, _u8p, _i64p, _ip, _i64p]),
    "poyhip_batch_create": (C.c_int, [_vp, C.c_int, C.c_int, _ip, _ip, _ip, C.c_size_t,
                                      C.POINTER(_vp)]),
    "poyhip_batch_destroy": (C.c_int, [_vp]),
    "poyhip_batch_out_bytes": (C.c_size_t, [_vp]),
    "poyhip_batch_out_offset": (C.c_int64, [_vp, C.c_int]),
    "poyhip_batch_out_cap": (C.c_int, [_vp, C.c_int]),
    "poyhip_batch_out_layout": (C.c_int, [_vp, C.POINTER(C.c_int64), C.POINTER(C.c_int)]),
    "poyhip_batch_run": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, C.c_int]),
    "poyhip_batch_stats": (C.c_int, [_vp, _i64p, _i64p, _ip, _ip]),
    "poyhip_batch_use_graph": (C.c_int, [_vp, C.c_int]),
    "poyhip_set_tuning": (C.c_int, [_vp, C.c_int, C.c_int]),
    "poyhip_get_tuning": (C.c_int, [_vp, C.c_int]),
    "poyhip_batch_last_fill_ms": (C.c_float, [_vp]),
    "poyhip_batch_last_total_ms": (C.c_float, [_vp]),
    "poyhip_batch_chunks": (C.c_int, [_vp]),
    "poyhip_sync": (C.c_int, [_vp]),
    "poyhip_stream": (_vp, [_vp]),
    "poyhip_gen_pairs": (C.c_int, [C.c_int, C.c_int, C.c_uint64, _u8p, _i64p, _ip, _i64p]),
}


def lib():
    """Loads libpoyhip.so; raises ImportError when it has not been built."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise ImportError(
                    f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                    "(no CPU fallback exists)")
            L = C.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                f = getattr(L, name)
                f.restype = res
                f.argtypes = args
            _lib = L
        return _lib


def check(rc, ctx=None):
    if rc != POYHIP_OK:
        msg = ""
        if ctx is not None:
            m = lib().poyhip_last_error(ctx)
            msg = m.decode() if m else ""
        raise PoyhipError(rc, msg)
    return rc
