"""ctypes binding of the library's C API (include/GraphBLAS.h, include/gm_kernels.h).

This is how a Python user reaches the drop-in ``GrB_mxm``: the same calls a
C program makes (GrB_Matrix_new, GxB_Matrix_pack_CSR, GrB_mxm,
GrB_Matrix_reduce_INT64, GxB_Matrix_unpack_CSR).  The shared library is the
in-tree build ``graphblas_amd/lib/libgraphblas_mi355x.so``; importing this
module without it raises, and every compute call needs a HIP device (there is
no CPU fallback anywhere in the product).
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

import numpy as np

from .workloads import Csr

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libgraphblas_mi355x.so")

GrB_SUCCESS = 0
GrB_NO_VALUE = 1
GrB_NULL_POINTER = -2
GrB_INVALID_VALUE = -3
GrB_DOMAIN_MISMATCH = -5
GrB_DIMENSION_MISMATCH = -6
GrB_OUTPUT_NOT_EMPTY = -7
GrB_NOT_IMPLEMENTED = -8
GrB_PANIC = -101
GrB_OUT_OF_MEMORY = -102
GrB_INVALID_OBJECT = -104

# gm_type codes (include/gm_kernels.h)
GM_TYPES = {"bool": 1, "int8": 2, "uint8": 3, "int16": 4, "uint16": 5, "int32": 6,
            "uint32": 7, "int64": 8, "uint64": 9, "fp32": 10, "fp64": 11}
NP_TYPES = {"bool": np.bool_, "int32": np.int32, "int64": np.int64, "fp64": np.float64,
            "uint8": np.uint8, "int8": np.int8, "fp32": np.float32, "uint32": np.uint32,
            "uint64": np.uint64, "int16": np.int16, "uint16": np.uint16}
GRB_TYPE_NAMES = {"bool": "GrB_BOOL", "int8": "GrB_INT8", "uint8": "GrB_UINT8",
                  "int16": "GrB_INT16", "uint16": "GrB_UINT16", "int32": "GrB_INT32",
                  "uint32": "GrB_UINT32", "int64": "GrB_INT64", "uint64": "GrB_UINT64",
                  "fp32": "GrB_FP32", "fp64": "GrB_FP64"}

# semiring handle name -> (result type, A/B value type or None for pattern-only)
SEMIRINGS = {
    "GxB_PLUS_PAIR_INT64": ("int64", None),
    "GxB_PLUS_PAIR_INT32": ("int32", None),
    "GxB_ANY_PAIR_BOOL": ("bool", None),
    "GxB_LOR_PAIR_BOOL": ("bool", None),
    "GrB_PLUS_TIMES_SEMIRING_FP64": ("fp64", "fp64"),
    "GrB_PLUS_TIMES_SEMIRING_INT64": ("int64", "int64"),
    "GrB_MIN_PLUS_SEMIRING_INT32": ("int32", "int32"),
    "GrB_MIN_PLUS_SEMIRING_INT64": ("int64", "int64"),
    "GrB_MIN_PLUS_SEMIRING_FP64": ("fp64", "fp64"),
    "GrB_MAX_PLUS_SEMIRING_INT32": ("int32", "int32"),
    "GxB_ANY_SECONDI_INT64": ("int64", None),
    "GxB_ANY_FIRSTJ_INT64": ("int64", None),
}

DESCRIPTORS = ["T1", "T0", "T0T1", "C", "CT1", "CT0", "CT0T1", "S", "ST1", "ST0", "ST0T1",
               "SC", "SCT1", "SCT0", "SCT0T1", "R", "RT1", "RT0", "RT0T1", "RC", "RCT1",
               "RCT0", "RCT0T1", "RS", "RST1", "RST0", "RST0T1", "RSC", "RSCT1", "RSCT0",
               "RSCT0T1"]


class GrBError(RuntimeError):
    def __init__(self, info: int, msg: str):
        super().__init__(f"GrB error {info}: {msg}")
        self.info = info


class GmCsr(ctypes.Structure):
    _fields_ = [("nrows", ctypes.c_int64), ("ncols", ctypes.c_int64),
                ("p", ctypes.c_void_p), ("j", ctypes.c_void_p), ("x", ctypes.c_void_p),
                ("iso", ctypes.c_int32), ("type", ctypes.c_int32),
                ("aux", ctypes.c_void_p)]


class GmDense(ctypes.Structure):
    _fields_ = [("nrows", ctypes.c_int64), ("ncols", ctypes.c_int64),
                ("b", ctypes.c_void_p), ("x", ctypes.c_void_p),
                ("iso", ctypes.c_int32), ("type", ctypes.c_int32)]


class GmDenseOut(ctypes.Structure):
    _fields_ = [("nrows", ctypes.c_int64), ("ncols", ctypes.c_int64), ("nvals", ctypes.c_int64),
                ("b", ctypes.c_void_p), ("x", ctypes.c_void_p),
                ("iso", ctypes.c_int32), ("type", ctypes.c_int32)]


class GmCsrOut(ctypes.Structure):
    _fields_ = [("nrows", ctypes.c_int64), ("ncols", ctypes.c_int64), ("nvals", ctypes.c_int64),
                ("p", ctypes.c_void_p), ("j", ctypes.c_void_p), ("x", ctypes.c_void_p),
                ("iso", ctypes.c_int32), ("type", ctypes.c_int32)]


class GmSemiring(ctypes.Structure):
    _fields_ = [("add_op", ctypes.c_int32), ("mult_op", ctypes.c_int32),
                ("ztype", ctypes.c_int32)]


class GmDot3Params(ctypes.Structure):
    _fields_ = [("thread_max", ctypes.c_int64), ("team16_max", ctypes.c_int64),
                ("wave_max", ctypes.c_int64), ("bitmap_min_len", ctypes.c_int64),
                ("order", ctypes.c_int64), ("row_min", ctypes.c_int64)]


class GmSaxpy3Params(ctypes.Structure):
    _fields_ = [("wave_small_max", ctypes.c_int64), ("wave_large_max", ctypes.c_int64),
                ("block_max", ctypes.c_int64), ("mode", ctypes.c_int64),
                ("table", ctypes.c_int64)]


class GmDot3Request(ctypes.Structure):
    _fields_ = [("M", GmCsr), ("A", GmCsr), ("B", GmCsr)]


class GmSaxpy3Request(ctypes.Structure):
    _fields_ = [("M", GmCsr), ("A", GmCsr)]


class GmServerConfig(ctypes.Structure):
    _fields_ = [("workers", ctypes.c_int32), ("max_batch_reqs", ctypes.c_int32),
                ("flop_budget", ctypes.c_int64), ("mem_budget", ctypes.c_int64)]


class GmServerRequestOpts(ctypes.Structure):
    _fields_ = [("max_steps", ctypes.c_int64), ("keep_visited", ctypes.c_int32),
                ("pad", ctypes.c_int32)]


class GmServerResult(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("pad", ctypes.c_int32), ("k", ctypes.c_int64),
                ("nvals", ctypes.c_int64), ("flops", ctypes.c_int64),
                ("latency_us", ctypes.c_double), ("queue_us", ctypes.c_double),
                ("batch_id", ctypes.c_int64), ("batch_size", ctypes.c_int64),
                ("steps", ctypes.c_int64), ("vnvals", ctypes.c_int64)]


class GmServerCounters(ctypes.Structure):
    _fields_ = [("completed", ctypes.c_int64), ("batches", ctypes.c_int64),
                ("pending", ctypes.c_int64), ("max_in_flight", ctypes.c_int64),
                ("mean_batch", ctypes.c_double), ("flops", ctypes.c_double),
                ("levels", ctypes.c_int64), ("oom_splits", ctypes.c_int64)]


class GmProfRecord(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 48), ("launches", ctypes.c_int64),
                ("total_ms", ctypes.c_double)]


class GmDeviceInfo(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 64), ("arch", ctypes.c_char * 32),
                ("cu_count", ctypes.c_int32), ("lds_per_block", ctypes.c_int32),
                ("hbm_bytes", ctypes.c_int64), ("hbm_peak_gbs", ctypes.c_double)]


_lib = None
_libc = ctypes.CDLL("libc.so.6")
_libc.malloc.restype = ctypes.c_void_p
_libc.malloc.argtypes = [ctypes.c_size_t]
_libc.free.restype = None
_libc.free.argtypes = [ctypes.c_void_p]


def lib():
    """The loaded C library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                           "(there is no fallback path)")
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    PP = ctypes.POINTER(ctypes.c_void_p)
    U64 = ctypes.c_uint64
    I32, I64 = ctypes.c_int32, ctypes.c_int64
    sig = {
        "GrB_init": ([ctypes.c_int], ctypes.c_int),
        "GrB_finalize": ([], ctypes.c_int),
        "GrB_Matrix_new": ([PP, P, U64, U64], ctypes.c_int),
        "GrB_Matrix_free": ([PP], ctypes.c_int),
        "GrB_Matrix_dup": ([PP, P], ctypes.c_int),
        "GrB_Matrix_clear": ([P], ctypes.c_int),
        "GrB_Matrix_nvals": ([ctypes.POINTER(U64), P], ctypes.c_int),
        "GrB_Matrix_nrows": ([ctypes.POINTER(U64), P], ctypes.c_int),
        "GrB_Matrix_ncols": ([ctypes.POINTER(U64), P], ctypes.c_int),
        "GxB_Matrix_iso": ([ctypes.POINTER(ctypes.c_bool), P], ctypes.c_int),
        "GxB_Matrix_pack_CSR": ([P, PP, PP, PP, U64, U64, U64, ctypes.c_bool, ctypes.c_bool, P],
                                ctypes.c_int),
        "GxB_Matrix_unpack_CSR": ([P, PP, PP, PP, ctypes.POINTER(U64), ctypes.POINTER(U64),
                                   ctypes.POINTER(U64), ctypes.POINTER(ctypes.c_bool),
                                   ctypes.POINTER(ctypes.c_bool), P], ctypes.c_int),
        "GrB_mxm": ([P, P, P, P, P, P, P], ctypes.c_int),
        "GrB_Matrix_reduce_INT64": ([ctypes.POINTER(I64), P, P, P, P], ctypes.c_int),
        "GrB_Matrix_reduce_INT32": ([ctypes.POINTER(I32), P, P, P, P], ctypes.c_int),
        "GrB_Matrix_reduce_FP64": ([ctypes.POINTER(ctypes.c_double), P, P, P, P], ctypes.c_int),
        "GrB_Matrix_error": ([ctypes.POINTER(ctypes.c_char_p), P], ctypes.c_int),
        "GrB_Descriptor_new": ([PP], ctypes.c_int),
        "GrB_Descriptor_set": ([P, ctypes.c_int, ctypes.c_int], ctypes.c_int),
        "GrB_Descriptor_free": ([PP], ctypes.c_int),
        "GrB_Semiring_new": ([PP, P, P], ctypes.c_int),
        "GrB_Semiring_free": ([PP], ctypes.c_int),
        "GxB_Global_Option_set_INT32": ([ctypes.c_int, I32], ctypes.c_int),
        "GrB_Matrix_build_INT64": ([P, P, P, P, U64, P], ctypes.c_int),
        "GrB_Matrix_build_FP64": ([P, P, P, P, U64, P], ctypes.c_int),
        "GrB_Matrix_build_INT32": ([P, P, P, P, U64, P], ctypes.c_int),
        "GrB_Matrix_build_BOOL": ([P, P, P, P, U64, P], ctypes.c_int),
        "GrB_Matrix_extractTuples_INT64": ([P, P, P, ctypes.POINTER(U64), P], ctypes.c_int),
        "GrB_Matrix_extractTuples_FP64": ([P, P, P, ctypes.POINTER(U64), P], ctypes.c_int),
        "GrB_Matrix_extractTuples_INT32": ([P, P, P, ctypes.POINTER(U64), P], ctypes.c_int),
        "GrB_Matrix_extractTuples_BOOL": ([P, P, P, ctypes.POINTER(U64), P], ctypes.c_int),
        "GrB_Matrix_reduce_BOOL": ([ctypes.POINTER(ctypes.c_bool), P, P, P, P], ctypes.c_int),
        "GxB_Matrix_type": ([PP, P], ctypes.c_int),
        "GxB_Type_size": ([ctypes.POINTER(ctypes.c_size_t), P], ctypes.c_int),
        "GrB_Type_get_size": ([ctypes.POINTER(ctypes.c_size_t), P], ctypes.c_int),
        "GxB_Global_Option_get_INT32": ([ctypes.c_int, ctypes.POINTER(I32)], ctypes.c_int),
        "GrB_BinaryOp_free": ([PP], ctypes.c_int),
        "GrB_Monoid_free": ([PP], ctypes.c_int),
        "GB_mxm_select": ([ctypes.c_bool, ctypes.c_bool, ctypes.c_bool, ctypes.c_bool,
                           ctypes.c_int], ctypes.c_int),
        "GB_semiring_codes": ([P, ctypes.POINTER(GmSemiring)], ctypes.c_bool),
        "GB_mxm_prefer_dot3": ([ctypes.c_int64] * 6, ctypes.c_bool),
        "GB_AxB_dot2_control": ([ctypes.c_double] * 5, ctypes.c_bool),
        # kernel layer
        "gm_init": ([], ctypes.c_int),
        "gm_finalize": ([], None),
        "gm_last_error": ([], ctypes.c_char_p),
        "gm_device_count": ([], ctypes.c_int),
        "gm_get_device_info": ([ctypes.POINTER(GmDeviceInfo)], ctypes.c_int),
        "gm_malloc": ([ctypes.c_size_t], P),
        "gm_free": ([P], None),
        "gm_arena_trim": ([], ctypes.c_int),
        "gm_arena_bytes_in_use": ([], I64),
        "gm_arena_bytes_cached": ([], I64),
        "gm_debug_fail_after": ([I64], None),
        "gm_debug_dot3_slice_max": ([I64], None),
        "gm_memcpy_h2d": ([P, P, ctypes.c_size_t], ctypes.c_int),
        "gm_memcpy_d2h": ([P, P, ctypes.c_size_t], ctypes.c_int),
        "gm_memcpy_d2d": ([P, P, ctypes.c_size_t], ctypes.c_int),
        "gm_memset": ([P, ctypes.c_int, ctypes.c_size_t], ctypes.c_int),
        "gm_synchronize": ([], ctypes.c_int),
        "gm_set_stream": ([P], ctypes.c_int),
        "gm_get_stream": ([], P),
        "gm_prof_enable": ([ctypes.c_int], None),
        "gm_prof_reset": ([], None),
        "gm_prof_get": ([ctypes.POINTER(GmProfRecord), ctypes.c_int], ctypes.c_int),
        "gm_dot3_branch": ([ctypes.POINTER(GmCsr), ctypes.c_bool, ctypes.POINTER(GmCsr),
                            ctypes.POINTER(GmCsr), ctypes.POINTER(GmSemiring)], ctypes.c_bool),
        "gm_dot3": ([ctypes.POINTER(GmCsrOut), ctypes.POINTER(GmCsr), ctypes.c_bool,
                     ctypes.POINTER(GmCsr), ctypes.POINTER(GmCsr), ctypes.POINTER(GmSemiring),
                     ctypes.POINTER(GmDot3Params)], ctypes.c_int),
        "gm_AxB_dot3_branch": ([ctypes.POINTER(GmCsr), ctypes.c_bool, ctypes.POINTER(GmCsr),
                                ctypes.POINTER(GmCsr), ctypes.POINTER(GmSemiring), ctypes.c_bool],
                               ctypes.c_bool),
        "gm_AxB_dot3": ([ctypes.POINTER(GmCsrOut), ctypes.POINTER(GmCsr), ctypes.c_bool,
                         ctypes.POINTER(GmCsr), ctypes.POINTER(GmCsr), ctypes.POINTER(GmSemiring),
                         ctypes.c_bool], ctypes.c_int),
        "gm_saxpy3_branch": ([ctypes.POINTER(GmCsr), ctypes.c_bool, ctypes.c_bool,
                              ctypes.POINTER(GmCsr), ctypes.POINTER(GmCsr),
                              ctypes.POINTER(GmSemiring)], ctypes.c_bool),
        "gm_saxpy3": ([ctypes.POINTER(GmCsrOut), ctypes.POINTER(GmCsr), ctypes.c_bool,
                       ctypes.c_bool, ctypes.POINTER(GmCsr), ctypes.POINTER(GmCsr),
                       ctypes.POINTER(GmSemiring), ctypes.POINTER(GmSaxpy3Params)], ctypes.c_int),
        "gm_reduce_scalar": ([P, I32, I32, P, I64, I32], ctypes.c_int),
        "gm_reduce_scalar_noalloc": ([P, I32, I32, P, I64, I32, P], ctypes.c_int),
        "gm_csr_aux_build": ([PP, ctypes.POINTER(GmCsr)], ctypes.c_int),
        "gm_dense_to_csr": ([ctypes.POINTER(GmCsrOut), ctypes.POINTER(GmDense)], ctypes.c_int),
        "gm_dot3_work": ([ctypes.POINTER(I64), ctypes.POINTER(I64), ctypes.POINTER(GmCsr),
                          ctypes.POINTER(GmCsr), ctypes.POINTER(GmCsr)], ctypes.c_int),
        "gm_csr_to_dense": ([ctypes.POINTER(GmDenseOut), ctypes.POINTER(GmCsr), ctypes.c_bool],
                            ctypes.c_int),
        "gm_saxbit": ([ctypes.POINTER(GmDenseOut), ctypes.POINTER(GmCsr), ctypes.POINTER(GmDense),
                       ctypes.c_bool, ctypes.c_bool, ctypes.POINTER(GmCsr), ctypes.POINTER(GmDense),
                       ctypes.POINTER(GmCsr), ctypes.POINTER(GmDense), ctypes.POINTER(GmSemiring),
                       ctypes.POINTER(GmDense)], ctypes.c_int),
        "GxB_Matrix_Option_set_INT32": ([P, ctypes.c_int, I32], ctypes.c_int),
        "GxB_Matrix_Option_get_INT32": ([P, ctypes.c_int, ctypes.POINTER(I32)], ctypes.c_int),
        "GxB_Matrix_pack_BitmapR": ([P, PP, PP, U64, U64, ctypes.c_bool, U64, P], ctypes.c_int),
        "GxB_Matrix_pack_FullR": ([P, PP, U64, ctypes.c_bool, P], ctypes.c_int),
        "GxB_Matrix_unpack_BitmapR": ([P, PP, PP, ctypes.POINTER(U64), ctypes.POINTER(U64),
                                       ctypes.POINTER(ctypes.c_bool), ctypes.POINTER(U64), P],
                                      ctypes.c_int),
        "GxB_Matrix_unpack_FullR": ([P, PP, ctypes.POINTER(U64), ctypes.POINTER(ctypes.c_bool), P],
                                    ctypes.c_int),
        "gm_csr_aux_free": ([P], None),
        "gm_batch_dot3": ([ctypes.POINTER(GmCsrOut), P, P, ctypes.POINTER(GmDot3Request), I64,
                           ctypes.c_bool, ctypes.POINTER(GmSemiring),
                           ctypes.POINTER(GmDot3Params)], ctypes.c_int),
        "gm_batch_saxpy3": ([ctypes.POINTER(GmCsrOut), P, ctypes.POINTER(GmSaxpy3Request), I64,
                             ctypes.c_bool, ctypes.c_bool, ctypes.POINTER(GmCsr),
                             ctypes.POINTER(GmSemiring), ctypes.POINTER(GmSaxpy3Params)],
                            ctypes.c_int),
        "gm_batch_split": ([ctypes.POINTER(GmCsrOut), ctypes.POINTER(GmCsrOut), P, P, I64],
                           ctypes.c_int),
        "GxB_Matrix_serialize": ([PP, ctypes.POINTER(U64), P, P], ctypes.c_int),
        "GrB_Matrix_serialize": ([P, ctypes.POINTER(U64), P], ctypes.c_int),
        "GrB_Matrix_serializeSize": ([ctypes.POINTER(U64), P], ctypes.c_int),
        "GxB_Matrix_deserialize": ([PP, P, P, U64, P], ctypes.c_int),
        "GrB_Matrix_deserialize": ([PP, P, P, U64], ctypes.c_int),
        "GrB_Vector_new": ([PP, P, U64], ctypes.c_int),
        "GrB_Vector_free": ([PP], ctypes.c_int),
        "GrB_Vector_nvals": ([ctypes.POINTER(U64), P], ctypes.c_int),
        "GrB_Vector_size": ([ctypes.POINTER(U64), P], ctypes.c_int),
        "GrB_Vector_clear": ([P], ctypes.c_int),
        "GrB_Vector_build_INT64": ([P, P, P, U64, P], ctypes.c_int),
        "GrB_Vector_build_INT32": ([P, P, P, U64, P], ctypes.c_int),
        "GrB_Vector_build_FP64": ([P, P, P, U64, P], ctypes.c_int),
        "GrB_Vector_build_BOOL": ([P, P, P, U64, P], ctypes.c_int),
        "GrB_Vector_extractTuples_INT64": ([P, P, ctypes.POINTER(U64), P], ctypes.c_int),
        "GrB_Vector_extractTuples_INT32": ([P, P, ctypes.POINTER(U64), P], ctypes.c_int),
        "GrB_Vector_extractTuples_FP64": ([P, P, ctypes.POINTER(U64), P], ctypes.c_int),
        "GrB_Vector_extractTuples_BOOL": ([P, P, ctypes.POINTER(U64), P], ctypes.c_int),
        "GrB_Vector_reduce_INT64": ([ctypes.POINTER(I64), P, P, P, P], ctypes.c_int),
        "GrB_Vector_reduce_INT32": ([ctypes.POINTER(I32), P, P, P, P], ctypes.c_int),
        "GrB_Vector_reduce_FP64": ([ctypes.POINTER(ctypes.c_double), P, P, P, P], ctypes.c_int),
        "GrB_Vector_reduce_BOOL": ([ctypes.POINTER(ctypes.c_bool), P, P, P, P], ctypes.c_int),
        "GrB_Vector_dup": ([PP, P], ctypes.c_int),
        "GrB_Vector_error": ([ctypes.POINTER(ctypes.c_char_p), P], ctypes.c_int),
        "GrB_vxm": ([P, P, P, P, P, P, P], ctypes.c_int),
        "GrB_mxv": ([P, P, P, P, P, P, P], ctypes.c_int),
        "gm_transpose": ([ctypes.POINTER(GmCsrOut), ctypes.POINTER(GmCsr)], ctypes.c_int),
        "gm_flopcount": ([ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(GmCsr), ctypes.POINTER(GmCsr)],
                         ctypes.c_int),
        "gm_nonempty_rows": ([ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(GmCsr)], ctypes.c_int),
        "gm_dot2": ([ctypes.POINTER(GmCsrOut), ctypes.POINTER(GmCsr), ctypes.c_bool, ctypes.c_bool,
                     ctypes.POINTER(GmCsr), ctypes.POINTER(GmCsr), ctypes.POINTER(GmSemiring),
                     ctypes.POINTER(GmDot3Params)], ctypes.c_int),
        "gm_accum_mask": ([ctypes.POINTER(GmCsrOut), ctypes.POINTER(GmCsr), ctypes.POINTER(GmCsr),
                           ctypes.c_bool, ctypes.c_bool, ctypes.c_bool, ctypes.POINTER(GmCsr),
                           I32], ctypes.c_int),
        "gm_dot3_set_defaults": ([ctypes.POINTER(GmDot3Params)], ctypes.c_int),
        "gm_saxpy3_set_defaults": ([ctypes.POINTER(GmSaxpy3Params)], ctypes.c_int),
        "gm_load_tuning": ([ctypes.c_char_p], ctypes.c_int),
        "gm_server_create": ([ctypes.POINTER(GmCsr), ctypes.POINTER(GmSemiring),
                              ctypes.POINTER(GmServerConfig)], P),
        "gm_server_submit": ([P, P, I64, P, P], I64),
        "gm_server_submit_bfs": ([P, P, I64, P, P, ctypes.POINTER(GmServerRequestOpts)], I64),
        "gm_server_fetch_visited": ([P, I64, P, P, P], ctypes.c_int),
        "gm_server_step": ([P], I64),
        "gm_server_wait": ([P, I64, ctypes.POINTER(GmServerResult)], ctypes.c_int),
        "gm_server_fetch": ([P, I64, P, P, P], ctypes.c_int),
        "gm_server_release": ([P, I64], ctypes.c_int),
        "gm_server_stats": ([P, ctypes.POINTER(GmServerCounters)], ctypes.c_int),
        "gm_server_destroy": ([P], None),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def handle(name: str) -> ctypes.c_void_p:
    """Value of a global handle such as GxB_PLUS_PAIR_INT64 or GrB_DESC_ST1."""
    return ctypes.c_void_p.in_dll(lib(), name).value


class Descriptor:
    """A GrB_Descriptor built field by field, e.g. Descriptor(mask_struct=True,
    axb=GxB_AxB_SAXPY); free() releases it."""

    def __init__(self, mask_struct=False, mask_comp=False, replace=False, axb=0):
        self.h = ctypes.c_void_p()
        _check(lib().GrB_Descriptor_new(ctypes.byref(self.h)), "GrB_Descriptor_new")
        fields = []
        if replace:
            fields.append((0, 1))                     # GrB_OUTP = GrB_REPLACE
        if mask_struct or mask_comp:
            fields.append((1, (4 if mask_struct else 0) + (2 if mask_comp else 0)))
        if axb:
            fields.append((7090, axb))                # GxB_AxB_METHOD
        for f, v in fields:
            _check(lib().GrB_Descriptor_set(self.h, f, v), "GrB_Descriptor_set")

    def free(self):
        if self.h:
            lib().GrB_Descriptor_free(ctypes.byref(self.h))
            self.h = ctypes.c_void_p()


GxB_AxB_GUSTAVSON, GxB_AxB_DOT, GxB_AxB_HASH, GxB_AxB_SAXPY = 7081, 7083, 7084, 7085


def desc(name):
    """A descriptor argument: None, a Descriptor, or the suffix of a
    predefined GrB_DESC_* ("S", "RSC", "T1", ...)."""
    if name is None:
        return None
    if isinstance(name, Descriptor):
        return name.h
    return handle("GrB_DESC_" + name)


_initialized = False


def init():
    global _initialized
    if not _initialized:
        rc = lib().GrB_init(0)
        if rc not in (GrB_SUCCESS, GrB_INVALID_VALUE):
            raise GrBError(rc, "GrB_init")
        _initialized = True


def _check(rc: int, what: str, obj=None):
    if rc != GrB_SUCCESS:
        msg = what
        if obj is not None:
            s = ctypes.c_char_p()
            lib().GrB_Matrix_error(ctypes.byref(s), obj)
            if s.value:
                msg += ": " + s.value.decode(errors="replace")
        raise GrBError(rc, msg)


def _malloc_copy(a: np.ndarray) -> int:
    a = np.ascontiguousarray(a)
    nb = max(a.nbytes, 8)
    p = _libc.malloc(nb)
    if not p:
        raise MemoryError("malloc")
    ctypes.memmove(p, a.ctypes.data, a.nbytes)
    return p


class Matrix:
    """A GrB_Matrix handle (CSR in HBM)."""

    def __init__(self, type_name: str, nrows: int, ncols: int):
        init()
        self.type_name = type_name
        self.h = ctypes.c_void_p()
        _check(lib().GrB_Matrix_new(ctypes.byref(self.h), handle(GRB_TYPE_NAMES[type_name]),
                                    nrows, ncols), "GrB_Matrix_new")
        self.nrows, self.ncols = nrows, ncols

    def __del__(self):
        try:
            if self.h and self.h.value:
                lib().GrB_Matrix_free(ctypes.byref(self.h))
        except Exception:
            pass

    def clear(self):
        _check(lib().GrB_Matrix_clear(self.h), "GrB_Matrix_clear", self.h)

    @property
    def nvals(self) -> int:
        v = ctypes.c_uint64()
        _check(lib().GrB_Matrix_nvals(ctypes.byref(v), self.h), "nvals", self.h)
        return v.value

    @property
    def iso(self) -> bool:
        v = ctypes.c_bool()
        _check(lib().GxB_Matrix_iso(ctypes.byref(v), self.h), "iso", self.h)
        return v.value

    # storage formats (GxB_SPARSITY_STATUS / GxB_SPARSITY_CONTROL)
    SPARSE, BITMAP, FULL = 2, 4, 8

    @property
    def sparsity(self) -> int:
        v = ctypes.c_int32()
        _check(lib().GxB_Matrix_Option_get_INT32(self.h, 7034, ctypes.byref(v)), "sparsity", self.h)
        return v.value

    def set_sparsity_control(self, ctrl: int):
        _check(lib().GxB_Matrix_Option_set_INT32(self.h, 7036, ctrl), "sparsity control", self.h)

    @classmethod
    def from_dense(cls, x: np.ndarray, present: Optional[np.ndarray], type_name: str) -> "Matrix":
        """Pack a row-major dense array as a bitmap (present given) or full
        matrix (present None)."""
        nrows, ncols = x.shape
        M = cls(type_name, nrows, ncols)
        dt = NP_TYPES[type_name]
        xv = np.ascontiguousarray(x, dtype=dt)
        xp = ctypes.c_void_p(_malloc_copy(xv))
        if present is None:
            rc = lib().GxB_Matrix_pack_FullR(M.h, ctypes.byref(xp), max(xv.nbytes, 1), False, None)
            bp = None
        else:
            bv = np.ascontiguousarray(present, dtype=np.int8)
            bp = ctypes.c_void_p(_malloc_copy(bv))
            rc = lib().GxB_Matrix_pack_BitmapR(M.h, ctypes.byref(bp), ctypes.byref(xp),
                                               max(bv.nbytes, 1), max(xv.nbytes, 1), False,
                                               int(np.count_nonzero(bv)), None)
        if rc != GrB_SUCCESS:
            for q in (xp, bp):
                if q is not None and q.value:
                    _libc.free(q)
        _check(rc, "GxB_Matrix_pack", M.h)
        return M

    def to_dense(self):
        """(values, present) as row-major arrays (the matrix is unchanged)."""
        D = ctypes.c_void_p()
        _check(lib().GrB_Matrix_dup(ctypes.byref(D), self.h), "GrB_Matrix_dup", self.h)
        bp, xp = ctypes.c_void_p(), ctypes.c_void_p()
        bs, xs, nv = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        iso = ctypes.c_bool()
        rc = lib().GxB_Matrix_unpack_BitmapR(D, ctypes.byref(bp), ctypes.byref(xp), ctypes.byref(bs),
                                             ctypes.byref(xs), ctypes.byref(iso), ctypes.byref(nv),
                                             None)
        try:
            _check(rc, "GxB_Matrix_unpack_BitmapR", D)
            mn = self.nrows * self.ncols
            dt = NP_TYPES[self.type_name]
            b = np.frombuffer(bytes((ctypes.c_uint8 * max(mn, 1)).from_address(bp.value)),
                              dtype=np.int8)[:mn].copy()
            nx = 1 if iso.value else max(mn, 1)
            xb = (ctypes.c_uint8 * (nx * np.dtype(dt).itemsize)).from_address(xp.value)
            x = np.frombuffer(bytes(xb), dtype=dt)
            x = np.full(mn, x[0], dtype=dt) if iso.value else x[:mn].copy()
            return x.reshape(self.nrows, self.ncols), b.reshape(self.nrows, self.ncols) != 0
        finally:
            for q in (bp, xp):
                if q.value:
                    _libc.free(q)
            lib().GrB_Matrix_free(ctypes.byref(D))

    @classmethod
    def from_csr(cls, m: Csr, type_name: str, iso_value=None) -> "Matrix":
        """Pack host CSR arrays (copied) into a new matrix.  Pattern-only
        matrices (m.x is None) become iso with value 1 / true."""
        M = cls(type_name, m.nrows, m.ncols)
        dt = NP_TYPES[type_name]
        if m.x is None:
            x = np.array([1 if iso_value is None else iso_value], dtype=dt)
            iso = True
        elif m.iso:
            x = np.asarray(m.x[:1], dtype=dt)
            iso = True
        else:
            x = np.asarray(m.x, dtype=dt)
            iso = False
        pp = ctypes.c_void_p(_malloc_copy(np.asarray(m.p, dtype=np.int64)))
        jp = ctypes.c_void_p(_malloc_copy(np.asarray(m.j, dtype=np.int64)))
        xp = ctypes.c_void_p(_malloc_copy(x))
        rc = lib().GxB_Matrix_pack_CSR(M.h, ctypes.byref(pp), ctypes.byref(jp), ctypes.byref(xp),
                                       8 * (m.nrows + 1), max(8 * m.nnz, 8), max(x.nbytes, 1),
                                       iso, False, None)
        if rc != GrB_SUCCESS:
            for q in (pp, jp, xp):
                if q.value:
                    _libc.free(q)
        _check(rc, "GxB_Matrix_pack_CSR", M.h)
        return M

    def to_csr(self) -> Csr:
        """Copy to host CSR (the matrix is unchanged)."""
        D = ctypes.c_void_p()
        _check(lib().GrB_Matrix_dup(ctypes.byref(D), self.h), "GrB_Matrix_dup", self.h)
        pp, jp, xp = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        ps, js, xs = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        iso, jum = ctypes.c_bool(), ctypes.c_bool()
        rc = lib().GxB_Matrix_unpack_CSR(D, ctypes.byref(pp), ctypes.byref(jp), ctypes.byref(xp),
                                         ctypes.byref(ps), ctypes.byref(js), ctypes.byref(xs),
                                         ctypes.byref(iso), ctypes.byref(jum), None)
        try:
            _check(rc, "GxB_Matrix_unpack_CSR", D)
            n = self.nrows
            p = np.ctypeslib.as_array((ctypes.c_int64 * (n + 1)).from_address(pp.value)).copy()
            nz = int(p[n])
            j = np.ctypeslib.as_array((ctypes.c_int64 * max(nz, 1)).from_address(jp.value))[:nz].copy()
            dt = NP_TYPES[self.type_name]
            nx = 1 if iso.value else max(nz, 1)
            xb = (ctypes.c_uint8 * (nx * np.dtype(dt).itemsize)).from_address(xp.value)
            x = np.frombuffer(bytes(xb), dtype=dt)
            x = np.full(nz, x[0], dtype=dt) if iso.value else x[:nz].copy()
            return Csr(n, self.ncols, p, j, x, False)
        finally:
            for q in (pp, jp, xp):
                if q.value:
                    _libc.free(q)
            lib().GrB_Matrix_free(ctypes.byref(D))


_VT = {"int64": "INT64", "int32": "INT32", "fp64": "FP64", "bool": "BOOL"}


class Vector:
    """A GrB_Vector (held in HBM as a 1 x n CSR row)."""

    def __init__(self, type_name: str, n: int):
        init()
        self.type_name = type_name
        self.n = n
        h = ctypes.c_void_p()
        _check(lib().GrB_Vector_new(ctypes.byref(h), handle(GRB_TYPE_NAMES[type_name]), n),
               "GrB_Vector_new")
        self.h = h

    def __del__(self):
        try:
            if self.h:
                lib().GrB_Vector_free(ctypes.byref(self.h))
        except Exception:
            pass

    @classmethod
    def from_arrays(cls, idx, vals, type_name: str, n: int) -> "Vector":
        v = cls(type_name, n)
        I = np.ascontiguousarray(idx, dtype=np.uint64)
        X = np.ascontiguousarray(vals, dtype=NP_TYPES[type_name])
        f = getattr(lib(), "GrB_Vector_build_" + _VT[type_name])
        _check(f(v.h, I.ctypes.data, X.ctypes.data, len(I), None), "GrB_Vector_build")
        return v

    @property
    def nvals(self) -> int:
        n = ctypes.c_uint64()
        _check(lib().GrB_Vector_nvals(ctypes.byref(n), self.h), "GrB_Vector_nvals")
        return n.value

    def to_arrays(self):
        nv = self.nvals
        I = np.empty(max(nv, 1), dtype=np.uint64)
        X = np.empty(max(nv, 1), dtype=NP_TYPES[self.type_name])
        n = ctypes.c_uint64(nv)
        f = getattr(lib(), "GrB_Vector_extractTuples_" + _VT[self.type_name])
        _check(f(I.ctypes.data, X.ctypes.data, ctypes.byref(n), self.h), "GrB_Vector_extractTuples")
        return I[:n.value].astype(np.int64), X[:n.value]


def vxm(w: Vector, mask: Optional[Vector], semiring: str, u: Vector, A: Matrix,
        descriptor: Optional[str] = None, accum: Optional[str] = None) -> int:
    return lib().GrB_vxm(w.h, mask.h if mask is not None else None,
                         handle(accum) if accum else None, handle(semiring), u.h, A.h,
                         desc(descriptor))


def mxv(w: Vector, mask: Optional[Vector], semiring: str, A: Matrix, u: Vector,
        descriptor: Optional[str] = None, accum: Optional[str] = None) -> int:
    return lib().GrB_mxv(w.h, mask.h if mask is not None else None,
                         handle(accum) if accum else None, handle(semiring), A.h, u.h,
                         desc(descriptor))


def mxm(C: Matrix, M: Optional[Matrix], semiring: str, A: Matrix, B: Matrix,
        descriptor=None, accum=None) -> int:
    """GrB_mxm(C, M, accum, semiring, A, B, desc); returns the GrB_Info code."""
    return lib().GrB_mxm(C.h, M.h if M is not None else None, accum, handle(semiring),
                         A.h, B.h, desc(descriptor))


def mxm_checked(C, M, semiring, A, B, descriptor=None, accum=None):
    _check(mxm(C, M, semiring, A, B, descriptor, handle(accum) if accum else None), "GrB_mxm", C.h)


def reduce_sum_int64(A: Matrix) -> int:
    v = ctypes.c_int64(0)
    _check(lib().GrB_Matrix_reduce_INT64(ctypes.byref(v), None, handle("GrB_PLUS_MONOID_INT64"),
                                         A.h, None), "reduce", A.h)
    return v.value


def reduce_sum_fp64(A: Matrix) -> float:
    v = ctypes.c_double(0)
    _check(lib().GrB_Matrix_reduce_FP64(ctypes.byref(v), None, handle("GrB_PLUS_MONOID_FP64"),
                                        A.h, None), "reduce", A.h)
    return v.value


def set_burble(on: bool):
    lib().GxB_Global_Option_set_INT32(7019, 1 if on else 0)


def last_error() -> str:
    s = lib().gm_last_error()
    return s.decode(errors="replace") if s else ""


def device_info() -> dict:
    info = GmDeviceInfo()
    rc = lib().gm_get_device_info(ctypes.byref(info))
    if rc != 0:
        raise GrBError(rc, last_error())
    return {"name": info.name.decode(), "arch": info.arch.decode(), "cu_count": info.cu_count,
            "lds_per_block": info.lds_per_block, "hbm_bytes": info.hbm_bytes,
            "hbm_peak_gbs_hip_attrs": info.hbm_peak_gbs}


def load_tuning(path: str):
    """Process-wide bin cutoffs from an autotuner file (graphblas_amd/tune.py)."""
    rc = lib().gm_load_tuning(path.encode())
    if rc != 0:
        raise GrBError(rc, last_error())


def reset_tuning():
    lib().gm_dot3_set_defaults(None)
    lib().gm_saxpy3_set_defaults(None)


def prof_enable(on=True):
    lib().gm_prof_enable(1 if on else 0)


def prof_reset():
    lib().gm_prof_reset()


def prof_get() -> dict:
    recs = (GmProfRecord * 64)()
    n = lib().gm_prof_get(recs, 64)
    return {recs[t].name.decode(): (recs[t].launches, recs[t].total_ms) for t in range(min(n, 64))}


class DeviceCsr:
    """A host CSR copied into gm_malloc'd HBM; .view is its gm_csr struct."""

    def __init__(self, m: Csr, type_code: int, iso=False):
        L = lib()
        self.bufs = []
        self.m = m

        def up(a):
            a = np.ascontiguousarray(a)
            d = L.gm_malloc(max(a.nbytes, 8))
            if not d:
                raise GrBError(-102, last_error())
            if a.nbytes and L.gm_memcpy_h2d(d, a.ctypes.data, a.nbytes) != 0:
                raise GrBError(-101, last_error())
            self.bufs.append(d)
            return d

        p = up(np.asarray(m.p, dtype=np.int64))
        j = up(np.asarray(m.j, dtype=np.int64)) if m.nnz > 0 else None
        x = None
        if m.x is not None:
            x = up(m.x[:1] if iso else m.x)
        self.view = GmCsr(m.nrows, m.ncols, p, j, x, 1 if iso else 0, type_code)

    def free(self):
        for d in self.bufs:
            lib().gm_free(d)
        self.bufs = []


class Server:
    """Served frontier expansion on one resident graph (gm_server_*).

    graph: a gm_csr view whose device arrays outlive the server."""

    def __init__(self, graph: GmCsr, semiring=(4, 4, 8), workers=2, max_batch_reqs=256,
                 flop_budget=1 << 24, mem_budget=0):
        self._graph = graph
        self._sr = GmSemiring(*semiring)
        self._zt = semiring[2]
        cfg = GmServerConfig(workers, max_batch_reqs, flop_budget, mem_budget)
        self.h = lib().gm_server_create(ctypes.byref(self._graph), ctypes.byref(self._sr),
                                        ctypes.byref(cfg))
        if not self.h:
            raise GrBError(-1, last_error())

    def submit(self, sources, vis=None, max_steps=1, keep_visited=False) -> int:
        """Queue a BFS of max_steps levels (<= 0: until the frontier is empty)
        from the sources; vis = (vis_p, vis_j) marks vertices visited."""
        src = np.ascontiguousarray(sources, dtype=np.int64)
        vp = vj = None
        if vis is not None:
            vp = np.ascontiguousarray(vis[0], dtype=np.int64)
            vj = np.ascontiguousarray(vis[1], dtype=np.int64)
            if len(vj) == 0:
                vj = np.zeros(1, dtype=np.int64)
        opts = GmServerRequestOpts(max_steps, 1 if keep_visited else 0, 0)
        rid = lib().gm_server_submit_bfs(self.h, src.ctypes.data, len(src),
                                         vp.ctypes.data if vp is not None else None,
                                         vj.ctypes.data if vj is not None else None,
                                         ctypes.byref(opts))
        if rid <= 0:
            raise GrBError(int(rid), last_error())
        return rid

    def wait(self, rid) -> GmServerResult:
        res = GmServerResult()
        rc = lib().gm_server_wait(self.h, rid, ctypes.byref(res))
        if rc != 0:
            raise GrBError(rc, last_error())
        return res

    def fetch(self, rid, res: GmServerResult):
        """(row pointers, column indices, values) of a completed request."""
        dt = {8: np.int64, 6: np.int32, 1: np.uint8, 11: np.float64}[self._zt]
        p = np.empty(res.k + 1, dtype=np.int64)
        j = np.empty(max(res.nvals, 1), dtype=np.int64)
        x = np.empty(max(res.nvals, 1), dtype=dt)
        rc = lib().gm_server_fetch(self.h, rid, p.ctypes.data, j.ctypes.data, x.ctypes.data)
        if rc != 0:
            raise GrBError(rc, last_error())
        return p, j[:res.nvals], x[:res.nvals]

    def fetch_visited(self, rid, res: GmServerResult):
        """(row pointers, column indices, parents) of a completed request's
        visited set (submitted with keep_visited)."""
        dt = {8: np.int64, 6: np.int32, 1: np.uint8, 11: np.float64}[self._zt]
        p = np.empty(res.k + 1, dtype=np.int64)
        j = np.empty(max(res.vnvals, 1), dtype=np.int64)
        x = np.empty(max(res.vnvals, 1), dtype=dt)
        rc = lib().gm_server_fetch_visited(self.h, rid, p.ctypes.data, j.ctypes.data, x.ctypes.data)
        if rc != 0:
            raise GrBError(rc, last_error())
        return p, j[:res.vnvals], x[:res.vnvals]

    def step(self) -> int:
        """Without worker threads: run one batch here; its request count."""
        n = lib().gm_server_step(self.h)
        if n < 0:
            raise GrBError(int(n), last_error())
        return n

    def release(self, rid):
        lib().gm_server_release(self.h, rid)

    def stats(self) -> dict:
        c = GmServerCounters()
        lib().gm_server_stats(self.h, ctypes.byref(c))
        return {f: getattr(c, f) for f, _ in GmServerCounters._fields_}

    def close(self):
        if self.h:
            lib().gm_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
