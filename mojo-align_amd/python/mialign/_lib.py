"""ctypes binding of libmialign.so (the C ABI declared in include/mialign.h).

The library is loaded from this package directory; a missing or unloadable
library raises MialignError — there is no CPU fallback for the product path.
"""

from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libmialign.so"

MA_OK = 0
MA_ERR_NOMEM = -4
MA_ENC_ASCII = 0
MA_ENC_2BIT = 1
MA_BS_NONE = 0
MA_BS_C2T = 1
MA_BS_G2A = 2

BS_MODES = {"": MA_BS_NONE, "none": MA_BS_NONE, "C2T": MA_BS_C2T, "G2A": MA_BS_G2A}


class MialignError(RuntimeError):
    """Raised when a C-ABI call fails (the reference raises a Mojo Error)."""


class Params(C.Structure):
    _fields_ = [
        ("k", C.c_int32),
        ("stride", C.c_int32),
        ("max_occ", C.c_int32),
        ("vote_occ", C.c_int32),
        ("vote_fast", C.c_int32),
        ("max_diff", C.c_int32),
        ("max_soft", C.c_int32),
        ("max_indel", C.c_int32),
        ("do_rc", C.c_int32),
    ]


class BuildOpts(C.Structure):
    _fields_ = [("keep_dense", C.c_int32), ("partition_bits", C.c_int32), ("chunk_windows", C.c_int64)]


class Recs(C.Structure):
    _fields_ = [
        ("cid", C.c_void_p),
        ("pos", C.c_void_p),
        ("mapq", C.c_void_p),
        ("flag", C.c_void_p),
        ("sl", C.c_void_p),
        ("sr", C.c_void_p),
        ("nm", C.c_void_p),
        ("iop", C.c_void_p),
        ("iat", C.c_void_p),
        ("ilen", C.c_void_p),
    ]


BAM_FIELDS = ("flag", "tid", "pos", "mapq", "sl", "sr", "nm", "ntid", "npos", "tlen")


class BamFields(C.Structure):
    _fields_ = [(f, C.c_void_p) for f in BAM_FIELDS]


class BamReads(C.Structure):
    _fields_ = [
        ("names", C.c_void_p),
        ("name_off", C.c_void_p),
        ("seq", C.c_void_p),
        ("qual", C.c_void_p),
        ("lens", C.c_void_p),
        ("qual_lens", C.c_void_p),
        ("stride", C.c_int64),
    ]


REC_FIELDS = ("cid", "pos", "mapq", "flag", "sl", "sr", "nm", "iop", "iat", "ilen")
REC_DTYPES = {
    "cid": np.int32,
    "pos": np.uint32,
    "mapq": np.uint32,
    "flag": np.int32,
    "sl": np.uint32,
    "sr": np.uint32,
    "nm": np.uint32,
    "iop": np.uint32,
    "iat": np.uint32,
    "ilen": np.uint32,
}

P = C.c_void_p
I32 = C.c_int32
I64 = C.c_int64

# name -> (restype, argtypes); every symbol declared in include/mialign.h
SIGNATURES = {
    "ma_params_default": (None, [C.POINTER(Params)]),
    "ma_last_error": (C.c_char_p, []),
    "ma_version": (C.c_char_p, []),
    "ma_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "ma_device_info": (C.c_int, [C.c_int, C.c_char_p, C.c_int, C.POINTER(I64), C.POINTER(I64),
                                 C.POINTER(I32), C.POINTER(I32)]),
    "ma_dev_alloc": (C.c_int, [C.c_int, I64, C.POINTER(P)]),
    "ma_dev_free": (C.c_int, [P]),
    "ma_memcpy_h2d": (C.c_int, [P, P, I64, P]),
    "ma_memcpy_d2h": (C.c_int, [P, P, I64, P]),
    "ma_memcpy_d2d": (C.c_int, [P, P, I64, P]),
    "ma_stream_create": (C.c_int, [C.c_int, C.POINTER(P)]),
    "ma_stream_destroy": (C.c_int, [P]),
    "ma_stream_sync": (C.c_int, [P]),
    "ma_event_create": (C.c_int, [C.POINTER(P)]),
    "ma_event_destroy": (C.c_int, [P]),
    "ma_event_record": (C.c_int, [P, P]),
    "ma_event_sync": (C.c_int, [P]),
    "ma_event_elapsed_ms": (C.c_int, [P, P, C.POINTER(C.c_float)]),
    "ma_event_query": (C.c_int, [P]),
    "ma_host_alloc": (C.c_int, [I64, C.POINTER(P)]),
    "ma_host_free": (C.c_int, [P]),
    "ma_index_open": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(P)]),
    "ma_index_from_dense": (C.c_int, [C.c_int, I32, P, I64, P, P, I64, P, I64, P, I64, C.POINTER(P)]),
    "ma_index_build": (C.c_int, [C.c_int, I32, P, I64, P, I64, C.POINTER(P)]),
    "ma_build_opts_default": (None, [C.POINTER(BuildOpts)]),
    "ma_index_build_opts": (C.c_int, [C.c_int, I32, P, I64, P, I64, C.POINTER(BuildOpts), C.POINTER(P)]),
    "ma_index_stats": (C.c_int, [P, C.POINTER(I64), C.POINTER(I64), C.POINTER(I64), C.POINTER(I64),
                                 C.POINTER(I64), C.POINTER(I32)]),
    "ma_index_hash_stats": (C.c_int, [P, C.POINTER(I64), C.POINTER(I64), C.POINTER(I32), C.POINTER(I32)]),
    "ma_index_export_dense": (C.c_int, [P, P, P, P]),
    "ma_index_close": (C.c_int, [P]),
    "ma_workspace_create": (C.c_int, [P, I64, I32, C.POINTER(P)]),
    "ma_workspace_destroy": (C.c_int, [P]),
    "ma_workspace_set_persistent": (C.c_int, [P, I32]),
    "ma_workspace_tune": (C.c_int, [P, I32, I32]),
    "ma_workspace_bytes": (I64, [I64, I32]),
    "ma_map_forward": (C.c_int, [P, P, I64, P, I64, I32, I32, I32, I64, C.POINTER(Params), C.POINTER(Recs)]),
    "ma_map_forward_dev": (C.c_int, [P, P, P, P, I64, P, I64, I32, I32, I32, I64, C.POINTER(Params),
                                     C.POINTER(Recs)]),
    "ma_map_batch": (C.c_int, [P, P, I64, P, I64, I32, I32, I32, I32, I64, C.POINTER(Params),
                               C.POINTER(Recs), C.POINTER(Recs), C.POINTER(Recs)]),
    "ma_map_batch_dev": (C.c_int, [P, P, P, P, I64, P, I64, I32, I32, I32, I32, I64,
                                   C.POINTER(Params), C.POINTER(Recs), C.POINTER(Recs),
                                   C.POINTER(Recs)]),
    "ma_map_batch_dev_noalloc": (C.c_int, [P, P, P, I64, P, I64, I32, I32, I32, I32, I64,
                                           C.POINTER(Params), C.POINTER(Recs), C.POINTER(Recs),
                                           C.POINTER(Recs)]),
    "ma_sort_keys_se_dev": (C.c_int, [P, I64, C.POINTER(Recs), P, P, I32, I64, P, P, P, P, P]),
    "ma_mate_rescue_dev": (C.c_int, [P, P, P, I64, P, I64, I32, I32, I32, I32, I32, I32, C.POINTER(Recs)]),
    "ma_sort_dev_status": (C.c_int, [P, I64, P, I64]),
    "ma_tune_set": (C.c_int, [I32, I64]),
    "ma_graph_begin": (C.c_int, [P]),
    "ma_graph_end": (C.c_int, [P, C.POINTER(P), C.POINTER(I64)]),
    "ma_graph_launch": (C.c_int, [P, P]),
    "ma_graph_destroy": (C.c_int, [P]),
    "ma_tune_get": (C.c_int, [I32, C.POINTER(I64), C.POINTER(I64)]),
    "ma_sort_markdup": (C.c_int, [C.c_int, I64, P, P, P, P, P, P, P, I32, P]),
    "ma_sort_scratch_bytes": (I64, [I64]),
    "ma_sort_keys_pe_dev": (C.c_int, [P, I64, C.POINTER(Recs), P, P, I32, I64, P, P, P, P, P]),
    "ma_sort_markdup_dev": (C.c_int, [P, I64, P, P, P, P, P, P, P, I32, P, P, I64]),
    "ma_radix_sort_u64_dev": (C.c_int, [P, I64, P, P, P, I64]),
    "ma_seed_kmers": (C.c_int, [C.c_int, P, I64, I32, I32, P]),
    "ma_bam_scratch_bytes": (I64, [I64]),
    "ma_bam_record_bound": (I64, [I32, I32, I32]),
    "ma_bam_pack_se_dev": (C.c_int, [P, I64, C.POINTER(Recs), P, I32, C.POINTER(BamReads), C.c_char_p,
                                     I32, P, P, I32, P, I64, P, P, I64]),
    "ma_bam_pack_pe_dev": (C.c_int, [P, I64, C.POINTER(Recs), P, I32, C.POINTER(BamReads), C.c_char_p,
                                     I32, P, P, I32, P, I64, P, P, I64]),
    "ma_bam_pack_dev": (C.c_int, [P, I64, C.POINTER(BamFields), C.POINTER(BamReads), C.c_char_p, I32, P,
                                  P, I32, P, I64, P, P, I64]),
    "ma_fastq_scratch_bytes": (I64, [I64, I64]),
    "ma_fastq_parse_dev": (C.c_int, [P, P, I64, I32, I64, I32, I32, P, P, P, P, P, P, I64, P, P, P, I64]),
    "ma_bgzf_bound": (I64, [I64]),
    "ma_bgzf_dev_scratch_bytes": (I64, [I64]),
    "ma_bgzf_compress_dev": (C.c_int, [P, P, I64, I32, I32, P, I64, P, P, I64]),
    "ma_bgzf_compress": (C.c_int, [P, I64, I32, I32, I32, P, I64, C.POINTER(I64)]),
    "ma_bgzf_inflate": (C.c_int, [P, I64, I32, P, I64, C.POINTER(I64), C.POINTER(I64)]),
    "ma_gz_open": (C.c_int, [C.c_char_p, C.POINTER(P)]),
    "ma_gz_read": (C.c_int, [P, P, I64, C.POINTER(I64)]),
    "ma_gz_close": (C.c_int, [P]),
    "ma_bgzf_open": (C.c_int, [C.c_char_p, I32, I32, C.POINTER(P)]),
    "ma_bgzf_write": (C.c_int, [P, P, I64]),
    "ma_bgzf_close": (C.c_int, [P, C.POINTER(I64)]),
    "ma_serve_open": (C.c_int, [P, I32, I64, I32, C.POINTER(P)]),
    "ma_serve_tune": (C.c_int, [P, I64, C.c_double, I32, I32]),
    "ma_serve_submit": (I64, [P, P, I64, P, I64, I32, I32, I32, I32, I64, C.POINTER(Params),
                              C.POINTER(Recs)]),
    "ma_serve_submit_sorted": (I64, [P, P, I64, P, I64, I32, I32, I32, I32, I64, C.POINTER(Params),
                                     C.POINTER(Recs), I32, I32, P, P]),
    "ma_serve_set_targets": (C.c_int, [P, P, I32]),
    "ma_serve_memory": (C.c_int, [P, C.POINTER(I64), C.POINTER(I64), C.POINTER(I64)]),
    "ma_serve_poll": (C.c_int, [P, I64, I32, C.POINTER(C.c_double)]),
    "ma_serve_reap": (C.c_int, [P, C.POINTER(I64), I32, I32]),
    "ma_serve_close": (C.c_int, [P]),
}

_lib = None


def lib() -> C.CDLL:
    """Load libmialign.so once; raise if it is missing (build it first)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.is_file():
            raise MialignError(
                f"{LIB_PATH} is missing: run __graft_entry__.build() (hipcc, gfx950) first"
            )
        try:
            handle = C.CDLL(str(LIB_PATH), mode=os.RTLD_NOW | os.RTLD_GLOBAL)
        except OSError as e:  # pragma: no cover - loader failure text is platform specific
            raise MialignError(f"cannot load {LIB_PATH}: {e}") from e
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def last_error() -> str:
    msg = lib().ma_last_error()
    return msg.decode() if msg else ""


def check(rc: int, what: str) -> int:
    if rc < 0:
        raise MialignError(f"{what} failed ({rc}): {last_error()}")
    return rc


def default_params(**over) -> Params:
    p = Params()
    lib().ma_params_default(C.byref(p))
    for k, v in over.items():
        setattr(p, k, int(v))
    return p


def device_count() -> int:
    n = C.c_int(0)
    lib().ma_device_count(C.byref(n))
    return n.value


def ptr(a: np.ndarray | None) -> int | None:
    if a is None:
        return None
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return a.ctypes.data


def host_recs(n: int) -> dict[str, np.ndarray]:
    return {f: np.zeros(n, dtype=REC_DTYPES[f]) for f in REC_FIELDS}


def recs_struct(arrs: dict[str, np.ndarray] | None) -> Recs | None:
    if arrs is None:
        return None
    r = Recs()
    for f in REC_FIELDS:
        a = arrs.get(f)
        setattr(r, f, ptr(a) if a is not None else None)
    return r
