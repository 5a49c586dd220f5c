"""ctypes binding of libhashinator_amd.so (the C-ABI in include/hashinator_amd.h).

The library must have been built in-tree (``python -m hashinator_amd.build``);
there is no fallback: loading fails loudly when it is missing.
"""
from __future__ import annotations

import ctypes as C
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libhashinator_amd.so")

HN_STATUS_SUCCESS, HN_STATUS_FAIL, HN_STATUS_INVALID = 0, 1, 2
HN_INFO_ORDERED = 2
HN_BUCKET_OVERFLOW = 64
HN_EINVAL, HN_ESCRATCH, HN_ENOLIB = -1, -2, -3
HN_PRED_ALL, HN_PRED_VALID, HN_PRED_MASKEQ, HN_PRED_LT, HN_PRED_GE = 0, 1, 2, 3, 4


class hn_info(C.Structure):
    """Byte-for-byte Hashinator::Info (defaults.h:44-54)."""

    _fields_ = [
        ("size_power", C.c_int32),
        ("reserved0", C.c_int32),
        ("fill", C.c_uint64),
        ("current_max_bucket_overflow", C.c_uint64),
        ("tombstone_counter", C.c_uint64),
        ("err", C.c_int32),
        ("reserved1", C.c_int32),
    ]


class hn_tune(C.Structure):
    _fields_ = [
        ("probe_width", C.c_int32),
        ("block_size", C.c_int32),
        ("grid_blocks", C.c_int32),
        ("reserved", C.c_int32),
    ]


class hn_batch_map(C.Structure):
    _fields_ = [
        ("buckets", C.c_void_p),
        ("info", C.c_void_p),
        ("ins_off", C.c_uint64),
        ("ins_len", C.c_uint64),
        ("qry_off", C.c_uint64),
        ("qry_len", C.c_uint64),
        ("ers_off", C.c_uint64),
        ("ers_len", C.c_uint64),
        ("scratch_off", C.c_uint64),
        ("scratch_cap", C.c_uint64),
    ]


HN_BATCH_INSERT, HN_BATCH_RETRIEVE, HN_BATCH_ERASE, HN_BATCH_CLEAN = 1, 2, 4, 8
HN_BATCH_ALL = 15


class hn_pred(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("field", C.c_int32),
        ("mask", C.c_uint64),
        ("value", C.c_uint64),
    ]


assert C.sizeof(hn_info) == 40

P = C.c_void_p
S = C.c_size_t
I = C.c_int
TUNE = C.POINTER(hn_tune)
PRED = C.POINTER(hn_pred)

# name -> (restype, argtypes); every entry point declared in include/hashinator_amd.h
SIGNATURES = {
    "hn_info_init": (I, [P, I, P]),
    "hn_table_init_u32": (I, [P, S, P]),
    "hn_table_init_u64": (I, [P, S, P]),
    "hn_insert_u32": (I, [P, P, P, P, S, P, TUNE]),
    "hn_insert_u64": (I, [P, P, P, P, S, P, TUNE]),
    "hn_insert_index_u32": (I, [P, P, P, S, P, TUNE]),
    "hn_insert_index_u64": (I, [P, P, P, S, P, TUNE]),
    "hn_insert_pairs_u32": (I, [P, P, P, S, P, TUNE]),
    "hn_insert_pairs_u64": (I, [P, P, P, S, P, TUNE]),
    "hn_retrieve_u32": (I, [P, P, P, P, P, S, P, TUNE]),
    "hn_retrieve_u64": (I, [P, P, P, P, P, S, P, TUNE]),
    "hn_retrieve_pairs_u32": (I, [P, P, P, P, S, P, TUNE]),
    "hn_retrieve_pairs_u64": (I, [P, P, P, P, S, P, TUNE]),
    "hn_erase_u32": (I, [P, P, P, S, P, TUNE]),
    "hn_erase_u64": (I, [P, P, P, S, P, TUNE]),
    "hn_reset_all_u32": (I, [P, P, S, P]),
    "hn_reset_all_u64": (I, [P, P, S, P]),
    "hn_reset_u32": (I, [P, P, P, S, P]),
    "hn_reset_u64": (I, [P, P, P, S, P]),
    "hn_clean_tombstones_scratch_bytes_u32": (S, [I, S]),
    "hn_clean_tombstones_scratch_bytes_u64": (S, [I, S]),
    "hn_clean_tombstones_min_scratch": (S, []),
    "hn_clean_tombstones_u32": (I, [P, P, I, P, S, P]),
    "hn_clean_tombstones_u64": (I, [P, P, I, P, S, P]),
    "hn_rehash_scratch_bytes": (S, []),
    "hn_rehash_u32": (I, [P, I, P, I, P, P, P]),
    "hn_rehash_u64": (I, [P, I, P, I, P, P, P]),
    "hn_reorder_scratch_bytes": (S, [I]),
    "hn_reorder_u32": (I, [P, P, I, P, S, P]),
    "hn_reorder_u64": (I, [P, P, I, P, S, P]),
    "hn_build_scratch_bytes_u32": (S, [I, S]),
    "hn_build_scratch_bytes_u64": (S, [I, S]),
    "hn_build_u32": (I, [P, P, P, P, S, I, P, S, P, TUNE]),
    "hn_build_u64": (I, [P, P, P, P, S, I, P, S, P, TUNE]),
    "hn_build_pairs_u32": (I, [P, P, P, S, I, P, S, P, TUNE]),
    "hn_build_pairs_u64": (I, [P, P, P, S, I, P, S, P, TUNE]),
    "hn_rehash_build_scratch_bytes_u32": (S, [I, I]),
    "hn_rehash_build_scratch_bytes_u64": (S, [I, I]),
    "hn_rehash_build_u32": (I, [P, I, P, I, P, P, S, P]),
    "hn_rehash_build_u64": (I, [P, I, P, I, P, P, S, P]),
    "hn_merge_scratch_bytes_u32": (S, [I, S]),
    "hn_merge_scratch_bytes_u64": (S, [I, S]),
    "hn_merge_u32": (I, [P, P, P, P, P, S, I, P, S, P]),
    "hn_merge_u64": (I, [P, P, P, P, P, S, I, P, S, P]),
    "hn_merge_pairs_u32": (I, [P, P, P, P, S, I, P, S, P]),
    "hn_merge_pairs_u64": (I, [P, P, P, P, S, I, P, S, P]),
    "hn_compact_scratch_bytes": (S, [S]),
    "hn_copy_if_pairs_u32": (I, [P, S, PRED, P, P, P, S, P]),
    "hn_copy_if_pairs_u64": (I, [P, S, PRED, P, P, P, S, P]),
    "hn_copy_if_pairs_noalloc_u64": (I, [P, S, PRED, P, P, P]),
    "hn_copy_keys_if_pairs_u32": (I, [P, S, PRED, P, P, P, S, P]),
    "hn_copy_keys_if_pairs_u64": (I, [P, S, PRED, P, P, P, S, P]),
    "hn_copy_if_flags": (I, [P, P, S, S, P, P, P, S, P]),
    "hn_copy_keys_if_flags_u32": (I, [P, P, S, P, P, P, S, P]),
    "hn_copy_if_bits": (I, [P, P, S, S, P, P, P, S, P]),
    "hn_copy_keys_if_bits_u32": (I, [P, P, S, P, P, P, S, P]),
    "hn_copy_keys_if_bits_u64": (I, [P, P, S, P, P, P, S, P]),
    "hn_copy_keys_if_flags_u64": (I, [P, P, S, P, P, P, S, P]),
    "hn_exclusive_scan_u32": (I, [P, P, S, P, S, P]),
    "hn_exclusive_scan_u64": (I, [P, P, S, P, S, P]),
    "hn_batch_scratch_bytes": (S, [C.c_uint32]),
    "hn_batch_init_u32": (I, [P, C.c_uint32, P, P]),
    "hn_batch_init_u64": (I, [P, C.c_uint32, P, P]),
    "hn_batch_run_u32": (I, [P, C.c_uint32, C.c_uint32, P, P, P, P, P, P, P, P, P, TUNE]),
    "hn_batch_run_u64": (I, [P, C.c_uint32, C.c_uint32, P, P, P, P, P, P, P, P, P, TUNE]),
    "hn_version": (C.c_char_p, []),
    # include/hashinator_amd_calib.h
    "hn_calib_random_gather": (I, [P, S, I, S, I, C.c_uint64, P, P]),
    "hn_calib_random_gather_lane": (I, [P, S, I, S, C.c_uint64, P, P]),
    "hn_calib_stream_copy": (I, [P, P, S, P]),
}

_LIB = None


class HnError(RuntimeError):
    pass


def load():
    """Load the in-tree library (raises if it was not built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise HnError(
                f"{LIB_PATH} not found: build it with `python -m hashinator_amd.build` "
                "(there is no CPU fallback)"
            )
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _LIB = lib
    return _LIB


def check(rc: int, what: str = "") -> None:
    """SPLIT_CHECK_ERR-style: a nonzero return is an error (split_allocators.h:41-50)."""
    if rc != 0:
        names = {HN_EINVAL: "HN_EINVAL", HN_ESCRATCH: "HN_ESCRATCH", HN_ENOLIB: "HN_ENOLIB"}
        raise HnError(f"{what} failed: {names.get(rc, f'hipError {rc}')}")
