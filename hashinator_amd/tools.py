"""split::tools over the C-ABI: stable compaction and exclusive prefix scan.

Mirrors kstppd/hashinator include/splitvector/split_tools.h: ``copy_if``
(1128-1223), ``copy_keys_if`` (1094-1191), ``split_prefix_scan`` (363-418) and
``estimateMemoryForCompaction`` (1064-1089).  A user functor cannot cross a C
ABI; the built-in predicates (``hn_pred``) cover the reference's own rules,
and any other rule is evaluated into a flags tensor by the caller and passed
to :func:`copy_if_flags`.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib
from ._lib import check, hn_pred
from .hashmap import _ptr, _stream, on_stream

_TDT = {32: torch.int32, 64: torch.int64}


def estimateMemoryForCompaction(n: int) -> int:
    """Scratch bytes for one compaction or scan of n elements."""
    return int(_lib.load().hn_compact_scratch_bytes(n))


def _scratch(n, device, scratch):
    need = estimateMemoryForCompaction(n)
    if scratch is None:
        return torch.empty(need, dtype=torch.uint8, device=device), need
    if scratch.numel() * scratch.element_size() < need:
        raise ValueError(f"scratch of {scratch.numel() * scratch.element_size()} B < {need} B needed")
    return scratch, scratch.numel() * scratch.element_size()


def pred(kind=_lib.HN_PRED_VALID, field=0, mask=0, value=0) -> hn_pred:
    return hn_pred(kind, field, mask, value)


@on_stream
def copy_if_pairs(pairs: torch.Tensor, p: hn_pred, key_bits: int, stream=None, out=None, scratch=None,
                  sync=True):
    """Stable copy_if over an (n, 2) tensor of hash_pair<K,K>.  Returns the
    selected pairs (a view of ``out``) when ``sync``, else (out, d_count)."""
    L = _lib.load()
    n = pairs.shape[0]
    out = out if out is not None else torch.empty_like(pairs)
    cnt = torch.zeros(1, dtype=torch.int64, device=pairs.device)
    sc, sb = _scratch(n, pairs.device, scratch)
    fn = L.hn_copy_if_pairs_u32 if key_bits == 32 else L.hn_copy_if_pairs_u64
    check(fn(_ptr(pairs), n, C.byref(p), _ptr(out), _ptr(cnt), _ptr(sc), sb, _stream(stream)), "copy_if")
    if not sync:
        return out, cnt
    return out[: int(cnt.item())]


@on_stream
def copy_keys_if_pairs(pairs: torch.Tensor, p: hn_pred, key_bits: int, stream=None, out=None, scratch=None,
                       sync=True):
    """copy_keys_if: the .first of every selected pair, in input order."""
    L = _lib.load()
    n = pairs.shape[0]
    out = out if out is not None else torch.empty(n, dtype=_TDT[key_bits], device=pairs.device)
    cnt = torch.zeros(1, dtype=torch.int64, device=pairs.device)
    sc, sb = _scratch(n, pairs.device, scratch)
    fn = L.hn_copy_keys_if_pairs_u32 if key_bits == 32 else L.hn_copy_keys_if_pairs_u64
    check(fn(_ptr(pairs), n, C.byref(p), _ptr(out), _ptr(cnt), _ptr(sc), sb, _stream(stream)), "copy_keys_if")
    if not sync:
        return out, cnt
    return out[: int(cnt.item())]


@on_stream
def copy_if_flags(elems: torch.Tensor, flags: torch.Tensor, stream=None, out=None, scratch=None, sync=True):
    """copy_if with the rule pre-evaluated: keep elems[i] where flags[i] != 0.
    ``elems`` is (n,) of 4- or 8-byte items or (n, 2) of 8-byte pairs."""
    L = _lib.load()
    n = elems.shape[0]
    eb = elems.element_size() * (elems.shape[1] if elems.dim() == 2 else 1)
    out = out if out is not None else torch.empty_like(elems)
    cnt = torch.zeros(1, dtype=torch.int64, device=elems.device)
    sc, sb = _scratch(n, elems.device, scratch)
    f = flags.to(torch.uint8) if flags.dtype != torch.uint8 else flags
    check(L.hn_copy_if_flags(_ptr(elems), _ptr(f), n, eb, _ptr(out), _ptr(cnt), _ptr(sc), sb, _stream(stream)),
          "copy_if(flags)")
    if not sync:
        return out, cnt
    return out[: int(cnt.item())]


@on_stream
def split_prefix_scan(inp: torch.Tensor, out: torch.Tensor | None = None, stream=None, scratch=None, sync=True):
    """Exclusive prefix sum with wraparound (uint32 via int32 bit patterns, or uint64 via int64)."""
    L = _lib.load()
    n = inp.numel()
    out = out if out is not None else torch.empty_like(inp)
    sc, sb = _scratch(n, inp.device, scratch)
    if inp.element_size() == 4:
        fn = L.hn_exclusive_scan_u32
    elif inp.element_size() == 8:
        fn = L.hn_exclusive_scan_u64
    else:
        raise TypeError("scan supports 4- and 8-byte elements")
    check(fn(_ptr(inp), _ptr(out), n, _ptr(sc), sb, _stream(stream)), "split_prefix_scan")
    if sync:
        torch.cuda.current_stream().synchronize()
    return out


# reference-style aliases
copy_if = copy_if_pairs
copy_keys_if = copy_keys_if_pairs
