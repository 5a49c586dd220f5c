"""Many independent maps: the one-launch batched path and the per-map path.

``BatchedMaps`` lays out a set of maps (sizes 2^sp each) and their per-map
inputs in device memory and runs, for every map, insert -> retrieve -> erase
-> clean_tombstones either

* batched: one persistent launch (``hn_batch_run_*``), graph-capturable, or
* per map: the plain C-ABI calls, one launch (or a few) per map per operation,
  which is how the reference's API is used (hashers.h:46-163).

Both write the same outputs (``qry_vals``/``qry_hit``) and leave every map in
a state with the same key/value set.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from ._lib import check, hn_batch_map, hn_tune

_TDT = {32: torch.int32, 64: torch.int64}
_NDT = {32: np.uint32, 64: np.uint64}


def _p(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def _concat(parts, dtype):
    lens = np.array([len(p) for p in parts], dtype=np.uint64)
    offs = np.concatenate([[0], np.cumsum(lens)[:-1]]).astype(np.uint64) if len(parts) else np.zeros(0, np.uint64)
    flat = np.concatenate(parts).astype(dtype) if len(parts) and lens.sum() else np.zeros(0, dtype)
    return flat, offs, lens


class BatchedMaps:
    """A batch of Hashmap<K,K> (K = uint32 or uint64) with per-map inputs.

    sps[m]        size power of map m
    ins_keys[m]   keys to insert (unique within the map), ins_vals[m] values
    qry_keys[m]   keys to retrieve after the insert
    ers_keys[m]   keys to erase after the retrieve
    """

    def __init__(self, sps, ins_keys, ins_vals, qry_keys, ers_keys, key_bits=32, device="cuda",
                 largest_first=True):
        self.kb = key_bits
        self.dev = torch.device(device)
        self.L = _lib.load()
        self.sfx = "u32" if key_bits == 32 else "u64"
        self.n = len(sps)
        self.sps = np.asarray(sps, dtype=np.int32)
        nd = _NDT[key_bits]
        ik, self.ins_off, self.ins_len = _concat(list(ins_keys), nd)
        iv, _, _ = _concat(list(ins_vals), nd)
        qk, self.qry_off, self.qry_len = _concat(list(qry_keys), nd)
        ek, self.ers_off, self.ers_len = _concat(list(ers_keys), nd)
        iview = np.int32 if key_bits == 32 else np.int64
        self.ins_keys = torch.from_numpy(ik.view(iview)).to(self.dev)
        self.ins_vals = torch.from_numpy(iv.view(iview)).to(self.dev)
        self.qry_keys = torch.from_numpy(qk.view(iview)).to(self.dev)
        self.ers_keys = torch.from_numpy(ek.view(iview)).to(self.dev)
        self.qry_vals = torch.zeros_like(self.qry_keys)
        self.qry_hit = torch.zeros(self.qry_keys.numel(), dtype=torch.uint8, device=self.dev)
        # bucket arena and Info array
        slots = (np.ones(self.n, dtype=np.uint64) << self.sps.astype(np.uint64))
        self.slot_off = np.concatenate([[0], np.cumsum(slots)[:-1]]).astype(np.uint64)
        self.total_slots = int(slots.sum())
        self.arena = torch.empty((self.total_slots, 2), dtype=_TDT[key_bits], device=self.dev)
        self.infos = torch.zeros((self.n, 5), dtype=torch.int64, device=self.dev)
        # clean scratch: a slot per possibly-live element (maps start empty)
        self.scr_off = np.concatenate([[0], np.cumsum(self.ins_len)[:-1]]).astype(np.uint64) if self.n else \
            np.zeros(0, np.uint64)
        self.clean_scratch = torch.empty((max(1, int(self.ins_len.sum())), 2), dtype=_TDT[key_bits],
                                         device=self.dev)
        self.scratch = torch.zeros(int(self.L.hn_batch_scratch_bytes(self.n)), dtype=torch.uint8, device=self.dev)
        self.order = np.argsort(-self.sps, kind="stable") if largest_first else np.arange(self.n)
        slot_bytes = 8 if key_bits == 32 else 16
        descs = (hn_batch_map * self.n)()
        for j, m in enumerate(self.order):
            d = descs[j]
            d.buckets = self.arena.data_ptr() + int(self.slot_off[m]) * slot_bytes
            d.info = self.infos.data_ptr() + int(m) * 40
            d.ins_off, d.ins_len = int(self.ins_off[m]), int(self.ins_len[m])
            d.qry_off, d.qry_len = int(self.qry_off[m]), int(self.qry_len[m])
            d.ers_off, d.ers_len = int(self.ers_off[m]), int(self.ers_len[m])
            d.scratch_off, d.scratch_cap = int(self.scr_off[m]), int(self.ins_len[m])
        raw = np.frombuffer(bytes(descs), dtype=np.uint8).copy() if self.n else np.zeros(0, np.uint8)
        self.descs = torch.from_numpy(raw).to(self.dev)
        self.sps_dev = torch.from_numpy(self.sps[self.order].copy()).to(self.dev)
        self._percall_scratch = None

    # ------------------------------------------------------------ device steps (asynchronous)
    def _s(self, stream):
        return C.c_void_p((stream or torch.cuda.current_stream()).cuda_stream)

    def init(self, stream=None):
        """Every map: a fresh table and Info(sp), one launch."""
        check(getattr(self.L, f"hn_batch_init_{self.sfx}")(_p(self.descs), self.n, _p(self.sps_dev),
                                                           self._s(stream)), "hn_batch_init")

    def run(self, ops=_lib.HN_BATCH_ALL, tune: hn_tune | None = None, stream=None):
        """All requested phases on every map in one persistent launch."""
        check(getattr(self.L, f"hn_batch_run_{self.sfx}")(
            _p(self.descs), self.n, ops, _p(self.ins_keys), _p(self.ins_vals), _p(self.qry_keys),
            _p(self.qry_vals), _p(self.qry_hit), _p(self.ers_keys), _p(self.clean_scratch), _p(self.scratch),
            self._s(stream), C.byref(tune) if tune is not None else None), "hn_batch_run")

    def init_per_map(self, stream=None):
        """The per-map way to start: one table-init and one Info-init launch per map."""
        s = self._s(stream)
        slot_bytes = 8 if self.kb == 32 else 16
        tinit = getattr(self.L, f"hn_table_init_{self.sfx}")
        for m in range(self.n):
            check(tinit(C.c_void_p(self.arena.data_ptr() + int(self.slot_off[m]) * slot_bytes),
                        1 << int(self.sps[m]), s), "table_init")
            check(self.L.hn_info_init(C.c_void_p(self.infos.data_ptr() + m * 40), int(self.sps[m]), s), "info_init")

    def run_per_map(self, ops=_lib.HN_BATCH_ALL, tune: hn_tune | None = None, stream=None):
        """The same phases through the plain single-map C-ABI, map by map."""
        s = self._s(stream)
        L, sfx = self.L, self.sfx
        slot_bytes = 8 if self.kb == 32 else 16
        es = self.ins_keys.element_size()
        t = C.byref(tune) if tune is not None else None
        if ops & _lib.HN_BATCH_CLEAN and self._percall_scratch is None:
            sb = int(getattr(L, f"hn_clean_tombstones_scratch_bytes_{sfx}")(int(self.sps.max()), 0))
            self._percall_scratch = torch.empty(sb, dtype=torch.uint8, device=self.dev)
        for m in self.order:
            b = C.c_void_p(self.arena.data_ptr() + int(self.slot_off[m]) * slot_bytes)
            info = C.c_void_p(self.infos.data_ptr() + int(m) * 40)
            if ops & _lib.HN_BATCH_INSERT:
                o, n = int(self.ins_off[m]), int(self.ins_len[m])
                check(getattr(L, f"hn_insert_{sfx}")(C.c_void_p(self.ins_keys.data_ptr() + o * es),
                                                     C.c_void_p(self.ins_vals.data_ptr() + o * es), b, info, n, s, t),
                      "insert")
            if ops & _lib.HN_BATCH_RETRIEVE:
                o, n = int(self.qry_off[m]), int(self.qry_len[m])
                check(getattr(L, f"hn_retrieve_{sfx}")(C.c_void_p(self.qry_keys.data_ptr() + o * es),
                                                       C.c_void_p(self.qry_vals.data_ptr() + o * es),
                                                       C.c_void_p(self.qry_hit.data_ptr() + o), b, info, n, s, t),
                      "retrieve")
            if ops & _lib.HN_BATCH_ERASE:
                o, n = int(self.ers_off[m]), int(self.ers_len[m])
                check(getattr(L, f"hn_erase_{sfx}")(C.c_void_p(self.ers_keys.data_ptr() + o * es), b, info, n, s, t),
                      "erase")
            if ops & _lib.HN_BATCH_CLEAN:
                sb = self._percall_scratch.numel()
                check(getattr(L, f"hn_clean_tombstones_{sfx}")(b, info, int(self.sps[m]),
                                                               _p(self._percall_scratch), sb, s), "clean")

    # ------------------------------------------------------------ host views (synchronising)
    def info(self, m):
        return _lib.hn_info.from_buffer_copy(self.infos[m].cpu().numpy().tobytes())

    def all_infos(self):
        raw = self.infos.cpu().numpy()
        return raw  # (n, 5) int64: sp|reserved, fill, overflow, tombstones, err|reserved

    def buckets(self, m):
        lo = int(self.slot_off[m])
        hi = lo + (1 << int(self.sps[m]))
        return self.arena[lo:hi].cpu().numpy().view(_NDT[self.kb])

    def results(self, m):
        lo, n = int(self.qry_off[m]), int(self.qry_len[m])
        v = self.qry_vals[lo:lo + n].cpu().numpy().view(_NDT[self.kb])
        h = self.qry_hit[lo:lo + n].cpu().numpy().astype(bool)
        return v, h

    def counts(self):
        return int(self.ins_len.sum()), int(self.qry_len.sum()), int(self.ers_len.sum())
