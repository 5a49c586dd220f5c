"""Host-side mirror of Hashinator's accelerated map API over the C-ABI.

``Hashmap`` follows ``Hashinator::Hashmap<K,V>`` (kstppd/hashinator
include/hashinator/hashinator.h:59-1802): same method names, argument meaning
and synchronous behaviour (every accelerated call ends with a stream
synchronize, hashers.h:55-161).  Device memory is torch tensors; torch is the
allocator and stream source only, every operation runs the gfx950 kernels of
libhashinator_amd.so.

Keys and values are unsigned integers of 32 or 64 bits; torch tensors of
int32/int64 (or uint32/uint64) carry their bit patterns.
"""
from __future__ import annotations

import contextlib
import ctypes as C
import functools
import inspect
import math

import numpy as np
import torch

from . import _lib
from ._lib import check, hn_info, hn_pred, hn_tune

EMPTY = {32: 0xFFFFFFFF, 64: 0xFFFFFFFFFFFFFFFF}
TOMBSTONE = {32: 0xFFFFFFFE, 64: 0xFFFFFFFFFFFFFFFE}
_TDT = {32: torch.int32, 64: torch.int64}
_NDT = {32: np.uint32, 64: np.uint64}
_SIZE_MAX = (1 << 64) - 1


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def _stream(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def on_stream(fn):
    """Run a method with its ``stream`` argument as torch's current stream.

    Then every allocation of the call comes from that stream's pool, every
    table initialisation and Info write is enqueued on it, and every read
    back (``.cpu()``, ``.item()``) waits for it: the kernels launched on
    ``stream`` and the work around them are ordered by the stream itself.
    """
    idx = list(inspect.signature(fn).parameters).index("stream")

    @functools.wraps(fn)
    def call(*a, **kw):
        s = kw["stream"] if "stream" in kw else (a[idx] if len(a) > idx else None)
        ctx = torch.cuda.stream(s) if s is not None else contextlib.nullcontext()
        with ctx:
            return fn(*a, **kw)

    return call


def _same_device(a: torch.device, b: torch.device) -> bool:
    if a.type != b.type:
        return False
    if a.type != "cuda":
        return True
    ia = a.index if a.index is not None else torch.cuda.current_device()
    ib = b.index if b.index is not None else torch.cuda.current_device()
    return ia == ib


def out_tensor(t, n, dtype, device, what):
    """Check a caller-provided output tensor before a kernel writes to it:
    on ``device``, of ``dtype`` (or its unsigned twin), contiguous, with at
    least ``n`` elements.  Returns it viewed as ``dtype``."""
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{what} must be a torch tensor, got {type(t).__name__}")
    twin = {torch.int32: torch.uint32, torch.int64: torch.uint64}.get(dtype)
    if t.dtype == twin:
        t = t.view(dtype)
    if t.dtype != dtype:
        raise TypeError(f"{what} must be {dtype}, got {t.dtype}")
    if not _same_device(t.device, torch.device(device)):
        raise ValueError(f"{what} is on {t.device}, the map is on {device}")
    if not t.is_contiguous():
        raise ValueError(f"{what} must be contiguous")
    if t.numel() < n:
        raise ValueError(f"{what} holds {t.numel()} elements, {n} needed")
    return t


def as_device(a, key_bits, device="cuda"):
    """numpy / torch / list of unsigned integers -> contiguous device tensor of bit patterns."""
    if isinstance(a, torch.Tensor):
        t = a
        if t.dtype in (torch.uint32, torch.uint64):
            t = t.view(_TDT[key_bits])
        if t.dtype != _TDT[key_bits]:
            raise TypeError(f"expected {key_bits}-bit integer tensor, got {t.dtype}")
        return t.to(device).contiguous()
    arr = np.ascontiguousarray(np.asarray(a, dtype=_NDT[key_bits]))
    return torch.from_numpy(arr.view(np.int32 if key_bits == 32 else np.int64)).to(device)


def to_numpy(t, key_bits):
    """device tensor of bit patterns -> numpy unsigned array."""
    return t.detach().cpu().contiguous().numpy().view(_NDT[key_bits])


class targets:
    """Hashinator::targets (common.h:90)."""

    host, device, automatic = 0, 1, 2


class status:
    """Hashinator::status (common.h:85)."""

    success, fail, invalid = 0, 1, 2


class Hashmap:
    """Hashinator::Hashmap<K,V> with K = V = uint32_t or uint64_t, accelerated API."""

    def __init__(self, sizePower: int = 5, key_bits: int = 32, device="cuda", tune: hn_tune | None = None,
                 auto_reorder: float = 1.0 / 16, auto_build: float = 1.0 / 16, build_min_keys: int = 1 << 17,
                 merge_min: float = 0.25):
        """``auto_reorder``: after an insert of at least this fraction of the
        bucket count, and after clean_tombstones and device_rehash, run the
        Robin Hood reorder pass (``reorder()``), which makes later misses
        stop early; 0 turns it off.  ``auto_build``: an insert of at least
        this fraction of the bucket count into an empty table takes the bulk
        path (hn_build_*: partition + per-region sort, no probing), which
        writes the reordered image directly; 0 turns it off.
        ``build_min_keys``: and holds at least this many keys -- the build's
        dozen small launches take about 0.09 ms however few the keys, so
        below about 2^17 keys at low load factors the probing insert and
        reorder are faster (one MI355X, u32: 0.088 against 0.090 ms at
        131,072 keys and LF 0.5, 0.10 against 0.19 ms at 235,929 keys and LF
        0.9, 0.10 against 0.26 ms at 524,288 keys and LF 0.5).  Growing a
        map of at least this many keys rehashes it by the build too.
        ``merge_min``: an insert of at least this fraction of the bucket count
        (and of ``build_min_keys``) into a table that holds keys and no
        tombstones takes the merge (hn_merge_*), which writes the insert +
        reorder image into a fresh bucket array; the merge reads and writes
        the whole table (about 7 ms at 2^28 u64 slots, against 0.18 ns per
        key for the probing insert there), so it pays from about a quarter of
        the table.  A table not in Robin Hood order is reordered first when
        the batch is at least 4/3 of this fraction.  0 turns it off."""
        if key_bits not in (32, 64):
            raise ValueError("key_bits must be 32 or 64")
        self.kb = key_bits
        self.auto_reorder = auto_reorder
        self.auto_build = auto_build
        self.build_min_keys = build_min_keys
        self.merge_min = merge_min
        self.device = torch.device(device)
        self.tune = tune
        self._L = _lib.load()
        self._sfx = "u32" if key_bits == 32 else "u64"
        self._sp = int(sizePower)
        self.buckets = self._new_table(self._sp)
        self._info = torch.zeros(5, dtype=torch.int64, device=self.device)  # 40 bytes == Info
        check(self._L.hn_info_init(_ptr(self._info), self._sp, _stream()), "hn_info_init")
        self._sync()

    # ----------------------------------------------------------- plumbing
    def _fn(self, name):
        return getattr(self._L, f"hn_{name}_{self._sfx}")

    def _tune(self):
        return C.byref(self.tune) if self.tune is not None else None

    def _new_table(self, sp, init=True):
        t = torch.empty((1 << sp, 2), dtype=_TDT[self.kb], device=self.device)
        if init:
            check(self._fn("table_init")(_ptr(t), 1 << sp, _stream()), "hn_table_init")
        return t

    def _sync(self, stream=None):
        (stream or torch.cuda.current_stream()).synchronize()

    def _dev(self, a):
        return as_device(a, self.kb, self.device)

    # ----------------------------------------------------------- metadata (hashinator.h:434-620)
    def info(self) -> hn_info:
        """Host copy of the map's Info (a device-to-host read, like touching managed Info)."""
        raw = self._info.cpu().numpy().tobytes()
        return hn_info.from_buffer_copy(raw)

    def info_ptr(self):
        return self._info.data_ptr()

    def bucket_ptr(self):
        return self.buckets.data_ptr()

    def getSizePower(self) -> int:
        return self._sp

    def size(self) -> int:
        return int(self.info().fill)

    def bucket_count(self) -> int:
        return 1 << self._sp

    def load_factor(self) -> float:
        return float(np.float32(self.size()) / np.float32(self.bucket_count()))

    def tombstone_count(self) -> int:
        return int(self.info().tombstone_counter)

    def tombstone_ratio(self) -> float:
        t = self.tombstone_count()
        return 0.0 if t == 0 else float(np.float32(t) / np.float32(self.bucket_count()))

    def peek_status(self) -> int:
        """Read err and reset it to invalid (hashinator.h:434-438)."""
        i = self.info()
        self._info.view(torch.int32)[8] = status.invalid
        return int(i.err)

    def get_emptybucket(self) -> int:
        return EMPTY[self.kb]

    def get_tombstone(self) -> int:
        return TOMBSTONE[self.kb]

    def stats(self) -> str:
        i = self.info()
        return (
            f"Hashinator Stats \nBucket size= {self.bucket_count()}\nFill= {i.fill}, "
            f"LoadFactor={self.load_factor():f} \nTombstones= {i.tombstone_counter}\n"
            f"Overflow= {i.current_max_bucket_overflow}\n"
        )

    def host_buckets(self) -> np.ndarray:
        """(N, 2) unsigned copy of the bucket array."""
        return to_numpy(self.buckets, self.kb).reshape(-1, 2)

    # ----------------------------------------------------------- accelerated API
    def _grow_for(self, n, targetLF, stream):
        """Auto-grow before an insert (hashinator.h:1329-1332)."""
        fill = self.size()
        needed = math.ceil(math.log2((fill + n) * (1.0 / targetLF)))
        if needed > self._sp:
            self.resize(needed, targets.device, stream)

    @on_stream
    def insert(self, keys, vals=None, targetLF: float = 0.5, stream=None):
        """insert(keys, vals, len, targetLF, s) (hashinator.h:1318-1335), or
        insert(pairs, len, targetLF, s) (1358-1374) when ``vals`` is None and
        ``keys`` is an (n, 2) array of pairs."""
        k = self._dev(keys)
        if vals is None:
            k = k.reshape(-1, 2)
            n = k.shape[0]
        else:
            v = self._dev(vals)
            n = k.numel()
            if v.numel() != n:
                raise ValueError("keys and vals differ in length")
        if n == 0:
            self._info.view(torch.int32)[8] = status.success
            return
        self._grow_for(n, targetLF, stream)
        if self._try_build(k, None if vals is None else v, n, vals is None, stream):
            return
        if self._try_merge(k, None if vals is None else v, n, vals is None, stream):
            return
        if vals is None:
            rc = self._fn("insert_pairs")(_ptr(k), _ptr(self.buckets), _ptr(self._info), n, _stream(stream), self._tune())
        else:
            rc = self._fn("insert")(_ptr(k), _ptr(v), _ptr(self.buckets), _ptr(self._info), n, _stream(stream),
                                    self._tune())
        check(rc, "insert")
        self._after_insert(n, stream)

    def _try_build(self, k, v, n, pairs, stream) -> bool:
        """Bulk path for a large insert into an empty table.  False when it
        does not apply or the device declined it (err = INVALID: the table was
        not empty and is untouched; the probing insert runs instead).  Strongly
        skewed keys are placed by the device's own probing fallback inside the
        same call (err = SUCCESS, table not marked ordered)."""
        N = self.bucket_count()
        if not self.auto_build or n < self.auto_build * N or n < self.build_min_keys or n >= N:
            return False
        i = self.info()
        if i.fill != 0 or i.tombstone_counter != 0:
            return False
        sb = getattr(self._L, f"hn_build_scratch_bytes_{self._sfx}")(self._sp, n)
        if sb == 0:
            return False
        scratch = torch.empty(sb, dtype=torch.uint8, device=self.device)
        if pairs:
            rc = self._fn("build_pairs")(_ptr(k), _ptr(self.buckets), _ptr(self._info), n, self._sp, _ptr(scratch),
                                         sb, _stream(stream), self._tune())
        else:
            rc = self._fn("build")(_ptr(k), _ptr(v) if v is not None else None, _ptr(self.buckets), _ptr(self._info),
                                   n, self._sp, _ptr(scratch), sb, _stream(stream), self._tune())
        check(rc, "build")
        self._sync(stream)
        del scratch
        return self.info().err == status.success

    def _try_merge(self, k, v, n, pairs, stream) -> bool:
        """Large insert into a table that holds keys: hn_merge_* writes the
        insert + reorder image into a fresh bucket array, which then replaces
        the old one.  Where the merge itself declines (keys crowding one
        region, say), the device writes the probing insert's result into the
        new array instead, so the insert is complete either way."""
        N = self.bucket_count()
        if not self.merge_min or n < self.merge_min * N or n < self.build_min_keys or n >= N or self._sp > 30:
            return False
        i = self.info()
        if i.fill == 0 or i.tombstone_counter != 0:
            return False
        if not (i.reserved1 & _lib.HN_INFO_ORDERED):
            if n < self.merge_min * N * 4 / 3:
                return False
            self._reorder(stream)
        sb = getattr(self._L, f"hn_merge_scratch_bytes_{self._sfx}")(self._sp, n)
        if sb == 0:
            return False
        scratch = torch.empty(sb, dtype=torch.uint8, device=self.device)
        nb = self._new_table(self._sp, init=False)  # the merge (or its fallback) writes every slot
        if pairs:
            rc = self._fn("merge_pairs")(_ptr(k), _ptr(self.buckets), _ptr(nb), _ptr(self._info), n, self._sp,
                                         _ptr(scratch), sb, _stream(stream))
        else:
            rc = self._fn("merge")(_ptr(k), _ptr(v) if v is not None else None, _ptr(self.buckets), _ptr(nb),
                                   _ptr(self._info), n, self._sp, _ptr(scratch), sb, _stream(stream))
        check(rc, "merge")
        old = self.buckets  # read by the merge until the stream is synchronised
        self.buckets = nb
        self._sync(stream)
        del old, scratch
        if self.auto_reorder and not self.ordered() and n >= self.auto_reorder * N:
            self._reorder(stream)  # the fallback inserted in arrival order
            self._sync(stream)
        return True

    def _after_insert(self, n, stream):
        if self.auto_reorder and n >= self.auto_reorder * self.bucket_count():
            self._reorder(stream)
        self._sync(stream)

    @on_stream
    def reorder(self, stream=None) -> bool:
        """Robin Hood reorder of every cluster (no counterpart in the
        reference).  Same live set and occupancy; retrieve and erase of absent
        keys stop early afterwards.  Returns whether the table is now marked
        ordered (False with tombstones present)."""
        self._reorder(stream)
        self._sync(stream)
        return self.ordered()

    def _reorder(self, stream):
        sb = self._L.hn_reorder_scratch_bytes(self._sp)
        scratch = torch.empty(sb, dtype=torch.uint8, device=self.device)
        check(self._fn("reorder")(_ptr(self.buckets), _ptr(self._info), self._sp, _ptr(scratch), sb,
                                  _stream(stream)), "reorder")
        self._scratch_keep = scratch  # freed after the next synchronisation point at the earliest

    def ordered(self) -> bool:
        """True when the table carries the Robin Hood order flag."""
        return bool(self.info().reserved1 & _lib.HN_INFO_ORDERED)

    @on_stream
    def insertIndex(self, keys, targetLF: float = 0.5, stream=None):
        """insertIndex (hashinator.h:1338-1355): value = position in ``keys``."""
        k = self._dev(keys)
        n = k.numel()
        if n == 0:
            self._info.view(torch.int32)[8] = status.success
            return
        self._grow_for(n, targetLF, stream)
        if self._try_build(k, None, n, False, stream):
            return
        if self._try_merge(k, None, n, False, stream):
            return
        check(self._fn("insert_index")(_ptr(k), _ptr(self.buckets), _ptr(self._info), n, _stream(stream),
                                       self._tune()), "insertIndex")
        self._after_insert(n, stream)

    @on_stream
    def retrieve(self, keys, vals=None, hit=None, stream=None):
        """retrieve(keys, vals, len, s) (hashinator.h:1377-1384).  ``vals`` is
        updated in place on hits and left untouched on misses; with ``vals``
        None, (keys, vals, hit) device tensors are returned, vals zero-filled.
        With ``keys`` an (n, 2) pairs tensor and ``vals`` None, the pairs
        variant (1387-1394) writes .second in place."""
        k = self._dev(keys)
        if k.dim() == 2 and vals is None:
            n = k.shape[0]
            if hit is not None:
                hit = out_tensor(hit, n, torch.uint8, self.device, "hit")
            h = hit if hit is not None else torch.empty(n, dtype=torch.uint8, device=self.device)
            check(self._fn("retrieve_pairs")(_ptr(k), _ptr(h), _ptr(self.buckets), _ptr(self._info), n,
                                             _stream(stream), self._tune()), "retrieve(pairs)")
            self._sync(stream)
            return k, h
        n = k.numel()
        if vals is not None:
            vals = out_tensor(vals, n, _TDT[self.kb], self.device, "vals")
        if hit is not None:
            hit = out_tensor(hit, n, torch.uint8, self.device, "hit")
        v = vals if vals is not None else torch.zeros(n, dtype=_TDT[self.kb], device=self.device)
        h = hit if hit is not None else torch.empty(n, dtype=torch.uint8, device=self.device)
        check(self._fn("retrieve")(_ptr(k), _ptr(v), _ptr(h), _ptr(self.buckets), _ptr(self._info), n,
                                   _stream(stream), self._tune()), "retrieve")
        self._sync(stream)
        return k, v, h

    @on_stream
    def erase(self, keys, stream=None):
        """erase(keys, len, s) (hashinator.h:1397-1409)."""
        k = self._dev(keys)
        check(self._fn("erase")(_ptr(k), _ptr(self.buckets), _ptr(self._info), k.numel(), _stream(stream),
                                self._tune()), "erase")
        self._sync(stream)

    @on_stream
    def clean_tombstones(self, stream=None, scratch_budget: int = 0):
        """clean_tombstones(s) (hashinator.h:1261-1315).  ``scratch_budget``
        bounds the scratch in bytes (0: the default, a few MB of segment lists
        plus a pool for segments too long for LDS).  A table with every slot
        occupied, and a retry after the pool ran short (err = FAIL with
        tombstones left), get room for a region clean over the whole table."""
        q = getattr(self._L, f"hn_clean_tombstones_scratch_bytes_{self._sfx}")
        i = self.info()
        if i.tombstone_counter == 0:
            return
        full = i.fill + i.tombstone_counter >= self.bucket_count()
        for attempt in range(2):
            budget = _SIZE_MAX if (full or attempt) else scratch_budget
            sb = q(self._sp, budget)
            scratch = torch.empty(sb, dtype=torch.uint8, device=self.device)
            check(self._fn("clean_tombstones")(_ptr(self.buckets), _ptr(self._info), self._sp, _ptr(scratch), sb,
                                               _stream(stream)), "clean_tombstones")
            self._sync(stream)
            del scratch
            i = self.info()
            if full or scratch_budget or i.tombstone_counter == 0 or i.err != status.fail:
                break
        if self.auto_reorder and not self.ordered():
            self._reorder(stream)  # the small-scratch region path reinserts in arrival order
            self._sync(stream)

    @on_stream
    def device_rehash(self, newSizePower: int, stream=None):
        """device_rehash(p, s) (hashinator.h:284-336)."""
        if newSizePower > 32:
            raise OverflowError("Hashmap ran into rehashing catastrophe and exceeded 32bit buckets.")
        fill = self.size()
        # insert(valid, n, targetLF=1) grows further when the elements do not fit (hashinator.h:332, 1329-1332)
        if fill > 0:
            newSizePower = max(newSizePower, math.ceil(math.log2(fill)))
        nb = self._new_table(newSizePower, init=False)  # both rehash paths write every slot of it
        # Growing a map of many keys: one bulk build from the old table's live
        # slots (Robin Hood order directly); else reinsert them by probing.
        sb = 0
        if self.auto_build and newSizePower > self._sp and fill >= self.build_min_keys:
            sb = getattr(self._L, f"hn_rehash_build_scratch_bytes_{self._sfx}")(self._sp, newSizePower)
        if sb:
            scratch = torch.empty(sb, dtype=torch.uint8, device=self.device)
            check(self._fn("rehash_build")(_ptr(self.buckets), self._sp, _ptr(nb), newSizePower, _ptr(self._info),
                                           _ptr(scratch), sb, _stream(stream)), "device_rehash")
        else:
            scratch = torch.empty(self._L.hn_rehash_scratch_bytes(), dtype=torch.uint8, device=self.device)
            check(self._fn("rehash")(_ptr(self.buckets), self._sp, _ptr(nb), newSizePower, _ptr(self._info),
                                     _ptr(scratch), _stream(stream)), "device_rehash")
        old = self.buckets  # read by the rehash until the stream is synchronised
        self.buckets = nb
        self._sp = newSizePower
        if sb:
            self._sync(stream)  # the order flag is read next (skewed keys leave the table unordered)
        if self.auto_reorder and not (sb and self.ordered()):
            self._reorder(stream)
        self._sync(stream)
        del old, scratch

    @on_stream
    def resize(self, newSizePower: int, t: int = targets.device, stream=None):
        """resize(p, targets, s) (hashinator.h:552-566).  Every target runs
        the device rehash here; the host rehash of the reference is its
        CPU-only path, not part of the accelerated API."""
        self.device_rehash(newSizePower, stream)

    @on_stream
    def resize_to_lf(self, targetLF: float = 0.5, t: int = targets.device, stream=None):
        """resize_to_lf (hashinator.h:530-546)."""
        while self.load_factor() > targetLF:
            self.device_rehash(self._sp + 1, stream)

    @on_stream
    def clear(self, t: int = targets.device, stream=None, length: int = 0):
        """clear(targets, s, len) (hashinator.h:491-518).  targets.device:
        keys -> EMPTY and fill = 0 only; otherwise a fresh table and Info."""
        if t == targets.device:
            n = length or self.bucket_count()
            check(self._fn("reset_all")(_ptr(self.buckets), _ptr(self._info), n, _stream(stream)), "clear")
        else:
            self.buckets = self._new_table(self._sp)
            check(self._L.hn_info_init(_ptr(self._info), self._sp, _stream(stream)), "clear")
        self._sync(stream)

    @on_stream
    def performCleanupTasks(self, stream=None):
        """GPU performCleanupTasks (hashinator.h:642-650)."""
        if self.tombstone_ratio() > 0.025:
            self.clean_tombstones(stream)
        while self.info().current_max_bucket_overflow > _lib.HN_BUCKET_OVERFLOW:
            self.device_rehash(self._sp + 1, stream)

    @on_stream
    def download(self, stream=None):
        """download() (hashinator.h:1457-1467): rehash when the overflow grew
        beyond BUCKET_OVERFLOW, else clean tombstones if any."""
        if self.info().current_max_bucket_overflow > _lib.HN_BUCKET_OVERFLOW:
            self.device_rehash(self._sp + 1, stream)
        elif self.tombstone_count() > 0:
            self.clean_tombstones(stream)

    # ----------------------------------------------------------- extraction (hashinator.h:1154-1259)
    @on_stream
    def extractPattern(self, pred: hn_pred | None = None, stream=None):
        """Pairs of the bucket array selected by ``pred`` (default: valid keys),
        in bucket order, as an (n, 2) device tensor."""
        from .tools import copy_if_pairs

        return copy_if_pairs(self.buckets, pred if pred is not None else valid_pred(), self.kb, stream)

    @on_stream
    def extractKeysByPattern(self, pred: hn_pred | None = None, stream=None):
        from .tools import copy_keys_if_pairs

        return copy_keys_if_pairs(self.buckets, pred if pred is not None else valid_pred(), self.kb, stream)

    @on_stream
    def extractAllKeys(self, stream=None):
        return self.extractKeysByPattern(valid_pred(), stream)

    # The *Loop variants (hashinator.h:1186-1259): the reference runs them as
    # one block looping over the buckets; the same elements and order come
    # from the single-pass compaction here.
    @on_stream
    def extractPatternLoop(self, pred: hn_pred | None = None, stream=None):
        return self.extractPattern(pred, stream)

    @on_stream
    def extractLoop(self, stream=None):
        return self.extractPattern(valid_pred(), stream)

    @on_stream
    def extractKeysByPatternLoop(self, pred: hn_pred | None = None, stream=None):
        return self.extractKeysByPattern(pred, stream)

    @on_stream
    def extractAllKeysLoop(self, stream=None):
        return self.extractAllKeys(stream)


def valid_pred() -> hn_pred:
    return hn_pred(_lib.HN_PRED_VALID, 0, 0, 0)
