"""A continuous-batching request loop over maps resident in HBM.

No counterpart in the reference; each request is one of the reference's
accelerated calls (insert / retrieve / erase, hashinator.h:1318-1409) on one
map, with the reference's auto-grow (a resize before an insert that would
cross the load-factor ceiling, 1329-1332 -- here at 0.85).

* Requests arrive over time, each for one map.  Per map they run in arrival
  order (a map has at most one request on the device), so every result is
  the one the same request gives when the stream is replayed alone.
* Requests for different maps run concurrently on a pool of HIP streams; a
  request is admitted when a stream is free and its device bytes (inputs,
  outputs and, for a growing insert, the new table) fit in the HBM budget
  next to the requests already in flight.
* Queued retrieves (or erases) for the same map are merged into one launch
  when they are next in line together -- both are order-insensitive among
  themselves -- which is the batching; inserts are not merged (a key repeated
  across two inserts must keep the later value).
* Maps of at least ``build_rehash_min`` keys (default 2^20) grow by the bulk
  rehash, and inserts of at least ``merge_min_keys`` keys (default 2^20) and
  ``merge_min_frac`` of their table go through the reorder and the merge
  into a fresh array; ``None`` turns either off (growth by reinsertion,
  inserts by probing).  Both give the same map contents as the probing
  paths.

Latency is completion minus arrival; the loop reports p50/p99 and Mops/s.
"""
from __future__ import annotations

import ctypes as C
import math
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from ._lib import check

INSERT, RETRIEVE, ERASE = "insert", "retrieve", "erase"
_TDT = {32: torch.int32, 64: torch.int64}
_NDT = {32: np.uint32, 64: np.uint64}


@dataclass
class Request:
    rid: int
    map_id: int
    kind: str
    keys: np.ndarray
    vals: np.ndarray | None = None
    arrival: float = 0.0  # seconds from the start of the run
    # host buffers the keys / values live in, when pinned (see pin_request)
    pinned_keys: torch.Tensor | None = None
    pinned_vals: torch.Tensor | None = None
    # filled in by the engine
    out_vals: np.ndarray | None = None
    out_hit: np.ndarray | None = None
    t_admit: float = -1.0
    t_done: float = -1.0

    @property
    def latency(self):
        return self.t_done - self.arrival


def pin_request(r: Request) -> Request:
    """Move a request's keys (and values) into pinned host memory, as a
    server receives them into pinned buffers: the engine then copies them to
    the device straight from there, with no host-side staging copy."""
    kb = 64 if r.keys.dtype.itemsize == 8 else 32
    tk = torch.empty(r.keys.size, dtype=_TDT[kb], pin_memory=True)
    a = tk.numpy().view(_NDT[kb])
    a[:] = r.keys
    r.keys, r.pinned_keys = a, tk
    if r.vals is not None:
        tv = torch.empty(r.vals.size, dtype=_TDT[kb], pin_memory=True)
        b = tv.numpy().view(_NDT[kb])
        b[:] = r.vals
        r.vals, r.pinned_vals = b, tv
    return r


@dataclass
class _Map:
    sp: int
    buckets: torch.Tensor
    info: torch.Tensor
    fill_upper: int  # an upper bound of the fill, known without reading the device
    info_host: torch.Tensor | None = None  # pinned copy of the Info, written after each request
    busy: bool = False
    queue: deque = field(default_factory=deque)


class ServingEngine:
    def __init__(self, sps, key_bits=64, n_streams=8, hbm_budget_bytes=32 << 30, grow_lf=0.85, device="cuda",
                 tune=None, build_rehash_min=1 << 20, merge_min_keys=1 << 20, merge_min_frac=0.25):
        self.kb = key_bits
        self.sfx = "u32" if key_bits == 32 else "u64"
        self.slot = 8 if key_bits == 32 else 16
        self.es = 4 if key_bits == 32 else 8
        self.L = _lib.load()
        self.dev = torch.device(device)
        self.grow_lf = grow_lf
        # grow maps of at least this many keys by the bulk rehash (None: never,
        # always by reinsertion; its scratch, ~16 B per old slot, comes from
        # the launch stream's warmed allocator pool)
        self.build_rehash_min = build_rehash_min
        # inserts of at least this many keys and merge_min_frac of the table
        # (after any growth) take the reorder + merge into a fresh array
        # (hn_merge_*: the insert + reorder image; a table the merge cannot
        # take gets the probing insert's result from the same call); None: never
        self.merge_min_keys = merge_min_keys
        self.merge_min_frac = merge_min_frac
        self.budget = hbm_budget_bytes
        self.tune = C.byref(tune) if tune is not None else None
        self.maps = []
        s0 = torch.cuda.current_stream()
        for sp in sps:
            b = torch.empty((1 << int(sp), 2), dtype=_TDT[key_bits], device=self.dev)
            info = torch.zeros(5, dtype=torch.int64, device=self.dev)
            check(getattr(self.L, f"hn_table_init_{self.sfx}")(b.data_ptr(), 1 << int(sp), s0.cuda_stream), "init")
            check(self.L.hn_info_init(info.data_ptr(), int(sp), s0.cuda_stream), "info")
            self.maps.append(_Map(int(sp), b, info, 0, torch.zeros(5, dtype=torch.int64, pin_memory=True)))
        torch.cuda.synchronize()
        self.streams = [torch.cuda.Stream(device=self.dev) for _ in range(n_streams)]
        self._warm_up()
        self.free = list(range(n_streams))
        self.inflight = []  # (event, stream index, map id, [requests], bytes, keepalive tensors, t_admit)
        self.bytes_inflight = 0
        self.rehashes = 0
        self.merges = 0
        self.launches = 0
        self.completed = 0

    def _warm_up(self):
        """Run every entry point the loop uses once on a small scratch map, so
        that loading code objects and the launch-shape queries done on a
        kernel's first call do not land on the first requests."""
        L, sfx, s = self.L, self.sfx, torch.cuda.current_stream().cuda_stream
        sp = 10
        b = torch.empty((1 << sp, 2), dtype=_TDT[self.kb], device=self.dev)
        nb = torch.empty((1 << (sp + 1), 2), dtype=_TDT[self.kb], device=self.dev)
        info = torch.zeros(5, dtype=torch.int64, device=self.dev)
        k = torch.arange(1, 513, dtype=_TDT[self.kb], device=self.dev)
        v = torch.empty_like(k)
        h = torch.empty(k.numel(), dtype=torch.uint8, device=self.dev)
        sc = torch.empty(L.hn_rehash_scratch_bytes(), dtype=torch.uint8, device=self.dev)
        check(getattr(L, f"hn_table_init_{sfx}")(b.data_ptr(), 1 << sp, s), "warm-up")
        check(L.hn_info_init(info.data_ptr(), sp, s), "warm-up")
        check(getattr(L, f"hn_insert_{sfx}")(k.data_ptr(), k.data_ptr(), b.data_ptr(), info.data_ptr(), k.numel(), s,
                                              self.tune), "warm-up")
        check(getattr(L, f"hn_retrieve_{sfx}")(k.data_ptr(), v.data_ptr(), h.data_ptr(), b.data_ptr(),
                                                info.data_ptr(), k.numel(), s, self.tune), "warm-up")
        check(getattr(L, f"hn_erase_{sfx}")(k.data_ptr(), b.data_ptr(), info.data_ptr(), 16, s, self.tune), "warm-up")
        check(getattr(L, f"hn_rehash_{sfx}")(b.data_ptr(), sp, nb.data_ptr(), sp + 1, info.data_ptr(), sc.data_ptr(), s),
              "warm-up")
        nb2 = torch.empty((1 << (sp + 2), 2), dtype=_TDT[self.kb], device=self.dev)
        sb2 = getattr(L, f"hn_rehash_build_scratch_bytes_{sfx}")(sp + 1, sp + 2)
        sc2 = torch.empty(sb2, dtype=torch.uint8, device=self.dev)
        check(getattr(L, f"hn_rehash_build_{sfx}")(nb.data_ptr(), sp + 1, nb2.data_ptr(), sp + 2, info.data_ptr(),
                                                   sc2.data_ptr(), sb2, s), "warm-up")
        rsb = L.hn_reorder_scratch_bytes(sp + 2)
        rsc = torch.empty(rsb, dtype=torch.uint8, device=self.dev)
        check(getattr(L, f"hn_reorder_{sfx}")(nb2.data_ptr(), info.data_ptr(), sp + 2, rsc.data_ptr(), rsb, s),
              "warm-up")
        nb3 = torch.empty_like(nb2)
        msb = getattr(L, f"hn_merge_scratch_bytes_{sfx}")(sp + 2, k.numel())
        msc = torch.empty(msb, dtype=torch.uint8, device=self.dev)
        check(getattr(L, f"hn_merge_{sfx}")(k.data_ptr(), k.data_ptr(), nb2.data_ptr(), nb3.data_ptr(),
                                             info.data_ptr(), k.numel(), sp + 2, msc.data_ptr(), msb, s), "warm-up")
        torch.cuda.synchronize()
        # pinned result buffers come from torch's caching host allocator: fill
        # its size classes (powers of two, 4 KiB .. 64 MiB, two of each) now
        held = [torch.empty(1 << e, dtype=torch.uint8, pin_memory=True) for e in range(12, 27) for _ in range(2)]
        del held
        # Device-side request buffers, fresh bucket arrays and scratch come
        # from torch's caching allocator, which reuses a freed block only on
        # the stream that allocated it: leave one cached segment in every
        # launch stream's pool (up to 2 GiB, within the in-flight budget), so
        # that the first requests on a stream do not wait for hipMalloc
        # (~1.5 ms per new segment of tens of MB on the host's launch path).
        reserve = max(min(self.budget // len(self.streams), 2 << 30), 64 << 20)
        for st in self.streams:
            with torch.cuda.stream(st):
                t = torch.empty(reserve, dtype=torch.uint8, device=self.dev)
                u = torch.empty(1 << 20, dtype=torch.uint8, device=self.dev)  # the small-block pool (<= 1 MiB)
            del t, u
        torch.cuda.synchronize()

    # ------------------------------------------------------------ preload (not timed)
    def preload(self, map_id, keys, vals):
        m = self.maps[map_id]
        s = torch.cuda.current_stream()
        dk = torch.from_numpy(np.ascontiguousarray(keys).view(np.int64 if self.kb == 64 else np.int32)).to(self.dev)
        dv = torch.from_numpy(np.ascontiguousarray(vals).view(np.int64 if self.kb == 64 else np.int32)).to(self.dev)
        check(getattr(self.L, f"hn_insert_{self.sfx}")(dk.data_ptr(), dv.data_ptr(), m.buckets.data_ptr(),
                                                        m.info.data_ptr(), dk.numel(), s.cuda_stream, self.tune),
              "preload")
        m.fill_upper += dk.numel()

    # ------------------------------------------------------------ loop
    def _request_bytes(self, m: _Map, reqs):
        n = sum(r.keys.size for r in reqs)
        kind = reqs[0].kind
        b = n * self.es * (2 if kind != ERASE else 1) + (n if kind == RETRIEVE else 0)
        if kind == INSERT and self._needs_grow(m, n):
            new_sp = self._grow_sp(m, n)
            b += (1 << new_sp) * self.slot + self._rehash_scratch(m, new_sp)[0]
        if kind == INSERT:
            sp = self._grow_sp(m, n) if self._needs_grow(m, n) else m.sp
            if self._merges(sp, n):
                b += (1 << sp) * self.slot + self._merge_scratch(sp, n)
        return b

    def _merges(self, sp, n):
        return self.merge_min_keys is not None and n >= self.merge_min_keys and n >= self.merge_min_frac * (1 << sp)

    def _merge_scratch(self, sp, n):
        return (getattr(self.L, f"hn_merge_scratch_bytes_{self.sfx}")(sp, n)
                + (self.L.hn_reorder_scratch_bytes(sp) + 255) // 256 * 256)

    def _needs_grow(self, m: _Map, n):
        return m.fill_upper + n > self.grow_lf * (1 << m.sp)

    def _grow_sp(self, m: _Map, n):
        new_sp = m.sp + 1
        while m.fill_upper + n > self.grow_lf * (1 << new_sp):
            new_sp += 1
        return new_sp

    def _rehash_scratch(self, m: _Map, new_sp):
        """(bytes, bulk) of a growth's scratch: a map of many keys grows by the
        bulk rehash (one build from its live slots), a small one by reinsertion."""
        if self.build_rehash_min is not None and m.fill_upper >= self.build_rehash_min:
            sb = getattr(self.L, f"hn_rehash_build_scratch_bytes_{self.sfx}")(m.sp, new_sp)
            if sb:
                return sb, True
        return self.L.hn_rehash_scratch_bytes(), False

    def _take(self, m: _Map):
        """Next group of requests for a map: one insert, or a run of same-kind retrieves/erases."""
        head = m.queue[0]
        group = [head]
        if head.kind != INSERT:
            for r in list(m.queue)[1:]:
                if r.kind != head.kind:
                    break
                group.append(r)
        return group

    def _host_input(self, reqs, what):
        """The group's keys or values in pinned host memory: a single pinned
        request's own buffer, else one pinned buffer filled in order."""
        if len(reqs) == 1 and getattr(reqs[0], f"pinned_{what}") is not None:
            return getattr(reqs[0], f"pinned_{what}")
        parts = [getattr(r, what) for r in reqs]
        h = torch.empty(sum(p.size for p in parts), dtype=_TDT[self.kb], pin_memory=True)
        np.concatenate(parts, out=h.numpy().view(_NDT[self.kb]))
        return h

    def _launch(self, map_id, reqs, si, now, nbytes):
        m = self.maps[map_id]
        s = self.streams[si]
        L, sfx = self.L, self.sfx
        kind = reqs[0].kind
        keep = []
        with torch.cuda.stream(s):
            hk = self._host_input(reqs, "keys")
            n = hk.numel()
            dk = hk.to(self.dev, non_blocking=True)
            keep += [hk, dk]
            if kind == INSERT:
                if self._needs_grow(m, n):
                    new_sp = self._grow_sp(m, n)
                    nb = torch.empty((1 << new_sp, 2), dtype=_TDT[self.kb], device=self.dev)
                    sb, bulk = self._rehash_scratch(m, new_sp)
                    sc = torch.empty(sb, dtype=torch.uint8, device=self.dev)
                    if bulk:
                        check(getattr(L, f"hn_rehash_build_{sfx}")(m.buckets.data_ptr(), m.sp, nb.data_ptr(), new_sp,
                                                                   m.info.data_ptr(), sc.data_ptr(), sb, s.cuda_stream),
                              "rehash")
                    else:
                        check(getattr(L, f"hn_rehash_{sfx}")(m.buckets.data_ptr(), m.sp, nb.data_ptr(), new_sp,
                                                             m.info.data_ptr(), sc.data_ptr(), s.cuda_stream), "rehash")
                    m.buckets.record_stream(s)
                    keep += [m.buckets, sc]
                    m.buckets, m.sp = nb, new_sp
                    self.rehashes += 1
                hv = self._host_input(reqs, "vals")
                dv = hv.to(self.dev, non_blocking=True)
                keep += [hv, dv]
                if self._merges(m.sp, n):
                    # reorder (a no-op on an ordered table), then the merge into a fresh array
                    rsb = (L.hn_reorder_scratch_bytes(m.sp) + 255) // 256 * 256
                    msb = getattr(L, f"hn_merge_scratch_bytes_{sfx}")(m.sp, n)
                    sc = torch.empty(rsb + msb, dtype=torch.uint8, device=self.dev)
                    check(getattr(L, f"hn_reorder_{sfx}")(m.buckets.data_ptr(), m.info.data_ptr(), m.sp,
                                                           sc.data_ptr(), rsb, s.cuda_stream), "reorder")
                    nb = torch.empty((1 << m.sp, 2), dtype=_TDT[self.kb], device=self.dev)
                    check(getattr(L, f"hn_merge_{sfx}")(dk.data_ptr(), dv.data_ptr(), m.buckets.data_ptr(),
                                                         nb.data_ptr(), m.info.data_ptr(), n, m.sp,
                                                         sc.data_ptr() + rsb, msb, s.cuda_stream), "merge")
                    m.buckets.record_stream(s)
                    keep += [m.buckets, sc]
                    m.buckets = nb
                    self.merges += 1
                else:
                    check(getattr(L, f"hn_insert_{sfx}")(dk.data_ptr(), dv.data_ptr(), m.buckets.data_ptr(),
                                                          m.info.data_ptr(), n, s.cuda_stream, self.tune), "insert")
                m.fill_upper += n
            elif kind == RETRIEVE:
                dv = torch.zeros(n, dtype=_TDT[self.kb], device=self.dev)
                dh = torch.empty(n, dtype=torch.uint8, device=self.dev)
                check(getattr(L, f"hn_retrieve_{sfx}")(dk.data_ptr(), dv.data_ptr(), dh.data_ptr(),
                                                        m.buckets.data_ptr(), m.info.data_ptr(), n, s.cuda_stream,
                                                        self.tune), "retrieve")
                hv = torch.empty(n, dtype=_TDT[self.kb], pin_memory=True)
                hh = torch.empty(n, dtype=torch.uint8, pin_memory=True)
                hv.copy_(dv, non_blocking=True)
                hh.copy_(dh, non_blocking=True)
                keep += [dv, dh, hv, hh]
            else:
                check(getattr(L, f"hn_erase_{sfx}")(dk.data_ptr(), m.buckets.data_ptr(), m.info.data_ptr(), n,
                                                     s.cuda_stream, self.tune), "erase")
            m.info_host.copy_(m.info, non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(s)
        self.launches += 1
        for r in reqs:
            r.t_admit = now
            m.queue.popleft()
        m.busy = True
        self.bytes_inflight += nbytes
        self.inflight.append((ev, si, map_id, reqs, nbytes, keep))

    def _complete(self, now):
        still = []
        for item in self.inflight:
            ev, si, map_id, reqs, nbytes, keep = item
            if not ev.query():
                still.append(item)
                continue
            if reqs[0].kind == RETRIEVE:
                # results stay in the pinned buffers they arrived in (views, no copy)
                hv, hh = keep[-2], keep[-1]
                vals = hv.numpy().view(_NDT[self.kb])
                hit = hh.numpy().view(np.bool_)
                off = 0
                for r in reqs:
                    k = r.keys.size
                    r.out_vals = vals[off:off + k]
                    r.out_hit = hit[off:off + k]
                    off += k
            for r in reqs:
                r.t_done = now
            self.completed += len(reqs)
            m = self.maps[map_id]
            m.fill_upper = int(m.info_host[1])  # the device's fill, now that the map is idle
            m.busy = False
            self.free.append(si)
            self.bytes_inflight -= nbytes
        self.inflight = still

    def serve(self, requests, realtime=True):
        """Run `requests` (sorted by arrival) through the loop; returns wall seconds."""
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        t0 = time.perf_counter()
        done0 = self.completed
        done = 0
        total = len(pending)
        while done < total:
            now = time.perf_counter() - t0
            while pending and (pending[0].arrival <= now or not realtime):
                r = pending.popleft()
                self.maps[r.map_id].queue.append(r)
            self._complete(now)
            dispatched = False
            # maps in the order their oldest waiting request arrived: with the
            # streams all busy, a freed stream goes to the longest wait
            ready = sorted((m.queue[0].arrival, map_id) for map_id, m in enumerate(self.maps)
                           if m.queue and not m.busy)
            for _, map_id in ready:
                m = self.maps[map_id]
                if not self.free:
                    break
                reqs = self._take(m)
                need = self._request_bytes(m, reqs)
                if self.inflight and self.bytes_inflight + need > self.budget:
                    continue  # wait for memory to free up
                self._launch(map_id, reqs, self.free.pop(), now, need)
                dispatched = True
            done = self.completed - done0
            if not dispatched and self.inflight:
                # poll every in-flight launch: blocking on the oldest one would
                # hold back the completions of faster launches behind it
                time.sleep(0)
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    def serve_alone(self, requests):
        """Every request by itself, in arrival order, one at a time (the reference's synchronous use)."""
        t0 = time.perf_counter()
        for r in sorted(requests, key=lambda r: r.arrival):
            m = self.maps[r.map_id]
            self.maps[r.map_id].queue.append(r)
            need = self._request_bytes(m, [r])
            self._launch(r.map_id, [r], self.free.pop(), time.perf_counter() - t0, need)
            self.inflight[-1][0].synchronize()
            self._complete(time.perf_counter() - t0)
            assert not m.busy
        return time.perf_counter() - t0

    # ------------------------------------------------------------ reporting
    def live_set(self, map_id):
        m = self.maps[map_id]
        img = m.buckets.cpu().numpy().view(_NDT[self.kb]).reshape(-1, 2)
        k = img[:, 0]
        e = np.iinfo(_NDT[self.kb]).max
        rows = img[(k != e) & (k != e - 1)]
        return rows[np.argsort(rows[:, 0], kind="stable")]

    def infos(self):
        return np.stack([m.info.cpu().numpy() for m in self.maps])


def make_requests(engine_sps, n_requests, rate_hz, seed, key_bits=64, min_batch=1000, max_batch=4_000_000,
                  mix=(0.45, 0.45, 0.10), repeat_frac=0.3, live_keys=None, pinned=True):
    """A Poisson stream of requests over the maps: kind by `mix`, batch size
    log-uniform in [min_batch, max_batch], keys uniform (unique within a
    batch) with `repeat_frac` drawn from keys already live in the map;
    values are a function of (key, request id).  `pinned`: the keys and
    values are delivered in pinned host memory (pin_request)."""
    from . import workloads as W

    rng = np.random.default_rng(seed)
    nd = _NDT[key_bits]
    t = 0.0
    reqs = []
    live = [list(x) if x is not None else [] for x in (live_keys or [None] * len(engine_sps))]
    next_idx = 10_000_000
    for rid in range(n_requests):
        t += rng.exponential(1.0 / rate_hz)
        mid = int(rng.integers(0, len(engine_sps)))
        kind = rng.choice([INSERT, RETRIEVE, ERASE], p=list(mix))
        n = int(math.exp(rng.uniform(math.log(min_batch), math.log(max_batch))))
        n_rep = int(repeat_frac * n) if live[mid] else 0
        fresh = W.unique_keys(n - n_rep, key_bits, seed * 7919 + mid, start=next_idx)
        next_idx += n + 2
        if n_rep:
            pool = live[mid][int(rng.integers(0, len(live[mid])))]
            rep = pool[rng.integers(0, pool.size, size=min(n_rep, pool.size))]
            keys = np.unique(np.concatenate([fresh, rep]).astype(nd))
            keys = rng.permutation(keys)
        else:
            keys = fresh.astype(nd)
        vals = W.values(keys, rid).astype(nd) if kind == INSERT else None
        if kind == INSERT:
            live[mid].append(keys)
        r = Request(rid, mid, str(kind), keys, vals, t)
        reqs.append(pin_request(r) if pinned else r)
    return reqs


def latency_report(requests, wall_s):
    lat = np.array([r.latency for r in requests])
    ops = sum(r.keys.size for r in requests)
    return {
        "requests": len(requests),
        "ops": int(ops),
        "wall_s": round(wall_s, 4),
        "Mops": round(ops / wall_s / 1e6, 1),
        "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
        "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3),
        "max_ms": round(float(lat.max()) * 1e3, 3),
    }
