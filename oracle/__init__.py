"""CPU restatement of Hashinator's semantics -- TEST INFRASTRUCTURE ONLY.

This package is the checker for the HIP path and the CPU timing baseline.
Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
may import it; the shipped package ``hashinator_amd`` never does.

``liboracle.so`` is built from ``hn_oracle.c`` (see the Makefile beside it).
``ref()`` loads ``_ref/libhashref.so``, the reference's own host-only map,
when it was built in this tree (it needs the reference sources to build).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

EMPTY32 = np.uint32(0xFFFFFFFF)
TOMB32 = np.uint32(0xFFFFFFFE)
EMPTY64 = np.uint64(0xFFFFFFFFFFFFFFFF)
TOMB64 = np.uint64(0xFFFFFFFFFFFFFFFE)


class Info(C.Structure):
    """Byte layout of Hashinator::Info (defaults.h:44-54)."""

    _fields_ = [
        ("size_power", C.c_int32),
        ("_pad0", C.c_int32),
        ("fill", C.c_uint64),
        ("overflow", C.c_uint64),
        ("tombstones", C.c_uint64),
        ("err", C.c_int32),
        ("_pad1", C.c_int32),
    ]

    def as_tuple(self):
        return (self.size_power, self.fill, self.overflow, self.tombstones, self.err)


class Pred(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("field", C.c_int32),
        ("mask", C.c_uint64),
        ("value", C.c_uint64),
        ("empty", C.c_uint64),
        ("tombstone", C.c_uint64),
    ]


PRED_ALL, PRED_VALID, PRED_MASKEQ, PRED_LT, PRED_GE = 0, 1, 2, 3, 4


def make_pred(kind, field=0, mask=0, value=0, width=64):
    e = (1 << width) - 1
    return Pred(kind, field, mask, value, e, e - 1)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        _LIB = C.CDLL(path)
        _declare(_LIB)
    return _LIB


def ref_available() -> bool:
    return os.path.exists(os.path.join(_HERE, "_ref", "libhashref.so"))


def ref():
    """The reference's own CPU-only map (kstppd/hashinator, HASHINATOR_CPU_ONLY_MODE)."""
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "libhashref.so")
        _REF = C.CDLL(path)
        P = C.c_void_p
        for w, K in ((32, C.c_uint32), (64, C.c_uint64)):
            f = getattr(_REF, f"ref_fibhash{w}")
            f.restype, f.argtypes = K, [K, C.c_int]
            g = lambda name: getattr(_REF, f"ref_map{w}_{name}")
            g("new").restype, g("new").argtypes = P, [C.c_int]
            g("free").argtypes = [P]
            g("info").argtypes = [P, C.POINTER(Info)]
            g("nbuckets").restype, g("nbuckets").argtypes = C.c_size_t, [P]
            g("copy_buckets").argtypes = [P, P]
            g("insert").argtypes = [P, P, P, C.c_size_t]
            g("find").argtypes = [P, P, P, P, C.c_size_t]
            g("erase").restype, g("erase").argtypes = C.c_size_t, [P, P, C.c_size_t]
            g("rehash").argtypes = [P, C.c_int]
    return _REF


def _declare(L):
    P = C.c_void_p
    IP = C.POINTER(Info)
    L.or_fibhash32.restype, L.or_fibhash32.argtypes = C.c_uint32, [C.c_uint32, C.c_int]
    L.or_fibhash64.restype, L.or_fibhash64.argtypes = C.c_uint64, [C.c_uint64, C.c_int]
    for w in (32, 64):
        getattr(L, f"or_fibhash{w}_n").argtypes = [P, P, P, C.c_size_t]
        g = lambda name: getattr(L, f"or_map{w}_{name}")
        g("new").restype, g("new").argtypes = P, [C.c_int]
        g("free").argtypes = [P]
        g("info").restype, g("info").argtypes = IP, [P]
        g("nbuckets").restype, g("nbuckets").argtypes = C.c_size_t, [P]
        g("copy_buckets").argtypes = [P, P]
        g("insert").argtypes = [P, P, P, C.c_size_t]
        g("find").argtypes = [P, P, P, P, C.c_size_t]
        g("erase").restype, g("erase").argtypes = C.c_size_t, [P, P, C.c_size_t]
        g("rehash").argtypes = [P, C.c_int]
        h = lambda name: getattr(L, f"or_gpu{w}_{name}")
        h("init").argtypes = [P, IP, C.c_int]
        h("insert").restype, h("insert").argtypes = C.c_int, [P, IP, P, P, C.c_size_t]
        h("insert_index").restype, h("insert_index").argtypes = C.c_int, [P, IP, P, C.c_size_t]
        h("insert_pairs").restype, h("insert_pairs").argtypes = C.c_int, [P, IP, P, C.c_size_t]
        h("retrieve").argtypes = [P, IP, P, P, P, C.c_size_t]
        h("erase").argtypes = [P, IP, P, C.c_size_t]
        h("reset_all").argtypes = [P, IP, C.c_size_t]
        h("reset").argtypes = [P, IP, P, C.c_size_t]
        h("clean_tombstones").restype, h("clean_tombstones").argtypes = C.c_size_t, [P, IP]
        h("rehash").restype, h("rehash").argtypes = C.c_int, [P, C.c_int, P, C.c_int, IP]
        h("max_probe").restype, h("max_probe").argtypes = C.c_uint64, [P, C.c_int]
        h("reorder").restype, h("reorder").argtypes = C.c_int, [P, IP]
        for nm in (f"or_copy_if_pairs{w}", f"or_copy_keys_if_pairs{w}"):
            getattr(L, nm).restype = C.c_size_t
            getattr(L, nm).argtypes = [P, C.c_size_t, C.POINTER(Pred), P]
    L.or_copy_if_flags.restype = C.c_size_t
    L.or_copy_if_flags.argtypes = [P, P, C.c_size_t, C.c_size_t, P]
    L.or_exclusive_scan_u32.argtypes = [P, P, C.c_size_t]
    L.or_exclusive_scan_u64.argtypes = [P, P, C.c_size_t]


def _p(a):
    return C.c_void_p(a.ctypes.data) if a is not None else None


def _kdt(width):
    return np.uint32 if width == 32 else np.uint64


# ---------------------------------------------------------------- hashing
def fibhash(keys, sps, width):
    keys = np.ascontiguousarray(keys, dtype=_kdt(width))
    sps = np.ascontiguousarray(np.broadcast_to(np.asarray(sps, dtype=np.int32), keys.shape))
    out = np.empty_like(keys)
    getattr(lib(), f"or_fibhash{width}_n")(_p(keys), _p(sps), _p(out), keys.size)
    return out


# ---------------------------------------------------------------- host-only map
class CpuMap:
    """The reference's HASHINATOR_CPU_ONLY_MODE map, restated in C."""

    def __init__(self, size_power, width=32, impl=None):
        self.w = width
        self.L = impl if impl is not None else lib()
        self.pre = ("or_map" if impl is None else "ref_map") + str(width) + "_"
        self.h = getattr(self.L, self.pre + "new")(size_power)

    def _f(self, name):
        return getattr(self.L, self.pre + name)

    def close(self):
        if self.h:
            self._f("free")(self.h)
            self.h = None

    __del__ = close

    def info(self):
        if self.pre.startswith("ref"):
            i = Info()
            self._f("info")(self.h, C.byref(i))
            return i
        return self._f("info")(self.h).contents

    def nbuckets(self):
        return self._f("nbuckets")(self.h)

    def buckets(self):
        out = np.empty(2 * self.nbuckets(), dtype=_kdt(self.w))
        self._f("copy_buckets")(self.h, _p(out))
        return out

    def insert(self, keys, vals):
        k = np.ascontiguousarray(keys, dtype=_kdt(self.w))
        v = np.ascontiguousarray(vals, dtype=_kdt(self.w))
        self._f("insert")(self.h, _p(k), _p(v), k.size)

    def find(self, keys):
        k = np.ascontiguousarray(keys, dtype=_kdt(self.w))
        v = np.zeros_like(k)
        hit = np.zeros(k.size, dtype=np.uint8)
        self._f("find")(self.h, _p(k), _p(v), _p(hit), k.size)
        return v, hit.astype(bool)

    def erase(self, keys):
        k = np.ascontiguousarray(keys, dtype=_kdt(self.w))
        return self._f("erase")(self.h, _p(k), k.size)

    def rehash(self, sp):
        self._f("rehash")(self.h, sp)


def ref_map(size_power, width=32):
    return CpuMap(size_power, width, impl=ref())


# ---------------------------------------------------------------- accelerated-path model
class GpuModel:
    """Sequential statement of the accelerated kernels' per-key semantics.

    ``buckets`` is an interleaved (key, value) array of 2*N words and ``info``
    an :class:`Info`, exactly as the device holds them.
    """

    def __init__(self, size_power, width=32, buckets=None, info=None):
        self.w = width
        self.L = lib()
        if buckets is None:
            self.b = np.empty(2 << size_power, dtype=_kdt(width))
            self.info = Info()
            self._g("init")(_p(self.b), C.byref(self.info), size_power)
        else:
            self.b = np.array(buckets, dtype=_kdt(width), copy=True)
            self.info = Info.from_buffer_copy(bytes(info))

    def _g(self, name):
        return getattr(self.L, f"or_gpu{self.w}_{name}")

    def _k(self, a):
        return np.ascontiguousarray(a, dtype=_kdt(self.w))

    def insert(self, keys, vals):
        k, v = self._k(keys), self._k(vals)
        return self._g("insert")(_p(self.b), C.byref(self.info), _p(k), _p(v), k.size)

    def insert_index(self, keys):
        k = self._k(keys)
        return self._g("insert_index")(_p(self.b), C.byref(self.info), _p(k), k.size)

    def insert_pairs(self, pairs):
        pr = self._k(pairs)
        return self._g("insert_pairs")(_p(self.b), C.byref(self.info), _p(pr), pr.size // 2)

    def retrieve(self, keys, vals_in=None):
        k = self._k(keys)
        v = np.zeros_like(k) if vals_in is None else self._k(vals_in).copy()
        hit = np.zeros(k.size, dtype=np.uint8)
        self._g("retrieve")(_p(self.b), C.byref(self.info), _p(k), _p(v), _p(hit), k.size)
        return v, hit.astype(bool)

    def erase(self, keys):
        k = self._k(keys)
        self._g("erase")(_p(self.b), C.byref(self.info), _p(k), k.size)

    def reset_all(self, len=None):
        n = self.b.size // 2 if len is None else len
        self._g("reset_all")(_p(self.b), C.byref(self.info), n)

    def reset(self, pairs):
        pr = self._k(pairs)
        self._g("reset")(_p(self.b), C.byref(self.info), _p(pr), pr.size // 2)

    def clean_tombstones(self):
        return self._g("clean_tombstones")(_p(self.b), C.byref(self.info))

    def rehash(self, new_sp):
        nb = np.empty(2 << new_sp, dtype=_kdt(self.w))
        rc = self._g("rehash")(_p(self.b), self.info.size_power, _p(nb), new_sp, C.byref(self.info))
        self.b = nb
        return rc

    def reorder(self):
        return self._g("reorder")(_p(self.b), C.byref(self.info))

    def max_probe(self):
        return self._g("max_probe")(_p(self.b), self.info.size_power)


# ---------------------------------------------------------------- split::tools
def copy_if_pairs(pairs, pred, width):
    a = np.ascontiguousarray(pairs, dtype=_kdt(width))
    out = np.empty_like(a)
    n = getattr(lib(), f"or_copy_if_pairs{width}")(_p(a), a.size // 2, C.byref(pred), _p(out))
    return out[: 2 * n]


def copy_keys_if_pairs(pairs, pred, width):
    a = np.ascontiguousarray(pairs, dtype=_kdt(width))
    out = np.empty(a.size // 2, dtype=a.dtype)
    n = getattr(lib(), f"or_copy_keys_if_pairs{width}")(_p(a), a.size // 2, C.byref(pred), _p(out))
    return out[:n]


def copy_if_flags(elems, flags):
    e = np.ascontiguousarray(elems)
    f = np.ascontiguousarray(flags, dtype=np.uint8)
    out = np.empty_like(e)
    eb = e.itemsize * (e.size // max(1, f.size)) if f.size else e.itemsize
    n = lib().or_copy_if_flags(_p(e), _p(f), f.size, eb, _p(out))
    per = e.size // max(1, f.size)
    return out[: n * per]


def exclusive_scan(x):
    x = np.ascontiguousarray(x)
    out = np.empty_like(x)
    fn = lib().or_exclusive_scan_u32 if x.dtype == np.uint32 else lib().or_exclusive_scan_u64
    fn(_p(x), _p(out), x.size)
    return out
