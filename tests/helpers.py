"""Shared helpers for the parity tests (the oracle is the checker, never the product)."""
from __future__ import annotations

import numpy as np

import oracle as O
from hashinator_amd import workloads as W

NDT = {32: np.uint32, 64: np.uint64}
EMPTY = {32: np.uint32(0xFFFFFFFF), 64: np.uint64(0xFFFFFFFFFFFFFFFF)}
TOMB = {32: np.uint32(0xFFFFFFFE), 64: np.uint64(0xFFFFFFFFFFFFFFFE)}


def keyset(n, bits, seed, start=0):
    return W.unique_keys(n, bits, seed, start)


def live_pairs(img, bits):
    """Sorted (key, value) rows of the live slots of an (N, 2) bucket image."""
    img = np.asarray(img).reshape(-1, 2)
    k = img[:, 0]
    live = (k != EMPTY[bits]) & (k != TOMB[bits])
    rows = img[live]
    return rows[np.argsort(rows[:, 0], kind="stable")]


def occupancy(img, bits):
    """Non-EMPTY slot mask (TOMBSTONEs count as occupied)."""
    return np.asarray(img).reshape(-1, 2)[:, 0] != EMPTY[bits]


def model_from(img, info_tuple, bits):
    """An oracle GpuModel holding a copy of a device bucket image and Info."""
    sp, fill, ovf, tomb, err = info_tuple
    i = O.Info(sp, 0, fill, ovf, tomb, err, 0)
    return O.GpuModel(sp, bits, buckets=np.asarray(img).reshape(-1), info=i)


def fresh_model(sp, bits):
    return O.GpuModel(sp, bits)


def info_tuple(hm):
    i = hm.info()
    return (i.size_power, i.fill, i.current_max_bucket_overflow, i.tombstone_counter, i.err)


def reachable_all(img, bits, sp):
    """Every live key's probe path from its home slot crosses no EMPTY slot."""
    img = np.asarray(img).reshape(-1, 2)
    n = img.shape[0]
    keys = img[:, 0]
    occ = keys != EMPTY[bits]
    live = occ & (keys != TOMB[bits])
    pos = np.nonzero(live)[0]
    home = O.fibhash(keys[pos], sp, bits).astype(np.int64) & (n - 1)
    # number of EMPTY slots in the cyclic interval [home, pos]
    empt = np.concatenate([[0], np.cumsum(~occ)])
    dist = (pos - home) % n
    end = home + dist
    cnt = np.where(end < n, empt[np.minimum(end, n - 1) + 1] - empt[home],
                   (empt[n] - empt[home]) + empt[(end - n) + 1])
    return bool(np.all(cnt == 0))


def next_overflow(x, w=64):
    return x if x % w == 0 else x + (w - x % w)
