"""Robin Hood reorder pass and the early-stopping lookups it enables.

Linear probing without deletions fills the same slot set whatever the
arrival order, so the reordered image (every cluster sorted by home, ties by
key) is a function of the key set alone: the device image after a parallel
insert + reorder must equal the oracle's reorder of its sequential insert,
bit for bit.  The oracle's reorder is itself pinned (tests/test_oracle.py) to
the reference's own map fed the keys in sorted-home order.
"""
import numpy as np
import pytest
import torch

import oracle as O
from hashinator_amd import Hashmap, hn_tune, to_numpy
from hashinator_amd import _lib
from tests.helpers import NDT, info_tuple, keyset, live_pairs, occupancy, reachable_all

pytestmark = pytest.mark.gpu


def build(bits, sp, keys, vals, width=0):
    hm = Hashmap(sp, bits, tune=hn_tune(width, 256, 0, 0) if width else None, auto_reorder=0, auto_build=0)
    hm.insert(keys, vals, targetLF=1.0)
    model = O.GpuModel(sp, bits)
    model.insert(keys, vals)
    return hm, model


def ordered_flag(hm):
    return bool(hm.info().reserved1 & _lib.HN_INFO_ORDERED)


def check_exact(hm, model, bits):
    before_ovf = hm.info().current_max_bucket_overflow
    assert hm.reorder()
    assert ordered_flag(hm)
    assert model.reorder() == 1
    img = hm.host_buckets()
    assert np.array_equal(img.reshape(-1), model.b)
    sp, fill, ovf, tomb, err = info_tuple(hm)
    assert (fill, tomb) == (model.info.fill, 0)
    # the overflow is recomputed from the new layout: the oracle's, never above the arrival-order bound
    assert ovf == model.info.overflow <= before_ovf


@pytest.mark.parametrize("bits", [32, 64])
@pytest.mark.parametrize("sp,lf", [(3, 0.9), (6, 0.99), (10, 0.9), (12, 0.5), (14, 0.9), (16, 0.95), (18, 0.9)])
def test_reorder_image_exact(gpu, bits, sp, lf):
    n = max(1, int(lf * (1 << sp)))
    keys = keyset(n, bits, seed=sp * 31 + bits)
    vals = keys ^ NDT[bits](0x5A5A)
    hm, model = build(bits, sp, keys, vals)
    check_exact(hm, model, bits)
    # idempotent
    assert hm.reorder()
    assert np.array_equal(hm.host_buckets().reshape(-1), model.b)


def concentrated_keys(bits, sp, n, span, seed):
    """n distinct keys whose home slots all fall in [0, span): one long cluster."""
    out = []
    rng = np.random.default_rng(seed)
    while sum(len(o) for o in out) < n:
        c = rng.integers(1, 2**31 if bits == 32 else 2**62, 1 << 20, dtype=np.uint64).astype(NDT[bits])
        h = O.fibhash(c, sp, bits).astype(np.uint64) & np.uint64((1 << sp) - 1)
        out.append(c[h < span])
    k = np.unique(np.concatenate(out))[:n]
    return k[np.random.default_rng(seed + 1).permutation(k.size)]


@pytest.mark.parametrize("bits", [32, 64])
def test_reorder_long_cluster_second_pass(gpu, bits):
    # a cluster of ~3000 slots: longer than the first pass's segments, within the second's
    sp = 14
    keys = concentrated_keys(bits, sp, 3000, 256, seed=7)
    hm, model = build(bits, sp, keys, keys)
    check_exact(hm, model, bits)
    _, v, h = hm.retrieve(keys)
    assert h.cpu().numpy().all() and np.array_equal(to_numpy(v, bits), keys)


@pytest.mark.parametrize("bits", [32, 64])
def test_reorder_too_long_cluster_left_unordered(gpu, bits):
    # a cluster of ~9000 slots exceeds both passes: not marked ordered, still a valid table
    sp = 14
    keys = concentrated_keys(bits, sp, 9000, 512, seed=11)
    hm, model = build(bits, sp, keys, keys)
    before = hm.host_buckets()
    assert not hm.reorder()
    img = hm.host_buckets()
    assert np.array_equal(live_pairs(img, bits), live_pairs(before, bits))
    assert np.array_equal(occupancy(img, bits), occupancy(before, bits))
    assert reachable_all(img, bits, sp)
    absent = keyset(4000, bits, seed=3, start=1 << 20)
    absent = absent[~np.isin(absent, keys)]
    q = np.concatenate([keys, absent])
    _, v, h = hm.retrieve(q)
    ev, eh = model.retrieve(q)
    assert np.array_equal(h.cpu().numpy().astype(bool), eh)


@pytest.mark.parametrize("bits", [32, 64])
@pytest.mark.parametrize("width", [1, 4, 8, 16, 64])
def test_ordered_retrieve_and_erase_exact(gpu, bits, width):
    sp = 13
    keys = keyset(int(0.9 * (1 << sp)), bits, seed=width + bits)
    vals = keys * NDT[bits](3)
    hm, model = build(bits, sp, keys, vals, width)
    check_exact(hm, model, bits)
    absent = keyset(keys.size, bits, seed=99, start=1 << 24)
    absent = absent[~np.isin(absent, keys)]
    q = np.concatenate([keys, absent])
    np.random.default_rng(width).shuffle(q)
    sentinel = NDT[bits](777)
    vin = np.full(q.size, sentinel, dtype=NDT[bits])
    _, v, h = hm.retrieve(q, torch.from_numpy(vin.view(np.int32 if bits == 32 else np.int64)).cuda())
    ev, eh = model.retrieve(q, vin)
    assert np.array_equal(h.cpu().numpy().astype(bool), eh)
    assert np.array_equal(to_numpy(v, bits), ev)
    # erase a mix of present and absent keys on the ordered table: image exact
    e = np.concatenate([keys[::3], absent[::2]])
    hm.erase(e)
    model.erase(e)
    assert np.array_equal(hm.host_buckets().reshape(-1), model.b)
    assert info_tuple(hm)[1:4:2] == (model.info.fill, model.info.tombstones)
    assert ordered_flag(hm)  # tombstones keep the order valid
    _, v, h = hm.retrieve(q)
    ev, eh = model.retrieve(q)
    assert np.array_equal(h.cpu().numpy().astype(bool), eh)
    assert np.array_equal(to_numpy(v, bits)[eh], ev[eh])


@pytest.mark.parametrize("bits", [32, 64])
def test_flag_cleared_by_insert_clean_rehash(gpu, bits):
    sp = 16  # large enough that 64 KiB of clean scratch is below the rebuild's need
    keys = keyset(int(0.8 * (1 << sp)), bits, seed=5)
    hm, model = build(bits, sp, keys[:-100], keys[:-100])
    assert hm.reorder() and ordered_flag(hm)
    hm.insert(keys[-100:], keys[-100:], targetLF=1.0)
    assert not ordered_flag(hm)
    _, v, h = hm.retrieve(keys)
    assert h.cpu().numpy().all()
    assert hm.reorder()
    hm.erase(keys[:50])
    assert ordered_flag(hm)
    before = hm.host_buckets()
    assert hm.reorder()  # tombstones present: no-op; the flag set before stays valid
    assert np.array_equal(hm.host_buckets(), before)
    hm.clean_tombstones(scratch_budget=1 << 16)  # bounded scratch: region-by-region reinsertion
    assert not ordered_flag(hm)
    assert hm.reorder()
    hm.erase(keys[50:60])
    hm.clean_tombstones()  # whole-table scratch: rebuilt in Robin Hood order
    assert ordered_flag(hm)
    hm.device_rehash(sp + 1)
    assert not ordered_flag(hm)
    _, v, h = hm.retrieve(keys)
    assert np.array_equal(h.cpu().numpy().astype(bool), ~np.isin(keys, keys[:60]))


@pytest.mark.parametrize("bits", [32, 64])
def test_auto_reorder_policy(gpu, bits):
    sp = 12
    keys = keyset(2000, bits, seed=8)
    hm = Hashmap(sp, bits)  # default: reorder after inserts of >= 1/16 of the buckets
    hm.insert(keys, keys, targetLF=1.0)
    assert ordered_flag(hm)
    more = keyset(10, bits, seed=9, start=1 << 20)
    hm.insert(more, more, targetLF=1.0)  # small batch: no reorder
    assert not ordered_flag(hm)


@pytest.mark.parametrize("bits", [32, 64])
def test_reorder_full_table_is_noop(gpu, bits):
    sp = 6
    keys = keyset(1 << sp, bits, seed=2)
    hm, model = build(bits, sp, keys, keys)
    before = hm.host_buckets()
    assert not hm.reorder()
    assert np.array_equal(hm.host_buckets(), before)


def test_reorder_invalid_args(gpu):
    L = _lib.load()
    assert L.hn_reorder_u64(None, None, 10, None, 0, None) == _lib.HN_EINVAL
    hm = Hashmap(10, 64, auto_reorder=0, auto_build=0)
    sb = L.hn_reorder_scratch_bytes(10)
    s = torch.empty(sb, dtype=torch.uint8, device="cuda")
    assert L.hn_reorder_u64(hm.buckets.data_ptr(), hm._info.data_ptr(), 10, s.data_ptr(), sb - 1, None) == \
        _lib.HN_ESCRATCH
    assert L.hn_reorder_u64(hm.buckets.data_ptr(), hm._info.data_ptr(), 0, s.data_ptr(), sb, None) == _lib.HN_EINVAL
