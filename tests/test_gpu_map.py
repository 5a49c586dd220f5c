"""Parity of the gfx950 map kernels with the oracle, through the C-ABI.

Exact where the operation is deterministic (retrieve and erase on a given
bucket image, reset_all); for the parallel inserts -- whose slot choice
depends on scheduling, as in the reference -- the checks are the ones the
mathematics fixes: the key/value set, the occupied-slot set (order
independent under linear probing), reachability of every key, fill, and the
overflow bound derived from the device's own image.
"""
import ctypes as C

import numpy as np
import pytest
import torch

import oracle as O
from hashinator_amd import Hashmap, hn_tune, to_numpy
from tests.helpers import (EMPTY, TOMB, NDT, fresh_model, info_tuple, keyset, live_pairs, model_from,
                           next_overflow, occupancy, reachable_all)

pytestmark = pytest.mark.gpu

WIDTHS = [1, 4, 8, 16, 64]


def upload(hm, model):
    """Put an oracle bucket image and Info into a device map."""
    img = model.b.reshape(-1, 2)
    t = torch.from_numpy(img.view(np.int32 if hm.kb == 32 else np.int64).copy()).to(hm.device)
    hm.buckets = t
    hm._sp = model.info.size_power
    raw = np.frombuffer(bytes(model.info), dtype=np.int64).copy()
    hm._info.copy_(torch.from_numpy(raw))
    torch.cuda.synchronize()


def check_insert_invariants(hm, model, bits):
    img = hm.host_buckets()
    sp, fill, ovf, tomb, err = info_tuple(hm)
    assert err == 0
    assert np.array_equal(live_pairs(img, bits), live_pairs(model.b, bits))
    assert np.array_equal(occupancy(img, bits), occupancy(model.b, bits))
    assert fill == model.info.fill
    assert tomb == model.info.tombstones
    assert reachable_all(img, bits, sp)
    mp = getattr(O.lib(), f"or_gpu{bits}_max_probe")(np.ascontiguousarray(img.reshape(-1)).ctypes.data, sp)
    assert ovf % 64 == 0 and ovf >= mp


@pytest.mark.parametrize("bits", [32, 64])
@pytest.mark.parametrize("width", WIDTHS)
def test_retrieve_exact_on_oracle_image(gpu, bits, width):
    sp = 12
    rng = np.random.default_rng(bits * 100 + width)
    model = fresh_model(sp, bits)
    keys = keyset(int(0.9 * (1 << sp)), bits, seed=bits + width)
    vals = rng.integers(0, 2**31, keys.size).astype(NDT[bits])
    model.insert(keys, vals)
    model.erase(keys[::7])  # tombstones on the probe paths
    hm = Hashmap(sp, bits, tune=hn_tune(width, 256, 0, 0), auto_build=0)
    upload(hm, model)
    absent = keyset(keys.size, bits, seed=bits + width, start=keys.size + 2)
    q = np.concatenate([keys, absent])
    rng.shuffle(q)
    sentinel = NDT[bits](12345)
    vin = np.full(q.size, sentinel, dtype=NDT[bits])
    _, v, h = hm.retrieve(q, torch.from_numpy(vin.view(np.int32 if bits == 32 else np.int64)).cuda())
    v, h = to_numpy(v, bits), h.cpu().numpy().astype(bool)
    ev, eh = model.retrieve(q, vin)
    assert np.array_equal(h, eh)
    assert np.array_equal(v, ev)  # misses left untouched
    # the reference host find agrees on membership
    cpu = O.CpuMap(sp, bits)
    cpu.insert(keys, vals)
    cpu.erase(keys[::7])
    cv, ch = cpu.find(q)
    assert np.array_equal(ch, eh) and np.array_equal(cv[ch], ev[eh])


@pytest.mark.parametrize("bits", [32, 64])
@pytest.mark.parametrize("width", WIDTHS)
@pytest.mark.parametrize("lf", [0.5, 0.9, 0.99])
def test_insert_set_and_occupancy(gpu, bits, width, lf):
    sp = 14
    keys = keyset(int(lf * (1 << sp)), bits, seed=int(lf * 100) + width)
    vals = (keys * NDT[bits](3) + NDT[bits](1)).astype(NDT[bits])
    hm = Hashmap(sp, bits, tune=hn_tune(width, 256, 0, 0), auto_reorder=0, auto_build=0)
    hm.insert(keys, vals, targetLF=1.0)
    model = fresh_model(sp, bits)
    model.insert(keys, vals)
    check_insert_invariants(hm, model, bits)
    # overflow is exactly the rounded longest probe of the device's own (arrival-order) image
    img = hm.host_buckets().reshape(-1)
    mp = getattr(O.lib(), f"or_gpu{bits}_max_probe")(np.ascontiguousarray(img).ctypes.data, sp)
    assert info_tuple(hm)[2] == max(64, next_overflow(mp))


@pytest.mark.parametrize("bits", [32, 64])
@pytest.mark.parametrize("width", [1, 4, 8, 64])
def test_insert_over_empty_slots_with_garbage_values(gpu, bits, width):
    # EMPTY slots may hold any value bits (a table the build wrote over
    # uninitialised memory, or user memory): the 32-bit claim compares the
    # whole 8-byte slot against the value its window load saw, so it must
    # still claim such slots, and no EMPTY slot's value may change.
    sp = 12
    rng = np.random.default_rng(40 + width)
    model = fresh_model(sp, bits)
    model.b[1::2] = rng.integers(0, np.iinfo(NDT[bits]).max, (1 << sp), dtype=NDT[bits])
    hm = Hashmap(sp, bits, tune=hn_tune(width, 256, 0, 0), auto_reorder=0, auto_build=0)
    upload(hm, model)
    before = hm.host_buckets().copy()
    keys = keyset(int(0.9 * (1 << sp)), bits, seed=width + bits)
    vals = (keys * NDT[bits](5) + NDT[bits](2)).astype(NDT[bits])
    hm.insert(keys, vals, targetLF=1.0)
    model.insert(keys, vals)
    check_insert_invariants(hm, model, bits)
    img = hm.host_buckets()
    still_empty = img[:, 0] == EMPTY[bits]
    assert np.array_equal(img[still_empty], before[still_empty])


@pytest.mark.parametrize("bits", [32, 64])
def test_insert_duplicates_and_overwrite(gpu, bits):
    sp = 12
    rng = np.random.default_rng(5)
    keys = keyset(1500, bits, seed=99)
    dup = np.concatenate([keys, keys[:700], keys[:300]])
    rng.shuffle(dup)
    vals = (dup ^ NDT[bits](0x5A5A)).astype(NDT[bits])  # value a function of the key
    hm = Hashmap(sp, bits, auto_build=0)
    hm.insert(dup, vals, targetLF=1.0)
    model = fresh_model(sp, bits)
    model.insert(dup, vals)
    check_insert_invariants(hm, model, bits)
    # second batch overwrites values of present keys and adds new ones
    more = np.concatenate([keys[:400], keyset(300, bits, seed=99, start=2000)])
    mv = (more + NDT[bits](7)).astype(NDT[bits])
    hm.insert(more, mv, targetLF=1.0)
    model.insert(more, mv)
    check_insert_invariants(hm, model, bits)


@pytest.mark.parametrize("bits", [32, 64])
def test_insert_index_and_pairs(gpu, bits):
    sp = 11
    keys = keyset(1200, bits, seed=3)
    hm = Hashmap(sp, bits, auto_build=0)
    hm.insertIndex(keys, targetLF=1.0)
    model = fresh_model(sp, bits)
    model.insert_index(keys)
    check_insert_invariants(hm, model, bits)
    pairs = np.stack([keyset(500, bits, seed=4), np.arange(500, dtype=NDT[bits])], axis=1)
    hm2 = Hashmap(sp, bits, auto_build=0)
    hm2.insert(pairs, None, targetLF=1.0)
    m2 = fresh_model(sp, bits)
    m2.insert_pairs(pairs.reshape(-1))
    check_insert_invariants(hm2, m2, bits)
    # pairs retrieve writes .second in place
    q = np.stack([pairs[:, 0], np.zeros(500, dtype=NDT[bits])], axis=1)
    out, h = hm2.retrieve(torch.from_numpy(q.view(np.int32 if bits == 32 else np.int64)).cuda())
    out = to_numpy(out, bits).reshape(-1, 2)
    assert h.cpu().numpy().all()
    assert np.array_equal(out[:, 1], pairs[:, 1])


@pytest.mark.parametrize("bits", [32, 64])
def test_insert_into_tombstoned_map(gpu, bits):
    sp = 12
    keys = keyset(3000, bits, seed=21)
    vals = keys.astype(NDT[bits])
    hm = Hashmap(sp, bits, auto_build=0)
    hm.insert(keys, vals, targetLF=1.0)
    hm.erase(keys[::3])
    img = hm.host_buckets()
    model = model_from(img, info_tuple(hm), bits)
    new = keyset(600, bits, seed=21, start=3002)
    back = keys[::3][:200]  # erased keys come back (tombstones are not reused)
    ins = np.concatenate([new, back])
    iv = (ins + NDT[bits](11)).astype(NDT[bits])
    hm.insert(ins, iv, targetLF=1.0)
    model.insert(ins, iv)
    check_insert_invariants(hm, model, bits)


@pytest.mark.parametrize("bits", [32, 64])
@pytest.mark.parametrize("width", [1, 4, 16, 64])
def test_erase_exact(gpu, bits, width):
    sp = 13
    keys = keyset(int(0.85 * (1 << sp)), bits, seed=31 + width)
    hm = Hashmap(sp, bits, tune=hn_tune(width, 256, 0, 0), auto_build=0)
    hm.insert(keys, keys, targetLF=1.0)
    model = model_from(hm.host_buckets(), info_tuple(hm), bits)
    rng = np.random.default_rng(width)
    er = np.concatenate([rng.choice(keys, keys.size // 4, replace=False), keyset(500, bits, 31 + width, 9000)])
    er = np.concatenate([er, er[:100]])  # repeats in one batch count once
    hm.erase(er)
    model.erase(er)
    assert np.array_equal(hm.host_buckets().reshape(-1), model.b)
    assert info_tuple(hm)[:4] == model.info.as_tuple()[:4]


@pytest.mark.parametrize("bits", [32, 64])
def test_reset_all_exact(gpu, bits):
    sp = 10
    keys = keyset(700, bits, seed=41)
    hm = Hashmap(sp, bits, auto_build=0)
    hm.insert(keys, keys, targetLF=1.0)
    hm.erase(keys[:100])
    model = model_from(hm.host_buckets(), info_tuple(hm), bits)
    hm.clear()
    model.reset_all()
    assert np.array_equal(hm.host_buckets().reshape(-1), model.b)
    assert info_tuple(hm)[:4] == model.info.as_tuple()[:4]


def _tombstoned(bits, sp, seed):
    keys = keyset(int(0.95 * (1 << sp)), bits, seed=seed)
    model = fresh_model(sp, bits)
    model.insert(keys, (keys ^ NDT[bits](0x5A5A)).astype(NDT[bits]))
    model.erase(keys[::5])
    model.info.err = 0
    return model


def _call(hm, name, *args):
    rc = hm._fn(name)(*args, C.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0, (name, rc)
    torch.cuda.synchronize()


@pytest.mark.parametrize("bits", [32, 64])
@pytest.mark.parametrize("n_len", [0, 1, 333, 1024])
def test_reset_all_lengths_exact(gpu, bits, n_len):
    """reset_all over the first len slots of a tombstoned table with values
    and a raised overflow: image and Info as the oracle's (len 0: no change)."""
    sp = 10
    model = _tombstoned(bits, sp, 71)
    hm = Hashmap(sp, bits, auto_build=0)
    upload(hm, model)
    _call(hm, "reset_all", C.c_void_p(hm.bucket_ptr()), C.c_void_p(hm.info_ptr()), n_len)
    model.reset_all(n_len)
    assert np.array_equal(hm.host_buckets().reshape(-1), model.b)
    assert info_tuple(hm) == model.info.as_tuple()


@pytest.mark.parametrize("bits", [32, 64])
@pytest.mark.parametrize("sp", [6, 11, 16])
def test_reset_exact(gpu, bits, sp):
    """Hasher::reset (hashers.h:144-152): present keys, absent keys and
    repeats on a tombstoned table -- image and Info equal the oracle's."""
    model = _tombstoned(bits, sp, 73 + sp)
    hm = Hashmap(sp, bits, auto_build=0)
    upload(hm, model)
    img0 = model.b.reshape(-1, 2)
    present = img0[(img0[:, 0] != EMPTY[bits]) & (img0[:, 0] != TOMB[bits])]
    rng = np.random.default_rng(sp)
    pick = present[rng.choice(present.shape[0], present.shape[0] // 3, replace=False)]
    absent = keyset(40, bits, seed=sp, start=10**6)
    src = np.concatenate([pick, np.stack([absent, absent], axis=1), pick[:7]]).astype(NDT[bits])
    rng.shuffle(src)
    d = torch.from_numpy(src.view(np.int32 if bits == 32 else np.int64).copy()).cuda()
    _call(hm, "reset", C.c_void_p(d.data_ptr()), C.c_void_p(hm.bucket_ptr()), C.c_void_p(hm.info_ptr()),
          src.shape[0])
    model.reset(src)
    assert np.array_equal(hm.host_buckets().reshape(-1), model.b)
    assert info_tuple(hm) == model.info.as_tuple()


@pytest.mark.parametrize("bits", [32, 64])
def test_reset_first_copy_in_probe_order(gpu, bits):
    """Two copies of a key, the first in probe order behind an EMPTY home slot
    and across the table's end: the device empties the same slots as the
    oracle, one call at a time."""
    from tests.test_oracle import _key_with_home

    sp = 3
    E = EMPTY[bits]
    a = _key_with_home(bits, sp, 6)
    img = np.array([[E, 0], [a, 11], [E, 0], [E, 0], [E, 0], [E, 0], [E, 0], [a, 77]], dtype=NDT[bits])
    model = model_from(img, (sp, 2, 64, 0, 0), bits)
    hm = Hashmap(sp, bits, auto_build=0)
    upload(hm, model)
    src = np.array([[a, 0]], dtype=NDT[bits])
    d = torch.from_numpy(src.view(np.int32 if bits == 32 else np.int64).copy()).cuda()
    for _ in range(2):
        _call(hm, "reset", C.c_void_p(d.data_ptr()), C.c_void_p(hm.bucket_ptr()), C.c_void_p(hm.info_ptr()), 1)
        model.reset(src)
        assert np.array_equal(hm.host_buckets().reshape(-1), model.b)
        assert info_tuple(hm) == model.info.as_tuple()


@pytest.mark.parametrize("bits", [32, 64])
def test_reset_then_reinsert(gpu, bits):
    """clean_tombstones' own use of reset (hashinator.h:1299-1311): the
    displaced live elements are reset and inserted again; live set, occupied
    slots and Info equal the oracle's."""
    sp = 12
    model = _tombstoned(bits, sp, 79)
    hm = Hashmap(sp, bits, auto_build=0, auto_reorder=0)
    upload(hm, model)
    img0 = model.b.reshape(-1, 2)
    live = (img0[:, 0] != EMPTY[bits]) & (img0[:, 0] != TOMB[bits])
    home = O.fibhash(img0[:, 0], sp, bits).astype(np.int64) & ((1 << sp) - 1)
    disp = img0[live & (home != np.arange(1 << sp))].copy()
    assert disp.shape[0] > 100
    d = torch.from_numpy(disp.view(np.int32 if bits == 32 else np.int64).copy()).cuda()
    _call(hm, "reset", C.c_void_p(d.data_ptr()), C.c_void_p(hm.bucket_ptr()), C.c_void_p(hm.info_ptr()),
          disp.shape[0])
    model.reset(disp)
    assert np.array_equal(hm.host_buckets().reshape(-1), model.b)
    hm.insert(d, targetLF=1.0)
    model.insert_pairs(disp)
    img = hm.host_buckets()
    assert np.array_equal(live_pairs(img, bits), live_pairs(model.b, bits))
    assert np.array_equal(occupancy(img, bits), occupancy(model.b, bits))
    sp_, fill, ovf, tomb, err = info_tuple(hm)
    assert (fill, tomb, err) == (model.info.fill, model.info.tombstones, 0)


@pytest.mark.parametrize("bits", [32, 64])
def test_insert_copy_in_window_wins_over_hole(gpu, bits):
    """A key whose copy sits behind an EMPTY hole in the same probe window is
    updated in place, not stored twice (insert_kernel, kernels_AMD.h:339-349)."""
    from tests.test_oracle import _key_with_home

    sp = 3
    E = EMPTY[bits]
    a = _key_with_home(bits, sp, 6)
    img = np.array([[E, 0], [E, 0], [E, 0], [E, 0], [E, 0], [E, 0], [E, 0], [a, 77]], dtype=NDT[bits])
    model = model_from(img, (sp, 1, 64, 0, 0), bits)
    hm = Hashmap(sp, bits, auto_build=0, auto_reorder=0, tune=hn_tune(8, 256, 0, 0))
    upload(hm, model)
    hm.insert(np.array([a], dtype=NDT[bits]), np.array([5], dtype=NDT[bits]), targetLF=1.0)
    exp = img.copy()
    exp[7, 1] = 5
    assert np.array_equal(hm.host_buckets(), exp)
    assert info_tuple(hm)[1] == 1


@pytest.mark.parametrize("bits", [32, 64])
@pytest.mark.parametrize("budget", [0, 1 << 16])
def test_clean_tombstones(gpu, bits, budget):
    sp = 14
    keys = keyset(int(0.9 * (1 << sp)), bits, seed=51)
    vals = (keys ^ NDT[bits](77)).astype(NDT[bits])
    hm = Hashmap(sp, bits, auto_reorder=0, auto_build=0)  # the overflow check below is on clean's own image
    hm.insert(keys, vals, targetLF=1.0)
    rng = np.random.default_rng(budget)
    er = rng.choice(keys, keys.size // 4, replace=False)
    hm.erase(er)
    model = model_from(hm.host_buckets(), info_tuple(hm), bits)
    model.clean_tombstones()
    hm.clean_tombstones(scratch_budget=budget)
    img = hm.host_buckets()
    sp_, fill, ovf, tomb, err = info_tuple(hm)
    assert tomb == 0 and not np.any(img[:, 0] == TOMB[bits])
    assert fill == model.info.fill
    assert np.array_equal(live_pairs(img, bits), live_pairs(model.b, bits))
    assert np.array_equal(occupancy(img, bits), occupancy(model.b, bits))
    assert reachable_all(img, bits, sp)
    mp = getattr(O.lib(), f"or_gpu{bits}_max_probe")(np.ascontiguousarray(img.reshape(-1)).ctypes.data, sp)
    assert ovf == max(64, next_overflow(mp)) and err == 0
    # every surviving key is found, every erased key is not
    _, v, h = hm.retrieve(np.concatenate([np.setdiff1d(keys, er), er]))
    h = h.cpu().numpy().astype(bool)
    assert h[: keys.size - er.size].all() and not h[keys.size - er.size:].any()


@pytest.mark.parametrize("bits", [32, 64])
def test_clean_no_tombstones_is_noop(gpu, bits):
    hm = Hashmap(10, bits, auto_build=0)
    keys = keyset(600, bits, seed=61)
    hm.insert(keys, keys, targetLF=1.0)
    before = hm.host_buckets().copy()
    i0 = info_tuple(hm)
    hm.clean_tombstones()
    assert np.array_equal(before, hm.host_buckets())
    assert info_tuple(hm) == i0


@pytest.mark.parametrize("bits", [32, 64])
@pytest.mark.parametrize("new_sp", [12, 13, 15])
def test_device_rehash(gpu, bits, new_sp):
    sp = 13
    keys = keyset(int(0.7 * (1 << sp)), bits, seed=71)
    vals = (keys * NDT[bits](5)).astype(NDT[bits])
    hm = Hashmap(sp, bits, auto_build=0)
    hm.insert(keys, vals, targetLF=1.0)
    hm.erase(keys[:500])
    live = live_pairs(hm.host_buckets(), bits)
    hm.device_rehash(new_sp)
    img = hm.host_buckets()
    sp_, fill, ovf, tomb, err = info_tuple(hm)
    assert sp_ == max(new_sp, int(np.ceil(np.log2(live.shape[0])))) == hm.getSizePower()
    assert err == 0 and tomb == 0 and fill == live.shape[0]
    assert np.array_equal(live_pairs(img, bits), live)
    model = fresh_model(sp_, bits)
    model.insert(live[:, 0].copy(), live[:, 1].copy())
    assert np.array_equal(occupancy(img, bits), occupancy(model.b, bits))
    assert reachable_all(img, bits, sp_)


@pytest.mark.parametrize("bits", [32, 64])
def test_insert_overflow_sets_fail(gpu, bits):
    sp = 6
    keys = keyset(80, bits, seed=81)  # more keys than slots
    hm = Hashmap(sp, bits, auto_build=0)
    hm.insert(keys, keys, targetLF=100.0)  # no auto-grow
    sp_, fill, ovf, tomb, err = info_tuple(hm)
    assert fill == 64 and err == 1
    assert not np.any(hm.host_buckets()[:, 0] == EMPTY[bits])


@pytest.mark.parametrize("bits", [32, 64])
def test_auto_grow_on_insert(gpu, bits):
    hm = Hashmap(5, bits)
    keys = keyset(1000, bits, seed=91)
    hm.insert(keys, keys)  # targetLF 0.5 -> grows to 2^11
    assert hm.getSizePower() == 11
    _, v, h = hm.retrieve(keys)
    assert h.cpu().numpy().all() and np.array_equal(to_numpy(v, bits), keys)


@pytest.mark.parametrize("bits", [32, 64])
def test_tiny_tables(gpu, bits):
    for sp in (1, 2, 3):
        hm = Hashmap(sp, bits, tune=hn_tune(64, 256, 0, 0), auto_build=0)
        keys = keyset((1 << sp) - 1, bits, seed=sp)
        hm.insert(keys, keys, targetLF=100.0)
        _, v, h = hm.retrieve(keys)
        assert h.cpu().numpy().all() and np.array_equal(to_numpy(v, bits), keys)
        assert hm.size() == keys.size and hm.peek_status() == 0


@pytest.mark.parametrize("bits", [32, 64])
def test_empty_batches(gpu, bits):
    hm = Hashmap(8, bits, auto_build=0)
    hm.insert(np.zeros(0, NDT[bits]), np.zeros(0, NDT[bits]))
    assert hm.peek_status() == 0
    _, v, h = hm.retrieve(np.zeros(0, NDT[bits]))
    hm.erase(np.zeros(0, NDT[bits]))
    assert hm.size() == 0


@pytest.mark.parametrize("bits", [32, 64])
def test_side_stream_growth_rehash_and_compaction(gpu, bits):
    """Every call on a non-default stream: an insert that grows the map (new
    table and scratch on that stream), retrieve, device_rehash, clean and
    compaction read back on it -- results as on the default stream."""
    from hashinator_amd import tools

    s = torch.cuda.Stream()
    keys = keyset(5000, bits, seed=91)
    vals = (keys ^ NDT[bits](3)).astype(NDT[bits])
    hm = Hashmap(8, bits, auto_build=0)
    hm.insert(keys, vals, targetLF=0.5, stream=s)  # 2^8 -> 2^14 slots
    assert hm.getSizePower() == 14 and hm.size() == keys.size
    _, v, h = hm.retrieve(keys, stream=s)
    assert h.cpu().numpy().all() and np.array_equal(to_numpy(v, bits), vals)
    hm.erase(keys[::3], stream=s)
    hm.device_rehash(15, stream=s)
    hm.clean_tombstones(stream=s)
    live = np.stack([np.delete(keys, np.s_[::3]), np.delete(vals, np.s_[::3])], axis=1)
    live = live[np.argsort(live[:, 0])]
    assert np.array_equal(live_pairs(hm.host_buckets(), bits), live)
    got = to_numpy(hm.extractPattern(stream=s), bits).reshape(-1, 2)
    assert np.array_equal(got[np.argsort(got[:, 0])], live)
    flags = torch.from_numpy((np.arange(hm.bucket_count()) % 3 == 0).astype(np.uint8)).cuda()
    torch.cuda.synchronize()  # the flags are copied on the default stream
    out = tools.copy_if_flags(hm.buckets, flags, stream=s)
    snap = hm.host_buckets()
    assert np.array_equal(to_numpy(out, bits).reshape(-1, 2), snap[::3])
