"""device_rehash into a larger table as one bulk build (hn_rehash_build_*):
the new table's image equals the oracle's fresh insert of the old table's
live pairs followed by its Robin Hood reorder, bit for bit (the reordered
image is a function of the key set), and the Info matches (size power, fill,
overflow, no tombstones, ordered flag, err = SUCCESS).  The map's
device_rehash takes this path when it grows a map of many keys."""
import numpy as np
import pytest
import torch

import oracle as O
from hashinator_amd import Hashmap, to_numpy
from hashinator_amd import _lib
from tests.helpers import NDT, keyset, live_pairs, occupancy, reachable_all
from tests.test_gpu_build import concentrated_keys

pytestmark = pytest.mark.gpu

TDT = {32: torch.int32, 64: torch.int64}


def old_map(bits, sp, keys, erase_frac, seed):
    vals = (keys * NDT[bits](7) + NDT[bits](1)).astype(NDT[bits])
    hm = Hashmap(sp, bits, auto_build=0, auto_reorder=0)
    if keys.size:
        hm.insert(keys, vals, targetLF=1.0)
    if erase_frac:
        er = keys[np.random.default_rng(seed).choice(keys.size, int(erase_frac * keys.size), replace=False)]
        hm.erase(er)
    return hm


def rehash_build(bits, hm, new_sp, seed=0):
    """Direct C-ABI call into a table holding garbage; returns (image, info words, rc)."""
    L = _lib.load()
    sfx = "u32" if bits == 32 else "u64"
    old_sp = hm.getSizePower()
    junk = np.random.default_rng(seed).integers(0, 2**62, size=(1 << new_sp) * 2, dtype=np.uint64)
    nb = torch.from_numpy(junk.astype(NDT[bits]).view(np.int32 if bits == 32 else np.int64)).cuda().reshape(-1, 2)
    sb = getattr(L, f"hn_rehash_build_scratch_bytes_{sfx}")(old_sp, new_sp)
    assert sb > 0
    sc = torch.empty(sb, dtype=torch.uint8, device="cuda")
    rc = getattr(L, f"hn_rehash_build_{sfx}")(hm.bucket_ptr(), old_sp, nb.data_ptr(), new_sp, hm.info_ptr(),
                                              sc.data_ptr(), sb, None)
    torch.cuda.synchronize()
    return to_numpy(nb, bits).reshape(-1), hm._info.cpu().numpy(), rc


@pytest.mark.parametrize("bits", [32, 64])
# 2^22 old slots: the first partition pass reads the old table in the 8,192-pair tiles.
@pytest.mark.parametrize("old_sp,new_sp,lf,erase", [(3, 4, 0.75, 0.0), (10, 11, 0.9, 0.25), (12, 14, 0.7, 0.0),
                                                     (14, 15, 0.95, 0.3), (16, 17, 0.85, 0.1), (20, 21, 0.9, 0.25),
                                                     (12, 13, 0.0, 0.0), (22, 23, 0.85, 0.2)])
def test_rehash_build_image_exact(gpu, bits, old_sp, new_sp, lf, erase):
    keys = keyset(int(lf * (1 << old_sp)), bits, seed=old_sp + bits)
    hm = old_map(bits, old_sp, keys, erase, seed=old_sp)
    live = live_pairs(hm.host_buckets(), bits)
    img, inf, rc = rehash_build(bits, hm, new_sp, seed=new_sp)
    assert rc == 0
    m = O.GpuModel(new_sp, bits)
    if live.shape[0]:
        m.insert(live[:, 0].copy(), live[:, 1].copy())
    assert m.reorder() == 1
    assert np.array_equal(img, m.b)
    assert int(inf[0]) == new_sp
    assert int(inf[1]) == live.shape[0] == m.info.fill
    assert int(inf[2]) == m.info.overflow
    assert int(inf[3]) == 0
    assert (int(inf[4]) & 0xFFFFFFFF) == 0  # err = SUCCESS: fill preserved
    assert (int(inf[4]) >> 32) == _lib.HN_INFO_ORDERED


@pytest.mark.parametrize("bits", [32, 64])
def test_rehash_build_overfull_region_falls_back(gpu, bits):
    # 6000 keys homed, at the new size, in one 2048-slot region: over the
    # region capacity, so the new table is filled by the probing insert
    new_sp = 15
    keys = concentrated_keys(bits, new_sp, 6000, 2048, seed=8)
    hm = old_map(bits, 14, keys, 0.0, seed=1)
    live = live_pairs(hm.host_buckets(), bits)
    assert live.shape[0] == keys.size
    img, inf, rc = rehash_build(bits, hm, new_sp, seed=3)
    assert rc == 0
    assert (int(inf[4]) & 0xFFFFFFFF) == 0 and int(inf[1]) == keys.size and int(inf[3]) == 0
    assert (int(inf[4]) >> 32) == 0  # arrival order: not marked ordered
    m = O.GpuModel(new_sp, bits)
    m.insert(live[:, 0].copy(), live[:, 1].copy())
    assert np.array_equal(live_pairs(img, bits), live)
    assert np.array_equal(occupancy(img, bits), occupancy(m.b, bits))
    assert reachable_all(img, bits, new_sp)


@pytest.mark.parametrize("bits", [32, 64])
def test_map_device_rehash_uses_build(gpu, bits):
    sp = 14
    keys = keyset(int(0.85 * (1 << sp)), bits, seed=12)
    vals = (keys ^ NDT[bits](0x55)).astype(NDT[bits])
    hm = Hashmap(sp, bits, build_min_keys=0, auto_build=1.0 / 16)
    hm.insert(keys, vals, targetLF=1.0)
    hm.erase(keys[::5])
    live = live_pairs(hm.host_buckets(), bits)
    hm.device_rehash(sp + 2)
    assert hm.getSizePower() == sp + 2 and hm.ordered()
    m = O.GpuModel(sp + 2, bits)
    m.insert(live[:, 0].copy(), live[:, 1].copy())
    assert m.reorder() == 1
    assert np.array_equal(hm.host_buckets().reshape(-1), m.b)
    i = hm.info()
    assert (i.fill, i.tombstone_counter, i.err) == (live.shape[0], 0, 0)
    _, v, h = hm.retrieve(live[:, 0].copy())
    assert h.cpu().numpy().all() and np.array_equal(to_numpy(v, bits), live[:, 1])
    # a map below the key threshold grows by the probing reinsert (then reordered)
    hm2 = Hashmap(sp, bits)  # build_min_keys = 2^17
    hm2.insert(keys, vals, targetLF=1.0)
    hm2.erase(keys[::5])
    hm2.device_rehash(sp + 2)
    assert np.array_equal(hm2.host_buckets().reshape(-1), m.b)


def test_rehash_build_invalid_args(gpu):
    L = _lib.load()
    for sfx in ("u32", "u64"):
        sb = getattr(L, f"hn_rehash_build_scratch_bytes_{sfx}")
        assert sb(10, 10) == 0 and sb(10, 9) == 0 and sb(0, 4) == 0 and sb(10, 33) == 0
        assert sb(10, 11) > 0 and sb(10, 12) == sb(10, 12)
    hm = Hashmap(10, 64)
    b2 = torch.empty((1 << 11, 2), dtype=torch.int64, device="cuda")
    sb = L.hn_rehash_build_scratch_bytes_u64(10, 11)
    sc = torch.empty(sb, dtype=torch.uint8, device="cuda")
    f = L.hn_rehash_build_u64
    assert f(hm.bucket_ptr(), 10, b2.data_ptr(), 10, hm.info_ptr(), sc.data_ptr(), sb, None) == _lib.HN_EINVAL  # same size
    assert f(hm.bucket_ptr(), 10, b2.data_ptr(), 9, hm.info_ptr(), sc.data_ptr(), sb, None) == _lib.HN_EINVAL  # shrink
    assert f(hm.bucket_ptr(), 10, hm.bucket_ptr(), 11, hm.info_ptr(), sc.data_ptr(), sb, None) == _lib.HN_EINVAL  # overlap
    assert f(hm.bucket_ptr(), 10, b2.data_ptr(), 11, hm.info_ptr(), sc.data_ptr(), sb - 1, None) == _lib.HN_ESCRATCH
    assert f(hm.bucket_ptr(), 10, b2.data_ptr(), 11, None, sc.data_ptr(), sb, None) == _lib.HN_EINVAL
    torch.cuda.synchronize()
    assert hm.getSizePower() == 10 and hm.info().size_power == 10  # nothing ran
