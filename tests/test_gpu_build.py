"""Bulk insert into an empty table (hn_build_*): the image equals the oracle's
serial insert followed by its Robin Hood reorder, bit for bit (the reordered
image is a function of the key set; the oracle's reorder is pinned to the
reference's own map in tests/test_oracle.py), and the Info matches
(fill, overflow recomputed, ordered flag, err)."""
import ctypes as C

import numpy as np
import pytest
import torch

import oracle as O
from hashinator_amd import Hashmap, hn_tune, to_numpy
from hashinator_amd import _lib
from tests.helpers import NDT, keyset, live_pairs, occupancy, reachable_all

pytestmark = pytest.mark.gpu

TDT = {32: torch.int32, 64: torch.int64}
VDT = {32: np.int32, 64: np.int64}


def dev(a, bits):
    return torch.from_numpy(np.ascontiguousarray(a).view(VDT[bits])).cuda()


def run_build(bits, sp, keys, vals=None, pairs=False, digit_bits=0, info=None, table=None):
    """Direct C-ABI call on a fresh table; returns (image, info words, rc)."""
    L = _lib.load()
    sfx = "u32" if bits == 32 else "u64"
    N = 1 << sp
    b = table if table is not None else torch.empty((N, 2), dtype=TDT[bits], device="cuda")
    if table is None:
        _lib.check(getattr(L, f"hn_table_init_{sfx}")(b.data_ptr(), N, None), "init")
    inf = info if info is not None else torch.zeros(5, dtype=torch.int64, device="cuda")
    if info is None:
        _lib.check(L.hn_info_init(inf.data_ptr(), sp, None), "info")
    sb = getattr(L, f"hn_build_scratch_bytes_{sfx}")(sp, keys.size)
    sc = torch.empty(sb, dtype=torch.uint8, device="cuda")
    tune = C.byref(hn_tune(0, 0, 0, digit_bits)) if digit_bits else None
    if pairs:
        pr = np.stack([keys, vals], axis=1)
        dp = dev(pr, bits)
        rc = getattr(L, f"hn_build_pairs_{sfx}")(dp.data_ptr(), b.data_ptr(), inf.data_ptr(), keys.size, sp,
                                                 sc.data_ptr(), sb, None, tune)
    else:
        dk = dev(keys, bits)
        dv = dev(vals, bits) if vals is not None else None
        rc = getattr(L, f"hn_build_{sfx}")(dk.data_ptr(), dv.data_ptr() if dv is not None else None, b.data_ptr(),
                                           inf.data_ptr(), keys.size, sp, sc.data_ptr(), sb, None, tune)
    torch.cuda.synchronize()
    img = to_numpy(b, bits).reshape(-1)
    return img, inf.cpu().numpy(), rc


def expect(bits, sp, keys, vals):
    m = O.GpuModel(sp, bits)
    m.insert(keys, vals)
    assert m.reorder() == 1
    return m


def check(img, inf, m):
    assert np.array_equal(img, m.b)
    assert int(inf[1]) == m.info.fill
    assert int(inf[2]) == m.info.overflow
    assert int(inf[3]) == 0
    assert (int(inf[4]) & 0xFFFFFFFF) == 0  # err = SUCCESS
    assert (int(inf[4]) >> 32) == _lib.HN_INFO_ORDERED


@pytest.mark.parametrize("bits", [32, 64])
# sp 23 at LF 0.9 (7.5 M keys) takes the 8,192-pair partition tiles (inputs of 2^22 pairs and more).
@pytest.mark.parametrize("sp,lf", [(1, 0.5), (3, 0.9), (8, 0.99), (11, 0.9), (12, 0.5), (14, 0.9), (16, 0.95),
                                   (20, 0.9), (21, 0.9), (23, 0.9)])
def test_build_image_exact(gpu, bits, sp, lf):
    n = max(1, min((1 << sp) - 1, int(lf * (1 << sp))))
    keys = keyset(n, bits, seed=sp * 7 + bits)
    vals = (keys ^ NDT[bits](0x1234)).astype(NDT[bits])
    img, inf, rc = run_build(bits, sp, keys, vals)
    assert rc == 0
    check(img, inf, expect(bits, sp, keys, vals))


@pytest.mark.parametrize("bits", [32, 64])
@pytest.mark.parametrize("digit_bits", [1, 2, 4])
def test_build_multi_pass(gpu, bits, digit_bits):
    # 2^16 slots = 32 regions (5 region bits): digits of 1/2/4 bits need 3/3/2 passes
    # (3 passes is the cap; 1-bit digits would need 5 and are refused)
    sp = 16
    keys = keyset(int(0.9 * (1 << sp)), bits, seed=digit_bits)
    vals = np.arange(keys.size).astype(NDT[bits])
    img, inf, rc = run_build(bits, sp, keys, vals, digit_bits=digit_bits)
    if digit_bits == 1:
        assert rc == _lib.HN_EINVAL
        return
    assert rc == 0
    check(img, inf, expect(bits, sp, keys, vals))


@pytest.mark.parametrize("bits", [32, 64])
def test_build_index_and_pairs(gpu, bits):
    sp = 15
    keys = keyset(int(0.85 * (1 << sp)), bits, seed=3)
    img, inf, rc = run_build(bits, sp, keys, None)  # vals NULL: index as value
    assert rc == 0
    check(img, inf, expect(bits, sp, keys, np.arange(keys.size).astype(NDT[bits])))
    vals = (keys * NDT[bits](7)).astype(NDT[bits])
    img, inf, rc = run_build(bits, sp, keys, vals, pairs=True)
    assert rc == 0
    check(img, inf, expect(bits, sp, keys, vals))


@pytest.mark.parametrize("bits", [32, 64])
@pytest.mark.parametrize("sp,nb,nr", [(14, 9000, 3000), (23, 4_000_000, 1_000_000)])  # small and large partition tiles
def test_build_duplicates_last_wins(gpu, bits, sp, nb, nr):
    base = keyset(nb, bits, seed=4)
    rng = np.random.default_rng(4)
    rep = base[rng.choice(base.size, nr, replace=True)]
    keys = np.concatenate([base, rep])
    rng.shuffle(keys)
    vals = np.arange(keys.size).astype(NDT[bits])
    img, inf, rc = run_build(bits, sp, keys, vals)
    assert rc == 0
    m = expect(bits, sp, keys, vals)  # serial insert: the last occurrence's value
    check(img, inf, m)
    assert m.info.fill == np.unique(keys).size


@pytest.mark.parametrize("bits", [32, 64])
def test_build_refuses_nonempty_table(gpu, bits):
    sp = 12
    keys = keyset(2000, bits, seed=5)
    hm = Hashmap(sp, bits, auto_build=0, auto_reorder=0)
    hm.insert(keys[:100], keys[:100], targetLF=1.0)
    before = hm.host_buckets().copy()
    img, inf, rc = run_build(bits, sp, keys[100:], keys[100:], info=hm._info, table=hm.buckets)
    assert rc == 0
    assert (int(inf[4]) & 0xFFFFFFFF) == _lib.HN_STATUS_INVALID
    assert int(inf[1]) == 100
    assert np.array_equal(hm.host_buckets(), before)


def concentrated_keys(bits, sp, n, span, seed):
    out = []
    rng = np.random.default_rng(seed)
    while sum(len(o) for o in out) < n:
        c = rng.integers(1, 2**31 if bits == 32 else 2**62, 1 << 20, dtype=np.uint64).astype(NDT[bits])
        h = O.fibhash(c, sp, bits).astype(np.uint64) & np.uint64((1 << sp) - 1)
        out.append(c[h < span])
    k = np.unique(np.concatenate(out))[:n]
    return k[np.random.default_rng(seed + 1).permutation(k.size)]


@pytest.mark.parametrize("bits", [32, 64])
def test_build_overfull_region_falls_back_on_device(gpu, bits):
    # 3000 keys homed in one 2048-slot region: over the 2560-key region
    # capacity, so the table is filled by the probing insert instead
    sp = 14
    keys = concentrated_keys(bits, sp, 3000, 2048, seed=6)
    vals = (keys ^ NDT[bits](3)).astype(NDT[bits])
    img, inf, rc = run_build(bits, sp, keys, vals)
    assert rc == 0
    assert (int(inf[4]) & 0xFFFFFFFF) == 0 and int(inf[1]) == keys.size and int(inf[3]) == 0
    assert (int(inf[4]) >> 32) == 0  # arrival order: not marked ordered
    m = O.GpuModel(sp, bits)
    m.insert(keys, vals)
    assert np.array_equal(live_pairs(img, bits), live_pairs(m.b, bits))
    assert np.array_equal(occupancy(img, bits), occupancy(m.b, bits))
    assert reachable_all(img, bits, sp)
    mp = getattr(O.lib(), f"or_gpu{bits}_max_probe")(np.ascontiguousarray(img).ctypes.data, sp)
    assert int(inf[2]) >= mp
    hm = Hashmap(sp, bits)
    hm.insert(keys, vals, targetLF=1.0)
    _, v, h = hm.retrieve(keys)
    assert h.cpu().numpy().all() and np.array_equal(to_numpy(v, bits), vals)


@pytest.mark.parametrize("bits", [32, 64])
@pytest.mark.parametrize("case", ["uniform", "overfull"])
def test_build_needs_no_initialised_table(gpu, bits, case):
    # Every slot is written (keys and EMPTY slots; the overfull fallback empties
    # the table itself), so bucket memory holding garbage, keys included, gives
    # the image an initialised table gives.
    sp = 14
    if case == "uniform":
        keys = keyset(int(0.9 * (1 << sp)), bits, seed=41)
    else:
        keys = concentrated_keys(bits, sp, 3000, 2048, seed=6)
    vals = (keys ^ NDT[bits](0x77)).astype(NDT[bits])
    ref_img, ref_inf, rc = run_build(bits, sp, keys, vals)
    assert rc == 0
    rng = np.random.default_rng(5)
    junk = rng.integers(0, 2**62, size=(1 << sp) * 2, dtype=np.uint64).astype(NDT[bits])
    junk[0:64:2] = keys[:32]  # live-looking keys too
    table = dev(junk, bits).reshape(-1, 2)
    img, inf, rc = run_build(bits, sp, keys, vals, table=table)
    assert rc == 0
    if case == "uniform":
        assert np.array_equal(img, ref_img)
        assert np.array_equal(inf, ref_inf)
    else:  # the probing fallback places in arrival order: same live set and occupancy
        assert np.array_equal(live_pairs(img, bits), live_pairs(ref_img, bits))
        assert np.array_equal(occupancy(img, bits), occupancy(ref_img, bits))
        assert int(inf[1]) == keys.size and (int(inf[4]) & 0xFFFFFFFF) == 0


def _tombstoned_map(bits, sp, lf, seed=0, ordered=False):
    """A probing insert at load factor lf (the bulk build, in Robin Hood
    order, when `ordered`), then a quarter of the keys erased; with the
    oracle's model of the image."""
    keys = keyset(int(lf * (1 << sp)), bits, seed=sp + bits + seed)
    vals = (keys * NDT[bits](9)).astype(NDT[bits])
    hm = Hashmap(sp, bits, auto_build=1.0 / 16 if ordered else 0, auto_reorder=0, build_min_keys=1)
    hm.insert(keys, vals, targetLF=1.0)
    assert hm.ordered() == ordered
    er = keys[np.random.default_rng(sp).choice(keys.size, keys.size // 4, replace=False)]
    hm.erase(er)
    i = hm.info()
    m = O.GpuModel(sp, bits, buckets=hm.host_buckets().reshape(-1).copy(),
                   info=O.Info(sp, 0, i.fill, i.current_max_bucket_overflow, i.tombstone_counter, i.err, 0))
    return hm, m


def _check_clean_exact(hm, m, bits):
    """Every key slot equals the oracle's clean followed by its reorder; values
    equal in live slots (an EMPTY slot's value is unspecified: the oracle keeps
    the erased value under the EMPTY key, the device writes 0); Info too."""
    m.clean_tombstones()
    assert m.reorder() == 1
    img, ref = hm.host_buckets(), m.b.reshape(-1, 2)
    assert np.array_equal(img[:, 0], ref[:, 0])
    live = ref[:, 0] != np.iinfo(NDT[bits]).max
    assert np.array_equal(img[live, 1], ref[live, 1])
    i = hm.info()
    assert (i.fill, i.tombstone_counter, i.current_max_bucket_overflow) == (m.info.fill, 0, m.info.overflow)
    assert i.err == 0 and (i.reserved1 & _lib.HN_INFO_ORDERED)


@pytest.mark.parametrize("ordered", [False, True])
@pytest.mark.parametrize("bits", [32, 64])
@pytest.mark.parametrize("sp,lf", [(10, 0.9), (14, 0.9), (16, 0.97), (18, 0.98), (20, 0.95), (21, 0.85), (22, 0.9),
                                   (12, 0.5)])
def test_clean_image_exact(gpu, bits, sp, lf, ordered):
    """clean_tombstones by EMPTY-bounded segments: the image is the oracle's
    clean followed by its reorder, bit for bit, from a table in arrival order
    (the sorting pass) and from one in Robin Hood order (the scan without a
    sort).  LF 0.97 and 0.98 have clusters longer than the 8,192 slots LDS
    holds (the pass over global memory; at 0.98 the default pool runs short
    and the map retries with a larger one)."""
    hm, m = _tombstoned_map(bits, sp, lf, ordered=ordered)
    assert hm.ordered() == ordered
    hm.clean_tombstones()
    _check_clean_exact(hm, m, bits)
    # no tombstones: a no-op, Info untouched
    before, info = hm.host_buckets().copy(), bytes(hm.info())
    hm.clean_tombstones()
    assert np.array_equal(hm.host_buckets(), before) and bytes(hm.info()) == info


def _skewed_keys(bits, sp, n, lo, width, seed):
    """n distinct keys whose home slots lie in [lo, lo + width): one cluster of about n slots."""
    out = []
    start = 0
    while sum(a.size for a in out) < n:
        cand = keyset(1 << 22, bits, seed=seed, start=start)
        start += (1 << 22) + 2
        h = O.fibhash(cand, sp, bits).astype(np.int64) & ((1 << sp) - 1)
        out.append(cand[(h >= lo) & (h < lo + width)])
    return np.concatenate(out)[:n]


@pytest.mark.parametrize("ordered", [False, True])
@pytest.mark.parametrize("bits", [32, 64])
def test_clean_pool_too_small_then_enough(gpu, bits, ordered):
    """A cluster of about 12,000 slots (keys with homes in a 64-slot window)
    is longer than LDS holds.  Through the C-ABI with a 64 KiB pool it does
    not fit: that segment is left as it was (same live set, every key
    reachable), its tombstones stay counted, err = FAIL, the table is not
    marked ordered.  A second call with scratch_bytes(sp, SIZE_MAX) finishes
    the image exactly as the oracle's clean + reorder."""
    sp = 14
    keys = _skewed_keys(bits, sp, 12000, 100, 64, seed=5)
    more = keyset(2000, bits, seed=6, start=10**7)  # ordinary keys elsewhere
    allk = np.concatenate([keys, more])
    vals = (allk * NDT[bits](3)).astype(NDT[bits])
    hm = Hashmap(sp, bits, auto_build=0, auto_reorder=0)
    hm.insert(allk, vals, targetLF=1.0)
    if ordered:  # a complete clean leaves Robin Hood order at any cluster length
        hm.erase(allk[:64])
        hm.clean_tombstones()
        assert hm.ordered()
    hm.erase(allk[np.random.default_rng(1).choice(allk.size, allk.size // 4, replace=False)])
    i = hm.info()
    m = O.GpuModel(sp, bits, buckets=hm.host_buckets().reshape(-1).copy(),
                   info=O.Info(sp, 0, i.fill, i.current_max_bucket_overflow, i.tombstone_counter, i.err, 0))
    L = _lib.load()
    sfx = "u32" if bits == 32 else "u64"
    q = getattr(L, f"hn_clean_tombstones_scratch_bytes_{sfx}")
    fixed = q(sp, 0) - (1 << 20)  # the default pool at this size is 1 MiB
    tiny = fixed + (64 << 10)
    assert q(sp, tiny) == tiny
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    clean = getattr(L, f"hn_clean_tombstones_{sfx}")
    live0 = live_pairs(hm.host_buckets(), bits)
    tomb0 = hm.tombstone_count()
    sc = torch.empty(tiny, dtype=torch.uint8, device="cuda")
    assert clean(C.c_void_p(hm.bucket_ptr()), C.c_void_p(hm.info_ptr()), sp, C.c_void_p(sc.data_ptr()), tiny, s) == 0
    torch.cuda.synchronize()
    i = hm.info()
    img = hm.host_buckets()
    assert np.array_equal(live_pairs(img, bits), live0) and reachable_all(img, bits, sp)
    assert i.err == _lib.HN_STATUS_FAIL and 0 < i.tombstone_counter < tomb0
    assert i.tombstone_counter == int(np.sum(img[:, 0] == np.iinfo(NDT[bits]).max - 1))
    assert not (i.reserved1 & _lib.HN_INFO_ORDERED)
    big = q(sp, (1 << 64) - 1)
    sc = torch.empty(big, dtype=torch.uint8, device="cuda")
    assert clean(C.c_void_p(hm.bucket_ptr()), C.c_void_p(hm.info_ptr()), sp, C.c_void_p(sc.data_ptr()), big, s) == 0
    torch.cuda.synchronize()
    _check_clean_exact(hm, m, bits)


@pytest.mark.parametrize("bits", [32, 64])
def test_clean_order_flag_without_order(gpu, bits):
    """A table in arrival order whose Info claims Robin Hood order: the scan
    pass checks every segment's order and hands the segments that break it
    to the sorting pass, so the result keeps the live set, puts every key on
    an unbroken probe path, occupies exactly the slots the oracle's clean
    occupies, and lists each cluster in home order."""
    sp, lf = 18, 0.9
    hm, m = _tombstoned_map(bits, sp, lf, seed=3)
    live0 = live_pairs(hm.host_buckets(), bits)
    hm._info.view(torch.int32)[9] = _lib.HN_INFO_ORDERED  # reserved1: a claim the table does not meet
    hm.clean_tombstones()
    m.clean_tombstones()
    assert m.reorder() == 1
    img, ref = hm.host_buckets(), m.b.reshape(-1, 2)
    assert np.array_equal(live_pairs(img, bits), live0) and reachable_all(img, bits, sp)
    assert np.array_equal(occupancy(img, bits), occupancy(ref, bits))
    i = hm.info()
    assert (i.fill, i.tombstone_counter, i.err) == (m.info.fill, 0, 0)
    # within every run of occupied slots the homes (relative to the run's start) do not decrease
    n = 1 << sp
    k = img[:, 0]
    occ = k != np.iinfo(NDT[bits]).max
    start = np.nonzero(~occ)[0][0]
    idx = (np.arange(n) + start) % n
    home = (O.fibhash(k[idx], sp, bits).astype(np.int64) - start) % n
    run = np.cumsum(~occ[idx])  # run id: EMPTY slots so far
    o = occ[idx]
    h, r = home[o], run[o]
    assert np.all((np.diff(h) >= 0) | (np.diff(r) != 0))


@pytest.mark.parametrize("bits", [32, 64])
def test_clean_full_table(gpu, bits):
    """Every slot occupied (live or TOMBSTONE): no segment boundary, the region
    clean over the whole table; then the same image as the oracle's."""
    sp = 12
    n = 1 << sp
    keys = keyset(n, bits, seed=77)
    hm = Hashmap(sp, bits, auto_build=0, auto_reorder=0)
    hm.insert(keys, keys, targetLF=1.0)
    hm.erase(keys[::3])
    i = hm.info()
    assert i.fill + i.tombstone_counter == n
    m = O.GpuModel(sp, bits, buckets=hm.host_buckets().reshape(-1).copy(),
                   info=O.Info(sp, 0, i.fill, i.current_max_bucket_overflow, i.tombstone_counter, i.err, 0))
    hm.auto_reorder = 1.0 / 16
    hm.clean_tombstones()
    _check_clean_exact(hm, m, bits)


@pytest.mark.parametrize("bits", [32, 64])
def test_map_insert_uses_build(gpu, bits):
    sp = 16
    keys = keyset(int(0.9 * (1 << sp)), bits, seed=9)
    vals = (keys + NDT[bits](5)).astype(NDT[bits])
    hm = Hashmap(sp, bits, build_min_keys=0)  # build even for this small a batch
    hm.insert(keys, vals, targetLF=1.0)
    m = expect(bits, sp, keys, vals)
    assert np.array_equal(hm.host_buckets().reshape(-1), m.b)
    assert hm.peek_status() == 0 and hm.size() == keys.size
    absent = keyset(keys.size, bits, seed=10, start=1 << 24)
    absent = absent[~np.isin(absent, keys)]
    q = np.concatenate([keys, absent])
    _, v, h = hm.retrieve(q)
    ev, eh = m.retrieve(q)
    assert np.array_equal(h.cpu().numpy().astype(bool), eh)
    assert np.array_equal(to_numpy(v, bits)[eh], ev[eh])


def test_build_invalid_args(gpu):
    L = _lib.load()
    assert L.hn_build_scratch_bytes_u64(10, 1 << 10) == 0  # len must be < 2^sp
    assert L.hn_build_scratch_bytes_u64(10, 1000) > 0
    assert L.hn_build_u64(None, None, None, None, 10, 10, None, 0, None, None) == _lib.HN_EINVAL
    b = torch.empty((1024, 2), dtype=torch.int64, device="cuda")
    inf = torch.zeros(5, dtype=torch.int64, device="cuda")
    k = torch.zeros(10, dtype=torch.int64, device="cuda")
    sb = L.hn_build_scratch_bytes_u64(10, 10)
    sc = torch.empty(sb, dtype=torch.uint8, device="cuda")
    assert L.hn_build_u64(k.data_ptr(), None, b.data_ptr(), inf.data_ptr(), 10, 10, sc.data_ptr(), sb - 1, None,
                          None) == _lib.HN_ESCRATCH
    assert L.hn_build_u64(k.data_ptr(), None, b.data_ptr(), inf.data_ptr(), 0, 10, sc.data_ptr(), sb, None,
                          None) == 0


def test_build_rejects_misaligned_scratch(gpu):
    """Scratch that does not start on a 256-byte boundary is refused before
    any launch (HN_EINVAL), for the build, the rehash build and the merge;
    the tables and Info stay as they were."""
    L = _lib.load()
    sp = 12
    N = 1 << sp
    keys = keyset(N // 2, 64, seed=5)
    b = torch.empty((N, 2), dtype=torch.int64, device="cuda")
    _lib.check(L.hn_table_init_u64(b.data_ptr(), N, None), "init")
    inf = torch.zeros(5, dtype=torch.int64, device="cuda")
    _lib.check(L.hn_info_init(inf.data_ptr(), sp, None), "info")
    dk = dev(keys, 64)
    sb = L.hn_build_scratch_bytes_u64(sp, keys.size)
    sc = torch.empty(sb + 256, dtype=torch.uint8, device="cuda")
    assert sc.data_ptr() % 256 == 0
    img0, inf0 = b.clone(), inf.clone()
    rc = L.hn_build_u64(dk.data_ptr(), dk.data_ptr(), b.data_ptr(), inf.data_ptr(), keys.size, sp, sc.data_ptr() + 16,
                        sb, None, None)
    torch.cuda.synchronize()
    assert rc == _lib.HN_EINVAL and torch.equal(b, img0) and torch.equal(inf, inf0)
    # the same table built with aligned scratch, then grown and merged into with misaligned scratch
    rc = L.hn_build_u64(dk.data_ptr(), dk.data_ptr(), b.data_ptr(), inf.data_ptr(), keys.size, sp, sc.data_ptr(), sb,
                        None, None)
    torch.cuda.synchronize()
    assert rc == 0
    img1, inf1 = b.clone(), inf.clone()
    nb = torch.empty((2 * N, 2), dtype=torch.int64, device="cuda")
    rsb = L.hn_rehash_build_scratch_bytes_u64(sp, sp + 1)
    rs = torch.empty(rsb + 256, dtype=torch.uint8, device="cuda")
    rc = L.hn_rehash_build_u64(b.data_ptr(), sp, nb.data_ptr(), sp + 1, inf.data_ptr(), rs.data_ptr() + 128, rsb,
                               None)
    torch.cuda.synchronize()
    assert rc == _lib.HN_EINVAL and torch.equal(inf, inf1)
    nk = dev(keyset(N // 8, 64, seed=6) | np.uint64(1 << 62), 64)
    fresh = torch.empty_like(b)
    msb = L.hn_merge_scratch_bytes_u64(sp, nk.numel())
    ms = torch.empty(msb + 256, dtype=torch.uint8, device="cuda")
    rc = L.hn_merge_u64(nk.data_ptr(), nk.data_ptr(), b.data_ptr(), fresh.data_ptr(), inf.data_ptr(), nk.numel(), sp,
                        ms.data_ptr() + 64, msb, None)
    torch.cuda.synchronize()
    assert rc == _lib.HN_EINVAL and torch.equal(b, img1) and torch.equal(inf, inf1)
