"""The 2^28-slot u64 workload at its full size (241,591,910 keys at load
factor 0.9, half uniform, half runs of consecutive ids, plus as many absent
keys), checked through properties that hold at any size: every present key
retrieves its own value, no absent key touches its output, erase/clean keep
exactly the survivors, compaction returns the valid slots in bucket order,
and a scan of ones is the identity ramp.  The bit-exact image comparisons
against the oracle run at smaller sizes in the other test files."""
import numpy as np
import pytest
import torch

from hashinator_amd import Hashmap, tools
from hashinator_amd import _lib
from hashinator_amd import workloads as W
from tests.helpers import live_pairs

pytestmark = pytest.mark.gpu

SP = 28
MISS = -0x2152411021524111  # an output value no retrieve writes here
CHUNK = 1 << 24  # boolean-mask selections in chunks of this many rows


@pytest.fixture(scope="module")
def big(gpu):
    n = int(0.9 * (1 << SP))
    seed = W.SEED_BASE + 2
    keys = W.mixed64(n, seed)
    vals = W.values(keys, seed)
    absent = W.absent64(n, seed, n)
    dk = torch.from_numpy(keys.view(np.int64)).cuda()
    dv = torch.from_numpy(vals.view(np.int64)).cuda()
    da = torch.from_numpy(absent.view(np.int64)).cuda()
    del keys, vals, absent
    g = torch.Generator(device="cuda").manual_seed(1234)
    perm = torch.randperm(2 * n, device="cuda", generator=g)
    yield n, dk, dv, da, perm
    del dk, dv, da, perm
    torch.cuda.empty_cache()


def built(n, dk, dv):
    hm = Hashmap(SP, 64)
    hm.insert(dk, dv, targetLF=1.0)  # an insert this large into an empty map takes the bulk build
    i = hm.info()
    assert (i.fill, i.tombstone_counter, i.err) == (n, 0, 0)
    assert hm.ordered()
    return hm


def check_retrieve(hm, q, expect_vals, expect_hit):
    out = torch.full_like(q, MISS)
    hit = torch.empty(q.numel(), dtype=torch.uint8, device="cuda")
    hm.retrieve(q, out, hit)
    assert torch.equal(hit.bool(), expect_hit)
    assert torch.equal(out, expect_vals)


def test_fullsize_build_then_retrieve(big):
    n, dk, dv, da, perm = big
    hm = built(n, dk, dv)
    q = torch.cat([dk, da])[perm]
    expect = torch.cat([dv, torch.full((n,), MISS, dtype=torch.int64, device="cuda")])[perm]
    check_retrieve(hm, q, expect, perm < n)


def test_fullsize_erase_clean_then_retrieve(big):
    n, dk, dv, da, perm = big
    hm = built(n, dk, dv)
    g = torch.Generator(device="cuda").manual_seed(99)
    idx = torch.randperm(n, device="cuda", generator=g)[: n // 4]
    gone = torch.zeros(n, dtype=torch.bool, device="cuda")
    gone[idx] = True
    hm.erase(dk[idx])
    i = hm.info()
    assert (i.fill, i.tombstone_counter) == (n - n // 4, n // 4)
    expect = torch.where(gone, torch.full_like(dv, MISS), dv)
    check_retrieve(hm, dk, expect, ~gone)
    hm.clean_tombstones()
    i = hm.info()
    assert (i.fill, i.tombstone_counter, i.err) == (n - n // 4, 0, 0)
    assert hm.ordered()
    check_retrieve(hm, dk, expect, ~gone)


def test_fullsize_extract_pattern_and_scan(big):
    n, dk, dv, da, perm = big
    hm = built(n, dk, dv)
    got = hm.extractPattern()
    assert got.shape == (n, 2)
    # stable: the valid slots in bucket order
    b = hm.buckets
    pos = 0
    for c0 in range(0, b.shape[0], CHUNK):
        blk = b[c0:c0 + CHUNK]
        sel = blk[(blk[:, 0] != -1) & (blk[:, 0] != -2)]
        assert torch.equal(got[pos:pos + sel.shape[0]], sel)
        pos += sel.shape[0]
    assert pos == n
    # the same pairs as inserted
    ks, order = torch.sort(got[:, 0])
    ik = torch.argsort(dk)
    assert torch.equal(ks, dk[ik])
    assert torch.equal(got[:, 1][order], dv[ik])
    del got, ks, order, ik
    ones = torch.ones(1 << SP, dtype=torch.int32, device="cuda")
    ramp = tools.split_prefix_scan(ones)
    assert torch.equal(ramp, torch.arange(1 << SP, dtype=torch.int32, device="cuda"))


@pytest.mark.parametrize("lf,build", [(0.5, True), (0.75, True), (0.5, False), (0.75, False)])
def test_fullsize_other_load_factors(gpu, lf, build):
    """2^28 slots at load factors 0.5 and 0.75, through the bulk build and
    through the probing insert (+ reorder): every present key retrieves its
    value, no absent key touches its output, Info is exact."""
    n = int(lf * (1 << SP))
    seed = W.SEED_BASE + 2
    keys = W.mixed64(n, seed)
    dk = torch.from_numpy(keys.view(np.int64)).cuda()
    dv = torch.from_numpy(W.values(keys, seed).view(np.int64)).cuda()
    da = torch.from_numpy(W.absent64(n, seed, n).view(np.int64)).cuda()
    del keys
    hm = Hashmap(SP, 64, auto_build=1.0 / 16 if build else 0)
    hm.insert(dk, dv, targetLF=1.0)
    i = hm.info()
    assert (hm.getSizePower(), i.fill, i.tombstone_counter, i.err) == (SP, n, 0, 0)
    assert hm.ordered()  # the build writes Robin Hood order; the probing path reorders after a large insert
    g = torch.Generator(device="cuda").manual_seed(int(lf * 100))
    perm = torch.randperm(2 * n, device="cuda", generator=g)
    q = torch.cat([dk, da])[perm]
    expect = torch.cat([dv, torch.full((n,), MISS, dtype=torch.int64, device="cuda")])[perm]
    check_retrieve(hm, q, expect, perm < n)
    del hm, q, expect, perm, dk, dv, da
    torch.cuda.empty_cache()


def test_fullsize_u32_probing_insert(gpu):
    """2^26 u32 slots at load factor 0.9 through the probing insert, whose
    32-bit claim stores key and value with one 64-bit CAS of the slot: every
    key retrieves its value, and the table holds exactly the inserted pairs
    (checked on the host)."""
    sp = 26
    n = int(0.9 * (1 << sp))
    keys = W.unique_keys(n, 32, 11)
    dk = torch.from_numpy(keys.view(np.int32)).cuda()
    dv = dk * 3 + 1
    hm = Hashmap(sp, 32, auto_build=0, auto_reorder=0)
    hm.insert(dk, dv, targetLF=1.0)
    i = hm.info()
    assert (hm.getSizePower(), i.fill, i.tombstone_counter, i.err) == (sp, n, 0, 0)
    out = torch.full_like(dk, 12345)
    hit = torch.empty(n, dtype=torch.uint8, device="cuda")
    hm.retrieve(dk, out, hit)
    assert bool(hit.bool().all()) and torch.equal(out, dv)
    rows = live_pairs(hm.host_buckets(), 32)
    order = np.argsort(keys)
    exp = np.stack([keys[order], dv.cpu().numpy().view(np.uint32)[order]], axis=1)
    assert np.array_equal(rows, exp)
    del hm, dk, dv, out, hit
    torch.cuda.empty_cache()


def test_fullsize_merge_equals_probing_insert_then_reorder(big):
    """The merge at its full size: 120 M keys (5% of them keys already in the
    map, with new values; no key twice, so the probing insert's result does
    not depend on which of two copies wins) into the 2^28-slot table holding
    2^27 keys in Robin Hood order.  Its table equals, slot for slot, the
    probing insert of the same batch into a copy of the table followed by the
    reorder -- the path the image tests pin to the oracle at smaller sizes --
    with the same fill and overflow; every key of the union retrieves its
    last value; the old table is only read."""
    n, dk, dv, da, perm = big
    L = _lib.load()
    N = 1 << SP
    n_old, n_new = N // 2, 120_000_000
    n_ow = n_new // 20
    ok_, ov_ = dk[:n_old], dv[:n_old]
    rng = np.random.default_rng(2028)
    ow = rng.choice(n_old, n_ow, replace=False)  # indices of present keys the batch overwrites
    hk = np.concatenate([da[: n_new - n_ow].cpu().numpy(), ok_.cpu().numpy()[ow]])
    order = rng.permutation(n_new)
    hk = hk[order]
    hv = (hk ^ np.int64(0x5BD1E9955BD1E995)) + 1  # new values, not the keys' old ones
    nk = torch.from_numpy(hk).cuda()
    nv = torch.from_numpy(hv).cuda()

    old = torch.empty((N, 2), dtype=torch.int64, device="cuda")
    info0 = torch.zeros(5, dtype=torch.int64, device="cuda")
    _lib.check(L.hn_info_init(info0.data_ptr(), SP, None), "info_init")
    bsb = L.hn_build_scratch_bytes_u64(SP, n_old)
    bs = torch.empty(bsb, dtype=torch.uint8, device="cuda")
    _lib.check(L.hn_build_u64(ok_.data_ptr(), ov_.data_ptr(), old.data_ptr(), info0.data_ptr(), n_old, SP,
                              bs.data_ptr(), bsb, None, None), "build")
    torch.cuda.synchronize()
    del bs
    assert (int(info0[4].item()) >> 32) == _lib.HN_INFO_ORDERED
    old_copy = old.clone()

    # the merge, into a second array
    nb = torch.empty_like(old)
    inf_m = info0.clone()
    msb = L.hn_merge_scratch_bytes_u64(SP, n_new)
    ms = torch.empty(msb, dtype=torch.uint8, device="cuda")
    _lib.check(L.hn_merge_u64(nk.data_ptr(), nv.data_ptr(), old.data_ptr(), nb.data_ptr(), inf_m.data_ptr(), n_new,
                              SP, ms.data_ptr(), msb, None), "merge")
    torch.cuda.synchronize()
    del ms
    im = inf_m.cpu().numpy()
    assert (int(im[1]), int(im[3])) == (n_old + n_new - n_ow, 0)
    assert (int(im[4]) & 0xFFFFFFFF) == 0 and (int(im[4]) >> 32) == _lib.HN_INFO_ORDERED

    # the probing insert of the same batch into a copy, then the reorder
    b2 = old.clone()
    inf_p = info0.clone()
    _lib.check(L.hn_insert_u64(nk.data_ptr(), nv.data_ptr(), b2.data_ptr(), inf_p.data_ptr(), n_new, None, None),
               "insert")
    rsb = L.hn_reorder_scratch_bytes(SP)
    rs = torch.empty(max(rsb, 1), dtype=torch.uint8, device="cuda")
    _lib.check(L.hn_reorder_u64(b2.data_ptr(), inf_p.data_ptr(), SP, rs.data_ptr(), rsb, None), "reorder")
    torch.cuda.synchronize()
    del rs
    ip = inf_p.cpu().numpy()
    assert (int(ip[1]), int(ip[2]), int(ip[3])) == (int(im[1]), int(im[2]), int(im[3]))

    # slot for slot: keys everywhere, values where a key lives
    for c0 in range(0, N, CHUNK):
        a, b = nb[c0:c0 + CHUNK], b2[c0:c0 + CHUNK]
        assert torch.equal(a[:, 0], b[:, 0]), c0
        live = a[:, 0] != -1
        assert torch.equal(torch.where(live, a[:, 1], 0), torch.where(live, b[:, 1], 0)), c0
        assert torch.equal(old[c0:c0 + CHUNK], old_copy[c0:c0 + CHUNK]), c0
    del b2, old_copy

    # every key of the union retrieves its last value
    exp_old = ov_.cpu().numpy().copy()
    exp_old[ow] = hv[np.argsort(order)][n_new - n_ow:]
    for q, e in ((ok_, torch.from_numpy(exp_old).cuda()), (nk, nv)):
        out = torch.full_like(q, MISS)
        hit = torch.empty(q.numel(), dtype=torch.uint8, device="cuda")
        _lib.check(L.hn_retrieve_u64(q.data_ptr(), out.data_ptr(), hit.data_ptr(), nb.data_ptr(), inf_m.data_ptr(),
                                     q.numel(), None, None), "retrieve")
        torch.cuda.synchronize()
        assert bool(hit.bool().all())
        assert torch.equal(out, e)
    del nb, old, nk, nv
    torch.cuda.empty_cache()


def test_fullsize_growth_by_rehash_build(gpu):
    """Growth at full size: a 2^27-slot map at load factor 0.85 with a fifth
    of its keys erased (tombstones in place) grown to 2^28 slots by the bulk
    rehash (hn_rehash_build_u64, the old table as the partition's source,
    live slots only).  The new table equals, slot for slot, the bulk build of
    the old map's live pairs into an empty 2^28 table (the Robin Hood image
    is a function of the key set), with the same fill and overflow."""
    L = _lib.load()
    old_sp, new_sp = SP - 1, SP
    n = int(0.85 * (1 << old_sp))
    seed = W.SEED_BASE + 5
    keys = W.mixed64(n, seed)
    dk = torch.from_numpy(keys.view(np.int64)).cuda()
    dv = torch.from_numpy(W.values(keys, seed).view(np.int64)).cuda()
    er = torch.from_numpy(keys[np.random.default_rng(5).choice(n, n // 5, replace=False)].view(np.int64)).cuda()
    del keys
    hm = Hashmap(old_sp, 64)
    hm.insert(dk, dv, targetLF=1.0)
    hm.erase(er)
    i = hm.info()
    assert (i.fill, i.tombstone_counter) == (n - n // 5, n // 5)
    live = hm.extractPattern()  # the live pairs, in bucket order
    assert live.shape[0] == n - n // 5
    del dk, dv, er

    nb = torch.empty((1 << new_sp, 2), dtype=torch.int64, device="cuda")
    sb = L.hn_rehash_build_scratch_bytes_u64(old_sp, new_sp)
    sc = torch.empty(sb, dtype=torch.uint8, device="cuda")
    rc = L.hn_rehash_build_u64(hm.bucket_ptr(), old_sp, nb.data_ptr(), new_sp, hm.info_ptr(), sc.data_ptr(), sb, None)
    torch.cuda.synchronize()
    assert rc == 0
    inf = hm._info.cpu().numpy()
    del sc, hm

    ref = Hashmap(new_sp, 64)
    ref.insert(live[:, 0].contiguous(), live[:, 1].contiguous(), targetLF=1.0)  # the bulk build
    r = ref.info()
    assert ref.ordered() and (r.fill, r.tombstone_counter, r.err) == (n - n // 5, 0, 0)
    assert (int(inf[0]), int(inf[1]), int(inf[2]), int(inf[3])) == (new_sp, r.fill, r.current_max_bucket_overflow, 0)
    assert (int(inf[4]) & 0xFFFFFFFF) == 0 and (int(inf[4]) >> 32) == _lib.HN_INFO_ORDERED
    b = ref.buckets
    for c0 in range(0, 1 << new_sp, CHUNK):
        x, y = nb[c0:c0 + CHUNK], b[c0:c0 + CHUNK]
        assert torch.equal(x[:, 0], y[:, 0]), c0
        lv = x[:, 0] != -1
        assert torch.equal(torch.where(lv, x[:, 1], 0), torch.where(lv, y[:, 1], 0)), c0
    del nb, ref, live
    torch.cuda.empty_cache()


def test_build_with_nine_bit_second_digit_u32(gpu):
    """2^29 u32 slots: 18 region bits, two partition passes of 9-bit digits,
    so the second pass counts its digits from the pairs (a digit that fits a
    byte comes from the byte array the first pass writes; 2^28 slots and less
    take that path).  The bulk build of 0.6 of the slots equals the probing
    insert followed by the reorder slot for slot, with the same Info, and
    every key retrieves its value."""
    sp = 29
    n = int(0.6 * (1 << sp))
    k = (np.arange(n, dtype=np.uint64) * np.uint64(2654435761)) & np.uint64(0xFFFFFFFF)  # distinct: odd multiplier
    k = k[k < np.uint64(0xFFFFFFFE)].astype(np.uint32)  # no sentinel keys
    n = k.size
    dk = torch.from_numpy(k.view(np.int32)).cuda()
    dv = torch.from_numpy((k ^ np.uint32(0x5BD1E995)).view(np.int32)).cuda()
    del k
    hb = Hashmap(sp, 32)
    hb.insert(dk, dv, targetLF=1.0)  # the bulk build
    hp = Hashmap(sp, 32, auto_build=0)
    hp.insert(dk, dv, targetLF=1.0)  # the probing insert, then the reorder
    ib, ip = hb.info(), hp.info()
    assert hb.ordered() and hp.ordered()
    assert (ib.fill, ib.tombstone_counter, ib.err) == (n, 0, 0)
    assert (ip.fill, ip.current_max_bucket_overflow, ip.tombstone_counter) == (
        ib.fill, ib.current_max_bucket_overflow, ib.tombstone_counter)
    for c0 in range(0, 1 << sp, CHUNK):
        x, y = hb.buckets[c0:c0 + CHUNK], hp.buckets[c0:c0 + CHUNK]
        assert torch.equal(x[:, 0], y[:, 0]), c0
        lv = x[:, 0] != -1
        assert torch.equal(torch.where(lv, x[:, 1], 0), torch.where(lv, y[:, 1], 0)), c0
    del hp
    out = torch.zeros_like(dk)
    hit = torch.empty(n, dtype=torch.uint8, device="cuda")
    hb.retrieve(dk, out, hit)
    assert bool(hit.bool().all()) and torch.equal(out, dv)
    del hb, dk, dv, out, hit
    torch.cuda.empty_cache()
