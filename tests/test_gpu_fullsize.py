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
