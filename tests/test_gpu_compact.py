"""Stable compaction and exclusive scan on the GPU vs the oracle, bit-exact."""
import numpy as np
import pytest
import torch

import oracle as O
from hashinator_amd import Hashmap, hn_pred, to_numpy, tools
from hashinator_amd import _lib
from tests.helpers import NDT, keyset

pytestmark = pytest.mark.gpu

# around one compaction tile (8192 elements) and a few tiles, plus sizes of many tiles
SIZES = [0, 1, 63, 2047, 2048, 2049, 8191, 8192, 8193, 24_577, 100_003, 1 << 20]


def _t(a, bits):
    return torch.from_numpy(np.ascontiguousarray(a).view(np.int32 if bits == 32 else np.int64)).cuda()


def opred(p, bits):
    return O.make_pred(p.kind, p.field, p.mask, p.value, width=bits)


@pytest.mark.parametrize("bits", [32, 64])
@pytest.mark.parametrize("n", SIZES)
def test_copy_if_pairs(gpu, bits, n):
    rng = np.random.default_rng(n + bits)
    pairs = rng.integers(0, np.iinfo(NDT[bits]).max, size=(n, 2), dtype=NDT[bits], endpoint=True)
    for p in (hn_pred(_lib.HN_PRED_MASKEQ, 1, 1, 1), hn_pred(_lib.HN_PRED_LT, 0, 0, 1 << (bits - 2)),
              hn_pred(_lib.HN_PRED_ALL, 0, 0, 0), hn_pred(_lib.HN_PRED_GE, 1, 0, (1 << bits) - 1)):
        out = tools.copy_if_pairs(_t(pairs, bits).reshape(-1, 2), p, bits)
        exp = O.copy_if_pairs(pairs.reshape(-1), opred(p, bits), bits).reshape(-1, 2)
        assert np.array_equal(to_numpy(out, bits).reshape(-1, 2), exp)
        keys = tools.copy_keys_if_pairs(_t(pairs, bits).reshape(-1, 2), p, bits)
        assert np.array_equal(to_numpy(keys, bits), O.copy_keys_if_pairs(pairs.reshape(-1), opred(p, bits), bits))


@pytest.mark.parametrize("eb", [4, 8, 16])
@pytest.mark.parametrize("n", [5, 4096, 333_333])
def test_copy_if_flags_clustered(gpu, eb, n):
    rng = np.random.default_rng(eb * n)
    # alternating keep/drop runs, geometric lengths of mean 64
    lens = rng.geometric(1 / 64, size=n // 8 + 8)
    flags = np.repeat(np.arange(lens.size) % 2, lens)[:n].astype(np.uint8)
    if flags.size < n:
        flags = np.concatenate([flags, np.zeros(n - flags.size, np.uint8)])
    if eb == 16:
        elems = rng.integers(0, 2**63, size=(n, 2), dtype=np.uint64)
        t = torch.from_numpy(elems.view(np.int64)).cuda()
    else:
        dt = np.uint32 if eb == 4 else np.uint64
        elems = rng.integers(0, np.iinfo(dt).max, size=n, dtype=dt)
        t = torch.from_numpy(elems.view(np.int32 if eb == 4 else np.int64)).cuda()
    out = tools.copy_if_flags(t, torch.from_numpy(flags).cuda())
    exp = elems[flags.astype(bool)]
    assert np.array_equal(out.cpu().numpy().view(elems.dtype), exp)


@pytest.mark.parametrize("bits", [32, 64])
# around one scan tile (65536 u32, 32768 u64 elements), two tiles, many tiles
@pytest.mark.parametrize("n", [1, 15, 16, 4095, 4096, 4097, 32_767, 32_768, 32_769, 65_535, 65_536, 65_537,
                               131_073, 1_000_003, 1 << 22])
def test_exclusive_scan(gpu, bits, n):
    rng = np.random.default_rng(n * bits)
    x = rng.integers(0, np.iinfo(NDT[bits]).max, size=n, dtype=NDT[bits], endpoint=True)  # wraps
    out = tools.split_prefix_scan(_t(x, bits))
    assert np.array_equal(to_numpy(out, bits), O.exclusive_scan(x))
    # in place
    t = _t(x, bits)
    tools.split_prefix_scan(t, t)
    assert np.array_equal(to_numpy(t, bits), O.exclusive_scan(x))


@pytest.mark.parametrize("bits", [32, 64])
# pointers off 16-byte alignment take element loads and tiles of 4 rows
# (16,384 u32 or 8,192 u64 elements): sizes around one and two such tiles
@pytest.mark.parametrize("n", [1, 1000, 8191, 8192, 8193, 16_383, 16_384, 16_385, 32_769, 1_000_003])
def test_exclusive_scan_unaligned(gpu, bits, n):
    rng = np.random.default_rng(n * bits + 7)
    x = rng.integers(0, np.iinfo(NDT[bits]).max, size=n, dtype=NDT[bits], endpoint=True)
    src = torch.empty(n + 1, dtype=torch.int32 if bits == 32 else torch.int64, device="cuda")
    src[1:].copy_(_t(x, bits))
    dst = torch.zeros(n + 3, dtype=src.dtype, device="cuda")
    inp, out = src[1:], dst[3:]  # 4 / 12 (u32) or 8 / 24 (u64) bytes off the allocations' alignment
    assert inp.data_ptr() % 16 and out.data_ptr() % 16
    tools.split_prefix_scan(inp, out)
    assert np.array_equal(to_numpy(out, bits), O.exclusive_scan(x))
    assert int(dst[:3].abs().sum()) == 0  # nothing written before the output


def test_scan_flags_closed_form(gpu):
    n = 1 << 24
    flags = np.ones(n, dtype=np.uint32)
    out = tools.split_prefix_scan(torch.from_numpy(flags.view(np.int32)).cuda())
    assert np.array_equal(to_numpy(out, 32), np.arange(n, dtype=np.uint32))


@pytest.mark.parametrize("bits", [32, 64])
def test_extract_pattern_on_map_snapshot(gpu, bits):
    hm = Hashmap(14, bits)
    keys = keyset(12000, bits, seed=5)
    hm.insert(keys, keys ^ NDT[bits](1), targetLF=1.0)
    hm.erase(keys[::5])
    snap = hm.host_buckets().reshape(-1).copy()
    got = to_numpy(hm.extractPattern(), bits).reshape(-1, 2)
    exp = O.copy_if_pairs(snap, O.make_pred(O.PRED_VALID, width=bits), bits).reshape(-1, 2)
    assert np.array_equal(got, exp) and got.shape[0] == hm.size()
    ks = to_numpy(hm.extractAllKeys(), bits)
    assert np.array_equal(ks, exp[:, 0])
    lt = hn_pred(_lib.HN_PRED_LT, 0, 0, 1 << (bits - 1))
    got = to_numpy(hm.extractKeysByPattern(lt), bits)
    assert np.array_equal(got, O.copy_keys_if_pairs(snap, O.make_pred(O.PRED_LT, 0, 0, 1 << (bits - 1), bits), bits))
    # the *Loop variants return the same elements in the same order
    assert np.array_equal(to_numpy(hm.extractLoop(), bits).reshape(-1, 2), exp)
    assert np.array_equal(to_numpy(hm.extractPatternLoop(), bits).reshape(-1, 2), exp)
    assert np.array_equal(to_numpy(hm.extractAllKeysLoop(), bits), exp[:, 0])
    assert np.array_equal(to_numpy(hm.extractKeysByPatternLoop(lt), bits), got)
