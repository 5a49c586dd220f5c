"""Insert into a non-empty table by merging (hn_merge_*): the new bucket array
equals the oracle's serial insert into the old table followed by its Robin
Hood reorder, key for key and value for value in live slots, with the Info
(fill, overflow, ordered flag, err).  Batches mix new keys, keys already in
the table (their value is overwritten) and keys repeated within the batch
(the last value stays).  Where the merge does not apply, the output is the
probing insert's, made on the device."""
import numpy as np
import pytest
import torch

import oracle as O
from hashinator_amd import to_numpy
from hashinator_amd import _lib
from tests.helpers import EMPTY, NDT, keyset, live_pairs, occupancy, reachable_all
from tests.test_gpu_build import TDT, dev, run_build

pytestmark = pytest.mark.gpu


def ordered_table(bits, sp, keys, vals):
    """A table built by the bulk build (Robin Hood order), on the device."""
    N = 1 << sp
    b = torch.empty((N, 2), dtype=TDT[bits], device="cuda")
    inf = torch.zeros(5, dtype=torch.int64, device="cuda")
    _lib.check(_lib.load().hn_info_init(inf.data_ptr(), sp, None), "info")
    img, info, rc = run_build(bits, sp, keys, vals, info=inf, table=b)
    assert rc == 0 and (int(info[4]) >> 32) == _lib.HN_INFO_ORDERED
    return b, inf


def batch(bits, old_keys, n_new, seed, overwrite=0.1, repeat=0.05):
    """n_new fresh keys (disjoint from old_keys), plus a share of old keys and
    of repeats of the batch's own keys, shuffled; values all distinct."""
    rng = np.random.default_rng(seed)
    fresh = keyset(old_keys.size + n_new, bits, seed=seed)
    fresh = np.setdiff1d(fresh, old_keys)[:n_new]
    ow = rng.choice(old_keys, int(overwrite * n_new), replace=False) if old_keys.size else old_keys[:0]
    base = np.concatenate([fresh, ow])
    rep = rng.choice(base, int(repeat * n_new), replace=True) if base.size else base[:0]
    keys = rng.permutation(np.concatenate([base, rep])).astype(NDT[bits])
    vals = (np.arange(keys.size, dtype=np.uint64) * 7 + 11).astype(NDT[bits])
    return keys, vals


def run_merge(bits, sp, old, inf, keys, vals, pairs=False, index=False):
    L = _lib.load()
    sfx = "u32" if bits == 32 else "u64"
    nb = torch.full_like(old, 0x5A5A)
    sb = getattr(L, f"hn_merge_scratch_bytes_{sfx}")(sp, keys.size)
    assert sb > 0
    sc = torch.empty(sb, dtype=torch.uint8, device="cuda")
    if pairs:
        dp = dev(np.stack([keys, vals], axis=1), bits)
        rc = getattr(L, f"hn_merge_pairs_{sfx}")(dp.data_ptr(), old.data_ptr(), nb.data_ptr(), inf.data_ptr(),
                                                 keys.size, sp, sc.data_ptr(), sb, None)
    else:
        dk = dev(keys, bits)
        dv = None if index else dev(vals, bits)
        rc = getattr(L, f"hn_merge_{sfx}")(dk.data_ptr(), None if dv is None else dv.data_ptr(), old.data_ptr(),
                                           nb.data_ptr(), inf.data_ptr(), keys.size, sp, sc.data_ptr(), sb, None)
    torch.cuda.synchronize()
    return rc, to_numpy(nb, bits).reshape(-1, 2), inf.cpu().numpy()


def expected(bits, sp, old_img, info, keys, vals):
    m = O.GpuModel(sp, bits, buckets=old_img.reshape(-1).copy(),
                   info=O.Info(sp, 0, int(info[1]), int(info[2]), 0, 0, 0))
    assert m.insert(keys, vals) == 0
    assert m.reorder() == 1
    return m


def check_merge(bits, img, inf, m):
    ref = m.b.reshape(-1, 2)
    assert np.array_equal(img[:, 0], ref[:, 0])
    live = ref[:, 0] != EMPTY[bits]
    assert np.array_equal(img[live, 1], ref[live, 1])
    assert np.all(img[~live, 1] == 0)
    assert (int(inf[1]), int(inf[2]), int(inf[3])) == (m.info.fill, m.info.overflow, 0)
    assert (int(inf[4]) & 0xFFFFFFFF) == 0 and (int(inf[4]) >> 32) == _lib.HN_INFO_ORDERED


@pytest.mark.parametrize("bits", [32, 64])
# sp 23: a batch of ~4.3 M keys, partitioned in the 8,192-pair tiles of inputs of 2^22 pairs and more.
@pytest.mark.parametrize("sp,old_lf,new_lf", [(4, 0.5, 0.3), (6, 0.6, 0.3), (8, 0.8, 0.15), (10, 0.5, 0.4),
                                              (11, 0.45, 0.45), (12, 0.5, 0.45), (13, 0.9, 0.05), (14, 0.3, 0.6),
                                              (16, 0.7, 0.2), (18, 0.5, 0.45), (20, 0.5, 0.45), (23, 0.4, 0.45)])
def test_merge_image_exact(gpu, bits, sp, old_lf, new_lf):
    N = 1 << sp
    old_keys = keyset(int(old_lf * N), bits, seed=3 * sp + bits)
    old_vals = (old_keys ^ NDT[bits](0x77)).astype(NDT[bits])
    old, inf = ordered_table(bits, sp, old_keys, old_vals)
    old_img = to_numpy(old, bits).reshape(-1, 2).copy()
    info0 = inf.cpu().numpy()
    keys, vals = batch(bits, old_keys, int(new_lf * N), seed=sp + 100 * bits)
    m = expected(bits, sp, old_img, info0, keys, vals)
    rc, img, inf1 = run_merge(bits, sp, old, inf, keys, vals)
    assert rc == 0
    check_merge(bits, img, inf1, m)
    assert np.array_equal(to_numpy(old, bits).reshape(-1, 2), old_img)  # the old table is only read


@pytest.mark.parametrize("bits", [32, 64])
def test_merge_wraps_past_the_end(gpu, bits):
    """Keys whose homes crowd the last slots: the union's last cluster runs
    past the table's end into region 0 (one region at sp 10, several at
    sp 14)."""
    for sp in (10, 14):
        N = 1 << sp
        cand = keyset(40 * N, bits, seed=91 + sp)
        home = O.fibhash(cand, sp, bits).astype(np.int64) & (N - 1)
        tail = cand[home >= N - 48]
        rest = cand[home < N - 48]
        old_keys = np.concatenate([tail[:30], rest[: N // 3]])
        old, inf = ordered_table(bits, sp, old_keys, old_keys)
        old_img = to_numpy(old, bits).reshape(-1, 2).copy()
        keys = np.concatenate([tail[30:60], rest[N // 3: N // 3 + N // 4]]).astype(NDT[bits])
        vals = (keys + NDT[bits](5)).astype(NDT[bits])
        m = expected(bits, sp, old_img, inf.cpu().numpy(), keys, vals)
        assert m.b.reshape(-1, 2)[0, 0] != EMPTY[bits]  # the wrap reaches slot 0
        rc, img, inf1 = run_merge(bits, sp, old, inf, keys, vals)
        assert rc == 0
        check_merge(bits, img, inf1, m)


@pytest.mark.parametrize("bits", [32, 64])
def test_merge_pairs_and_index(gpu, bits):
    sp = 15
    N = 1 << sp
    old_keys = keyset(N // 2, bits, seed=5)
    old, inf = ordered_table(bits, sp, old_keys, old_keys)
    old_img = to_numpy(old, bits).reshape(-1, 2).copy()
    info0 = inf.cpu().numpy().copy()
    keys, vals = batch(bits, old_keys, N // 3, seed=6)
    m = expected(bits, sp, old_img, info0, keys, vals)
    rc, img, inf1 = run_merge(bits, sp, old, inf, keys, vals, pairs=True)
    assert rc == 0
    check_merge(bits, img, inf1, m)
    # values = batch index
    inf2 = torch.from_numpy(info0).cuda()
    idx = np.arange(keys.size, dtype=NDT[bits])
    m2 = expected(bits, sp, old_img, info0, keys, idx)
    rc, img2, inf3 = run_merge(bits, sp, old, inf2, keys, None, index=True)
    assert rc == 0
    check_merge(bits, img2, inf3, m2)


@pytest.mark.parametrize("bits", [32, 64])
def test_merge_falls_back_to_probing(gpu, bits):
    """Where the merge does not apply -- a table with TOMBSTONEs, one without
    the order flag, a batch whose keys crowd one region past its 1,536-pair
    capacity -- the output is the probing insert's: the old table untouched,
    the new one holding the oracle insert's live set and occupancy, every key
    on an unbroken probe path, fill and tombstones as the oracle's, err =
    SUCCESS and no order flag."""
    sp = 14
    N = 1 << sp
    old_keys = keyset(N // 2, bits, seed=8)
    keys, vals = batch(bits, old_keys, N // 8, seed=9)
    L = _lib.load()
    sfx = "u32" if bits == 32 else "u64"

    def fell_back(old, inf, k, v):
        # which copy of a repeated key the parallel probing insert keeps is
        # not fixed (as in the reference's kernel): values that follow the key
        v = (k ^ NDT[bits](0x5A5A)).astype(NDT[bits])
        before_img = to_numpy(old, bits).reshape(-1, 2).copy()
        before = inf.cpu().numpy().copy()
        m = O.GpuModel(sp, bits, buckets=before_img.reshape(-1).copy(),
                       info=O.Info(sp, 0, int(before[1]), int(before[2]), int(before[3]), 0, 0))
        assert m.insert(k, v) == 0
        rc, img, after = run_merge(bits, sp, old, inf, k, v)
        assert rc == 0
        assert np.array_equal(to_numpy(old, bits).reshape(-1, 2), before_img)
        assert np.array_equal(live_pairs(img, bits), live_pairs(m.b, bits))
        assert np.array_equal(occupancy(img, bits), occupancy(m.b, bits)) and reachable_all(img, bits, sp)
        assert (int(after[1]), int(after[3])) == (m.info.fill, m.info.tombstones)
        assert (int(after[4]) & 0xFFFFFFFF) == 0 and (int(after[4]) >> 32) == 0

    old, inf = ordered_table(bits, sp, old_keys, old_keys)
    dk = dev(old_keys[:100], bits)
    _lib.check(getattr(L, f"hn_erase_{sfx}")(dk.data_ptr(), old.data_ptr(), inf.data_ptr(), 100, None, None), "erase")
    fell_back(old, inf, keys, vals)

    old, inf = ordered_table(bits, sp, old_keys, old_keys)
    inf.view(torch.int32)[9] = 0  # no order flag
    fell_back(old, inf, keys, vals)

    old, inf = ordered_table(bits, sp, old_keys, old_keys)
    cand = keyset(4000 * 64, bits, seed=10, start=10**7)
    home = O.fibhash(cand, sp, bits).astype(np.int64) & (N - 1)
    crowd = np.setdiff1d(cand[(home >= 4096) & (home < 4096 + 2048)], old_keys)[:1700]
    assert crowd.size == 1700
    fell_back(old, inf, crowd, crowd)


@pytest.mark.parametrize("bits", [32, 64])
def test_merge_large_then_lookups(gpu, bits):
    """2^22 slots: a quarter of the table merged into a half-full one, 20% of
    the batch overwriting keys already present and 10% repeats: the live
    (key, value) set is the dict of old pairs updated by the batch in order."""
    sp = 22
    N = 1 << sp
    old_keys = keyset(N // 2, bits, seed=12)
    old, inf = ordered_table(bits, sp, old_keys, old_keys)
    keys, vals = batch(bits, old_keys, N // 4, seed=13, overwrite=0.2, repeat=0.1)
    rc, img, inf1 = run_merge(bits, sp, old, inf, keys, vals)
    assert rc == 0 and (int(inf1[4]) & 0xFFFFFFFF) == 0
    final = dict(zip(old_keys.tolist(), old_keys.tolist()))
    final.update(zip(keys.tolist(), vals.tolist()))
    rows = live_pairs(img, bits)
    assert rows.shape[0] == len(final) == int(inf1[1])
    exp = np.array(sorted(final.items()), dtype=NDT[bits])
    assert np.array_equal(rows, exp)


def _final(old_keys, old_vals, keys, vals):
    d = dict(zip(old_keys.tolist(), old_vals.tolist()))
    d.update(zip(keys.tolist(), vals.tolist()))
    return d


def _check_map(hm, bits, final):
    rows = live_pairs(hm.host_buckets(), bits)
    exp = np.array(sorted(final.items()), dtype=NDT[bits]).reshape(-1, 2)
    assert np.array_equal(rows, exp)
    i = hm.info()
    assert (i.fill, i.tombstone_counter, i.err) == (len(final), 0, 0)
    q = exp[:, 0]
    got = torch.zeros(q.size, dtype=TDT[bits], device="cuda")
    hit = torch.zeros(q.size, dtype=torch.uint8, device="cuda")
    hm.retrieve(q, got, hit)
    assert bool(hit.bool().all()) and np.array_equal(to_numpy(got, bits), exp[:, 1])


@pytest.mark.parametrize("bits", [32, 64])
def test_map_insert_merges(gpu, bits):
    """Hashmap.insert of a quarter of the table into a half-full ordered map
    takes the merge: the map ends ordered, with the dict-update live set, and
    retrieves every value; a map in arrival order is reordered first when
    the batch is a third of the table."""
    from hashinator_amd import Hashmap
    sp = 16
    N = 1 << sp
    old_keys = keyset(N // 2, bits, seed=21)
    keys, vals = batch(bits, old_keys, N // 4, seed=22)
    hm = Hashmap(sp, bits, build_min_keys=1)
    hm.insert(old_keys, old_keys, targetLF=1.0)
    assert hm.ordered()
    b0 = hm.buckets
    hm.insert(keys, vals, targetLF=1.0)
    assert hm.buckets is not b0 and hm.ordered()  # a fresh array: the merge ran
    _check_map(hm, bits, _final(old_keys, old_keys, keys, vals))

    hm = Hashmap(sp, bits, auto_build=0, auto_reorder=0, build_min_keys=1)
    hm.insert(old_keys, old_keys, targetLF=1.0)
    assert not hm.ordered()
    keys2, vals2 = batch(bits, old_keys, N // 3, seed=23)
    hm.insert(keys2, vals2, targetLF=1.0)
    assert hm.ordered()
    _check_map(hm, bits, _final(old_keys, old_keys, keys2, vals2))


@pytest.mark.parametrize("bits", [32, 64])
def test_map_grow_then_merge(gpu, bits):
    """An insert that grows the map: the bulk rehash (ordered) and then the
    merge into the larger table (repeated keys keep their last value, which
    the probing insert would not promise)."""
    from hashinator_amd import Hashmap
    sp = 12
    N = 1 << sp
    old_keys = keyset(N // 2, bits, seed=31)
    hm = Hashmap(sp, bits, build_min_keys=1)
    hm.insert(old_keys, old_keys, targetLF=1.0)
    keys, vals = batch(bits, old_keys, 3 * N // 4, seed=32)
    hm.insert(keys, vals, targetLF=0.9)  # grows to 2^13: the batch is over a quarter of it
    assert hm.getSizePower() == sp + 1 and hm.ordered()
    _check_map(hm, bits, _final(old_keys, old_keys, keys, vals))
