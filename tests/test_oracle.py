"""The oracle pinned: against the reference's own CPU-only build (golden
fixtures, and the live build when present), against known answers, and
against what the mathematics fixes (order independence of linear-probing
occupancy, brute-force compaction and scan).  CPU only."""
import glob
import os

import numpy as np
import pytest

import oracle as O
from hashinator_amd import workloads as W
from tests.helpers import EMPTY, NDT, TOMB, live_pairs, occupancy, reachable_all

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# ------------------------------------------------------------------ hash (hashfunctions.h:39-56)
def test_fibhash_known_answers():
    L = O.lib()
    assert L.or_fibhash32(1, 20) == 648055
    assert L.or_fibhash32(0xDEADBEEF, 20) == 609062230  # unmasked high bits kept (hashfunctions.h:41)
    assert L.or_fibhash32(0xDEADBEEF, 20) & ((1 << 20) - 1) == 888150
    assert L.or_fibhash64(1, 28) == 165902235


@pytest.mark.parametrize("bits", [32, 64])
def test_fibhash_golden(bits):
    g = np.load(os.path.join(GOLD, f"fibhash{bits}.npz"))
    assert np.array_equal(O.fibhash(g["keys"], g["sps"], bits), g["hash"])


@pytest.mark.parametrize("bits", [32, 64])
def test_fibhash_closed_form(bits):
    rng = np.random.default_rng(bits)
    keys = rng.integers(0, 2**bits - 1, size=2000, dtype=np.uint64, endpoint=True)
    for sp in (1, 7, 20, 31):
        got = O.fibhash(keys.astype(NDT[bits]), sp, bits)
        for k, h in zip(keys[:200].tolist(), got[:200].tolist()):
            if bits == 32:
                k2 = k ^ (k >> (32 - sp))
                exp = ((k2 * 2654435769) >> (32 - sp)) & 0xFFFFFFFF
            else:
                k2 = k ^ (k >> (64 - sp))
                exp = ((k2 * 11400714819323198485) % 2**64) >> (64 - sp)
            assert h == exp


# ------------------------------------------------------------------ host-only map vs the reference build
FIX = sorted(glob.glob(os.path.join(GOLD, "cpumap*_sp*.npz")))


@pytest.mark.parametrize("path", FIX, ids=[os.path.basename(p) for p in FIX])
def test_cpu_map_golden_images(path):
    bits = 32 if "cpumap32" in path else 64
    sp = int(path.split("_sp")[1].split(".")[0])
    g = np.load(path)
    m = O.CpuMap(sp, bits)
    m.insert(g["k1"], g["v1"])
    assert np.array_equal(m.buckets(), g["img_insert"])
    assert np.array_equal(np.array(m.info().as_tuple(), dtype=np.uint64), g["info_insert"])
    assert m.erase(g["er"]) == int(g["n_erased"][0])
    assert np.array_equal(m.buckets(), g["img_erase"])
    assert np.array_equal(np.array(m.info().as_tuple(), dtype=np.uint64), g["info_erase"])
    m.insert(g["k2"], g["v2"])  # tombstone reuse with the look-ahead (hashinator.h:362-392)
    assert np.array_equal(m.buckets(), g["img_reinsert"])
    assert np.array_equal(np.array(m.info().as_tuple(), dtype=np.uint64), g["info_reinsert"])
    v, h = m.find(g["q"])
    assert np.array_equal(h, g["find_hit"]) and np.array_equal(np.where(h, v, 0), g["find_vals"])
    m.rehash(sp + 1)
    assert np.array_equal(m.buckets(), g["img_rehash"])
    assert np.array_equal(np.array(m.info().as_tuple(), dtype=np.uint64), g["info_rehash"])
    v, h = m.find(g["q"])
    assert np.array_equal(h, g["find2_hit"]) and np.array_equal(np.where(h, v, 0), g["find2_vals"])


@pytest.mark.skipif(not O.ref_available(), reason="reference CPU-only build not present (build container only)")
@pytest.mark.parametrize("bits", [32, 64])
def test_cpu_map_vs_live_reference(bits):
    rng = np.random.default_rng(bits)
    for trial in range(25):
        sp = int(rng.integers(3, 12))
        a, b = O.CpuMap(sp, bits), O.ref_map(sp, bits)
        for step in range(5):
            n = int(rng.integers(1, 1 << sp))
            keys = rng.integers(0, 4 << sp, n).astype(NDT[bits])
            vals = rng.integers(0, 2**31, n).astype(NDT[bits])
            a.insert(keys, vals)
            b.insert(keys, vals)
            er = keys[rng.random(n) < 0.4]
            assert a.erase(er) == b.erase(er)
            assert np.array_equal(a.buckets(), b.buckets())
            assert a.info().as_tuple() == b.info().as_tuple()
            if step == 2:
                a.rehash(a.info().size_power + 1)
                b.rehash(b.info().size_power + 1)


# ------------------------------------------------------------------ accelerated-path model
@pytest.mark.parametrize("path", FIX, ids=[os.path.basename(p) for p in FIX])
def test_gpu_model_equals_reference_without_tombstones(path):
    """With no tombstones, the accelerated insert (first EMPTY in probe order,
    overwrite on match) applied serially is the reference's _at: same image."""
    bits = 32 if "cpumap32" in path else 64
    sp = int(path.split("_sp")[1].split(".")[0])
    g = np.load(path)
    m = O.GpuModel(sp, bits)
    m.insert(g["k1"], g["v1"])
    assert np.array_equal(m.b, g["img_insert"])
    assert m.info.fill == int(g["info_insert"][1])
    # lookups on the reference's post-erase image agree with the reference's find
    m2 = O.GpuModel(sp, bits, buckets=g["img_reinsert"],
                    info=O.Info(sp, 0, int(g["info_reinsert"][1]), 1 << sp, int(g["info_reinsert"][3]), 0, 0))
    v, h = m2.retrieve(g["q"])
    assert np.array_equal(h, g["find_hit"]) and np.array_equal(np.where(h, v, 0), g["find_vals"])


@pytest.mark.parametrize("bits", [32, 64])
def test_gpu_model_occupancy_is_order_independent(bits):
    sp = 12
    keys = W.unique_keys(int(0.93 * (1 << sp)), bits, 17)
    base = O.GpuModel(sp, bits)
    base.insert(keys, keys)
    rng = np.random.default_rng(1)
    for _ in range(3):
        p = rng.permutation(keys.size)
        m = O.GpuModel(sp, bits)
        m.insert(keys[p], keys[p])
        assert np.array_equal(occupancy(m.b, bits), occupancy(base.b, bits))
        assert np.array_equal(live_pairs(m.b, bits), live_pairs(base.b, bits))
        assert reachable_all(m.b, bits, sp)


@pytest.mark.parametrize("bits", [32, 64])
def test_gpu_model_set_semantics_vs_dict(bits):
    sp = 12  # room for every key of the 0..2999 range
    rng = np.random.default_rng(bits + 3)
    m = O.GpuModel(sp, bits)
    ref = {}
    for rnd in range(6):
        keys = rng.integers(0, 3000, 300).astype(NDT[bits])
        vals = (keys.astype(np.uint64) * 7 + rnd).astype(NDT[bits])  # one value per key within a batch
        m.insert(keys, vals)
        ref.update(zip(keys.tolist(), vals.tolist()))
        er = rng.integers(0, 3000, 100).astype(NDT[bits])
        m.erase(er)
        for k in set(er.tolist()):
            ref.pop(k, None)
        if rnd % 2:
            m.clean_tombstones()
            assert m.info.tombstones == 0 and not np.any(m.b.reshape(-1, 2)[:, 0] == TOMB[bits])
        assert m.info.fill == len(ref)
        lp = live_pairs(m.b, bits)
        assert dict(zip(lp[:, 0].tolist(), lp[:, 1].tolist())) == ref
        q = np.arange(3000, dtype=NDT[bits])
        v, h = m.retrieve(q)
        assert [int(x) for x in q[h]] == sorted(ref)
        assert all(ref[int(k)] == int(x) for k, x in zip(q[h], v[h]))


@pytest.mark.parametrize("bits", [32, 64])
def test_gpu_model_clean_and_rehash_postconditions(bits):
    sp = 12
    keys = W.unique_keys(int(0.9 * (1 << sp)), bits, 23)
    m = O.GpuModel(sp, bits)
    m.insert(keys, keys)
    m.erase(keys[::4])
    live = live_pairs(m.b, bits)
    m.clean_tombstones()
    assert np.array_equal(live_pairs(m.b, bits), live)
    fresh = O.GpuModel(sp, bits)
    fresh.insert(live[:, 0].copy(), live[:, 1].copy())
    assert np.array_equal(occupancy(m.b, bits), occupancy(fresh.b, bits))
    assert reachable_all(m.b, bits, sp)
    assert m.info.overflow == max(64, -(-m.max_probe() // 64) * 64)
    assert m.rehash(sp + 1) == 0
    assert np.array_equal(live_pairs(m.b, bits), live) and m.info.size_power == sp + 1
    assert reachable_all(m.b, bits, sp + 1)


# ------------------------------------------------------------------ clear(device) and reset_to_empty
def _tombstoned_model(bits, sp=10, seed=61):
    """A table with live keys, TOMBSTONEs, non-zero values and an overflow above 64."""
    keys = W.unique_keys(int(0.95 * (1 << sp)), bits, seed)
    m = O.GpuModel(sp, bits)
    m.insert(keys, (keys ^ NDT[bits](0x5A5A)).astype(NDT[bits]))
    m.erase(keys[::5])
    assert m.info.tombstones > 0 and m.info.overflow > 64
    m.info.err = 0
    return m, keys


def reset_all_closed_form(b0, info0, n_len, bits):
    """kernels_AMD.h:26-43 as the closed form: slots [0, len) get key EMPTY,
    every value stays, fill = 0; sizePower, overflow, tombstones and err stay.
    len == 0 launches no block (hashers.h:155-163), so nothing changes."""
    img = np.asarray(b0).reshape(-1, 2).copy()
    t = info0.as_tuple()
    if n_len == 0:
        return img, t
    img[:n_len, 0] = EMPTY[bits]
    return img, (t[0], 0, t[2], t[3], t[4])


@pytest.mark.parametrize("bits", [32, 64])
def test_reset_all_closed_form(bits):
    m0, _ = _tombstoned_model(bits)
    n = 1 << m0.info.size_power
    for n_len in (0, 1, 333, n - 1, n):
        m = O.GpuModel(m0.info.size_power, bits, buckets=m0.b, info=m0.info)
        m.reset_all(n_len)
        exp_img, exp_info = reset_all_closed_form(m0.b, m0.info, n_len, bits)
        assert np.array_equal(m.b.reshape(-1, 2), exp_img), n_len
        assert m.info.as_tuple() == exp_info, n_len


@pytest.mark.parametrize("bits", [32, 64])
def test_reset_all_check_catches_planted_mistakes(bits):
    """clear(host) semantics (hashinator.h:494-498: fresh buckets {EMPTY, V()}
    and Info(sizePower)) and other near misses fail the closed form."""
    m0, _ = _tombstoned_model(bits)
    n = 1 << m0.info.size_power
    m = O.GpuModel(m0.info.size_power, bits, buckets=m0.b, info=m0.info)
    m.reset_all(n)
    exp_img, exp_info = reset_all_closed_form(m0.b, m0.info, n, bits)
    got_img, got_info = m.b.reshape(-1, 2), m.info.as_tuple()
    assert np.array_equal(got_img, exp_img) and got_info == exp_info
    # tombstone counter reset, as clear(host) does
    assert (got_info[0], 0, got_info[2], 0, got_info[4]) != exp_info
    # overflow back to BUCKET_OVERFLOW, as Info(sizePower) has it
    assert (got_info[0], 0, 64, got_info[3], got_info[4]) != exp_info
    # values reset to V()
    bad = got_img.copy()
    bad[:, 1] = 0
    assert not np.array_equal(bad, exp_img)
    # only live keys reset (TOMBSTONE slots left as they were)
    bad = got_img.copy()
    tomb = np.asarray(m0.b).reshape(-1, 2)[:, 0] == TOMB[bits]
    bad[tomb, 0] = TOMB[bits]
    assert tomb.any() and not np.array_equal(bad, exp_img)
    # an off-by-one in the slot range
    m1 = O.GpuModel(m0.info.size_power, bits, buckets=m0.b, info=m0.info)
    m1.reset_all(334)
    assert not np.array_equal(m1.b.reshape(-1, 2), reset_all_closed_form(m0.b, m0.info, 333, bits)[0])


def _key_with_home(bits, sp, home, skip=0):
    """The smallest keys (after `skip` of them) whose home slot is `home`."""
    k = 0
    found = 0
    while True:
        if int(O.fibhash(np.array([k], dtype=NDT[bits]), sp, bits)[0]) & ((1 << sp) - 1) == home:
            if found == skip:
                return NDT[bits](k)
            found += 1
        k += 1


@pytest.mark.parametrize("bits", [32, 64])
def test_reset_known_answers(bits):
    """Hand-built 8-slot images with the outcome read off kernels_AMD.h:65-134:
    the first copy in probe order from the home slot is emptied, across EMPTY
    slots and around the end of the table; the value stays; fill -= len."""
    sp = 3
    E = EMPTY[bits]
    a = _key_with_home(bits, sp, 6)
    b = _key_with_home(bits, sp, 2)
    c = _key_with_home(bits, sp, 6, skip=1)  # never stored: an absent key
    img = np.array([[E, 0], [a, 11], [b, 22], [E, 0], [E, 0], [E, 0], [E, 0], [a, 77]], dtype=NDT[bits])
    # a's home is slot 6 (EMPTY); probing 6, 7, 0, 1, ... finds the copy in slot 7 first
    info = O.Info(sp, 0, 3, 64, 0, 0, 0)
    m = O.GpuModel(sp, bits, buckets=img.reshape(-1), info=info)
    m.reset(np.array([[a, 999]], dtype=NDT[bits]))
    exp = img.copy()
    exp[7, 0] = E  # slot 7, not slot 1 (the lower index) and not nothing (EMPTY at 6 does not stop)
    assert np.array_equal(m.b.reshape(-1, 2), exp)
    assert m.info.fill == 2 and m.info.as_tuple()[2:] == (64, 0, 0)
    # the second reset of a finds the copy in slot 1, past EMPTY slots 6 and 7 and the wrap
    m.reset(np.array([[a, 0]], dtype=NDT[bits]))
    exp[1, 0] = E
    assert np.array_equal(m.b.reshape(-1, 2), exp) and m.info.fill == 1
    # an absent key changes no slot but still counts in fill -= len
    m.reset(np.array([[c, 0], [b, 0]], dtype=NDT[bits]))
    exp[2, 0] = E
    assert np.array_equal(m.b.reshape(-1, 2), exp)
    assert m.info.fill == np.uint64(2**64 - 1)  # 1 - 2 wraps, as size_t does
    # len == 0 changes nothing
    before = m.b.copy(), m.info.as_tuple()
    m.reset(np.empty((0, 2), dtype=NDT[bits]))
    assert np.array_equal(m.b, before[0]) and m.info.as_tuple() == before[1]


@pytest.mark.parametrize("bits", [32, 64])
def test_reset_set_postconditions(bits):
    """On a consistent table (each key at most once): exactly the slots of the
    reset keys that are present become EMPTY, values and every other slot stay,
    fill drops by len (absent keys and repeats included)."""
    m0, keys = _tombstoned_model(bits, sp=11, seed=67)
    img0 = m0.b.reshape(-1, 2).copy()
    rng = np.random.default_rng(bits)
    present = img0[(img0[:, 0] != EMPTY[bits]) & (img0[:, 0] != TOMB[bits]), 0]
    pick = rng.choice(present, present.size // 3, replace=False)
    absent = W.unique_keys(50, bits, 68, start=10**6)
    src_keys = np.concatenate([pick, absent, pick[:5]])  # five repeats: found once, counted twice
    rng.shuffle(src_keys)
    src = np.stack([src_keys, np.zeros_like(src_keys)], axis=1)
    m = O.GpuModel(m0.info.size_power, bits, buckets=m0.b, info=m0.info)
    m.reset(src)
    img = m.b.reshape(-1, 2)
    changed = np.nonzero(img[:, 0] != img0[:, 0])[0]
    assert np.array_equal(np.sort(img0[changed, 0]), np.sort(pick))
    assert np.all(img[changed, 0] == EMPTY[bits])
    assert np.array_equal(img[:, 1], img0[:, 1])
    assert m.info.fill == m0.info.fill - src_keys.size
    assert m.info.as_tuple()[2:] == m0.info.as_tuple()[2:]


def test_checks_catch_planted_mistakes():
    """The invariants used on the GPU results fail on a broken image."""
    bits, sp = 64, 10
    keys = W.unique_keys(900, bits, 29)
    m = O.GpuModel(sp, bits)
    m.insert(keys, keys)
    img = m.b.reshape(-1, 2).copy()
    # an EMPTY hole planted inside a probe path breaks reachability
    occ = np.nonzero(img[:, 0] != EMPTY[bits])[0]
    home = O.fibhash(img[occ, 0], sp, bits).astype(np.int64) & ((1 << sp) - 1)
    displaced = occ[home != occ]
    victim = displaced[0] - 1
    bad = img.copy()
    bad[victim % (1 << sp), 0] = EMPTY[bits]
    assert not reachable_all(bad, bits, sp)
    # swapping two keys' values changes the live set
    bad = img.copy()
    a, b = occ[0], occ[1]
    bad[a, 1], bad[b, 1] = img[b, 1] + 1, img[a, 1]
    assert not np.array_equal(live_pairs(bad, bits), live_pairs(img, bits))
    # a wrong hash (dropped xor-shift term) changes the image
    wrong = np.empty_like(img)
    wrong[:] = [EMPTY[bits], 0]
    for k in keys:
        h = ((int(k) * 11400714819323198485) % 2**64) >> (64 - sp)
        while wrong[h, 0] != EMPTY[bits]:
            h = (h + 1) % (1 << sp)
        wrong[h] = [k, k]
    assert not np.array_equal(wrong, img)


# ------------------------------------------------------------------ Robin Hood reorder
RH = sorted(glob.glob(os.path.join(GOLD, "rhorder*_sp*.npz")))


@pytest.mark.parametrize("path", RH, ids=[os.path.basename(p) for p in RH])
def test_reorder_golden(path):
    """The reference's own map, fed the keys in sorted (home, key) order,
    produces the Robin Hood image; the oracle reaches it from any arrival order."""
    g = np.load(path)
    bits = 32 if "rhorder32" in path else 64
    sp = int(g["sp"][0])
    m = O.GpuModel(sp, bits)
    m.insert(g["keys"], g["vals"])
    assert m.reorder() == 1
    assert np.array_equal(m.b, g["img"])


@pytest.mark.skipif(not O.ref_available(), reason="reference CPU-only build not present (build container only)")
@pytest.mark.parametrize("bits", [32, 64])
def test_reorder_vs_live_reference(bits):
    hits = 0
    for sp in (5, 8, 11):
        for seed in range(12):
            keys = W.unique_keys(int(0.7 * (1 << sp)), bits, 500 + 16 * seed + sp)
            m = O.GpuModel(sp, bits)
            m.insert(keys, keys)
            assert m.reorder() == 1
            home = O.fibhash(keys, sp, bits).astype(np.uint64) & np.uint64((1 << sp) - 1)
            order = np.lexsort((keys, home))
            r = O.ref_map(sp, bits)
            r.insert(keys[order], keys[order])
            if r.info().size_power == sp and m.b.reshape(-1, 2)[-1, 0] == EMPTY[bits]:
                assert np.array_equal(r.buckets(), m.b)
                hits += 1
    assert hits >= 6


def rh_sorted(img, bits, sp):
    """Every cluster in non-decreasing (home offset from cluster start, key) order."""
    img = np.asarray(img).reshape(-1, 2)
    n = img.shape[0]
    occ = img[:, 0] != EMPTY[bits]
    e0 = int(np.nonzero(~occ)[0][0])
    prev = None
    for u in range(e0 + 1, e0 + 1 + n):
        q = u % n
        if not occ[q]:
            prev = None
            continue
        if prev is None:
            start = u
        h = int(O.fibhash(img[q:q + 1, 0], sp, bits)[0]) & (n - 1)
        cur = ((h - start) % n, int(img[q, 0]))
        if prev is not None and cur < prev:
            return False
        prev = cur
    return True


@pytest.mark.parametrize("bits", [32, 64])
@pytest.mark.parametrize("sp,lf", [(4, 0.9), (9, 0.97), (12, 0.85)])
def test_reorder_invariants(bits, sp, lf):
    """Same live set and occupancy, clusters sorted, longest probe not longer,
    idempotent; wrapping clusters included (high load, small tables)."""
    for seed in range(3):
        keys = W.unique_keys(int(lf * (1 << sp)), bits, 40 + seed)
        m = O.GpuModel(sp, bits)
        m.insert(keys, keys)
        before = m.b.copy()
        mp = m.max_probe()
        assert not rh_sorted(before, bits, sp) or sp == 4
        assert m.reorder() == 1
        assert rh_sorted(m.b, bits, sp)
        assert np.array_equal(live_pairs(m.b, bits), live_pairs(before, bits))
        assert np.array_equal(occupancy(m.b, bits), occupancy(before, bits))
        assert reachable_all(m.b, bits, sp)
        assert m.max_probe() <= mp
        assert m.info.overflow == max(64, -(-m.max_probe() // 64) * 64)
        again = m.b.copy()
        assert m.reorder() == 1 and np.array_equal(m.b, again)
        # lookups with the Robin Hood early stop (a sequential restatement) agree
        absent = W.unique_keys(keys.size, bits, 40 + seed, start=keys.size + 5)
        q = np.concatenate([keys, absent])
        v, h = m.retrieve(q)
        assert np.array_equal(h, np.isin(q, keys))
        assert np.array_equal(rh_lookup(m.b, bits, sp, q), h)


def rh_lookup(img, bits, sp, q):
    """Probe stops at EMPTY or at an occupant displaced less than the offset."""
    img = np.asarray(img).reshape(-1, 2)
    n = img.shape[0]
    out = np.zeros(q.size, dtype=bool)
    homes = O.fibhash(q, sp, bits).astype(np.int64) & (n - 1)
    okeys = img[:, 0]
    ohome = O.fibhash(okeys, sp, bits).astype(np.int64) & (n - 1)
    for j, (k, h) in enumerate(zip(q, homes)):
        for off in range(n):
            s = (h + off) % n
            if okeys[s] == k:
                out[j] = True
                break
            if okeys[s] == EMPTY[bits]:
                break
            if okeys[s] != TOMB[bits] and (s - ohome[s]) % n < off:
                break
    return out


def test_reorder_skips_tombstoned_and_full_tables():
    bits, sp = 64, 8
    keys = W.unique_keys(200, bits, 3)
    m = O.GpuModel(sp, bits)
    m.insert(keys, keys)
    m.erase(keys[:5])
    b = m.b.copy()
    assert m.reorder() == 0 and np.array_equal(m.b, b)
    full = O.GpuModel(4, bits)
    k = W.unique_keys(16, bits, 4)
    full.insert(k, k)
    b = full.b.copy()
    assert full.reorder() == 0 and np.array_equal(full.b, b)


def test_reorder_check_catches_planted_mistake():
    """A layout sorted by key alone (home ignored) fails the golden image."""
    g = np.load(RH[0])
    bits = 32 if "rhorder32" in RH[0] else 64
    sp = int(g["sp"][0])
    m = O.GpuModel(sp, bits)
    m.insert(g["keys"], g["vals"])
    img = m.b.reshape(-1, 2).copy()
    occ = img[:, 0] != EMPTY[bits]
    rows = img[occ]
    img[occ] = rows[np.argsort(rows[:, 0], kind="stable")]
    assert not np.array_equal(img.reshape(-1), g["img"])
    assert not rh_sorted(img, bits, sp)


# ------------------------------------------------------------------ split::tools
@pytest.mark.parametrize("n", [0, 1, 1000, 100_001])
def test_copy_if_and_scan_brute_force(n):
    rng = np.random.default_rng(n)
    pairs = rng.integers(0, 2**64 - 1, size=(n, 2), dtype=np.uint64, endpoint=True)
    p = O.make_pred(O.PRED_MASKEQ, 1, 1, 1, 64)
    out = O.copy_if_pairs(pairs.reshape(-1), p, 64).reshape(-1, 2)
    assert np.array_equal(out, pairs[(pairs[:, 1] & np.uint64(1)) == 1])
    keys = O.copy_keys_if_pairs(pairs.reshape(-1), p, 64)
    assert np.array_equal(keys, pairs[(pairs[:, 1] & np.uint64(1)) == 1][:, 0])
    x = rng.integers(0, 2**32 - 1, size=n, dtype=np.uint32, endpoint=True)
    ex = O.exclusive_scan(x)
    ref = np.concatenate([[0], np.cumsum(x.astype(np.uint64))[:-1]]).astype(np.uint64) & np.uint64(0xFFFFFFFF) \
        if n else np.zeros(0, np.uint64)
    assert np.array_equal(ex.astype(np.uint64), ref)
    f = (rng.random(n) < 0.3).astype(np.uint8)
    assert np.array_equal(O.copy_if_flags(x, f), x[f.astype(bool)])


def test_cpu_baseline_stages_small():
    """bench.py's CPU leg at a small size: every stage runs, checks itself
    and reports."""
    import bench
    from hashinator_amd import workloads as Wl
    r = bench.cpu_baseline(sp_sample=12)
    st = r["stages"]
    for k in ("insert_Mops", "retrieve_Mops", "erase_Mops", "clean_ms", "copy_if_GBps", "scan_GBps"):
        assert st[k] > 0
    assert r["host"]["nproc"] >= 1 and r["cores"] == 1
    if O.ref_available():
        b = bench.cpu_batched(Wl.batched_workload(n_maps=16, sp_hi=12))
        assert b["ok"] and b["threads"] >= 1
