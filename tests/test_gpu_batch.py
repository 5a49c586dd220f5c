"""The one-launch batched path: every map's results equal the oracle's
sequential application of insert -> retrieve -> erase -> clean, and equal the
per-map path's; graph replays repeat the eager result."""
import numpy as np
import pytest
import torch

import oracle as O
from hashinator_amd import _lib, hn_tune
from hashinator_amd import workloads as W
from hashinator_amd.batch import BatchedMaps
from tests.helpers import live_pairs, next_overflow, occupancy, reachable_all

pytestmark = pytest.mark.gpu


def small_batch(key_bits=32, n_maps=96, seed=7):
    rng = np.random.default_rng(seed)
    sps = rng.integers(1, 13, size=n_maps)
    ins, vals, qry, ers = [], [], [], []
    for m, sp in enumerate(sps):
        n = int(rng.uniform(0.3, 0.95) * (1 << int(sp)))
        k = W.unique_keys(n, key_bits, seed * 1000 + m)
        a = W.unique_keys(max(1, n // 2), key_bits, seed * 1000 + m, start=n + 2)
        ins.append(k)
        vals.append(W.values(k, m).astype(k.dtype))
        qry.append(rng.permutation(np.concatenate([k, a])))
        ers.append(rng.choice(k, int(0.25 * n), replace=False) if n else k[:0])
    return sps, ins, vals, qry, ers


def check_against_oracle(bm, wl, key_bits, ops=_lib.HN_BATCH_ALL):
    sps, ins, vals, qry, ers = wl
    for m, sp in enumerate(sps):
        model = O.GpuModel(int(sp), key_bits)
        if ops & _lib.HN_BATCH_INSERT:
            model.insert(ins[m], vals[m])
        if ops & _lib.HN_BATCH_RETRIEVE:
            ev, eh = model.retrieve(qry[m])
            v, h = bm.results(m)
            assert np.array_equal(h, eh), m
            assert np.array_equal(v[h], ev[eh]), m
        if ops & _lib.HN_BATCH_ERASE:
            model.erase(ers[m])
        if ops & _lib.HN_BATCH_CLEAN:
            model.clean_tombstones()
        img = bm.buckets(m)
        info = bm.info(m)
        assert np.array_equal(live_pairs(img, key_bits), live_pairs(model.b, key_bits)), m
        assert np.array_equal(occupancy(img, key_bits), occupancy(model.b, key_bits)), m
        assert info.fill == model.info.fill and info.tombstone_counter == model.info.tombstones, m
        assert info.err == 0, m
        assert reachable_all(img, key_bits, int(sp)), m
        mp = getattr(O.lib(), f"or_gpu{key_bits}_max_probe")(np.ascontiguousarray(img.reshape(-1)).ctypes.data,
                                                              int(sp))
        assert info.current_max_bucket_overflow >= max(64, next_overflow(mp)), m


def check_bulk(bm, wl, key_bits, ops=_lib.HN_BATCH_ALL):
    """check_against_oracle with one download of every device array (many maps)."""
    sps, ins, vals, qry, ers = wl
    nd = np.uint32 if key_bits == 32 else np.uint64
    arena = bm.arena.cpu().numpy().view(nd)
    qv = bm.qry_vals.cpu().numpy().view(nd)
    qh = bm.qry_hit.cpu().numpy().astype(bool)
    infos = bm.all_infos()
    mp_fn = getattr(O.lib(), f"or_gpu{key_bits}_max_probe")
    for m, sp in enumerate(sps):
        sp = int(sp)
        model = O.GpuModel(sp, key_bits)
        if ops & _lib.HN_BATCH_INSERT:
            model.insert(ins[m], vals[m])
        if ops & _lib.HN_BATCH_RETRIEVE:
            ev, eh = model.retrieve(qry[m])
            lo, n = int(bm.qry_off[m]), int(bm.qry_len[m])
            h, v = qh[lo:lo + n], qv[lo:lo + n]
            assert np.array_equal(h, eh), m
            assert np.array_equal(v[h], ev[eh]), m
        if ops & _lib.HN_BATCH_ERASE:
            model.erase(ers[m])
        if ops & _lib.HN_BATCH_CLEAN:
            model.clean_tombstones()
        lo = int(bm.slot_off[m])
        img = np.ascontiguousarray(arena[lo:lo + (1 << sp)])
        fill, ovf, tomb = int(infos[m, 1]), int(infos[m, 2]), int(infos[m, 3])
        err = int(infos[m, 4]) & 0xFFFFFFFF
        assert np.array_equal(live_pairs(img, key_bits), live_pairs(model.b, key_bits)), m
        assert np.array_equal(occupancy(img, key_bits), occupancy(model.b, key_bits)), m
        assert fill == model.info.fill and tomb == model.info.tombstones and err == 0, m
        assert reachable_all(img, key_bits, sp), m
        assert ovf >= max(64, next_overflow(mp_fn(img.reshape(-1).ctypes.data, sp))), m


@pytest.mark.parametrize("key_bits", [32, 64])
@pytest.mark.parametrize("width,block", [(8, 512), (4, 256), (16, 1024), (1, 64)])
def test_batch_all_ops_vs_oracle(gpu, key_bits, width, block):
    wl = small_batch(key_bits, seed=width + block)
    bm = BatchedMaps(*wl, key_bits=key_bits)
    bm.init()
    bm.run(_lib.HN_BATCH_ALL, hn_tune(width, block, 0, 0))
    torch.cuda.synchronize()
    check_against_oracle(bm, wl, key_bits)


@pytest.mark.parametrize("key_bits", [32, 64])
def test_batch_phase_by_phase_and_per_map_agree(gpu, key_bits):
    wl = small_batch(key_bits, n_maps=40, seed=3)
    a = BatchedMaps(*wl, key_bits=key_bits)
    a.init()
    for op in (_lib.HN_BATCH_INSERT, _lib.HN_BATCH_RETRIEVE, _lib.HN_BATCH_ERASE, _lib.HN_BATCH_CLEAN):
        a.run(op)
    b = BatchedMaps(*wl, key_bits=key_bits)
    b.init_per_map()
    b.run_per_map(_lib.HN_BATCH_ALL)
    torch.cuda.synchronize()
    check_against_oracle(a, wl, key_bits)
    for m in range(len(wl[0])):
        va, ha = a.results(m)
        vb, hb = b.results(m)
        assert np.array_equal(ha, hb) and np.array_equal(va[ha], vb[hb])
        assert np.array_equal(live_pairs(a.buckets(m), key_bits), live_pairs(b.buckets(m), key_bits))


def test_batch_velocity_grid_workload_small(gpu):
    wl = W.batched_workload(n_maps=64, sp_lo=8, sp_hi=13, seed=11)
    bm = BatchedMaps(*wl, key_bits=32)
    bm.init()
    bm.run()
    torch.cuda.synchronize()
    check_against_oracle(bm, wl, 32)


def test_batch_velocity_grid_workload_full_size(gpu):
    """The full batched workload: 4,096 u32 maps of 2^8..2^16 slots at LF 0.6
    (35.7 M keys), every map's retrieve results, live set, occupancy and Info
    against the oracle's sequential insert -> retrieve -> erase -> clean."""
    wl = W.batched_workload()
    bm = BatchedMaps(*wl, key_bits=32)
    bm.init()
    bm.run()
    torch.cuda.synchronize()
    assert len(wl[0]) == 4096 and sum(len(k) for k in wl[1]) > 35_000_000
    check_bulk(bm, wl, 32)


@pytest.mark.parametrize("key_bits", [32, 64])
def test_batch_skewed_query_units(gpu, key_bits):
    """One map with 2 M queries (489 retrieve units) among 3,000 maps, most
    with no queries at all: every unit is served once, every map's results
    equal the oracle's."""
    rng = np.random.default_rng(17)
    n_maps = 3000
    sps = np.concatenate([[18], rng.integers(2, 9, size=n_maps - 1)])
    ins, vals, qry, ers = [], [], [], []
    for m, sp in enumerate(sps):
        n = int(0.7 * (1 << int(sp)))
        k = W.unique_keys(n, key_bits, 300 + m)
        ins.append(k)
        vals.append(W.values(k, m).astype(k.dtype))
        if m == 0:
            a = W.unique_keys(n, key_bits, 300 + m, start=n + 2)
            q = np.concatenate([np.tile(k, 8), a])[:2_000_000]
            qry.append(rng.permutation(q))
        elif m % 3 == 0:
            qry.append(rng.permutation(np.concatenate([k, W.unique_keys(5, key_bits, 300 + m, start=n + 2)])))
        else:
            qry.append(k[:0])
        ers.append(k[: n // 5])
    wl = (sps, ins, vals, qry, ers)
    bm = BatchedMaps(*wl, key_bits=key_bits)
    bm.init()
    bm.run()
    torch.cuda.synchronize()
    check_bulk(bm, wl, key_bits)


def test_batch_graph_replays_match_eager(gpu):
    wl = small_batch(32, n_maps=48, seed=5)
    bm = BatchedMaps(*wl, key_bits=32)
    bm.init()
    bm.run()
    torch.cuda.synchronize()
    eager_vals, eager_hit = bm.qry_vals.clone(), bm.qry_hit.clone()
    eager_infos = bm.infos.clone()
    eager_live = [live_pairs(bm.buckets(m), 32) for m in range(len(wl[0]))]
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        bm.init(stream=s)  # warm-up on the capture stream
        bm.run(stream=s)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    with torch.cuda.graph(g):
        bm.init()
        bm.run()
    for rep in range(3):
        bm.qry_vals.zero_()
        bm.qry_hit.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(bm.qry_hit, eager_hit), rep
        assert torch.equal(bm.qry_vals[eager_hit.bool()], eager_vals[eager_hit.bool()]), rep
        inf = bm.infos.cpu().numpy()
        ref = eager_infos.cpu().numpy()
        assert np.array_equal(inf[:, 1], ref[:, 1]) and np.array_equal(inf[:, 3], ref[:, 3]), rep
        for m in range(len(wl[0])):
            assert np.array_equal(live_pairs(bm.buckets(m), 32), eager_live[m]), (rep, m)


def test_batch_clean_without_enough_scratch_fails_cleanly(gpu):
    wl = small_batch(32, n_maps=4, seed=9)
    bm = BatchedMaps(*wl, key_bits=32)
    # shrink the first listed map's scratch capacity below its fill
    raw = bm.descs.cpu().numpy().copy()
    rec = raw[:80].view(np.uint64)
    rec[9] = 0  # scratch_cap
    bm.descs.copy_(torch.from_numpy(raw))
    bm.init()
    bm.run(_lib.HN_BATCH_INSERT | _lib.HN_BATCH_ERASE)
    torch.cuda.synchronize()
    m0 = int(bm.order[0])
    assert bm.info(m0).tombstone_counter > 0  # precondition: there is something to clean
    before = bm.buckets(m0).copy()
    bm.run(_lib.HN_BATCH_CLEAN)
    torch.cuda.synchronize()
    assert np.array_equal(bm.buckets(m0), before)
    assert bm.info(m0).err == 1 and bm.info(m0).tombstone_counter > 0
    for m in range(4):
        if m != m0:
            assert bm.info(m).tombstone_counter == 0
