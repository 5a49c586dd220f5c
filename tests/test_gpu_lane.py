"""One-key-per-lane lookups, the default retrieve/erase path (tune NULL or
probe_width 0): 8 lanes load each 128-byte window together, the windows are
staged through LDS and every lane scans its own.  Results must equal the
oracle's serial probe and the group-probe path's, bit for bit, on the same
tables: plain and Robin Hood ordered, tiny tables (below one window, which take
the group loop), full tables (no EMPTY slot: probes run to the bound) and
tables with tombstones.
"""
import numpy as np
import pytest
import torch

import oracle as O
from hashinator_amd import Hashmap, hn_tune, to_numpy
from hashinator_amd import _lib
from tests.helpers import NDT, info_tuple, keyset, model_from

pytestmark = pytest.mark.gpu

TD = {32: np.int32, 64: np.int64}


def make(bits, sp, lf, seed, ordered, tune=None):
    n = max(1, int(lf * (1 << sp)))
    keys = keyset(n, bits, seed)
    vals = keys ^ NDT[bits](0x3C3C)
    hm = Hashmap(sp, bits, tune=tune, auto_reorder=0, auto_build=0)
    hm.insert(keys, vals, targetLF=1.0)
    if ordered:
        seq = O.GpuModel(sp, bits)
        seq.insert(keys, vals)
        assert hm.reorder() == bool(seq.reorder())
    # the oracle probes a copy of the device image (a parallel insert places
    # keys in arrival order; the serial model would place them differently)
    return hm, model_from(hm.host_buckets(), info_tuple(hm), bits), keys


def queries(keys, bits, seed):
    absent = keyset(max(64, keys.size), bits, seed=seed + 1000, start=1 << 28)
    absent = absent[~np.isin(absent, keys)]
    q = np.concatenate([keys, absent, keys[: keys.size // 3]])  # duplicates too
    np.random.default_rng(seed).shuffle(q)
    return q


def retrieve_np(hm, q, bits, sentinel):
    vin = torch.from_numpy(np.full(q.size, sentinel, dtype=NDT[bits]).view(TD[bits])).cuda()
    _, v, h = hm.retrieve(q, vin)
    return to_numpy(v, bits), h.cpu().numpy().astype(bool)


@pytest.mark.parametrize("bits", [32, 64])
@pytest.mark.parametrize("ordered", [False, True])
@pytest.mark.parametrize("sp,lf", [(1, 1.0), (2, 0.75), (3, 1.0), (4, 0.9), (5, 1.0), (6, 0.99), (8, 0.9),
                                   (12, 0.95), (16, 0.9)])
def test_lane_retrieve_exact(gpu, bits, ordered, sp, lf):
    hm, model, keys = make(bits, sp, lf, seed=sp * 7 + bits, ordered=ordered)
    q = queries(keys, bits, seed=sp)
    sentinel = NDT[bits](12345)
    v, h = retrieve_np(hm, q, bits, sentinel)
    ev, eh = model.retrieve(q, np.full(q.size, sentinel, dtype=NDT[bits]))
    assert np.array_equal(h, eh)
    assert np.array_equal(v, ev)
    # the group-probe path on the same table gives the same answers
    hm.tune = hn_tune(8, 256, 0, 0)
    v2, h2 = retrieve_np(hm, q, bits, sentinel)
    assert np.array_equal(h2, eh) and np.array_equal(v2, ev)


@pytest.mark.parametrize("bits", [32, 64])
def test_lane_retrieve_pairs(gpu, bits):
    sp = 14
    hm, model, keys = make(bits, sp, 0.9, seed=3, ordered=True)
    q = queries(keys, bits, seed=4)
    pairs = np.zeros((q.size, 2), dtype=NDT[bits])
    pairs[:, 0] = q
    pairs[:, 1] = 77
    t = torch.from_numpy(pairs.view(TD[bits])).cuda()
    hm.retrieve(t)
    out = to_numpy(t, bits).reshape(-1, 2)
    ev, eh = model.retrieve(q, np.full(q.size, 77, dtype=NDT[bits]))
    assert np.array_equal(out[:, 0], q)
    assert np.array_equal(out[:, 1], ev)


@pytest.mark.parametrize("bits", [32, 64])
@pytest.mark.parametrize("ordered", [False, True])
@pytest.mark.parametrize("sp", [3, 10, 15])
def test_lane_erase_exact(gpu, bits, ordered, sp):
    hm, model, keys = make(bits, sp, 0.9, seed=sp + bits, ordered=ordered)
    absent = keyset(500, bits, seed=9, start=1 << 29)
    e = np.concatenate([keys[::2], absent, keys[:40]])  # absent keys and repeats
    np.random.default_rng(5).shuffle(e)
    hm.erase(e)
    model.erase(e)
    assert np.array_equal(hm.host_buckets().reshape(-1), model.b)
    _, fill, _, tomb, _ = info_tuple(hm)
    assert (fill, tomb) == (model.info.fill, model.info.tombstones)
    # lookups over tombstones
    q = queries(keys, bits, seed=11)
    v, h = retrieve_np(hm, q, bits, NDT[bits](5))
    ev, eh = model.retrieve(q, np.full(q.size, 5, dtype=NDT[bits]))
    assert np.array_equal(h, eh) and np.array_equal(v, ev)


@pytest.mark.parametrize("bits", [32, 64])
def test_lane_large_ordered(gpu, bits):
    sp = 21
    hm, model, keys = make(bits, sp, 0.9, seed=21, ordered=True)
    q = queries(keys, bits, seed=22)
    v, h = retrieve_np(hm, q, bits, NDT[bits](0))
    ev, eh = model.retrieve(q, np.zeros(q.size, dtype=NDT[bits]))
    assert np.array_equal(h, eh) and np.array_equal(v, ev)


@pytest.mark.parametrize("grid", [1, 3, 1000])
def test_lane_grid_sizes(gpu, grid):
    # one one-wave block for all queries, an odd count of blocks, more blocks than queries
    bits, sp = 64, 12
    hm, model, keys = make(bits, sp, 0.9, seed=grid, ordered=False, tune=hn_tune(0, 0, grid, 0))
    q = queries(keys, bits, seed=grid)[:2000]
    v, h = retrieve_np(hm, q, bits, NDT[bits](1))
    ev, eh = model.retrieve(q, np.ones(q.size, dtype=NDT[bits]))
    assert np.array_equal(h, eh) and np.array_equal(v, ev)


def test_lane_ignores_block_size(gpu):
    # probe_width 0 selects the lane kernel, which runs one wave per block whatever block_size says
    bits, sp = 64, 10
    hm, model, keys = make(bits, sp, 0.9, seed=2, ordered=False, tune=hn_tune(0, 512, 0, 0))
    q = queries(keys, bits, seed=2)
    v, h = retrieve_np(hm, q, bits, NDT[bits](1))
    ev, eh = model.retrieve(q, np.ones(q.size, dtype=NDT[bits]))
    assert np.array_equal(h, eh) and np.array_equal(v, ev)


@pytest.mark.parametrize("bits", [32, 64])
@pytest.mark.parametrize("nq", [1, 2, 63, 64, 65, 1000])
def test_lane_query_counts(gpu, bits, nq):
    # query counts around one wave's 64 lanes, and fewer queries than lanes
    hm, model, keys = make(bits, 12, 0.9, seed=nq + bits, ordered=True)
    q = queries(keys, bits, seed=nq)[:nq]
    v, h = retrieve_np(hm, q, bits, NDT[bits](9))
    ev, eh = model.retrieve(q, np.full(q.size, 9, dtype=NDT[bits]))
    assert np.array_equal(h, eh) and np.array_equal(v, ev)
    hm.erase(q)
    model.erase(q)
    assert np.array_equal(hm.host_buckets().reshape(-1), model.b)
