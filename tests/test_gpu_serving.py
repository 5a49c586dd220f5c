"""The serving loop: every request's result is the same in the concurrent mix
as when the stream is replayed one request at a time, and both agree with the
oracle's sequential replay (retrieve values and hit flags are exact; the
final key/value sets and fills are exact)."""
import numpy as np
import pytest

import oracle as O
from hashinator_amd import workloads as W
from hashinator_amd.serving import INSERT, RETRIEVE, ERASE, ServingEngine, latency_report, make_requests
from tests.helpers import live_pairs

pytestmark = pytest.mark.gpu

SPS = [10, 11, 12, 13, 12, 11]


def build(seed=1):
    pre = []
    for m, sp in enumerate(SPS):
        k = W.unique_keys(int(0.4 * (1 << sp)), 64, 500 + m)
        pre.append((k, W.values(k, 99)))
    reqs = make_requests(SPS, 240, rate_hz=2e5, seed=seed, min_batch=20, max_batch=3000,
                         live_keys=[[k] for k, _ in pre])
    return pre, reqs


def engine(pre, n_streams=8, budget=1 << 30, bulk_rehash_min=1 << 19, merge_min_keys=None):
    e = ServingEngine(SPS, key_bits=64, n_streams=n_streams, hbm_budget_bytes=budget, build_rehash_min=bulk_rehash_min,
                      merge_min_keys=merge_min_keys)
    for m, (k, v) in enumerate(pre):
        e.preload(m, k, v)
    return e


def clone(reqs):
    from hashinator_amd.serving import Request

    return [Request(r.rid, r.map_id, r.kind, r.keys, r.vals, r.arrival) for r in reqs]


@pytest.mark.parametrize("realtime,budget,bulk_rehash_min,merge_min_keys",
                         [(False, 1 << 30, 1 << 19, None), (True, 1 << 30, 1 << 19, None),
                          (False, 200_000, 1 << 19, None), (False, 1 << 30, 0, None), (False, 1 << 30, 0, 64),
                          (True, 1 << 30, 0, 64)])
def test_mix_equals_alone_equals_oracle(gpu, realtime, budget, bulk_rehash_min, merge_min_keys):
    # bulk_rehash_min 0: every growth is the bulk rehash (one build from the live slots);
    # merge_min_keys 64: inserts of a quarter of a table take the reorder + merge
    pre, reqs = build(seed=3)
    mixed = clone(reqs)
    alone = clone(reqs)
    e1 = engine(pre, budget=budget, bulk_rehash_min=bulk_rehash_min, merge_min_keys=merge_min_keys)
    wall = e1.serve(mixed, realtime=realtime)
    rep = latency_report(mixed, wall)
    assert rep["requests"] == len(reqs) and all(r.t_done >= r.arrival for r in mixed)
    e2 = engine(pre, n_streams=1)
    e2.serve_alone(alone)
    # oracle: sequential per map on a table large enough never to grow
    models = []
    for m, (k, v) in enumerate(pre):
        g = O.GpuModel(18, 64)
        g.insert(k, v)
        models.append(g)
    expect = {}
    for r in sorted(reqs, key=lambda r: r.arrival):
        g = models[r.map_id]
        if r.kind == INSERT:
            g.insert(r.keys, r.vals)
        elif r.kind == ERASE:
            g.erase(r.keys)
        else:
            expect[r.rid] = g.retrieve(r.keys)
    for a, b in zip(mixed, alone):
        assert a.rid == b.rid
        if a.kind == RETRIEVE:
            ev, eh = expect[a.rid]
            assert np.array_equal(a.out_hit, eh) and np.array_equal(b.out_hit, eh)
            assert np.array_equal(a.out_vals[eh], ev[eh]) and np.array_equal(b.out_vals[eh], ev[eh])
    i1, i2 = e1.infos(), e2.infos()
    for m in range(len(SPS)):
        la, lb = e1.live_set(m), e2.live_set(m)
        assert np.array_equal(la, lb)
        assert np.array_equal(la, live_pairs(models[m].b, 64))
        assert i1[m][1] == i2[m][1] == models[m].info.fill
        assert (i1[m][4] & 0xFFFFFFFF) == 0
    assert e1.rehashes >= 1  # the stream crosses the 0.85 ceiling somewhere
    if merge_min_keys is not None:
        assert e1.merges >= 1


def test_serving_full_size(gpu):
    """The serving workload at full size: 64 resident u64 maps of 2^20..2^26
    slots preloaded to LF 0.4 (500 M keys), 300 Poisson requests of 1k..4M
    keys at 4,000 requests/s with growth at LF 0.85.  Every retrieve's hit
    flags and values and every map's final fill equal a sequential host
    replay of the stream in set semantics (workloads.ServingReplay)."""
    import torch

    rng = np.random.default_rng(5)
    sps = [int(x) for x in rng.integers(20, 27, size=64)]
    pre = [W.unique_keys(int(0.4 * (1 << sp)), 64, 7000 + m) for m, sp in enumerate(sps)]
    eng = ServingEngine(sps, key_bits=64, n_streams=8, hbm_budget_bytes=64 << 30)
    for m, k in enumerate(pre):
        eng.preload(m, k, W.values(k, 1))
    torch.cuda.synchronize()
    reqs = make_requests(sps, 300, rate_hz=4000.0, seed=4000, live_keys=[[k] for k in pre])
    eng.serve(reqs, realtime=True)
    infos = eng.infos()
    assert eng.rehashes >= 1
    assert not np.any(infos[:, 4] & 0xFFFFFFFF)  # err == SUCCESS on every map
    del pre
    rp = W.ServingReplay([(int(0.4 * (1 << sp)), 7000 + m, 1) for m, sp in enumerate(sps)])
    bad, checked = rp.check(reqs, infos[:, 1])
    assert not bad, bad[:10]
    assert checked > 10_000_000
