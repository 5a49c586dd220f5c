"""The serving loop's full-size checker (workloads.ServingReplay) against the
oracle's sequential map on a small request stream: every retrieve and every
final fill agree, and planted mistakes are caught.  CPU only."""
import numpy as np

import oracle as O
from hashinator_amd import workloads as W
from hashinator_amd.serving import ERASE, INSERT, RETRIEVE, make_requests

SPS = [10, 11, 12, 9]


def _stream(seed=5):
    pre = []
    for m, sp in enumerate(SPS):
        pre.append(W.unique_keys(int(0.4 * (1 << sp)), 64, 500 + m))
    reqs = make_requests(SPS, 300, rate_hz=1e5, seed=seed, min_batch=10, max_batch=2000,
                         live_keys=[[k] for k in pre], pinned=False)
    return pre, reqs


def _replay_both(pre, reqs, value_seed=99, drop_erases=False):
    rp = W.ServingReplay([(k.size, 500 + m, 99) for m, k in enumerate(pre)])
    rp.pre = [(n, ks, value_seed) for n, ks, _ in rp.pre]
    models = []
    for k in pre:
        g = O.GpuModel(18, 64)
        g.insert(k, W.values(k, 99))
        models.append(g)
    agree = True
    for r in sorted(reqs, key=lambda r: r.arrival):
        g = models[r.map_id]
        if r.kind == INSERT:
            g.insert(r.keys, r.vals)
            rp.insert(r.map_id, r.keys, r.vals)
        elif r.kind == ERASE:
            g.erase(r.keys)
            if not drop_erases:
                rp.erase(r.map_id, r.keys)
        else:
            ev, eh = g.retrieve(r.keys)
            v, h = rp.retrieve(r.map_id, r.keys)
            agree &= bool(np.array_equal(h, eh) and np.array_equal(v[h], ev[eh]))
    fills = all(rp.fill(m) == models[m].info.fill for m in range(len(SPS)))
    return agree, fills


def test_replay_equals_oracle():
    pre, reqs = _stream()
    kinds = {r.kind for r in reqs}
    assert kinds == {INSERT, RETRIEVE, ERASE}
    agree, fills = _replay_both(pre, reqs)
    assert agree and fills


def test_replay_catches_planted_mistakes():
    pre, reqs = _stream()
    assert _replay_both(pre, reqs, value_seed=98)[0] is False  # wrong preload values
    assert _replay_both(pre, reqs, drop_erases=True) != (True, True)  # erases forgotten


def test_preload_membership_by_inverse_bijection():
    k = W.unique_keys(50_000, 64, 7003)
    assert np.array_equal(W.inv_mix64(W.mix64(k)), k)
    assert W.in_unique_keys64(k, k.size, 7003).all()
    assert not W.in_unique_keys64(W.unique_keys(50_000, 64, 7003, start=50_002), k.size, 7003).any()
    assert not W.in_unique_keys64(W.unique_keys(5_000, 64, 7004), k.size, 7003).any()
