"""The compaction entry point that allocates nothing
(hn_copy_if_pairs_noalloc_u64) against the one that takes a scratch buffer
(hn_copy_if_pairs_u64), through the C-ABI: same elements in the same order
and the same count, on two calls in a row (the second smaller, so state left
by the first must not leak into it), and sizes beyond its static scratch
rejected before any launch."""
import ctypes as C

import numpy as np
import pytest
import torch

from hashinator_amd import _lib
from hashinator_amd import workloads as W

pytestmark = pytest.mark.gpu


def _p(t):
    return C.c_void_p(t.data_ptr())


def _pairs(n, seed):
    keys = W.unique_keys(n, 64, seed)
    vals = W.values(keys, seed)
    a = np.stack([keys, vals], axis=1)
    return torch.from_numpy(a.view(np.int64).copy()).cuda()


def _with_scratch(L, src, pred, s):
    n = src.shape[0]
    out = torch.full_like(src, -1)
    cnt = torch.full((1,), -1, dtype=torch.int64, device=src.device)
    sb = L.hn_compact_scratch_bytes(n)
    sc = torch.empty(sb, dtype=torch.uint8, device=src.device)
    _lib.check(L.hn_copy_if_pairs_u64(_p(src), n, C.byref(pred), _p(out), _p(cnt), _p(sc), sb, s), "copy_if")
    torch.cuda.synchronize()
    return out, int(cnt.item())


def _no_alloc(L, src, pred, s):
    n = src.shape[0]
    out = torch.full_like(src, -1)
    cnt = torch.full((1,), -1, dtype=torch.int64, device=src.device)
    _lib.check(L.hn_copy_if_pairs_noalloc_u64(_p(src), n, C.byref(pred), _p(out), _p(cnt), s), "copy_if noalloc")
    torch.cuda.synchronize()
    return out, int(cnt.item())


@pytest.mark.parametrize("kind", [_lib.HN_PRED_VALID, _lib.HN_PRED_MASKEQ])
def test_no_alloc_copy_if_equals_scratch_variant(gpu, kind):
    L = _lib.load()
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    pred = _lib.hn_pred(kind, 1, 1, 1)  # MASKEQ: odd values; VALID: keys that are not sentinels
    for n, seed in ((3_000_000, 1), (12_345, 2)):  # the second call is smaller
        src = _pairs(n, seed)
        if kind == _lib.HN_PRED_VALID:
            src[::7, 0] = -1  # EMPTYBUCKET
            src[3::11, 0] = -2  # TOMBSTONE
        ref, rc = _with_scratch(L, src, pred, s)
        got, gc = _no_alloc(L, src, pred, s)
        assert gc == rc and 0 < gc < n
        assert torch.equal(got[:gc], ref[:rc])
        assert torch.equal(got[gc:], torch.full_like(got[gc:], -1))  # nothing written past the count


def test_no_alloc_rejects_sizes_beyond_its_scratch(gpu):
    L = _lib.load()
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    src = _pairs(16, 3)
    out = torch.empty_like(src)
    cnt = torch.zeros(1, dtype=torch.int64, device=src.device)
    pred = _lib.hn_pred(_lib.HN_PRED_ALL, 0, 0, 0)
    rc = L.hn_copy_if_pairs_noalloc_u64(_p(src), (1 << 32) + 1, C.byref(pred), _p(out), _p(cnt), s)
    assert rc == _lib.HN_ESCRATCH
    # an empty input is accepted and counts zero
    cnt.fill_(7)
    torch.cuda.synchronize()
    assert L.hn_copy_if_pairs_noalloc_u64(_p(src), 0, C.byref(pred), _p(out), _p(cnt), s) == 0
    torch.cuda.synchronize()
    assert int(cnt.item()) == 0
