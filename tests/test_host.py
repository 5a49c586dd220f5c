"""Host-side checks that need no GPU: the C-ABI library loads and exports
every entry point the header declares, host-only helpers behave, and the
benchmark's algorithmic byte model matches its closed forms."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from hashinator_amd import _lib
from hashinator_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in ("hashinator_amd.h", "hashinator_amd_calib.h")]


def declared_symbols():
    txt = "".join(open(h).read() for h in HEADERS)
    return sorted(set(re.findall(r"^HN_API\s+[\w\s\*]+?\b(hn_\w+)\s*\(", txt, flags=re.M)))


def test_header_declares_the_entry_points():
    names = declared_symbols()
    assert len(names) >= 30
    assert set(names) == set(_lib.SIGNATURES), "binding table and header differ"


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing
    assert _lib.load().hn_version().decode().startswith("hashinator_amd")


def test_info_layout_matches_reference():
    # Hashinator::Info: int sizePower@0, size_t fill@8, overflow@16, tombstones@24, status err@32 (40 B)
    I = _lib.hn_info
    assert C.sizeof(I) == 40
    assert [I.size_power.offset, I.fill.offset, I.current_max_bucket_overflow.offset,
            I.tombstone_counter.offset, I.err.offset] == [0, 8, 16, 24, 32]


def test_scratch_sizing_host_only():
    L = _lib.load()
    for sp in (10, 20, 28):
        table = (1 << sp) * 16
        default = L.hn_clean_tombstones_scratch_bytes_u64(sp, 0)
        # the segment pass: per-tile lists (12 B per 1024 slots), the reorder's, and a long-segment pool
        # for 1/32 of the table -- a tenth of the table or less from 2^20 slots on, never twice the table
        assert default >= (1 << sp) // 1024 * 12 + (1 << 20)
        assert sp < 20 or default <= table // 8
        anything = L.hn_clean_tombstones_scratch_bytes_u64(sp, (1 << 64) - 1)  # a full table included
        assert anything >= table and anything >= default
        assert L.hn_clean_tombstones_scratch_bytes_u64(sp, default + 4096) == default + 4096  # a budget is used whole
        b = L.hn_clean_tombstones_scratch_bytes_u64(sp, 1 << 16)  # too small for the segment pass: regions
        assert 0 < b <= 1 << 16
    assert L.hn_clean_tombstones_scratch_bytes_u32(0, 0) == 0  # invalid size power
    assert L.hn_compact_scratch_bytes(1 << 28) >= (1 << 28) // 8192 * 16  # two 8-byte state words per 8192-element tile
    assert L.hn_rehash_scratch_bytes() >= 8
    # rehash as a build: growing only; about one pair per old slot plus bookkeeping
    assert L.hn_rehash_build_scratch_bytes_u64(20, 20) == 0 and L.hn_rehash_build_scratch_bytes_u64(20, 19) == 0
    assert L.hn_rehash_build_scratch_bytes_u64(20, 21) >= (1 << 20) * 16
    assert L.hn_rehash_build_scratch_bytes_u32(20, 22) >= (1 << 20) * 8


def test_invalid_arguments_are_rejected_without_a_gpu():
    L = _lib.load()
    assert L.hn_clean_tombstones_u64(None, None, 10, None, 0, None) == _lib.HN_EINVAL
    assert L.hn_rehash_u32(None, 10, None, 11, None, None, None) == _lib.HN_EINVAL
    assert L.hn_rehash_build_u64(None, 10, None, 11, None, None, 0, None) == _lib.HN_EINVAL
    assert L.hn_insert_u64(None, None, None, None, 10, None, None) == _lib.HN_EINVAL
    t = _lib.hn_tune(3, 256, 0, 0)  # probe width must be a power of two
    buf = (C.c_uint64 * 8)()
    assert L.hn_retrieve_u64(buf, buf, None, buf, buf, 0, None, C.byref(t)) == 0  # len 0: no launch
    assert L.hn_copy_if_flags(buf, None, 4, 8, buf, buf, buf, 64, None) == _lib.HN_EINVAL


@pytest.mark.parametrize("bits", [32, 64])
def test_workload_keys_unique_and_sentinel_free(bits):
    k = W.unique_keys(200_000, bits, 5)
    a = W.unique_keys(200_000, bits, 5, start=200_002)
    both = np.concatenate([k, a])
    assert np.unique(both).size == both.size
    assert not np.any(both >= np.iinfo(both.dtype).max - 1)


def test_mixed64_and_absent():
    k = W.mixed64(300_000, 9)
    a = W.absent64(300_000, 9, 300_000)
    assert np.unique(k).size == k.size
    assert np.intersect1d(k, a).size == 0
    runs = k[k.size // 2:]
    assert np.all(runs < np.uint64(1 << 63)) and np.all(k[: k.size // 2] >= np.uint64(1 << 63))


def test_velocity_grid_keys():
    rng = np.random.default_rng(0)
    k = W.velocity_grid_keys(5000, (100, 100, 100), 20.0, rng)
    assert k.size == 5000 and np.unique(k).size == 5000 and k.max() < 200**3


def test_algorithmic_bytes_model():
    import bench

    # worked numbers of the byte model at alpha = 0.9, 16 B slots: hit 152 B, miss 872 B, insert 216 B
    assert abs(bench.alg_bytes_retrieve(1, 0, 0.9) - 152) < 1e-9
    assert abs(bench.alg_bytes_retrieve(0, 1, 0.9) - 872) < 1e-9
    assert abs(bench.alg_bytes_insert(1, 0.9) - 216) < 1e-9
    n = int(0.9 * (1 << 28))
    assert abs(bench.alg_bytes_retrieve(n, n, 0.9) / 1e9 - 247.39) < 0.01


def test_output_tensors_are_checked_before_a_launch():
    """retrieve's caller-provided vals/hit are rejected unless they are device
    tensors of the right dtype, contiguous and long enough."""
    import torch

    from hashinator_amd.hashmap import out_tensor

    dev = "cuda:0"
    with pytest.raises(ValueError, match="is on cpu"):
        out_tensor(torch.zeros(8, dtype=torch.int64), 8, torch.int64, dev, "vals")
    with pytest.raises(TypeError, match="must be torch.int64"):
        out_tensor(torch.zeros(8, dtype=torch.int32), 8, torch.int64, dev, "vals")
    with pytest.raises(TypeError, match="must be a torch tensor"):
        out_tensor(np.zeros(8, dtype=np.uint64), 8, torch.int64, dev, "vals")
    # on the CPU the remaining checks are reached with a CPU "map device"
    with pytest.raises(ValueError, match="contiguous"):
        out_tensor(torch.zeros(16, dtype=torch.uint8)[::2], 8, torch.uint8, "cpu", "hit")
    with pytest.raises(ValueError, match="8 needed"):
        out_tensor(torch.zeros(7, dtype=torch.uint8), 8, torch.uint8, "cpu", "hit")
    t = torch.zeros(9, dtype=torch.uint64)
    assert out_tensor(t, 8, torch.int64, "cpu", "vals").dtype == torch.int64
