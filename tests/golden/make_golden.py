"""Write the golden fixtures from the reference's own CPU-only build.

    make -C oracle ref && python tests/golden/make_golden.py

oracle/_ref/libhashref.so is kstppd/hashinator's Hashmap compiled with
-DHASHINATOR_CPU_ONLY_MODE (oracle/ref_driver.cpp).  This script only records
inputs and outputs; the fixtures are data.  It needs the reference tree, so it
runs in the build container only; the tests read the .npz files.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402
from hashinator_amd import workloads as W  # noqa: E402

NDT = {32: np.uint32, 64: np.uint64}


def hash_fixture(bits, n=4096, seed=1):
    rng = np.random.default_rng(seed + bits)
    maxsp = 32 if bits == 32 else 32  # Hashmap::hash returns uint32_t; sizePower <= 32 (hashinator.h:237)
    keys = rng.integers(0, np.iinfo(NDT[bits]).max, size=n, dtype=NDT[bits], endpoint=True)
    sps = rng.integers(1, maxsp + 1, size=n).astype(np.int32)
    R = O.ref()
    f = getattr(R, f"ref_fibhash{bits}")
    out = np.array([f(int(k), int(s)) for k, s in zip(keys, sps)], dtype=NDT[bits])
    return keys, sps, out


def map_fixture(bits, sp, seed):
    """insert (LF ~0.7) -> erase 30% -> insert fresh keys and half the erased
    ones (tombstone reuse) -> find -> rehash(sp+1) -> find, recording the
    bucket image and Info after every phase."""
    rng = np.random.default_rng(seed)
    N = 1 << sp
    k1 = W.unique_keys(int(0.7 * N), bits, seed)
    v1 = rng.integers(0, np.iinfo(NDT[bits]).max, size=k1.size, dtype=NDT[bits])
    er = rng.choice(k1, int(0.3 * k1.size), replace=False)
    k2 = np.concatenate([W.unique_keys(int(0.1 * N), bits, seed, start=k1.size + 2), er[: er.size // 2]])
    rng.shuffle(k2)
    v2 = rng.integers(0, np.iinfo(NDT[bits]).max, size=k2.size, dtype=NDT[bits])
    q = np.concatenate([k1, k2, W.unique_keys(N // 4, bits, seed, start=k1.size + k2.size + 4)])
    rng.shuffle(q)
    m = O.ref_map(sp, bits)
    rec = {"k1": k1, "v1": v1, "er": er, "k2": k2, "v2": v2, "q": q}
    m.insert(k1, v1)
    rec["img_insert"], rec["info_insert"] = m.buckets(), np.array(m.info().as_tuple(), dtype=np.uint64)
    rec["n_erased"] = np.array([m.erase(er)], dtype=np.uint64)
    rec["img_erase"], rec["info_erase"] = m.buckets(), np.array(m.info().as_tuple(), dtype=np.uint64)
    m.insert(k2, v2)
    rec["img_reinsert"], rec["info_reinsert"] = m.buckets(), np.array(m.info().as_tuple(), dtype=np.uint64)
    v, h = m.find(q)
    rec["find_vals"], rec["find_hit"] = np.where(h, v, 0).astype(NDT[bits]), h
    m.rehash(sp + 1)
    rec["img_rehash"], rec["info_rehash"] = m.buckets(), np.array(m.info().as_tuple(), dtype=np.uint64)
    v, h = m.find(q)
    rec["find2_vals"], rec["find2_hit"] = np.where(h, v, 0).astype(NDT[bits]), h
    return rec


def reorder_fixture(bits, sp, lf, seed):
    """The reference's own map fed keys in sorted (home, key) order: the Robin
    Hood layout.  Seeds are skipped until no cluster wraps past the last slot
    and the map kept its size, so the image is the canonical one."""
    R = O.ref()
    f = getattr(R, f"ref_fibhash{bits}")
    N = 1 << sp
    while True:
        keys = W.unique_keys(int(lf * N), bits, seed)
        vals = keys ^ NDT[bits](0x9E37)
        home = np.array([f(int(k), sp) for k in keys], dtype=np.uint64) & np.uint64(N - 1)
        order = np.lexsort((keys, home))
        m = O.ref_map(sp, bits)
        m.insert(keys[order], vals[order])
        img = m.buckets()
        if m.info().size_power == sp and img.reshape(-1, 2)[-1, 0] == np.iinfo(NDT[bits]).max:
            perm = np.random.default_rng(seed).permutation(keys.size)
            return {"keys": keys[perm], "vals": vals[perm], "sp": np.array([sp]), "img": img}
        seed += 1


def main():
    if not O.ref_available():
        sys.exit("oracle/_ref/libhashref.so is missing: run `make -C oracle ref` first")
    for bits in (32, 64):
        k, s, h = hash_fixture(bits)
        np.savez_compressed(os.path.join(HERE, f"fibhash{bits}.npz"), keys=k, sps=s, hash=h)
        for sp in (8, 10, 12):
            rec = map_fixture(bits, sp, seed=1000 * bits + sp)
            np.savez_compressed(os.path.join(HERE, f"cpumap{bits}_sp{sp}.npz"), **rec)
        for sp in (10, 12):
            rec = reorder_fixture(bits, sp, 0.85, seed=77 * bits + sp)
            np.savez_compressed(os.path.join(HERE, f"rhorder{bits}_sp{sp}.npz"), **rec)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
