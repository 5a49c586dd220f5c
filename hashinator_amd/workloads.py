"""Seeded synthetic inputs for the map, compaction and batched workloads.

Keys come from bijections of consecutive indices, so a key set is unique by
construction (no dedupe pass) and a disjoint "absent" set is just another
index range of the same bijection.  Neither sentinel (EMPTYBUCKET = all ones,
TOMBSTONE = all ones - 1) is ever produced.

Seeds: ``SEED_BASE + workload id`` (0x48415348494E4154 = "HASHINAT").
"""
from __future__ import annotations

import numpy as np

SEED_BASE = 0x48415348494E4154
M64 = np.uint64(0xFFFFFFFFFFFFFFFF)
M63 = np.uint64((1 << 63) - 1)


def _u64(x):
    return np.asarray(x, dtype=np.uint64)


def mix64(x):
    """SplitMix64 finalizer: a bijection of uint64."""
    x = _u64(x).copy()
    with np.errstate(over="ignore"):
        x ^= x >> np.uint64(30)
        x *= np.uint64(0xBF58476D1CE4E5B9)
        x ^= x >> np.uint64(27)
        x *= np.uint64(0x94D049BB133111EB)
        x ^= x >> np.uint64(31)
    return x


def _bij_bits(x, bits):
    """A bijection of the integers [0, 2^bits) (odd multiplies and xorshifts mod 2^bits)."""
    m = np.uint64((1 << bits) - 1)
    x = _u64(x) & m
    h = np.uint64(bits // 2)
    with np.errstate(over="ignore"):
        x ^= x >> h
        x = (x * np.uint64(0x9E3779B97F4A7C15)) & m
        x ^= x >> h
        x = (x * np.uint64(0xC2B2AE3D27D4EB4F)) & m
        x ^= x >> h
    return x


def mix32(x):
    """A bijection of uint32 (xorshift-multiply), computed in 64-bit lanes."""
    x = _u64(x) & np.uint64(0xFFFFFFFF)
    m = np.uint64(0xFFFFFFFF)
    with np.errstate(over="ignore"):
        x ^= x >> np.uint64(16)
        x = (x * np.uint64(0x7FEB352D)) & m
        x ^= x >> np.uint64(15)
        x = (x * np.uint64(0x846CA68B)) & m
        x ^= x >> np.uint64(16)
    return x.astype(np.uint32)


def unique_keys(n, bits, seed, start=0):
    """n distinct keys (no sentinels), uniform over the key space.  Index
    ranges [start, start+n+2) of one seed never collide with other ranges."""
    idx = np.arange(start, start + n + 2, dtype=np.uint64)
    if bits == 32:
        off = np.uint64(mix64(np.uint64(seed)) & np.uint64(0xFFFFFFFF))
        with np.errstate(over="ignore"):
            k = mix32((idx + off) & np.uint64(0xFFFFFFFF))
        k = k[k < np.uint32(0xFFFFFFFE)]
    else:
        off = mix64(np.uint64(seed))
        with np.errstate(over="ignore"):
            idx += off
        k = mix64(idx)
        if (k >= np.uint64(0xFFFFFFFFFFFFFFFE)).any():
            k = k[k < np.uint64(0xFFFFFFFFFFFFFFFE)]
    return k[:n]


def values(keys, seed):
    """Uniform values, a function of (key, seed)."""
    v = mix64(_u64(keys) ^ mix64(np.uint64(seed) + np.uint64(1)))
    return v.astype(keys.dtype) if keys.dtype == np.uint32 else v


def mixed64(n, seed, start=0):
    """Keys of the large-map workload: half uniform uint64 (upper half of the
    key space, a 63-bit bijection), half contiguous runs of 8..512 ids in the
    lower half, each run in its own 1024-aligned block (a 53-bit bijection), so
    all keys are distinct by construction.  Returned unshuffled."""
    n_u = n // 2
    n_r = n - n_u
    idx = np.arange(start, start + n_u + 2, dtype=np.uint64)
    with np.errstate(over="ignore"):
        u = _bij_bits(idx + (mix64(np.uint64(seed)) & M63), 63) | np.uint64(1 << 63)
    u = u[u < np.uint64(0xFFFFFFFFFFFFFFFE)][:n_u]
    rng = np.random.default_rng(int(seed) & 0xFFFFFFFF)
    lens = rng.integers(8, 513, size=n_r // 8 + 2, dtype=np.int64)
    csum = np.cumsum(lens)
    n_runs = int(np.searchsorted(csum, n_r)) + 1
    lens = lens[:n_runs]
    lens[-1] -= int(csum[n_runs - 1] - n_r)
    with np.errstate(over="ignore"):
        starts = _bij_bits(np.arange(n_runs, dtype=np.uint64) + (mix64(np.uint64(seed) ^ np.uint64(0x5555)) &
                                                                   np.uint64((1 << 53) - 1)), 53) << np.uint64(10)
    run_id = np.repeat(np.arange(n_runs), lens)
    within = np.arange(n_r, dtype=np.uint64) - np.repeat((np.cumsum(lens) - lens).astype(np.uint64), lens)
    r = starts[run_id] + within
    return np.concatenate([u, r])


def absent64(n, seed, n_used):
    """Keys guaranteed absent from mixed64(n_used, seed): the next index range
    of the same 63-bit bijection (upper half)."""
    return mixed64_uniform_range(n, seed, start=n_used // 2 + 2)


def mixed64_uniform_range(n, seed, start):
    idx = np.arange(start, start + n + 2, dtype=np.uint64)
    with np.errstate(over="ignore"):
        u = _bij_bits(idx + (mix64(np.uint64(seed)) & M63), 63) | np.uint64(1 << 63)
    return u[u < np.uint64(0xFFFFFFFFFFFFFFFE)][:n]


def shuffled(a, seed):
    return a[np.random.default_rng(int(seed) & 0xFFFFFFFF).permutation(a.size)]


def velocity_grid_keys(n, center, sigma, rng, grid=200):
    """Keys of one batched-workload map: linear indices x + grid*y + grid^2*z
    of cells drawn around `center` with i.i.d. N(0, sigma^2) offsets, inside
    the grid, without duplicates."""
    out = np.empty(0, dtype=np.uint32)
    while out.size < n:
        m = max(64, 2 * (n - out.size))
        d = np.rint(rng.normal(0.0, sigma, size=(m, 3))).astype(np.int64) + np.asarray(center, dtype=np.int64)
        ok = np.all((d >= 0) & (d < grid), axis=1)
        d = d[ok]
        k = (d[:, 0] + grid * d[:, 1] + grid * grid * d[:, 2]).astype(np.uint32)
        out = np.unique(np.concatenate([out, k]))
    return rng.permutation(out)[:n]


def batched_workload(n_maps=4096, sp_lo=8, sp_hi=16, lf=0.6, grid=200, sigma=20.0, erase_frac=0.2,
                     seed=SEED_BASE + 4):
    """Many uneven uint32 maps (the batched workload): size powers uniform in
    [sp_lo, sp_hi] (log-uniform slot counts), each filled to `lf` with
    velocity-grid cell ids drawn around a per-map random centre
    (Maxwellian-like shell, `sigma` cells); queries are the keys plus as many
    absent in-grid cells, shuffled; `erase_frac` of the keys are erased.
    Returns (sps, ins_keys, ins_vals, qry_keys, ers_keys) as lists of arrays."""
    rng = np.random.default_rng(int(seed) & 0xFFFFFFFF)
    sps = rng.integers(sp_lo, sp_hi + 1, size=n_maps)
    ins, vals, qry, ers = [], [], [], []
    cells = grid ** 3
    member = np.zeros(cells, dtype=bool)  # the current map's keys, cleared after each map
    for m in range(n_maps):
        n = int(lf * (1 << int(sps[m])))
        center = rng.integers(0, grid, size=3)
        k = velocity_grid_keys(n, center, sigma, rng, grid)
        v = values(k, int(seed) + m)
        member[k] = True
        a = np.empty(0, dtype=np.uint32)
        while a.size < n:
            cand = rng.integers(0, cells, size=n - a.size + n // 4 + 16).astype(np.uint32)
            cand = cand[~member[cand]]
            a = np.unique(np.concatenate([a, cand]))
        member[k] = False
        a = rng.permutation(a)[:n]
        q = rng.permutation(np.concatenate([k, a]))
        e = rng.choice(k, int(erase_frac * n), replace=False) if n else np.zeros(0, np.uint32)
        ins.append(k)
        vals.append(v.astype(np.uint32))
        qry.append(q)
        ers.append(e.astype(np.uint32))
    return sps, ins, vals, qry, ers


_MIX_C1 = 0xBF58476D1CE4E5B9
_MIX_C2 = 0x94D049BB133111EB


def inv_mix64(y):
    """The inverse of mix64: each xor-shift and odd multiply undone in reverse order."""
    y = _u64(y).copy()
    i1 = np.uint64(pow(_MIX_C1, -1, 1 << 64))
    i2 = np.uint64(pow(_MIX_C2, -1, 1 << 64))
    with np.errstate(over="ignore"):
        y ^= (y >> np.uint64(31)) ^ (y >> np.uint64(62))
        y *= i2
        y ^= (y >> np.uint64(27)) ^ (y >> np.uint64(54))
        y *= i1
        y ^= (y >> np.uint64(30)) ^ (y >> np.uint64(60))
    return y


def in_unique_keys64(keys, n, seed):
    """Membership of `keys` in unique_keys(n, 64, seed) (start 0), without
    generating the set: the key's index under the inverse bijection is below
    n, plus one for each sentinel the generator skipped before that point."""
    off = mix64(np.uint64(seed))
    with np.errstate(over="ignore"):
        idx = inv_mix64(keys) - off
        skipped = np.sort(inv_mix64(np.array([0xFFFFFFFFFFFFFFFE, 0xFFFFFFFFFFFFFFFF], dtype=np.uint64)) - off)
    cut = n
    for s in skipped.tolist():
        if s < cut:
            cut += 1
    k = _u64(keys)
    return (idx < np.uint64(cut)) & (k < np.uint64(0xFFFFFFFFFFFFFFFE))


class ServingReplay:
    """Sequential replay, in set semantics, of a serving request stream over
    u64 maps, for checking the loop's results at full size.

    Map m starts with the keys unique_keys(n0, 64, key_seed) holding
    values(key, value_seed); that set is never materialised (membership by
    in_unique_keys64).  What the requests change is kept per map as sorted
    arrays: the key, its last inserted value and whether it is live."""

    def __init__(self, preloads):
        """preloads: one (n0, key_seed, value_seed) per map."""
        self.pre = [(int(n), int(ks), int(vs)) for n, ks, vs in preloads]
        self.dk = [np.empty(0, np.uint64) for _ in self.pre]
        self.dv = [np.empty(0, np.uint64) for _ in self.pre]
        self.live = [np.empty(0, bool) for _ in self.pre]

    def preloaded(self, m, keys):
        n, ks, _ = self.pre[m]
        return in_unique_keys64(keys, n, ks)

    def _find(self, m, skeys):
        """Positions of sorted keys in the map's change arrays, and which are there."""
        dk = self.dk[m]
        idx = np.searchsorted(dk, skeys)
        ok = idx < dk.size
        found = np.zeros(skeys.size, dtype=bool)
        found[ok] = dk[idx[ok]] == skeys[ok]
        return found, idx

    def _add(self, m, skeys, vals, idx, live):
        """Insert sorted keys absent from the change arrays at their positions."""
        self.dk[m] = np.insert(self.dk[m], idx, skeys)
        self.dv[m] = np.insert(self.dv[m], idx, vals)
        self.live[m] = np.insert(self.live[m], idx, np.full(skeys.size, live))

    def insert(self, m, keys, vals):
        o = np.argsort(_u64(keys))
        sk, sv = _u64(keys)[o], _u64(vals)[o]
        found, idx = self._find(m, sk)
        self.dv[m][idx[found]] = sv[found]
        self.live[m][idx[found]] = True
        self._add(m, sk[~found], sv[~found], idx[~found], True)

    def erase(self, m, keys):
        sk = np.sort(_u64(keys))
        found, idx = self._find(m, sk)
        self.live[m][idx[found]] = False
        new = ~found
        new[new] = self.preloaded(m, sk[new])
        self._add(m, sk[new], np.zeros(int(new.sum()), np.uint64), idx[new], False)

    def retrieve(self, m, keys):
        """(values, hit) the map gives for `keys` now; values where hit."""
        keys = _u64(keys)
        o = np.argsort(keys)
        sk = keys[o]
        found, idx = self._find(m, sk)
        hit_s = np.empty(sk.size, dtype=bool)
        val_s = np.empty(sk.size, dtype=np.uint64)
        hit_s[found] = self.live[m][idx[found]]
        val_s[found] = self.dv[m][idx[found]]
        rest = ~found
        hit_s[rest] = self.preloaded(m, sk[rest])
        val_s[rest] = values(sk[rest], self.pre[m][2])
        hit = np.empty_like(hit_s)
        val = np.empty_like(val_s)
        hit[o] = hit_s
        val[o] = val_s
        return val, hit

    def check(self, requests, fills):
        """Replay served `requests` (objects with rid, map_id, kind, keys, vals,
        arrival, and out_vals/out_hit for retrieves) in arrival order and
        compare every retrieve's hit flags and hit values, then each map's
        final fill against `fills`.  Returns the mismatches found (empty when
        everything agrees) and the number of retrieved keys checked."""
        bad, checked = [], 0
        for r in sorted(requests, key=lambda r: r.arrival):
            if r.kind == "insert":
                self.insert(r.map_id, r.keys, r.vals)
            elif r.kind == "erase":
                self.erase(r.map_id, r.keys)
            else:
                ev, eh = self.retrieve(r.map_id, r.keys)
                h = np.asarray(r.out_hit, dtype=bool)
                v = np.asarray(r.out_vals).view(np.uint64)
                checked += r.keys.size
                if not np.array_equal(h, eh):
                    bad.append(f"request {r.rid}: {int(np.count_nonzero(h != eh))} hit flags differ")
                elif not np.array_equal(v[h], ev[eh]):
                    bad.append(f"request {r.rid}: {int(np.count_nonzero(v[h] != ev[eh]))} values differ")
        for m in range(len(self.pre)):
            if int(fills[m]) != self.fill(m):
                bad.append(f"map {m}: fill {int(fills[m])}, replay {self.fill(m)}")
        return bad, checked

    def fill(self, m):
        pre = self.preloaded(m, self.dk[m])
        return self.pre[m][0] + int(np.count_nonzero(self.live[m] & ~pre)) - int(np.count_nonzero(~self.live[m] & pre))
