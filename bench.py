"""Benchmark: Mops/s of insert+retrieve on one large map at load factor 0.9.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Workload (one step): a fresh Hashmap<uint64_t,uint64_t> of 2^28 slots
(4 GiB of buckets), insert floor(0.9 * 2^28) = 241,591,910 distinct keys (half
uniform uint64, half runs of 8..512 consecutive ids), then retrieve all of
them plus as many absent keys (shuffled together).  Ops per step = 3 * n.
Inputs are synthetic, generated once, and resident in HBM before timing
starts.  The insert takes the path the library's maps take for a batch this
large into an empty map: the bulk build (hn_build_u64), which writes the
Robin Hood ordered table directly, every slot of it, so the fresh map needs
only its Info initialised (hn_info_init; no bucket pass).  --no-build
initialises the buckets (hn_table_init_u64) and times the probing insert
kernel followed by the reorder pass instead; --no-build --no-reorder the
reference's own layout (arrival order: every miss probes to an EMPTY slot).

Besides the headline line it reports, from the same process: per-stage
timings, the rest of the large-map sequence (erase 25%, clean_tombstones,
retrieve after clean), the other load factors, the HBM roofline of the
dominant kernel (retrieve) and the CPU baseline (the reference's host-only
algorithm on a bounded sample of the same workload, one core).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mops/s insert+retrieve @LF0.9; batched vs per-map; % HBM roofline"
PEAK_HBM_GBS = 8000.0
SECTOR = 64  # bytes per random HBM sector in the algorithmic model
MISS_VAL = -0x2152411021524111  # 0xDEADBEEFDEADBEEF: an absent key's output value, never written


def knuth_hit(a):
    return 0.5 * (1 + 1 / (1 - a))


def knuth_miss(a):
    return 0.5 * (1 + 1 / (1 - a) ** 2)


def rh_miss(a):
    """Expected slots examined by a miss in a Robin Hood ordered table: the
    home slot, the elements hashed to it (Poisson, mean a) and the elements
    carried past it from earlier homes (total displacement / N =
    a * (L_hit - 1) = a^2 / (2 (1 - a)))."""
    return 1 + a + a * a / (2 * (1 - a))


def sectors(L, slot):
    """Expected 64 B sectors spanned by L consecutive slots at a random start."""
    return 1 + (L - 1) * slot / SECTOR


def alg_bytes_retrieve(n_hit, n_miss, alpha, kv=16, slot=16, ordered=False):
    miss = rh_miss(alpha) if ordered else knuth_miss(alpha)
    return n_hit * (kv + SECTOR * sectors(knuth_hit(alpha), slot)) + \
        n_miss * (kv + SECTOR * sectors(miss, slot))


def alg_bytes_insert(n, alpha, kv=16, slot=16):
    return n * (kv + SECTOR * sectors(knuth_hit(alpha), slot) + SECTOR)


def read_traffic(name):
    """HBM bytes per launch measured by rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE,
    calibrated on a copy kernel) recorded in profiles/<name>."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("TRAFFIC_BYTES_PER_LAUNCH:"):
            return float(line.split(":", 1)[1])
    return None


def host_cpu():
    """nproc and the CPU model of the host this runs on."""
    model = "unknown"
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "cpu_model": model}


def cpu_threads():
    """Worker threads for the all-cores CPU leg: the host's cores, at most 16
    (the CPU share of one GPU on the measurement hosts)."""
    return max(1, min(os.cpu_count() or 1, 16))


def cpu_baseline(sp_sample=24, alpha=0.9, seed=0x48415348494E4154 + 2):
    """The reference's host-only algorithm on the large-map workload at
    2^sp_sample slots, one core, stage by stage: serial _at insert
    (hashinator.h:340-403), const find of as many present as absent keys
    (742-768), erase(key) of a quarter of the keys (842-850; its cleanup
    rehashes only above 25% tombstones per slot, and CPU-only inserts leave
    the overflow at 32, 631-639), and the CPU-only clean, a rehash at the
    same size (635-638); then compaction and scan of 2^sp_sample elements by
    the C restatement (std::copy_if / std::exclusive_scan semantics).  The
    map stages use the reference's own CPU-only build when it was built into
    oracle/_ref, else the C restatement."""
    import oracle as O
    from hashinator_amd import workloads as Wl

    n = int(alpha * (1 << sp_sample))
    keys = Wl.shuffled(Wl.mixed64(n, seed), seed)
    vals = Wl.values(keys, seed)
    absent = Wl.absent64(n, seed, n)
    q = Wl.shuffled(np.concatenate([keys, absent]), seed + 1)
    er = keys[np.random.default_rng(12345).choice(n, n // 4, replace=False)]
    kind = "reference" if O.ref_available() else "port"
    m = O.ref_map(sp_sample, 64) if kind == "reference" else O.CpuMap(sp_sample, 64)
    t0 = time.perf_counter()
    m.insert(keys, vals)
    t1 = time.perf_counter()
    v, h = m.find(q)
    t2 = time.perf_counter()
    assert int(h.sum()) == n
    erased = m.erase(er)
    t3 = time.perf_counter()
    assert erased == n // 4 and m.info().size_power == sp_sample
    m.rehash(sp_sample)
    t4 = time.perf_counter()
    i = m.info()
    assert i.fill == n - n // 4 and i.tombstones == 0
    m.close()
    # compaction: pairs with Bernoulli(0.3) keep in bit 0 of the value; scan: u32
    ne = 1 << sp_sample
    rng = np.random.default_rng(seed)
    pairs = np.empty((ne, 2), dtype=np.uint64)
    pairs[:, 0] = np.arange(ne, dtype=np.uint64)
    pairs[:, 1] = (rng.integers(0, 2**63, size=ne, dtype=np.uint64) << np.uint64(1)) | \
        (rng.random(ne) < 0.3).astype(np.uint64)
    pred = O.make_pred(O.PRED_MASKEQ, field=1, mask=1, value=1, width=64)
    t5 = time.perf_counter()
    kept = O.copy_if_pairs(pairs, pred, 64)
    t6 = time.perf_counter()
    x = rng.integers(0, 16, size=ne, dtype=np.uint32)
    t7 = time.perf_counter()
    y = O.exclusive_scan(x)
    t8 = time.perf_counter()
    assert y[-1] == np.uint32(int(x[:-1].sum()) & 0xFFFFFFFF)
    ops = n + q.size
    cb_bytes = ne * 16 + kept.shape[0] * 16
    return {
        "value": round(ops / (t2 - t0) / 1e6, 2),
        "unit": "Mops/s",
        "cores": 1,
        "kind": kind,
        "sample": (f"Hashmap<u64,u64>(sp={sp_sample}) at LF {alpha}: insert {n} keys (_at), then find {n} present + "
                   f"{n} absent keys (const find); same key distribution as the GPU workload at 1/{1 << (28 - sp_sample)} "
                   f"scale; {t2 - t0:.1f} s of CPU time (the stages below {t8 - t0:.1f} s)"),
        "host": host_cpu(),
        "stages": {
            "insert_Mops": round(n / (t1 - t0) / 1e6, 2),
            "retrieve_Mops": round(q.size / (t2 - t1) / 1e6, 2),
            "erase_Mops": round((n // 4) / (t3 - t2) / 1e6, 2),
            "clean_ms": round((t4 - t3) * 1e3, 1),
            "clean_note": "CPU-only mode cleans tombstones by a rehash at the same size (hashinator.h:635-638)",
            "copy_if_ms": round((t6 - t5) * 1e3, 1),
            "copy_if_GBps": round(cb_bytes / (t6 - t5) / 1e9, 2),
            "scan_ms": round((t8 - t7) * 1e3, 1),
            "scan_GBps": round(2 * ne * 4 / (t8 - t7) / 1e9, 2),
            "compaction_scan_kind": "port",
            "compaction_scan_sample": f"copy_if of {ne} hash_pair<u64,u64> (keep 0.3), exclusive scan of {ne} u32",
        },
    }


def cpu_batched(wl):
    """The 4,096-map workload on the host's cores: each map serial through the
    reference's CPU-only build (insert, const find, erase(key) of the erase
    list, clean as a rehash at the same size), maps spread over a thread pool
    (ctypes releases the GIL for the calls), largest first."""
    import oracle as O
    from concurrent.futures import ThreadPoolExecutor

    if not O.ref_available():
        return None
    sps, ins, vals, qry, ers = wl
    order = sorted(range(len(sps)), key=lambda m: -int(sps[m]))
    threads = cpu_threads()

    def one(mi):
        m = O.ref_map(int(sps[mi]), 32)
        m.insert(ins[mi], vals[mi])
        _, h = m.find(qry[mi])
        m.erase(ers[mi])
        m.rehash(int(sps[mi]))
        ok = int(h.sum()) == ins[mi].size and m.info().fill == ins[mi].size - ers[mi].size
        m.close()
        return ok

    O.ref()  # load the library before the clock starts
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=threads) as ex:
        oks = list(ex.map(one, order))
    dt = time.perf_counter() - t0
    ops = sum(a.size for a in ins) + sum(a.size for a in qry) + sum(a.size for a in ers)
    return {"Mops": round(ops / dt / 1e6, 2), "ms": round(dt * 1e3, 1), "threads": threads, "kind": "reference",
            "ok": all(oks), "host": host_cpu()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--sp", type=int, default=28)
    ap.add_argument("--lf", type=float, default=0.9)
    ap.add_argument("--probe-width", type=int, default=0)
    ap.add_argument("--no-reorder", action="store_true",
                    help="with --no-build: skip the Robin Hood reorder pass (arrival order, as in the reference)")
    ap.add_argument("--no-build", action="store_true",
                    help="insert with the probing kernel (hn_insert_u64) instead of the bulk build")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-extra", action="store_true", help="skip the erase/clean and load-factor extras")
    ap.add_argument("--no-batched", action="store_true", help="skip the many-maps (batched vs per-map) comparison")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving-loop measurement")
    ap.add_argument("--no-compaction", action="store_true", help="skip the 2^28-element compaction/scan measurement")
    ap.add_argument("--batched-only", action="store_true",
                    help="only the many-maps batched call, 3 runs with probe width 8 and 1024 threads per map "
                         "(for a kernel profile of that call)")
    ap.add_argument("--calib-sweep", action="store_true",
                    help="only measure the random-gather ceiling (16..256 B reads, 1/4/16 dependent reads) and exit")
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run one plain 1 GiB device copy (counter calibration for rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "bench.py runs on one GPU; --gpus > 1 is not supported "
                                   "(a map would shard by key hash across GPUs, one shard per GPU)"}))
        sys.exit(2)

    import torch

    from hashinator_amd import _lib, hn_tune
    from hashinator_amd import workloads as Wl

    L = _lib.load()
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    stream = torch.cuda.current_stream()
    S = C.c_void_p(stream.cuda_stream)

    def P(t):
        return C.c_void_p(t.data_ptr())

    if args.calib_sweep:
        print(json.dumps(calib_sweep(torch, L, dev, S, P)), flush=True)
        return
    if args.batched_only:
        print(json.dumps(batched_only(torch)), flush=True)
        return

    tune = C.byref(hn_tune(args.probe_width, 0, 0, 0)) if args.probe_width else None
    sp, alpha = args.sp, args.lf
    N = 1 << sp
    seed = Wl.SEED_BASE + 2

    def make_inputs(alpha):
        n = int(alpha * N)
        keys = Wl.mixed64(n, seed)
        vals = Wl.values(keys, seed)
        absent = Wl.absent64(n, seed, n)
        g = torch.Generator(device=dev).manual_seed(seed & 0x7FFFFFFF)
        dk = torch.from_numpy(keys.view(np.int64)).to(dev)
        dv = torch.from_numpy(vals.view(np.int64)).to(dev)
        perm = torch.randperm(n, device=dev, generator=g)
        dk, dv = dk[perm].contiguous(), dv[perm].contiguous()
        q = torch.cat([dk, torch.from_numpy(absent.view(np.int64)).to(dev)])
        qperm = torch.randperm(q.numel(), device=dev, generator=g)
        make_inputs.qperm = qperm  # query i is key qperm[i] of [present..., absent...]
        return n, dk, dv, q[qperm].contiguous()

    t_gen = time.perf_counter()
    n, dk, dv, q = make_inputs(alpha)
    buckets = torch.empty((N, 2), dtype=torch.int64, device=dev)
    info = torch.zeros(5, dtype=torch.int64, device=dev)
    # The step's retrieve writes values only, as the reference's retrieve does
    # (a miss leaves its value untouched): `out` starts as MISS_VAL, so after
    # any number of steps it must equal the present keys' values with
    # MISS_VAL at the absent keys.
    out = torch.full_like(q, MISS_VAL)
    expect = torch.cat([dv, torch.full((q.numel() - n,), MISS_VAL, dtype=torch.int64, device=dev)])[make_inputs.qperm]
    hit = torch.empty(q.numel(), dtype=torch.uint8, device=dev)  # for the extra measurements
    rsb = L.hn_reorder_scratch_bytes(sp)
    rscratch = torch.empty(rsb, dtype=torch.uint8, device=dev)
    use_build = not args.no_build
    reorder = not args.no_reorder and not use_build
    bsb = L.hn_build_scratch_bytes_u64(sp, n) if use_build else 0
    bscratch = torch.empty(bsb, dtype=torch.uint8, device=dev) if use_build else None
    torch.cuda.synchronize()
    t_gen = time.perf_counter() - t_gen

    def ev():
        return torch.cuda.Event(enable_timing=True)

    def step(rec=None):
        e = [ev() for _ in range(5)]
        e[0].record(stream)
        if not use_build:  # the bulk build writes every slot: a fresh map needs only its Info
            _lib.check(L.hn_table_init_u64(P(buckets), N, S), "table_init")
        _lib.check(L.hn_info_init(P(info), sp, S), "info_init")
        e[1].record(stream)
        if use_build:
            _lib.check(L.hn_build_u64(P(dk), P(dv), P(buckets), P(info), n, sp, P(bscratch), bsb, S, None), "build")
        else:
            _lib.check(L.hn_insert_u64(P(dk), P(dv), P(buckets), P(info), n, S, tune), "insert")
        e[2].record(stream)
        if reorder:
            _lib.check(L.hn_reorder_u64(P(buckets), P(info), sp, P(rscratch), rsb, S), "reorder")
        e[3].record(stream)
        _lib.check(L.hn_retrieve_u64(P(q), P(out), None, P(buckets), P(info), q.numel(), S, tune), "retrieve")
        e[4].record(stream)
        if rec is not None:
            rec.append(e)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    recs = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(recs)
    torch.cuda.synchronize()
    t_total = time.perf_counter() - t0

    # correctness: every present key's value, no absent key touched
    ok = torch.equal(out, expect)
    del expect
    inf = info.cpu().numpy()
    assert ok and int(inf[1]) == n and (int(inf[4]) & 0xFFFFFFFF) == 0, "result check failed"
    ordered = bool((int(inf[4]) >> 32) & _lib.HN_INFO_ORDERED)
    assert ordered == (reorder or use_build), "order flag"

    t_init = np.mean([a[0].elapsed_time(a[1]) for a in recs])
    t_ins = np.mean([a[1].elapsed_time(a[2]) for a in recs])
    t_ro = np.mean([a[2].elapsed_time(a[3]) for a in recs])
    t_ret = np.mean([a[3].elapsed_time(a[4]) for a in recs])
    ms_step = t_total * 1e3 / args.steps
    ops = 3 * n
    value = ops / (ms_step * 1e-3) / 1e6

    ret_bytes = alg_bytes_retrieve(n, n, n / N, ordered=ordered)
    ins_bytes = alg_bytes_insert(n, n / N)
    achieved = ret_bytes / (t_ret * 1e-3) / 1e9
    # measured for the default (lane) kernel only
    traffic = None if args.probe_width else read_traffic("retrieve_u64_sp28_lf09_rh_traffic.md" if ordered else
                                                         "retrieve_u64_sp28_lf09_traffic.md")
    kname = ("k_lookup_lane<u64> (retrieve, one key per lane" if not args.probe_width else
             f"k_lookup<u64,{args.probe_width}> (retrieve, {args.probe_width} lanes per key")
    roofline = {
        "bound": "hbm",
        "kernel": kname + (", Robin Hood ordered table)" if ordered else ", arrival order)"),
        "achieved": round(achieved, 1),
        "peak": PEAK_HBM_GBS,
        "unit": "GB/s",
        "frac": round(achieved / PEAK_HBM_GBS, 4),
        "traffic": traffic,
        "alg_bytes_per_launch": round(ret_bytes),
        "avg_launch_ms": round(float(t_ret), 3),
    }
    stages = {
        "table_init_ms": round(float(t_init), 3),
        "insert_path": "bulk build (hn_build_u64)" if use_build else "probing insert (hn_insert_u64)",
        "insert_ms": round(float(t_ins), 3),
        "insert_Mops": round(n / t_ins / 1e3, 1),
        "insert_GBps_alg": round(ins_bytes / (t_ins * 1e-3) / 1e9, 1),
        "reorder_ms": round(float(t_ro), 3),
        "reorder_GBps_alg": round(2 * N * 16 / (t_ro * 1e-3) / 1e9, 1) if reorder else None,
        "retrieve_ms": round(float(t_ret), 3),
        "retrieve_Mops": round(2 * n / t_ret / 1e3, 1),
        "overflow": int(inf[2]),
    }

    # Random-access ceiling of this device, measured now on the table's own
    # 4 GiB: independent random reads of 64 B sectors and 128 B lines
    # (hn_calib_random_gather), 4 GiB of reads each, best of 3.
    sink = torch.zeros(1, dtype=torch.int64, device=dev)
    ceil = {}
    for gran in (64, 128):
        best = None
        for rep in range(3):
            e0, e1 = ev(), ev()
            e0.record(stream)
            _lib.check(L.hn_calib_random_gather(P(buckets), N * 16, gran, (N * 16) // gran, 1, 1234 + rep, P(sink), S),
                       "calib")
            e1.record(stream)
            torch.cuda.synchronize()
            t = e0.elapsed_time(e1)
            best = t if best is None else min(best, t)
        ceil[gran] = N * 16 / (best * 1e-3) / 1e9
    roofline["random_ceiling_GBps"] = {"64B": round(ceil[64], 1), "128B": round(ceil[128], 1)}
    # algorithmic (64 B-sector) rate against the random 64 B-sector ceiling; above 1 when the
    # kernel reads whole 128 B lines, which cost one request each, like a 64 B sector
    roofline["frac_of_random_64B_ceiling"] = round(achieved / ceil[64], 4)
    if traffic:
        roofline["traffic_GBps"] = round(traffic / (t_ret * 1e-3) / 1e9, 1)
        roofline["traffic_frac_of_random_128B_ceiling"] = round(traffic / (t_ret * 1e-3) / 1e9 / ceil[128], 4)

    if args.calib_copy:
        src = torch.ones(1 << 27, dtype=torch.int64, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    extra = {}
    if not args.no_extra:
        extra = run_extras(L, torch, dev, S, P, tune, buckets, info, dk, dv, q, out, hit, n, N, sp, make_inputs,
                           rscratch if (reorder or use_build) else None, bscratch)
    if bscratch is not None:
        del bscratch
    if not args.no_extra:
        extra["merge_insert"] = run_merge_insert(torch, L, dev, S, P, dk, dv, n, N, sp, seed)

    if not args.no_compaction:
        extra["compaction"] = run_compaction(torch, L, dev, S, P, buckets, int(info.cpu().numpy()[1]))
    if not args.no_batched:
        del buckets, out, hit, q, dk, dv
        torch.cuda.empty_cache()
        extra["batched"] = run_batched(torch, with_cpu=not args.no_cpu)
    if not args.no_serving:
        torch.cuda.empty_cache()
        extra["serving"] = run_serving(torch)
    if not args.no_extra:
        extra["small_map_u32"] = run_small_map(torch, L, dev, S, P, with_cpu=not args.no_cpu)

    line = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "Mops/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {
            "workload": f"single Hashmap<u64,u64> sp={sp} ({N} slots), LF {alpha}: insert {n} keys + retrieve "
                        f"{n} present + {n} absent keys per step",
            "size_power": sp,
            "load_factor": alpha,
            "keys": "50% uniform u64, 50% runs of 8-512 consecutive ids",
            "probe_width": args.probe_width or "default",
            "insert_path": "build" if use_build else "probe",
            "robin_hood_order": ordered,
        },
        "roofline": roofline,
        "stages": stages,
        "extra": extra,
        "input_gen_s": round(t_gen, 1),
    }
    if not args.no_cpu:
        cb = cpu_baseline()
        bc = extra.get("batched", {}).get("cpu_reference_all_cores")
        if bc:
            cb["stages"]["batched_4096_maps_all_cores"] = bc
        line["cpu_baseline"] = cb
    print(json.dumps(line), flush=True)


def calib_sweep(torch, L, dev, S, P):
    """hn_calib_random_gather over a 4 GiB buffer: 16 GiB of reads per point,
    best of 3, for read sizes 16..256 B and 1, 4 or 16 dependent reads per
    lane group; then hn_calib_random_gather_lane (one lane per read) for
    16..128 B.  GB/s and G reads/s."""
    from hashinator_amd import _lib

    buf = torch.ones(1 << 29, dtype=torch.int64, device=dev)  # 4 GiB
    sink = torch.zeros(1, dtype=torch.int64, device=dev)
    stream = torch.cuda.current_stream()
    out = {"buffer_bytes": 1 << 32, "read_bytes_per_point": 1 << 34, "points": []}
    for gran in (16, 32, 64, 128, 256):
        for chain in (1, 4, 16):
            n_access = (1 << 34) // gran
            best = None
            for rep in range(3):
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                _lib.check(L.hn_calib_random_gather(P(buf), 1 << 32, gran, n_access, chain, 99 + rep, P(sink), S),
                           "calib")
                e1.record(stream)
                torch.cuda.synchronize()
                t = e0.elapsed_time(e1)
                best = t if best is None else min(best, t)
            out["points"].append({"gran": gran, "chain": chain, "GBps": round((1 << 34) / (best * 1e-3) / 1e9, 1),
                                  "Greads_per_s": round(n_access / (best * 1e-3) / 1e9, 2)})
    # one lane reads the whole chunk with gran/16 consecutive 16-byte loads
    for gran in (16, 32, 64, 128):
        n_access = (1 << 34) // gran
        best = None
        for rep in range(3):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            _lib.check(L.hn_calib_random_gather_lane(P(buf), 1 << 32, gran, n_access, 77 + rep, P(sink), S), "calib")
            e1.record(stream)
            torch.cuda.synchronize()
            t = e0.elapsed_time(e1)
            best = t if best is None else min(best, t)
        out["points"].append({"gran": gran, "per_lane": True, "GBps": round((1 << 34) / (best * 1e-3) / 1e9, 1),
                              "Greads_per_s": round(n_access / (best * 1e-3) / 1e9, 2)})
    return out


def run_merge_insert(torch, L, dev, S, P, dk, dv, n, N, sp, seed):
    """A large insert into a half-full map: about 0.45 of the table (120 M
    keys at 2^28 slots) merged into a table holding N/2 keys in Robin Hood
    order (the bulk build's), hn_merge_u64 into a second array, best of 3
    after a warm-up; beside it the probing insert of the same batch into a
    copy of the same table.  Checks: fill, the order flag, and every key of
    the union retrieving its value from the merged table."""
    from hashinator_amd import _lib
    from hashinator_amd import workloads as Wl

    n_old = N // 2
    n_new = min(120_000_000, int(0.447 * N))
    ok_, ov_ = dk[:n_old], dv[:n_old]
    absent = Wl.absent64(n_new, seed, n)  # disjoint from the map's keys
    nk = torch.from_numpy(absent.view(np.int64)).to(dev)
    nv = torch.from_numpy(Wl.values(absent, seed + 1).view(np.int64)).to(dev)
    del absent
    old = torch.empty((N, 2), dtype=torch.int64, device=dev)
    info = torch.zeros(5, dtype=torch.int64, device=dev)
    _lib.check(L.hn_info_init(P(info), sp, S), "info_init")
    bsb = L.hn_build_scratch_bytes_u64(sp, n_old)
    bs = torch.empty(bsb, dtype=torch.uint8, device=dev)
    _lib.check(L.hn_build_u64(P(ok_), P(ov_), P(old), P(info), n_old, sp, P(bs), bsb, S, None), "build")
    torch.cuda.synchronize()
    del bs
    info0 = info.clone()
    nb = torch.empty_like(old)
    msb = L.hn_merge_scratch_bytes_u64(sp, n_new)
    ms = torch.empty(msb, dtype=torch.uint8, device=dev)
    stream = torch.cuda.current_stream()

    def merge():
        info.copy_(info0)
        _lib.check(L.hn_merge_u64(P(nk), P(nv), P(old), P(nb), P(info), n_new, sp, P(ms), msb, S), "merge")

    merge()  # first call: code objects, occupancy queries
    torch.cuda.synchronize()
    best = float("inf")
    for _ in range(3):
        info.copy_(info0)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        _lib.check(L.hn_merge_u64(P(nk), P(nv), P(old), P(nb), P(info), n_new, sp, P(ms), msb, S), "merge")
        e1.record(stream)
        torch.cuda.synchronize()
        best = min(best, e0.elapsed_time(e1))
    inf = info.cpu().numpy()
    fill, ordered = int(inf[1]), bool((int(inf[4]) >> 32) & _lib.HN_INFO_ORDERED)
    qk = torch.cat([ok_, nk])
    qv = torch.cat([ov_, nv])
    out = torch.full_like(qk, MISS_VAL)
    hit = torch.empty(qk.numel(), dtype=torch.uint8, device=dev)
    _lib.check(L.hn_retrieve_u64(P(qk), P(out), P(hit), P(nb), P(info), qk.numel(), S, None), "retrieve")
    torch.cuda.synchronize()
    ok = fill == n_old + n_new and ordered and bool(hit.bool().all()) and bool(torch.equal(out, qv))
    del out, hit, qk, qv, ms
    b2 = old.clone()
    info.copy_(info0)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    _lib.check(L.hn_insert_u64(P(nk), P(nv), P(b2), P(info), n_new, S, None), "insert")
    e1.record(stream)
    torch.cuda.synchronize()
    t_probe = e0.elapsed_time(e1)
    pinf = info.cpu().numpy()
    return {
        "workload": f"insert {n_new} new keys into Hashmap<u64,u64> sp={sp} holding {n_old} keys (Robin Hood order)",
        "merge_ms": round(best, 3), "merge_Gkeys_per_s": round(n_new / best / 1e6, 2),
        "merge_scratch_MiB": round(msb / 2**20, 1), "overflow_after_merge": int(inf[2]),
        "probing_insert_ms": round(t_probe, 3), "probing_Gkeys_per_s": round(n_new / t_probe / 1e6, 2),
        "overflow_after_probing": int(pinf[2]), "speedup": round(t_probe / best, 2),
        "fill_after": fill, "ordered_after": ordered, "ok": ok and int(pinf[1]) == n_old + n_new,
    }


def run_compaction(torch, L, dev, S, P, buckets, fill):
    """split::tools on 2^28 hash_pair<u64,u64> {i, random with bit 0 = keep},
    next to a plain streaming copy of the same 4 GiB (hn_calib_stream_copy):
    stable copy_if with the rule (second & 1) == 1, keep drawn Bernoulli(0.3)
    and in alternating runs of Geometric(1/64) lengths; exclusive scan of the
    2^28 uint32 keep flags; extractPattern (valid keys) over the large map's
    table.  Results are checked against torch on the same device; times are
    HIP events, best of 3."""
    import ctypes as C

    from hashinator_amd import _lib

    n = 1 << 28
    stream = torch.cuda.current_stream()
    g = torch.Generator(device=dev).manual_seed(0x48415348494E4154 & 0x7FFFFFFF)
    pairs = torch.empty((n, 2), dtype=torch.int64, device=dev)
    pairs[:, 0] = torch.arange(n, dtype=torch.int64, device=dev)
    out = torch.empty_like(pairs)
    cnt = torch.zeros(1, dtype=torch.int64, device=dev)
    sb = L.hn_compact_scratch_bytes(n)
    sc = torch.empty(sb, dtype=torch.uint8, device=dev)
    pred = _lib.hn_pred(_lib.HN_PRED_MASKEQ, 1, 1, 1)

    def best(fn):
        t = None
        for _ in range(3):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            fn()
            e1.record(stream)
            torch.cuda.synchronize()
            dt = e0.elapsed_time(e1)
            t = dt if t is None else min(t, dt)
        return t

    def copy_if():
        _lib.check(L.hn_copy_if_pairs_u64(P(pairs), n, C.byref(pred), P(out), P(cnt), P(sc), sb, S), "copy_if")

    # the streaming ceiling of this device: a plain 16-byte-vector copy of the
    # 4 GiB pair array (read + write)
    t_copy = best(lambda: _lib.check(L.hn_calib_stream_copy(P(pairs), P(out), n * 16, S), "stream copy"))
    copy_rate = 2 * n * 16 / (t_copy * 1e-3)
    res = {"stream_copy": {"ms": round(t_copy, 3), "bytes": 2 * n * 16, "GBps": round(copy_rate / 1e9, 1),
                           "frac_of_8TBps": round(copy_rate / 8e12, 4)}}
    rng = np.random.default_rng(64)
    for mode in ("bernoulli", "clustered"):
        if mode == "bernoulli":
            keep = torch.rand(n, generator=g, device=dev) < 0.3
        else:
            runs = rng.geometric(1 / 64, size=(n // 32) + 16)
            runs = runs[: np.searchsorted(np.cumsum(runs), n) + 1]
            bit = (np.arange(runs.size) & 1).astype(np.uint8)
            keep = torch.from_numpy(np.repeat(bit, runs)[:n].astype(bool)).to(dev)
        rnd = torch.randint(-(2**62), 2**62, (n,), generator=g, device=dev)
        pairs[:, 1] = (rnd & ~1) | keep.to(torch.int64)
        t = best(copy_if)
        k = int(keep.sum().item())
        ok = int(cnt.item()) == k
        # chunked check (torch's own boolean indexing is not reliable on a 4 GiB tensor)
        base = 0
        for c0 in range(0, n, 1 << 24):
            kc = keep[c0:c0 + (1 << 24)]
            m = int(kc.sum().item())
            ok = ok and torch.equal(out[base:base + m], pairs[c0:c0 + (1 << 24)][kc])
            base += m
        alg = n * 16 + k * 16
        res[f"copy_if_{mode}"] = {"ms": round(t, 3), "kept": k, "GBps_alg": round(alg / (t * 1e-3) / 1e9, 1),
                                  "frac_of_8TBps": round(alg / (t * 1e-3) / 8e12, 4),
                                  "frac_of_stream_copy": round(alg / (t * 1e-3) / copy_rate, 4), "ok": bool(ok)}
    flags = keep.to(torch.int32)
    sout = torch.empty_like(flags)
    t = best(lambda: _lib.check(L.hn_exclusive_scan_u32(P(flags), P(sout), n, P(sc), sb, S), "scan"))
    ok = True
    run = 0
    for c0 in range(0, n, 1 << 24):
        fc = flags[c0:c0 + (1 << 24)].to(torch.int64)
        ref = torch.cumsum(fc, 0) - fc + run
        ok = ok and torch.equal(sout[c0:c0 + (1 << 24)].to(torch.int64), ref)
        run += int(fc.sum().item())
    res["exclusive_scan_u32"] = {"ms": round(t, 3), "GBps_alg": round(8 * n / (t * 1e-3) / 1e9, 1),
                                 "frac_of_8TBps": round(8 * n / (t * 1e-3) / 8e12, 4),
                                 "frac_of_stream_copy": round(8 * n / (t * 1e-3) / copy_rate, 4), "ok": bool(ok)}
    del flags, sout, keep, pairs
    N = buckets.shape[0]
    vpred = _lib.hn_pred(_lib.HN_PRED_VALID, 0, 0, 0)
    t = best(lambda: _lib.check(L.hn_copy_if_pairs_u64(P(buckets), N, C.byref(vpred), P(out), P(cnt), P(sc), sb, S),
                                "extractPattern"))
    alg = N * 16 + fill * 16
    res["extractPattern_valid"] = {"ms": round(t, 3), "slots": N, "extracted": int(cnt.item()),
                                   "GBps_alg": round(alg / (t * 1e-3) / 1e9, 1), "ok": int(cnt.item()) == fill}
    del out, sc
    torch.cuda.empty_cache()
    res["dropin_copy_if"] = run_dropin_compaction()
    return res


def run_dropin_compaction(log2n=28, p=0.3):
    """The same Bernoulli(0.3) compaction of 2^28 hash_pair<u64,u64> through
    the drop-in C++ API: split::tools::copy_if(SplitVector&, SplitVector&,
    rule functor) on device-resident vectors (split_device_allocator), timed
    as a whole call, next to the raw C-ABI call on the same input
    (tests/cpp/dropin_compaction.cpp, built by __graft_entry__.build())."""
    import json as _json
    import subprocess

    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "cpp", "_bin", "dropin_compaction")
    if not os.path.exists(exe):
        return {"error": "tests/cpp/_bin/dropin_compaction not built"}
    r = subprocess.run([exe, str(log2n), str(p)], capture_output=True, text=True, timeout=120)
    try:
        out = _json.loads(r.stdout.strip().splitlines()[-1])
    except (ValueError, IndexError):
        return {"error": f"rc {r.returncode}: {r.stderr[-300:]}"}
    out["ok"] = r.returncode == 0 and bool(out.get("same_as_abi"))
    return out


def run_small_map(torch, L, dev, S, P, with_cpu=True, reps=20):
    """The small single map: Hashmap<u32,u32> of 2^20 slots, insert 524,288
    distinct uniform keys (uniform values), then retrieve them plus 524,288
    absent keys.  The insert takes the maps' policy for a batch of half the
    table into an empty map (the bulk build).  The 8 MiB table sits in the
    caches: no HBM claim.  GPU: best of `reps` HIP-event-timed runs; beside it
    the reference's CPU-only build (oracle/_ref, else the C restatement) on
    the same inputs, one core."""
    from hashinator_amd import _lib
    from hashinator_amd import workloads as Wl

    sp, n = 20, 524_288
    seed = Wl.SEED_BASE + 1
    keys = Wl.unique_keys(n, 32, seed)
    vals = Wl.values(keys, seed)
    absent = Wl.unique_keys(n, 32, seed, start=n + 2)
    q = Wl.shuffled(np.concatenate([keys, absent]), seed)
    dk = torch.from_numpy(keys.view(np.int32)).to(dev)
    dv = torch.from_numpy(vals.view(np.int32)).to(dev)
    dq = torch.from_numpy(q.view(np.int32)).to(dev)
    out = torch.empty_like(dq)
    hit = torch.empty(dq.numel(), dtype=torch.uint8, device=dev)
    b = torch.empty((1 << sp, 2), dtype=torch.int32, device=dev)
    info = torch.zeros(5, dtype=torch.int64, device=dev)
    bsb = L.hn_build_scratch_bytes_u32(sp, n)
    bs = torch.empty(bsb, dtype=torch.uint8, device=dev)
    stream = torch.cuda.current_stream()
    t_ins, t_ret = [], []
    for _ in range(reps):
        e = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
        e[0].record(stream)
        _lib.check(L.hn_info_init(P(info), sp, S), "info_init")
        _lib.check(L.hn_build_u32(P(dk), P(dv), P(b), P(info), n, sp, P(bs), bsb, S, None), "build")
        e[1].record(stream)
        _lib.check(L.hn_retrieve_u32(P(dq), P(out), P(hit), P(b), P(info), dq.numel(), S, None), "retrieve")
        e[2].record(stream)
        torch.cuda.synchronize()
        t_ins.append(e[0].elapsed_time(e[1]))
        t_ret.append(e[1].elapsed_time(e[2]))
    h = hit.cpu().numpy().astype(bool)
    got = out.cpu().numpy().view(np.uint32)
    exp_hit = np.isin(q, keys)
    order = np.argsort(keys)
    exp_val = vals[order][np.searchsorted(keys[order], q[exp_hit])]
    ok = bool(np.array_equal(h, exp_hit) and np.array_equal(got[h], exp_val))
    ti, tr = min(t_ins), min(t_ret)
    res = {"workload": f"Hashmap<u32,u32>(sp={sp}): insert {n} keys, retrieve {n} present + {n} absent; "
                       f"8 MiB table, cache-resident", "insert_ms": round(ti, 4), "retrieve_ms": round(tr, 4),
           "insert_retrieve_Mops": round(3 * n / (ti + tr) / 1e3, 1), "ok": ok}
    if with_cpu:
        import oracle as O

        kind = "reference" if O.ref_available() else "port"
        m = O.ref_map(sp, 32) if kind == "reference" else O.CpuMap(sp, 32)
        t0 = time.perf_counter()
        m.insert(keys, vals)
        t1 = time.perf_counter()
        cv, ch = m.find(q)
        t2 = time.perf_counter()
        assert np.array_equal(ch, exp_hit) and np.array_equal(cv[ch], exp_val)
        res["cpu"] = {"kind": kind, "cores": 1, "insert_ms": round((t1 - t0) * 1e3, 2),
                      "retrieve_ms": round((t2 - t1) * 1e3, 2),
                      "insert_retrieve_Mops": round(3 * n / (t2 - t0) / 1e6, 2)}
    return res


def batched_only(torch):
    """The batched call alone on the many-maps workload (see run_batched):
    fresh tables, then insert -> retrieve -> erase -> clean in one call, 3
    times; HIP-event time of each call."""
    from hashinator_amd import _lib, hn_tune
    from hashinator_amd import workloads as Wl
    from hashinator_amd.batch import BatchedMaps

    bm = BatchedMaps(*Wl.batched_workload(), key_bits=32)
    ops = _lib.HN_BATCH_INSERT | _lib.HN_BATCH_RETRIEVE | _lib.HN_BATCH_ERASE | _lib.HN_BATCH_CLEAN
    tune = hn_tune(8, 1024, 0, 0)
    stream = torch.cuda.current_stream()
    times = []
    for _ in range(3):
        bm.init()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        bm.run(ops=ops, tune=tune)
        e1.record(stream)
        torch.cuda.synchronize()
        times.append(round(e0.elapsed_time(e1), 3))
    return {"workload": "4096 Hashmap<u32,u32>, sizes 2^8..2^16, LF 0.6: insert + retrieve + erase 20% + clean",
            "tune": {"probe_width": 8, "block_size": 1024}, "batched_ms": times}


def run_batched(torch, with_cpu=False):
    """Many uneven maps: 4096 Hashmap<u32,u32> of 2^8..2^16 slots at LF 0.6
    (velocity-grid keys): insert -> retrieve (keys + as many absent cells)
    -> erase 20% -> clean_tombstones, batched in one persistent launch (eager
    and replayed from a HIP graph) against the per-map calls."""
    from hashinator_amd import _lib
    from hashinator_amd import workloads as Wl
    from hashinator_amd.autotune import tune_batch
    from hashinator_amd.batch import BatchedMaps

    t0 = time.perf_counter()
    wl = Wl.batched_workload()
    bm = BatchedMaps(*wl, key_bits=32)
    cpu = cpu_batched(wl) if with_cpu else None
    gen_s = time.perf_counter() - t0
    n_ins, n_qry, n_ers = bm.counts()
    ops = n_ins + n_qry + n_ers
    stream = torch.cuda.current_stream()

    def timed(fn, reps=3):
        best = float("inf")
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            fn()
            e1.record(stream)
            torch.cuda.synchronize()
            best = min(best, e0.elapsed_time(e1))
        return best

    tune, table = tune_batch(bm, candidates=[(w, b) for w in (4, 8, 16) for b in (256, 512, 1024)])

    def per_map():
        bm.init_per_map()
        bm.run_per_map()

    t_per = timed(per_map, reps=2)
    ok_per = _batch_ok(bm, n_ins, n_ers)

    def batched():
        bm.init()
        bm.run(tune=tune)

    t_bat = timed(batched)
    ok_bat = _batch_ok(bm, n_ins, n_ers)
    g = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(stream)
    with torch.cuda.stream(side):
        batched()
    stream.wait_stream(side)
    torch.cuda.synchronize()
    with torch.cuda.graph(g):
        batched()
    t_graph = timed(g.replay)
    ok_graph = _batch_ok(bm, n_ins, n_ers)
    return {
        "workload": f"{bm.n} Hashmap<u32,u32>, sizes 2^8..2^16 ({bm.total_slots} slots), LF 0.6, velocity-grid keys: "
                    f"insert {n_ins} + retrieve {n_qry} + erase {n_ers} + clean",
        "per_map_ms": round(t_per, 3), "batched_ms": round(t_bat, 3), "batched_graph_ms": round(t_graph, 3),
        "per_map_Mops": round(ops / t_per / 1e3, 1), "batched_Mops": round(ops / t_bat / 1e3, 1),
        "batched_graph_Mops": round(ops / t_graph / 1e3, 1),
        "speedup_vs_per_map": round(t_per / min(t_bat, t_graph), 2),
        "tune": {"probe_width": tune.probe_width, "block_size": tune.block_size},
        "tune_table_ms": {k: round(v, 3) for k, v in table.items()},
        "ok": bool(ok_per and ok_bat and ok_graph), "input_gen_s": round(gen_s, 1),
        "cpu_reference_all_cores": cpu,
    }


def run_serving(torch, rates=(1000.0, 4000.0), n_requests=300):
    """The serving loop: 64 resident Hashmap<u64,u64> of 2^20..2^26 slots
    preloaded to LF 0.4; a Poisson stream of insert/retrieve/erase batches
    (45/45/10%, log-uniform 1k..4M keys, 30% repeats of live keys) from host
    memory; growth at LF 0.85; admission by a 64 GiB in-flight budget.  Each
    rate runs twice: inserts by probing and growth by reinsertion, then
    inserts of at least 2^20 keys and a quarter of their table by the
    reorder + merge and growth of maps of 2^20+ keys by the bulk rehash."""
    from hashinator_amd import workloads as Wl
    from hashinator_amd.serving import ServingEngine, latency_report, make_requests

    rng = np.random.default_rng(5)
    sps = [int(x) for x in rng.integers(20, 27, size=64)]
    out = {"maps": "64 Hashmap<u64,u64>, 2^20..2^26 slots, preloaded to LF 0.4", "runs": []}
    t0 = time.perf_counter()
    pre = []
    for m, sp in enumerate(sps):
        k = Wl.unique_keys(int(0.4 * (1 << sp)), 64, 7000 + m)
        pre.append(k)
    gen = time.perf_counter() - t0
    for rate, big in [(r, b) for r in rates for b in (False, True)]:
        eng = ServingEngine(sps, key_bits=64, n_streams=8, hbm_budget_bytes=64 << 30,
                            build_rehash_min=(1 << 20) if big else None, merge_min_keys=(1 << 20) if big else None)
        for m, k in enumerate(pre):
            eng.preload(m, k, Wl.values(k, 1))
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        reqs = make_requests(sps, n_requests, rate_hz=rate, seed=int(rate), live_keys=[[k] for k in pre])
        gen += time.perf_counter() - t1
        wall = eng.serve(reqs, realtime=True)
        rep = latency_report(reqs, wall)
        rep.update({"rate_hz": rate, "bulk_paths": big, "rehashes": eng.rehashes, "merge_inserts": eng.merges,
                    "launches": eng.launches, "merged_requests": n_requests - eng.launches})
        # every retrieve and every final fill against a sequential host replay (not timed)
        t2 = time.perf_counter()
        infos = eng.infos()
        replay = Wl.ServingReplay([(k.size, 7000 + m, 1) for m, k in enumerate(pre)])
        bad, checked = replay.check(reqs, infos[:, 1])
        bad += [f"map {m}: err {int(e)}" for m, e in enumerate(infos[:, 4] & 0xFFFFFFFF) if e]
        rep.update({"checked": not bad, "checked_keys": checked, "check_s": round(time.perf_counter() - t2, 1)})
        if bad:
            rep["mismatches"] = bad[:5]
        out["runs"].append(rep)
        del eng
        torch.cuda.empty_cache()
    out["input_gen_s"] = round(gen, 1)
    return out


def _batch_ok(bm, n_ins, n_ers):
    inf = bm.all_infos()
    hits = int(bm.qry_hit.sum().item())
    return (hits == n_ins and int(inf[:, 1].sum()) == n_ins - n_ers and int(inf[:, 3].sum()) == 0
            and not np.any(inf[:, 4] & 0xFFFFFFFF))


def run_extras(L, torch, dev, S, P, tune, buckets, info, dk, dv, q, out, hit, n, N, sp, make_inputs, rscratch,
               bscratch):
    """Rest of the large-map sequence on the map left by the last step, then
    the other load factors (one step each), then (with the bulk build as the
    headline path) the probing insert + reorder on the same keys."""
    from hashinator_amd import _lib

    extra = {}
    stream = torch.cuda.current_stream()

    def timed(fn):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1)

    g = torch.Generator(device=dev).manual_seed(12345)
    er = dk[torch.randperm(n, device=dev, generator=g)[: n // 4]].contiguous()
    t_er = timed(lambda: _lib.check(L.hn_erase_u64(P(er), P(buckets), P(info), er.numel(), S, tune), "erase"))
    sb = L.hn_clean_tombstones_scratch_bytes_u64(sp, 0)
    scratch = torch.empty(sb, dtype=torch.uint8, device=dev)
    # the clean's first call (code-object load, occupancy queries) on a small
    # empty table, outside the timed region
    wsp = 12
    wb = torch.full((1 << wsp, 2), -1, dtype=torch.int64, device=dev)
    wi = torch.zeros(5, dtype=torch.int64, device=dev)
    _lib.check(L.hn_info_init(P(wi), wsp, S), "info_init")
    _lib.check(L.hn_clean_tombstones_u64(P(wb), P(wi), wsp, P(scratch), sb, S), "clean warm-up")
    torch.cuda.synchronize()
    t_cl = timed(lambda: _lib.check(L.hn_clean_tombstones_u64(P(buckets), P(info), sp, P(scratch), sb, S), "clean"))

    # the map's policy: reorder after clean unless the clean (whole-table
    # scratch: a rebuild) already left the table in Robin Hood order
    ordered_after_clean = bool((int(info.cpu().numpy()[4]) >> 32) & _lib.HN_INFO_ORDERED)

    def ro():
        if rscratch is not None and not ordered_after_clean:
            _lib.check(L.hn_reorder_u64(P(buckets), P(info), sp, P(rscratch), rscratch.numel(), S), "reorder")

    t_ro = timed(ro)
    t_ra = timed(lambda: _lib.check(L.hn_retrieve_u64(P(q), P(out), P(hit), P(buckets), P(info), q.numel(), S, tune),
                                    "retrieve"))
    inf = info.cpu().numpy()
    n_hit = int(hit.sum().item())
    extra["erase25_clean_retrieve"] = {
        "erase_ms": round(t_er, 3), "erase_Mops": round(er.numel() / t_er / 1e3, 1),
        "clean_ms": round(t_cl, 3), "clean_scratch_MiB": round(sb / 2**20, 1),
        "ordered_after_clean": ordered_after_clean, "reorder_ms": round(t_ro, 3),
        "retrieve_after_ms": round(t_ra, 3), "retrieve_after_Mops": round(q.numel() / t_ra / 1e3, 1),
        "hits_after": n_hit, "expected_hits": n - n // 4, "tombstones_after": int(inf[3]),
        "ok": n_hit == n - n // 4 and int(inf[3]) == 0 and int(inf[1]) == n - n // 4,
    }
    del scratch
    for a in (0.5, 0.75):
        n2, dk2, dv2, q2 = make_inputs(a)
        out2 = torch.empty_like(q2)
        hit2 = torch.empty(q2.numel(), dtype=torch.uint8, device=dev)

        def one():
            L.hn_table_init_u64(P(buckets), N, S)
            L.hn_info_init(P(info), sp, S)

        timed(one)
        if bscratch is not None:
            ti = timed(lambda: _lib.check(L.hn_build_u64(P(dk2), P(dv2), P(buckets), P(info), n2, sp, P(bscratch),
                                                         bscratch.numel(), S, None), "build"))
        else:
            ti = timed(lambda: _lib.check(L.hn_insert_u64(P(dk2), P(dv2), P(buckets), P(info), n2, S, tune),
                                          "insert"))
            ti += timed(lambda: _lib.check(L.hn_reorder_u64(P(buckets), P(info), sp, P(rscratch), rscratch.numel(),
                                                            S), "reorder") if rscratch is not None else None)
        tr = timed(lambda: _lib.check(L.hn_retrieve_u64(P(q2), P(out2), P(hit2), P(buckets), P(info), q2.numel(),
                                                        S, tune), "retrieve"))
        extra[f"lf{a}"] = {"n": n2, "insert_Mops": round(n2 / ti / 1e3, 1),
                           "retrieve_Mops": round(q2.numel() / tr / 1e3, 1),
                           "insert_retrieve_Mops": round(3 * n2 / (ti + tr) / 1e3, 1),
                           "hits_ok": int(hit2.sum().item()) == n2}
        del dk2, dv2, q2, out2, hit2
    if bscratch is not None:
        timed(lambda: (L.hn_table_init_u64(P(buckets), N, S), L.hn_info_init(P(info), sp, S)))
        ti = timed(lambda: _lib.check(L.hn_insert_u64(P(dk), P(dv), P(buckets), P(info), n, S, tune), "insert"))
        tr = timed(lambda: _lib.check(L.hn_reorder_u64(P(buckets), P(info), sp, P(rscratch), rscratch.numel(), S),
                                      "reorder"))
        inf = info.cpu().numpy()
        extra["probing_insert_path"] = {
            "insert_ms": round(ti, 3), "reorder_ms": round(tr, 3), "insert_Mops": round(n / ti / 1e3, 1),
            "ok": int(inf[1]) == n and (int(inf[4]) & 0xFFFFFFFF) == 0,
        }
    return extra


if __name__ == "__main__":
    main()
