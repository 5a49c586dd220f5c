// hn_build.hip -- bulk insert into an empty table, written directly in Robin
// Hood order.
//
// An insert of n keys into an empty linear-probing table (kernels_AMD.h
// insert_kernel on a fresh map) followed by the reorder pass gives one image
// that depends on the key set alone: in every cluster the keys stand in order
// of home slot, ties by key.  That image is also what a max-plus scan gives
// over the keys sorted by (home, key):  position_j = max(home_j,
// position_{j-1} + 1), taken cyclically.  This file computes it without
// probing:
//
//   1. stable radix partition of the (key, value) pairs by region
//      (region = home >> log2 R, R = 2048 slots): per pass a per-tile digit
//      count, an exclusive scan of the digit-major count matrix, and a stable
//      multisplit scatter through LDS (digits of at most 9 bits, 1-3 passes);
//   2. region boundaries in the partitioned array (a binary search per
//      region);
//   3. per region, in LDS: sort by (home, key), drop all but the last
//      occurrence of a repeated key (the value a serial insert leaves), and
//      summarise the region as the carry function g(c) = max(a, c + b), where
//      c is the number of keys spilling into the region from before it;
//   4. a workgroup walks 16 consecutive regions, handing the carry from one
//      to the next (carry out = max(a, c + b); the walk starts from the
//      region before it, taken with no carry in), places each region's keys
//      with a max-scan and writes its slot range (keys and EMPTY slots)
//      once, coalesced;
//   5. one scan composes the carry functions (g2 after g1 is
//      (max(a2, a1 + b2), b1 + b2)) into the true carries (the carry into
//      region 0 is the least fixed point of the whole-table composition,
//      max(A, c + B) with B = n - N < 0, i.e. c0 = A), and the regions whose
//      carry was wrong, or whose range wraps past the table's end, are
//      placed again.
//
// Traffic per key: the first partition pass reads the key for its count,
// each pass reads and writes the 16 (u64) or 8 (u32) byte pair, and a pass
// whose successor's digit fits a byte also writes that digit at the pair's
// output position, so the successor counts from one byte per pair instead of
// the pair; the region step reads the pair once more and the table is
// written once.  No atomics on the table, no probing.
//
// The table's memory is the output of the first partition pass when there
// are two, so the scratch holds one pair array, not two.  A region holding
// more than kRegionCap keys (only with strongly skewed keys) abandons the
// placement on the device: the table is emptied and the probing insert
// places the partitioned pairs (k_build_fallback_prep, then k_insert), so the
// call still ends with err = SUCCESS (or FAIL if a key found no slot), the
// same live set in arrival order, and HN_INFO_ORDERED not set.  A table that
// is not empty (fill or tombstones nonzero) is answered with err = INVALID
// and left untouched.
#include <algorithm>
#include <atomic>

#include "hn_device.hpp"
#include "hn_host.hpp"
#include "hn_probe.hpp"

namespace hn {

constexpr int kRegionLog = 11;   // 2048 slots per region
constexpr int kPartBS = 256;     // count workgroup
constexpr int kPartBSS = 256, kPartRoundsS = 16;  // 4,096-pair scatter tiles: threads, pairs per thread
constexpr int kPartBSL = 1024, kPartRoundsL = 8;  // 8,192-pair scatter tiles
// Scatter tiles: 8,192 pairs (512 threads, 16-pair runs per 9-bit digit,
// one block per CU) for inputs of at least kPartLargeMin pairs, 4,096 (256
// threads, two blocks per CU) below, where the large tile would leave CUs
// idle.  At 2^28 slots the large tile builds in 9.19 ms against 9.64 ms:
// longer runs per digit, fewer partial lines; a 2,048-pair tile with three
// blocks per CU took 10.7 ms.
constexpr uint32_t kPartTileS = 4096;
constexpr uint32_t kPartTileL = 8192;
constexpr uint64_t kPartLargeMin = 1ull << 22;
constexpr int kPartSub = 8;      // tiles per partition block at most (fewer for small inputs)
constexpr int kMaxDigitBits = 9; // 512 digits per pass at most
constexpr int kRegionBS = 512;
constexpr int kRegionIPT = 5;
constexpr int kRegionCap = kRegionBS * kRegionIPT; // keys per region held in LDS

struct BuildHeader {
   uint32_t go;        // the table was empty: the build runs
   uint32_t fail;      // a region over capacity: the build is abandoned
   uint32_t max_probe; // longest probe (displacement + 1) of the placed keys
   uint32_t pad0;
   int64_t c0;         // carry into region 0
   uint64_t n_distinct;
   uint32_t n_fix;     // regions placed again after the carries are known
   int32_t fb_fail;    // the fallback insert found no slot for a key
   uint64_t n;         // pairs after the first partition pass (the live slots when the input is a table)
   uint64_t fb_len;    // pairs the fallback insert takes from the partitioned array (0: none)
   uint64_t prior_fill; // rehash: the old table's fill, which the new one must hold
   uint32_t pad[48];
};
static_assert(sizeof(BuildHeader) == 256, "header is one 256-byte line");

struct RegionSum {
   int64_t a, b; // carry out = max(a, carry in + b)
};

struct BuildPlan {
   int sp;
   uint64_t n;          // input pairs
   uint64_t N;          // slots
   int rlog;            // log2 slots per region
   uint64_t R;          // slots per region
   uint64_t NR;         // regions
   int rbits;           // log2 NR
   int passes;          // partition passes
   int dbits[3];        // digit bits per pass (low digit first)
   uint32_t nblocks;    // partition blocks
   uint32_t tile;       // pairs per scatter tile (kPartTileS or kPartTileL)
   uint32_t sub;        // tiles per partition block
   size_t off_pairs, off_mat, off_scan, off_start, off_word, off_spec, off_carry, off_mark, off_list, off_blk, off_dig,
       total;
};

inline size_t align256(size_t x) { return (x + 255) / 256 * 256; }

// The scratch layouts place their arrays at 256-byte offsets from the start
// and read them with 16-byte loads, so the start must be 256-byte aligned
// (hipMalloc's and hipMallocAsync's are).
inline bool scratch_aligned(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 255u) == 0; }

// dcap: widest partition digit (1..9 bits); narrower digits mean more passes.
template <typename K>
bool build_plan(int sp, uint64_t n, int dcap, BuildPlan* P) {
   if (dcap < 1 || dcap > kMaxDigitBits) return false;
   if (sp < 1 || sp > std::min(max_size_power<K>(), 32)) return false;
   P->sp = sp;
   P->n = n;
   P->N = 1ull << sp;
   if (n == 0 || n >= P->N) return false;
   P->rlog = std::min(sp, kRegionLog);
   P->R = 1ull << P->rlog;
   P->rbits = sp - P->rlog;
   P->NR = 1ull << P->rbits;
   P->passes = std::max(1, (P->rbits + dcap - 1) / dcap);
   if (P->passes > 3) return false;
   int left = P->rbits;
   for (int i = 0; i < P->passes; ++i) {
      P->dbits[i] = left / (P->passes - i) + ((left % (P->passes - i)) ? 1 : 0);
      left -= P->dbits[i];
   }
   // Tiles per block: kPartSub, or as few as keep the blocks within one
   // resident round (2 small-tile or 1 large-tile scatter blocks per CU):
   // more blocks than that would leave a partial second round.
   P->tile = n >= kPartLargeMin ? kPartTileL : kPartTileS;
   const uint64_t tiles = (n + P->tile - 1) / P->tile;
   const uint64_t slots = static_cast<uint64_t>(cu_count()) * (P->tile == kPartTileL ? 1 : 2);
   const uint64_t per = (tiles + slots - 1) / slots;
   P->sub = static_cast<uint32_t>(per < 1 ? 1 : (per > kPartSub ? kPartSub : per));
   const uint64_t super = static_cast<uint64_t>(P->tile) * P->sub;
   const uint64_t nb = (n + super - 1) / super;
   if (nb > 0xFFFFFFFFull) return false;
   P->nblocks = static_cast<uint32_t>(nb);
   // Sized for the default digit width, so any narrower one fits the same scratch.
   const uint64_t mat = (1ull << kMaxDigitBits) * nb;
   size_t o = sizeof(BuildHeader);
   P->off_pairs = o;
   o = align256(o + n * sizeof(Slot<K>));
   P->off_mat = o;
   o = align256(o + mat * sizeof(uint32_t));
   P->off_scan = o;
   o = align256(o + hn_compact_scratch_bytes(mat));
   P->off_start = o;
   o = align256(o + (P->NR + 1) * sizeof(uint64_t));
   P->off_word = o;
   o = align256(o + P->NR * sizeof(uint64_t));
   P->off_spec = o;
   o = align256(o + P->NR * sizeof(int64_t));
   P->off_carry = o;
   o = align256(o + P->NR * sizeof(int64_t));
   P->off_mark = o;
   o = align256(o + P->NR * sizeof(uint32_t));
   P->off_list = o;
   o = align256(o + P->NR * sizeof(uint32_t));
   P->off_blk = o;
   o = align256(o + ((P->NR + 1023) / 1024) * sizeof(RegionSum));
   // One byte per pair: the next pass's digit, written by a scatter pass
   // whose successor's digit fits a byte (reserved for every plan of more
   // than one region, so the size does not depend on the digit width).
   P->off_dig = o;
   if (P->NR > 1) o = align256(o + n);
   P->total = o;
   return true;
}

// Input of a partition pass: the caller's arrays (pass 0) or a pair array.
template <typename K>
struct PartIn {
   const K* keys;        // SRC_SOA / SRC_INDEX
   const K* vals;        // SRC_SOA
   const Slot<K>* pairs; // SRC_PAIRS / SRC_TABLE (a pair array)
};

template <typename K, int SRC>
__device__ __forceinline__ K part_key(const PartIn<K>& in, uint64_t e) {
   if constexpr (SRC == SRC_SOA || SRC == SRC_INDEX)
      return in.keys[e];
   else
      return in.pairs[e].key;
}

template <typename K, int SRC>
__device__ __forceinline__ Slot<K> part_pair(const PartIn<K>& in, uint64_t e) {
   Slot<K> s;
   if constexpr (SRC == SRC_SOA) {
      s.key = in.keys[e];
      s.val = in.vals[e];
   } else if constexpr (SRC == SRC_INDEX) {
      s.key = in.keys[e];
      s.val = static_cast<K>(e);
   } else {
      s = in.pairs[e];
   }
   return s;
}

template <typename K>
__device__ __forceinline__ uint32_t region_of(K key, int sp, int rlog) {
   const uint64_t mask = (1ull << sp) - 1;
   return static_cast<uint32_t>((fib_home(key, sp) & mask) >> rlog);
}

// Build from caller arrays into an empty table, (rehash) from the live
// slots of another, smaller table into a new one, or (merge) from caller
// arrays and an ordered table into a second array of the table's size.
enum BuildMode { BUILD_EMPTY = 0, BUILD_REHASH = 2, BUILD_MERGE = 3 };

__global__ void k_build_begin(const hn_info* info, BuildHeader* h, int sp, uint64_t len, int mode) {
   if (threadIdx.x == 0) {
      const bool sp_ok = info->size_power == sp;
      if (mode == BUILD_EMPTY)
         h->go = (sp_ok && info->fill == 0 && info->tombstone_counter == 0) ? 1u : 0u;
      else if (mode == BUILD_MERGE)
         h->go = sp_ok ? 1u : 0u; // the batch is partitioned whether the merge or its fallback runs
      else
         h->go = 1u; // k_rehash_info has just set the Info up for the new table
      h->fail = 0;
      h->max_probe = 0;
      h->c0 = 0;
      h->n_distinct = 0;
      h->fb_fail = 0;
      h->fb_len = 0;
      h->n = mode == BUILD_REHASH ? 0 : len; // a rehash counts the live slots in its first pass
   }
}

template <typename K>
__device__ __forceinline__ bool live_key(K k) {
   return k != empty_key<K>() && k != tomb_key<K>();
}

// Per-block digit counts, written digit-major: mat[d * nblocks + block].  A
// block covers `sub` consecutive tiles of `tile` pairs, which keeps the
// count matrix -- and its scan -- small.
// DIG: the digits come from the byte array the previous scatter pass wrote
// (one byte per pair, 16 per load) instead of from the pairs' keys.
template <typename K, int SRC, bool FILTER, bool DIG = false>
__global__ __launch_bounds__(kPartBS) void k_part_count(PartIn<K> in, uint64_t n_in, const uint64_t* d_n, uint32_t* mat,
                                                        uint32_t nblocks, uint32_t tile, uint32_t nsub, int sp,
                                                        int rlog, int dshift, int dbits, BuildHeader* h,
                                                        const uint8_t* dig) {
   __shared__ uint32_t hist[1 << kMaxDigitBits];
   __shared__ uint32_t live_s;
   if (!h->go) return;
   const uint64_t n = d_n ? *d_n : n_in;
   const uint32_t D = 1u << dbits, dmask = D - 1;
   for (uint32_t d = threadIdx.x; d < D; d += kPartBS) hist[d] = 0;
   if (threadIdx.x == 0) live_s = 0;
   __syncthreads();
   const uint64_t super = static_cast<uint64_t>(tile) * nsub;
   const uint64_t t0 = static_cast<uint64_t>(blockIdx.x) * super;
   const uint64_t t1 = std::min<uint64_t>(t0 + super, n);
   uint32_t live = 0;
   if constexpr (DIG) {
      static_assert(!FILTER, "digit bytes follow a pass that kept only live pairs");
      // t0 is a multiple of the block's span, so the 16-byte loads are aligned.
      for (uint64_t e = t0 + 16ull * threadIdx.x; e < t1; e += 16ull * kPartBS) {
         if (e + 16 <= t1) {
            const uint4 w = *reinterpret_cast<const uint4*>(dig + e);
            const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
            for (int q = 0; q < 4; ++q)
#pragma unroll
               for (int b = 0; b < 4; ++b) atomicAdd(&hist[(ws[q] >> (8 * b)) & 0xFFu], 1u);
         } else {
            for (uint64_t x = e; x < t1; ++x) atomicAdd(&hist[dig[x]], 1u);
         }
      }
   } else {
      // Four keys loaded before their four counts: more loads in flight per thread.
      constexpr int U = 4;
      for (uint64_t e0 = t0 + threadIdx.x; e0 < t1; e0 += U * kPartBS) {
         K key[U];
#pragma unroll
         for (int u = 0; u < U; ++u) {
            const uint64_t e = e0 + static_cast<uint64_t>(u) * kPartBS;
            if (e < t1) key[u] = part_key<K, SRC>(in, e);
         }
#pragma unroll
         for (int u = 0; u < U; ++u) {
            const uint64_t e = e0 + static_cast<uint64_t>(u) * kPartBS;
            if (e >= t1 || (FILTER && !live_key(key[u]))) continue;
            atomicAdd(&hist[(region_of(key[u], sp, rlog) >> dshift) & dmask], 1u);
            ++live;
         }
      }
   }
   if (FILTER) {
      live = wave_sum(live);
      if (__lane_id() == 0 && live) atomicAdd(&live_s, live);
   }
   __syncthreads();
   for (uint32_t d = threadIdx.x; d < D; d += kPartBS) mat[static_cast<uint64_t>(d) * nblocks + blockIdx.x] = hist[d];
   if (FILTER && threadIdx.x == 0 && live_s) atomic_add_relaxed(&h->n, static_cast<uint64_t>(live_s));
}

// Stable multisplit.  A block walks its `nsub` tiles in order, keeping the
// running output offset of every digit in LDS.  In a tile, wave w ranks the
// pairs [64*ROUNDS*w, 64*ROUNDS*(w+1)) in ROUNDS rounds of 64: lanes with equal
// digits are found with one ballot per digit bit, and the wave's own LDS
// counters carry the earlier rounds (no block barrier per round).  Per-wave
// counts then give each pair's rank among equal digits in tile order; the
// tile is staged in LDS in digit order and leaves in runs.  The per-wave
// counters are 16-bit (a count or an offset within a tile is below 2^16),
// which leaves room in LDS for 16 waves beside an 8,192-pair tile.
template <typename K, int SRC, bool FILTER, int BS, int ROUNDS>
__global__ __launch_bounds__(BS) void k_part_scatter(PartIn<K> in, uint64_t n_in, const uint64_t* d_n,
                                                          Slot<K>* out, const uint32_t* offs, uint32_t nblocks,
                                                          uint32_t nsub, int sp, int rlog, int dshift, int dbits,
                                                          const BuildHeader* h, uint8_t* dig, int ndshift,
                                                          uint32_t nmask) {
   constexpr int NW = BS / kWave;
   constexpr int TILE = BS * ROUNDS;
   constexpr uint32_t DMAX = 1u << kMaxDigitBits;
   constexpr int DPT = DMAX >= BS ? DMAX / BS : 1; // digits per thread in the digit loops
   static_assert(DPT * BS >= DMAX && TILE < 65536, "digit split; 16-bit counters");
   __shared__ Slot<K> tile_s[TILE];
   __shared__ uint16_t wcnt_s[NW][DMAX];
   __shared__ uint32_t lstart_s[DMAX];
   __shared__ uint32_t goff_s[DMAX];
   __shared__ uint32_t wsum_s[NW];
   if (!h->go) return;
   const uint64_t n = d_n ? *d_n : n_in;
   const uint32_t D = 1u << dbits, dmask = D - 1;
   const int lane = __lane_id(), wave = threadIdx.x / kWave;
   const uint64_t lt = (1ull << lane) - 1ull;
   const uint32_t d0 = threadIdx.x * DPT; // this thread's digits in the digit loops
   for (uint32_t d = d0; d < d0 + DPT && d < D; ++d) goff_s[d] = offs[static_cast<uint64_t>(d) * nblocks + blockIdx.x];
   const uint64_t base = static_cast<uint64_t>(blockIdx.x) * TILE * nsub;
   // The pairs of tile `sub` are loaded while tile sub-1 is written out.
   Slot<K> el[ROUNDS];
   auto load_tile = [&](int sub) {
      const uint64_t w0 = base + static_cast<uint64_t>(sub) * TILE + static_cast<uint64_t>(wave) * (TILE / NW) +
                          static_cast<uint64_t>(lane);
#pragma unroll
      for (int r = 0; r < ROUNDS; ++r) {
         const uint64_t e = w0 + static_cast<uint64_t>(r) * kWave;
         if (e < n) el[r] = part_pair<K, SRC>(in, e);
      }
   };
   load_tile(0);
   for (int sub = 0; sub < static_cast<int>(nsub); ++sub) {
      const uint64_t t0 = base + static_cast<uint64_t>(sub) * TILE;
      if (t0 >= n) break;
      for (uint32_t d = d0; d < d0 + DPT && d < D; ++d) {
#pragma unroll
         for (int w = 0; w < NW; ++w) wcnt_s[w][d] = 0;
      }
      __syncthreads();
      uint32_t dg[ROUNDS], rk[ROUNDS];
      bool ok[ROUNDS];
      const uint64_t w0 = t0 + static_cast<uint64_t>(wave) * (TILE / NW) + lane;
#pragma unroll
      for (int r = 0; r < ROUNDS; ++r) {
         const uint64_t e = w0 + static_cast<uint64_t>(r) * kWave;
         dg[r] = 0;
         ok[r] = false;
         if (e < n) {
            ok[r] = !FILTER || live_key(el[r].key);
            dg[r] = (region_of(el[r].key, sp, rlog) >> dshift) & dmask;
         }
      }
      uint16_t* wc = wcnt_s[wave];
#pragma unroll
      for (int r = 0; r < ROUNDS; ++r) {
         const bool valid = ok[r];
         uint64_t m = __ballot(valid);
         for (int b = 0; b < dbits; ++b) {
            const bool bit = (dg[r] >> b) & 1u;
            const uint64_t bal = __ballot(bit);
            m &= bit ? bal : ~bal;
         }
         // All lanes read the counter before its one writer (the group's
         // first lane) updates it; a wave's LDS operations run in order.
         if (valid) rk[r] = wc[dg[r]] + static_cast<uint32_t>(__popcll(m & lt));
         if (valid && first_bit(m) == lane) wc[dg[r]] += static_cast<uint16_t>(__popcll(m));
      }
      __syncthreads();
      // Per digit: waves' counts -> exclusive offsets over waves; tile totals -> exclusive scan.
      uint32_t tot[DPT];
      uint32_t acc = 0;
#pragma unroll
      for (int i = 0; i < DPT; ++i) {
         tot[i] = 0;
         const uint32_t d = d0 + i;
         if (d < D) {
#pragma unroll
            for (int w = 0; w < NW; ++w) {
               const uint32_t c = wcnt_s[w][d];
               wcnt_s[w][d] = static_cast<uint16_t>(tot[i]);
               tot[i] += c;
            }
         }
         acc += tot[i];
      }
      {
         uint32_t inc = acc;
#pragma unroll
         for (int o = 1; o < kWave; o <<= 1) {
            const uint32_t y = __shfl_up(inc, o, kWave);
            if (lane >= o) inc += y;
         }
         if (lane == kWave - 1) wsum_s[wave] = inc;
         __syncthreads();
         uint32_t bse = inc - acc;
         for (int w = 0; w < wave; ++w) bse += wsum_s[w];
#pragma unroll
         for (int i = 0; i < DPT; ++i) {
            if (d0 + i < D) lstart_s[d0 + i] = bse;
            bse += tot[i];
         }
      }
      __syncthreads();
#pragma unroll
      for (int r = 0; r < ROUNDS; ++r) {
         if (ok[r]) {
            const uint32_t p = lstart_s[dg[r]] + wcnt_s[wave][dg[r]] + rk[r];
            tile_s[p] = el[r];
         }
      }
      if (sub + 1 < static_cast<int>(nsub)) load_tile(sub + 1);
      __syncthreads();
      uint32_t cnt = 0; // pairs of this tile that are written: the digit scan's total
      for (int w = 0; w < NW; ++w) cnt += wsum_s[w];
      for (uint32_t p = threadIdx.x; p < cnt; p += BS) {
         const Slot<K> e = tile_s[p];
         const uint32_t rg = region_of(e.key, sp, rlog);
         const uint32_t d = (rg >> dshift) & dmask; // recomputed: cheaper than LDS for it
         const uint64_t pos = static_cast<uint64_t>(goff_s[d]) + (p - lstart_s[d]);
         out[pos] = e;
         if (nmask) dig[pos] = static_cast<uint8_t>((rg >> ndshift) & nmask); // the next pass's digit
      }
      __syncthreads();
#pragma unroll
      for (int i = 0; i < DPT; ++i)
         if (d0 + i < D) goff_s[d0 + i] += tot[i];
   }
}

// Exclusive scan of x[0..m) in LDS by a whole workgroup; returns the total.
template <int BS>
__device__ uint32_t lds_exclusive_scan(uint32_t* x, uint32_t m, uint32_t* wsum) {
   const int lane = __lane_id(), wave = threadIdx.x / kWave;
   const uint32_t per = (m + BS - 1) / BS;
   const uint32_t c0 = std::min<uint32_t>(threadIdx.x * per, m), c1 = std::min<uint32_t>(c0 + per, m);
   uint32_t acc = 0;
   for (uint32_t i = c0; i < c1; ++i) acc += x[i];
   uint32_t inc = acc;
#pragma unroll
   for (int o = 1; o < kWave; o <<= 1) {
      const uint32_t y = __shfl_up(inc, o, kWave);
      if (lane >= o) inc += y;
   }
   if (lane == kWave - 1) wsum[wave] = inc;
   __syncthreads();
   uint32_t run = inc - acc, total = 0;
   for (int w = 0; w < BS / kWave; ++w) {
      if (w < wave) run += wsum[w];
      total += wsum[w];
   }
   for (uint32_t i = c0; i < c1; ++i) {
      const uint32_t v = x[i];
      x[i] = run;
      run += v;
   }
   __syncthreads();
   return total;
}

// Carry functions g(c) = max(a, c + b): `first` applied, then `then`.
constexpr int64_t kNegInf = INT64_MIN / 4; // identity: max(-inf, c + 0) = c
__host__ __device__ __forceinline__ RegionSum compose(RegionSum first, RegionSum then) {
   RegionSum g;
   g.a = first.a + then.b > then.a ? first.a + then.b : then.a;
   g.b = first.b + then.b;
   return g;
}

// Carry bookkeeping over regions: each region's carry function, the carry
// it was placed with (spec), the true carry, and the regions to place again.
struct RegionLB {
   uint64_t* word;  // [NR] the region's carry function, packed
   int64_t* spec;   // [NR] carry the region was placed with
   int64_t* carry;  // [NR] true carry
   uint32_t* mark;  // [NR] 1: place again
   uint32_t* list;  // [NR] regions to place again (count in the header)
};

// A region's carry function in one 64-bit word: a <= keys in the region
// < 2^24; b + R in [0, 2^24).
__device__ __forceinline__ uint64_t pack_sum(RegionSum g, int64_t R) {
   return (1ull << 63) | (static_cast<uint64_t>(g.b + R) << 24) | static_cast<uint64_t>(g.a);
}
__device__ __forceinline__ RegionSum unpack_sum(uint64_t w, int64_t R) {
   RegionSum g;
   g.a = static_cast<int64_t>(w & 0xFFFFFFull);
   g.b = static_cast<int64_t>((w >> 24) & 0xFFFFFFull) - R;
   return g;
}
__device__ __forceinline__ RegionSum load_sum(const RegionLB& lb, uint64_t r, int64_t R) {
   return unpack_sum(lb.word[r], R);
}

// start[r] = first index of region r in the partitioned array (the first
// index whose region is >= r, so an empty region starts where the next one
// does), start[NR] = n: a binary search per region over the sorted array.
// Neighbouring lanes search neighbouring regions and share the upper levels
// of their searches in cache; the array itself is not streamed.  The same
// launch clears the region bookkeeping of the placing step.
template <typename K>
__global__ __launch_bounds__(256) void k_region_starts(const Slot<K>* S, uint64_t* start, RegionLB lb, uint64_t NR,
                                                       int sp, int rlog, BuildHeader* h) {
   if (!h->go) return;
   const uint64_t n = h->n;
   const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
   for (uint64_t r = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; r <= NR; r += stride) {
      uint64_t lo = 0, hi = n;
      while (lo < hi) {
         const uint64_t mid = lo + (hi - lo) / 2;
         if (region_of(S[mid].key, sp, rlog) < r) lo = mid + 1;
         else hi = mid;
      }
      start[r] = lo;
      if (r < NR) {
         lb.word[r] = 0;
         lb.mark[r] = 0;
      }
   }
   if (blockIdx.x == 0 && threadIdx.x == 0) h->n_fix = 0;
}

constexpr int kRegionNW = kRegionBS / kWave;
constexpr uint32_t kRegionMax = 1u << kRegionLog;
constexpr uint32_t kSpanMax = kRegionMax + kRegionCap; // slots one region writes at most
constexpr uint32_t kSpanWords = (kSpanMax + 31) / 32;

// LDS of one region: the sort tables are dead once each key knows its sorted
// index, so the placement arrays reuse their space.
template <typename K>
struct RegionLds {
   K key[kRegionCap];
   uint16_t order[kRegionCap];
   union {
      uint8_t dup[kRegionCap];
      uint32_t occ[kSpanWords]; // slots of the write span holding a key
   } u1;
   union {
      struct {
         uint32_t cnt[kRegionMax + 1];  // keys per home offset, then their exclusive starts
         uint32_t scnt[kRegionMax + 1]; // the same over the surviving keys
      } c;
      int32_t v[kRegionCap]; // by sorted index f: home offset - f, then its running max
   } u2;
   uint32_t wsum[kRegionNW];
   int32_t wmax[kRegionNW];
   uint32_t r;
   int64_t carry;
};
static_assert(sizeof(uint32_t) * kSpanWords <= kRegionCap, "occupancy bits fit the dup table");
static_assert(sizeof(int32_t) * kRegionCap <= 2 * sizeof(uint32_t) * (kRegionMax + 1), "v fits the count tables");

// Sort region keys S[lo, lo + m) in LDS by (home, key), keeping the last
// occurrence of a repeated key.  Thread t owns keys t + k*kRegionBS; ff[k] =
// its sorted index among the survivors, or -1.  Leaves v[f] = home offset - f.
// Returns the number of survivors.  Called by the whole workgroup.
//
// A counting sort by home slot puts the keys, with their input indices, in
// home order in LDS; a key's rank among its home's keys is then one pass over
// that home's few entries (0.9 others on average at load factor 0.9), which
// also spots repeated keys.  Without repeats in the region (the block agrees
// with one __syncthreads_or) every key survives and the sorted index is
// home start + rank; otherwise the survivors are counted per home and ranked
// again among survivors.  PRE: kk/vv already hold the pairs (loaded ahead by
// the caller), S is not read.
template <typename K, bool PRE = false>
__device__ uint32_t region_sort(const Slot<K>* S, uint64_t lo, uint32_t m, int sp, uint64_t R, RegionLds<K>& L,
                                K* kk, K* vv, uint32_t* uu, int32_t* ff) {
   __syncthreads(); // the previous user of L is done
   for (uint32_t i = threadIdx.x; i <= R; i += kRegionBS) L.u2.c.cnt[i] = 0;
   __syncthreads();
   uint32_t sub[kRegionIPT];
#pragma unroll
   for (int k = 0; k < kRegionIPT; ++k) {
      const uint32_t e = threadIdx.x + k * kRegionBS;
      ff[k] = -1;
      if (e < m) {
         if (!PRE) {
            const Slot<K> s = S[lo + e];
            kk[k] = s.key;
            vv[k] = s.val;
         }
         uu[k] = static_cast<uint32_t>((fib_home(kk[k], sp) & ((1ull << sp) - 1)) & (R - 1));
         sub[k] = atomicAdd(&L.u2.c.cnt[uu[k]], 1u);
      }
   }
   __syncthreads();
   lds_exclusive_scan<kRegionBS>(L.u2.c.cnt, static_cast<uint32_t>(R) + 1, L.wsum);
   // keys and input indices in home order
#pragma unroll
   for (int k = 0; k < kRegionIPT; ++k) {
      const uint32_t e = threadIdx.x + k * kRegionBS;
      if (e < m) {
         const uint32_t pos = L.u2.c.cnt[uu[k]] + sub[k];
         L.key[pos] = kk[k];
         L.order[pos] = static_cast<uint16_t>(e);
      }
   }
   __syncthreads();
   // rank among the home's keys; a later copy of the same key drops this one
   uint32_t dupm = 0; // bit k: key k is dropped
   int repeats = 0;
#pragma unroll
   for (int k = 0; k < kRegionIPT; ++k) {
      const uint32_t e = threadIdx.x + k * kRegionBS;
      if (e < m) {
         const uint32_t j0 = L.u2.c.cnt[uu[k]], j1 = L.u2.c.cnt[uu[k] + 1];
         int32_t rank = 0;
         for (uint32_t j = j0; j < j1; ++j) {
            const K kj = L.key[j];
            const uint32_t oj = L.order[j];
            rank += kj < kk[k] ? 1 : 0;
            if (kj == kk[k] && oj != e) {
               repeats = 1;
               if (oj > e) dupm |= 1u << k;
            }
         }
         ff[k] = rank;
      }
   }
   uint32_t ms;
   if (!__syncthreads_or(repeats)) {
      ms = m;
#pragma unroll
      for (int k = 0; k < kRegionIPT; ++k)
         if (ff[k] >= 0) ff[k] += static_cast<int32_t>(L.u2.c.cnt[uu[k]]);
   } else {
      for (uint32_t i = threadIdx.x; i <= R; i += kRegionBS) L.u2.c.scnt[i] = 0;
#pragma unroll
      for (int k = 0; k < kRegionIPT; ++k) {
         const uint32_t e = threadIdx.x + k * kRegionBS;
         if (e < m) L.u1.dup[e] = (dupm >> k) & 1u;
      }
      __syncthreads();
#pragma unroll
      for (int k = 0; k < kRegionIPT; ++k)
         if (ff[k] >= 0 && !((dupm >> k) & 1u)) atomicAdd(&L.u2.c.scnt[uu[k]], 1u);
      __syncthreads();
      ms = lds_exclusive_scan<kRegionBS>(L.u2.c.scnt, static_cast<uint32_t>(R) + 1, L.wsum);
#pragma unroll
      for (int k = 0; k < kRegionIPT; ++k) {
         if (ff[k] < 0) continue;
         if ((dupm >> k) & 1u) {
            ff[k] = -1;
            continue;
         }
         int32_t rank = 0;
         for (uint32_t j = L.u2.c.cnt[uu[k]]; j < L.u2.c.cnt[uu[k] + 1]; ++j)
            rank += (!L.u1.dup[L.order[j]] && L.key[j] < kk[k]) ? 1 : 0;
         ff[k] = static_cast<int32_t>(L.u2.c.scnt[uu[k]]) + rank;
      }
   }
   __syncthreads(); // count tables dead from here; v takes their place
#pragma unroll
   for (int k = 0; k < kRegionIPT; ++k)
      if (ff[k] >= 0) L.u2.v[ff[k]] = static_cast<int32_t>(uu[k]) - ff[k];
   __syncthreads();
   return ms;
}

template <typename K>
__device__ RegionSum region_summary(RegionLds<K>& L, uint32_t ms, uint64_t R, const int32_t* ff) {
   int32_t M = INT32_MIN;
#pragma unroll
   for (int k = 0; k < kRegionIPT; ++k)
      if (ff[k] >= 0) M = std::max(M, L.u2.v[ff[k]]);
   M = wave_max(M);
   if (__lane_id() == 0) L.wmax[threadIdx.x / kWave] = M;
   __syncthreads();
   int32_t Mb = INT32_MIN;
   for (int w = 0; w < kRegionNW; ++w) Mb = std::max(Mb, L.wmax[w]);
   __syncthreads();
   RegionSum g;
   const int64_t mp = static_cast<int64_t>(ms), Ri = static_cast<int64_t>(R);
   g.b = mp - Ri;
   g.a = mp > 0 ? std::max<int64_t>(0, mp - Ri + Mb) : 0;
   return g;
}

// Place a sorted region after carry c: key f at p = f + max(running max of v, c),
// and write the span [c, max(R, last p + 1)) of slots (keys and EMPTY) at
// base + offset.  `keep` picks by unwrapped position: 0 = below N only, 1 =
// all, 2 = N and above only.  Returns this thread's longest probe.
template <typename K>
__device__ uint32_t region_place(RegionLds<K>& L, uint32_t ms, int64_t c, uint64_t base, int sp, uint64_t R,
                                 int keep, const K* kk, const K* vv, const uint32_t* uu, const int32_t* ff,
                                 Slot<K>* b) {
   const int lane = __lane_id(), wave = threadIdx.x / kWave;
   const uint64_t N = 1ull << sp, mask = N - 1;
   {
      const uint32_t per = (ms + kRegionBS - 1) / kRegionBS;
      const uint32_t j0 = std::min<uint32_t>(threadIdx.x * per, ms), j1 = std::min<uint32_t>(j0 + per, ms);
      int32_t acc = INT32_MIN;
      for (uint32_t j = j0; j < j1; ++j) acc = std::max(acc, L.u2.v[j]);
      int32_t inc = acc;
#pragma unroll
      for (int o = 1; o < kWave; o <<= 1) {
         const int32_t y = __shfl_up(inc, o, kWave);
         if (lane >= o) inc = std::max(inc, y);
      }
      if (lane == kWave - 1) L.wmax[wave] = inc;
      __syncthreads();
      int32_t run = __shfl_up(inc, 1, kWave);
      if (lane == 0) run = INT32_MIN;
      for (int w = 0; w < wave; ++w) run = std::max(run, L.wmax[w]);
      for (uint32_t j = j0; j < j1; ++j) {
         run = std::max(run, L.u2.v[j]);
         L.u2.v[j] = run;
      }
   }
   __syncthreads();
   const int64_t last = ms ? static_cast<int64_t>(ms - 1) + std::max<int64_t>(L.u2.v[ms - 1], c) : -1;
   const int64_t end = std::max<int64_t>(static_cast<int64_t>(R), last + 1);
   const int64_t span = end - c;
   const int64_t words = span > 0 ? (span + 31) / 32 : 0;
   for (int64_t w = threadIdx.x; w < words; w += kRegionBS) L.u1.occ[w] = 0;
   __syncthreads();
   auto kept = [&](uint64_t pos) { return keep == 1 || (keep == 0 ? pos < N : pos >= N); };
   uint32_t probe = 0;
#pragma unroll
   for (int k = 0; k < kRegionIPT; ++k) {
      if (ff[k] >= 0) {
         const int64_t p = ff[k] + std::max<int64_t>(L.u2.v[ff[k]], c);
         const int64_t q = p - c;
         atomicOr(&L.u1.occ[q >> 5], 1u << (q & 31));
         const uint64_t pos = base + static_cast<uint64_t>(p);
         if (kept(pos)) {
            Slot<K> s;
            s.key = kk[k];
            s.val = vv[k];
            b[pos & mask] = s;
         }
         probe = std::max<uint32_t>(probe, static_cast<uint32_t>(p - static_cast<int64_t>(uu[k]) + 1));
      }
   }
   __syncthreads();
   for (int64_t q = threadIdx.x; q < span; q += kRegionBS) {
      if (!((L.u1.occ[q >> 5] >> (q & 31)) & 1u)) {
         const uint64_t pos = base + static_cast<uint64_t>(c + q);
         if (kept(pos)) {
            Slot<K> s;
            s.key = empty_key<K>();
            s.val = 0;
            b[pos & mask] = s;
         }
      }
   }
   return probe;
}

template <typename K>
__device__ void probe_to_header(RegionLds<K>& L, uint32_t probe, BuildHeader* h) {
   probe = wave_max(probe);
   __syncthreads();
   if (__lane_id() == 0) L.wmax[threadIdx.x / kWave] = static_cast<int32_t>(probe);
   __syncthreads();
   if (threadIdx.x == 0) {
      uint32_t pb = 0;
      for (int w = 0; w < kRegionNW; ++w) pb = std::max(pb, static_cast<uint32_t>(L.wmax[w]));
      if (pb > __hip_atomic_load(&h->max_probe, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
         atomic_max_relaxed(&h->max_probe, pb);
   }
}

// One workgroup walks `chunk` consecutive regions in order: sort a
// region, record its carry function, place it with the carry handed on from
// the region before, and hand on max(a, c + b).  The walk starts from the
// carry out of the region before the chunk taken with no carry in (that
// region is sorted once more for its summary).  Slots at unwrapped positions
// >= N are skipped.  Nothing waits on another workgroup; the carry kernels
// finds the regions whose carry was wrong (a spill crossing a whole region
// into a chunk, or the wrap past the table's end) and k_region_fix places
// them again.  The carries do not depend on the chunk: a shorter one only
// means more walks and more regions sorted twice.
constexpr uint32_t kRegionChunk = 16;

// Regions per walk: kRegionChunk, or fewer when that would leave the device
// short of workgroups (tables below 2^24 slots): a 2^20-slot table has 512
// regions, 32 walks of 16 but 512 of 1.
inline uint32_t region_chunk(uint64_t NR) {
   const uint64_t c = NR / (static_cast<uint64_t>(cu_count()) * 2);
   return static_cast<uint32_t>(c < 1 ? 1 : (c > kRegionChunk ? kRegionChunk : c));
}

// The next region's pairs are loaded while the current one is placed (no
// walk waits on another, so nothing depends on the order of completion);
// held to 128 VGPRs for two workgroups per CU: 8.68 ms against 9.19 ms for
// the 2^28-slot build, 10.2 ms when the extra registers left one per CU.
template <typename K>
__global__ __launch_bounds__(kRegionBS) __attribute__((amdgpu_waves_per_eu(4))) void k_region_build(
    const Slot<K>* S, const uint64_t* start, RegionLB lb, Slot<K>* b, int sp, int rlog, uint64_t NR, uint32_t chunk,
    BuildHeader* h) {
   __shared__ RegionLds<K> L;
   if (!h->go) return;
   const uint64_t R = 1ull << rlog;
   const int64_t Ri = static_cast<int64_t>(R);
   const uint64_t r0 = static_cast<uint64_t>(blockIdx.x) * chunk;
   const uint64_t r1 = std::min<uint64_t>(r0 + chunk, NR);
   K kk[kRegionIPT], vv[kRegionIPT];
   K nk[kRegionIPT], nv[kRegionIPT]; // the next region's pairs, loaded while this one is placed
   uint32_t uu[kRegionIPT];
   int32_t ff[kRegionIPT];
   uint32_t probe = 0;
   int64_t c = 0;
   auto prefetch = [&](uint64_t r) {
      const uint64_t lo = start[r], m = start[r + 1] - lo;
#pragma unroll
      for (int k = 0; k < kRegionIPT; ++k) {
         const uint32_t e = threadIdx.x + k * kRegionBS;
         if (e < m && m <= static_cast<uint64_t>(kRegionCap)) {
            const Slot<K> s = S[lo + e];
            nk[k] = s.key;
            nv[k] = s.val;
         }
      }
   };
   if (r0 < r1) prefetch(r0);
   if (r0 > 0) {
      // Seed the walk with the carry out of the region before the chunk, that
      // region taken with no carry in: exact unless a spill crosses it whole.
      const uint64_t lo = start[r0 - 1], m = start[r0] - lo;
      if (m <= static_cast<uint64_t>(kRegionCap)) {
         const uint32_t ms = region_sort<K>(S, lo, static_cast<uint32_t>(m), sp, R, L, kk, vv, uu, ff);
         const RegionSum g = region_summary<K>(L, ms, R, ff);
         c = std::max(g.a, g.b);
      }
   }
   for (uint64_t r = r0; r < r1; ++r) {
      const uint64_t lo = start[r], m = start[r + 1] - lo;
      if (m > static_cast<uint64_t>(kRegionCap)) { // abandon the build
         if (threadIdx.x == 0) __hip_atomic_store(&h->fail, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
         return;
      }
#pragma unroll
      for (int k = 0; k < kRegionIPT; ++k) {
         kk[k] = nk[k];
         vv[k] = nv[k];
      }
      const uint32_t ms = region_sort<K, true>(S, lo, static_cast<uint32_t>(m), sp, R, L, kk, vv, uu, ff);
      if (r + 1 < r1) prefetch(r + 1);
      probe = std::max(probe, region_place<K>(L, ms, c, r * R, sp, R, 0, kk, vv, uu, ff, b));
      // The region's summary from the prefix max the placement left in v:
      // v[ms - 1] = max over its keys of (home offset - sorted index).
      RegionSum g;
      g.b = static_cast<int64_t>(ms) - Ri;
      g.a = ms > 0 ? std::max<int64_t>(0, static_cast<int64_t>(ms) - Ri + L.u2.v[ms - 1]) : 0;
      if (threadIdx.x == 0) {
         lb.word[r] = pack_sum(g, Ri);
         lb.spec[r] = c;
      }
      c = std::max(g.a, c + g.b);
   }
   probe_to_header<K>(L, probe, h);
}

// True carries.  The carry into region 0 is the least fixed point of the
// whole-table composition max(A, c + B), B = n - N < 0: c0 = A; region r's
// carry is the composition over regions 0..r-1 applied to c0.  Three small
// kernels: per block of kCarryTile regions the composition (k_carry_local),
// the blocks' prefixes and c0 (k_carry_top, one workgroup), each region's
// carry (k_carry_apply); then k_carry_mark lists the regions to place again.
constexpr int kCarryBS = 256;
constexpr int kCarryPer = 4;
constexpr uint64_t kCarryTile = kCarryBS * kCarryPer;

// Block-wide inclusive scan (composition) of one RegionSum per thread, in LDS.
__device__ RegionSum carry_block_scan(RegionSum mine, int64_t* fa, int64_t* fb) {
   fa[threadIdx.x] = mine.a;
   fb[threadIdx.x] = mine.b;
   __syncthreads();
   for (int o = 1; o < static_cast<int>(blockDim.x); o <<= 1) {
      const RegionSum cur{fa[threadIdx.x], fb[threadIdx.x]};
      RegionSum prev{kNegInf, 0};
      if (static_cast<int>(threadIdx.x) >= o) prev = RegionSum{fa[threadIdx.x - o], fb[threadIdx.x - o]};
      __syncthreads();
      if (static_cast<int>(threadIdx.x) >= o) {
         const RegionSum c = compose(prev, cur);
         fa[threadIdx.x] = c.a;
         fb[threadIdx.x] = c.b;
      }
      __syncthreads();
   }
   return RegionSum{fa[threadIdx.x], fb[threadIdx.x]};
}

__global__ __launch_bounds__(kCarryBS) void k_carry_local(RegionLB lb, RegionSum* blk, uint64_t NR, int64_t R,
                                                           const BuildHeader* h) {
   __shared__ int64_t fa[kCarryBS], fb[kCarryBS];
   if (!h->go || h->fail) return;
   const uint64_t r0 = blockIdx.x * kCarryTile + threadIdx.x * kCarryPer;
   RegionSum g{kNegInf, 0};
   for (uint64_t r = r0; r < r0 + kCarryPer && r < NR; ++r) g = compose(g, load_sum(lb, r, R));
   const RegionSum inc = carry_block_scan(g, fa, fb);
   if (threadIdx.x == kCarryBS - 1) blk[blockIdx.x] = inc;
}

// One workgroup: blk[i] becomes the carry into block i (as RegionSum{c, 0}).
__global__ __launch_bounds__(1024) void k_carry_top(RegionSum* blk, uint64_t nblk, uint64_t N, BuildHeader* h) {
   __shared__ int64_t fa[1024], fb[1024];
   if (!h->go || h->fail) return;
   const uint64_t per = (nblk + 1023) / 1024;
   const uint64_t b0 = std::min<uint64_t>(threadIdx.x * per, nblk), b1 = std::min<uint64_t>(b0 + per, nblk);
   RegionSum g{kNegInf, 0};
   for (uint64_t i = b0; i < b1; ++i) g = compose(g, blk[i]);
   const RegionSum inc = carry_block_scan(g, fa, fb);
   const RegionSum T{fa[1023], fb[1023]};
   const int64_t c0 = std::max<int64_t>(T.a, 0);
   RegionSum P{kNegInf, 0}; // composition over the blocks before b0
   if (threadIdx.x > 0) P = RegionSum{fa[threadIdx.x - 1], fb[threadIdx.x - 1]};
   (void)inc;
   __syncthreads();
   int64_t c = std::max(P.a, c0 + P.b);
   for (uint64_t i = b0; i < b1; ++i) {
      const RegionSum gi = blk[i];
      blk[i] = RegionSum{c, 0};
      c = std::max(gi.a, c + gi.b);
   }
   if (threadIdx.x == 0) {
      h->c0 = c0;
      h->n_distinct = static_cast<uint64_t>(T.b + static_cast<int64_t>(N));
   }
}

__global__ __launch_bounds__(kCarryBS) void k_carry_apply(RegionLB lb, const RegionSum* blk, uint64_t NR, int64_t R,
                                                           const BuildHeader* h) {
   __shared__ int64_t fa[kCarryBS], fb[kCarryBS];
   if (!h->go || h->fail) return;
   const uint64_t r0 = blockIdx.x * kCarryTile + threadIdx.x * kCarryPer;
   RegionSum gs[kCarryPer];
   RegionSum g{kNegInf, 0};
#pragma unroll
   for (int k = 0; k < kCarryPer; ++k) {
      gs[k] = r0 + k < NR ? load_sum(lb, r0 + k, R) : RegionSum{kNegInf, 0};
      g = compose(g, gs[k]);
   }
   carry_block_scan(g, fa, fb);
   RegionSum P{kNegInf, 0};
   if (threadIdx.x > 0) P = RegionSum{fa[threadIdx.x - 1], fb[threadIdx.x - 1]};
   const int64_t cin = blk[blockIdx.x].a;
   int64_t c = std::max(P.a, cin + P.b);
#pragma unroll
   for (int k = 0; k < kCarryPer; ++k) {
      if (r0 + k < NR) {
         lb.carry[r0 + k] = c;
         c = std::max(gs[k].a, c + gs[k].b);
      }
   }
}

// Placed again: every region whose spec carry differs, the regions whose
// true slot ranges cover the slots such a region wrote wrongly
// ([rR + spec, rR + carry)), and the regions whose range reaches past N
// (those slots were skipped).
__global__ __launch_bounds__(256) void k_carry_mark(RegionLB lb, uint64_t NR, int64_t R, int64_t N, BuildHeader* h) {
   if (!h->go || h->fail) return;
   const int64_t c0 = h->c0;
   auto start_of = [&](uint64_t r) { return static_cast<int64_t>(r) * R + lb.carry[r]; };
   auto mark = [&](uint64_t r) {
      if (atomicExch(&lb.mark[r], 1u) == 0) lb.list[atomicAdd(&h->n_fix, 1u)] = static_cast<uint32_t>(r);
   };
   const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
   for (uint64_t r = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; r < NR; r += stride) {
      const int64_t cr = lb.carry[r], sr = lb.spec[r];
      const int64_t next = r + 1 < NR ? start_of(r + 1) : N + c0; // end of r's true range
      if (next > N) mark(r);                                      // reaches past N
      if (cr != sr) {
         mark(r);
         const int64_t lo = static_cast<int64_t>(r) * R + sr;
         for (int64_t q = static_cast<int64_t>(r) - 1; q >= 0; --q) {
            mark(static_cast<uint64_t>(q));
            if (start_of(static_cast<uint64_t>(q)) <= lo) break;
         }
      }
   }
}

// Place the listed regions again with their true carries, all slots of
// their ranges (wrapping past N).
template <typename K>
__global__ __launch_bounds__(kRegionBS) void k_region_fix(const Slot<K>* S, const uint64_t* start, RegionLB lb,
                                                         Slot<K>* b, int sp, int rlog, BuildHeader* h) {
   __shared__ RegionLds<K> L;
   if (!h->go || h->fail) return;
   const uint64_t R = 1ull << rlog;
   K kk[kRegionIPT], vv[kRegionIPT];
   uint32_t uu[kRegionIPT];
   int32_t ff[kRegionIPT];
   uint32_t probe = 0;
   const uint32_t nfix = h->n_fix;
   for (uint32_t i = blockIdx.x; i < nfix; i += gridDim.x) {
      const uint32_t r = lb.list[i];
      const uint64_t lo = start[r];
      const uint32_t ms = region_sort<K>(S, lo, static_cast<uint32_t>(start[r + 1] - lo), sp, R, L, kk, vv, uu, ff);
      probe = std::max(probe, region_place<K>(L, ms, lb.carry[r], static_cast<uint64_t>(r) * R, sp, R, 1, kk, vv,
                                              uu, ff, b));
   }
   probe_to_header<K>(L, probe, h);
}

// A region over capacity abandons the placement: the table is emptied and
// the probing insert takes the partitioned pairs (all of them, in S).
template <typename K>
__global__ __launch_bounds__(256) void k_build_fallback_prep(Slot<K>* b, uint64_t N, hn_info* info, BuildHeader* h) {
   const bool fb = h->go && h->fail;
   if (blockIdx.x == 0 && threadIdx.x == 0) {
      h->fb_len = fb ? h->n : 0;
      if (fb) {
         info->fill = 0;
         info->current_max_bucket_overflow = kBucketOverflow;
      }
   }
   if (!fb) return;
   const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
   for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < N; i += stride) {
      Slot<K> s;
      s.key = empty_key<K>();
      s.val = 0;
      b[i] = s;
   }
}

// Finish the Info.  Not run (go = 0): an empty-table build reports
// err = INVALID with the table untouched.  Placed: fill, overflow of the layout, ordered.  Fallback: the
// insert has set fill and overflow.
__global__ void k_build_end(hn_info* info, const BuildHeader* h, int mode) {
   if (threadIdx.x != 0) return;
   if (!h->go) {
      if (mode == BUILD_EMPTY) info->err = HN_STATUS_INVALID;
      return;
   }
   info->reserved0 = 0;
   info->tombstone_counter = 0;
   if (h->fail) {
      info->err = h->fb_fail ? HN_STATUS_FAIL : HN_STATUS_SUCCESS;
      info->reserved1 = 0;
   } else {
      info->fill = h->n_distinct;
      info->current_max_bucket_overflow =
          std::max<uint64_t>(kBucketOverflow, next_overflow(h->max_probe, kBucketOverflow));
      info->err = HN_STATUS_SUCCESS;
      info->reserved1 = kInfoOrdered;
   }
   // device_rehash reports success iff the new table holds the old fill (hashinator.h:333)
   if (mode == BUILD_REHASH && info->err == HN_STATUS_SUCCESS && info->fill != h->prior_fill)
      info->err = HN_STATUS_FAIL;
}

// Rehash: remember the old fill and give the Info the new table's size
// (Info(newSizePower), hashinator.h:328).  The build writes every slot of the
// new table, so it needs no initialisation.
__global__ void k_rehash_info(hn_info* info, int new_sp, BuildHeader* h) {
   if (threadIdx.x != 0) return;
   h->prior_fill = info->fill;
   info->size_power = new_sp;
   info->reserved0 = 0;
   info->fill = 0;
   info->current_max_bucket_overflow = kBucketOverflow;
   info->tombstone_counter = 0;
   info->err = HN_STATUS_SUCCESS;
   info->reserved1 = 0;
}

// A pass's digits come from the byte array when the pass before it wrote
// them: a pass after the first whose digit fits a byte.
inline bool digits_in_bytes(const BuildPlan& P, int pass) { return pass > 0 && pass < P.passes && P.dbits[pass] <= 8; }

template <typename K, int SRC, bool FILTER>
int launch_part(const BuildPlan& P, int pass, PartIn<K> in, uint64_t n_in, const uint64_t* d_n, Slot<K>* out,
                uint32_t* mat, void* scan_scratch, size_t scan_bytes, BuildHeader* h, uint8_t* dig, hipStream_t s) {
   int dshift = 0;
   for (int i = 0; i < pass; ++i) dshift += P.dbits[i];
   const int dbits = P.dbits[pass];
   const uint64_t D = 1ull << dbits;
   const int ndshift = dshift + dbits;
   const uint32_t nmask = digits_in_bytes(P, pass + 1) ? (1u << P.dbits[pass + 1]) - 1u : 0u;
   if constexpr (SRC == SRC_TABLE && !FILTER) {
      if (digits_in_bytes(P, pass)) {
         hipLaunchKernelGGL((k_part_count<K, SRC, FILTER, true>), dim3(P.nblocks), dim3(kPartBS), 0, s, in, n_in,
                            d_n, mat, P.nblocks, P.tile, P.sub, P.sp, P.rlog, dshift, dbits, h, dig);
      } else {
         hipLaunchKernelGGL((k_part_count<K, SRC, FILTER>), dim3(P.nblocks), dim3(kPartBS), 0, s, in, n_in, d_n,
                            mat, P.nblocks, P.tile, P.sub, P.sp, P.rlog, dshift, dbits, h, dig);
      }
   } else {
      hipLaunchKernelGGL((k_part_count<K, SRC, FILTER>), dim3(P.nblocks), dim3(kPartBS), 0, s, in, n_in, d_n, mat,
                         P.nblocks, P.tile, P.sub, P.sp, P.rlog, dshift, dbits, h, dig);
   }
   int rc = hn_exclusive_scan_u32(mat, mat, D * P.nblocks, scan_scratch, scan_bytes, reinterpret_cast<hn_stream_t>(s));
   if (rc) return rc;
   static_assert(kPartBSL * kPartRoundsL == kPartTileL && kPartBSS * kPartRoundsS == kPartTileS, "scatter tiles");
   if (P.tile == kPartTileL)
      hipLaunchKernelGGL((k_part_scatter<K, SRC, FILTER, kPartBSL, kPartRoundsL>), dim3(P.nblocks), dim3(kPartBSL), 0,
                         s, in, n_in, d_n, out, mat, P.nblocks, P.sub, P.sp, P.rlog, dshift, dbits, h, dig, ndshift,
                         nmask);
   else
      hipLaunchKernelGGL((k_part_scatter<K, SRC, FILTER, kPartBSS, kPartRoundsS>), dim3(P.nblocks), dim3(kPartBSS), 0,
                         s, in, n_in, d_n, out, mat, P.nblocks, P.sub, P.sp, P.rlog, dshift, dbits, h, dig, ndshift,
                         nmask);
   return static_cast<int>(hipGetLastError());
}

// Defined with the insert kernels (hn_map.hip): the probing insert of
// pairs[0, *d_len), failures OR-ed into *fail_sink, Info left to the caller.
template <typename K>
int launch_insert_pairs_dlen(const Slot<K>* pairs, Slot<K>* buckets, hn_info* info, uint64_t max_len,
                             const uint64_t* d_len, int32_t* fail_sink, hipStream_t s);

// The partition passes and the region starts: the input pairs end in S,
// stably partitioned by region, with start[r] the first of region r.  src0
// picks the first pass's input kind; n_in is its element count (the old
// table's slot count for a rehash, whose first pass keeps live slots only).
// Pass i writes S when (passes - 1 - i) is even, `alt` (memory of the same
// size that is written later anyway: the output table) otherwise, so the
// last pass lands in S.
template <typename K>
int partition_regions(const BuildPlan& P, int src0, PartIn<K> in0, uint64_t n_in, Slot<K>* alt, void* scratch,
                      hipStream_t s) {
   char* base = static_cast<char*>(scratch);
   BuildHeader* h = reinterpret_cast<BuildHeader*>(base);
   Slot<K>* S = reinterpret_cast<Slot<K>*>(base + P.off_pairs);
   uint32_t* mat = reinterpret_cast<uint32_t*>(base + P.off_mat);
   void* scan_scratch = base + P.off_scan;
   const size_t scan_bytes = P.off_start - P.off_scan;
   uint64_t* start = reinterpret_cast<uint64_t*>(base + P.off_start);
   uint8_t* dig = reinterpret_cast<uint8_t*>(base + P.off_dig);
   RegionLB lb;
   lb.word = reinterpret_cast<uint64_t*>(base + P.off_word);
   lb.spec = reinterpret_cast<int64_t*>(base + P.off_spec);
   lb.carry = reinterpret_cast<int64_t*>(base + P.off_carry);
   lb.mark = reinterpret_cast<uint32_t*>(base + P.off_mark);
   lb.list = reinterpret_cast<uint32_t*>(base + P.off_list);
   const Slot<K>* src = nullptr;
   for (int pass = 0; pass < P.passes; ++pass) {
      Slot<K>* dst = ((P.passes - 1 - pass) % 2 == 0) ? S : alt;
      int rc;
      if (pass == 0) {
         switch (src0) {
         case SRC_PAIRS:
            rc = launch_part<K, SRC_PAIRS, false>(P, 0, in0, n_in, nullptr, dst, mat, scan_scratch, scan_bytes, h, dig, s);
            break;
         case SRC_SOA:
            rc = launch_part<K, SRC_SOA, false>(P, 0, in0, n_in, nullptr, dst, mat, scan_scratch, scan_bytes, h, dig, s);
            break;
         case SRC_INDEX:
            rc = launch_part<K, SRC_INDEX, false>(P, 0, in0, n_in, nullptr, dst, mat, scan_scratch, scan_bytes, h, dig, s);
            break;
         default:
            rc = launch_part<K, SRC_TABLE, true>(P, 0, in0, n_in, nullptr, dst, mat, scan_scratch, scan_bytes, h, dig,
                                                 s);
            break;
         }
      } else {
         PartIn<K> in{nullptr, nullptr, src};
         rc = launch_part<K, SRC_TABLE, false>(P, pass, in, 0, &h->n, dst, mat, scan_scratch, scan_bytes, h, dig, s);
      }
      if (rc) return rc;
      src = dst;
   }
   hipLaunchKernelGGL(k_region_starts<K>, dim3(stream_blocks(P.NR + 1, 256)), dim3(256), 0, s, S, start, lb, P.NR,
                      P.sp, P.rlog, h);
   return static_cast<int>(hipGetLastError());
}

// Everything after the first partition pass's input is chosen (see
// partition_regions for src0 and n_in).
template <typename K>
int build_core(const BuildPlan& P, int mode, int src0, PartIn<K> in0, uint64_t n_in, void* buckets, hn_info* info,
               void* scratch, hipStream_t s) {
   char* base = static_cast<char*>(scratch);
   BuildHeader* h = reinterpret_cast<BuildHeader*>(base);
   Slot<K>* S = reinterpret_cast<Slot<K>*>(base + P.off_pairs);
   uint64_t* start = reinterpret_cast<uint64_t*>(base + P.off_start);
   RegionLB lb;
   lb.word = reinterpret_cast<uint64_t*>(base + P.off_word);
   lb.spec = reinterpret_cast<int64_t*>(base + P.off_spec);
   lb.carry = reinterpret_cast<int64_t*>(base + P.off_carry);
   lb.mark = reinterpret_cast<uint32_t*>(base + P.off_mark);
   lb.list = reinterpret_cast<uint32_t*>(base + P.off_list);
   Slot<K>* T = static_cast<Slot<K>*>(buckets);

   hipLaunchKernelGGL(k_build_begin, dim3(1), dim3(64), 0, s, info, h, P.sp, static_cast<uint64_t>(P.n), mode);
   int rc = partition_regions<K>(P, src0, in0, n_in, T, scratch, s);
   if (rc) return rc;
   const uint32_t chunk = region_chunk(P.NR);
   hipLaunchKernelGGL(k_region_build<K>, dim3(static_cast<uint32_t>((P.NR + chunk - 1) / chunk)), dim3(kRegionBS), 0,
                      s, S, start, lb, T, P.sp, P.rlog, P.NR, chunk, h);
   RegionSum* blk = reinterpret_cast<RegionSum*>(base + P.off_blk);
   const uint64_t nblk = (P.NR + kCarryTile - 1) / kCarryTile;
   const int64_t Ri = static_cast<int64_t>(P.R), Ni = static_cast<int64_t>(P.N);
   hipLaunchKernelGGL(k_carry_local, dim3(static_cast<uint32_t>(nblk)), dim3(kCarryBS), 0, s, lb, blk, P.NR, Ri, h);
   hipLaunchKernelGGL(k_carry_top, dim3(1), dim3(1024), 0, s, blk, nblk, P.N, h);
   hipLaunchKernelGGL(k_carry_apply, dim3(static_cast<uint32_t>(nblk)), dim3(kCarryBS), 0, s, lb, blk, P.NR, Ri, h);
   hipLaunchKernelGGL(k_carry_mark, dim3(stream_blocks(P.NR, 256)), dim3(256), 0, s, lb, P.NR, Ri, Ni, h);
   const uint32_t gfix = static_cast<uint32_t>(std::min<uint64_t>(P.NR, static_cast<uint64_t>(cu_count()) * 2));
   hipLaunchKernelGGL(k_region_fix<K>, dim3(gfix), dim3(kRegionBS), 0, s, S, start, lb, T, P.sp, P.rlog, h);
   hipLaunchKernelGGL(k_build_fallback_prep<K>, dim3(stream_blocks(P.N, 256)), dim3(256), 0, s, T, P.N, info, h);
   rc = launch_insert_pairs_dlen<K>(S, T, info, P.n, &h->fb_len, &h->fb_fail, s);
   if (rc) return rc;
   hipLaunchKernelGGL(k_build_end, dim3(1), dim3(64), 0, s, info, h, mode);
   return static_cast<int>(hipGetLastError());
}

template <typename K>
int build(const void* keys, const void* vals, const void* pairs, void* buckets, hn_info* info, size_t len, int sp,
          void* scratch, size_t bytes, hipStream_t s, const hn_tune* t) {
   if (!buckets || !info) return HN_EINVAL;
   if (len == 0) return 0;
   if (!keys && !pairs) return HN_EINVAL;
   const int dcap = (t && t->reserved) ? t->reserved : kMaxDigitBits;
   BuildPlan P;
   if (!build_plan<K>(sp, len, dcap, &P)) return HN_EINVAL;
   if (!scratch || bytes < P.total) return HN_ESCRATCH;
   if (!scratch_aligned(scratch)) return HN_EINVAL;
   PartIn<K> in{static_cast<const K*>(keys), static_cast<const K*>(vals), static_cast<const Slot<K>*>(pairs)};
   const int src0 = pairs ? SRC_PAIRS : (vals ? SRC_SOA : SRC_INDEX);
   return build_core<K>(P, BUILD_EMPTY, src0, in, len, buckets, info, scratch, s);
}

template <typename K>
size_t build_scratch_bytes(int sp, size_t len) {
   BuildPlan P;
   return build_plan<K>(sp, len, kMaxDigitBits, &P) ? P.total : 0;
}

// device_rehash into a larger table as a build: the first partition pass
// reads the old table's live slots (extractPattern + insert in the reference,
// hashinator.h:284-336, without the intermediate array).  The new table is
// Robin Hood ordered.  Growing only: the live pairs then fill at most half of
// the new table.
template <typename K>
bool rehash_plan(int old_sp, int new_sp, BuildPlan* P) {
   if (old_sp < 1 || new_sp <= old_sp || new_sp > std::min(max_size_power<K>(), 32)) return false;
   return build_plan<K>(new_sp, 1ull << old_sp, kMaxDigitBits, P);
}

template <typename K>
size_t rehash_build_scratch_bytes(int old_sp, int new_sp) {
   BuildPlan P;
   return rehash_plan<K>(old_sp, new_sp, &P) ? P.total : 0;
}

template <typename K>
int rehash_build(const void* old_b, int old_sp, void* new_b, int new_sp, hn_info* info, void* scratch, size_t bytes,
                 hipStream_t s) {
   if (!old_b || !new_b || !info) return HN_EINVAL;
   BuildPlan P;
   if (!rehash_plan<K>(old_sp, new_sp, &P)) return HN_EINVAL;
   const char* o0 = static_cast<const char*>(old_b);
   const char* n0 = static_cast<const char*>(new_b);
   if (o0 < n0 + P.N * sizeof(Slot<K>) && n0 < o0 + (1ull << old_sp) * sizeof(Slot<K>)) return HN_EINVAL;
   if (!scratch || bytes < P.total) return HN_ESCRATCH;
   if (!scratch_aligned(scratch)) return HN_EINVAL;
   hipLaunchKernelGGL(k_rehash_info, dim3(1), dim3(64), 0, s, info, new_sp, static_cast<BuildHeader*>(scratch));
   PartIn<K> in{nullptr, nullptr, static_cast<const Slot<K>*>(old_b)};
   return build_core<K>(P, BUILD_REHASH, SRC_TABLE, in, 1ull << old_sp, new_b, info, scratch, s);
}


// ------------------------------------------------------------------ merge insert
//
// Hasher::insert (hashers.h:55-70, kernels_AMD.h insert_kernel) into a table
// that already holds keys, followed by the reorder: the image as a function
// of the live set, written into a second bucket array of the same size.  The
// old table must be free of TOMBSTONEs and in Robin Hood order; then its live
// keys stand in slot order sorted by (home, key), unwrapped from an EMPTY
// slot, and the keys whose home lies in region r fill one run of slots
// [s_r, s_{r+1}): s_r is the first slot at or after rR that is EMPTY or holds
// a key whose home is at or after rR (k_merge_bounds).  The new pairs take
// the build's partition passes.  Then one workgroup per region, in ticket
// order (k_merge_region), loads both, sorts the union by (home offset, key)
// with a home histogram in LDS, keeps only the newest copy of a key (old
// before new, new in batch order: an insert of a present key overwrites its
// value, a key repeated in the batch keeps its last value) and places it:
// p_j = j + max(v_j, carry), v the prefix max of (home offset - j) and carry
// the slots the earlier regions spill into this one, passed along by a
// decoupled look-back over the maps x -> max(A, x + C) (a region: A the
// position of its last key with no carry, C its key count).  A region
// writes its span [carry, max(R, last + 1)) of the output, keys and EMPTY
// slots, so the spans tile the table, and positions past the end wrap into
// region 0's carry-in slots, which region 0 does not write.  Region 0's
// carry comes from that wrap, known only at the end, so region 0 is parked in
// scratch and placed by k_merge_end.  When the table has TOMBSTONEs or no
// order flag, a region exceeds its LDS capacity (3,072 old slots or 1,536 new
// pairs), the old table is not in order after all (a key outside its
// region's run), the wrap would change what region 0 spills into region 1,
// or the union would fill the table, the output is made the other way, on
// the device: the old table copied, then the probing insert of the
// partitioned batch (the partition keeps the batch order of equal keys), so
// the output is hn_insert_*'s result and the order flag is cleared.
constexpr int kMergeBS = 512;
// 2,560 old slots per 2,048-slot region: a region's run is its 2,048 slots
// plus the difference of the spills across its two ends (tens of slots at
// load factor 0.9); a run longer than that sends the call to the fallback.
// Six per thread (3,072) cost 11 spilled VGPRs and took 8.99 ms instead of
// 8.50 ms for 120 M keys into a half-full 2^28 table.
constexpr int kMergeEO = 5; // old slots per thread
constexpr int kMergeEN = 3; // new pairs per thread
constexpr uint32_t kMergeCapO = kMergeBS * kMergeEO;
constexpr uint32_t kMergeCapN = kMergeBS * kMergeEN;
constexpr uint32_t kMergeCap = kMergeCapO + kMergeCapN;
constexpr uint32_t kMergeSpanWords = (kRegionMax + kMergeCap + 31) / 32;
constexpr uint32_t kMgBias = 1u << 16; // keeps home offset - sorted index non-negative
static_assert(kMergeCap < 0x8000, "element ids leave the top bit of a 16-bit order entry free");

struct MergeHeader {
   uint32_t ticket;     // next region
   uint32_t fail;       // a region gave up, or a bound was not found
   uint32_t max_probe;  // longest probe of the keys placed by the regions
   uint32_t n0;         // keys of region 0, parked
   uint32_t ineligible; // TOMBSTONEs or no order flag: the regions do not run
   int32_t fb_fail;     // the fallback insert found no slot for a key
   uint64_t fb_len;     // pairs the fallback insert takes (0: the merge stands)
   uint32_t pad[56];
};
static_assert(sizeof(MergeHeader) == 256, "header is one 256-byte line");

// A parked key of region 0: its prefix-max term and home offset.
template <typename K>
struct MergePark {
   K key, val;
   int32_t v;
   uint32_t hr;
};

// Region state for the look-back: 2 state bits, A + 1 and C in 31 bits each
// (table size power at most 30).
constexpr uint64_t kMgAgg = 1ull << 62, kMgInc = 2ull << 62;
__device__ __forceinline__ uint64_t mg_pack(uint64_t st, RegionSum g) {
   return st | (static_cast<uint64_t>(g.a + 1) << 31) | static_cast<uint64_t>(g.b);
}
__device__ __forceinline__ RegionSum mg_unpack(uint64_t w) {
   RegionSum g;
   g.a = static_cast<int64_t>((w >> 31) & 0x7FFFFFFFull) - 1;
   g.b = static_cast<int64_t>(w & 0x7FFFFFFFull);
   return g;
}
// No keys: (A, C) = (-1, 0) acts as the identity for region maps (A >= C - 1).
constexpr RegionSum kMgNone = {-1, 0};

__global__ void k_merge_begin(const hn_info* info, MergeHeader* mh) {
   if (threadIdx.x == 0) {
      mh->ticket = 0;
      mh->fail = 0;
      mh->max_probe = 0;
      mh->n0 = 0;
      mh->ineligible = (info->tombstone_counter != 0 || !(info->reserved1 & kInfoOrdered)) ? 1u : 0u;
      mh->fb_fail = 0;
      mh->fb_len = 0;
   }
}

// sb[r] = s_r (unwrapped, < rR + kMergeCapO), sb[NR] = N + s_0; clears the
// region states.  One wave per region.
template <typename K>
__global__ __launch_bounds__(256) void k_merge_bounds(const Slot<K>* ob, int sp, int rlog, uint64_t NR, uint64_t* sb,
                                                      uint64_t* status, MergeHeader* mh, const BuildHeader* h) {
   if (!h->go || mh->ineligible) return;
   const int lane = __lane_id();
   const uint64_t N = 1ull << sp, mask = N - 1;
   const uint64_t waves = static_cast<uint64_t>(gridDim.x) * (blockDim.x / kWave);
   for (uint64_t r = static_cast<uint64_t>(blockIdx.x) * (blockDim.x / kWave) + threadIdx.x / kWave; r < NR;
        r += waves) {
      const int64_t lo = static_cast<int64_t>(r << rlog);
      int64_t found = -1;
      for (uint32_t o = 0; o < kMergeCapO; o += kWave) {
         const int64_t p = lo + o + lane;
         const K k = ob[static_cast<uint64_t>(p) & mask].key;
         bool stop = k == empty_key<K>();
         if (!stop) {
            const int64_t d = static_cast<int64_t>((static_cast<uint64_t>(p) - fib_home(k, sp)) & mask);
            stop = p - d >= lo;
         }
         const uint64_t m = __ballot(stop);
         if (m) {
            found = lo + o + first_bit(m);
            break;
         }
      }
      if (lane == 0) {
         if (found < 0) {
            __hip_atomic_store(&mh->fail, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            found = lo;
         }
         sb[r] = static_cast<uint64_t>(found);
         if (r == 0) sb[NR] = N + static_cast<uint64_t>(found);
         status[r] = 0;
      }
   }
}

// Inclusive prefix max of x[0, m) in place by the whole workgroup.
template <int BS>
__device__ void lds_incl_max16(int16_t* x, uint32_t m, int32_t* wmax) {
   const int lane = __lane_id(), wave = threadIdx.x / kWave;
   const uint32_t per = (m + BS - 1) / BS;
   const uint32_t c0 = std::min<uint32_t>(threadIdx.x * per, m), c1 = std::min<uint32_t>(c0 + per, m);
   int32_t acc = INT32_MIN;
   for (uint32_t i = c0; i < c1; ++i) acc = std::max<int32_t>(acc, x[i]);
   int32_t inc = acc;
#pragma unroll
   for (int o = 1; o < kWave; o <<= 1) {
      const int32_t y = __shfl_up(inc, o, kWave);
      if (lane >= o) inc = std::max(inc, y);
   }
   if (lane == kWave - 1) wmax[wave] = inc;
   __syncthreads();
   int32_t run = __shfl_up(inc, 1, kWave);
   if (lane == 0) run = INT32_MIN;
   for (int w = 0; w < wave; ++w) run = std::max(run, wmax[w]);
   for (uint32_t i = c0; i < c1; ++i) {
      run = std::max<int32_t>(run, x[i]);
      x[i] = static_cast<int16_t>(run);
   }
   __syncthreads();
}

template <typename K>
struct MergeLds {
   K key[kMergeCap];        // by element id: old slot offset, or kMergeCapO + new index
   uint16_t order[kMergeCap]; // ids by home offset (top bit: a dropped copy)
   int16_t v[kMergeCap];    // prefix-max terms by sorted index
   uint32_t cnt[kRegionMax + 1];
   // With repeats: the group starts over all copies, kept for the re-ranking
   // of the survivors (in LDS rather than per-element registers: with them in
   // registers the u64 kernel spilled 28 VGPRs and merged 10.18 ms, not 8.99).
   uint32_t cnt0[kRegionMax + 1];
   uint32_t occ[kMergeSpanWords];
   uint32_t wsum[kMergeBS / kWave];
   int32_t wmax[kMergeBS / kWave];
   uint32_t wa[kMergeBS / kWave];  // per wave: max of (home offset - sorted index) + kMgBias
   int64_t lb_a[kMergeBS / kWave]; // look-back: each wave's composed window
   int64_t lb_b[kMergeBS / kWave];
   int32_t lb_inc[kMergeBS / kWave];
   uint32_t ticket;
   uint32_t probe;
   int64_t xin;
};

// Rank of one element in its home group order[g0, g1): keys below it, and
// whether a newer copy of its key (a larger id) drops it.
template <typename K>
__device__ __forceinline__ uint32_t merge_rank(const MergeLds<K>& L, uint32_t id, uint32_t g0, uint32_t g1, bool& drop,
                                               bool& rep) {
   const K key = L.key[id];
   uint32_t rank = 0;
   for (uint32_t k = g0; k < g1; ++k) {
      const uint32_t id2 = L.order[k];
      const K k2 = L.key[id2];
      rank += k2 < key ? 1u : 0u;
      if (k2 == key && id2 != id) {
         rep = true;
         drop = drop || id2 > id;
      }
   }
   return rank;
}
// The same over the copies that stay (dropped ones carry the top bit).
template <typename K>
__device__ __forceinline__ uint32_t merge_rank_kept(const MergeLds<K>& L, uint32_t id, uint32_t g0, uint32_t g1) {
   const K key = L.key[id];
   uint32_t rank = 0;
   for (uint32_t k = g0; k < g1; ++k) {
      const uint32_t e = L.order[k];
      if (!(e & 0x8000u)) rank += L.key[e] < key ? 1u : 0u;
   }
   return rank;
}

// A thread holds kMergeEO old slots (ids threadIdx.x + e * kMergeBS) and
// kMergeEN new pairs (ids kMergeCapO + threadIdx.x + e * kMergeBS); element
// state is packed to keep registers low: hs = home offset << 16 | arrival
// rank within the home (kNoElem: no element), g = group start | end << 16,
// j = sorted index, one drop bit per element.  Keys stay in LDS; values are
// read again from memory when written.
constexpr int kMergeE = kMergeEO + kMergeEN;
constexpr uint32_t kNoElem = 0xFFFFFFFFu;

template <typename K>
__global__ __launch_bounds__(kMergeBS) __attribute__((amdgpu_waves_per_eu(4))) void k_merge_region(const Slot<K>* ob, const Slot<K>* S, const uint64_t* start,
                                                          const uint64_t* sb, uint64_t* status, Slot<K>* nb,
                                                          MergePark<K>* park, MergeHeader* mh, const BuildHeader* h,
                                                          int sp, int rlog, uint64_t NR) {
   __shared__ MergeLds<K> L;
   if (!h->go || mh->ineligible) return; // the same for every workgroup: no look-back waits on a skipped region
   const uint64_t N = 1ull << sp, mask = N - 1, R = 1ull << rlog;
   const int lane = __lane_id(), wave = threadIdx.x / kWave;
   if (threadIdx.x == 0) L.probe = 0;
   uint32_t max_probe = 0;
   auto elem_id = [&](int e) -> uint32_t {
      return e < kMergeEO ? threadIdx.x + e * kMergeBS : kMergeCapO + threadIdx.x + (e - kMergeEO) * kMergeBS;
   };
   for (;;) {
      __syncthreads(); // LDS of the previous region is no longer read
      if (threadIdx.x == 0) L.ticket = atomic_add_relaxed(&mh->ticket, 1u);
      __syncthreads();
      const uint64_t r = L.ticket;
      if (r >= NR) break;
      const uint64_t rb = r << rlog;
      const uint64_t o0 = sb[r], n0 = start[r];
      const uint64_t Lo = sb[r + 1] - o0, Ln = start[r + 1] - n0;
      auto give_up = [&]() { // the call fails; successors still get a state
         if (threadIdx.x == 0) {
            __hip_atomic_store(&mh->fail, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&status[r], mg_pack(kMgInc, kMgNone), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
         }
      };
      if (Lo > kMergeCapO || Ln > kMergeCapN) {
         give_up();
         continue;
      }
      for (uint32_t i = threadIdx.x; i <= R; i += kMergeBS) L.cnt[i] = 0;
      __syncthreads();
      // the value of element e, read again at write time
      auto value = [&](int e) -> K {
         return e < kMergeEO ? ob[(o0 + threadIdx.x + e * kMergeBS) & mask].val
                             : S[n0 + threadIdx.x + (e - kMergeEO) * kMergeBS].val;
      };
      uint32_t hs[kMergeE], jj[kMergeE];
      bool bad = false;
#pragma unroll
      for (int e = 0; e < kMergeE; ++e) {
         hs[e] = kNoElem;
         const uint32_t x = e < kMergeEO ? threadIdx.x + e * kMergeBS : threadIdx.x + (e - kMergeEO) * kMergeBS;
         if (e < kMergeEO ? x < Lo : x < Ln) {
            const K key = e < kMergeEO ? ob[(o0 + x) & mask].key : S[n0 + x].key;
            if (e >= kMergeEO || key != empty_key<K>()) {
               const uint64_t hr = (fib_home(key, sp) - rb) & mask;
               if (hr >= R || (e < kMergeEO && key == tomb_key<K>())) {
                  bad = true;
               } else {
                  L.key[elem_id(e)] = key;
                  hs[e] = static_cast<uint32_t>(hr << 16) | atomicAdd(&L.cnt[hr], 1u);
               }
            }
         }
      }
      if (__syncthreads_or(bad)) {
         give_up();
         continue;
      }
      lds_exclusive_scan<kMergeBS>(L.cnt, static_cast<uint32_t>(R) + 1, L.wsum);
#pragma unroll
      for (int e = 0; e < kMergeE; ++e)
         if (hs[e] != kNoElem) L.order[L.cnt[hs[e] >> 16] + (hs[e] & 0xFFFFu)] = static_cast<uint16_t>(elem_id(e));
      __syncthreads();
      // sorted index: group start + rank by key; a newer copy drops an element
      uint32_t drop = 0;
      bool rep = false;
      uint32_t dmax = 0; // max of (home offset - sorted index) + kMgBias over this thread's elements
#pragma unroll
      for (int e = 0; e < kMergeE; ++e) {
         if (hs[e] != kNoElem) {
            const uint32_t g0 = L.cnt[hs[e] >> 16], g1 = L.cnt[(hs[e] >> 16) + 1];
            bool d = false;
            jj[e] = g0 + merge_rank(L, elem_id(e), g0, g1, d, rep);
            drop |= d ? (1u << e) : 0u;
            dmax = std::max<uint32_t>(dmax, (hs[e] >> 16) + kMgBias - jj[e]);
         }
      }
      dmax = wave_max(dmax);
      if (lane == 0) L.wa[wave] = dmax;
      uint32_t ms;
      bool published = false;
      if (!__syncthreads_or(rep)) {
         ms = L.cnt[R];
         // Every element stays, so the region's map x -> max(A, x + C) is
         // known now: C = ms, A = rb + ms - 1 + max(home offset - sorted
         // index), the last key's position with no carry.  Published before
         // the prefix max, it holds up the regions behind it for less time
         // (10.34 -> 10.15 ms for 120 M keys into a half-full 2^28 table).
         if (threadIdx.x == 0) {
            RegionSum ea = kMgNone;
            if (ms) {
               uint32_t m = 0;
               for (int w = 0; w < kMergeBS / kWave; ++w) m = std::max(m, L.wa[w]);
               ea.a = static_cast<int64_t>(rb) + static_cast<int64_t>(ms - 1) + static_cast<int64_t>(m) - kMgBias;
               ea.b = ms;
            }
            __hip_atomic_store(&status[r], mg_pack(r == 0 ? kMgInc : kMgAgg, ea), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
         }
         published = true;
      } else {
         // mark the dropped copies, then count and rank the kept ones
#pragma unroll
         for (int e = 0; e < kMergeE; ++e)
            if (hs[e] != kNoElem && ((drop >> e) & 1u)) L.order[L.cnt[hs[e] >> 16] + (hs[e] & 0xFFFFu)] |= 0x8000u;
         for (uint32_t i = threadIdx.x; i <= R; i += kMergeBS) L.cnt0[i] = L.cnt[i];
         __syncthreads();
         for (uint32_t i = threadIdx.x; i <= R; i += kMergeBS) L.cnt[i] = 0;
         __syncthreads();
#pragma unroll
         for (int e = 0; e < kMergeE; ++e)
            if (hs[e] != kNoElem && !((drop >> e) & 1u)) atomicAdd(&L.cnt[hs[e] >> 16], 1u);
         __syncthreads();
         ms = lds_exclusive_scan<kMergeBS>(L.cnt, static_cast<uint32_t>(R) + 1, L.wsum);
#pragma unroll
         for (int e = 0; e < kMergeE; ++e)
            if (hs[e] != kNoElem && !((drop >> e) & 1u))
               jj[e] = L.cnt[hs[e] >> 16] +
                       merge_rank_kept(L, elem_id(e), L.cnt0[hs[e] >> 16], L.cnt0[(hs[e] >> 16) + 1]);
      }
#pragma unroll
      for (int e = 0; e < kMergeE; ++e)
         if (hs[e] != kNoElem && !((drop >> e) & 1u))
            L.v[jj[e]] = static_cast<int16_t>(static_cast<int32_t>(hs[e] >> 16) - static_cast<int32_t>(jj[e]));
      __syncthreads();
      lds_incl_max16<kMergeBS>(L.v, ms, L.wmax);
      RegionSum agg = kMgNone;
      if (ms) {
         agg.a = static_cast<int64_t>(rb) + static_cast<int64_t>(ms - 1) + L.v[ms - 1];
         agg.b = ms;
      }
      // look-back: the carry is the position of the last key of regions < r.
      // Every wave reads 64 predecessors, so one round of loads covers the
      // kMergeBS regions in flight (lane i of wave w: region top - 64w - i).
      if (threadIdx.x == 0 && !published)
         __hip_atomic_store(&status[r], mg_pack(r == 0 ? kMgInc : kMgAgg, agg), __ATOMIC_RELAXED,
                            __HIP_MEMORY_SCOPE_AGENT);
      RegionSum excl = kMgNone;
      if (r > 0) {
         int64_t top = static_cast<int64_t>(r) - 1;
         for (;;) {
            const int64_t q = top - static_cast<int64_t>(threadIdx.x);
            uint64_t w = kMgInc;
            if (q >= 0) w = __hip_atomic_load(&status[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            while (__ballot((w >> 62) == 0)) {
               if ((w >> 62) == 0) w = __hip_atomic_load(&status[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            const uint64_t incm = __ballot((w >> 62) == 2);
            const int k = incm ? first_bit(incm) : kWave - 1;
            RegionSum gs = lane <= k ? mg_unpack(w) : kMgNone;
#pragma unroll
            for (int d = 1; d < kWave; d <<= 1) { // lane i + d holds earlier regions than lane i
               RegionSum y;
               y.a = __shfl_down(gs.a, d, kWave);
               y.b = __shfl_down(gs.b, d, kWave);
               if (lane + d < kWave) gs = compose(y, gs);
            }
            if (lane == 0) {
               L.lb_a[wave] = gs.a;
               L.lb_b[wave] = gs.b;
               L.lb_inc[wave] = incm ? 1 : 0;
            }
            __syncthreads();
            // waves in order from the nearest; stop at the first holding an inclusive state
            bool found = false;
            RegionSum win = kMgNone;
            for (int v = 0; v < kMergeBS / kWave && !found; ++v) {
               RegionSum gv;
               gv.a = L.lb_a[v];
               gv.b = L.lb_b[v];
               win = compose(gv, win);
               found = L.lb_inc[v] != 0;
            }
            excl = compose(win, excl);
            __syncthreads();
            if (found) break;
            top -= kMergeBS;
         }
         if (threadIdx.x == 0)
            __hip_atomic_store(&status[r], mg_pack(kMgInc, compose(excl, agg)), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
      }
      if (threadIdx.x == 0) L.xin = excl.a;
      __syncthreads();
      const int64_t carry = L.xin + 1 - static_cast<int64_t>(rb);
      if (r == 0) { // parked until the wrap is known
#pragma unroll
         for (int e = 0; e < kMergeE; ++e)
            if (hs[e] != kNoElem && !((drop >> e) & 1u))
               park[jj[e]] = MergePark<K>{L.key[elem_id(e)], value(e), L.v[jj[e]], hs[e] >> 16};
         if (threadIdx.x == 0) mh->n0 = ms;
         continue;
      }
      const int64_t last = ms ? static_cast<int64_t>(ms - 1) + std::max<int64_t>(L.v[ms - 1], carry) : -1;
      const int64_t c0 = std::max<int64_t>(carry, 0);
      const int64_t span = std::max<int64_t>(static_cast<int64_t>(R), last + 1) - c0;
      const int64_t words = span > 0 ? (span + 31) / 32 : 0;
      for (int64_t w = threadIdx.x; w < words; w += kMergeBS) L.occ[w] = 0;
      __syncthreads();
#pragma unroll
      for (int e = 0; e < kMergeE; ++e) {
         if (hs[e] != kNoElem && !((drop >> e) & 1u)) {
            const int64_t p = static_cast<int64_t>(jj[e]) + std::max<int64_t>(L.v[jj[e]], carry);
            const int64_t q = p - c0;
            atomicOr(&L.occ[q >> 5], 1u << (q & 31));
            Slot<K> sl;
            sl.key = L.key[elem_id(e)];
            sl.val = value(e);
            nb[(rb + static_cast<uint64_t>(p)) & mask] = sl;
            max_probe = std::max<uint32_t>(max_probe, static_cast<uint32_t>(p - static_cast<int64_t>(hs[e] >> 16) + 1));
         }
      }
      __syncthreads();
      for (int64_t q = threadIdx.x; q < span; q += kMergeBS) {
         if (!((L.occ[q >> 5] >> (q & 31)) & 1u)) {
            Slot<K> sl;
            sl.key = empty_key<K>();
            sl.val = 0;
            nb[(rb + static_cast<uint64_t>(c0 + q)) & mask] = sl;
         }
      }
   }
   max_probe = wave_max(max_probe);
   if (lane == 0 && max_probe) atomicMax(&L.probe, max_probe);
   __syncthreads();
   if (threadIdx.x == 0 && L.probe > __hip_atomic_load(&mh->max_probe, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
      atomic_max_relaxed(&mh->max_probe, L.probe);
}

// Region 0 with its carry from the wrap, then the Info; or, when the merge
// does not stand, the fallback's length.  One workgroup.
template <typename K>
__global__ __launch_bounds__(kMergeBS) void k_merge_end(const MergePark<K>* park, const uint64_t* status, Slot<K>* nb,
                                                       hn_info* info, MergeHeader* mh, const BuildHeader* h,
                                                       int sp, int rlog, uint64_t NR) {
   __shared__ uint32_t occ[kMergeSpanWords];
   __shared__ uint32_t probe_s;
   const uint64_t N = 1ull << sp, mask = N - 1, R = 1ull << rlog;
   if (!h->go) { // a size power that is not the table's
      if (threadIdx.x == 0) info->err = HN_STATUS_INVALID;
      return;
   }
   bool fail = mh->ineligible || mh->fail;
   const RegionSum tot = fail ? kMgNone : mg_unpack(status[NR - 1]);
   fail = fail || tot.b >= static_cast<int64_t>(N);
   const uint32_t n0 = fail ? 0 : mh->n0;
   const int64_t c0 = std::max<int64_t>(tot.a + 1 - static_cast<int64_t>(N), 0); // slots the wrap fills in region 0
   const int64_t vlast = n0 ? park[n0 - 1].v : 0;
   const int64_t last = n0 ? static_cast<int64_t>(n0 - 1) + std::max<int64_t>(vlast, c0) : c0 - 1;
   const int64_t spec = n0 ? static_cast<int64_t>(n0 - 1) + vlast : -1; // what region 1 was placed after
   if (NR > 1 && std::max<int64_t>(last + 1 - static_cast<int64_t>(R), 0) !=
                     std::max<int64_t>(spec + 1 - static_cast<int64_t>(R), 0))
      fail = true;
   if (fail) {
      if (threadIdx.x == 0) mh->fb_len = h->n;
      return;
   }
   const int64_t span = std::max<int64_t>(static_cast<int64_t>(R), last + 1) - c0;
   for (int64_t w = threadIdx.x; w < (span + 31) / 32; w += kMergeBS) occ[w] = 0;
   if (threadIdx.x == 0) probe_s = 0;
   __syncthreads();
   uint32_t probe = 0;
   for (uint32_t j = threadIdx.x; j < n0; j += kMergeBS) {
      const MergePark<K> e = park[j];
      const int64_t p = static_cast<int64_t>(j) + std::max<int64_t>(e.v, c0);
      const int64_t q = p - c0;
      atomicOr(&occ[q >> 5], 1u << (q & 31));
      Slot<K> sl;
      sl.key = e.key;
      sl.val = e.val;
      nb[static_cast<uint64_t>(p) & mask] = sl;
      probe = std::max<uint32_t>(probe, static_cast<uint32_t>(p - static_cast<int64_t>(e.hr) + 1));
   }
   __syncthreads();
   for (int64_t q = threadIdx.x; q < span; q += kMergeBS) {
      if (!((occ[q >> 5] >> (q & 31)) & 1u)) {
         Slot<K> sl;
         sl.key = empty_key<K>();
         sl.val = 0;
         nb[static_cast<uint64_t>(c0 + q) & mask] = sl;
      }
   }
   probe = wave_max(probe);
   if (__lane_id() == 0 && probe) atomicMax(&probe_s, probe);
   __syncthreads();
   if (threadIdx.x == 0) {
      const uint64_t mp = std::max<uint32_t>(probe_s, mh->max_probe);
      info->fill = static_cast<uint64_t>(tot.b);
      info->tombstone_counter = 0;
      info->current_max_bucket_overflow = std::max<uint64_t>(kBucketOverflow, next_overflow(mp, kBucketOverflow));
      info->reserved0 = 0;
      info->reserved1 = kInfoOrdered;
      info->err = HN_STATUS_SUCCESS;
   }
}

// Fallback: the old table into the output, then (the probing insert of the
// partitioned batch runs between) err and the flags.
template <typename K>
__global__ __launch_bounds__(256) void k_merge_copy(const Slot<K>* ob, Slot<K>* nb, uint64_t N, const MergeHeader* mh) {
   if (!mh->fb_len) return;
   const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
   for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < N; i += stride) nb[i] = ob[i];
}
__global__ void k_merge_fin(hn_info* info, const MergeHeader* mh) {
   if (threadIdx.x != 0 || !mh->fb_len) return;
   info->err = mh->fb_fail ? HN_STATUS_FAIL : HN_STATUS_SUCCESS;
   info->reserved0 = 0;
   info->reserved1 = 0;
}

template <typename K>
struct MergePlan {
   BuildPlan P;
   size_t off_mh, off_sb, off_status, off_park, total;
};

template <typename K>
bool merge_plan(int sp, uint64_t len, MergePlan<K>* M) {
   if (sp < 1 || sp > 30 || len == 0) return false;
   if (!build_plan<K>(sp, len, kMaxDigitBits, &M->P)) return false;
   size_t o = align256(M->P.total);
   M->off_mh = o;
   o = align256(o + sizeof(MergeHeader));
   M->off_sb = o;
   o = align256(o + (M->P.NR + 1) * sizeof(uint64_t));
   M->off_status = o;
   o = align256(o + M->P.NR * sizeof(uint64_t));
   M->off_park = o;
   o = align256(o + kMergeCap * sizeof(MergePark<K>));
   M->total = o;
   return true;
}

template <typename K>
size_t merge_scratch_bytes(int sp, size_t len) {
   MergePlan<K> M;
   return merge_plan<K>(sp, len, &M) ? M.total : 0;
}

template <typename K>
int merge(const void* keys, const void* vals, const void* pairs, const void* old_b, void* new_b, hn_info* info,
          size_t len, int sp, void* scratch, size_t bytes, hipStream_t s) {
   if (!old_b || !new_b || !info) return HN_EINVAL;
   if (len > 0 && !keys && !pairs) return HN_EINVAL;
   if (sp < 1 || sp > 30) return HN_EINVAL;
   const uint64_t N = 1ull << sp;
   const char* o0 = static_cast<const char*>(old_b);
   const char* n0 = static_cast<const char*>(new_b);
   if (o0 < n0 + N * sizeof(Slot<K>) && n0 < o0 + N * sizeof(Slot<K>)) return HN_EINVAL;
   if (len == 0) // nothing to merge: the output is the table
      return static_cast<int>(hipMemcpyAsync(new_b, old_b, N * sizeof(Slot<K>), hipMemcpyDeviceToDevice, s));
   MergePlan<K> M;
   if (!merge_plan<K>(sp, len, &M)) return HN_EINVAL;
   if (!scratch || bytes < M.total) return HN_ESCRATCH;
   if (!scratch_aligned(scratch)) return HN_EINVAL;
   const BuildPlan& P = M.P;
   char* base = static_cast<char*>(scratch);
   BuildHeader* h = reinterpret_cast<BuildHeader*>(base);
   const Slot<K>* S = reinterpret_cast<const Slot<K>*>(base + P.off_pairs);
   const uint64_t* start = reinterpret_cast<const uint64_t*>(base + P.off_start);
   MergeHeader* mh = reinterpret_cast<MergeHeader*>(base + M.off_mh);
   uint64_t* sb = reinterpret_cast<uint64_t*>(base + M.off_sb);
   uint64_t* status = reinterpret_cast<uint64_t*>(base + M.off_status);
   MergePark<K>* park = reinterpret_cast<MergePark<K>*>(base + M.off_park);
   const Slot<K>* ob = static_cast<const Slot<K>*>(old_b);
   Slot<K>* nb = static_cast<Slot<K>*>(new_b);

   hipLaunchKernelGGL(k_build_begin, dim3(1), dim3(64), 0, s, info, h, sp, static_cast<uint64_t>(len),
                      static_cast<int>(BUILD_MERGE));
   hipLaunchKernelGGL(k_merge_begin, dim3(1), dim3(64), 0, s, info, mh);
   PartIn<K> in{static_cast<const K*>(keys), static_cast<const K*>(vals), static_cast<const Slot<K>*>(pairs)};
   const int src0 = pairs ? SRC_PAIRS : (vals ? SRC_SOA : SRC_INDEX);
   int rc = partition_regions<K>(P, src0, in, len, nb, scratch, s);
   if (rc) return rc;
   hipLaunchKernelGGL(k_merge_bounds<K>, dim3(static_cast<uint32_t>(std::min<uint64_t>((P.NR + 3) / 4, 65536))),
                      dim3(256), 0, s, ob, sp, P.rlog, P.NR, sb, status, mh, h);
   static std::atomic<int> occ{0};
   const int per_cu = resident_blocks_per_cu(k_merge_region<K>, kMergeBS, occ);
   const uint32_t grid =
       static_cast<uint32_t>(std::min<uint64_t>(P.NR, static_cast<uint64_t>(cu_count()) * static_cast<uint64_t>(per_cu)));
   hipLaunchKernelGGL(k_merge_region<K>, dim3(grid), dim3(kMergeBS), 0, s, ob, S, start, sb, status, nb, park, mh, h,
                      sp, P.rlog, P.NR);
   hipLaunchKernelGGL(k_merge_end<K>, dim3(1), dim3(kMergeBS), 0, s, park, status, nb, info, mh, h, sp, P.rlog, P.NR);
   hipLaunchKernelGGL(k_merge_copy<K>, dim3(stream_blocks(N, 256)), dim3(256), 0, s, ob, nb, N, mh);
   rc = launch_insert_pairs_dlen<K>(S, nb, info, len, &mh->fb_len, &mh->fb_fail, s);
   if (rc) return rc;
   hipLaunchKernelGGL(k_merge_fin, dim3(1), dim3(64), 0, s, info, mh);
   return static_cast<int>(hipGetLastError());
}

} // namespace hn

using namespace hn;
extern "C" {
size_t hn_rehash_build_scratch_bytes_u32(int old_sp, int new_sp) {
   return rehash_build_scratch_bytes<uint32_t>(old_sp, new_sp);
}
size_t hn_rehash_build_scratch_bytes_u64(int old_sp, int new_sp) {
   return rehash_build_scratch_bytes<uint64_t>(old_sp, new_sp);
}
int hn_rehash_build_u32(const void* ob, int osp, void* nb, int nsp, hn_info* i, void* sc, size_t bytes, hn_stream_t s) {
   return rehash_build<uint32_t>(ob, osp, nb, nsp, i, sc, bytes, s);
}
int hn_rehash_build_u64(const void* ob, int osp, void* nb, int nsp, hn_info* i, void* sc, size_t bytes, hn_stream_t s) {
   return rehash_build<uint64_t>(ob, osp, nb, nsp, i, sc, bytes, s);
}
size_t hn_build_scratch_bytes_u32(int sp, size_t len) { return build_scratch_bytes<uint32_t>(sp, len); }
size_t hn_build_scratch_bytes_u64(int sp, size_t len) { return build_scratch_bytes<uint64_t>(sp, len); }
int hn_build_u32(const void* k, const void* v, void* b, hn_info* i, size_t n, int sp, void* sc, size_t sb,
                 hn_stream_t s, const hn_tune* t) {
   return build<uint32_t>(k, v, nullptr, b, i, n, sp, sc, sb, s, t);
}
int hn_build_u64(const void* k, const void* v, void* b, hn_info* i, size_t n, int sp, void* sc, size_t sb,
                 hn_stream_t s, const hn_tune* t) {
   return build<uint64_t>(k, v, nullptr, b, i, n, sp, sc, sb, s, t);
}
int hn_build_pairs_u32(const void* p, void* b, hn_info* i, size_t n, int sp, void* sc, size_t sb, hn_stream_t s,
                       const hn_tune* t) {
   return build<uint32_t>(nullptr, nullptr, p, b, i, n, sp, sc, sb, s, t);
}
int hn_build_pairs_u64(const void* p, void* b, hn_info* i, size_t n, int sp, void* sc, size_t sb, hn_stream_t s,
                       const hn_tune* t) {
   return build<uint64_t>(nullptr, nullptr, p, b, i, n, sp, sc, sb, s, t);
}
size_t hn_merge_scratch_bytes_u32(int sp, size_t len) { return merge_scratch_bytes<uint32_t>(sp, len); }
size_t hn_merge_scratch_bytes_u64(int sp, size_t len) { return merge_scratch_bytes<uint64_t>(sp, len); }
int hn_merge_u32(const void* k, const void* v, const void* ob, void* nb, hn_info* i, size_t n, int sp, void* sc,
                 size_t sb, hn_stream_t s) {
   return merge<uint32_t>(k, v, nullptr, ob, nb, i, n, sp, sc, sb, s);
}
int hn_merge_u64(const void* k, const void* v, const void* ob, void* nb, hn_info* i, size_t n, int sp, void* sc,
                 size_t sb, hn_stream_t s) {
   return merge<uint64_t>(k, v, nullptr, ob, nb, i, n, sp, sc, sb, s);
}
int hn_merge_pairs_u32(const void* p, const void* ob, void* nb, hn_info* i, size_t n, int sp, void* sc, size_t sb,
                       hn_stream_t s) {
   return merge<uint32_t>(nullptr, nullptr, p, ob, nb, i, n, sp, sc, sb, s);
}
int hn_merge_pairs_u64(const void* p, const void* ob, void* nb, hn_info* i, size_t n, int sp, void* sc, size_t sb,
                       hn_stream_t s) {
   return merge<uint64_t>(nullptr, nullptr, p, ob, nb, i, n, sp, sc, sb, s);
}
} // extern "C"
