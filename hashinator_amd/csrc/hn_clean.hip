// hn_clean.hip -- clean_tombstones as one pass over EMPTY-bounded segments.
//
// Hashmap::clean_tombstones (hashinator.h:1261-1315) turns TOMBSTONEs into
// EMPTY slots and places the displaced live elements again; the result holds
// the same key -> value set with the occupancy a fresh insert of the live
// set gives.  Followed by the Robin Hood reorder (hn_reorder.hip) it is one
// image, a function of the live set alone: every cluster in (home, key)
// order.  This pass writes that image directly.
//
// Segments.  Tile t owns [b_t, b_{t+1}), b_t the first EMPTY slot at or after
// t*T (unwrapped; the last segment wraps), as in the reorder pass.  This pass
// turns slots EMPTY, so every b_t is found by a kernel of its own before any
// slot changes.  A live
// key's probe path [home, position] crosses no EMPTY slot, so a segment holds
// every live key whose home lies in it.  Inserting a subset of a key set
// never occupies a slot the whole set leaves EMPTY, so the live keys of a
// segment, placed from its leading EMPTY slot, stay inside it: segments are
// independent and each starts with no carry.
//
// Placement.  With the segment's live keys sorted by (relative home r, key)
// as j = 0, 1, ..., linear probing in that order puts key j at
//     p_j = max(r_j, p_{j-1} + 1) = j + max_{k <= j} (r_k - k),
// a prefix-max scan.  A workgroup loads its segment into LDS, sorts by a
// histogram scan over r with ties ranked by key, scans, and writes every slot
// whose content changes (moved keys, and EMPTY where a key or TOMBSTONE
// stood).  Traffic: the table read once, the changed slots written once.
//
// A segment longer than the first pass's LDS capacity (2,048 slots) goes to a
// second pass with 4x the capacity, and one longer still to a third pass that
// keeps the same arrays in a scratch pool in global memory.  A table with no
// EMPTY slot at all has no segment boundary: it goes to the region clean of
// hn_map.hip (whose kernels read the flag left here) and the reorder pass.
#include <algorithm>
#include <atomic>

#include "hn_device.hpp"
#include "hn_host.hpp"

namespace hn {

constexpr uint64_t kSegTile = 1024;

struct SegCleanHeader {
   uint32_t n_over;     // segments for the second pass
   uint32_t n_long;     // segments for the pass over global memory
   uint32_t skip;       // no tombstones: nothing to do
   uint32_t max_probe;  // longest probe (displacement + 1) of the placed keys
   uint64_t tombs_left; // TOMBSTONEs in segments left as they were
   uint64_t pool_used;  // bytes of the long-segment pool handed out
   uint32_t no_empty;   // every slot occupied: no segment boundary exists
   uint32_t failed;     // a segment left as it was (pool too small, or keys off their probe paths)
   uint32_t displaced;  // a live key stood away from its home slot
   uint32_t run_region; // the region clean (and the reorder after it) must run
   uint32_t region_ok;  // the caller's scratch holds a whole-table region clean
   uint32_t ordered;    // the table is in Robin Hood order (HN_INFO_ORDERED)
   uint32_t pad[50];
   // followed by bnd[tiles] (uint64_t), over[tiles] and long_list[tiles] (uint32_t)
};
static_assert(sizeof(SegCleanHeader) == 256, "header is one 256-byte line");

inline uint64_t seg_tile(int sp) { return std::min<uint64_t>(kSegTile, 1ull << sp); }

__host__ __device__ constexpr uint64_t align256(uint64_t b) { return (b + 255) & ~static_cast<uint64_t>(255); }

// Header and per-tile lists.
size_t seg_clean_fixed_bytes(int sp) {
   if (sp < 1 || sp > 32) return 0;
   const uint64_t tiles = (1ull << sp) / seg_tile(sp);
   return align256(sizeof(SegCleanHeader) + tiles * (sizeof(uint64_t) + 2 * sizeof(uint32_t)));
}

// Pool bytes one segment of L slots takes in the pass over global memory:
// its pairs, relative homes, home counts, ranks, the sorted order, the
// prefix-max terms and an occupancy bit map.
template <typename K>
__host__ __device__ inline uint64_t long_seg_bytes(uint64_t L) {
   return align256(L * sizeof(Slot<K>)) + 5 * align256((L + 1) * 4) + align256(((L + 31) / 32) * 4);
}

// The flag the region clean and the reorder read after the segment passes.
const uint32_t* seg_clean_gate(const void* scratch) {
   return &static_cast<const SegCleanHeader*>(scratch)->run_region;
}

__global__ void k_seg_clean_begin(const hn_info* info, SegCleanHeader* h, uint64_t n, uint32_t region_ok) {
   if (threadIdx.x == 0) {
      h->n_over = 0;
      h->n_long = 0;
      h->max_probe = 0;
      h->tombs_left = 0;
      h->pool_used = 0;
      h->failed = 0;
      h->displaced = 0;
      h->run_region = 0;
      h->region_ok = region_ok;
      h->skip = info->tombstone_counter == 0 ? 1u : 0u;
      h->no_empty = (!h->skip && info->fill + info->tombstone_counter >= n) ? 1u : 0u;
      h->ordered = (info->reserved1 & kInfoOrdered) ? 1u : 0u;
   }
}

constexpr uint64_t kSegNoPos = ~0ull;

// bnd[t] = the first EMPTY slot in tile t, or kSegNoPos.  One wave per tile.
template <typename K>
__global__ __launch_bounds__(256) void k_seg_bounds(const Slot<K>* b, uint64_t tile, uint64_t n_tiles,
                                                    const SegCleanHeader* h, uint64_t* bnd) {
   if (h->skip || h->no_empty) return;
   const int lane = __lane_id();
   const uint64_t waves = static_cast<uint64_t>(gridDim.x) * (blockDim.x / kWave);
   for (uint64_t t = static_cast<uint64_t>(blockIdx.x) * (blockDim.x / kWave) + threadIdx.x / kWave; t < n_tiles;
        t += waves) {
      const uint64_t lo = t * tile;
      uint64_t r = kSegNoPos;
      for (uint64_t o = 0; o < tile; o += kWave) {
         const bool e = (o + lane < tile) && b[lo + o + lane].key == empty_key<K>();
         const uint64_t m = __ballot(e);
         if (m) {
            r = lo + o + first_bit(m);
            break;
         }
      }
      if (lane == 0) bnd[t] = r;
   }
}

// Exclusive scan (sum) of x[0, m) in place by the whole workgroup; `wsum`
// holds BS / kWave words.  Returns the total.
template <int BS>
__device__ uint32_t block_excl_sum(uint32_t* x, uint32_t m, uint32_t* wsum) {
   const int lane = __lane_id();
   const int wave = threadIdx.x / kWave;
   const uint32_t per = (m + BS - 1) / BS;
   const uint32_t c0 = std::min<uint32_t>(threadIdx.x * per, m), c1 = std::min<uint32_t>(c0 + per, m);
   uint32_t acc = 0;
   for (uint32_t i = c0; i < c1; ++i) acc += x[i];
   uint32_t inc = acc;
#pragma unroll
   for (int d = 1; d < kWave; d <<= 1) {
      const uint32_t y = __shfl_up(inc, d, kWave);
      if (lane >= d) inc += y;
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

// Inclusive prefix max of x[0, m) in place by the whole workgroup.
template <int BS>
__device__ void block_incl_max(int16_t* x, uint32_t m, int32_t* wmax) {
   const int lane = __lane_id();
   const int wave = threadIdx.x / kWave;
   const uint32_t per = (m + BS - 1) / BS;
   const uint32_t c0 = std::min<uint32_t>(threadIdx.x * per, m), c1 = std::min<uint32_t>(c0 + per, m);
   constexpr int32_t kLow = -(1 << 30);
   int32_t acc = kLow;
   for (uint32_t i = c0; i < c1; ++i) acc = std::max<int32_t>(acc, x[i]);
   int32_t inc = acc;
#pragma unroll
   for (int d = 1; d < kWave; d <<= 1) {
      const int32_t y = __shfl_up(inc, d, kWave);
      if (lane >= d) inc = std::max(inc, y);
   }
   if (lane == kWave - 1) wmax[wave] = inc;
   __syncthreads();
   int32_t run = kLow;
   for (int w = 0; w < wave; ++w) run = std::max(run, wmax[w]);
   const int32_t before = __shfl_up(inc, 1, kWave); // inclusive max of the lanes below
   if (lane > 0) run = std::max(run, before);
   for (uint32_t i = c0; i < c1; ++i) {
      run = std::max<int32_t>(run, x[i]);
      x[i] = static_cast<int16_t>(run);
   }
   __syncthreads();
}

// One workgroup per segment; capacity BS*E slots.  LDS per slot of capacity:
// the key, a home count, the index of the key in (home) order, and the
// prefix-max term (16 bits: relative homes and ranks are below 2^15).
template <typename K, int BS, int E>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(4))) void k_seg_clean(Slot<K>* b, int sp, SegCleanHeader* h, const uint64_t* bnd,
                                                  uint32_t* over, uint32_t* long_list, uint64_t tile, uint32_t n_tiles,
                                                  int second) {
   constexpr uint32_t CAP = BS * E;
   static_assert(CAP <= 16384, "16-bit LDS indices and prefix-max terms");
   __shared__ K key_s[CAP];
   __shared__ uint32_t cnt_s[CAP + 1];
   __shared__ uint16_t order_s[CAP];
   __shared__ int16_t dmax_s[CAP];
   __shared__ uint32_t occ_s[CAP / 32];
   __shared__ uint64_t bounds_s[2];
   __shared__ uint32_t wsum_s[BS / kWave];
   __shared__ int32_t wmax_s[BS / kWave];
   __shared__ uint32_t probe_s;
   __shared__ uint32_t moved_s;
   if (h->skip || h->no_empty || (!second && h->ordered)) return;
   const uint64_t n = 1ull << sp, mask = n - 1;
   const int lane = __lane_id();
   const uint32_t count = second ? h->n_over : n_tiles;
   if (threadIdx.x == 0) {
      probe_s = 0;
      moved_s = 0;
   }
   uint32_t max_probe = 0;
   bool moved_any = false;
   // Wave 0 reads the starts of tiles t .. t + kb (lane j: tile t + j,
   // unwrapped past the table's end), one segment ahead of its use: the
   // segment is [start of t, first later start) unless that is beyond CAP.
   const int wave = threadIdx.x / kWave;
   const uint64_t kb = std::min<uint64_t>(std::min<uint64_t>(CAP / tile + 2, n_tiles + 1), kWave - 1);
   auto starts = [&](uint32_t it) -> uint64_t {
      uint64_t v = kSegNoPos;
      if (it < count) {
         const uint64_t t = second ? over[it] : it;
         const uint64_t u = t + static_cast<uint64_t>(lane);
         if (static_cast<uint64_t>(lane) <= kb) {
            v = bnd[u % n_tiles];
            if (v != kSegNoPos) v += n * (u / n_tiles);
         }
      }
      return v;
   };
   uint64_t nxt = wave == 0 ? starts(blockIdx.x) : kSegNoPos;
   for (uint32_t it = blockIdx.x; it < count; it += gridDim.x) {
      __syncthreads(); // LDS of the previous segment is no longer read
      const uint64_t t = second ? over[it] : it;
      if (wave == 0) {
         const uint64_t cur = nxt;
         nxt = starts(it + gridDim.x);
         const uint64_t b0 = __shfl(cur, 0, kWave);
         const uint64_t later = __ballot(lane >= 1 && cur != kSegNoPos);
         const uint64_t b1 = (b0 == kSegNoPos || !later) ? kSegNoPos : __shfl(cur, first_bit(later), kWave);
         if (lane == 0) {
            bounds_s[0] = b0;
            bounds_s[1] = b1;
         }
      }
      __syncthreads();
      const uint64_t b0 = bounds_s[0], b1 = bounds_s[1];
      if (b0 == kSegNoPos) continue; // no EMPTY in this tile: another tile's segment
      if (b1 == kSegNoPos || b1 - b0 > CAP) {
         if (threadIdx.x == 0) {
            uint32_t* list = second ? long_list : over;
            list[atomic_add_relaxed(second ? &h->n_long : &h->n_over, 1u)] = static_cast<uint32_t>(t);
         }
         continue;
      }
      const uint32_t len = static_cast<uint32_t>(b1 - b0);
      for (uint32_t i = threadIdx.x; i <= len; i += BS) cnt_s[i] = 0;
      for (uint32_t i = threadIdx.x; i < (len + 31) / 32; i += BS) occ_s[i] = 0;
      __syncthreads();

      // live keys: relative home r (1 <= r <= i on a valid table) and their
      // arrival rank among the keys of the same home
      K kk[E], vv[E];
      uint32_t rr[E], sub[E];
      uint32_t tombs = 0;
      bool moved = false;
#pragma unroll
      for (int e = 0; e < E; ++e) {
         const uint32_t i = threadIdx.x + e * BS;
         kk[e] = empty_key<K>();
         if (i < len) {
            const Slot<K> s = b[(b0 + i) & mask];
            key_s[i] = s.key;
            if (s.key == tomb_key<K>()) {
               ++tombs;
            } else if (s.key != empty_key<K>()) {
               kk[e] = s.key;
               vv[e] = s.val;
               const uint32_t r = static_cast<uint32_t>(std::min<uint64_t>((fib_home(s.key, sp) - b0) & mask, i));
               rr[e] = r;
               moved = moved || r != i;
               sub[e] = atomicAdd(&cnt_s[r], 1u);
            }
         }
      }
      __syncthreads();
      block_excl_sum<BS>(cnt_s, len + 1, wsum_s); // cnt_s[r] = live keys with a smaller home
#pragma unroll
      for (int e = 0; e < E; ++e) {
         const uint32_t i = threadIdx.x + e * BS;
         if (i < len && kk[e] != empty_key<K>()) order_s[cnt_s[rr[e]] + sub[e]] = static_cast<uint16_t>(i);
      }
      __syncthreads();
      // sorted index j = (keys of smaller homes) + (same-home keys with smaller keys);
      // the prefix-max term of j is r_j - j
      uint32_t jj[E];
#pragma unroll
      for (int e = 0; e < E; ++e) {
         const uint32_t i = threadIdx.x + e * BS;
         if (i < len && kk[e] != empty_key<K>()) {
            const uint32_t s0 = cnt_s[rr[e]], s1 = cnt_s[rr[e] + 1];
            uint32_t tie = 0;
            for (uint32_t j = s0; j < s1; ++j) tie += key_s[order_s[j]] < kk[e] ? 1u : 0u;
            jj[e] = s0 + tie;
            dmax_s[jj[e]] = static_cast<int16_t>(static_cast<int32_t>(rr[e]) - static_cast<int32_t>(jj[e]));
         }
      }
      const uint32_t live = cnt_s[len];
      __syncthreads();
      block_incl_max<BS>(dmax_s, live, wmax_s);
      // p_j = j + max_{k <= j}(r_k - k).  A destination past the segment only
      // arises on a table whose keys are not where probing puts them (EMPTY
      // holes inside probe paths): the segment is then left as it is.
      bool out_of_range = false;
#pragma unroll
      for (int e = 0; e < E; ++e) {
         const uint32_t i = threadIdx.x + e * BS;
         if (i < len && kk[e] != empty_key<K>())
            out_of_range = out_of_range || jj[e] + static_cast<uint32_t>(static_cast<int32_t>(dmax_s[jj[e]])) >= len;
      }
      if (__syncthreads_or(out_of_range)) {
         const uint64_t t_all = wave_sum(static_cast<uint64_t>(tombs)); // this segment keeps its TOMBSTONEs
         if (__lane_id() == 0 && t_all) atomic_add_relaxed(&h->tombs_left, t_all);
         if (threadIdx.x == 0) __hip_atomic_store(&h->failed, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
         continue;
      }
      moved_any = moved_any || moved;
      // place, marking the slots taken
#pragma unroll
      for (int e = 0; e < E; ++e) {
         const uint32_t i = threadIdx.x + e * BS;
         if (i < len && kk[e] != empty_key<K>()) {
            const uint32_t dest = jj[e] + static_cast<uint32_t>(static_cast<int32_t>(dmax_s[jj[e]]));
            atomicOr(&occ_s[dest >> 5], 1u << (dest & 31));
            if (dest != i || key_s[i] != kk[e]) {
               Slot<K> s;
               s.key = kk[e];
               s.val = vv[e];
               b[(b0 + dest) & mask] = s;
            }
            max_probe = std::max<uint32_t>(max_probe, dest - rr[e] + 1);
         }
      }
      __syncthreads();
      // slots left without a key become EMPTY (TOMBSTONEs, and slots keys left)
#pragma unroll
      for (int e = 0; e < E; ++e) {
         const uint32_t i = threadIdx.x + e * BS;
         if (i < len && key_s[i] != empty_key<K>() && !((occ_s[i >> 5] >> (i & 31)) & 1u)) {
            Slot<K> s;
            s.key = empty_key<K>();
            s.val = 0;
            b[(b0 + i) & mask] = s;
         }
      }
   }
   // one update of the header per workgroup, and only when it changes something
   max_probe = wave_max(max_probe);
   __syncthreads();
   if (lane == 0) {
      if (max_probe) atomicMax(&probe_s, max_probe);
      if (moved_any) moved_s = 1;
   }
   __syncthreads();
   if (threadIdx.x == 0) {
      if (probe_s > __hip_atomic_load(&h->max_probe, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
         atomic_max_relaxed(&h->max_probe, probe_s);
      if (moved_s && !__hip_atomic_load(&h->displaced, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
         __hip_atomic_store(&h->displaced, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
   }
}

// Ordered tables.  In Robin Hood order every segment's live keys already
// stand in (relative home, key) order, and erasing keys keeps that, so the
// placement needs no sort: p_j = max(r_j, p_{j-1} + 1) runs over the live
// keys in slot order.  It is a scan of the maps x -> max(a, x + c) under
// composition, each live key contributing (r, 1); `m`, the largest relative
// home so far, checks the order (a key whose r is below an earlier key's
// sends the segment to the sorting pass, as does a home outside the segment
// or a destination past it).
constexpr int32_t kNegPos = -(1 << 30);
struct Place {
   int32_t a, c, m;
};
__device__ __forceinline__ Place place_join(Place l, Place r) { // l: the slots before r
   return Place{std::max(r.a, l.a + r.c), l.c + r.c, std::max(l.m, r.m)};
}
template <int CTRL, int ROWS>
__device__ __forceinline__ Place place_dpp(Place x) {
   Place y;
   y.a = __builtin_amdgcn_update_dpp(kNegPos, x.a, CTRL, ROWS, 0xf, false);
   y.c = __builtin_amdgcn_update_dpp(0, x.c, CTRL, ROWS, 0xf, false);
   y.m = __builtin_amdgcn_update_dpp(kNegPos, x.m, CTRL, ROWS, 0xf, false);
   return place_join(y, x);
}
// Inclusive scan over the wave's lanes: row shifts inside 16-lane rows, then
// row broadcasts; lanes a step does not reach keep the identity.
__device__ __forceinline__ Place place_wave_scan(Place x) {
   x = place_dpp<0x111, 0xf>(x); // row_shr:1
   x = place_dpp<0x112, 0xf>(x); // row_shr:2
   x = place_dpp<0x114, 0xf>(x); // row_shr:4
   x = place_dpp<0x118, 0xf>(x); // row_shr:8
   x = place_dpp<0x142, 0xa>(x); // row_bcast:15 -> rows 1, 3
   x = place_dpp<0x143, 0xc>(x); // row_bcast:31 -> rows 2, 3
   return x;
}
__device__ __forceinline__ uint64_t readlane64(uint64_t v, int l) {
   const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(static_cast<uint32_t>(v)), l));
   const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(static_cast<uint32_t>(v >> 32)), l));
   return (static_cast<uint64_t>(hi) << 32) | lo;
}

// One workgroup per segment at a time, capacity BS*E slots in rows of 64:
// row q = e * (BS / 64) + wave holds slots [64q, 64q + 64), so short segments
// keep every wave busy and each row is one coalesced read.  Every wave finds
// the segment's bounds itself, one segment ahead; the occupancy bit maps
// alternate between two buffers, so a segment takes three barriers.
template <typename K, int BS, int E>
__global__ __launch_bounds__(BS) void k_seg_clean_ord(Slot<K>* b, int sp, SegCleanHeader* h, const uint64_t* bnd,
                                                      uint32_t* over, uint64_t tile, uint32_t n_tiles) {
   constexpr int NW = BS / kWave;
   constexpr uint32_t CAP = BS * E;
   constexpr int NR = NW * E;
   static_assert(NR <= kWave, "one lane per row for the row prefix");
   __shared__ uint32_t occ_s[2][CAP / 32];
   __shared__ int32_t agg_s[3][NR];
   __shared__ uint32_t probe_s;
   __shared__ uint32_t moved_s;
   if (h->skip || h->no_empty || !h->ordered) return;
   const uint64_t n = 1ull << sp, mask = n - 1;
   const int lane = __lane_id();
   const int wave = threadIdx.x / kWave;
   if (threadIdx.x == 0) {
      probe_s = 0;
      moved_s = 0;
   }
   uint32_t max_probe = 0;
   bool moved_any = false;
   int par = 0;
   // lane j: the start of tile t + j (unwrapped past the table's end)
   const uint64_t kb = std::min<uint64_t>(std::min<uint64_t>(CAP / tile + 2, n_tiles + 1), kWave - 1);
   auto starts = [&](uint32_t it) -> uint64_t {
      uint64_t v = kSegNoPos;
      if (it < n_tiles && static_cast<uint64_t>(lane) <= kb) {
         const uint64_t u = static_cast<uint64_t>(it) + static_cast<uint64_t>(lane);
         v = bnd[u % n_tiles];
         if (v != kSegNoPos) v += n * (u / n_tiles);
      }
      return v;
   };
   uint64_t nxt = starts(blockIdx.x);
   for (uint32_t it = blockIdx.x; it < n_tiles; it += gridDim.x) {
      const uint64_t cur = nxt;
      nxt = starts(it + gridDim.x);
      const uint64_t b0 = readlane64(cur, 0);
      const uint64_t later = __ballot(lane >= 1 && cur != kSegNoPos);
      if (b0 == kSegNoPos) continue; // no EMPTY in this tile: another tile's segment
      const uint64_t b1 = later ? readlane64(cur, first_bit(later)) : kSegNoPos;
      if (b1 == kSegNoPos || b1 - b0 > CAP) {
         if (threadIdx.x == 0) over[atomic_add_relaxed(&h->n_over, 1u)] = it;
         continue;
      }
      const uint32_t len = static_cast<uint32_t>(b1 - b0);
      uint32_t* occ = occ_s[par];
      par ^= 1;
      for (uint32_t i = threadIdx.x; i < (len + 31) / 32; i += BS) occ[i] = 0;
      K kk[E], vv[E];
      int32_t rr[E];
      Place loc[E];
      bool bad = false;
#pragma unroll
      for (int e = 0; e < E; ++e) {
         const uint32_t q = static_cast<uint32_t>(e * NW + wave);
         const uint32_t i = q * kWave + static_cast<uint32_t>(lane);
         kk[e] = empty_key<K>();
         rr[e] = -1;
         Place x{kNegPos, 0, kNegPos};
         if (q * kWave < len) {
            if (i < len) {
               const Slot<K> s = b[(b0 + i) & mask];
               kk[e] = s.key;
               vv[e] = s.val;
               if (s.key != empty_key<K>() && s.key != tomb_key<K>()) {
                  const uint64_t rh = (fib_home(s.key, sp) - b0) & mask;
                  bad = bad || rh > i;
                  rr[e] = static_cast<int32_t>(std::min<uint64_t>(rh, i));
                  x = Place{rr[e], 1, rr[e]};
               }
            }
            x = place_wave_scan(x);
         }
         if (lane == kWave - 1) {
            agg_s[0][q] = x.a;
            agg_s[1][q] = x.c;
            agg_s[2][q] = x.m;
         }
         loc[e] = x;
      }
      __syncthreads();
      // the rows' aggregates scanned by every wave (lane q: row q)
      Place g{kNegPos, 0, kNegPos};
      if (lane < NR) g = Place{agg_s[0][lane], agg_s[1][lane], agg_s[2][lane]};
      g = place_wave_scan(g);
      int32_t pos[E];
#pragma unroll
      for (int e = 0; e < E; ++e) {
         const int q = e * NW + wave;
         pos[e] = -1;
         if (rr[e] >= 0) {
            Place pre{kNegPos, 0, kNegPos};
            if (q > 0)
               pre = Place{__builtin_amdgcn_readlane(g.a, q - 1), __builtin_amdgcn_readlane(g.c, q - 1),
                           __builtin_amdgcn_readlane(g.m, q - 1)};
            const Place f = place_join(pre, loc[e]);
            pos[e] = f.a;
            bad = bad || f.m != rr[e] || f.a >= static_cast<int32_t>(len);
         }
      }
      if (__syncthreads_or(bad)) { // the sorting pass takes it
         if (threadIdx.x == 0) over[atomic_add_relaxed(&h->n_over, 1u)] = it;
         continue;
      }
#pragma unroll
      for (int e = 0; e < E; ++e) {
         if (rr[e] >= 0) {
            const uint32_t i = static_cast<uint32_t>((e * NW + wave) * kWave + lane);
            const uint32_t p = static_cast<uint32_t>(pos[e]);
            atomicOr(&occ[p >> 5], 1u << (p & 31));
            if (p != i) {
               Slot<K> s;
               s.key = kk[e];
               s.val = vv[e];
               b[(b0 + p) & mask] = s;
            }
            max_probe = std::max<uint32_t>(max_probe, p - static_cast<uint32_t>(rr[e]) + 1);
            moved_any = moved_any || static_cast<uint32_t>(rr[e]) != i;
         }
      }
      __syncthreads();
      // slots left without a key become EMPTY (TOMBSTONEs, and slots keys left)
#pragma unroll
      for (int e = 0; e < E; ++e) {
         const uint32_t i = static_cast<uint32_t>((e * NW + wave) * kWave + lane);
         if (i < len && kk[e] != empty_key<K>() && !((occ[i >> 5] >> (i & 31)) & 1u)) {
            Slot<K> s;
            s.key = empty_key<K>();
            s.val = 0;
            b[(b0 + i) & mask] = s;
         }
      }
   }
   // one update of the header per workgroup, and only when it changes something
   max_probe = wave_max(max_probe);
   __syncthreads();
   if (lane == 0) {
      if (max_probe) atomicMax(&probe_s, max_probe);
      if (moved_any) moved_s = 1;
   }
   __syncthreads();
   if (threadIdx.x == 0) {
      if (probe_s > __hip_atomic_load(&h->max_probe, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
         atomic_max_relaxed(&h->max_probe, probe_s);
      if (moved_s && !__hip_atomic_load(&h->displaced, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
         __hip_atomic_store(&h->displaced, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
   }
}

// Segments too long for LDS: the same algorithm with its arrays in a pool of
// global memory, one workgroup per segment.  A segment the pool cannot hold
// is left as it is (failed).
constexpr int kLongBS = 1024;

// Exclusive sum of x[0, m) in place by the workgroup, in chunks; returns the total.
__device__ uint32_t block_excl_sum_g(uint32_t* x, uint64_t m, uint32_t* wsum, uint32_t* carry) {
   constexpr int BS = kLongBS;
   const int lane = __lane_id();
   const int wave = threadIdx.x / kWave;
   if (threadIdx.x == 0) *carry = 0;
   __syncthreads();
   for (uint64_t c = 0; c < m; c += BS) {
      const uint64_t i = c + threadIdx.x;
      const uint32_t v = i < m ? x[i] : 0u;
      uint32_t inc = v;
#pragma unroll
      for (int d = 1; d < kWave; d <<= 1) {
         const uint32_t y = __shfl_up(inc, d, kWave);
         if (lane >= d) inc += y;
      }
      if (lane == kWave - 1) wsum[wave] = inc;
      __syncthreads();
      uint32_t run = *carry + inc - v, total = 0;
      for (int w = 0; w < BS / kWave; ++w) {
         if (w < wave) run += wsum[w];
         total += wsum[w];
      }
      if (i < m) x[i] = run;
      __syncthreads();
      if (threadIdx.x == 0) *carry += total;
      __syncthreads();
   }
   return *carry;
}

// Inclusive prefix max of x[0, m) in place by the workgroup, in chunks.
__device__ void block_incl_max_g(int32_t* x, uint64_t m, int32_t* wmax, int32_t* carry) {
   constexpr int BS = kLongBS;
   const int lane = __lane_id();
   const int wave = threadIdx.x / kWave;
   constexpr int32_t kLow = -(1 << 30);
   if (threadIdx.x == 0) *carry = kLow;
   __syncthreads();
   for (uint64_t c = 0; c < m; c += BS) {
      const uint64_t i = c + threadIdx.x;
      const int32_t v = i < m ? x[i] : kLow;
      int32_t inc = v;
#pragma unroll
      for (int d = 1; d < kWave; d <<= 1) {
         const int32_t y = __shfl_up(inc, d, kWave);
         if (lane >= d) inc = std::max(inc, y);
      }
      if (lane == kWave - 1) wmax[wave] = inc;
      __syncthreads();
      int32_t run = std::max(*carry, inc), total = kLow;
      for (int w = 0; w < BS / kWave; ++w) {
         if (w < wave) run = std::max(run, wmax[w]);
         total = std::max(total, wmax[w]);
      }
      if (i < m) x[i] = run;
      __syncthreads();
      if (threadIdx.x == 0) *carry = std::max(*carry, total);
      __syncthreads();
   }
}

template <typename K>
__global__ __launch_bounds__(kLongBS) void k_seg_clean_long(Slot<K>* b, int sp, SegCleanHeader* h, const uint64_t* bnd,
                                                          const uint32_t* long_list, uint64_t tile, uint32_t n_tiles,
                                                          char* pool, uint64_t pool_bytes) {
   constexpr int BS = kLongBS;
   __shared__ uint64_t b0_s, len_s, off_s;
   __shared__ uint32_t wsum_s[BS / kWave];
   __shared__ int32_t wmax_s[BS / kWave];
   __shared__ uint32_t carry_s;
   __shared__ int32_t cmax_s;
   __shared__ uint32_t probe_s;
   __shared__ uint32_t moved_s;
   if (h->skip || h->no_empty) return;
   const uint64_t n = 1ull << sp, mask = n - 1;
   const uint32_t count = h->n_long;
   if (threadIdx.x == 0) {
      probe_s = 0;
      moved_s = 0;
   }
   uint32_t max_probe = 0;
   bool moved_any = false;
   for (uint32_t it = blockIdx.x; it < count; it += gridDim.x) {
      __syncthreads();
      if (threadIdx.x == 0) {
         const uint32_t t = long_list[it];
         const uint64_t b0 = bnd[t];
         uint64_t b1 = b0 + n; // the next tile with an EMPTY slot, a full cycle at most
         for (uint64_t u = t + 1; u < static_cast<uint64_t>(t) + n_tiles; ++u) {
            const uint64_t bu = bnd[u < n_tiles ? u : u - n_tiles];
            if (bu != kSegNoPos) {
               b1 = u < n_tiles ? bu : bu + n;
               break;
            }
         }
         const uint64_t L = b1 - b0;
         const uint64_t need = long_seg_bytes<K>(L);
         const uint64_t off = atomic_add_relaxed(&h->pool_used, need);
         b0_s = b0;
         len_s = L;
         off_s = off + need <= pool_bytes ? off : kSegNoPos;
         if (off_s == kSegNoPos) __hip_atomic_store(&h->failed, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      __syncthreads();
      if (off_s == kSegNoPos) { // left as it is: count its TOMBSTONEs
         uint64_t t = 0;
         for (uint64_t i = threadIdx.x; i < len_s; i += BS) t += b[(b0_s + i) & mask].key == tomb_key<K>() ? 1 : 0;
         t = wave_sum(t);
         if (__lane_id() == 0 && t) atomic_add_relaxed(&h->tombs_left, t);
         continue;
      }
      const uint64_t b0 = b0_s, L = len_s;
      char* p = pool + off_s;
      Slot<K>* kv = reinterpret_cast<Slot<K>*>(p);
      p += align256(L * sizeof(Slot<K>));
      uint32_t* rr = reinterpret_cast<uint32_t*>(p);
      p += align256((L + 1) * 4);
      uint32_t* cnt = reinterpret_cast<uint32_t*>(p);
      p += align256((L + 1) * 4);
      uint32_t* jx = reinterpret_cast<uint32_t*>(p);
      p += align256((L + 1) * 4);
      uint32_t* ord = reinterpret_cast<uint32_t*>(p);
      p += align256((L + 1) * 4);
      int32_t* dm = reinterpret_cast<int32_t*>(p);
      p += align256((L + 1) * 4);
      uint32_t* occ = reinterpret_cast<uint32_t*>(p);
      for (uint64_t i = threadIdx.x; i <= L; i += BS) cnt[i] = 0;
      for (uint64_t i = threadIdx.x; i < (L + 31) / 32; i += BS) occ[i] = 0;
      __syncthreads();
      uint32_t tombs = 0;
      bool moved = false;
      for (uint64_t i = threadIdx.x; i < L; i += BS) {
         const Slot<K> s = b[(b0 + i) & mask];
         kv[i] = s;
         if (s.key == tomb_key<K>()) {
            ++tombs;
         } else if (s.key != empty_key<K>()) {
            const uint64_t r = std::min<uint64_t>((fib_home(s.key, sp) - b0) & mask, i);
            rr[i] = static_cast<uint32_t>(r);
            moved = moved || r != i;
            jx[i] = __hip_atomic_fetch_add(&cnt[r], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
         }
      }
      __syncthreads();
      const uint32_t live = block_excl_sum_g(cnt, L + 1, wsum_s, &carry_s);
      for (uint64_t i = threadIdx.x; i < L; i += BS) {
         const K k = kv[i].key;
         if (k != empty_key<K>() && k != tomb_key<K>()) ord[cnt[rr[i]] + jx[i]] = static_cast<uint32_t>(i);
      }
      __syncthreads();
      for (uint64_t i = threadIdx.x; i < L; i += BS) {
         const K k = kv[i].key;
         if (k != empty_key<K>() && k != tomb_key<K>()) {
            const uint32_t s0 = cnt[rr[i]], s1 = cnt[rr[i] + 1];
            uint32_t tie = 0;
            for (uint32_t j = s0; j < s1; ++j) tie += kv[ord[j]].key < k ? 1u : 0u;
            const uint32_t j = s0 + tie;
            jx[i] = j;
            dm[j] = static_cast<int32_t>(rr[i]) - static_cast<int32_t>(j);
         }
      }
      __syncthreads();
      block_incl_max_g(dm, live, wmax_s, &cmax_s);
      bool out_of_range = false;
      for (uint64_t i = threadIdx.x; i < L; i += BS) {
         const K k = kv[i].key;
         if (k != empty_key<K>() && k != tomb_key<K>())
            out_of_range = out_of_range || static_cast<uint64_t>(jx[i] + dm[jx[i]]) >= L;
      }
      if (__syncthreads_or(out_of_range)) {
         const uint64_t t_all = wave_sum(static_cast<uint64_t>(tombs)); // this segment keeps its TOMBSTONEs
         if (__lane_id() == 0 && t_all) atomic_add_relaxed(&h->tombs_left, t_all);
         if (threadIdx.x == 0) __hip_atomic_store(&h->failed, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
         continue;
      }
      moved_any = moved_any || moved;
      for (uint64_t i = threadIdx.x; i < L; i += BS) {
         const Slot<K> s = kv[i];
         if (s.key != empty_key<K>() && s.key != tomb_key<K>()) {
            const uint64_t dest = static_cast<uint64_t>(jx[i] + dm[jx[i]]);
            __hip_atomic_fetch_or(&occ[dest >> 5], 1u << (dest & 31), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (dest != i) b[(b0 + dest) & mask] = s;
            max_probe = std::max<uint32_t>(max_probe, static_cast<uint32_t>(dest - rr[i] + 1));
         }
      }
      __syncthreads();
      for (uint64_t i = threadIdx.x; i < L; i += BS) {
         if (kv[i].key != empty_key<K>() && !((occ[i >> 5] >> (i & 31)) & 1u)) {
            Slot<K> s;
            s.key = empty_key<K>();
            s.val = 0;
            b[(b0 + i) & mask] = s;
         }
      }
   }
   max_probe = wave_max(max_probe);
   __syncthreads();
   if (__lane_id() == 0) {
      if (max_probe) atomicMax(&probe_s, max_probe);
      if (moved_any) moved_s = 1;
   }
   __syncthreads();
   if (threadIdx.x == 0) {
      if (probe_s) atomic_max_relaxed(&h->max_probe, probe_s);
      if (moved_s) __hip_atomic_store(&h->displaced, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
   }
}

// Info after the passes.  Every segment cleaned: tombstones 0, the overflow
// of the new layout, Robin Hood order, err = SUCCESS when a live key was
// placed again (the reference's reinsert sets it; with nothing displaced it
// is left as it was).  A segment left as it was keeps its probes (the
// overflow only grows) and its TOMBSTONEs (err = FAIL while any remain).  A
// table with no EMPTY slot goes to the region clean when the scratch holds
// the whole table, else err = FAIL and nothing changes.
__global__ void k_seg_clean_end(hn_info* info, SegCleanHeader* h) {
   if (threadIdx.x != 0 || h->skip) return;
   if (h->no_empty) {
      if (h->region_ok)
         h->run_region = 1;
      else
         info->err = HN_STATUS_FAIL;
      return;
   }
   const uint64_t left = h->failed ? h->tombs_left : 0;
   const uint64_t o = std::max<uint64_t>(kBucketOverflow, next_overflow(h->max_probe, kBucketOverflow));
   info->tombstone_counter = left;
   info->reserved0 = 0;
   if (h->failed) {
      if (o > info->current_max_bucket_overflow) info->current_max_bucket_overflow = o;
      info->reserved1 = 0;
      if (left) info->err = HN_STATUS_FAIL;
      else if (h->displaced) info->err = HN_STATUS_SUCCESS;
   } else {
      info->current_max_bucket_overflow = o;
      info->reserved1 = kInfoOrdered;
      if (h->displaced) info->err = HN_STATUS_SUCCESS;
   }
}

// First pass: segments of up to 2,048 slots, 512 threads x 4 (the register
// arrays of 8 elements per thread held two waves per SIMD for u64 keys), the
// ordered kernel or the sorting one; second pass (sorting): up to 8,192 slots.
// The ordered kernel holds 2 slots per thread (58 VGPRs for u64 keys, 8
// waves per SIMD).
constexpr int kSegBS1 = 512, kSegE1 = 4, kSegBS2 = 1024, kSegE2 = 8;
constexpr int kOrdBS = 1024, kOrdE = 2;
constexpr size_t kSegPoolMin = 64 << 10;

// The segment passes.  `scratch` holds seg_clean_fixed_bytes(sp) bytes;
// `pool` (pool_bytes) takes the segments too long for LDS.  `region_ok`: the
// caller will run the region clean over the whole table when the table has
// no EMPTY slot (it reads seg_clean_gate(scratch)).
template <typename K>
int seg_clean(void* buckets, hn_info* info, int sp, void* scratch, void* pool, size_t pool_bytes, bool region_ok,
              hipStream_t s) {
   const uint64_t n = 1ull << sp, tile = seg_tile(sp), tiles = n / tile;
   SegCleanHeader* h = static_cast<SegCleanHeader*>(scratch);
   uint64_t* bnd = reinterpret_cast<uint64_t*>(h + 1);
   uint32_t* over = reinterpret_cast<uint32_t*>(bnd + tiles);
   uint32_t* long_list = over + tiles;
   Slot<K>* b = static_cast<Slot<K>*>(buckets);
   hipLaunchKernelGGL(k_seg_clean_begin, dim3(1), dim3(64), 0, s, info, h, n, region_ok ? 1u : 0u);
   hipLaunchKernelGGL(k_seg_bounds<K>, dim3(static_cast<uint32_t>(std::min<uint64_t>((tiles + 3) / 4, 65536))), dim3(256),
                      0, s, b, tile, tiles, h, bnd);
   // as many workgroups as stay resident, each looping over segments: the
   // header then sees one update per workgroup, not one per segment
   static std::atomic<int> occ0{0}, occ1{0};
   const int per_cu0 = resident_blocks_per_cu(k_seg_clean_ord<K, kOrdBS, kOrdE>, kOrdBS, occ0);
   const uint32_t grid0 = static_cast<uint32_t>(std::min<uint64_t>(tiles, static_cast<uint64_t>(cu_count()) * per_cu0));
   hipLaunchKernelGGL((k_seg_clean_ord<K, kOrdBS, kOrdE>), dim3(grid0), dim3(kOrdBS), 0, s, b, sp, h, bnd, over, tile,
                      static_cast<uint32_t>(tiles));
   const int per_cu = resident_blocks_per_cu(k_seg_clean<K, kSegBS1, kSegE1>, kSegBS1, occ1);
   const uint32_t grid1 = static_cast<uint32_t>(std::min<uint64_t>(tiles, static_cast<uint64_t>(cu_count()) * per_cu));
   hipLaunchKernelGGL((k_seg_clean<K, kSegBS1, kSegE1>), dim3(grid1), dim3(kSegBS1), 0, s, b, sp, h, bnd, over, long_list,
                      tile, static_cast<uint32_t>(tiles), 0);
   const uint32_t grid2 = static_cast<uint32_t>(std::min<uint64_t>(tiles, static_cast<uint64_t>(cu_count())));
   hipLaunchKernelGGL((k_seg_clean<K, kSegBS2, kSegE2>), dim3(grid2), dim3(kSegBS2), 0, s, b, sp, h, bnd, over,
                      long_list, tile, static_cast<uint32_t>(tiles), 1);
   hipLaunchKernelGGL(k_seg_clean_long<K>, dim3(grid2), dim3(kLongBS), 0, s, b, sp, h, bnd, long_list, tile,
                      static_cast<uint32_t>(tiles), static_cast<char*>(pool), static_cast<uint64_t>(pool_bytes));
   hipLaunchKernelGGL(k_seg_clean_end, dim3(1), dim3(64), 0, s, info, h);
   return static_cast<int>(hipGetLastError());
}
template int seg_clean<uint32_t>(void*, hn_info*, int, void*, void*, size_t, bool, hipStream_t);
template int seg_clean<uint64_t>(void*, hn_info*, int, void*, void*, size_t, bool, hipStream_t);

// Default pool: room for long segments covering 1/32 of the table (1 MiB at least).
template <typename K>
size_t seg_clean_default_pool(int sp) {
   const uint64_t n = 1ull << sp;
   return std::max<size_t>(1 << 20, long_seg_bytes<K>(n / 32));
}
// A pool no table exhausts: one segment over the whole table.
template <typename K>
size_t seg_clean_full_pool(int sp) {
   return long_seg_bytes<K>(1ull << sp);
}
template size_t seg_clean_default_pool<uint32_t>(int);
template size_t seg_clean_default_pool<uint64_t>(int);
template size_t seg_clean_full_pool<uint32_t>(int);
template size_t seg_clean_full_pool<uint64_t>(int);
size_t seg_clean_min_pool() { return kSegPoolMin; }

} // namespace hn
