// hn_reorder.hip -- Robin Hood reorder pass over a linear-probing table.
//
// The accelerated insert (kernels_AMD.h insert_kernel) leaves each cluster in
// the order the keys happened to arrive.  This pass rewrites every cluster (a
// maximal run of non-EMPTY slots) so that its elements stand in order of home
// slot, ties broken by key: the layout linear probing produces when the keys
// arrive in sorted-home order.  The live set, the occupancy of every slot,
// fill and tombstone count are unchanged; the longest probe cannot grow
// (within a run of fixed slots, home order minimises the largest
// displacement), and currentMaxBucketOverflow is recomputed from the new
// layout (at LF 0.9 it falls from ~700 to ~64).  In that order a lookup may stop at the first occupant whose
// own displacement is smaller than the probe offset: every slot after it holds
// keys whose home lies after the query's.  The lookup kernels do so when the
// table's Info carries the ordered bit, which this pass sets and every insert
// clears.
//
// Work split.  Tile t owns the segment [b_t, b_{t+1}) where b_t is the first
// EMPTY slot at or after t*T (positions unwrapped, b_{N/T} = b_0 + N), so
// segments are whole clusters and tile the table.  A workgroup stages its
// segment's keys in LDS, computes each element's rank from a packed
// histogram-and-occupancy scan, and writes only the elements that move.  A
// segment longer than the first pass's LDS capacity goes to a list that a
// second pass with 4x the capacity works through; one longer than that is
// left as it is and the table is then not marked ordered.
#include <algorithm>

#include "hn_device.hpp"
#include "hn_host.hpp"

namespace hn {

constexpr uint64_t kReorderTile = 1024;

struct ReorderHeader {
   uint32_t n_over;    // segments handed to the second pass
   uint32_t fail;      // a segment too long for the second pass
   uint32_t skip;      // tombstones present or no EMPTY slot: nothing is done
   uint32_t max_probe; // longest probe (displacement + 1) over the reordered segments
   uint32_t pad[60];
};
static_assert(sizeof(ReorderHeader) == 256, "header is one 256-byte line");

inline uint64_t reorder_tile(int sp) { return std::min<uint64_t>(kReorderTile, 1ull << sp); }

// `gate` (nullable): run only when *gate is nonzero (a flag another pass leaves).
__global__ void k_reorder_begin(const hn_info* info, ReorderHeader* h, uint64_t n, const uint32_t* gate) {
   if (threadIdx.x == 0) {
      h->n_over = 0;
      h->fail = 0;
      h->max_probe = 0;
      h->skip = ((gate && *gate == 0) || info->tombstone_counter != 0 || info->fill >= n ||
                 (info->reserved1 & kInfoOrdered))
                    ? 1u
                    : 0u; // an ordered table is its own reorder
   }
}

// Unwrapped position of the first EMPTY slot in [from, from + limit), or kNoPos.
// Called by a whole wave.
constexpr uint64_t kNoPos = ~0ull;
template <typename K>
__device__ uint64_t first_empty(const Slot<K>* b, uint64_t mask, uint64_t from, uint64_t limit) {
   const int lane = __lane_id();
   for (uint64_t o = 0; o < limit; o += kWave) {
      const bool e = (o + lane < limit) && b[(from + o + lane) & mask].key == empty_key<K>();
      const uint64_t m = __ballot(e);
      if (m) return from + o + first_bit(m);
   }
   return kNoPos;
}

// One workgroup per segment.  LDS per slot of capacity: the key, a packed
// (histogram << 16 | occupied) scan word, and two 16-bit index tables.
template <typename K, int BS, int E>
__global__ __launch_bounds__(BS) void k_reorder(Slot<K>* b, int sp, hn_info* info, ReorderHeader* h,
                                                uint32_t* over, uint64_t tile, uint32_t n_tiles, int second) {
   constexpr uint32_t CAP = BS * E;
   static_assert(CAP < 65536, "16-bit LDS indices and packed scan halves");
   __shared__ K key_s[CAP];
   __shared__ uint32_t pre_s[CAP + 1];
   __shared__ uint16_t order_s[CAP];
   __shared__ uint16_t slot_of_s[CAP];
   __shared__ uint64_t bounds_s[2];
   __shared__ uint32_t wsum_s[BS / kWave];
   __shared__ uint32_t probe_s;
   if (h->skip) return;
   const uint64_t n = 1ull << sp, mask = n - 1;
   const int lane = __lane_id();
   const int wave = threadIdx.x / kWave;
   const uint32_t count = second ? h->n_over : n_tiles;
   if (threadIdx.x == 0) probe_s = 0;
   uint32_t max_probe = 0;
   for (uint32_t it = blockIdx.x; it < count; it += gridDim.x) {
      __syncthreads(); // LDS of the previous segment is no longer read
      const uint64_t t = second ? over[it] : it;
      const uint64_t lo = t * tile;
      if (wave < 2) {
         const uint64_t r = wave == 0 ? first_empty<K>(b, mask, lo, tile) : first_empty<K>(b, mask, lo + tile, CAP);
         if (lane == 0) bounds_s[wave] = r;
      }
      __syncthreads();
      const uint64_t b0 = bounds_s[0], b1 = bounds_s[1];
      // No EMPTY slot in this tile: the segment is empty (b_t == b_{t+1}).
      if (b0 == kNoPos || b0 >= lo + tile) continue;
      if (b1 == kNoPos || b1 - b0 > CAP) {
         if (threadIdx.x == 0) {
            if (second) {
               __hip_atomic_store(&h->fail, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            } else {
               const uint32_t k = atomic_add_relaxed(&h->n_over, 1u);
               over[k] = static_cast<uint32_t>(t);
            }
         }
         continue;
      }
      const uint32_t len = static_cast<uint32_t>(b1 - b0);
      for (uint32_t i = threadIdx.x; i <= len; i += BS) pre_s[i] = 0;
      __syncthreads();

      K kk[E], vv[E];
      uint32_t rr[E], sub[E];
#pragma unroll
      for (int e = 0; e < E; ++e) {
         const uint32_t i = threadIdx.x + e * BS;
         kk[e] = empty_key<K>();
         if (i < len) {
            const Slot<K> s = b[(b0 + i) & mask];
            kk[e] = s.key;
            vv[e] = s.val;
            key_s[i] = s.key;
            if (s.key != empty_key<K>()) {
               // Home relative to the segment start; a valid table has 1 <= r <= i.
               uint32_t r = static_cast<uint32_t>(std::min<uint64_t>((fib_home(s.key, sp) - b0) & mask, i));
               rr[e] = r;
               sub[e] = atomicAdd(&pre_s[r], 1u << 16) >> 16;
               atomicAdd(&pre_s[i], 1u);
            }
         }
      }
      __syncthreads();

      // Exclusive scan of pre_s[0..len]: high half = elements with a smaller
      // relative home, low half = occupied slots before.
      {
         const uint32_t m = len + 1;
         const uint32_t per = (m + BS - 1) / BS;
         const uint32_t c0 = std::min<uint32_t>(threadIdx.x * per, m), c1 = std::min<uint32_t>(c0 + per, m);
         uint32_t acc = 0;
         for (uint32_t x = c0; x < c1; ++x) acc += pre_s[x];
         uint32_t inc = acc;
#pragma unroll
         for (int d = 1; d < kWave; d <<= 1) {
            const uint32_t y = __shfl_up(inc, d, kWave);
            if (lane >= d) inc += y;
         }
         if (lane == kWave - 1) wsum_s[wave] = inc;
         __syncthreads();
         uint32_t run = inc - acc;
         for (int w = 0; w < wave; ++w) run += wsum_s[w];
         for (uint32_t x = c0; x < c1; ++x) {
            const uint32_t v = pre_s[x];
            pre_s[x] = run;
            run += v;
         }
      }
      __syncthreads();

#pragma unroll
      for (int e = 0; e < E; ++e) {
         const uint32_t i = threadIdx.x + e * BS;
         if (i < len && kk[e] != empty_key<K>()) {
            slot_of_s[pre_s[i] & 0xFFFFu] = static_cast<uint16_t>(i);
            order_s[(pre_s[rr[e]] >> 16) + sub[e]] = static_cast<uint16_t>(i);
         }
      }
      __syncthreads();

#pragma unroll
      for (int e = 0; e < E; ++e) {
         const uint32_t i = threadIdx.x + e * BS;
         if (i < len && kk[e] != empty_key<K>()) {
            const uint32_t s0 = pre_s[rr[e]] >> 16, s1 = pre_s[rr[e] + 1] >> 16;
            uint32_t tie = 0;
            for (uint32_t j = s0; j < s1; ++j) tie += key_s[order_s[j]] < kk[e] ? 1u : 0u;
            const uint32_t dest = slot_of_s[s0 + tie];
            if (dest != i) {
               Slot<K> s;
               s.key = kk[e];
               s.val = vv[e];
               b[(b0 + dest) & mask] = s;
            }
            if (dest >= rr[e]) max_probe = std::max<uint32_t>(max_probe, dest - rr[e] + 1);
         }
      }
   }
   // One atomic per block at most, and only while the running maximum grows.
   max_probe = wave_max(max_probe);
   __syncthreads();
   if (lane == 0 && max_probe) atomicMax(&probe_s, max_probe);
   __syncthreads();
   if (threadIdx.x == 0 && probe_s > __hip_atomic_load(&h->max_probe, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
      atomic_max_relaxed(&h->max_probe, probe_s);
}

// Every segment reordered: the overflow is the new layout's own.  Otherwise
// the segments left as they were keep their probes, so it can only grow.
__global__ void k_reorder_end(hn_info* info, const ReorderHeader* h) {
   if (threadIdx.x != 0 || h->skip) return;
   const uint64_t o = std::max<uint64_t>(kBucketOverflow, next_overflow(h->max_probe, kBucketOverflow));
   if (h->fail) {
      if (o > info->current_max_bucket_overflow) info->current_max_bucket_overflow = o;
   } else {
      info->current_max_bucket_overflow = o;
      info->reserved1 |= kInfoOrdered;
   }
}

size_t reorder_scratch_bytes(int sp) {
   if (sp < 1 || sp > 63) return 0;
   const uint64_t tiles = (1ull << sp) / reorder_tile(sp);
   return sizeof(ReorderHeader) + tiles * sizeof(uint32_t);
}

constexpr int kReorderBS1 = 256, kReorderBS2 = 1024, kReorderE = 8;

template <typename K>
int reorder(void* buckets, hn_info* info, int sp, void* scratch, size_t bytes, hipStream_t s,
            const uint32_t* gate = nullptr) {
   if (!buckets || !info || !scratch) return HN_EINVAL;
   if (sp < 1 || sp > max_size_power<K>()) return HN_EINVAL;
   if (bytes < reorder_scratch_bytes(sp)) return HN_ESCRATCH;
   const uint64_t n = 1ull << sp, tile = reorder_tile(sp), tiles = n / tile;
   if (tiles > 0xFFFFFFFFull) return HN_EINVAL;
   ReorderHeader* h = static_cast<ReorderHeader*>(scratch);
   uint32_t* over = reinterpret_cast<uint32_t*>(h + 1);
   Slot<K>* b = static_cast<Slot<K>*>(buckets);
   hipLaunchKernelGGL(k_reorder_begin, dim3(1), dim3(64), 0, s, info, h, n, gate);
   hipLaunchKernelGGL((k_reorder<K, kReorderBS1, kReorderE>), dim3(static_cast<uint32_t>(tiles)), dim3(kReorderBS1), 0,
                      s, b, sp, info, h, over, tile, static_cast<uint32_t>(tiles), 0);
   const uint32_t grid2 = static_cast<uint32_t>(std::min<uint64_t>(tiles, static_cast<uint64_t>(cu_count())));
   hipLaunchKernelGGL((k_reorder<K, kReorderBS2, kReorderE>), dim3(grid2), dim3(kReorderBS2), 0, s, b, sp, info, h,
                      over, tile, static_cast<uint32_t>(tiles), 1);
   hipLaunchKernelGGL(k_reorder_end, dim3(1), dim3(64), 0, s, info, h);
   return static_cast<int>(hipGetLastError());
}

// The reorder after the region clean of clean_tombstones, run only when the
// segment pass of hn_clean.hip left the table to that path.
template <typename K>
int reorder_gated(void* buckets, hn_info* info, int sp, void* scratch, size_t bytes, const uint32_t* gate,
                  hipStream_t s) {
   return reorder<K>(buckets, info, sp, scratch, bytes, s, gate);
}
template int reorder_gated<uint32_t>(void*, hn_info*, int, void*, size_t, const uint32_t*, hipStream_t);
template int reorder_gated<uint64_t>(void*, hn_info*, int, void*, size_t, const uint32_t*, hipStream_t);

} // namespace hn

using namespace hn;
extern "C" {
size_t hn_reorder_scratch_bytes(int sp) { return reorder_scratch_bytes(sp); }
int hn_reorder_u32(void* b, hn_info* i, int sp, void* sc, size_t sb, hn_stream_t s) {
   return reorder<uint32_t>(b, i, sp, sc, sb, s);
}
int hn_reorder_u64(void* b, hn_info* i, int sp, void* sc, size_t sb, hn_stream_t s) {
   return reorder<uint64_t>(b, i, sp, sc, sb, s);
}
} // extern "C"
