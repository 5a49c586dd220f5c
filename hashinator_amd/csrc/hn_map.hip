// hn_map.hip -- gfx950 kernels and C-ABI for the hash-map operations.
//
// The probe loops themselves (probe groups, aligned windows, block I/O) are
// in hn_probe.hpp; this file holds the single-map kernels around them, the
// full-table passes (clear, clean_tombstones, rehash) and the C-ABI.
//
// Reference: kstppd/hashinator include/hashinator/kernels_AMD.h (insert_kernel
// 272-405, insert_index_kernel 507-641, delete_kernel 412-505, retrieve_kernel
// 647-752, reset_all_to_empty 26-43, reset_to_empty 65-134) and the host
// wrappers in hashers.h / hashinator.h cited at each entry point.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>

#include "hn_device.hpp"
#include "hn_host.hpp"
#include "hn_probe.hpp"

namespace hn {

// The segment pass of clean_tombstones (hn_clean.hip) and the reorder that
// follows the region path when the segment pass hands a table to it
// (hn_reorder.hip).
size_t seg_clean_fixed_bytes(int sp);
template <typename K>
size_t seg_clean_default_pool(int sp);
template <typename K>
size_t seg_clean_full_pool(int sp);
size_t seg_clean_min_pool();
const uint32_t* seg_clean_gate(const void* scratch);
template <typename K>
int seg_clean(void* buckets, hn_info* info, int sp, void* scratch, void* pool, size_t pool_bytes, bool region_ok,
              hipStream_t s);
size_t reorder_scratch_bytes(int sp);
template <typename K>
int reorder_gated(void* buckets, hn_info* info, int sp, void* scratch, size_t bytes, const uint32_t* gate,
                  hipStream_t s);

// insert_kernel / insert_index_kernel (kernels_AMD.h:136-405, 507-641)
template <typename K, int G, int SRC>
__global__ __launch_bounds__(1024) void k_insert(InsertArgs<K> a) {
   const int sp = a.info->size_power;
   const bool sp_ok = sp >= 1 && sp <= max_size_power<K>();
   const uint64_t len = !sp_ok ? 0 : (a.d_len ? *a.d_len : a.len);
   const uint64_t gstride = static_cast<uint64_t>(gridDim.x) * (blockDim.x / G);
   const uint64_t blk = (static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / G;
   InsertCounters c;
   c.failed = sp_ok ? 0 : 1;
   if (sp_ok) insert_range<K, G, SRC>(a, sp, len, blk, gstride, c);
   if (a.fail_sink) {
      const uint64_t placed = wave_sum(c.placed);
      const uint64_t max_probe = wave_max(c.max_probe);
      const int failed = wave_max(c.failed);
      if (__lane_id() == 0) {
         if (placed) atomic_add_relaxed(&a.info->fill, placed);
         if (max_probe)
            atomic_max_relaxed(&a.info->current_max_bucket_overflow, next_overflow(max_probe, kBucketOverflow));
         if (failed) __hip_atomic_fetch_or(a.fail_sink, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
   } else {
      insert_fold(a.info, c.placed, c.max_probe, c.failed);
   }
}

// After an insert kernel: err = FAIL if a key found no slot (or the fill is
// not the expected one), else SUCCESS; reserved words cleared (an insert does
// not keep Robin Hood order).
__global__ void k_insert_finish(hn_info* info, const uint64_t* expected_fill) {
   int32_t f = info->reserved1 & kInfoFail;
   if (expected_fill && info->fill != *expected_fill) f = 1;
   info->err = f ? HN_STATUS_FAIL : HN_STATUS_SUCCESS;
   info->reserved0 = 0;
   info->reserved1 = 0;
}

// retrieve_kernel (kernels_AMD.h:647-752), delete_kernel (412-505), reset_to_empty (65-134)
template <typename K, int G, bool PAIRS, int MODE>
__global__ __launch_bounds__(1024) void k_lookup(LookupArgs<K> a) {
   const int sp = a.info->size_power;
   const bool sp_ok = sp >= 1 && sp <= max_size_power<K>();
   const uint64_t gstride = static_cast<uint64_t>(gridDim.x) * (blockDim.x / G);
   const uint64_t blk = (static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / G;
   uint64_t erased = 0;
   if (sp_ok)
      erased = lookup_range<K, G, PAIRS, MODE>(a, sp, a.info->current_max_bucket_overflow, a.len, blk, gstride,
                                               (a.info->reserved1 & kInfoOrdered) != 0);
   if constexpr (MODE == LK_ERASE) {
      erased = wave_sum(erased);
      if (__lane_id() == 0 && erased) {
         atomic_add_relaxed(&a.info->tombstone_counter, erased);
         __hip_atomic_fetch_sub(&a.info->fill, erased, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
   }
   if constexpr (MODE == LK_RESET) {
      if (blockIdx.x == 0 && threadIdx.x == 0)
         __hip_atomic_fetch_sub(&a.info->fill, a.len, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
   }
}

// Lane-window lookups (one key per lane, see lane_lookup): one wave per
// block (no block-level cooperation; single-wave blocks pack the LDS of a CU
// best), each with its own staging and its own contiguous share of the
// queries.  Tables smaller than one 128-byte window (and a zero probe bound)
// take the group loop with one lane per key.
constexpr int kLaneBlock = 64;
template <typename K, bool PAIRS, int MODE>
__global__ __launch_bounds__(kLaneBlock) void k_lookup_lane(LookupArgs<K> a) {
   constexpr int kWaves = kLaneBlock / kWave;
   constexpr int kRing = MODE == LK_RETRIEVE ? kLaneRing : 1;
   __shared__ __align__(16) char win_s[kWaves][kLaneWaveBytes];
   __shared__ uint32_t wb_s[kWaves][kWave];
   __shared__ K ring_v[kWaves][kRing];
   __shared__ uint8_t ring_s[kWaves][kRing];
   const int sp = a.info->size_power;
   const bool sp_ok = sp >= 1 && sp <= max_size_power<K>();
   const uint64_t ovf = a.info->current_max_bucket_overflow;
   const bool ordered = (a.info->reserved1 & kInfoOrdered) != 0;
   uint64_t erased = 0;
   if (sp_ok) {
      if (sp >= LaneWin<K>::kLog && ovf > 0) {
         const int wv = threadIdx.x / kWave;
         const uint64_t nwaves = static_cast<uint64_t>(gridDim.x) * (kLaneBlock / kWave);
         const uint64_t w = static_cast<uint64_t>(blockIdx.x) * (kLaneBlock / kWave) + wv;
         const uint64_t per = a.len / nwaves, rem = a.len % nwaves;
         const uint64_t q0 = w * per + (w < rem ? w : rem);
         const uint64_t q1 = q0 + per + (w < rem ? 1 : 0);
         erased = lane_lookup<K, PAIRS, MODE>(a, sp, ovf, ordered, q0, q1, win_s[wv], wb_s[wv], ring_v[wv], ring_s[wv]);
      } else {
         const uint64_t gstride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
         const uint64_t blk = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
         erased = lookup_range<K, 1, PAIRS, MODE>(a, sp, ovf, a.len, blk, gstride, ordered);
      }
   }
   if constexpr (MODE == LK_ERASE) {
      erased = wave_sum(erased);
      if (__lane_id() == 0 && erased) {
         atomic_add_relaxed(&a.info->tombstone_counter, erased);
         __hip_atomic_fetch_sub(&a.info->fill, erased, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
   }
}

// reset_all_to_empty (kernels_AMD.h:26-43): keys -> EMPTY, values untouched.
template <typename K>
__global__ __launch_bounds__(256) void k_reset_all(Slot<K>* b, hn_info* info, uint64_t len) {
   const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
   for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < len; i += stride) {
      Slot<K> s = b[i];
      s.key = empty_key<K>();
      b[i] = s;
   }
   if (blockIdx.x == 0 && threadIdx.x == 0) info->fill = 0;
}

// Fresh table: every slot {EMPTYBUCKET, V()} (splitvec.h:233-238).
template <typename K>
__global__ __launch_bounds__(256) void k_table_init(Slot<K>* b, uint64_t n) {
   const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
   for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
      Slot<K> s;
      s.key = empty_key<K>();
      s.val = 0;
      b[i] = s;
   }
}

__global__ void k_info_init(hn_info* info, int sp) {
   info->size_power = sp;
   info->reserved0 = 0;
   info->fill = 0;
   info->current_max_bucket_overflow = kBucketOverflow;
   info->tombstone_counter = 0;
   info->err = HN_STATUS_INVALID;
   info->reserved1 = 0;
}

// ------------------------------------------------------------ clean_tombstones

struct CleanHeader {
   uint64_t start;   // first slot of region 0 (an EMPTY slot), ~0 = none found yet
   uint64_t total;   // elements lifted out over all regions
   int32_t skip;     // no tombstones (or no usable start): do nothing
   int32_t fail;     // a reinsertion failed
   int32_t no_empty; // no EMPTY slot and scratch too small for the whole table
   int32_t pad;
   uint64_t counters[1]; // [n_regions]
};

constexpr uint64_t kNone = ~0ull;

// `gate` (nullable): run only when *gate is nonzero (the segment pass's flag).
__global__ void k_clean_begin(const hn_info* info, CleanHeader* h, uint32_t n_regions, const uint32_t* gate) {
   for (uint32_t r = threadIdx.x; r < n_regions; r += blockDim.x) h->counters[r] = 0;
   if (threadIdx.x == 0) {
      h->start = kNone;
      h->total = 0;
      h->skip = ((gate && *gate == 0) || info->tombstone_counter == 0) ? 1 : 0;
      h->fail = 0;
      h->no_empty = 0;
   }
}

// Any EMPTY slot will do as the start; the lowest one among the blocks that
// look is taken.  Blocks past an already-found slot stop at once.
template <typename K>
__global__ __launch_bounds__(256) void k_find_empty(const Slot<K>* b, uint64_t n, CleanHeader* h) {
   if (h->skip) return;
   const uint64_t chunk = static_cast<uint64_t>(blockDim.x) * 16;
   const uint64_t lo = static_cast<uint64_t>(blockIdx.x) * chunk;
   if (__hip_atomic_load(&h->start, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < lo) return;
   uint64_t best = kNone;
   for (uint64_t i = lo + threadIdx.x; i < lo + chunk && i < n; i += blockDim.x) {
      if (b[i].key == empty_key<K>()) {
         best = i;
         break;
      }
   }
   if (best != kNone) atomic_min_u64(&h->start, best);
}

__global__ void k_clean_commit(hn_info* info, CleanHeader* h, uint64_t n, uint64_t region_slots) {
   if (h->skip) return;
   if (h->start == kNone) {
      if (region_slots < n) { // cannot order the regions: leave the table as it is
         h->skip = 1;
         h->no_empty = 1;
         return;
      }
      h->start = 0;
   }
   info->tombstone_counter = 0;
   info->current_max_bucket_overflow = kBucketOverflow;
   info->reserved1 = 0; // the reinsertion does not keep Robin Hood order
}

// One region: TOMBSTONE -> EMPTY; a live element away from its home slot is
// copied to scratch and its slot emptied (the isOverflown rule at
// hashinator.h:1285-1297 followed by reset_to_empty).
template <typename K>
__global__ __launch_bounds__(256) void k_clean_extract(Slot<K>* b, int sp, hn_info* info, CleanHeader* h,
                                                        uint32_t region, uint64_t region_slots, Slot<K>* out) {
   if (h->skip) return;
   const uint64_t n = 1ull << sp, mask = n - 1;
   const uint64_t r0 = static_cast<uint64_t>(region) * region_slots;
   if (r0 >= n) return;
   const uint64_t rlen = std::min(region_slots, n - r0);
   const uint64_t start = h->start;
   const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
   const uint64_t iters = (rlen + stride - 1) / stride;
   uint64_t lifted = 0;
   for (uint64_t it = 0; it < iters; ++it) {
      const uint64_t i = it * stride + static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
      bool take = false;
      Slot<K> s{};
      uint64_t slot = 0;
      if (i < rlen) {
         slot = (start + r0 + i) & mask;
         s = b[slot];
         if (s.key == tomb_key<K>()) {
            Slot<K> e = s;
            e.key = empty_key<K>();
            b[slot] = e;
         } else if (s.key != empty_key<K>() && (fib_home(s.key, sp) & mask) != slot) {
            take = true;
         }
      }
      const uint64_t m = __ballot(take);
      if (m) {
         uint64_t wbase = 0;
         const int leader = first_bit(m);
         if (__lane_id() == leader) wbase = atomic_add_relaxed(&h->counters[region], static_cast<uint64_t>(__popcll(m)));
         wbase = __shfl(wbase, leader, kWave);
         if (take) {
            const uint64_t rank = __popcll(m & ((1ull << __lane_id()) - 1ull));
            out[wbase + rank] = s;
            Slot<K> e = s;
            e.key = empty_key<K>();
            b[slot] = e;
         }
         lifted += take ? 1 : 0;
      }
   }
   lifted = wave_sum(lifted);
   if (__lane_id() == 0 && lifted) {
      __hip_atomic_fetch_sub(&info->fill, lifted, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      atomic_add_relaxed(&h->total, lifted);
   }
}

__global__ void k_clean_end(hn_info* info, const CleanHeader* h) {
   if (h->no_empty) {
      info->err = HN_STATUS_FAIL;
      return;
   }
   if (h->skip || h->total == 0) return;
   info->err = h->fail ? HN_STATUS_FAIL : HN_STATUS_SUCCESS;
}

// ------------------------------------------------------------ device_rehash

struct RehashHeader {
   uint64_t prior_fill;
};

template <typename K>
__global__ __launch_bounds__(256) void k_rehash_begin(Slot<K>* nb, uint64_t n, hn_info* info, int new_sp,
                                                       RehashHeader* h) {
   const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
   for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
      Slot<K> s;
      s.key = empty_key<K>();
      s.val = 0;
      nb[i] = s;
   }
   if (blockIdx.x == 0 && threadIdx.x == 0) {
      h->prior_fill = info->fill;
      info->size_power = new_sp;
      info->reserved0 = 0;
      info->fill = 0;
      info->current_max_bucket_overflow = kBucketOverflow;
      info->tombstone_counter = 0;
      info->err = HN_STATUS_INVALID;
      info->reserved1 = 0;
   }
}

// ------------------------------------------------------------ launch helpers

template <typename K, int SRC>
using InsertKernel = void (*)(InsertArgs<K>);

template <typename K, int SRC>
InsertKernel<K, SRC> pick_insert(int G) {
   switch (G) {
   case 1: return k_insert<K, 1, SRC>;
   case 2: return k_insert<K, 2, SRC>;
   case 4: return k_insert<K, 4, SRC>;
   case 8: return k_insert<K, 8, SRC>;
   case 16: return k_insert<K, 16, SRC>;
   case 32: return k_insert<K, 32, SRC>;
   case 64: return k_insert<K, 64, SRC>;
   default: return nullptr;
   }
}

template <typename K, bool PAIRS, int MODE>
using LookupKernel = void (*)(LookupArgs<K>);

template <typename K, bool PAIRS, int MODE>
LookupKernel<K, PAIRS, MODE> pick_lookup(int G) {
   switch (G) {
   case 1: return k_lookup<K, 1, PAIRS, MODE>;
   case 2: return k_lookup<K, 2, PAIRS, MODE>;
   case 4: return k_lookup<K, 4, PAIRS, MODE>;
   case 8: return k_lookup<K, 8, PAIRS, MODE>;
   case 16: return k_lookup<K, 16, PAIRS, MODE>;
   case 32: return k_lookup<K, 32, PAIRS, MODE>;
   case 64: return k_lookup<K, 64, PAIRS, MODE>;
   default: return nullptr;
   }
}

template <typename K, int SRC>
int launch_insert(InsertArgs<K> a, uint64_t grid_len, hipStream_t s, const hn_tune* t, OpKind op) {
   Launch L;
   int rc = resolve_launch<K>(t, op, grid_len, &L);
   if (rc) return rc;
   InsertKernel<K, SRC> k = pick_insert<K, SRC>(L.probe_width);
   if (!k) return HN_EINVAL;
   hipLaunchKernelGGL(k, dim3(L.blocks), dim3(L.block_size), 0, s, a);
   if (!a.fail_sink) hipLaunchKernelGGL(k_insert_finish, dim3(1), dim3(1), 0, s, a.info, a.expected_fill);
   return static_cast<int>(hipGetLastError());
}

template <typename K, bool PAIRS, int MODE>
int launch_lookup(LookupArgs<K> a, hipStream_t s, const hn_tune* t, OpKind op) {
   if constexpr (MODE != LK_RESET) {
      if (!t || t->probe_width == 0) { // automatic: lane windows
         // block_size does not apply: the lane kernel runs one wave per block
         uint64_t need = (a.len + 255) / 256; // at least 256 queries per wave
         if (need == 0) need = 1;
         // one resident wave per block slot: the waves split the queries
         // statically, so a block that waits for a free slot would finish late
         static std::atomic<int> occ{0};
         const int per_cu = resident_blocks_per_cu(k_lookup_lane<K, PAIRS, MODE>, kLaneBlock, occ);
         uint64_t cap = (t && t->grid_blocks > 0) ? static_cast<uint64_t>(t->grid_blocks)
                                                  : static_cast<uint64_t>(cu_count()) * per_cu;
         if (cap > (1u << 20)) cap = 1u << 20;
         const uint32_t blocks = static_cast<uint32_t>(need < cap ? need : cap);
         hipLaunchKernelGGL((k_lookup_lane<K, PAIRS, MODE>), dim3(blocks), dim3(kLaneBlock), 0, s, a);
         return static_cast<int>(hipGetLastError());
      }
   }
   Launch L;
   int rc = resolve_launch<K>(t, op, a.len, &L);
   if (rc) return rc;
   LookupKernel<K, PAIRS, MODE> k = pick_lookup<K, PAIRS, MODE>(L.probe_width);
   if (!k) return HN_EINVAL;
   hipLaunchKernelGGL(k, dim3(L.blocks), dim3(L.block_size), 0, s, a);
   return static_cast<int>(hipGetLastError());
}

template <typename K>
int insert_soa(const K* keys, const K* vals, void* buckets, hn_info* info, size_t len, hipStream_t s,
               const hn_tune* t) {
   if (!buckets || !info || (len && (!keys || !vals))) return HN_EINVAL;
   InsertArgs<K> a{keys, vals, nullptr, static_cast<Slot<K>*>(buckets), info, len, nullptr, nullptr, nullptr};
   return launch_insert<K, SRC_SOA>(a, len, s, t, OP_INSERT);
}

template <typename K>
int insert_index(const K* keys, void* buckets, hn_info* info, size_t len, hipStream_t s, const hn_tune* t) {
   if (!buckets || !info || (len && !keys)) return HN_EINVAL;
   InsertArgs<K> a{keys, nullptr, nullptr, static_cast<Slot<K>*>(buckets), info, len, nullptr, nullptr, nullptr};
   return launch_insert<K, SRC_INDEX>(a, len, s, t, OP_INSERT);
}

template <typename K>
int insert_pairs(const void* pairs, void* buckets, hn_info* info, size_t len, hipStream_t s, const hn_tune* t) {
   if (!buckets || !info || (len && !pairs)) return HN_EINVAL;
   InsertArgs<K> a{nullptr, nullptr, static_cast<const Slot<K>*>(pairs), static_cast<Slot<K>*>(buckets), info, len,
                   nullptr, nullptr, nullptr};
   return launch_insert<K, SRC_PAIRS>(a, len, s, t, OP_INSERT);
}

template <typename K>
int retrieve_soa(const K* keys, K* vals, uint8_t* hit, const void* buckets, const hn_info* info, size_t len,
                 hipStream_t s, const hn_tune* t) {
   if (!buckets || !info || (len && (!keys || !vals))) return HN_EINVAL;
   if (len == 0) return 0;
   LookupArgs<K> a{keys, vals, nullptr, hit, const_cast<Slot<K>*>(static_cast<const Slot<K>*>(buckets)),
                   const_cast<hn_info*>(info), len};
   return launch_lookup<K, false, LK_RETRIEVE>(a, s, t, OP_RETRIEVE);
}

template <typename K>
int retrieve_pairs(void* pairs, uint8_t* hit, const void* buckets, const hn_info* info, size_t len, hipStream_t s,
                   const hn_tune* t) {
   if (!buckets || !info || (len && !pairs)) return HN_EINVAL;
   if (len == 0) return 0;
   LookupArgs<K> a{nullptr, nullptr, static_cast<Slot<K>*>(pairs), hit,
                   const_cast<Slot<K>*>(static_cast<const Slot<K>*>(buckets)), const_cast<hn_info*>(info), len};
   return launch_lookup<K, true, LK_RETRIEVE>(a, s, t, OP_RETRIEVE);
}

template <typename K>
int erase_keys(const K* keys, void* buckets, hn_info* info, size_t len, hipStream_t s, const hn_tune* t) {
   if (!buckets || !info || (len && !keys)) return HN_EINVAL;
   if (len == 0) return 0;
   LookupArgs<K> a{keys, nullptr, nullptr, nullptr, static_cast<Slot<K>*>(buckets), info, len};
   return launch_lookup<K, false, LK_ERASE>(a, s, t, OP_ERASE);
}

template <typename K>
int reset_keys(const void* pairs, void* buckets, hn_info* info, size_t len, hipStream_t s) {
   if (!buckets || !info || (len && !pairs)) return HN_EINVAL;
   if (len == 0) return 0;
   LookupArgs<K> a{nullptr, nullptr, const_cast<Slot<K>*>(static_cast<const Slot<K>*>(pairs)), nullptr,
                   static_cast<Slot<K>*>(buckets), info, len};
   return launch_lookup<K, true, LK_RESET>(a, s, nullptr, OP_ERASE);
}

template <typename K>
int reset_all(void* buckets, hn_info* info, size_t len, hipStream_t s) {
   if (!buckets || !info) return HN_EINVAL;
   if (len == 0) return 0; // the reference launches no block for len 0 (hashers.h:155-163): fill stays
   const uint32_t blocks = stream_blocks(len, 256);
   hipLaunchKernelGGL(k_reset_all<K>, dim3(blocks), dim3(256), 0, s, static_cast<Slot<K>*>(buckets), info,
                      static_cast<uint64_t>(len));
   return static_cast<int>(hipGetLastError());
}

template <typename K>
int table_init(void* buckets, size_t n, hipStream_t s) {
   if (!buckets && n) return HN_EINVAL;
   if (n == 0) return 0;
   hipLaunchKernelGGL(k_table_init<K>, dim3(stream_blocks(n, 256)), dim3(256), 0, s, static_cast<Slot<K>*>(buckets),
                      static_cast<uint64_t>(n));
   return static_cast<int>(hipGetLastError());
}

constexpr size_t kHeaderAlign = 256;

size_t clean_header_bytes(uint64_t n_regions) {
   size_t b = sizeof(CleanHeader) + (n_regions ? n_regions - 1 : 0) * sizeof(uint64_t);
   return (b + kHeaderAlign - 1) / kHeaderAlign * kHeaderAlign;
}

// Largest region (in slots) that fits `bytes` of scratch, and the header
// size that goes with it.
template <typename K>
bool clean_plan(uint64_t n, size_t bytes, uint64_t* region_slots, uint64_t* n_regions, size_t* header) {
   const size_t slot = sizeof(Slot<K>);
   for (uint64_t regions = 1; regions <= n; regions *= 2) {
      const uint64_t rs = (n + regions - 1) / regions;
      const size_t hb = clean_header_bytes(regions);
      if (hb + rs * slot <= bytes) {
         *region_slots = rs;
         *n_regions = regions;
         *header = hb;
         return true;
      }
   }
   return false;
}

// Scratch of clean_tombstones.  With room for the segment pass (its header
// and lists, the reorder's, and a pool of at least seg_clean_min_pool()
// bytes) it takes that path; the pool holds segments too long for LDS, and
// when it also holds a whole-table region clean, a table with no EMPTY slot
// is cleaned too.  With less, the region clean alone runs in what there is.
// budget 0: the segment pass with a pool for long segments covering 1/32 of
// the table.  budget SIZE_MAX: enough for any table -- a full one, or one
// segment over the whole table.
template <typename K>
size_t clean_scratch_bytes(int sp, size_t budget) {
   if (sp < 1 || sp > max_size_power<K>()) return 0;
   const uint64_t n = 1ull << sp;
   const size_t fixed = seg_clean_fixed_bytes(sp) + reorder_scratch_bytes(sp);
   const size_t whole = clean_header_bytes(1) + n * sizeof(Slot<K>);
   const size_t pool = seg_clean_default_pool<K>(sp);
   if (budget == 0) return fixed + pool;
   if (budget == SIZE_MAX) return fixed + std::max(std::max(pool, whole), seg_clean_full_pool<K>(sp));
   if (budget >= fixed + seg_clean_min_pool()) return budget;
   uint64_t rs, nr;
   size_t hb;
   for (size_t b = budget; b >= hn_clean_tombstones_min_scratch(); b /= 2)
      if (clean_plan<K>(n, b, &rs, &nr, &hb)) return hb + rs * sizeof(Slot<K>);
   return hn_clean_tombstones_min_scratch();
}

// Probing insert of pairs[0, *d_len) with failures OR-ed into *fail_sink
// (used by the bulk build's fallback, hn_build.hip).
template <typename K>
int launch_insert_pairs_dlen(const Slot<K>* pairs, Slot<K>* buckets, hn_info* info, uint64_t max_len,
                             const uint64_t* d_len, int32_t* fail_sink, hipStream_t s) {
   InsertArgs<K> a{nullptr, nullptr, pairs, buckets, info, max_len, d_len, nullptr, fail_sink};
   return launch_insert<K, SRC_PAIRS>(a, max_len, s, nullptr, OP_REINSERT);
}
template int launch_insert_pairs_dlen<uint32_t>(const Slot<uint32_t>*, Slot<uint32_t>*, hn_info*, uint64_t,
                                                const uint64_t*, int32_t*, hipStream_t);
template int launch_insert_pairs_dlen<uint64_t>(const Slot<uint64_t>*, Slot<uint64_t>*, hn_info*, uint64_t,
                                                const uint64_t*, int32_t*, hipStream_t);

// The region clean: TOMBSTONE -> EMPTY and displaced elements lifted out and
// reinserted, region by region from an EMPTY slot; `gate` as in k_clean_begin.
template <typename K>
int region_clean(void* buckets, hn_info* info, int sp, void* scratch, size_t bytes, const uint32_t* gate,
                 hipStream_t s) {
   const uint64_t n = 1ull << sp;
   uint64_t rs = 0, nr = 0;
   size_t hb = 0;
   if (!clean_plan<K>(n, bytes, &rs, &nr, &hb)) return HN_ESCRATCH;
   CleanHeader* h = static_cast<CleanHeader*>(scratch);
   Slot<K>* pairs = reinterpret_cast<Slot<K>*>(static_cast<char*>(scratch) + hb);
   Slot<K>* b = static_cast<Slot<K>*>(buckets);
   hipLaunchKernelGGL(k_clean_begin, dim3(1), dim3(256), 0, s, info, h, static_cast<uint32_t>(nr), gate);
   const uint64_t chunk = 256 * 16;
   hipLaunchKernelGGL(k_find_empty<K>, dim3(static_cast<uint32_t>((n + chunk - 1) / chunk)), dim3(256), 0, s, b, n, h);
   hipLaunchKernelGGL(k_clean_commit, dim3(1), dim3(1), 0, s, info, h, n, rs);
   for (uint64_t r = 0; r < nr; ++r) {
      hipLaunchKernelGGL(k_clean_extract<K>, dim3(stream_blocks(rs, 256)), dim3(256), 0, s, b, sp, info, h,
                         static_cast<uint32_t>(r), rs, pairs);
      InsertArgs<K> a{nullptr, nullptr, pairs, b, info, rs, &h->counters[r], nullptr, &h->fail};
      int rc = launch_insert<K, SRC_PAIRS>(a, rs, s, nullptr, OP_REINSERT);
      if (rc) return rc;
   }
   hipLaunchKernelGGL(k_clean_end, dim3(1), dim3(1), 0, s, info, h);
   return static_cast<int>(hipGetLastError());
}

// Hashmap::clean_tombstones (hashinator.h:1261-1315).
template <typename K>
int clean_tombstones(void* buckets, hn_info* info, int sp, void* scratch, size_t bytes, hipStream_t s) {
   if (!buckets || !info || !scratch) return HN_EINVAL;
   if (sp < 1 || sp > max_size_power<K>()) return HN_EINVAL;
   const uint64_t n = 1ull << sp;
   const size_t seg = seg_clean_fixed_bytes(sp), ro = reorder_scratch_bytes(sp);
   if (bytes < seg + ro + seg_clean_min_pool()) return region_clean<K>(buckets, info, sp, scratch, bytes, nullptr, s);
   char* base = static_cast<char*>(scratch);
   char* pool = base + seg + ro;
   const size_t pool_bytes = bytes - seg - ro;
   uint64_t rs = 0, nr = 0;
   size_t hb = 0;
   const bool region_ok = clean_plan<K>(n, pool_bytes, &rs, &nr, &hb) && nr == 1;
   int rc = seg_clean<K>(buckets, info, sp, base, pool, pool_bytes, region_ok, s);
   if (rc || !region_ok) return rc;
   // a table with no EMPTY slot: one region over the whole table, then the reorder
   const uint32_t* gate = seg_clean_gate(base);
   rc = region_clean<K>(buckets, info, sp, pool, pool_bytes, gate, s);
   if (rc) return rc;
   return reorder_gated<K>(buckets, info, sp, base + seg, ro, gate, s);
}

template <typename K>
int rehash(const void* old_b, int old_sp, void* new_b, int new_sp, hn_info* info, void* scratch, hipStream_t s) {
   if (!old_b || !new_b || !info || !scratch) return HN_EINVAL;
   if (old_sp < 1 || old_sp > max_size_power<K>() || new_sp < 1 || new_sp > max_size_power<K>()) return HN_EINVAL;
   if (new_sp > 32) return HN_EINVAL; // hashinator.h:288-290 throws beyond 32 bits
   const uint64_t on = 1ull << old_sp, nn = 1ull << new_sp;
   const char* o0 = static_cast<const char*>(old_b);
   const char* n0 = static_cast<const char*>(new_b);
   if (o0 < n0 + nn * sizeof(Slot<K>) && n0 < o0 + on * sizeof(Slot<K>)) return HN_EINVAL;
   RehashHeader* h = static_cast<RehashHeader*>(scratch);
   hipLaunchKernelGGL(k_rehash_begin<K>, dim3(stream_blocks(nn, 256)), dim3(256), 0, s, static_cast<Slot<K>*>(new_b),
                      nn, info, new_sp, h);
   InsertArgs<K> a{nullptr, nullptr, static_cast<const Slot<K>*>(old_b), static_cast<Slot<K>*>(new_b), info, on,
                   nullptr, &h->prior_fill, nullptr};
   return launch_insert<K, SRC_TABLE>(a, on, s, nullptr, OP_REINSERT);
}

} // namespace hn

using namespace hn;

extern "C" {

int hn_info_init(hn_info* info, int sp, hn_stream_t s) {
   if (!info) return HN_EINVAL;
   hipLaunchKernelGGL(k_info_init, dim3(1), dim3(1), 0, s, info, sp);
   return static_cast<int>(hipGetLastError());
}

int hn_table_init_u32(void* b, size_t n, hn_stream_t s) { return table_init<uint32_t>(b, n, s); }
int hn_table_init_u64(void* b, size_t n, hn_stream_t s) { return table_init<uint64_t>(b, n, s); }

int hn_insert_u32(const uint32_t* k, const uint32_t* v, void* b, hn_info* i, size_t n, hn_stream_t s,
                  const hn_tune* t) {
   return insert_soa<uint32_t>(k, v, b, i, n, s, t);
}
int hn_insert_u64(const uint64_t* k, const uint64_t* v, void* b, hn_info* i, size_t n, hn_stream_t s,
                  const hn_tune* t) {
   return insert_soa<uint64_t>(k, v, b, i, n, s, t);
}
int hn_insert_index_u32(const uint32_t* k, void* b, hn_info* i, size_t n, hn_stream_t s, const hn_tune* t) {
   return insert_index<uint32_t>(k, b, i, n, s, t);
}
int hn_insert_index_u64(const uint64_t* k, void* b, hn_info* i, size_t n, hn_stream_t s, const hn_tune* t) {
   return insert_index<uint64_t>(k, b, i, n, s, t);
}
int hn_insert_pairs_u32(const void* p, void* b, hn_info* i, size_t n, hn_stream_t s, const hn_tune* t) {
   return insert_pairs<uint32_t>(p, b, i, n, s, t);
}
int hn_insert_pairs_u64(const void* p, void* b, hn_info* i, size_t n, hn_stream_t s, const hn_tune* t) {
   return insert_pairs<uint64_t>(p, b, i, n, s, t);
}
int hn_retrieve_u32(const uint32_t* k, uint32_t* v, uint8_t* h, const void* b, const hn_info* i, size_t n,
                    hn_stream_t s, const hn_tune* t) {
   return retrieve_soa<uint32_t>(k, v, h, b, i, n, s, t);
}
int hn_retrieve_u64(const uint64_t* k, uint64_t* v, uint8_t* h, const void* b, const hn_info* i, size_t n,
                    hn_stream_t s, const hn_tune* t) {
   return retrieve_soa<uint64_t>(k, v, h, b, i, n, s, t);
}
int hn_retrieve_pairs_u32(void* p, uint8_t* h, const void* b, const hn_info* i, size_t n, hn_stream_t s,
                          const hn_tune* t) {
   return retrieve_pairs<uint32_t>(p, h, b, i, n, s, t);
}
int hn_retrieve_pairs_u64(void* p, uint8_t* h, const void* b, const hn_info* i, size_t n, hn_stream_t s,
                          const hn_tune* t) {
   return retrieve_pairs<uint64_t>(p, h, b, i, n, s, t);
}
int hn_erase_u32(const uint32_t* k, void* b, hn_info* i, size_t n, hn_stream_t s, const hn_tune* t) {
   return erase_keys<uint32_t>(k, b, i, n, s, t);
}
int hn_erase_u64(const uint64_t* k, void* b, hn_info* i, size_t n, hn_stream_t s, const hn_tune* t) {
   return erase_keys<uint64_t>(k, b, i, n, s, t);
}
int hn_reset_all_u32(void* b, hn_info* i, size_t n, hn_stream_t s) { return reset_all<uint32_t>(b, i, n, s); }
int hn_reset_u32(const void* p, void* b, hn_info* i, size_t n, hn_stream_t s) { return reset_keys<uint32_t>(p, b, i, n, s); }
int hn_reset_u64(const void* p, void* b, hn_info* i, size_t n, hn_stream_t s) { return reset_keys<uint64_t>(p, b, i, n, s); }
int hn_reset_all_u64(void* b, hn_info* i, size_t n, hn_stream_t s) { return reset_all<uint64_t>(b, i, n, s); }

size_t hn_clean_tombstones_min_scratch(void) { return 64 * 1024; }
size_t hn_clean_tombstones_scratch_bytes_u32(int sp, size_t budget) { return clean_scratch_bytes<uint32_t>(sp, budget); }
size_t hn_clean_tombstones_scratch_bytes_u64(int sp, size_t budget) { return clean_scratch_bytes<uint64_t>(sp, budget); }
int hn_clean_tombstones_u32(void* b, hn_info* i, int sp, void* sc, size_t sb, hn_stream_t s) {
   return clean_tombstones<uint32_t>(b, i, sp, sc, sb, s);
}
int hn_clean_tombstones_u64(void* b, hn_info* i, int sp, void* sc, size_t sb, hn_stream_t s) {
   return clean_tombstones<uint64_t>(b, i, sp, sc, sb, s);
}

size_t hn_rehash_scratch_bytes(void) { return 256; }
int hn_rehash_u32(const void* ob, int osp, void* nb, int nsp, hn_info* i, void* sc, hn_stream_t s) {
   return rehash<uint32_t>(ob, osp, nb, nsp, i, sc, s);
}
int hn_rehash_u64(const void* ob, int osp, void* nb, int nsp, hn_info* i, void* sc, hn_stream_t s) {
   return rehash<uint64_t>(ob, osp, nb, nsp, i, sc, s);
}

const char* hn_version(void) { return "hashinator_amd 0.1 gfx950"; }

} // extern "C"
