// hn_batch.hip -- many independent maps in a few persistent launches.
//
// No counterpart kernel in the reference (its API is one launch per map per
// operation, hashers.h:46-163); the per-map semantics are the reference's:
// for each map, in order, insert (insert_kernel, kernels_AMD.h:272-405),
// retrieve (retrieve_kernel 647-697), erase (delete_kernel 412-505), then
// clean_tombstones (hashinator.h:1261-1315).
//
// Insert, and erase + clean: one workgroup owns one map at a time and runs
// the phases on it back to back, separated by workgroup barriers with
// workgroup-scope fences (no other workgroup touches the map during the
// launch); maps are handed out by an atomic ticket in the order the caller
// lists them (largest first balances the grid).  Retrieve runs between the
// two as its own launch, one key per lane over units of every map's queries
// (k_batch_retrieve_lane).  The probe loops are the single-map ones
// (hn_probe.hpp), so a map's results are those of the plain calls.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "hn_device.hpp"
#include "hn_host.hpp"
#include "hn_probe.hpp"

namespace hn {

template <typename K>
struct BatchArgs {
   const hn_batch_map* maps;
   uint32_t n_maps;
   uint32_t ops;
   const K* ins_keys;
   const K* ins_vals;
   const K* qry_keys;
   K* qry_vals;
   uint8_t* qry_hit;
   const K* ers_keys;
   Slot<K>* clean_scratch;
   uint64_t* ticket;
};

__device__ __forceinline__ void phase_barrier() {
   __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
   __syncthreads();
   __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

template <typename K, int G>
__global__ __launch_bounds__(1024) void k_batch(BatchArgs<K> a) {
   __shared__ uint64_t s_map;
   __shared__ unsigned long long s_placed, s_max, s_erased, s_lifted;
   __shared__ int s_failed;
   const uint64_t groups = blockDim.x / G;
   const uint64_t g = threadIdx.x / G;
   const int lane = __lane_id();

   auto reset_shared = [&]() {
      s_placed = 0;
      s_max = 0;
      s_erased = 0;
      s_lifted = 0;
      s_failed = 0;
   };
   auto fold_insert = [&](const InsertCounters& c) {
      const uint64_t p = wave_sum(c.placed);
      const uint64_t m = wave_max(c.max_probe);
      const int f = wave_max(c.failed);
      if (lane == 0) {
         if (p) atomicAdd(&s_placed, static_cast<unsigned long long>(p));
         if (m) atomicMax(&s_max, static_cast<unsigned long long>(m));
         if (f) atomicOr(&s_failed, 1);
      }
   };

   for (;;) {
      if (threadIdx.x == 0) {
         s_map = __hip_atomic_fetch_add(a.ticket, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
         reset_shared();
      }
      __syncthreads();
      const uint64_t mi = s_map;
      if (mi >= a.n_maps) break;
      const hn_batch_map M = a.maps[mi];
      hn_info* info = M.info;
      Slot<K>* b = static_cast<Slot<K>*>(M.buckets);
      const int sp = info->size_power;
      const bool sp_ok = sp >= 1 && sp <= max_size_power<K>();
      const uint64_t nsl = sp_ok ? (1ull << sp) : 0;

      if (a.ops & HN_BATCH_INSERT) {
         InsertCounters c;
         if (sp_ok) {
            InsertArgs<K> ia{a.ins_keys + M.ins_off, a.ins_vals + M.ins_off, nullptr, b, info, M.ins_len,
                             nullptr, nullptr, nullptr};
            insert_range<K, G, SRC_SOA>(ia, sp, M.ins_len, g, groups, c);
         }
         fold_insert(c);
         __syncthreads();
         if (threadIdx.x == 0) {
            info->fill += s_placed;
            if (s_max) {
               const uint64_t o = next_overflow(s_max, kBucketOverflow);
               if (o > info->current_max_bucket_overflow) info->current_max_bucket_overflow = o;
            }
            info->err = (s_failed || !sp_ok) ? HN_STATUS_FAIL : HN_STATUS_SUCCESS;
            info->reserved1 = 0; // inserts do not keep Robin Hood order
            s_placed = 0;
            s_max = 0;
            s_failed = 0;
         }
         phase_barrier();
      }
      if ((a.ops & HN_BATCH_RETRIEVE) && sp_ok) {
         LookupArgs<K> la{a.qry_keys + M.qry_off, a.qry_vals + M.qry_off, nullptr,
                          a.qry_hit ? a.qry_hit + M.qry_off : nullptr, b, info, M.qry_len};
         lookup_range<K, G, false, LK_RETRIEVE>(la, sp, info->current_max_bucket_overflow, M.qry_len, g, groups,
                                                (info->reserved1 & kInfoOrdered) != 0);
         phase_barrier();
      }
      if ((a.ops & HN_BATCH_ERASE) && sp_ok) {
         LookupArgs<K> la{a.ers_keys + M.ers_off, nullptr, nullptr, nullptr, b, info, M.ers_len};
         const uint64_t e =
             wave_sum(lookup_range<K, G, false, LK_ERASE>(la, sp, info->current_max_bucket_overflow, M.ers_len, g, groups,
                                                          (info->reserved1 & kInfoOrdered) != 0));
         if (lane == 0 && e) atomicAdd(&s_erased, static_cast<unsigned long long>(e));
         __syncthreads();
         if (threadIdx.x == 0) {
            info->tombstone_counter += s_erased;
            info->fill -= s_erased;
         }
         phase_barrier();
      }
      if ((a.ops & HN_BATCH_CLEAN) && sp_ok && info->tombstone_counter > 0) {
         if (info->fill > M.scratch_cap) {
            // not enough scratch to lift the displaced elements out: leave the map as it is
            if (threadIdx.x == 0) info->err = HN_STATUS_FAIL;
         } else {
            // TOMBSTONE -> EMPTY; live elements away from home are lifted into scratch (one region: the map)
            Slot<K>* out = a.clean_scratch + M.scratch_off;
            const uint64_t mask = nsl - 1;
            for (uint64_t base = 0; base < nsl; base += blockDim.x) {
               const uint64_t i = base + threadIdx.x;
               bool take = false;
               Slot<K> s{};
               if (i < nsl) {
                  s = b[i];
                  if (s.key == tomb_key<K>()) {
                     Slot<K> e = s;
                     e.key = empty_key<K>();
                     b[i] = e;
                  } else if (s.key != empty_key<K>() && (fib_home(s.key, sp) & mask) != i) {
                     take = true;
                  }
               }
               const uint64_t m = __ballot(take);
               if (m) {
                  unsigned long long wb = 0;
                  const int leader = first_bit(m);
                  if (lane == leader) wb = atomicAdd(&s_lifted, static_cast<unsigned long long>(__popcll(m)));
                  wb = __shfl(wb, leader, kWave);
                  if (take) {
                     out[wb + __popcll(m & ((1ull << lane) - 1ull))] = s;
                     Slot<K> e = s;
                     e.key = empty_key<K>();
                     b[i] = e;
                  }
               }
            }
            phase_barrier();
            const uint64_t lifted = s_lifted;
            if (threadIdx.x == 0) {
               info->tombstone_counter = 0;
               info->current_max_bucket_overflow = kBucketOverflow;
               info->fill -= lifted;
               info->reserved1 = 0; // the reinsertion does not keep Robin Hood order
            }
            phase_barrier();
            InsertCounters c;
            if (lifted) {
               InsertArgs<K> ia{nullptr, nullptr, out, b, info, lifted, nullptr, nullptr, nullptr};
               insert_range<K, G, SRC_PAIRS>(ia, sp, lifted, g, groups, c);
            }
            fold_insert(c);
            __syncthreads();
            if (threadIdx.x == 0) {
               info->fill += s_placed;
               if (s_max) {
                  const uint64_t o = next_overflow(s_max, kBucketOverflow);
                  if (o > info->current_max_bucket_overflow) info->current_max_bucket_overflow = o;
               }
               if (lifted) info->err = s_failed ? HN_STATUS_FAIL : HN_STATUS_SUCCESS;
            }
         }
         phase_barrier();
      }
      __syncthreads(); // s_map is rewritten by thread 0 next
   }
}

__global__ void k_zero_u64(uint64_t* p) { *p = 0; }

// ---------------------------------------------------------- retrieve phase
//
// The retrieve phase runs on its own, one key per lane (lane_lookup): every
// map's queries are cut into units of kBatchUnit, numbered map after map
// (map m owns units [unit_start[m], unit_start[m+1]), so units come in the
// caller's map order, largest maps first), and a wave takes the next unit by
// a 64-bit ticket.  No unit is empty and no ticket is wasted.  Results per
// map are those of the plain retrieve: maps are independent, and the phases
// still run insert -> retrieve -> erase -> clean for each of them.
constexpr uint64_t kBatchUnit = 4096;

struct BatchScratch {
   uint64_t ticket;     // map ticket of k_batch
   uint64_t qry_ticket; // unit ticket of the retrieve phase
   uint64_t units;      // units of all maps
   uint64_t reserved[5];
   // followed by unit_start[n_maps + 1]
};
static_assert(sizeof(BatchScratch) == 64, "header of hn_batch_scratch_bytes");

__host__ __device__ inline uint64_t* unit_start(BatchScratch* sc) { return reinterpret_cast<uint64_t*>(sc + 1); }

size_t batch_scratch_bytes(uint32_t n_maps) {
   const size_t b = sizeof(BatchScratch) + (static_cast<size_t>(n_maps) + 1) * sizeof(uint64_t);
   return (b + 255) & ~static_cast<size_t>(255);
}

// unit_start = exclusive prefix sum of ceil(qry_len / kBatchUnit) over the
// maps, units = the total; the unit ticket starts at 0.  One workgroup of
// 1024 threads, 1024 maps per step.
__global__ __launch_bounds__(1024) void k_batch_prep(const hn_batch_map* maps, uint32_t n_maps, BatchScratch* sc) {
   __shared__ uint64_t wsum[1024 / kWave];
   __shared__ uint64_t carry;
   const int lane = __lane_id();
   const int w = threadIdx.x / kWave;
   uint64_t* us = unit_start(sc);
   if (threadIdx.x == 0) carry = 0;
   __syncthreads();
   for (uint64_t base = 0; base < n_maps; base += 1024) {
      const uint64_t m = base + threadIdx.x;
      const uint64_t u = m < n_maps ? (maps[m].qry_len + kBatchUnit - 1) / kBatchUnit : 0;
      uint64_t x = u; // inclusive scan over the wave
      for (int o = 1; o < kWave; o <<= 1) {
         const uint64_t y = __shfl_up(x, o, kWave);
         if (lane >= o) x += y;
      }
      if (lane == kWave - 1) wsum[w] = x;
      __syncthreads();
      uint64_t off = carry;
      for (int i = 0; i < w; ++i) off += wsum[i];
      if (m < n_maps) us[m] = off + x - u;
      __syncthreads();
      if (threadIdx.x == 1023) carry = off + x;
      __syncthreads();
   }
   if (threadIdx.x == 0) {
      us[n_maps] = carry;
      sc->units = carry;
      sc->qry_ticket = 0;
   }
}

// The next work unit of a wave.  Lane 0 draws one ticket and readfirstlane
// hands it to the wave as a scalar, so the unit -- and every branch on it --
// is wave-uniform by construction and each wave takes exactly one ticket per
// unit.  Must be called with the whole wave active (lane 0 is then the first
// active lane).
__device__ __forceinline__ uint64_t next_unit(uint64_t* ticket) {
   uint64_t t = 0;
   if (__lane_id() == 0) t = __hip_atomic_fetch_add(ticket, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
   const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(t));
   const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(t >> 32));
   return (static_cast<uint64_t>(hi) << 32) | lo;
}

template <typename K>
__global__ __launch_bounds__(kWave) void k_batch_retrieve_lane(BatchArgs<K> a, BatchScratch* sc) {
   __shared__ __align__(16) char win_s[kLaneWaveBytes];
   __shared__ uint32_t wb_s[kWave];
   __shared__ K ring_v[kLaneRing];
   __shared__ uint8_t ring_s[kLaneRing];
   const uint64_t units = sc->units;
   const uint64_t* us = unit_start(sc);
   for (;;) {
      const uint64_t u = next_unit(&sc->qry_ticket);
      if (u >= units) break;
      // the map holding unit u: us[lo] <= u < us[hi] throughout, so map lo has units
      uint32_t lo = 0, hi = a.n_maps;
      while (hi - lo > 1) {
         const uint32_t mid = lo + (hi - lo) / 2;
         if (us[mid] <= u)
            lo = mid;
         else
            hi = mid;
      }
      const hn_batch_map M = a.maps[lo];
      const uint64_t q0 = (u - us[lo]) * kBatchUnit;
      hn_info* info = M.info;
      const int sp = info->size_power;
      if (sp >= 1 && sp <= max_size_power<K>()) {
         const uint64_t q1 = q0 + kBatchUnit < M.qry_len ? q0 + kBatchUnit : M.qry_len;
         const uint64_t ovf = info->current_max_bucket_overflow;
         const bool ordered = (info->reserved1 & kInfoOrdered) != 0;
         LookupArgs<K> la{a.qry_keys + M.qry_off, a.qry_vals + M.qry_off, nullptr,
                          a.qry_hit ? a.qry_hit + M.qry_off : nullptr, static_cast<Slot<K>*>(M.buckets), info,
                          M.qry_len};
         if (sp >= LaneWin<K>::kLog && ovf > 0) {
            lane_lookup<K, false, LK_RETRIEVE>(la, sp, ovf, ordered, q0, q1, win_s, wb_s, ring_v, ring_s);
         } else {
            lookup_range<K, 1, false, LK_RETRIEVE>(la, sp, ovf, q1, q0 + __lane_id(), kWave, ordered);
         }
         wave_lds_sync();
      }
   }
}

// Fresh tables and Info(sp) for every map of a batch, one launch.
template <typename K>
__global__ __launch_bounds__(256) void k_batch_init(const hn_batch_map* maps, uint32_t n_maps, const int32_t* sps) {
   for (uint32_t m = blockIdx.x; m < n_maps; m += gridDim.x) {
      const hn_batch_map M = maps[m];
      const int sp = sps[m];
      Slot<K>* b = static_cast<Slot<K>*>(M.buckets);
      const uint64_t n = 1ull << sp;
      for (uint64_t i = threadIdx.x; i < n; i += blockDim.x) {
         Slot<K> s;
         s.key = empty_key<K>();
         s.val = 0;
         b[i] = s;
      }
      if (threadIdx.x == 0) {
         M.info->size_power = sp;
         M.info->reserved0 = 0;
         M.info->fill = 0;
         M.info->current_max_bucket_overflow = kBucketOverflow;
         M.info->tombstone_counter = 0;
         M.info->err = HN_STATUS_INVALID;
         M.info->reserved1 = 0;
      }
   }
}

template <typename K>
using BatchKernel = void (*)(BatchArgs<K>);

template <typename K>
BatchKernel<K> pick_batch(int G) {
   switch (G) {
   case 1: return k_batch<K, 1>;
   case 2: return k_batch<K, 2>;
   case 4: return k_batch<K, 4>;
   case 8: return k_batch<K, 8>;
   case 16: return k_batch<K, 16>;
   case 32: return k_batch<K, 32>;
   case 64: return k_batch<K, 64>;
   default: return nullptr;
   }
}

template <typename K>
int batch_run(const hn_batch_map* maps, uint32_t n_maps, uint32_t ops, const K* ik, const K* iv, const K* qk, K* qv,
              uint8_t* qh, const K* ek, void* clean_scratch, void* scratch, hipStream_t s, const hn_tune* t) {
   if (!maps || !scratch) return HN_EINVAL;
   if ((ops & HN_BATCH_INSERT) && (!ik || !iv)) return HN_EINVAL;
   if ((ops & HN_BATCH_RETRIEVE) && (!qk || !qv)) return HN_EINVAL;
   if ((ops & HN_BATCH_ERASE) && !ek) return HN_EINVAL;
   if ((ops & HN_BATCH_CLEAN) && !clean_scratch) return HN_EINVAL;
   if (n_maps == 0) return 0;
   const int G = (t && t->probe_width) ? t->probe_width : default_probe_width<K>(OP_INSERT);
   const int bs = (t && t->block_size) ? t->block_size : 512;
   if (G < 1 || G > 64 || (G & (G - 1)) || bs < 64 || bs > 1024 || (bs % 64)) return HN_EINVAL;
   uint32_t blocks = (t && t->grid_blocks > 0) ? static_cast<uint32_t>(t->grid_blocks)
                                               : static_cast<uint32_t>(cu_count()) * (2048 / bs);
   if (blocks > n_maps) blocks = n_maps;
   BatchKernel<K> k = pick_batch<K>(G);
   BatchScratch* sc = static_cast<BatchScratch*>(scratch);
   uint64_t* ticket = &sc->ticket;
   BatchArgs<K> a{maps, n_maps, ops, ik, iv, qk, qv, qh, ek, static_cast<Slot<K>*>(clean_scratch), ticket};
   // insert; retrieve (one key per lane, its own launch); erase and clean
   if (ops & HN_BATCH_INSERT) {
      a.ops = HN_BATCH_INSERT;
      hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, s, ticket);
      hipLaunchKernelGGL(k, dim3(blocks), dim3(bs), 0, s, a);
   }
   if (ops & HN_BATCH_RETRIEVE) {
      static std::atomic<int> occ{0};
      const int per_cu = resident_blocks_per_cu(k_batch_retrieve_lane<K>, kWave, occ);
      a.ops = HN_BATCH_RETRIEVE;
      hipLaunchKernelGGL(k_batch_prep, dim3(1), dim3(1024), 0, s, maps, n_maps, sc);
      hipLaunchKernelGGL(k_batch_retrieve_lane<K>, dim3(static_cast<uint32_t>(cu_count()) * per_cu), dim3(kWave), 0, s,
                         a, sc);
   }
   if (ops & (HN_BATCH_ERASE | HN_BATCH_CLEAN)) {
      a.ops = ops & (HN_BATCH_ERASE | HN_BATCH_CLEAN);
      hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, s, ticket);
      hipLaunchKernelGGL(k, dim3(blocks), dim3(bs), 0, s, a);
   }
   return static_cast<int>(hipGetLastError());
}

template <typename K>
int batch_init(const hn_batch_map* maps, uint32_t n_maps, const int32_t* sps, hipStream_t s) {
   if (!maps || !sps) return HN_EINVAL;
   if (n_maps == 0) return 0;
   const uint32_t blocks = n_maps < 65535u ? n_maps : 65535u;
   hipLaunchKernelGGL(k_batch_init<K>, dim3(blocks), dim3(256), 0, s, maps, n_maps, sps);
   return static_cast<int>(hipGetLastError());
}

} // namespace hn

using namespace hn;

extern "C" {

size_t hn_batch_scratch_bytes(uint32_t n_maps) { return batch_scratch_bytes(n_maps); }

int hn_batch_init_u32(const hn_batch_map* maps, uint32_t n, const int32_t* sps, hn_stream_t s) {
   return batch_init<uint32_t>(maps, n, sps, s);
}
int hn_batch_init_u64(const hn_batch_map* maps, uint32_t n, const int32_t* sps, hn_stream_t s) {
   return batch_init<uint64_t>(maps, n, sps, s);
}
int hn_batch_run_u32(const hn_batch_map* maps, uint32_t n, uint32_t ops, const uint32_t* ik, const uint32_t* iv,
                     const uint32_t* qk, uint32_t* qv, uint8_t* qh, const uint32_t* ek, void* cs, void* sc,
                     hn_stream_t s, const hn_tune* t) {
   return batch_run<uint32_t>(maps, n, ops, ik, iv, qk, qv, qh, ek, cs, sc, s, t);
}
int hn_batch_run_u64(const hn_batch_map* maps, uint32_t n, uint32_t ops, const uint64_t* ik, const uint64_t* iv,
                     const uint64_t* qk, uint64_t* qv, uint8_t* qh, const uint64_t* ek, void* cs, void* sc,
                     hn_stream_t s, const hn_tune* t) {
   return batch_run<uint64_t>(maps, n, ops, ik, iv, qk, qv, qh, ek, cs, sc, s, t);
}

} // extern "C"
