// hn_device.hpp -- device-side building blocks shared by the gfx950 kernels.
//
// Layout and arithmetic follow kstppd/hashinator:
//   slot        = hash_pair<K,V> {K first; V second;}          hash_pair.h:29-75
//   EMPTYBUCKET = max(K), TOMBSTONE = EMPTYBUCKET - 1          hashinator.h:59-60
//   home slot   = Fibonacci hash, masked by 2^sizePower - 1    hashfunctions.h:39-56
//   overflow    = nextOverflow(longest probe, 64)              common.h:52-57
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hashinator_amd.h"

namespace hn {

constexpr int kWave = 64;
// hn_info::reserved1 bits: 1 = an insert kernel saw a failure (transient,
// cleared when the insert's err is published); 2 = the table is in Robin
// Hood order (set by the reorder pass and the bulk build, cleared by every
// insert).
constexpr int32_t kInfoFail = 1;
constexpr int32_t kInfoOrdered = 2;
constexpr uint64_t kBucketOverflow = HN_BUCKET_OVERFLOW;

template <typename K>
struct Slot;
template <>
struct alignas(8) Slot<uint32_t> {
   uint32_t key, val;
};
template <>
struct alignas(16) Slot<uint64_t> {
   uint64_t key, val;
};

template <typename K>
__host__ __device__ constexpr K empty_key() {
   return static_cast<K>(~static_cast<K>(0));
}
template <typename K>
__host__ __device__ constexpr K tomb_key() {
   return static_cast<K>(~static_cast<K>(0) - 1);
}
// Hashmap::hash returns uint32_t (hashinator.h:77-81), so no table has more
// than 2^32 addressable slots; rehash refuses beyond (hashinator.h:288-290).
template <typename K>
__host__ __device__ constexpr int max_size_power() {
   return 32;
}

// Fibonacci hash (hashfunctions.h:39-43): 32-bit key, 64-bit product, shift,
// truncate to 32 bits.  The caller masks with 2^sp - 1.
__device__ __forceinline__ uint64_t fib_home(uint32_t key, int sp) {
   key ^= key >> (32 - sp);
   return static_cast<uint32_t>((static_cast<uint64_t>(key) * 2654435769ull) >> (32 - sp));
}
// hashfunctions.h:52-56, then Hashmap::hash's uint32_t return (hashinator.h:77-81).
__device__ __forceinline__ uint64_t fib_home(uint64_t key, int sp) {
   key ^= key >> (64 - sp);
   return static_cast<uint32_t>((key * 11400714819323198485ull) >> (64 - sp));
}

// The same hashes in 32-bit words for sp in [1, 32] (the probe loops' hot
// path).  64-bit key: key ^= key >> (64 - sp) only touches the low word, and
// bits [64 - sp, 64) of key * phi come from the high word of the product.
__device__ __forceinline__ uint32_t fib_home32(uint64_t key, int sp) {
   uint32_t lo = static_cast<uint32_t>(key);
   const uint32_t hi = static_cast<uint32_t>(key >> 32);
   lo ^= hi >> (32 - sp);
   const uint32_t ph = __umulhi(lo, 0x7F4A7C15u) + lo * 0x9E3779B9u + hi * 0x7F4A7C15u;
   return ph >> (32 - sp);
}
__device__ __forceinline__ uint32_t fib_home32(uint32_t key, int sp) {
   key ^= key >> (32 - sp);
   const uint32_t plo = key * 2654435769u, phi = __umulhi(key, 2654435769u);
   return __builtin_amdgcn_alignbit(phi, plo, static_cast<uint32_t>(32 - sp));
}
__device__ __forceinline__ uint32_t slot_mask32(int sp) { return sp >= 32 ? 0xFFFFFFFFu : ((1u << sp) - 1u); }

// The G-lane slice of a wave-wide ballot in 32-bit arithmetic (G <= 32).
template <int G>
__device__ __forceinline__ uint32_t group_bits32(uint64_t ballot, int gbase) {
   static_assert(G <= 32, "32-bit slice");
   const uint32_t half = gbase < 32 ? static_cast<uint32_t>(ballot) : static_cast<uint32_t>(ballot >> 32);
   if constexpr (G == 32) {
      return half;
   } else {
      return (half >> (gbase & 31)) & ((1u << G) - 1u);
   }
}

__host__ __device__ constexpr uint64_t next_overflow(uint64_t cur, uint64_t w) {
   return (cur % w) == 0 ? cur : cur + (w - cur % w);
}

// The G-lane slice of a wave-wide ballot that belongs to the caller's group.
template <int G>
__device__ __forceinline__ uint64_t group_bits(uint64_t ballot, int gbase) {
   if constexpr (G == 64) {
      return ballot;
   } else {
      return (ballot >> gbase) & ((1ull << G) - 1ull);
   }
}

__device__ __forceinline__ int first_bit(uint64_t m) { return __ffsll(static_cast<unsigned long long>(m)) - 1; }

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
   for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
   return v;
}
template <typename T>
__device__ __forceinline__ T wave_max(T v) {
   for (int o = 32; o > 0; o >>= 1) {
      T u = __shfl_xor(v, o, kWave);
      v = u > v ? u : v;
   }
   return v;
}

template <typename T>
__device__ __forceinline__ T atomic_add_relaxed(T* p, T v) {
   return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ T atomic_max_relaxed(T* p, T v) {
   return __hip_atomic_fetch_max(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t atomic_min_u64(uint64_t* p, uint64_t v) {
   return __hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ T atomic_cas(T* p, T expected, T desired) {
   __hip_atomic_compare_exchange_strong(p, &expected, desired, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                        __HIP_MEMORY_SCOPE_AGENT);
   return expected;
}

// Per-wave epilogue of the insert kernels: fold the wave's counters into the
// map's Info with relaxed device-scope atomics (a failure sets bit 0 of
// reserved1).  `err` is published by a one-thread kernel launched after the
// insert (k_insert_finish), so no wave needs a device-scope fence.  All
// lanes active.
__device__ __forceinline__ void insert_fold(hn_info* info, uint64_t placed, uint64_t max_probe, int failed) {
   placed = wave_sum(placed);
   max_probe = wave_max(max_probe);
   failed = wave_max(failed);
   if (__lane_id() == 0) {
      if (placed) atomic_add_relaxed(&info->fill, placed);
      if (max_probe) atomic_max_relaxed(&info->current_max_bucket_overflow, next_overflow(max_probe, kBucketOverflow));
      if (failed) __hip_atomic_fetch_or(&info->reserved1, kInfoFail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
   }
}

} // namespace hn
