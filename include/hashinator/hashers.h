/* hashers.h -- the DeviceHasher plugin of Hashmap, backed by libhashinator_amd.
 *
 * Same static, synchronous interface as kstppd/hashinator
 * include/hashinator/hashers.h:39-171 (Hashers::Hasher): every call launches
 * on stream `s` through the C-ABI and ends with a stream synchronize.  All
 * pointers (keys, values, buckets, Info) must be device-accessible; the
 * Info may be in device memory.  The
 * launch geometry is the library's (probe groups of 1..64 lanes), so the
 * WARP and elementsPerWarp parameters are accepted for source compatibility
 * and otherwise unused.
 */
#pragma once
#include <hashinator_amd.h>

#include <iostream>
#include <limits>
#include <type_traits>

#include "../splitvector/split_allocators.h"
#include "defaults.h"
#include "hash_pair.h"

namespace Hashinator {
namespace Hashers {

namespace detail {
inline void check(int rc) {
   if (rc > 0) SPLIT_CHECK_ERR(static_cast<hipError_t>(rc));
   if (rc < 0) {
      std::cerr << "Hashinator: libhashinator_amd returned " << rc << "\n";
      abort();
   }
}
inline hn_info* info_of(Info* i) { return reinterpret_cast<hn_info*>(i); }
inline hn_stream_t st(hipStream_t s) { return reinterpret_cast<hn_stream_t>(s); }
inline void warn_fail(Info* info, const char* what) {
#ifndef NDEBUG
   Info h;
   SPLIT_CHECK_ERR(hipMemcpy(&h, info, sizeof(Info), hipMemcpyDefault)); // info may live in device memory
   if (h.err == status::fail) {
      std::cerr << "***** Hashinator Runtime Warning ********" << std::endl;
      std::cerr << "Warning: Hashmap completely overflown in Device " << what
                << ".\nNot all elements were inserted!\nConsider resizing before calling insert" << std::endl;
      std::cerr << "******************************" << std::endl;
   }
#else
   (void)info;
   (void)what;
#endif
}
} // namespace detail

template <typename KEY_TYPE, typename VAL_TYPE, class HashFunction,
          KEY_TYPE EMPTYBUCKET = std::numeric_limits<KEY_TYPE>::max(), KEY_TYPE TOMBSTONE = EMPTYBUCKET - 1,
          int WARP = defaults::WARPSIZE, int elementsPerWarp = 1>
class Hasher {
   static_assert(std::is_unsigned<KEY_TYPE>::value && std::is_unsigned<VAL_TYPE>::value,
                 "the MI355X build instantiates unsigned keys and values");
   static_assert((sizeof(KEY_TYPE) == 4 || sizeof(KEY_TYPE) == 8) && sizeof(KEY_TYPE) == sizeof(VAL_TYPE),
                 "the MI355X build instantiates Hashmap<uint32_t,uint32_t> and Hashmap<uint64_t,uint64_t>");
   static_assert(EMPTYBUCKET == std::numeric_limits<KEY_TYPE>::max() && TOMBSTONE == EMPTYBUCKET - 1,
                 "the device kernels use the reference's default sentinels");
   static constexpr bool W64 = sizeof(KEY_TYPE) == 8;
   using pair = hash_pair<KEY_TYPE, VAL_TYPE>;

public:
   static void insert(KEY_TYPE* keys, VAL_TYPE* vals, pair* buckets, Info* info, size_t len, hipStream_t s = 0,
                      const hn_tune* t = nullptr) {
      detail::check(W64 ? hn_insert_u64(reinterpret_cast<const uint64_t*>(keys),
                                        reinterpret_cast<const uint64_t*>(vals), buckets, detail::info_of(info), len,
                                        detail::st(s), t)
                        : hn_insert_u32(reinterpret_cast<const uint32_t*>(keys),
                                        reinterpret_cast<const uint32_t*>(vals), buckets, detail::info_of(info), len,
                                        detail::st(s), t));
      SPLIT_CHECK_ERR(hipStreamSynchronize(s));
      detail::warn_fail(info, "Insert");
   }
   static void insertIndex(KEY_TYPE* keys, pair* buckets, Info* info, size_t len, hipStream_t s = 0,
                           const hn_tune* t = nullptr) {
      detail::check(W64 ? hn_insert_index_u64(reinterpret_cast<const uint64_t*>(keys), buckets,
                                              detail::info_of(info), len, detail::st(s), t)
                        : hn_insert_index_u32(reinterpret_cast<const uint32_t*>(keys), buckets,
                                              detail::info_of(info), len, detail::st(s), t));
      SPLIT_CHECK_ERR(hipStreamSynchronize(s));
      detail::warn_fail(info, "InsertIndex");
   }
   static void insert(pair* src, pair* buckets, Info* info, size_t len, hipStream_t s = 0,
                      const hn_tune* t = nullptr) {
      detail::check(W64 ? hn_insert_pairs_u64(src, buckets, detail::info_of(info), len, detail::st(s), t)
                        : hn_insert_pairs_u32(src, buckets, detail::info_of(info), len, detail::st(s), t));
      SPLIT_CHECK_ERR(hipStreamSynchronize(s));
      detail::warn_fail(info, "Insert");
   }
   static void retrieve(KEY_TYPE* keys, VAL_TYPE* vals, pair* buckets, Info* info, size_t len, hipStream_t s = 0,
                        const hn_tune* t = nullptr, uint8_t* hit = nullptr) {
      detail::check(W64 ? hn_retrieve_u64(reinterpret_cast<const uint64_t*>(keys), reinterpret_cast<uint64_t*>(vals),
                                          hit, buckets, detail::info_of(info), len, detail::st(s), t)
                        : hn_retrieve_u32(reinterpret_cast<const uint32_t*>(keys), reinterpret_cast<uint32_t*>(vals),
                                          hit, buckets, detail::info_of(info), len, detail::st(s), t));
      SPLIT_CHECK_ERR(hipStreamSynchronize(s));
   }
   static void retrieve(pair* src, pair* buckets, Info* info, size_t len, hipStream_t s = 0,
                        const hn_tune* t = nullptr, uint8_t* hit = nullptr) {
      detail::check(W64 ? hn_retrieve_pairs_u64(src, hit, buckets, detail::info_of(info), len, detail::st(s), t)
                        : hn_retrieve_pairs_u32(src, hit, buckets, detail::info_of(info), len, detail::st(s), t));
      SPLIT_CHECK_ERR(hipStreamSynchronize(s));
   }
   static void erase(KEY_TYPE* keys, pair* buckets, Info* info, size_t len, hipStream_t s = 0,
                     const hn_tune* t = nullptr) {
      detail::check(W64 ? hn_erase_u64(reinterpret_cast<const uint64_t*>(keys), buckets, detail::info_of(info), len,
                                       detail::st(s), t)
                        : hn_erase_u32(reinterpret_cast<const uint32_t*>(keys), buckets, detail::info_of(info), len,
                                       detail::st(s), t));
      SPLIT_CHECK_ERR(hipStreamSynchronize(s));
   }
   static void reset(pair* src, pair* dst, Info* info, size_t len, hipStream_t s = 0) {
      detail::check(W64 ? hn_reset_u64(src, dst, detail::info_of(info), len, detail::st(s))
                        : hn_reset_u32(src, dst, detail::info_of(info), len, detail::st(s)));
      SPLIT_CHECK_ERR(hipStreamSynchronize(s));
   }
   static void reset_all(pair* dst, Info* info, size_t len, hipStream_t s = 0) {
      detail::check(W64 ? hn_reset_all_u64(dst, detail::info_of(info), len, detail::st(s))
                        : hn_reset_all_u32(dst, detail::info_of(info), len, detail::st(s)));
      SPLIT_CHECK_ERR(hipStreamSynchronize(s));
   }
};

} // namespace Hashers
} // namespace Hashinator
