/* defaults.h -- constants and map metadata of the wave64 MI355X build.
 * kstppd/hashinator include/hashinator/defaults.h:24-54. */
#pragma once
#include <hashinator_amd.h>

#include "hashfunctions.h"

namespace Hashinator {
namespace defaults {
constexpr int WARPSIZE = 64;                            // wave64 on gfx950
constexpr int BUCKET_OVERFLOW = HN_BUCKET_OVERFLOW;     // 64, as the reference's wave64 build
constexpr int elementsPerWarp = 1;
constexpr int MAX_BLOCKSIZE = 1024;
template <typename T>
using DefaultHashFunction = HashFunctions::Fibonacci<T>;
} // namespace defaults

/* Same byte layout as the reference's Info; the two padding words are named
 * and must be zero (the device kernels use them to publish `err`). */
struct Info {
   Info() {}
   Info(int sz)
       : sizePower(sz), reserved0(0), fill(0), currentMaxBucketOverflow(defaults::BUCKET_OVERFLOW),
         tombstoneCounter(0), err(status::invalid), reserved1(0) {}
   int sizePower;
   int32_t reserved0;
   size_t fill;
   size_t currentMaxBucketOverflow;
   size_t tombstoneCounter;
   status err;
   int32_t reserved1;
};
static_assert(sizeof(Info) == sizeof(hn_info), "Info must match hn_info");
static_assert(offsetof(Info, err) == offsetof(hn_info, err), "Info must match hn_info");

} // namespace Hashinator
