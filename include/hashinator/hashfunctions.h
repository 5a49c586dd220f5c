/* hashfunctions.h -- Fibonacci hashing, same arithmetic as kstppd/hashinator
 * include/hashinator/hashfunctions.h:31-83 (the device kernels use the same
 * formulas; callers mask with 2^sizePower - 1). */
#pragma once
#include <type_traits>

#include "../common.h"

namespace Hashinator {
namespace HashFunctions {

template <typename T>
struct Fibonacci {
   HASHINATOR_HOSTDEVICE static constexpr uint32_t fibhash(uint32_t key, const int sizePower) {
      key ^= key >> (32 - sizePower);
      return static_cast<uint32_t>((static_cast<uint64_t>(key) * 2654435769ull) >> (32 - sizePower));
   }
   HASHINATOR_HOSTDEVICE static constexpr uint64_t fibhash(uint64_t key, int sizePower) {
      key ^= key >> (64 - sizePower);
      return (key * 11400714819323198485ull) >> (64 - sizePower);
   }
   HASHINATOR_HOSTDEVICE static constexpr uint32_t fibhash(int key, const int sizePower) {
      return fibhash(static_cast<uint32_t>(key), sizePower);
   }
   HASHINATOR_HOSTDEVICE static constexpr T _hash(T key, const int sizePower) {
      static_assert(std::is_integral<T>::value, "Hashinator only works for integral types");
      return fibhash(key, sizePower);
   }
};

} // namespace HashFunctions
} // namespace Hashinator
