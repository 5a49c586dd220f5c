/* common.h -- drop-in for kstppd/hashinator include/common.h on the MI355X build.
 *
 * Same names and meaning: nextPow2 (common.h:36-47), nextOverflow (52-57),
 * isDeviceAccessible (60-79), Hashinator::status / targets (84-91). */
#pragma once
#include <cstddef>
#include <cstdint>
#include <hip/hip_runtime.h>

#define HASHINATOR_DEVICEONLY __device__
#define HASHINATOR_HOSTDEVICE __host__ __device__

HASHINATOR_HOSTDEVICE
constexpr inline size_t nextPow2(size_t v) noexcept {
   v--;
   for (unsigned s = 1; s < 64; s <<= 1) v |= v >> s;
   return ++v;
}

HASHINATOR_HOSTDEVICE
constexpr inline size_t nextOverflow(size_t currentOverflow, size_t virtualWarp) noexcept {
   return (currentOverflow % virtualWarp) == 0 ? currentOverflow
                                               : currentOverflow + (virtualWarp - currentOverflow % virtualWarp);
}

inline bool isDeviceAccessible(void* ptr) {
   hipPointerAttribute_t a;
   if (hipPointerGetAttributes(&a, ptr) != hipSuccess) {
      (void)hipGetLastError();
      return false;
   }
   return a.type == hipMemoryTypeManaged || a.type == hipMemoryTypeDevice;
}

namespace Hashinator {
enum status { success, fail, invalid };
enum targets { host, device, automatic };
} // namespace Hashinator
