// hn_runtime.hip -- per-device facts the launch geometry needs.
#include <hip/hip_runtime.h>

#include "hn_host.hpp"

namespace hn {

int cu_count() {
   static int cached[64] = {0};
   int dev = 0;
   if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
   if (cached[dev] == 0) {
      int cus = 0;
      if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
      cached[dev] = cus;
   }
   return cached[dev];
}

} // namespace hn
