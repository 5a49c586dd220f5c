// residency.cpp -- the same insert + retrieve on a u64 map whose bucket array
// and Info live in device memory (hipMalloc) or in unified memory
// (hipMallocManaged, with and without advise + prefetch): results must agree,
// timings show what the residency costs on this device.  Usage: residency [sp]
#include <hashinator_amd.h>
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>
#include <vector>
#define CK(x) do { hipError_t e = (x); if (e != hipSuccess) { printf("HIP error %s at %d\n", hipGetErrorString(e), __LINE__); return 1; } } while (0)
int main(int argc, char** argv) {
   int dev = 0; CK(hipGetDevice(&dev));
   int mm = 0, cma = 0, pma = 0;
   hipDeviceGetAttribute(&mm, hipDeviceAttributeManagedMemory, dev);
   hipDeviceGetAttribute(&cma, hipDeviceAttributeConcurrentManagedAccess, dev);
   hipDeviceGetAttribute(&pma, hipDeviceAttributePageableMemoryAccess, dev);
   printf("managedMemory=%d concurrentManagedAccess=%d pageableMemoryAccess=%d\n", mm, cma, pma);
   const int sp = argc > 1 ? atoi(argv[1]) : 26; const size_t N = size_t(1) << sp, n = size_t(0.9 * N);
   std::vector<uint64_t> hk(n);
   for (size_t i = 0; i < n; ++i) hk[i] = (i * 0x9E3779B97F4A7C15ull) >> 1;
   for (int managed = 0; managed < 3; ++managed) {
      void *b, *keys, *vals, *info, *out;
      if (managed == 1) {
         CK(hipMallocManaged(&b, N * 16)); CK(hipMallocManaged(&info, 40));
         CK(hipMemAdvise(b, N * 16, hipMemAdviseSetPreferredLocation, dev));
         CK(hipMemPrefetchAsync(b, N * 16, dev, 0));
      } else if (managed == 2) {
         CK(hipMallocManaged(&b, N * 16)); CK(hipMallocManaged(&info, 40));
      } else { CK(hipMalloc(&b, N * 16)); CK(hipMalloc(&info, 40)); }
      CK(hipMalloc(&keys, n * 8)); CK(hipMalloc(&vals, n * 8)); CK(hipMalloc(&out, n * 8));
      CK(hipMemcpy(keys, hk.data(), n * 8, hipMemcpyHostToDevice)); CK(hipMemcpy(vals, hk.data(), n * 8, hipMemcpyHostToDevice));
      hipPointerAttribute_t at; CK(hipPointerGetAttributes(&at, b));
      hipEvent_t e0, e1, e2; hipEventCreate(&e0); hipEventCreate(&e1); hipEventCreate(&e2);
      float ti = 0, tr = 0;
      for (int rep = 0; rep < 3; ++rep) {
         if (hn_table_init_u64(b, N, 0) || hn_info_init((hn_info*)info, sp, 0)) return 2;
         hipEventRecord(e0, 0);
         if (hn_insert_u64((uint64_t*)keys, (uint64_t*)vals, b, (hn_info*)info, n, 0, nullptr)) return 3;
         hipEventRecord(e1, 0);
         if (hn_retrieve_u64((uint64_t*)keys, (uint64_t*)out, nullptr, b, (hn_info*)info, n, 0, nullptr)) return 4;
         hipEventRecord(e2, 0); CK(hipEventSynchronize(e2));
         hipEventElapsedTime(&ti, e0, e1); hipEventElapsedTime(&tr, e1, e2);
      }
      hn_info hi; CK(hipMemcpy(&hi, info, 40, hipMemcpyDefault));
      std::vector<uint64_t> got(n);
      CK(hipMemcpy(got.data(), out, n * 8, hipMemcpyDeviceToHost));
      for (size_t i = 0; i < n; ++i)
         if (got[i] != hk[i]) { printf("MISMATCH at %zu\n", i); return 5; }
      if (hi.fill != n || hi.err != 0) { printf("BAD INFO\n"); return 6; }
      printf("%s: ptr type %d  insert %.2f ms (%.0f Mops/s)  retrieve %.2f ms (%.0f Mops/s)  fill %lu err %d\n",
             managed == 0 ? "hipMalloc" : managed == 1 ? "managed+advise+prefetch" : "managed plain", (int)at.type,
             ti, n / ti / 1e3, tr, n / tr / 1e3, (unsigned long)hi.fill, hi.err);
      hipFree(b); hipFree(info); hipFree(keys); hipFree(vals); hipFree(out);
   }
   printf("residency ok\n");
   return 0;
}
