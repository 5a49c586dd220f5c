// dropin_compaction.cpp -- the stream-compaction workload through the drop-in
// C++ API, as a user of the reference calls it: split::tools::copy_if on two
// device-resident split::SplitVector<hash_pair<uint64_t,uint64_t>> with a user
// rule functor (split_tools.h:1128-1144), timed as a whole call (rule
// staging, the library's compaction, the count read back and the output's
// size update), next to the raw C-ABI call hn_copy_if_pairs_u64 with the
// same rule as a built-in predicate on the same input, kernel time only.
// The two outputs are compared element by element on the device.
//
//   dropin_compaction [log2 n] [keep probability]     -> one JSON line
#include <hashinator/hashinator.h>
#include <splitvector/split_tools.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>

using Pair = Hashinator::hash_pair<uint64_t, uint64_t>;
using DVec = split::SplitVector<Pair, split::split_device_allocator<Pair>>;

__host__ __device__ inline uint64_t splitmix(uint64_t x) {
   x += 0x9E3779B97F4A7C15ull;
   x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
   x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
   return x ^ (x >> 31);
}

// {i, uniform value whose bit 0 is the keep decision, Bernoulli(p)}; `all`
// keeps every element (p = 1 has no 64-bit threshold).
__global__ void make_input(Pair* p, size_t n, uint64_t threshold, int all) {
   const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
   for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
      const uint64_t keep = (all || splitmix(i ^ 0x48415348494E4154ull) < threshold) ? 1 : 0;
      p[i] = Pair(i, (splitmix(i + 12345) & ~1ull) | keep);
   }
}

__global__ void count_diff(const Pair* a, const Pair* b, size_t n, unsigned long long* diff) {
   const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
   unsigned long long d = 0;
   for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
      d += (a[i].first != b[i].first || a[i].second != b[i].second) ? 1 : 0;
   if (d) atomicAdd(diff, d);
}

struct KeepOdd {
   __host__ __device__ bool operator()(Pair& e) const { return (e.second & 1) != 0; }
};

static void check(int rc, const char* what) {
   if (rc != 0) {
      fprintf(stderr, "%s failed: %d\n", what, rc);
      exit(1);
   }
}

int main(int argc, char** argv) {
   const int lg = argc > 1 ? atoi(argv[1]) : 28;
   const double p = argc > 2 ? atof(argv[2]) : 0.3;
   const size_t n = size_t(1) << lg;
   const int all = p >= 1.0 ? 1 : 0;
   const uint64_t threshold = all || p <= 0.0 ? 0 : static_cast<uint64_t>(p * 18446744073709551616.0);
   hipStream_t s = 0;

   DVec in(n), out(n);
   hipLaunchKernelGGL(make_input, dim3(8192), dim3(256), 0, s, in.data(), n, threshold, all);
   SPLIT_CHECK_ERR(hipGetLastError());
   SPLIT_CHECK_ERR(hipDeviceSynchronize());

   hipEvent_t e0, e1;
   SPLIT_CHECK_ERR(hipEventCreate(&e0));
   SPLIT_CHECK_ERR(hipEventCreate(&e1));
   float best_dropin = 1e30f, best_abi = 1e30f;
   size_t kept = 0;
   for (int rep = 0; rep < 4; ++rep) { // the first is a warm-up
      out.resize(n);
      SPLIT_CHECK_ERR(hipEventRecord(e0, s));
      split::tools::copy_if(in, out, KeepOdd(), s);
      SPLIT_CHECK_ERR(hipEventRecord(e1, s));
      SPLIT_CHECK_ERR(hipEventSynchronize(e1));
      float ms = 0;
      SPLIT_CHECK_ERR(hipEventElapsedTime(&ms, e0, e1));
      if (rep) best_dropin = ms < best_dropin ? ms : best_dropin;
      kept = out.size();
   }

   // the raw C-ABI on the same input, the rule as a built-in predicate
   Pair* raw = nullptr;
   uint64_t* d_count = nullptr;
   void* scratch = nullptr;
   const size_t sb = hn_compact_scratch_bytes(n);
   SPLIT_CHECK_ERR(hipMalloc(&raw, n * sizeof(Pair)));
   SPLIT_CHECK_ERR(hipMalloc(&d_count, sizeof(uint64_t)));
   SPLIT_CHECK_ERR(hipMalloc(&scratch, sb));
   const hn_pred pred{HN_PRED_MASKEQ, 1, 1, 1};
   for (int rep = 0; rep < 4; ++rep) {
      SPLIT_CHECK_ERR(hipEventRecord(e0, s));
      check(hn_copy_if_pairs_u64(in.data(), n, &pred, raw, d_count, scratch, sb, s), "hn_copy_if_pairs_u64");
      SPLIT_CHECK_ERR(hipEventRecord(e1, s));
      SPLIT_CHECK_ERR(hipEventSynchronize(e1));
      float ms = 0;
      SPLIT_CHECK_ERR(hipEventElapsedTime(&ms, e0, e1));
      if (rep) best_abi = ms < best_abi ? ms : best_abi;
   }
   uint64_t raw_count = 0;
   SPLIT_CHECK_ERR(hipMemcpy(&raw_count, d_count, sizeof(raw_count), hipMemcpyDeviceToHost));
   unsigned long long* d_diff = nullptr;
   SPLIT_CHECK_ERR(hipMalloc(&d_diff, sizeof(unsigned long long)));
   SPLIT_CHECK_ERR(hipMemset(d_diff, 0, sizeof(unsigned long long)));
   const size_t cmp = kept < raw_count ? kept : raw_count;
   hipLaunchKernelGGL(count_diff, dim3(8192), dim3(256), 0, s, out.data(), raw, cmp, d_diff);
   unsigned long long diff = 0;
   SPLIT_CHECK_ERR(hipMemcpy(&diff, d_diff, sizeof(diff), hipMemcpyDeviceToHost));
   const bool same = kept == raw_count && diff == 0;
   const double alg = static_cast<double>(n) * sizeof(Pair) + static_cast<double>(kept) * sizeof(Pair);
   printf("{\"n\": %zu, \"keep_p\": %.3f, \"kept\": %zu, \"dropin_ms\": %.3f, \"abi_ms\": %.3f, "
          "\"dropin_GBps_alg\": %.1f, \"abi_GBps_alg\": %.1f, \"same_as_abi\": %s}\n",
          n, p, kept, best_dropin, best_abi, alg / (best_dropin * 1e-3) / 1e9, alg / (best_abi * 1e-3) / 1e9,
          same ? "true" : "false");
   SPLIT_CHECK_ERR(hipFree(raw));
   SPLIT_CHECK_ERR(hipFree(d_count));
   SPLIT_CHECK_ERR(hipFree(scratch));
   SPLIT_CHECK_ERR(hipFree(d_diff));
   return same ? 0 : 2;
}
