// hn_calib.hip -- random-access HBM calibration (no reference counterpart).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "hashinator_amd_calib.h"
#include "hn_device.hpp"
#include "hn_host.hpp"

namespace hn {

__device__ __forceinline__ uint64_t calib_mix(uint64_t x) {
   x ^= x >> 30;
   x *= 0xBF58476D1CE4E5B9ull;
   x ^= x >> 27;
   x *= 0x94D049BB133111EBull;
   x ^= x >> 31;
   return x;
}

// L lanes (gran / 16 bytes) read one aligned chunk together.  A lane group
// runs kStreams independent address streams at once (their reads are in
// flight together); within a stream, `chain` reads follow each other, the
// next address depending on the data read.  With a power-of-two chunk count
// the address is a mask, not a 64-bit modulo, so address arithmetic does not
// bound the read rate.
constexpr int kStreams = 4;

template <int L>
__global__ __launch_bounds__(256) void k_random_gather(const uint4* buf, uint64_t chunks, uint64_t n_access, int chain,
                                                       uint64_t seed, uint64_t* sink) {
   const int gl = __lane_id() & (L - 1);
   const uint64_t groups = static_cast<uint64_t>(gridDim.x) * (blockDim.x / L);
   const uint64_t g = (static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / L;
   const bool pow2 = (chunks & (chunks - 1)) == 0;
   const uint64_t mask = chunks - 1;
   const uint64_t per_iter = static_cast<uint64_t>(kStreams) * chain; // reads per group per iteration
   uint64_t acc = 0;
   for (uint64_t a = g; a * per_iter < n_access; a += groups) {
      uint64_t h[kStreams];
#pragma unroll
      for (int t = 0; t < kStreams; ++t) h[t] = calib_mix((a * kStreams + t) ^ seed);
      for (int c = 0; c < chain; ++c) {
         uint4 v[kStreams];
#pragma unroll
         for (int t = 0; t < kStreams; ++t) {
            const uint64_t chunk = pow2 ? (h[t] & mask) : (h[t] % chunks);
            v[t] = buf[chunk * L + gl];
         }
#pragma unroll
         for (int t = 0; t < kStreams; ++t) {
            acc += v[t].x ^ v[t].w;
            // the next offset depends on what was read (a dependent chain)
            h[t] = calib_mix(h[t] + (v[t].y & 1u));
         }
      }
   }
   acc = wave_sum(acc);
   if (__lane_id() == 0 && acc == 0x123456789ull) *sink = acc; // practically never; keeps the loads
}

// One lane reads a whole chunk on its own: NV consecutive 16-byte loads,
// issued back to back, for each of kLaneStreams independent chunks.  This is
// the access of a probe that keeps one key per lane and scans a window of
// slots in registers; the rate tells whether the L1 merges the NV loads of
// one line into one request.
constexpr int kLaneStreams = 2;

template <int NV>
__global__ __launch_bounds__(256) void k_random_gather_lane(const uint4* buf, uint64_t chunks, uint64_t n_access,
                                                            uint64_t seed, uint64_t* sink) {
   const uint64_t lanes = static_cast<uint64_t>(gridDim.x) * blockDim.x;
   const uint64_t g = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
   const uint64_t mask = chunks - 1; // chunks is a power of two
   uint64_t acc = 0;
   for (uint64_t a = g; a * kLaneStreams < n_access; a += lanes) {
      uint4 v[kLaneStreams][NV];
#pragma unroll
      for (int t = 0; t < kLaneStreams; ++t) {
         const uint64_t chunk = calib_mix((a * kLaneStreams + t) ^ seed) & mask;
#pragma unroll
         for (int j = 0; j < NV; ++j) v[t][j] = buf[chunk * NV + j];
      }
#pragma unroll
      for (int t = 0; t < kLaneStreams; ++t)
#pragma unroll
         for (int j = 0; j < NV; ++j) acc += v[t][j].x ^ v[t][j].w;
   }
   acc = wave_sum(acc);
   if (__lane_id() == 0 && acc == 0x123456789ull) *sink = acc;
}

} // namespace hn

namespace hn {
// Streaming copy: 16-byte vectors, one per thread and step, grid of 4
// workgroups per CU (the fastest plain copy measured on MI355X: 5.7 TB/s of
// read + write at 1 GiB).
__global__ __launch_bounds__(256) void k_stream_copy(const uint4* __restrict__ in, uint4* __restrict__ out, uint64_t n) {
   const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
   for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) out[i] = in[i];
}

} // namespace hn

using namespace hn;

extern "C" int hn_calib_stream_copy(const void* src, void* dst, size_t bytes, hn_stream_t s) {
   if (!src || !dst || bytes % 16 != 0) return HN_EINVAL;
   if (bytes == 0) return 0;
   const uint32_t blocks = static_cast<uint32_t>(cu_count()) * 4;
   hipLaunchKernelGGL(k_stream_copy, dim3(blocks), dim3(256), 0, s, static_cast<const uint4*>(src), static_cast<uint4*>(dst),
                      static_cast<uint64_t>(bytes / 16));
   return static_cast<int>(hipGetLastError());
}

extern "C" int hn_calib_random_gather_lane(const void* buf, size_t bytes, int gran, size_t n_access, uint64_t seed,
                                           uint64_t* sink, hn_stream_t s) {
   if (!buf || !sink) return HN_EINVAL;
   const uint64_t chunks = bytes / static_cast<uint64_t>(gran);
   if (chunks == 0 || (chunks & (chunks - 1)) != 0) return HN_EINVAL;
   const uint32_t blocks = static_cast<uint32_t>(cu_count()) * 8;
   const uint4* b = static_cast<const uint4*>(buf);
   switch (gran) {
   case 16: hipLaunchKernelGGL(k_random_gather_lane<1>, dim3(blocks), dim3(256), 0, s, b, chunks, n_access, seed, sink); break;
   case 32: hipLaunchKernelGGL(k_random_gather_lane<2>, dim3(blocks), dim3(256), 0, s, b, chunks, n_access, seed, sink); break;
   case 64: hipLaunchKernelGGL(k_random_gather_lane<4>, dim3(blocks), dim3(256), 0, s, b, chunks, n_access, seed, sink); break;
   case 128: hipLaunchKernelGGL(k_random_gather_lane<8>, dim3(blocks), dim3(256), 0, s, b, chunks, n_access, seed, sink); break;
   default: return HN_EINVAL;
   }
   return static_cast<int>(hipGetLastError());
}

extern "C" int hn_calib_random_gather(const void* buf, size_t bytes, int gran, size_t n_access, int chain,
                                      uint64_t seed, uint64_t* sink, hn_stream_t s) {
   if (!buf || !sink || chain < 1) return HN_EINVAL;
   const uint64_t chunks = bytes / static_cast<uint64_t>(gran);
   if (chunks == 0) return HN_EINVAL;
   const uint32_t blocks = static_cast<uint32_t>(cu_count()) * 8;
   const uint4* b = static_cast<const uint4*>(buf);
   switch (gran) {
   case 16: hipLaunchKernelGGL(k_random_gather<1>, dim3(blocks), dim3(256), 0, s, b, chunks, n_access, chain, seed, sink); break;
   case 32: hipLaunchKernelGGL(k_random_gather<2>, dim3(blocks), dim3(256), 0, s, b, chunks, n_access, chain, seed, sink); break;
   case 64: hipLaunchKernelGGL(k_random_gather<4>, dim3(blocks), dim3(256), 0, s, b, chunks, n_access, chain, seed, sink); break;
   case 128: hipLaunchKernelGGL(k_random_gather<8>, dim3(blocks), dim3(256), 0, s, b, chunks, n_access, chain, seed, sink); break;
   case 256: hipLaunchKernelGGL(k_random_gather<16>, dim3(blocks), dim3(256), 0, s, b, chunks, n_access, chain, seed, sink); break;
   default: return HN_EINVAL;
   }
   return static_cast<int>(hipGetLastError());
}
