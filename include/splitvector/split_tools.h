/* split_tools.h -- split::tools for the MI355X build.
 *
 * Host API of kstppd/hashinator include/splitvector/split_tools.h:
 * copy_if (1128-1223), copy_keys_if (1094-1191), split_prefix_scan (363-418),
 * splitStackArena (491-535), estimateMemoryForCompaction (1064-1089).
 *
 * A user Rule (a __host__ __device__ functor taking T&) cannot cross the
 * C-ABI, so it is evaluated here, once per element, by a small kernel
 * instantiated in the caller's translation unit that packs the answers one
 * bit per element; the stable single-pass compaction of libhashinator_amd
 * then moves the selected elements.  (The reference evaluates the rule twice,
 * in its count and its scatter pass; a rule with side effects on the element
 * sees them once here.)
 */
#pragma once
#include <hashinator_amd.h>

#include <cassert>

#include "splitvec.h"

namespace split {
namespace tools {

/* Bump allocator over device memory (split_tools.h:491-535): either owned
 * (stream-ordered allocation) or borrowed from the caller. */
class splitStackArena {
   size_t total_bytes = 0;
   size_t bytes_used = 0;
   void* _data = nullptr;
   hipStream_t s = 0;
   bool isOwner = false;

public:
   explicit splitStackArena(size_t bytes, hipStream_t str) : total_bytes(bytes), s(str), isOwner(true) {
      SPLIT_CHECK_ERR(hipMallocAsync(&_data, bytes ? bytes : 1, s));
   }
   explicit splitStackArena(void* ptr, size_t bytes) : total_bytes(bytes), _data(ptr), isOwner(false) {}
   splitStackArena() = delete;
   splitStackArena(const splitStackArena&) = delete;
   splitStackArena(splitStackArena&&) = delete;
   ~splitStackArena() {
      if (isOwner) SPLIT_CHECK_ERR(hipFreeAsync(_data, s));
   }
   void* allocate(const size_t bytes) {
      const size_t start = (bytes_used + 255) & ~size_t(255);
      assert(start + bytes <= total_bytes && "Mempool run out of space and crashed!");
      bytes_used = start + bytes;
      return static_cast<char*>(_data) + start;
   }
   void deallocate(const size_t bytes) { bytes_used -= bytes; }
   void reset() { bytes_used = 0; }
   const size_t& fill() const { return bytes_used; }
   const size_t& capacity() const { return total_bytes; }
   size_t free_space() const { return total_bytes - bytes_used; }
};

/* Device scratch needed to compact n elements with a user rule: the rule's
 * bit mask, the device-side count, and the look-back tile state. */
[[nodiscard]] inline size_t estimateMemoryForCompaction(const size_t inputSize) noexcept {
   return ((inputSize + 31) / 32) * 4 + 256 + 8 + 256 + hn_compact_scratch_bytes(inputSize) + 256;
}
template <typename T, class A>
[[nodiscard]] size_t estimateMemoryForCompaction(const SplitVector<T, A>& input) noexcept {
   return estimateMemoryForCompaction(input.size());
}

/* One rule evaluation per element, packed 32 per word (LSB first). */
template <typename T, typename Rule>
__global__ void split_rule_bits_kernel(T* input, size_t n, Rule rule, uint32_t* bits) {
   const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
   const bool keep = i < n ? rule(input[i]) : false;
   const unsigned long long m = __ballot(keep);
   const int lane = __lane_id();
   const size_t w0 = (i - lane) / 32; // first word of this wave
   if (lane == 0 && (i - lane) < n) bits[w0] = static_cast<uint32_t>(m);
   if (lane == 32 && i < n + 32 && (i - lane) + 32 < n) bits[w0 + 1] = static_cast<uint32_t>(m >> 32);
}

namespace detail {
template <typename T, typename Rule>
uint64_t* stage_rule(T* input, size_t n, Rule rule, splitStackArena& pool, uint32_t** bits_out, void** scratch,
                     size_t* scratch_bytes, hipStream_t s) {
   uint32_t* bits = static_cast<uint32_t*>(pool.allocate(((n + 63) / 64) * 8));
   uint64_t* d_count = static_cast<uint64_t*>(pool.allocate(8));
   *scratch_bytes = hn_compact_scratch_bytes(n);
   *scratch = pool.allocate(*scratch_bytes);
   if (n) {
      const unsigned blocks = static_cast<unsigned>((n + 255) / 256);
      hipLaunchKernelGGL((split_rule_bits_kernel<T, Rule>), dim3(blocks), dim3(256), 0, s, input, n, rule, bits);
      SPLIT_CHECK_ERR(hipGetLastError());
   }
   *bits_out = bits;
   return d_count;
}
inline size_t read_count(uint64_t* d_count, hipStream_t s) {
   uint64_t c = 0;
   SPLIT_CHECK_ERR(hipMemcpyAsync(&c, d_count, sizeof(c), hipMemcpyDeviceToHost, s));
   SPLIT_CHECK_ERR(hipStreamSynchronize(s));
   return static_cast<size_t>(c);
}
inline void check_abi(int rc) {
   if (rc > 0) SPLIT_CHECK_ERR(static_cast<hipError_t>(rc));
   if (rc < 0) {
      std::cerr << "split::tools: libhashinator_amd returned " << rc << "\n";
      abort();
   }
}
} // namespace detail

/* Stable compaction of raw device memory; returns the number of elements
 * written to `output` (copy_if_raw, split_tools.h:956-989). */
template <typename T, typename Rule>
size_t copy_if_raw(T* input, T* output, size_t size, Rule rule, splitStackArena& mPool, hipStream_t s = 0) {
   static_assert(sizeof(T) == 4 || sizeof(T) == 8 || sizeof(T) == 16, "element size must be 4, 8 or 16 bytes");
   uint32_t* bits;
   void* scratch;
   size_t sb;
   uint64_t* d_count = detail::stage_rule(input, size, rule, mPool, &bits, &scratch, &sb, s);
   detail::check_abi(hn_copy_if_bits(input, bits, size, sizeof(T), output, d_count, scratch, sb, s));
   return detail::read_count(d_count, s);
}

template <typename T, typename U, typename Rule>
size_t copy_keys_if_raw(T* input, U* output, size_t size, Rule rule, splitStackArena& mPool, hipStream_t s = 0) {
   static_assert(sizeof(U) == 4 || sizeof(U) == 8, "keys must be 4 or 8 bytes");
   static_assert(sizeof(T) == 2 * sizeof(U), "input must be hash_pair<U,U>");
   uint32_t* bits;
   void* scratch;
   size_t sb;
   uint64_t* d_count = detail::stage_rule(input, size, rule, mPool, &bits, &scratch, &sb, s);
   if (sizeof(U) == 4)
      detail::check_abi(hn_copy_keys_if_bits_u32(input, bits, size, reinterpret_cast<uint32_t*>(output), d_count,
                                                 scratch, sb, s));
   else
      detail::check_abi(hn_copy_keys_if_bits_u64(input, bits, size, reinterpret_cast<uint64_t*>(output), d_count,
                                                 scratch, sb, s));
   return detail::read_count(d_count, s);
}

/* copy_if (split_tools.h:1128-1144): output must already hold room for the
 * selected elements; its size becomes the count. */
template <typename T, typename Rule, class A>
void copy_if(SplitVector<T, A>& input, SplitVector<T, A>& output, Rule rule, hipStream_t s = 0) {
   splitStackArena mPool(estimateMemoryForCompaction(input.size()), s);
   const size_t len = copy_if_raw(input.data(), output.data(), input.size(), rule, mPool, s);
   output.erase(output.data() + len, output.end());
}
template <typename T, typename Rule, class A>
void copy_if(SplitVector<T, A>& input, SplitVector<T, A>& output, Rule rule, splitStackArena&& mPool,
             hipStream_t s = 0) {
   const size_t len = copy_if_raw(input.data(), output.data(), input.size(), rule, mPool, s);
   output.erase(output.data() + len, output.end());
}
template <typename T, typename Rule, class A>
void copy_if(SplitVector<T, A>& input, SplitVector<T, A>& output, Rule rule, void* stack, size_t max_size,
             hipStream_t s = 0) {
   assert(stack && "Invalid stack!");
   splitStackArena mPool(stack, max_size);
   const size_t len = copy_if_raw(input.data(), output.data(), input.size(), rule, mPool, s);
   output.erase(output.data() + len, output.end());
}
template <typename T, typename Rule>
size_t copy_if(T* input, T* output, size_t size, Rule rule, void* stack, size_t max_size, hipStream_t s = 0) {
   assert(stack && "Invalid stack!");
   splitStackArena mPool(stack, max_size);
   return copy_if_raw(input, output, size, rule, mPool, s);
}

/* copy_keys_if (split_tools.h:1094-1111, 1146-1191). */
template <typename T, typename U, typename Rule, class A, class B>
void copy_keys_if(SplitVector<T, A>& input, SplitVector<U, B>& output, Rule rule, hipStream_t s = 0) {
   splitStackArena mPool(estimateMemoryForCompaction(input.size()), s);
   const size_t len = copy_keys_if_raw(input.data(), output.data(), input.size(), rule, mPool, s);
   output.erase(output.data() + len, output.end());
}
template <typename T, typename U, typename Rule, class A, class B>
void copy_keys_if(SplitVector<T, A>& input, SplitVector<U, B>& output, Rule rule, splitStackArena&& mPool,
                  hipStream_t s = 0) {
   const size_t len = copy_keys_if_raw(input.data(), output.data(), input.size(), rule, mPool, s);
   output.erase(output.data() + len, output.end());
}
template <typename T, typename U, typename Rule, class A, class B>
void copy_keys_if(SplitVector<T, A>& input, SplitVector<U, B>& output, Rule rule, void* stack, size_t max_size,
                  hipStream_t s = 0) {
   assert(stack && "Invalid stack!");
   splitStackArena mPool(stack, max_size);
   const size_t len = copy_keys_if_raw(input.data(), output.data(), input.size(), rule, mPool, s);
   output.erase(output.data() + len, output.end());
}

/* copy_if_loop / copy_if_keys_loop (split_tools.h:994-1020): the entry points
 * the reference backs with one block looping over the input (loop_compact,
 * 718-881) for callers that avoid host round trips.  Same result as copy_if
 * (stable; output's size becomes the count), here through the single-pass
 * kernel; the size is set from the host after the count is read, so the call
 * ends with its stream synchronised.  The output grows to the input's size
 * first when it is smaller (the reference requires that capacity and
 * asserts it). */
template <typename T, typename Rule, class A>
void copy_if_loop(SplitVector<T, A>& input, SplitVector<T, A>& output, Rule rule, hipStream_t s = 0) {
   if (output.size() < input.size()) output.resize(input.size(), true, s);
   copy_if(input, output, rule, s);
}
template <typename T, typename U, typename Rule, class A, class B>
void copy_if_keys_loop(SplitVector<T, A>& input, SplitVector<U, B>& output, Rule rule, hipStream_t s = 0) {
   if (output.size() < input.size()) output.resize(input.size(), true, s);
   copy_keys_if(input, output, rule, s);
}

/* split_prefix_scan (split_tools.h:363-418): exclusive sum, wrapping in T.
 * Any size (the reference asserts a power of two). */
template <typename T, class A>
void split_prefix_scan(SplitVector<T, A>& input, SplitVector<T, A>& output, hipStream_t s = 0) {
   static_assert(std::is_integral<T>::value && (sizeof(T) == 4 || sizeof(T) == 8), "scan of 32/64-bit integers");
   const size_t n = input.size();
   if (output.size() < n) output.resize(n, true, s);
   const size_t sb = hn_compact_scratch_bytes(n);
   splitStackArena mPool(sb + 256, s);
   void* scratch = mPool.allocate(sb);
   if (sizeof(T) == 4)
      detail::check_abi(hn_exclusive_scan_u32(reinterpret_cast<const uint32_t*>(input.data()),
                                              reinterpret_cast<uint32_t*>(output.data()), n, scratch, sb, s));
   else
      detail::check_abi(hn_exclusive_scan_u64(reinterpret_cast<const uint64_t*>(input.data()),
                                              reinterpret_cast<uint64_t*>(output.data()), n, scratch, sb, s));
   SPLIT_CHECK_ERR(hipStreamSynchronize(s));
}

} // namespace tools
} // namespace split
