/* split_tools.h -- split::tools for the MI355X build.
 *
 * Host API of kstppd/hashinator include/splitvector/split_tools.h:
 * copy_if (1128-1223), copy_keys_if (1094-1191), split_prefix_scan (363-418),
 * splitStackArena (491-535), estimateMemoryForCompaction (1064-1089).
 *
 * A user Rule (a __host__ __device__ functor taking T&) cannot cross the
 * C-ABI, so copy_if and copy_keys_if instantiate the library's single-pass
 * compaction kernel (split_compact.h) here, in the caller's translation unit,
 * with the rule as its selector: the input is read once and each element's
 * rule evaluated once.  (The reference evaluates the rule twice, in its
 * count and its scatter pass.)  The rule receives a copy of the element; a
 * change it makes is stored back to the input, as the reference's rule,
 * which runs on the element in place, leaves it, and a selected element is
 * written to the output as the rule left it.
 */
#pragma once
#include <hashinator_amd.h>

#include <cassert>

#include "split_compact.h"
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

/* Device scratch needed to compact n elements with a user rule: the
 * device-side count and the look-back tile state. */
[[nodiscard]] inline size_t estimateMemoryForCompaction(const size_t inputSize) noexcept {
   return 256 + compact::scratch_bytes(inputSize) + 256;
}
template <typename T, class A>
[[nodiscard]] size_t estimateMemoryForCompaction(const SplitVector<T, A>& input) noexcept {
   return estimateMemoryForCompaction(input.size());
}

namespace detail {
/* Compaction of in[0, size) with a rule into out (element or key), counted
 * into device memory from the arena; returns the device count's address. */
template <typename T, typename Out, typename Rule, class Proj>
uint64_t* run_rule(T* input, Out* output, size_t size, Rule rule, Proj proj, splitStackArena& pool, hipStream_t s) {
   uint64_t* d_count = static_cast<uint64_t*>(pool.allocate(8));
   void* scratch = pool.allocate(compact::scratch_bytes(size));
   SPLIT_CHECK_ERR(compact::launch(input, size, output, d_count, scratch, s, compact::RuleSel<Rule>{rule}, proj));
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
   return detail::read_count(detail::run_rule(input, output, size, rule, compact::Whole{}, mPool, s), s);
}

/* Keys of the selected pairs (copy_keys_if_raw, split_tools.h:1022-1056). */
template <typename T, typename U, typename Rule>
size_t copy_keys_if_raw(T* input, U* output, size_t size, Rule rule, splitStackArena& mPool, hipStream_t s = 0) {
   return detail::read_count(detail::run_rule(input, output, size, rule, compact::FirstOf<U>{}, mPool, s), s);
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
