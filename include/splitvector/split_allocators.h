/* split_allocators.h -- allocators and the abort-on-error check of the
 * MI355X build.  kstppd/hashinator include/splitvector/split_allocators.h:
 * SPLIT_CHECK_ERR (31-50), split_unified_allocator (61-129),
 * split_host_allocator (142-201); split_device_allocator is an addition. */
#pragma once
#include <cassert>
#include <cstdlib>
#include <iostream>
#include <limits>
#include <new>
#include <type_traits>
#include <utility>

#include "archMacros.h"

namespace split {

#define SPLIT_CHECK_ERR(err) (split::hip_error((err), __FILE__, __LINE__))
/* Any HIP failure ends the program, naming the call site (the reference's
 * error convention for its containers). */
inline void hip_error(hipError_t err, const char* file, int line) {
   if (err == hipSuccess) return;
   std::cerr << "split: HIP call failed at " << file << ":" << line << ": " << hipGetErrorName(err) << " ("
             << hipGetErrorString(err) << ")" << std::endl;
   abort();
}

/* Whether an allocator's memory is reachable by kernels only (the host then
 * moves elements with copies). */
template <class A>
struct allocator_is_device_only : std::false_type {};

/* Unified (managed) memory, the reference's default for SplitVector and the
 * map metadata.  Allocations are advised to live on the current device. */
template <class T>
class split_unified_allocator {
public:
   typedef T value_type;
   typedef value_type* pointer;
   typedef const value_type* const_pointer;
   typedef value_type& reference;
   typedef const value_type& const_reference;
   typedef ptrdiff_t difference_type;
   typedef size_t size_type;
   template <class U>
   struct rebind {
      typedef split_unified_allocator<U> other;
   };
   split_unified_allocator() throw() {}
   template <class U>
   split_unified_allocator(split_unified_allocator<U> const&) throw() {}
   pointer address(reference x) const { return &x; }
   const_pointer address(const_reference x) const { return &x; }

   static void* allocate_raw(size_type bytes, const void* = 0) {
      void* ret = nullptr;
      SPLIT_CHECK_ERR(hipMallocManaged(&ret, bytes));
      if (ret == nullptr) throw std::bad_alloc();
      int dev = 0;
      if (hipGetDevice(&dev) == hipSuccess) (void)hipMemAdvise(ret, bytes, hipMemAdviseSetPreferredLocation, dev);
      (void)hipGetLastError();
      return ret;
   }
   pointer allocate(size_type n, const void* = 0) {
      assert(n && "allocate 0");
      return static_cast<pointer>(allocate_raw(n * sizeof(value_type)));
   }
   void deallocate(pointer p, size_type n) {
      if (n != 0 && p != 0) SPLIT_CHECK_ERR(hipFree(p));
   }
   static void deallocate(void* p, size_type n) {
      if (n != 0 && p != 0) SPLIT_CHECK_ERR(hipFree(p));
   }
   size_type max_size() const throw() { return std::numeric_limits<size_type>::max() / sizeof(value_type); }
   template <typename U, typename... Args>
   void construct(U* p, Args&&... args) {
      ::new (p) U(std::forward<Args>(args)...);
   }
   void destroy(pointer p) { p->~value_type(); }
};

/* Plain host memory (the reference's CPU-only allocator). */
template <class T>
class split_host_allocator {
public:
   typedef T value_type;
   typedef value_type* pointer;
   typedef const value_type* const_pointer;
   typedef value_type& reference;
   typedef const value_type& const_reference;
   typedef ptrdiff_t difference_type;
   typedef size_t size_type;
   template <class U>
   struct rebind {
      typedef split_host_allocator<U> other;
   };
   split_host_allocator() throw() {}
   template <class U>
   split_host_allocator(split_host_allocator<U> const&) throw() {}
   pointer allocate(size_type n, const void* = 0) {
      pointer ret = static_cast<pointer>(malloc(n * sizeof(value_type)));
      if (ret == nullptr) throw std::bad_alloc();
      return ret;
   }
   static void* allocate_raw(size_type bytes, const void* = 0) {
      void* ret = malloc(bytes);
      if (ret == nullptr) throw std::bad_alloc();
      return ret;
   }
   void deallocate(pointer p, size_type) { free(p); }
   static void deallocate(void* p, size_type) { free(p); }
   size_type max_size() const throw() { return std::numeric_limits<size_type>::max() / sizeof(value_type); }
   template <typename U, typename... Args>
   void construct(U* p, Args&&... args) {
      ::new (p) U(std::forward<Args>(args)...);
   }
   void destroy(pointer p) { p->~value_type(); }
};

/* Device memory (hipMalloc): the storage kernels reach fastest on a device
 * whose unified memory is host memory.  The host reaches the elements only
 * through copies; SplitVector's host members make them itself. */
template <class T>
class split_device_allocator {
public:
   typedef T value_type;
   typedef value_type* pointer;
   typedef const value_type* const_pointer;
   typedef value_type& reference;
   typedef const value_type& const_reference;
   typedef ptrdiff_t difference_type;
   typedef size_t size_type;
   template <class U>
   struct rebind {
      typedef split_device_allocator<U> other;
   };
   split_device_allocator() throw() {}
   template <class U>
   split_device_allocator(split_device_allocator<U> const&) throw() {}
   static void* allocate_raw(size_type bytes, const void* = 0) {
      void* ret = nullptr;
      SPLIT_CHECK_ERR(hipMalloc(&ret, bytes));
      if (ret == nullptr) throw std::bad_alloc();
      return ret;
   }
   pointer allocate(size_type n, const void* = 0) {
      assert(n && "allocate 0");
      return static_cast<pointer>(allocate_raw(n * sizeof(value_type)));
   }
   void deallocate(pointer p, size_type n) {
      if (n != 0 && p != 0) SPLIT_CHECK_ERR(hipFree(p));
   }
   static void deallocate(void* p, size_type n) {
      if (n != 0 && p != 0) SPLIT_CHECK_ERR(hipFree(p));
   }
   size_type max_size() const throw() { return std::numeric_limits<size_type>::max() / sizeof(value_type); }
};
template <class T>
struct allocator_is_device_only<split_device_allocator<T>> : std::true_type {};

} // namespace split
