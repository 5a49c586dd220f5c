/* splitvec.h -- split::SplitVector<T> for the MI355X build.
 *
 * The API of kstppd/hashinator include/splitvector/splitvec.h
 * (SplitVector<T,Allocator>, 83-1558): a std::vector-like container whose
 * storage kernels can reach, with explicit prefetch (optimizeGPU /
 * optimizeCPU, 505-538), a device copy of the object (upload, 482-500) and
 * members callable from kernels: size / capacity / operator[] / at / data /
 * front / back / iterators / erase / clear / pop_back (HOSTDEVICE, 617-669,
 * 940-1040, 1114-1260, 1472-1521) and the device mutators device_push_back,
 * device_resize and device_insert (1069-1097, 901-912, 1361-1462).  Growth
 * follows the reference: reserve() doubles unless `eco` (850-884), resize()
 * only grows the capacity when needed (885-900).
 *
 * Where the size lives.  The reference keeps size and capacity in unified
 * memory so host and kernels share them.  On a device without pageable
 * unified memory (hipDeviceAttributeManagedMemory == 0, the MI355X boxes
 * here) unified memory is host memory, and a kernel's atomics on it would
 * cross the host link.  Here the two words live in a 16-byte header in
 * device memory, made by the first upload(): kernels read and update it
 * there (device_push_back is one 64-bit atomic on it), and the host keeps a
 * copy that is refreshed from the header before every host read of the size
 * (one small device-to-host copy; the capacity only changes on the host) and
 * written back after every host change.  Before the first upload() there is
 * no header and the host copy is the only one.  As in the reference, a host
 * change that moves the
 * storage (a reallocation) needs a new upload() before kernels see it, and
 * kernels that change the size must have finished before the host reads it.
 *
 * Storage: split_unified_allocator (the default, as in the reference),
 * split_host_allocator, or split_device_allocator (device memory: the
 * fastest for kernels on a device without pageable unified memory; the host
 * then reaches elements only through copies, which the host-side members do
 * themselves).
 *
 * Differences from the reference's device members, all of them fixes:
 * device_push_back adds to the 64-bit size with a 64-bit atomic (the
 * reference adds to its low 32 bits), a capacity-0 vector rejects it, and it
 * keeps the reference's rule that a push succeeds while size + 1 <
 * capacity (one slot stays free); device_resize and device_insert return
 * false / end() instead of asserting when the capacity is too small; the
 * range device_insert moves the elements after the position (the reference
 * overwrites them), the count-and-value device_insert fills the inserted
 * range (the reference writes after it), and the one-element device_insert
 * writes nothing past the new end.
 */
#pragma once
#include <algorithm>
#include <cassert>
#include <cstring>
#include <initializer_list>
#include <iterator>
#include <new>
#include <stdexcept>
#include <type_traits>
#include <vector>

#include "split_allocators.h"

namespace split {

template <typename T>
using DefaultAllocator = split_unified_allocator<T>;

enum class Residency { host, device };

typedef struct SplitInfo {
   size_t size;
   size_t capacity;
} SplitInfo;

template <typename T>
__global__ void split_fill_kernel(T* p, size_t n, T val) {
   const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
   for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = val;
}

/* True when unified memory is device memory the host can page (HMM); on a
 * device without it hipMallocManaged is host memory and a host loop fills it
 * faster than a kernel. */
inline bool managed_memory_on_device() {
   static int cached = -1;
   if (cached < 0) {
      int dev = 0, v = 0;
      cached = (hipGetDevice(&dev) == hipSuccess &&
                hipDeviceGetAttribute(&v, hipDeviceAttributeManagedMemory, dev) == hipSuccess && v)
                   ? 1
                   : 0;
      (void)hipGetLastError();
   }
   return cached == 1;
}

template <typename T, class Allocator = DefaultAllocator<T>>
class SplitVector {
   static constexpr size_t kDeviceFillThreshold = 1 << 16;
   static constexpr bool kDeviceOnly = allocator_is_device_only<Allocator>::value;

   T* _data = nullptr;
   SplitInfo* _meta = nullptr;      // size and capacity where kernels reach them (device memory; upload() makes it)
   mutable SplitInfo _host{0, 0};   // the host's copy
   size_t _alloc_multiplier = 2;
   Allocator _allocator;
   Residency _location = Residency::host;
   SplitVector* d_vec = nullptr;    // the device copy of this object (upload())

   /* ------------------------------------------------ host copy of the header */
   void _pull() const {
      if (_meta) SPLIT_CHECK_ERR(hipMemcpy(&_host, _meta, sizeof(SplitInfo), hipMemcpyDeviceToHost));
   }
   void _push() {
      if (_meta) SPLIT_CHECK_ERR(hipMemcpy(_meta, &_host, sizeof(SplitInfo), hipMemcpyHostToDevice));
   }

   /* ------------------------------------------------ storage */
   void _fill_on_device(T* p, size_t n, const T& val) {
      const unsigned blocks = static_cast<unsigned>(std::min<size_t>((n + 255) / 256, 1u << 16));
      hipLaunchKernelGGL(split_fill_kernel<T>, dim3(blocks), dim3(256), 0, 0, p, n, val);
      SPLIT_CHECK_ERR(hipGetLastError());
      SPLIT_CHECK_ERR(hipDeviceSynchronize());
      _location = Residency::device;
   }
   void _fill(T* p, size_t n, const T& val) {
      if (n == 0) return;
      if constexpr (kDeviceOnly) {
         static_assert(std::is_trivially_copyable<T>::value,
                       "device-allocated SplitVector needs trivially copyable elements");
         _fill_on_device(p, n, val);
      } else {
         if (std::is_trivially_copyable<T>::value && n >= kDeviceFillThreshold &&
             std::is_same<Allocator, split_unified_allocator<T>>::value && managed_memory_on_device()) {
            _fill_on_device(p, n, val);
         } else {
            for (size_t i = 0; i < n; ++i) _allocator.construct(&p[i], val);
         }
      }
   }
   void _allocate(size_t n, const T& val) {
      _host.size = _host.capacity = n;
      if (n == 0) return;
      _data = _allocator.allocate(n);
      _fill(_data, n, val);
   }
   void _deallocate() {
      if (_data != nullptr) {
         if constexpr (!kDeviceOnly && !std::is_trivially_destructible<T>::value)
            for (size_t i = 0; i < _host.capacity; ++i) _allocator.destroy(&_data[i]);
         _allocator.deallocate(_data, _host.capacity);
         _data = nullptr;
      }
   }
   void _free_device_copy() {
      if (d_vec) SPLIT_CHECK_ERR(hipFree(d_vec));
      if (_meta) SPLIT_CHECK_ERR(hipFree(_meta));
      d_vec = nullptr;
      _meta = nullptr;
   }
   /* Element moves and stores from the host, for any storage. */
   void _move_range(size_t dst, size_t src, size_t n) {
      if (n == 0 || dst == src) return;
      if (!kDeviceOnly) {
         std::memmove(static_cast<void*>(_data + dst), static_cast<const void*>(_data + src), n * sizeof(T));
         return;
      }
      T* tmp = nullptr;
      SPLIT_CHECK_ERR(hipMalloc(&tmp, n * sizeof(T)));
      SPLIT_CHECK_ERR(hipMemcpy(tmp, _data + src, n * sizeof(T), hipMemcpyDeviceToDevice));
      SPLIT_CHECK_ERR(hipMemcpy(_data + dst, tmp, n * sizeof(T), hipMemcpyDeviceToDevice));
      SPLIT_CHECK_ERR(hipFree(tmp));
   }
   void _store(size_t i, const T& val) {
      if (kDeviceOnly)
         SPLIT_CHECK_ERR(hipMemcpy(_data + i, &val, sizeof(T), hipMemcpyHostToDevice));
      else
         _data[i] = val;
   }
   void _store_range(size_t i, const T* src, size_t n) {
      if (n == 0) return;
      if (kDeviceOnly)
         SPLIT_CHECK_ERR(hipMemcpy(_data + i, src, n * sizeof(T), hipMemcpyHostToDevice));
      else
         std::copy(src, src + n, _data + i);
   }
   __host__ __device__ void _rangeCheck(size_t index) const {
#if defined(__HIP_DEVICE_COMPILE__)
      assert(index < _meta->size && "SplitVector index out of range");
#else
      if (index >= size()) throw std::out_of_range("SplitVector index out of range");
#endif
   }

public:
   using value_type = T;

   /* Iterator over the elements (splitvec.h:1103-1226): a pointer with the
    * reference's data() accessor, usable on host and device. */
   template <typename P>
   class basic_iterator {
      P* _p = nullptr;

   public:
      using iterator_category = std::random_access_iterator_tag;
      using value_type = std::remove_const_t<P>;
      using difference_type = int64_t;
      using pointer = P*;
      using reference = P&;
      __host__ __device__ basic_iterator() {}
      __host__ __device__ basic_iterator(P* p) : _p(p) {}
      template <typename Q, typename = std::enable_if_t<std::is_convertible<Q*, P*>::value>>
      __host__ __device__ basic_iterator(const basic_iterator<Q>& o) : _p(o.data()) {}
      __host__ __device__ P* data() const noexcept { return _p; }
      __host__ __device__ P& operator*() const noexcept { return *_p; }
      __host__ __device__ P* operator->() const noexcept { return _p; }
      __host__ __device__ P& operator[](difference_type i) const noexcept { return _p[i]; }
      __host__ __device__ basic_iterator& operator++() noexcept { ++_p; return *this; }
      __host__ __device__ basic_iterator operator++(int) noexcept { basic_iterator t(*this); ++_p; return t; }
      __host__ __device__ basic_iterator& operator--() noexcept { --_p; return *this; }
      __host__ __device__ basic_iterator operator--(int) noexcept { basic_iterator t(*this); --_p; return t; }
      __host__ __device__ basic_iterator& operator+=(difference_type n) noexcept { _p += n; return *this; }
      __host__ __device__ basic_iterator& operator-=(difference_type n) noexcept { _p -= n; return *this; }
      __host__ __device__ basic_iterator operator+(difference_type n) const noexcept { return basic_iterator(_p + n); }
      __host__ __device__ basic_iterator operator-(difference_type n) const noexcept { return basic_iterator(_p - n); }
      __host__ __device__ difference_type operator-(const basic_iterator& o) const noexcept { return _p - o._p; }
      __host__ __device__ bool operator==(const basic_iterator& o) const noexcept { return _p == o._p; }
      __host__ __device__ bool operator!=(const basic_iterator& o) const noexcept { return _p != o._p; }
      __host__ __device__ bool operator<(const basic_iterator& o) const noexcept { return _p < o._p; }
      __host__ __device__ bool operator>(const basic_iterator& o) const noexcept { return _p > o._p; }
      __host__ __device__ bool operator<=(const basic_iterator& o) const noexcept { return _p <= o._p; }
      __host__ __device__ bool operator>=(const basic_iterator& o) const noexcept { return _p >= o._p; }
   };
   using iterator = basic_iterator<T>;
   using const_iterator = basic_iterator<const T>;

   /* ----------------------------------------------------------- construction */
   SplitVector() {}
   explicit SplitVector(size_t size) { _allocate(size, T()); }
   explicit SplitVector(size_t size, const T& val) { _allocate(size, val); }
   SplitVector(const SplitVector& other) {
      const size_t n = other.size();
      _allocate(n, T());
      if (n) SPLIT_CHECK_ERR(hipMemcpy(_data, other._data, n * sizeof(T), hipMemcpyDefault));
   }
   SplitVector(SplitVector&& other) noexcept
       : _data(other._data), _meta(other._meta), _host(other._host), _alloc_multiplier(other._alloc_multiplier),
         _location(other._location), d_vec(other.d_vec) {
      other._data = nullptr;
      other._meta = nullptr;
      other.d_vec = nullptr;
      other._host = SplitInfo{0, 0};
   }
   explicit SplitVector(std::initializer_list<T> init) {
      _allocate(init.size(), T());
      _store_range(0, init.begin(), init.size());
   }
   explicit SplitVector(const std::vector<T>& other) {
      _allocate(other.size(), T());
      _store_range(0, other.data(), other.size());
   }
   ~SplitVector() {
      _deallocate();
      _free_device_copy();
   }

   SplitVector& operator=(const SplitVector& other) {
      if (this == &other) return *this;
      const size_t n = other.size();
      _pull();
      if (_host.capacity < n) {
         _deallocate();
         _host.capacity = n;
         if (n) _data = _allocator.allocate(n);
      }
      _host.size = n;
      if (n) SPLIT_CHECK_ERR(hipMemcpy(_data, other._data, n * sizeof(T), hipMemcpyDefault));
      _push();
      return *this;
   }
   SplitVector& operator=(SplitVector&& other) noexcept {
      if (this == &other) return *this;
      _deallocate();
      _free_device_copy();
      _data = other._data;
      _meta = other._meta;
      _host = other._host;
      _location = other._location;
      d_vec = other.d_vec;
      other._data = nullptr;
      other._meta = nullptr;
      other.d_vec = nullptr;
      other._host = SplitInfo{0, 0};
      return *this;
   }

   /* Copy-assign on a stream (splitvec.h:376-415). */
   void overwrite(const SplitVector& other, hipStream_t stream = 0) {
      if (this == &other) return;
      const size_t n = other.size();
      _pull();
      if (_host.capacity < n) reallocate(n, stream);
      _host.size = n;
      _push();
      if (n) SPLIT_CHECK_ERR(hipMemcpyAsync(_data, other._data, n * sizeof(T), hipMemcpyDefault, stream));
   }

   /* ----------------------------------------------------------- device copy (splitvec.h:482-500) */
   /* Copy this object to device memory (made on the first call, reused
    * after) and return its device address: kernels call the HOSTDEVICE and
    * DEVICEONLY members through it.  The first call also moves the size and
    * capacity into device memory.  `optimize` prefetches unified storage to
    * the device as well. */
   template <bool optimize = true>
   SplitVector* upload(hipStream_t stream = 0) {
      if (!_meta) {
         SPLIT_CHECK_ERR(hipMalloc(reinterpret_cast<void**>(&_meta), sizeof(SplitInfo)));
         SPLIT_CHECK_ERR(hipMemcpy(_meta, &_host, sizeof(SplitInfo), hipMemcpyHostToDevice));
      }
      if (!d_vec) SPLIT_CHECK_ERR(hipMalloc(reinterpret_cast<void**>(&d_vec), sizeof(SplitVector)));
      SPLIT_CHECK_ERR(hipMemcpyAsync(d_vec, this, sizeof(SplitVector), hipMemcpyHostToDevice, stream));
      if (optimize) optimizeGPU(stream);
      SPLIT_CHECK_ERR(hipStreamSynchronize(stream)); // the object's bytes are read from the host before it changes
      return d_vec;
   }
   const SplitVector* device_pointer() const noexcept { return d_vec; }

   /* ----------------------------------------------------------- residency (splitvec.h:505-600) */
   void optimizeGPU(hipStream_t stream = 0) noexcept {
      _location = Residency::device;
      const size_t cap = capacity();
      if (cap == 0 || kDeviceOnly || !std::is_same<Allocator, split_unified_allocator<T>>::value) return;
      int device = 0;
      SPLIT_CHECK_ERR(hipGetDevice(&device));
      SPLIT_CHECK_ERR(hipMemPrefetchAsync(_data, cap * sizeof(T), device, stream));
   }
   void optimizeCPU(hipStream_t stream = 0) noexcept {
      _location = Residency::host;
      const size_t cap = capacity();
      if (cap == 0 || kDeviceOnly || !std::is_same<Allocator, split_unified_allocator<T>>::value) return;
      SPLIT_CHECK_ERR(hipMemPrefetchAsync(_data, cap * sizeof(T), hipCpuDeviceId, stream));
      SPLIT_CHECK_ERR(hipStreamSynchronize(stream));
   }
   void streamAttach(hipStream_t s, uint32_t flags = hipMemAttachSingle) {
      const size_t cap = capacity();
      if (cap && std::is_same<Allocator, split_unified_allocator<T>>::value)
         SPLIT_CHECK_ERR(hipStreamAttachMemAsync(s, _data, cap * sizeof(T), flags));
   }
   /* Size and capacity into *dst (splitvec.h:568-571): asynchronous on `s`
    * when they live on the device. */
   void copyMetadata(SplitInfo* dst, hipStream_t s = 0) {
      if (_meta)
         SPLIT_CHECK_ERR(hipMemcpyAsync(dst, _meta, sizeof(SplitInfo), hipMemcpyDeviceToHost, s));
      else
         *dst = _host;
   }
   void memAdvise(hipMemoryAdvise advice, int device = -1, hipStream_t = 0) {
      if (device == -1) SPLIT_CHECK_ERR(hipGetDevice(&device));
      const size_t cap = capacity();
      if (cap && std::is_same<Allocator, split_unified_allocator<T>>::value)
         SPLIT_CHECK_ERR(hipMemAdvise(_data, cap * sizeof(T), advice, device));
   }
   Residency residency() const noexcept { return _location; }

   /* Swaps contents, header and device copy, so each object's device copy
    * still describes it. */
   void swap(SplitVector& other) noexcept {
      std::swap(_data, other._data);
      std::swap(_meta, other._meta);
      std::swap(_host, other._host);
      std::swap(_location, other._location);
      std::swap(d_vec, other.d_vec);
   }

   /* ----------------------------------------------------------- access, host and device */
   __host__ __device__ const size_t& size() const noexcept {
#if defined(__HIP_DEVICE_COMPILE__)
      return _meta->size;
#else
      _pull();
      return _host.size;
#endif
   }
   /* Only the host changes the capacity, so its copy is always current. */
   __host__ __device__ size_t capacity() const noexcept {
#if defined(__HIP_DEVICE_COMPILE__)
      return _meta->capacity;
#else
      return _host.capacity;
#endif
   }
   __host__ __device__ T& operator[](size_t index) noexcept { return _data[index]; }
   __host__ __device__ const T& operator[](size_t index) const noexcept { return _data[index]; }
   __host__ __device__ T& at(size_t index) {
      _rangeCheck(index);
      return _data[index];
   }
   __host__ __device__ const T& at(size_t index) const {
      _rangeCheck(index);
      return _data[index];
   }
   __host__ __device__ T* data() noexcept { return _data; }
   __host__ __device__ const T* data() const noexcept { return _data; }
   __host__ __device__ T& back() noexcept { return _data[size() - 1]; }
   __host__ __device__ const T& back() const noexcept { return _data[size() - 1]; }
   __host__ __device__ T& front() noexcept { return _data[0]; }
   __host__ __device__ const T& front() const noexcept { return _data[0]; }
   __host__ __device__ bool empty() const noexcept { return size() == 0; }
   __host__ __device__ iterator begin() noexcept { return iterator(_data); }
   __host__ __device__ const_iterator begin() const noexcept { return const_iterator(_data); }
   __host__ __device__ iterator end() noexcept { return iterator(_data + size()); }
   __host__ __device__ const_iterator end() const noexcept { return const_iterator(_data + size()); }
   Allocator get_allocator() const noexcept { return _allocator; }

   /* ----------------------------------------------------------- removal, host and device (splitvec.h:940-975, 1472-1521) */
   __host__ __device__ void clear() noexcept {
#if defined(__HIP_DEVICE_COMPILE__)
      _meta->size = 0;
#else
      _pull();
      _host.size = 0;
      _push();
#endif
   }
   __host__ __device__ void remove_from_back(size_t n) noexcept {
#if defined(__HIP_DEVICE_COMPILE__)
      _meta->size = n > _meta->size ? 0 : _meta->size - n;
#else
      _pull();
      _host.size = n > _host.size ? 0 : _host.size - n;
      _push();
#endif
   }
   __host__ __device__ void pop_back() noexcept { remove_from_back(1); }
   __host__ __device__ iterator erase(iterator it) noexcept { return erase(it, it + 1); }
   __host__ __device__ iterator erase(iterator p0, iterator p1) noexcept {
      const size_t a = static_cast<size_t>(p0.data() - _data), b = static_cast<size_t>(p1.data() - _data);
#if defined(__HIP_DEVICE_COMPILE__)
      const size_t sz = _meta->size;
      if (b > a && b <= sz) {
         for (size_t i = a; i + (b - a) < sz; ++i) _data[i] = _data[i + (b - a)];
         _meta->size = sz - (b - a);
      }
#else
      _pull();
      if (b > a && b <= _host.size) {
         _move_range(a, b, _host.size - b);
         _host.size -= b - a;
         _push();
      }
#endif
      return iterator(_data + a);
   }

   /* ----------------------------------------------------------- device mutation (splitvec.h:901-912, 1069-1097, 1361-1462) */
   /* Append from any number of threads at once: one 64-bit atomic on the
    * size takes a slot.  Succeeds while the new size stays below the
    * capacity (the reference keeps one slot free); otherwise the size is
    * restored and false is returned. */
   __device__ bool device_push_back(const T& val) {
      size_t* sz = &_meta->size;
      const size_t old = __hip_atomic_fetch_add(sz, size_t(1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (old + 1 >= _meta->capacity) {
         __hip_atomic_fetch_sub(sz, size_t(1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
         return false;
      }
      _data[old] = val;
      return true;
   }
   __device__ bool device_push_back(T&& val) {
      const T v(static_cast<T&&>(val));
      return device_push_back(v);
   }
   /* One thread: set the size to newSize (value-initialising new elements
    * when `construct`); false, and no change, beyond the capacity. */
   __device__ bool device_resize(size_t newSize, bool construct = true) {
      if (newSize > _meta->capacity) return false;
      if (construct)
         for (size_t i = _meta->size; i < newSize; ++i) new (&_data[i]) T();
      _meta->size = newSize;
      return true;
   }
   /* One thread: insert before `it`, moving the elements after it; returns
    * the first inserted element, or end() (no change) when the position or
    * the capacity does not allow it. */
   __device__ iterator device_insert(iterator it, const T& val) noexcept { return device_insert(it, size_t(1), val); }
   __device__ iterator device_insert(iterator it, const size_t elements, const T& val) {
      const size_t sz = _meta->size;
      const size_t index = static_cast<size_t>(it.data() - _data);
      if (index > sz || sz + elements > _meta->capacity) return end();
      for (size_t i = sz; i > index; --i) _data[i - 1 + elements] = _data[i - 1];
      for (size_t i = 0; i < elements; ++i) _data[index + i] = val;
      _meta->size = sz + elements;
      return iterator(_data + index);
   }
   template <typename InputIterator, class = typename std::enable_if<!std::is_integral<InputIterator>::value>::type>
   __device__ iterator device_insert(iterator it, InputIterator p0, InputIterator p1) noexcept {
      const size_t sz = _meta->size;
      const size_t index = static_cast<size_t>(it.data() - _data);
      const size_t count = static_cast<size_t>(p1 - p0);
      if (index > sz || sz + count > _meta->capacity) return end();
      for (size_t i = sz; i > index; --i) _data[i - 1 + count] = _data[i - 1];
      size_t j = index;
      for (InputIterator p = p0; p != p1; ++p) _data[j++] = *p;
      _meta->size = sz + count;
      return iterator(_data + index);
   }

   /* ----------------------------------------------------------- capacity (splitvec.h:795-984) */
   void reallocate(size_t requested_space, hipStream_t stream = 0) {
      _pull();
      if (requested_space == 0) {
         _deallocate();
         _host.size = _host.capacity = 0;
         _push();
         return;
      }
      if (requested_space < _host.size) requested_space = _host.size;
      T* fresh = _allocator.allocate(requested_space);
      if (_host.size)
         SPLIT_CHECK_ERR(hipMemcpyAsync(fresh, _data, _host.size * sizeof(T), hipMemcpyDefault, stream));
      SPLIT_CHECK_ERR(hipStreamSynchronize(stream));
      _deallocate();
      _data = fresh;
      _host.capacity = requested_space;
      _push();
   }
   void reserve(size_t requested_space, bool eco = false, hipStream_t stream = 0) {
      if (requested_space <= capacity()) return;
      reallocate(eco ? requested_space : requested_space * _alloc_multiplier, stream);
   }
   void resize(size_t newSize, bool eco = false, hipStream_t stream = 0) {
      if (newSize > capacity()) reserve(newSize, eco, stream);
      _host.size = newSize;
      _push();
   }
   void grow(hipStream_t stream = 0) { reserve(capacity() + 1, false, stream); }
   void shrink_to_fit(hipStream_t stream = 0) {
      if (capacity() != _host.size) reallocate(_host.size, stream);
   }

   /* ----------------------------------------------------------- host mutation */
   void push_back(const T& val) {
      const size_t n = size();
      if (n + 1 > _host.capacity) reserve(n + 1);
      _store(n, val);
      _host.size = n + 1;
      _push();
   }
   template <typename... Args>
   void emplace_back(Args&&... args) {
      push_back(T(std::forward<Args>(args)...));
   }
   iterator insert(iterator it, const T& val) { return insert(it, size_t(1), val); }
   iterator insert(iterator it, size_t n, const T& val) {
      const size_t at = static_cast<size_t>(it.data() - _data);
      const size_t sz = size();
      if (at > sz) throw std::out_of_range("SplitVector insert position out of range");
      if (sz + n > _host.capacity) reserve(sz + n);
      _move_range(at + n, at, sz - at);
      if (kDeviceOnly) {
         std::vector<T> fill(n, val);
         _store_range(at, fill.data(), n);
      } else {
         for (size_t i = 0; i < n; ++i) _data[at + i] = val;
      }
      _host.size = sz + n;
      _push();
      return iterator(_data + at);
   }
   template <typename InputIterator, typename = typename std::enable_if<!std::is_integral<InputIterator>::value>::type>
   iterator insert(iterator it, InputIterator p0, InputIterator p1) {
      const size_t at = static_cast<size_t>(it.data() - _data);
      const std::vector<T> src(p0, p1);
      const size_t n = src.size();
      const size_t sz = size();
      if (at > sz) throw std::out_of_range("SplitVector insert position out of range");
      if (sz + n > _host.capacity) reserve(sz + n);
      _move_range(at + n, at, sz - at);
      _store_range(at, src.data(), n);
      _host.size = sz + n;
      _push();
      return iterator(_data + at);
   }
   template <class... Args>
   iterator emplace(iterator pos, Args&&... args) {
      return insert(pos, T(std::forward<Args>(args)...));
   }

   /* The elements as a host vector (a copy for any storage). */
   std::vector<T> to_host() const {
      std::vector<T> out(size());
      if (!out.empty()) SPLIT_CHECK_ERR(hipMemcpy(out.data(), _data, out.size() * sizeof(T), hipMemcpyDefault));
      return out;
   }
};

template <typename T, class Allocator>
inline bool operator==(const SplitVector<T, Allocator>& lhs, const SplitVector<T, Allocator>& rhs) {
   if (lhs.size() != rhs.size()) return false;
   const std::vector<T> a = lhs.to_host(), b = rhs.to_host();
   for (size_t i = 0; i < a.size(); i++)
      if (!(a[i] == b[i])) return false;
   return true;
}
template <typename T, class Allocator>
inline bool operator!=(const SplitVector<T, Allocator>& lhs, const SplitVector<T, Allocator>& rhs) {
   return !(rhs == lhs);
}

} // namespace split
