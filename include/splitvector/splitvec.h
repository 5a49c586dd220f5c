/* splitvec.h -- split::SplitVector<T> for the MI355X build.
 *
 * Host-side API of kstppd/hashinator include/splitvector/splitvec.h
 * (SplitVector<T,Allocator>, 83-1558): a std::vector-like container whose
 * storage is device-accessible unified memory, with explicit prefetch
 * (optimizeGPU / optimizeCPU, 505-538).  Growth follows the reference:
 * reserve() doubles unless `eco` (850-884), resize() only grows the
 * capacity when needed (885-900).
 *
 * Differences by design: the size and capacity are host members (the
 * device-side mutators that needed them in unified memory -- device_push_back,
 * device_insert, device_resize -- are not provided here); large
 * trivially-copyable fills run as a device kernel instead of a host loop.
 */
#pragma once
#include <algorithm>
#include <cassert>
#include <cstring>
#include <initializer_list>
#include <stdexcept>
#include <type_traits>
#include <vector>

#include "split_allocators.h"

namespace split {

template <typename T>
using DefaultAllocator = split_unified_allocator<T>;

enum class Residency { host, device };

struct SplitInfo {
   size_t size;
   size_t capacity;
};

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

   T* _data = nullptr;
   size_t _size = 0;
   size_t _capacity = 0;
   size_t _alloc_multiplier = 2;
   Allocator _allocator;
   Residency _location = Residency::host;

   void _fill(T* p, size_t n, const T& val) {
      if (n == 0) return;
      if (std::is_trivially_copyable<T>::value && n >= kDeviceFillThreshold &&
          std::is_same<Allocator, split_unified_allocator<T>>::value && managed_memory_on_device()) {
         const unsigned blocks = static_cast<unsigned>(std::min<size_t>((n + 255) / 256, 1u << 16));
         hipLaunchKernelGGL(split_fill_kernel<T>, dim3(blocks), dim3(256), 0, 0, p, n, val);
         SPLIT_CHECK_ERR(hipGetLastError());
         SPLIT_CHECK_ERR(hipDeviceSynchronize());
         _location = Residency::device;
      } else {
         for (size_t i = 0; i < n; ++i) _allocator.construct(&p[i], val);
      }
   }
   void _allocate(size_t n, const T& val) {
      _size = _capacity = n;
      if (n == 0) return;
      _data = _allocator.allocate(n);
      _fill(_data, n, val);
   }
   void _deallocate() {
      if (_data != nullptr) {
         if (!std::is_trivially_destructible<T>::value)
            for (size_t i = 0; i < _capacity; ++i) _allocator.destroy(&_data[i]);
         _allocator.deallocate(_data, _capacity);
         _data = nullptr;
      }
   }
   void _rangeCheck(size_t index) const {
      if (index >= _size) throw std::out_of_range("SplitVector index out of range");
   }

public:
   using value_type = T;
   using iterator = T*;
   using const_iterator = const T*;

   /* ----------------------------------------------------------- construction */
   SplitVector() {}
   explicit SplitVector(size_t size) { _allocate(size, T()); }
   explicit SplitVector(size_t size, const T& val) { _allocate(size, val); }
   SplitVector(const SplitVector& other) {
      _allocate(other._size, T());
      if (_size) SPLIT_CHECK_ERR(hipMemcpy(_data, other._data, _size * sizeof(T), hipMemcpyDefault));
   }
   SplitVector(SplitVector&& other) noexcept
       : _data(other._data), _size(other._size), _capacity(other._capacity), _location(other._location) {
      other._data = nullptr;
      other._size = other._capacity = 0;
   }
   explicit SplitVector(std::initializer_list<T> init) {
      _allocate(init.size(), T());
      std::copy(init.begin(), init.end(), _data);
   }
   explicit SplitVector(const std::vector<T>& other) {
      _allocate(other.size(), T());
      std::copy(other.begin(), other.end(), _data);
   }
   ~SplitVector() { _deallocate(); }

   SplitVector& operator=(const SplitVector& other) {
      if (this == &other) return *this;
      if (_capacity < other._size) {
         _deallocate();
         _capacity = other._size;
         if (_capacity) _data = _allocator.allocate(_capacity);
      }
      _size = other._size;
      if (_size) {
         SPLIT_CHECK_ERR(hipMemcpy(_data, other._data, _size * sizeof(T), hipMemcpyDefault));
      }
      return *this;
   }
   SplitVector& operator=(SplitVector&& other) noexcept {
      if (this == &other) return *this;
      _deallocate();
      _data = other._data;
      _size = other._size;
      _capacity = other._capacity;
      _location = other._location;
      other._data = nullptr;
      other._size = other._capacity = 0;
      return *this;
   }

   /* Copy-assign on a stream (splitvec.h:376-415). */
   void overwrite(const SplitVector& other, hipStream_t stream = 0) {
      if (this == &other) return;
      if (_capacity < other._size) reallocate(other._size, stream);
      _size = other._size;
      if (_size) SPLIT_CHECK_ERR(hipMemcpyAsync(_data, other._data, _size * sizeof(T), hipMemcpyDefault, stream));
   }

   /* ----------------------------------------------------------- residency (splitvec.h:505-600) */
   void optimizeGPU(hipStream_t stream = 0) noexcept {
      _location = Residency::device;
      if (_capacity == 0) return;
      int device = 0;
      SPLIT_CHECK_ERR(hipGetDevice(&device));
      SPLIT_CHECK_ERR(hipMemPrefetchAsync(_data, _capacity * sizeof(T), device, stream));
   }
   void optimizeCPU(hipStream_t stream = 0) noexcept {
      _location = Residency::host;
      if (_capacity == 0) return;
      SPLIT_CHECK_ERR(hipMemPrefetchAsync(_data, _capacity * sizeof(T), hipCpuDeviceId, stream));
      SPLIT_CHECK_ERR(hipStreamSynchronize(stream));
   }
   void streamAttach(hipStream_t s, uint32_t flags = hipMemAttachSingle) {
      if (_capacity) SPLIT_CHECK_ERR(hipStreamAttachMemAsync(s, _data, _capacity * sizeof(T), flags));
   }
   void copyMetadata(SplitInfo* dst, hipStream_t = 0) {
      dst->size = _size;
      dst->capacity = _capacity;
   }
   void memAdvise(hipMemoryAdvise advice, int device = -1, hipStream_t = 0) {
      if (device == -1) SPLIT_CHECK_ERR(hipGetDevice(&device));
      if (_capacity) SPLIT_CHECK_ERR(hipMemAdvise(_data, _capacity * sizeof(T), advice, device));
   }
   Residency residency() const noexcept { return _location; }

   void swap(SplitVector& other) noexcept {
      std::swap(_data, other._data);
      std::swap(_size, other._size);
      std::swap(_capacity, other._capacity);
      std::swap(_location, other._location);
   }

   /* ----------------------------------------------------------- access */
   const size_t& size() const noexcept { return _size; }
   size_t capacity() const noexcept { return _capacity; }
   T& operator[](size_t index) noexcept { return _data[index]; }
   const T& operator[](size_t index) const noexcept { return _data[index]; }
   T& at(size_t index) {
      _rangeCheck(index);
      return _data[index];
   }
   const T& at(size_t index) const {
      _rangeCheck(index);
      return _data[index];
   }
   T* data() noexcept { return _data; }
   const T* data() const noexcept { return _data; }
   T& back() noexcept { return _data[_size - 1]; }
   const T& back() const noexcept { return _data[_size - 1]; }
   T& front() noexcept { return _data[0]; }
   const T& front() const noexcept { return _data[0]; }
   bool empty() const noexcept { return _size == 0; }
   iterator begin() noexcept { return _data; }
   const_iterator begin() const noexcept { return _data; }
   iterator end() noexcept { return _data + _size; }
   const_iterator end() const noexcept { return _data + _size; }
   Allocator get_allocator() const noexcept { return _allocator; }

   /* ----------------------------------------------------------- capacity (splitvec.h:795-984) */
   void reallocate(size_t requested_space, hipStream_t stream = 0) {
      if (requested_space == 0) {
         _deallocate();
         _size = _capacity = 0;
         return;
      }
      if (requested_space < _size) requested_space = _size;
      T* fresh = _allocator.allocate(requested_space);
      if (_size) SPLIT_CHECK_ERR(hipMemcpyAsync(fresh, _data, _size * sizeof(T), hipMemcpyDefault, stream));
      SPLIT_CHECK_ERR(hipStreamSynchronize(stream));
      _deallocate();
      _data = fresh;
      _capacity = requested_space;
   }
   void reserve(size_t requested_space, bool eco = false, hipStream_t stream = 0) {
      if (requested_space <= _capacity) return;
      reallocate(eco ? requested_space : requested_space * _alloc_multiplier, stream);
   }
   void resize(size_t newSize, bool eco = false, hipStream_t stream = 0) {
      if (newSize <= _size) {
         _size = newSize;
         return;
      }
      reserve(newSize, eco, stream);
      _size = newSize;
   }
   void grow(hipStream_t stream = 0) { reserve(_capacity + 1, false, stream); }
   void shrink_to_fit(hipStream_t stream = 0) {
      if (_capacity != _size) reallocate(_size, stream);
   }
   void clear() noexcept { _size = 0; }
   void pop_back() noexcept {
      if (_size > 0) --_size;
   }
   void remove_from_back(size_t n) noexcept { _size = n > _size ? 0 : _size - n; }

   /* ----------------------------------------------------------- host mutation */
   void push_back(const T& val) {
      if (_size + 1 > _capacity) reserve(_size + 1);
      _data[_size++] = val;
   }
   template <typename... Args>
   void emplace_back(Args&&... args) {
      push_back(T(std::forward<Args>(args)...));
   }
   iterator insert(iterator it, const T& val) { return insert(it, size_t(1), val); }
   iterator insert(iterator it, size_t n, const T& val) {
      const size_t at = static_cast<size_t>(it - _data);
      if (at > _size) throw std::out_of_range("SplitVector insert position out of range");
      if (_size + n > _capacity) reserve(_size + n);
      std::memmove(_data + at + n, _data + at, (_size - at) * sizeof(T));
      for (size_t i = 0; i < n; ++i) _data[at + i] = val;
      _size += n;
      return _data + at;
   }
   template <typename InputIterator, typename = typename std::enable_if<!std::is_integral<InputIterator>::value>::type>
   iterator insert(iterator it, InputIterator p0, InputIterator p1) {
      const size_t at = static_cast<size_t>(it - _data);
      const size_t n = static_cast<size_t>(std::distance(p0, p1));
      if (at > _size) throw std::out_of_range("SplitVector insert position out of range");
      if (_size + n > _capacity) reserve(_size + n);
      std::memmove(_data + at + n, _data + at, (_size - at) * sizeof(T));
      std::copy(p0, p1, _data + at);
      _size += n;
      return _data + at;
   }
   iterator erase(iterator it) noexcept { return erase(it, it + 1); }
   iterator erase(iterator p0, iterator p1) noexcept {
      const size_t a = static_cast<size_t>(p0 - _data), b = static_cast<size_t>(p1 - _data);
      if (b > a && b <= _size) {
         std::memmove(_data + a, _data + b, (_size - b) * sizeof(T));
         _size -= b - a;
      }
      return _data + a;
   }
   template <class... Args>
   iterator emplace(iterator pos, Args&&... args) {
      return insert(pos, T(std::forward<Args>(args)...));
   }
};

template <typename T, class Allocator>
inline bool operator==(const SplitVector<T, Allocator>& lhs, const SplitVector<T, Allocator>& rhs) noexcept {
   if (lhs.size() != rhs.size()) return false;
   for (size_t i = 0; i < lhs.size(); i++)
      if (!(lhs[i] == rhs[i])) return false;
   return true;
}
template <typename T, class Allocator>
inline bool operator!=(const SplitVector<T, Allocator>& lhs, const SplitVector<T, Allocator>& rhs) noexcept {
   return !(rhs == lhs);
}

} // namespace split
