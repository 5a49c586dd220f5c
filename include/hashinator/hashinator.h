/* hashinator.h -- Hashinator::Hashmap for the MI355X (gfx950) build.
 *
 * Drop-in for the accelerated host API of kstppd/hashinator
 * include/hashinator/hashinator.h (Hashmap, 59-1802): the same class
 * template, method names, argument meaning, synchronous behaviour and
 * auto-grow rule, over libhashinator_amd's C-ABI through the DeviceHasher
 * plugin (hashers.h).
 *
 * Residency.  The reference keeps the buckets and the Info in unified
 * memory and reads them from the host.  MI355X without HMM serves
 * hipMallocManaged from host memory (hipDeviceAttributeManagedMemory = 0),
 * where the probe kernels run three orders of magnitude slower, so here the
 * bucket array and the Info live in device memory.  The host keeps a mirror
 * of the Info, refreshed after every call (each call synchronises its stream,
 * as in the reference), and host-side lookups (find, count, iteration) read
 * a host copy of the buckets taken on first use after a change.
 * expose_bucketdata() returns the device pointer; expose_mapinfo() the host
 * mirror (call pushInfo() after editing it).
 *
 * Keys, values and pairs handed to insert/retrieve/erase may be device
 * memory (fastest) or unified memory such as a split::SplitVector.
 *
 * Instantiations: Hashmap<uint32_t,uint32_t> and Hashmap<uint64_t,uint64_t>
 * with the default sentinels and hash.
 *
 * The reference's device-only API is provided (upload, the warp members,
 * device_find/device_insert/device_erase/set_element/read_element and device
 * iterators); not provided: the host insertion path (_at, operator[],
 * erase(key) on the host).
 */
#pragma once
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <limits>
#include <stdexcept>
#include <vector>

#include "../common.h"
#include "../splitvector/split_tools.h"
#include "../splitvector/splitvec.h"
#include "defaults.h"
#include "hash_pair.h"
#include "hashers.h"
#include "hashfunctions.h"

namespace Hashinator {

template <typename T>
using DefaultMetaAllocator = split::split_unified_allocator<T>;
#define DefaultHasher                                                                                                  \
   Hashers::Hasher<KEY_TYPE, VAL_TYPE, HashFunction, EMPTYBUCKET, TOMBSTONE, defaults::WARPSIZE,                       \
                   defaults::elementsPerWarp>

using MapInfo = Hashinator::Info;

template <typename KEY_TYPE, typename VAL_TYPE, KEY_TYPE EMPTYBUCKET = std::numeric_limits<KEY_TYPE>::max(),
          KEY_TYPE TOMBSTONE = EMPTYBUCKET - 1, class HashFunction = HashFunctions::Fibonacci<KEY_TYPE>,
          class DeviceHasher = DefaultHasher, class Meta_Allocator = DefaultMetaAllocator<MapInfo>>
class Hashmap {
   using pair_t = hash_pair<KEY_TYPE, VAL_TYPE>;
   static constexpr bool W64 = sizeof(KEY_TYPE) == 8;

   pair_t* _buckets = nullptr; // device memory, 2^sizePower slots
   size_t _nbuckets = 0;
   MapInfo* _dinfo = nullptr;  // device memory
   mutable MapInfo _info;      // host mirror of *_dinfo
   mutable std::vector<pair_t> _view;
   mutable bool _view_ok = false;
   Hashmap* _device_map = nullptr; // upload(): this object's copy in device memory
   hn_tune _tune{0, 0, 0, 0};
   bool _has_tune = false;
   float _auto_reorder = 1.0f / 16; // reorder after inserts of at least this fraction of the buckets
   float _auto_build = 1.0f / 16;   // bulk-build inserts of at least this fraction into an empty table
   size_t _build_min_keys = size_t(1) << 17; // ... and of at least this many keys (below, launches dominate)
   float _auto_merge = 0.0f; // merge inserts of at least this fraction into a table holding keys (0: off)

   uint32_t hash(KEY_TYPE in) const { return static_cast<uint32_t>(HashFunction::_hash(in, _info.sizePower)); }
   const hn_tune* tune() const { return _has_tune ? &_tune : nullptr; }
   hn_info* dinfo() const { return reinterpret_cast<hn_info*>(_dinfo); }
   static hn_stream_t st(hipStream_t s) { return reinterpret_cast<hn_stream_t>(s); }
   static void chk(int rc) { Hashers::detail::check(rc); }

   static pair_t* new_table(size_t n, hipStream_t s) {
      pair_t* p = nullptr;
      SPLIT_CHECK_ERR(hipMalloc(&p, n * sizeof(pair_t)));
      chk(W64 ? hn_table_init_u64(p, n, st(s)) : hn_table_init_u32(p, n, st(s)));
      return p;
   }
   void pull(hipStream_t s) const {
      SPLIT_CHECK_ERR(hipStreamSynchronize(s));
      SPLIT_CHECK_ERR(hipMemcpy(&_info, _dinfo, sizeof(MapInfo), hipMemcpyDeviceToHost));
   }
   void changed(hipStream_t s) {
      _view_ok = false;
      pull(s);
   }
   void set_status(status code) {
      _info.err = code;
      pushInfo();
   }
   void alloc(int sizePower) {
      _nbuckets = size_t(1) << sizePower;
      _buckets = new_table(_nbuckets, 0);
      SPLIT_CHECK_ERR(hipMalloc(&_dinfo, sizeof(MapInfo)));
      _info = MapInfo(sizePower);
      pushInfo();
   }
   void release() {
      if (_buckets) SPLIT_CHECK_ERR(hipFree(_buckets));
      if (_dinfo) SPLIT_CHECK_ERR(hipFree(_dinfo));
      if (_device_map) SPLIT_CHECK_ERR(hipFree(_device_map));
      _buckets = nullptr;
      _dinfo = nullptr;
      _device_map = nullptr;
      _nbuckets = 0;
   }
   const std::vector<pair_t>& view() const {
      if (!_view_ok) {
         _view.resize(_nbuckets);
         SPLIT_CHECK_ERR(hipDeviceSynchronize());
         SPLIT_CHECK_ERR(hipMemcpy(_view.data(), _buckets, _nbuckets * sizeof(pair_t), hipMemcpyDeviceToHost));
         _view_ok = true;
      }
      return _view;
   }

public:
   Hashmap() : Hashmap(5) {}
   Hashmap(int sizepower) { alloc(sizepower); }
   Hashmap(const Hashmap& other) {
      alloc(other._info.sizePower);
      SPLIT_CHECK_ERR(hipMemcpy(_buckets, other._buckets, _nbuckets * sizeof(pair_t), hipMemcpyDeviceToDevice));
      other.pull(0);
      _info = other._info;
      pushInfo();
      _tune = other._tune;
      _has_tune = other._has_tune;
   }
   Hashmap(Hashmap&& other) noexcept
       : _buckets(other._buckets), _nbuckets(other._nbuckets), _dinfo(other._dinfo), _info(other._info),
         _device_map(other._device_map), _tune(other._tune), _has_tune(other._has_tune) {
      other._buckets = nullptr;
      other._dinfo = nullptr;
      other._device_map = nullptr;
      other._nbuckets = 0;
   }
   Hashmap& operator=(const Hashmap& other) {
      if (this == &other) return *this;
      overwrite(other, 0);
      SPLIT_CHECK_ERR(hipStreamSynchronize(0));
      return *this;
   }
   Hashmap& operator=(Hashmap&& other) noexcept {
      if (this == &other) return *this;
      release();
      std::swap(_buckets, other._buckets);
      std::swap(_dinfo, other._dinfo);
      std::swap(_nbuckets, other._nbuckets);
      std::swap(_device_map, other._device_map);
      _info = other._info;
      _view_ok = false;
      return *this;
   }
   ~Hashmap() { release(); }

   /* Copy-assign on a stream (hashinator.h:165-174). */
   void overwrite(const Hashmap& other, hipStream_t stream = 0) {
      if (this == &other) return;
      if (_nbuckets != other._nbuckets) {
         if (_buckets) SPLIT_CHECK_ERR(hipFree(_buckets));
         _nbuckets = other._nbuckets;
         SPLIT_CHECK_ERR(hipMalloc(&_buckets, _nbuckets * sizeof(pair_t)));
      }
      SPLIT_CHECK_ERR(hipMemcpyAsync(_buckets, other._buckets, _nbuckets * sizeof(pair_t), hipMemcpyDeviceToDevice,
                                     stream));
      SPLIT_CHECK_ERR(hipMemcpyAsync(_dinfo, other._dinfo, sizeof(MapInfo), hipMemcpyDeviceToDevice, stream));
      changed(stream);
   }
   /* Copy of the Info to any host or device address (hashinator.h:228-230). */
   void copyMetadata(MapInfo* dst, hipStream_t s = 0) {
      SPLIT_CHECK_ERR(hipMemcpyAsync(dst, _dinfo, sizeof(MapInfo), hipMemcpyDefault, s));
   }
   /* Write the host mirror back to the device (after editing expose_mapinfo()). */
   void pushInfo() { SPLIT_CHECK_ERR(hipMemcpy(_dinfo, &_info, sizeof(MapInfo), hipMemcpyHostToDevice)); }

   /* Launch tuning for this map's calls (an addition): probe width 1..64
    * lanes, block size, grid cap; zeros mean the library defaults. */
   void setTune(const hn_tune& t) {
      _tune = t;
      _has_tune = true;
   }

   /* ------------------------------------------------------------ rehash / resize */

   /* Host rehash (hashinator.h:236-279): the serial algorithm with the
    * 32-probe placement limit, retried one size larger on failure; the result
    * is written back to the device. */
   void rehash(int newSizePower) {
      const std::vector<pair_t>& old = view();
      for (;;) {
         if (newSizePower > 32)
            throw std::out_of_range("Hashmap ran into rehashing catastrophe and exceeded 32bit buckets.");
         std::vector<pair_t> nb(size_t(1) << newSizePower, pair_t(EMPTYBUCKET, VAL_TYPE()));
         const size_t bitMask = nb.size() - 1;
         bool ok = true;
         for (size_t e = 0; e < old.size() && ok; ++e) {
            const pair_t el = old[e];
            if (el.first == EMPTYBUCKET || el.first == TOMBSTONE) continue;
            const uint32_t h = static_cast<uint32_t>(HashFunction::_hash(el.first, newSizePower));
            bool found = false;
            for (uint32_t i = 0; i < 32; ++i) { // the CPU BUCKET_OVERFLOW of the reference
               pair_t& c = nb[static_cast<uint32_t>(h + i) & bitMask];
               if (c.first == EMPTYBUCKET) {
                  c = el;
                  found = true;
                  break;
               }
            }
            ok = found;
         }
         if (!ok) {
            ++newSizePower;
            continue;
         }
         SPLIT_CHECK_ERR(hipFree(_buckets));
         _nbuckets = nb.size();
         SPLIT_CHECK_ERR(hipMalloc(&_buckets, _nbuckets * sizeof(pair_t)));
         SPLIT_CHECK_ERR(hipMemcpy(_buckets, nb.data(), _nbuckets * sizeof(pair_t), hipMemcpyHostToDevice));
         _info.sizePower = newSizePower;
         _info.currentMaxBucketOverflow = defaults::BUCKET_OVERFLOW;
         _info.tombstoneCounter = 0;
         _info.reserved1 = 0; // arrival order, not Robin Hood order
         pushInfo();
         _view.swap(nb);
         _view_ok = true;
         return;
      }
   }

   /* device_rehash (hashinator.h:284-336): every live element moves into a
    * fresh table of 2^newSizePower slots (a new allocation, also when the
    * size is unchanged); err = success iff the fill is preserved. */
   template <bool prefetches = true>
   void device_rehash(int newSizePower, hipStream_t s = 0) {
      if (newSizePower > 32)
         throw std::out_of_range("Hashmap ran into rehashing catastrophe and exceeded 32bit buckets.");
      pull(s);
      const size_t fill = _info.fill;
      // insert(valid, n, targetLF = 1) grows further when the elements do not fit (hashinator.h:332, 1329)
      if (fill > 0) {
         const int need = static_cast<int>(std::ceil(std::log2(static_cast<double>(fill))));
         if (need > newSizePower) newSizePower = need;
      }
      pair_t* fresh = nullptr;
      SPLIT_CHECK_ERR(hipMalloc(&fresh, (size_t(1) << newSizePower) * sizeof(pair_t)));
      const int osp = static_cast<int>(_info.sizePower);
      // Growing a map of many keys: one bulk build from the old table's live
      // slots (Robin Hood order directly); else reinsert them by probing.
      size_t sb = 0;
      if (_auto_build > 0.0f && newSizePower > osp && fill >= _build_min_keys)
         sb = W64 ? hn_rehash_build_scratch_bytes_u64(osp, newSizePower)
                  : hn_rehash_build_scratch_bytes_u32(osp, newSizePower);
      void* scratch = nullptr;
      if (sb) {
         SPLIT_CHECK_ERR(hipMallocAsync(&scratch, sb, s));
         chk(W64 ? hn_rehash_build_u64(_buckets, osp, fresh, newSizePower, dinfo(), scratch, sb, st(s))
                 : hn_rehash_build_u32(_buckets, osp, fresh, newSizePower, dinfo(), scratch, sb, st(s)));
      } else {
         SPLIT_CHECK_ERR(hipMallocAsync(&scratch, hn_rehash_scratch_bytes(), s));
         chk(W64 ? hn_rehash_u64(_buckets, osp, fresh, newSizePower, dinfo(), scratch, st(s))
                 : hn_rehash_u32(_buckets, osp, fresh, newSizePower, dinfo(), scratch, st(s)));
      }
      SPLIT_CHECK_ERR(hipFreeAsync(scratch, s));
      SPLIT_CHECK_ERR(hipStreamSynchronize(s));
      SPLIT_CHECK_ERR(hipFree(_buckets));
      _buckets = fresh;
      _nbuckets = size_t(1) << newSizePower;
      changed(s);
      if (_auto_reorder > 0.0f && !(sb && ordered())) reorder(s);
   }

   template <bool prefetches = true>
   void clear(targets t = targets::host, hipStream_t s = 0, size_t len = 0) {
      if (t == targets::device) {
         if (len == 0) len = _nbuckets;
         DeviceHasher::reset_all(_buckets, _dinfo, len, s);
      } else {
         chk(W64 ? hn_table_init_u64(_buckets, _nbuckets, st(s)) : hn_table_init_u32(_buckets, _nbuckets, st(s)));
         chk(hn_info_init(dinfo(), _info.sizePower, st(s)));
      }
      changed(s);
   }

   void resize_to_lf(float targetLF = 0.5, targets t = targets::host, hipStream_t s = 0) {
      while (load_factor() > targetLF) resize(_info.sizePower + 1, t, s);
   }
   void resize(int newSizePower, targets t = targets::host, hipStream_t s = 0) {
      if (t == targets::device)
         device_rehash(newSizePower, s);
      else
         rehash(newSizePower);
   }

   /* ------------------------------------------------------------ metadata (hashinator.h:434-620) */
   status peek_status(void) noexcept {
      const status r = _info.err;
      _info.err = status::invalid;
      pushInfo();
      return r;
   }
   int getSizePower(void) const noexcept { return _info.sizePower; }
   size_t size() const { return _info.fill; }
   size_t bucket_count() const { return _nbuckets; }
   constexpr KEY_TYPE get_emptybucket() const { return EMPTYBUCKET; }
   constexpr KEY_TYPE get_tombstone() const { return TOMBSTONE; }
   float load_factor() const { return (float)size() / bucket_count(); }
   size_t tombstone_count() const { return _info.tombstoneCounter; }
   float tombstone_ratio() const {
      if (tombstone_count() == 0) return 0.0;
      return (float)_info.tombstoneCounter / (float)_nbuckets;
   }
   template <bool warn = true>
   MapInfo* expose_mapinfo() noexcept {
      if (warn) printf("Warning, exposing Hashmap internal info struct!\n");
      return &_info;
   }
   template <bool warn = true>
   pair_t* expose_bucketdata() noexcept {
      if (warn) printf("Warning, exposing Hashmap internal bucket data!\n");
      _view_ok = false; // the caller may write through it, so the table is no longer known to be ordered
      if (_info.reserved1 & HN_INFO_ORDERED) {
         _info.reserved1 &= ~HN_INFO_ORDERED;
         pushInfo();
      }
      return _buckets;
   }
   MapInfo* device_info() noexcept { return _dinfo; }
   void stats() const {
      printf("Hashinator Stats \n");
      printf("Bucket size= %lu\n", (unsigned long)_nbuckets);
      printf("Fill= %lu, LoadFactor=%f \n", (unsigned long)_info.fill, load_factor());
      printf("Tombstones= %lu\n", (unsigned long)_info.tombstoneCounter);
      printf("Overflow= %lu\n", (unsigned long)_info.currentMaxBucketOverflow);
   }
   void swap(Hashmap& other) noexcept {
      std::swap(_buckets, other._buckets);
      std::swap(_nbuckets, other._nbuckets);
      std::swap(_dinfo, other._dinfo);
      std::swap(_device_map, other._device_map);
      std::swap(_info, other._info);
      _view_ok = other._view_ok = false;
   }

   /* ------------------------------------------------------------ device-side API (hashinator.h:853-1126)
    *
    * upload() / get_device_pointer() (hashinator.h:1416-1438) give user
    * kernels this map object in device memory; its warp* members run there,
    * one call per wave: all 64 lanes call together with the same key, w_tid =
    * the lane's index in the wave (0..63), and the lanes probe 64 consecutive
    * slots per step from the key's home.  Upload again after anything that
    * replaces the bucket array (an insert that grows, rehash, resize,
    * clear(host)); call download() after the kernels, before host-side use.
    * Loops are bounded by the table size.  Differences from the reference's
    * device code: a lookup takes the first match before the first EMPTY slot
    * in probe order (the reference lets a match later in the same 64-slot
    * window win; identical on tables these operations produce); the overflow
    * is the cyclic probe distance (the reference takes the whole table size
    * for a probe that wraps past the end); a new key clears the Robin Hood
    * order flag, which the library's host-side lookups use. */
   template <bool prefetches = true>
   Hashmap* upload(hipStream_t s = 0) {
      if (!_device_map) SPLIT_CHECK_ERR(hipMalloc(&_device_map, sizeof(Hashmap)));
      SPLIT_CHECK_ERR(hipMemcpyAsync(_device_map, this, sizeof(Hashmap), hipMemcpyHostToDevice, s));
      SPLIT_CHECK_ERR(hipStreamSynchronize(s)); // the source is this host object
      return _device_map;
   }
   Hashmap* get_device_pointer() { return _device_map; }

   /* warpInsert (hashinator.h:853-936): insert or overwrite candidateKey.
    * A key already present within the window wins over EMPTY slots; EMPTY
    * slots are claimed by CAS in lane order; tombstones are not reused. */
   template <bool skipOverWrites = false>
   HASHINATOR_DEVICEONLY void warpInsert(const KEY_TYPE& candidateKey, const VAL_TYPE& candidateVal,
                                         const size_t w_tid) noexcept {
      (void)warp_insert<skipOverWrites>(candidateKey, candidateVal, w_tid);
   }
   /* warpInsert_V (938-1025): the same, true when this call added the key. */
   template <bool skipOverWrites = false>
   HASHINATOR_DEVICEONLY bool warpInsert_V(const KEY_TYPE& candidateKey, const VAL_TYPE& candidateVal,
                                           const size_t w_tid) noexcept {
      return warp_insert<skipOverWrites>(candidateKey, candidateVal, w_tid);
   }
   /* warpFind (1027-1075): every lane's candidateVal becomes the key's value
    * when found, and is left as it was otherwise. */
   HASHINATOR_DEVICEONLY void warpFind(const KEY_TYPE& candidateKey, VAL_TYPE& candidateVal,
                                       const size_t w_tid) const noexcept {
      const int sp = _dinfo->sizePower;
      const uint64_t N = uint64_t(1) << sp, mask = N - 1;
      const uint64_t home = static_cast<uint64_t>(HashFunction::_hash(candidateKey, sp)) & mask;
      for (uint64_t i = 0; i < N; i += defaults::WARPSIZE) {
         const pair_t e = _buckets[(home + i + w_tid) & mask];
         const uint64_t hit = first_before_empty(__ballot(e.first == candidateKey), __ballot(e.first == EMPTYBUCKET));
         if (hit) {
            candidateVal = __shfl(e.second, __ffsll(static_cast<unsigned long long>(hit)) - 1, defaults::WARPSIZE);
            return;
         }
         if (__ballot(e.first == EMPTYBUCKET)) return;
      }
   }
   /* warpErase (1077-1126): the key's slot becomes TOMBSTONE (CAS, so a key
    * erased twice at once counts once); fill -1, tombstones +1. */
   HASHINATOR_DEVICEONLY void warpErase(const KEY_TYPE& candidateKey, const size_t w_tid) noexcept {
      const int sp = _dinfo->sizePower;
      const uint64_t N = uint64_t(1) << sp, mask = N - 1;
      const uint64_t home = static_cast<uint64_t>(HashFunction::_hash(candidateKey, sp)) & mask;
      for (uint64_t i = 0; i < N; i += defaults::WARPSIZE) {
         const uint64_t p = (home + i + w_tid) & mask;
         const KEY_TYPE k = _buckets[p].first;
         const uint64_t empty = __ballot(k == EMPTYBUCKET);
         const uint64_t hit = first_before_empty(__ballot(k == candidateKey), empty);
         if (hit) {
            if (static_cast<int>(w_tid) == __ffsll(static_cast<unsigned long long>(hit)) - 1) {
               KEY_TYPE expected = candidateKey;
               __hip_atomic_compare_exchange_strong(&_buckets[p].first, &expected, TOMBSTONE, __ATOMIC_RELAXED,
                                                    __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
               if (expected == candidateKey) {
                  __hip_atomic_fetch_sub(&_dinfo->fill, size_t(1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                  __hip_atomic_fetch_add(&_dinfo->tombstoneCounter, size_t(1), __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT);
               }
            }
            return;
         }
         if (empty) return;
      }
   }

   /* Single-thread device API (hashinator.h:1470-1769): iterators over the
    * live slots, device_find / device_count / device_erase, set_element /
    * device_insert (one thread per key, CAS from the home slot on; tombstones
    * are not reused), read_element.  A key absent from the map is undefined
    * for read_element, as in the reference (which asserts). */
   class device_iterator {
      size_t index;
      Hashmap* map;

   public:
      HASHINATOR_DEVICEONLY device_iterator(Hashmap& m, size_t i) : index(i), map(&m) {}
      HASHINATOR_DEVICEONLY size_t getIndex() const { return index; }
      HASHINATOR_DEVICEONLY device_iterator& operator++() {
         index = map->next_live(index + 1);
         return *this;
      }
      HASHINATOR_DEVICEONLY device_iterator operator++(int) {
         device_iterator t = *this;
         ++(*this);
         return t;
      }
      HASHINATOR_DEVICEONLY bool operator==(const device_iterator& o) const { return map == o.map && index == o.index; }
      HASHINATOR_DEVICEONLY bool operator!=(const device_iterator& o) const { return !(*this == o); }
      HASHINATOR_DEVICEONLY pair_t& operator*() const { return map->_buckets[index]; }
      HASHINATOR_DEVICEONLY pair_t* operator->() const { return &map->_buckets[index]; }
   };
   class const_device_iterator {
      size_t index;
      const Hashmap* map;

   public:
      HASHINATOR_DEVICEONLY explicit const_device_iterator(const Hashmap& m, size_t i) : index(i), map(&m) {}
      HASHINATOR_DEVICEONLY size_t getIndex() const { return index; }
      HASHINATOR_DEVICEONLY const_device_iterator& operator++() {
         index = map->next_live(index + 1);
         return *this;
      }
      HASHINATOR_DEVICEONLY const_device_iterator operator++(int) {
         const_device_iterator t = *this;
         ++(*this);
         return t;
      }
      HASHINATOR_DEVICEONLY bool operator==(const const_device_iterator& o) const {
         return map == o.map && index == o.index;
      }
      HASHINATOR_DEVICEONLY bool operator!=(const const_device_iterator& o) const { return !(*this == o); }
      HASHINATOR_DEVICEONLY const pair_t& operator*() const { return map->_buckets[index]; }
      HASHINATOR_DEVICEONLY const pair_t* operator->() const { return &map->_buckets[index]; }
   };

   HASHINATOR_DEVICEONLY size_t device_size() const { return _nbuckets; }
   HASHINATOR_DEVICEONLY device_iterator device_end() { return device_iterator(*this, _nbuckets); }
   HASHINATOR_DEVICEONLY const_device_iterator device_end() const { return const_device_iterator(*this, _nbuckets); }
   HASHINATOR_DEVICEONLY device_iterator device_begin() { return device_iterator(*this, next_live(0)); }
   HASHINATOR_DEVICEONLY const_device_iterator device_begin() const {
      return const_device_iterator(*this, next_live(0));
   }
   HASHINATOR_DEVICEONLY device_iterator device_find(KEY_TYPE key) { return device_iterator(*this, find_slot(key)); }
   HASHINATOR_DEVICEONLY const_device_iterator device_find(KEY_TYPE key) const {
      return const_device_iterator(*this, find_slot(key));
   }
   HASHINATOR_DEVICEONLY size_t device_count(const KEY_TYPE& key) const { return find_slot(key) != _nbuckets; }
   /* The slot becomes TOMBSTONE (CAS: erased once); fill -1, tombstones +1. */
   HASHINATOR_DEVICEONLY device_iterator device_erase(device_iterator pos) {
      const size_t i = pos.getIndex();
      if (i < _nbuckets) {
         KEY_TYPE item = _buckets[i].first;
         if (item != EMPTYBUCKET && item != TOMBSTONE) {
            KEY_TYPE expected = item;
            __hip_atomic_compare_exchange_strong(&_buckets[i].first, &expected, TOMBSTONE, __ATOMIC_RELAXED,
                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (expected == item) {
               __hip_atomic_fetch_sub(&_dinfo->fill, size_t(1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
               __hip_atomic_fetch_add(&_dinfo->tombstoneCounter, size_t(1), __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_AGENT);
            }
         }
      }
      return ++pos;
   }
   HASHINATOR_DEVICEONLY size_t device_erase(const KEY_TYPE& key) {
      const size_t i = find_slot(key);
      if (i == _nbuckets) return 0;
      device_erase(device_iterator(*this, i));
      return 1;
   }
   /* Insert or (unless skipOverWrites) overwrite; true for a new key.  The
    * overflow takes the probe length (rounded up to 64). */
   template <bool skipOverWrites = false>
   HASHINATOR_DEVICEONLY bool set_element(const KEY_TYPE& key, VAL_TYPE val) {
      const uint64_t mask = _nbuckets - 1;
      const uint64_t home = static_cast<uint64_t>(HashFunction::_hash(key, _dinfo->sizePower)) & mask;
      for (uint64_t i = 0; i < _nbuckets; ++i) {
         const uint64_t p = (home + i) & mask;
         KEY_TYPE old = EMPTYBUCKET;
         __hip_atomic_compare_exchange_strong(&_buckets[p].first, &old, key, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT);
         if (old == EMPTYBUCKET || old == key) {
            const bool added = old == EMPTYBUCKET;
            if (added || !skipOverWrites)
               __hip_atomic_exchange(&_buckets[p].second, val, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (added) {
               __hip_atomic_fetch_add(&_dinfo->fill, size_t(1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
               __hip_atomic_fetch_and(&_dinfo->reserved1, ~int32_t(HN_INFO_ORDERED), __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_AGENT);
            }
            __hip_atomic_fetch_max(&_dinfo->currentMaxBucketOverflow,
                                   nextOverflow(static_cast<size_t>(i + 1), defaults::WARPSIZE), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            return added;
         }
      }
      return false; // no EMPTY slot anywhere: not placed
   }
   template <bool skipOverWrites = false>
   HASHINATOR_DEVICEONLY hash_pair<device_iterator, bool> device_insert(pair_t newEntry) {
      const bool added = set_element<skipOverWrites>(newEntry.first, newEntry.second);
      return hash_pair<device_iterator, bool>(device_find(newEntry.first), added);
   }
   HASHINATOR_DEVICEONLY const VAL_TYPE& read_element(const KEY_TYPE& key) const {
      const size_t i = find_slot(key);
      return _buckets[i < _nbuckets ? i : 0].second;
   }

private:
   /* First live slot at or after i (the bucket count when none). */
   HASHINATOR_DEVICEONLY size_t next_live(size_t i) const {
      for (; i < _nbuckets; ++i) {
         const KEY_TYPE k = _buckets[i].first;
         if (k != EMPTYBUCKET && k != TOMBSTONE) return i;
      }
      return _nbuckets;
   }
   /* Slot of `key` (linear probe from its home, past tombstones, ending at
    * EMPTY), or the bucket count. */
   HASHINATOR_DEVICEONLY size_t find_slot(const KEY_TYPE& key) const {
      const uint64_t mask = _nbuckets - 1;
      const uint64_t home = static_cast<uint64_t>(HashFunction::_hash(key, _dinfo->sizePower)) & mask;
      for (uint64_t i = 0; i < _nbuckets; ++i) {
         const uint64_t p = (home + i) & mask;
         const KEY_TYPE k = _buckets[p].first;
         if (k == key) return static_cast<size_t>(p);
         if (k == EMPTYBUCKET) break;
      }
      return _nbuckets;
   }
   /* Bits of `match` below the lowest bit of `empty` (all of them when no EMPTY). */
   HASHINATOR_DEVICEONLY static uint64_t first_before_empty(uint64_t match, uint64_t empty) noexcept {
      return match & (empty ? ((empty & (0 - empty)) - 1) : ~uint64_t(0));
   }
   template <bool skipOverWrites>
   HASHINATOR_DEVICEONLY bool warp_insert(KEY_TYPE key, VAL_TYPE val, size_t w_tid) noexcept {
      MapInfo* info = _dinfo;
      const int sp = info->sizePower;
      const uint64_t N = uint64_t(1) << sp, mask = N - 1;
      const uint64_t home = static_cast<uint64_t>(HashFunction::_hash(key, sp)) & mask;
      const int lane = static_cast<int>(w_tid);
      for (uint64_t i = 0; i < N; i += defaults::WARPSIZE) {
         const uint64_t p = (home + i + w_tid) & mask;
         const KEY_TYPE k = _buckets[p].first;
         const uint64_t exists = __ballot(k == key);
         uint64_t empty = __ballot(k == EMPTYBUCKET);
         if (exists) {
            if (!skipOverWrites && lane == __ffsll(static_cast<unsigned long long>(exists)) - 1)
               __hip_atomic_exchange(&_buckets[p].second, val, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return false;
         }
         while (empty) {
            const int winner = __ffsll(static_cast<unsigned long long>(empty)) - 1;
            int done = 0; // 1: placed here, 2: the same key was placed here meanwhile
            if (lane == winner) {
               KEY_TYPE expected = EMPTYBUCKET;
               __hip_atomic_compare_exchange_strong(&_buckets[p].first, &expected, key, __ATOMIC_RELAXED,
                                                    __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
               if (expected == EMPTYBUCKET) {
                  __hip_atomic_exchange(&_buckets[p].second, val, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                  __hip_atomic_fetch_add(&info->fill, size_t(1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                  const size_t ovf = static_cast<size_t>((p - home) & mask) + 1;
                  if (ovf > info->currentMaxBucketOverflow)
                     __hip_atomic_fetch_max(&info->currentMaxBucketOverflow, nextOverflow(ovf, defaults::WARPSIZE),
                                            __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                  __hip_atomic_fetch_and(&info->reserved1, ~int32_t(HN_INFO_ORDERED), __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT);
                  done = 1;
               } else if (expected == key) {
                  if (!skipOverWrites)
                     __hip_atomic_exchange(&_buckets[p].second, val, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                  done = 2;
               }
            }
            if (__ballot(done == 1)) return true;
            if (__ballot(done == 2)) return false;
            empty &= ~(uint64_t(1) << winner);
         }
      }
      return false; // no EMPTY slot anywhere: not placed
   }

public:
   /* download() (hashinator.h:1457-1467), called after device-side work on the
    * map: read the Info back; grow one size up when the overflow exceeds
    * BUCKET_OVERFLOW, else remove tombstones if there are any.  The reference
    * grows with its host rehash; the device rehash here gives the same
    * key/value set. */
   void download(hipStream_t s = 0) {
      changed(s); // device-side work may have written the buckets
      if (_info.currentMaxBucketOverflow > static_cast<size_t>(defaults::BUCKET_OVERFLOW))
         device_rehash(_info.sizePower + 1, s);
      else if (_info.tombstoneCounter > 0)
         clean_tombstones(s);
   }

   /* GPU performCleanupTasks (hashinator.h:642-650). */
   template <bool prefetches = true>
   void performCleanupTasks(hipStream_t s = 0) {
      if (tombstone_ratio() > 0.025) clean_tombstones<prefetches>(s);
      while (_info.currentMaxBucketOverflow > static_cast<size_t>(defaults::BUCKET_OVERFLOW))
         device_rehash<prefetches>(_info.sizePower + 1, s);
   }

   /* ------------------------------------------------------------ host read-only lookups */
   class const_iterator {
      const Hashmap* h;
      size_t index;

   public:
      explicit const_iterator(const Hashmap& hashtable, size_t i) : h(&hashtable), index(i) {}
      const_iterator& operator++() {
         const std::vector<pair_t>& b = h->view();
         index++;
         while (index < b.size() && (b[index].first == EMPTYBUCKET || b[index].first == TOMBSTONE)) index++;
         return *this;
      }
      const_iterator operator++(int) {
         const_iterator t = *this;
         ++(*this);
         return t;
      }
      bool operator==(const_iterator o) const { return h == o.h && index == o.index; }
      bool operator!=(const_iterator o) const { return !(*this == o); }
      const pair_t& operator*() const { return h->view()[index]; }
      const pair_t* operator->() const { return &h->view()[index]; }
      size_t getIndex() { return index; }
   };

   /* const find (hashinator.h:742-768) over the host copy of the buckets. */
   const const_iterator find(KEY_TYPE key) const {
      const std::vector<pair_t>& b = view();
      const size_t bitMask = b.size() - 1;
      const uint32_t h = hash(key);
      for (size_t i = 0; i < b.size(); i++) {
         const size_t idx = (static_cast<size_t>(h) + i) & bitMask;
         const KEY_TYPE c = b[idx].first;
         if (c == TOMBSTONE) continue;
         if (c == key) return const_iterator(*this, idx);
         if (c == EMPTYBUCKET) return end();
      }
      return end();
   }
   size_t count(const KEY_TYPE& key) const { return find(key) != end() ? 1 : 0; }
   const_iterator begin() const {
      const std::vector<pair_t>& b = view();
      for (size_t i = 0; i < b.size(); i++)
         if (b[i].first != EMPTYBUCKET && b[i].first != TOMBSTONE) return const_iterator(*this, i);
      return end();
   }
   const_iterator end() const { return const_iterator(*this, _nbuckets); }

   /* ------------------------------------------------------------ residency (hashinator.h:1127-1444)
    * The table never leaves device memory; these keep the reference's names. */
   void optimizeGPU(hipStream_t stream = 0) noexcept { (void)stream; }
   void optimizeCPU(hipStream_t stream = 0) noexcept {
      SPLIT_CHECK_ERR(hipStreamSynchronize(stream));
      _view_ok = false;
      (void)view();
   }
   void memAdvise(hipMemoryAdvise, int, hipStream_t = 0) {}
   void streamAttach(hipStream_t, uint32_t = hipMemAttachSingle) {}

   /* ------------------------------------------------------------ extraction (hashinator.h:1154-1259) */
   template <bool prefetches = true, typename Rule, class A>
   size_t extractPattern(split::SplitVector<pair_t, A>& elements, Rule rule, hipStream_t s = 0) {
      elements.resize(_info.fill + 1, true);
      split::tools::splitStackArena mPool(split::tools::estimateMemoryForCompaction(_nbuckets), s);
      const size_t len = split::tools::copy_if_raw(_buckets, elements.data(), _nbuckets, rule, mPool, s);
      elements.erase(elements.data() + len, elements.end());
      return elements.size();
   }
   template <typename Rule, int BLOCKSIZE = 1024>
   size_t extractPattern(pair_t* elements, Rule rule, hipStream_t s = 0) {
      split::tools::splitStackArena mPool(split::tools::estimateMemoryForCompaction(_nbuckets), s);
      return split::tools::copy_if_raw(_buckets, elements, _nbuckets, rule, mPool, s);
   }
   template <bool prefetches = true, typename Rule, class A>
   size_t extractKeysByPattern(split::SplitVector<KEY_TYPE, A>& elements, Rule rule, hipStream_t s = 0) {
      elements.resize(_info.fill + 1, true);
      split::tools::splitStackArena mPool(split::tools::estimateMemoryForCompaction(_nbuckets), s);
      const size_t len = split::tools::copy_keys_if_raw(_buckets, elements.data(), _nbuckets, rule, mPool, s);
      elements.erase(elements.data() + len, elements.end());
      return elements.size();
   }
   template <bool prefetches = true, typename Rule, class A>
   size_t extractKeysByPattern(split::SplitVector<KEY_TYPE, A>& elements, Rule rule, void* stack, size_t max_size,
                               hipStream_t s = 0) {
      elements.resize(_info.fill + 1, true);
      split::tools::splitStackArena mPool(stack, max_size);
      const size_t len = split::tools::copy_keys_if_raw(_buckets, elements.data(), _nbuckets, rule, mPool, s);
      elements.erase(elements.data() + len, elements.end());
      return elements.size();
   }
   /* All live keys, through the library's built-in predicate (no rule staging). */
   template <bool prefetches = true, class A>
   size_t extractAllKeys(split::SplitVector<KEY_TYPE, A>& elements, hipStream_t s = 0) {
      elements.resize(_info.fill + 1, true);
      const size_t sb = hn_compact_scratch_bytes(_nbuckets);
      split::tools::splitStackArena mPool(sb + 512, s);
      uint64_t* d_count = static_cast<uint64_t*>(mPool.allocate(8));
      void* scratch = mPool.allocate(sb);
      const hn_pred p{HN_PRED_VALID, 0, 0, 0};
      chk(W64 ? hn_copy_keys_if_pairs_u64(_buckets, _nbuckets, &p, reinterpret_cast<uint64_t*>(elements.data()),
                                          d_count, scratch, sb, st(s))
              : hn_copy_keys_if_pairs_u32(_buckets, _nbuckets, &p, reinterpret_cast<uint32_t*>(elements.data()),
                                          d_count, scratch, sb, st(s)));
      const size_t len = split::tools::detail::read_count(d_count, s);
      elements.erase(elements.data() + len, elements.end());
      return elements.size();
   }
   template <bool prefetches = true, class A>
   size_t extractAllKeys(split::SplitVector<KEY_TYPE, A>& elements, void* stack, size_t max_size, hipStream_t s = 0) {
      auto rule = [] __host__ __device__(const pair_t& kval) -> bool {
         return kval.first != EMPTYBUCKET && kval.first != TOMBSTONE;
      };
      return extractKeysByPattern<prefetches>(elements, rule, stack, max_size, s);
   }
   /* The *Loop variants (hashinator.h:1186-1259): the reference runs them as
    * one block looping over the buckets (split::tools::copy_if_loop) and
    * resizes `elements` on the device.  Same elements and order here, through
    * the single-pass compaction; the size is set from the host after the
    * count is read, so the call ends with its stream synchronised. */
   template <typename Rule, class A>
   void extractPatternLoop(split::SplitVector<pair_t, A>& elements, Rule rule, hipStream_t s = 0) {
      extractPattern(elements, rule, s);
   }
   template <class A>
   void extractLoop(split::SplitVector<pair_t, A>& elements, hipStream_t s = 0) {
      auto rule = [] __host__ __device__(const pair_t& kval) -> bool {
         return kval.first != EMPTYBUCKET && kval.first != TOMBSTONE;
      };
      extractPatternLoop(elements, rule, s);
   }
   template <typename Rule, class A>
   void extractKeysByPatternLoop(split::SplitVector<KEY_TYPE, A>& elements, Rule rule, hipStream_t s = 0) {
      extractKeysByPattern(elements, rule, s);
   }
   template <class A>
   void extractAllKeysLoop(split::SplitVector<KEY_TYPE, A>& elements, hipStream_t s = 0) {
      extractAllKeys(elements, s);
   }

   /* ------------------------------------------------------------ Robin Hood order (an addition) */

   /* Rewrite every cluster in home-slot order (ties by key); afterwards
    * retrieve and erase stop a probe early, so absent keys cost about as much
    * as present ones.  Same live set and occupancy.  No-op while tombstones
    * are present.  Returns whether the table is now marked ordered. */
   bool reorder(hipStream_t s = 0) {
      const size_t sb = hn_reorder_scratch_bytes(_info.sizePower);
      void* scratch = nullptr;
      SPLIT_CHECK_ERR(hipMallocAsync(&scratch, sb, s));
      chk(W64 ? hn_reorder_u64(_buckets, dinfo(), _info.sizePower, scratch, sb, st(s))
              : hn_reorder_u32(_buckets, dinfo(), _info.sizePower, scratch, sb, st(s)));
      SPLIT_CHECK_ERR(hipFreeAsync(scratch, s));
      changed(s);
      return ordered();
   }
   bool ordered() const noexcept { return (_info.reserved1 & HN_INFO_ORDERED) != 0; }
   /* Fraction of the bucket count an insert must reach to be followed by
    * reorder() (default 1/16); 0 turns the automatic pass off. */
   void setAutoReorder(float fraction) noexcept { _auto_reorder = fraction; }

   /* Fraction of the bucket count an insert into an empty table must reach
    * to take the bulk path (hn_build_*: partition + per-region sort, no
    * probing; writes the reordered image directly), and the fewest keys it
    * must hold (below about 2^17 keys at low load factors the build's
    * launches cost more than the probing insert and reorder); fraction 0
    * turns it off.  A device_rehash that grows a map of at least min_keys
    * keys also takes the build. */
   void setAutoBuild(float fraction, size_t min_keys = size_t(1) << 17) noexcept {
      _auto_build = fraction;
      _build_min_keys = min_keys;
   }

   /* Fraction of the bucket count an insert into a table that already holds
    * keys (and no tombstones) must reach to take the merge (hn_merge_*: the
    * insert + reorder image written into a fresh bucket array, about a
    * table's read and write; it pays from about a quarter of the table);
    * a table not in Robin Hood order is reordered first when the batch is
    * at least 4/3 of the fraction.  Off by default: like a resize, the
    * merge replaces the bucket array, so a device copy from upload() must
    * be taken again after such an insert.  The batch must also hold the
    * build's min_keys. */
   void setAutoMerge(float fraction) noexcept { _auto_merge = fraction; }

private:
   /* A large insert into a table holding keys, by the merge into a fresh
    * array; false when it does not apply.  The device writes the probing
    * insert's result instead where the merge cannot take the table, so a
    * true return means the insert is done. */
   bool try_merge(const void* keys, const void* vals, const void* pairs, size_t len, hipStream_t s) {
      const double N = static_cast<double>(_nbuckets);
      const int sp = _info.sizePower;
      if (_auto_merge <= 0.0f || static_cast<double>(len) < _auto_merge * N || len < _build_min_keys ||
          len >= _nbuckets || sp > 30 || _info.fill == 0 || _info.tombstoneCounter != 0)
         return false;
      if (!ordered()) {
         if (static_cast<double>(len) < _auto_merge * N * 4.0 / 3.0) return false;
         reorder(s);
      }
      const size_t sb = W64 ? hn_merge_scratch_bytes_u64(sp, len) : hn_merge_scratch_bytes_u32(sp, len);
      if (sb == 0) return false;
      pair_t* fresh = nullptr;
      SPLIT_CHECK_ERR(hipMalloc(&fresh, _nbuckets * sizeof(pair_t)));
      void* scratch = nullptr;
      SPLIT_CHECK_ERR(hipMallocAsync(&scratch, sb, s));
      if (pairs)
         chk(W64 ? hn_merge_pairs_u64(pairs, _buckets, fresh, dinfo(), len, sp, scratch, sb, st(s))
                 : hn_merge_pairs_u32(pairs, _buckets, fresh, dinfo(), len, sp, scratch, sb, st(s)));
      else
         chk(W64 ? hn_merge_u64(keys, vals, _buckets, fresh, dinfo(), len, sp, scratch, sb, st(s))
                 : hn_merge_u32(keys, vals, _buckets, fresh, dinfo(), len, sp, scratch, sb, st(s)));
      SPLIT_CHECK_ERR(hipFreeAsync(scratch, s));
      SPLIT_CHECK_ERR(hipStreamSynchronize(s));
      SPLIT_CHECK_ERR(hipFree(_buckets));
      _buckets = fresh;
      changed(s);
      if (_auto_reorder > 0.0f && !ordered() && static_cast<double>(len) >= _auto_reorder * N) reorder(s);
      return true;
   }

   /* The bulk path; false when it does not apply or the device declined it
    * (then the table is empty and the probing insert runs). */
   bool try_build(const void* keys, const void* vals, const void* pairs, size_t len, hipStream_t s) {
      if (_auto_build <= 0.0f || static_cast<double>(len) < _auto_build * static_cast<double>(_nbuckets) ||
          len < _build_min_keys || len >= _nbuckets || _info.fill != 0 || _info.tombstoneCounter != 0)
         return false;
      const int sp = _info.sizePower;
      const size_t sb = W64 ? hn_build_scratch_bytes_u64(sp, len) : hn_build_scratch_bytes_u32(sp, len);
      if (sb == 0) return false;
      void* scratch = nullptr;
      SPLIT_CHECK_ERR(hipMallocAsync(&scratch, sb, s));
      if (pairs)
         chk(W64 ? hn_build_pairs_u64(pairs, _buckets, dinfo(), len, sp, scratch, sb, st(s), tune())
                 : hn_build_pairs_u32(pairs, _buckets, dinfo(), len, sp, scratch, sb, st(s), tune()));
      else
         chk(W64 ? hn_build_u64(keys, vals, _buckets, dinfo(), len, sp, scratch, sb, st(s), tune())
                 : hn_build_u32(keys, vals, _buckets, dinfo(), len, sp, scratch, sb, st(s), tune()));
      SPLIT_CHECK_ERR(hipFreeAsync(scratch, s));
      changed(s);
      return _info.err == status::success;
   }
   void after_insert(size_t len, hipStream_t s) {
      if (_auto_reorder > 0.0f && static_cast<double>(len) >= _auto_reorder * static_cast<double>(_nbuckets))
         reorder(s);
      else
         changed(s);
   }

public:
   /* ------------------------------------------------------------ accelerated API (hashinator.h:1261-1409) */
   template <bool prefetches = true>
   void clean_tombstones(hipStream_t s = 0) {
      if (_info.tombstoneCounter == 0) return;
      const int sp = _info.sizePower;
      // The segment pass (a few MB of lists plus a pool for the rare segments
      // too long for LDS); a table with every slot occupied needs room for a
      // region clean over the whole table, and so does a retry after the
      // pool ran short (err = FAIL with TOMBSTONEs left).
      const bool full = _info.fill + _info.tombstoneCounter >= (size_t(1) << sp);
      for (int attempt = 0; attempt < 2; ++attempt) {
         const size_t budget = (full || attempt) ? SIZE_MAX : 0;
         const size_t sb = W64 ? hn_clean_tombstones_scratch_bytes_u64(sp, budget)
                               : hn_clean_tombstones_scratch_bytes_u32(sp, budget);
         void* scratch = nullptr;
         SPLIT_CHECK_ERR(hipMallocAsync(&scratch, sb, s));
         chk(W64 ? hn_clean_tombstones_u64(_buckets, dinfo(), sp, scratch, sb, st(s))
                 : hn_clean_tombstones_u32(_buckets, dinfo(), sp, scratch, sb, st(s)));
         SPLIT_CHECK_ERR(hipFreeAsync(scratch, s));
         changed(s);
         if (full || _info.tombstoneCounter == 0 || _info.err != status::fail) break;
      }
      if (_auto_reorder > 0.0f && !ordered()) reorder(s);
   }

   template <bool prefetches = true>
   void insert(KEY_TYPE* keys, VAL_TYPE* vals, size_t len, float targetLF = 0.5, hipStream_t s = 0) {
      if (len == 0) {
         set_status(status::success);
         return;
      }
      const int64_t neededPowerSize = std::ceil(std::log2((_info.fill + len) * (1.0 / targetLF)));
      if (neededPowerSize > _info.sizePower) resize(static_cast<int>(neededPowerSize), targets::device, s);
      if (try_build(keys, vals, nullptr, len, s) || try_merge(keys, vals, nullptr, len, s)) return;
      DeviceHasher::insert(keys, vals, _buckets, _dinfo, len, s, tune());
      after_insert(len, s);
   }
   template <bool prefetches = true>
   void insertIndex(KEY_TYPE* keys, size_t len, float targetLF = 0.5, hipStream_t s = 0) {
      if (len == 0) {
         set_status(status::success);
         return;
      }
      const int64_t neededPowerSize = std::ceil(std::log2((_info.fill + len) * (1.0 / targetLF)));
      if (neededPowerSize > _info.sizePower) resize(static_cast<int>(neededPowerSize), targets::device, s);
      if (try_build(keys, nullptr, nullptr, len, s) || try_merge(keys, nullptr, nullptr, len, s)) return;
      DeviceHasher::insertIndex(keys, _buckets, _dinfo, len, s, tune());
      after_insert(len, s);
   }
   template <bool prefetches = true>
   void insert(pair_t* src, size_t len, float targetLF = 0.5, hipStream_t s = 0) {
      if (len == 0) {
         set_status(status::success);
         return;
      }
      const int64_t neededPowerSize = std::ceil(std::log2(((_info.fill) + len) * (1.0 / targetLF)));
      if (neededPowerSize > _info.sizePower) resize(static_cast<int>(neededPowerSize), targets::device, s);
      if (try_build(nullptr, nullptr, src, len, s) || try_merge(nullptr, nullptr, src, len, s)) return;
      DeviceHasher::insert(src, _buckets, _dinfo, len, s, tune());
      after_insert(len, s);
   }
   /* `hit` (optional, an addition) receives 1/0 per key. */
   template <bool prefetches = true>
   void retrieve(KEY_TYPE* keys, VAL_TYPE* vals, size_t len, hipStream_t s = 0, uint8_t* hit = nullptr) {
      DeviceHasher::retrieve(keys, vals, _buckets, _dinfo, len, s, tune(), hit);
   }
   template <bool prefetches = true>
   void retrieve(pair_t* src, size_t len, hipStream_t s = 0, uint8_t* hit = nullptr) {
      DeviceHasher::retrieve(src, _buckets, _dinfo, len, s, tune(), hit);
   }
   template <bool prefetches = true>
   void erase(KEY_TYPE* keys, size_t len, hipStream_t s = 0) {
      DeviceHasher::erase(keys, _buckets, _dinfo, len, s, tune());
      changed(s);
   }
};

} // namespace Hashinator
