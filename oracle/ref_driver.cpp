// ref_driver.cpp -- C entry points over the reference's own host-only map.
//
// TEST INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile against the
// reference's headers (kstppd/hashinator include/) with
// -DHASHINATOR_CPU_ONLY_MODE, into oracle/_ref/libhashref.so.  It exists to
// pin the C restatement in hn_oracle.c and to write the golden fixtures under
// tests/golden/; nothing shipped links it.
//
// Calls used (kstppd/hashinator):
//   HashFunctions::Fibonacci<K>::_hash        hashfunctions.h:79-82
//   Hashmap(int)                              hashinator.h:108-116
//   Hashmap::insert(K*, V*, len)  (CPU-only)  hashinator.h:1774-1778
//   Hashmap::find(K) const                    hashinator.h:742-768
//   Hashmap::erase(const K&)                  hashinator.h:842-850
//   Hashmap::resize(int) -> rehash            hashinator.h:236-279, 548
//   expose_mapinfo / expose_bucketdata        hashinator.h:469-482
#include <cstdint>
#include <cstdlib>
#include <stdlib.h>
#include <cstring>
#include <new>
#include <utility>
#include <stdexcept>
// The reference headers rely on these standard headers being included first.
#include <hashinator/hashinator.h>

using namespace Hashinator;

namespace {
template <typename K>
using Map = Hashmap<K, K>;

struct RefInfo {
   int32_t size_power;
   int32_t pad0;
   uint64_t fill, overflow, tombstones;
   int32_t err, pad1;
};

template <typename K>
void info_of(Map<K>* m, RefInfo* out) {
   MapInfo* in = m->template expose_mapinfo<false>();
   out->size_power = in->sizePower;
   out->pad0 = 0;
   out->fill = in->fill;
   out->overflow = in->currentMaxBucketOverflow;
   out->tombstones = in->tombstoneCounter;
   out->err = (int32_t)in->err;
   out->pad1 = 0;
}

template <typename K>
void find_n(Map<K>* m, const K* keys, K* vals, uint8_t* hit, size_t n) {
   const Map<K>& cm = *m;
   for (size_t i = 0; i < n; ++i) {
      auto it = cm.find(keys[i]);
      bool h = it != cm.end();
      if (h) vals[i] = it->second;
      hit[i] = h;
   }
}

template <typename K>
size_t erase_n(Map<K>* m, const K* keys, size_t n) {
   size_t c = 0;
   for (size_t i = 0; i < n; ++i) c += m->erase(keys[i]);
   return c;
}

template <typename K>
void dump(Map<K>* m, K* out) {
   hash_pair<K, K>* b = m->template expose_bucketdata<false>();
   size_t nb = m->bucket_count();
   for (size_t i = 0; i < nb; ++i) {
      out[2 * i] = b[i].first;
      out[2 * i + 1] = b[i].second;
   }
}
} // namespace

extern "C" {
uint32_t ref_fibhash32(uint32_t k, int sp) { return HashFunctions::Fibonacci<uint32_t>::_hash(k, sp); }
uint64_t ref_fibhash64(uint64_t k, int sp) { return HashFunctions::Fibonacci<uint64_t>::_hash(k, sp); }

void* ref_map32_new(int sp) { return new Map<uint32_t>(sp); }
void ref_map32_free(void* m) { delete static_cast<Map<uint32_t>*>(m); }
void ref_map32_info(void* m, RefInfo* out) { info_of(static_cast<Map<uint32_t>*>(m), out); }
size_t ref_map32_nbuckets(void* m) { return static_cast<Map<uint32_t>*>(m)->bucket_count(); }
void ref_map32_copy_buckets(void* m, uint32_t* out) { dump(static_cast<Map<uint32_t>*>(m), out); }
void ref_map32_insert(void* m, uint32_t* k, uint32_t* v, size_t n) { static_cast<Map<uint32_t>*>(m)->insert(k, v, n); }
void ref_map32_find(void* m, const uint32_t* k, uint32_t* v, uint8_t* h, size_t n) {
   find_n(static_cast<Map<uint32_t>*>(m), k, v, h, n);
}
size_t ref_map32_erase(void* m, const uint32_t* k, size_t n) { return erase_n(static_cast<Map<uint32_t>*>(m), k, n); }
void ref_map32_rehash(void* m, int sp) { static_cast<Map<uint32_t>*>(m)->resize(sp); }

void* ref_map64_new(int sp) { return new Map<uint64_t>(sp); }
void ref_map64_free(void* m) { delete static_cast<Map<uint64_t>*>(m); }
void ref_map64_info(void* m, RefInfo* out) { info_of(static_cast<Map<uint64_t>*>(m), out); }
size_t ref_map64_nbuckets(void* m) { return static_cast<Map<uint64_t>*>(m)->bucket_count(); }
void ref_map64_copy_buckets(void* m, uint64_t* out) { dump(static_cast<Map<uint64_t>*>(m), out); }
void ref_map64_insert(void* m, uint64_t* k, uint64_t* v, size_t n) { static_cast<Map<uint64_t>*>(m)->insert(k, v, n); }
void ref_map64_find(void* m, const uint64_t* k, uint64_t* v, uint8_t* h, size_t n) {
   find_n(static_cast<Map<uint64_t>*>(m), k, v, h, n);
}
size_t ref_map64_erase(void* m, const uint64_t* k, size_t n) { return erase_n(static_cast<Map<uint64_t>*>(m), k, n); }
void ref_map64_rehash(void* m, int sp) { static_cast<Map<uint64_t>*>(m)->resize(sp); }
}
