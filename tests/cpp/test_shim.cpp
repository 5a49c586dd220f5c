// test_shim.cpp -- the drop-in C++ headers (include/hashinator, include/splitvector)
// exercised the way the reference's own unit tests use them
// (kstppd/hashinator unit_tests/hashmap_unit_test/main.cu,
// unit_tests/stream_compaction/unit.cu), on the GPU, against host-side
// expectations.  Prints one line per case; exits nonzero on any failure.
#include <hashinator/hashinator.h>
#include <splitvector/split_tools.h>

#include <algorithm>
#include <cstdio>
#include <map>
#include <random>
#include <unordered_map>
#include <vector>

using namespace Hashinator;

static int failures = 0;
#define CHECK(cond)                                                                                                    \
   do {                                                                                                                \
      if (!(cond)) {                                                                                                   \
         printf("  check failed: %s (line %d)\n", #cond, __LINE__);                                                    \
         ++failures;                                                                                                   \
         return false;                                                                                                 \
      }                                                                                                                \
   } while (0)

template <typename K>
struct IsOdd {
   __host__ __device__ bool operator()(hash_pair<K, K>& e) const { return (e.second & 1) != 0; }
};
template <typename K>
struct KeyBelow {
   K limit;
   __host__ __device__ bool operator()(hash_pair<K, K>& e) const {
      return e.first < limit && e.first != std::numeric_limits<K>::max() && e.first != std::numeric_limits<K>::max() - 1;
   }
};
struct Below1000 {
   __host__ __device__ bool operator()(uint32_t& x) const { return x < 1000; }
};

template <typename K>
bool insert_retrieve_find(int power, float lf) {
   using Map = Hashmap<K, K>;
   const size_t N = size_t(1) << power;
   const size_t n = static_cast<size_t>(lf * N);
   split::SplitVector<K> keys(n), vals(n), out(2 * n), q(2 * n);
   std::mt19937_64 rng(power);
   for (size_t i = 0; i < n; ++i) {
      keys[i] = static_cast<K>(i * 2654435761ull + 17) & (std::numeric_limits<K>::max() >> 2);
      vals[i] = static_cast<K>(rng());
   }
   for (size_t i = 0; i < n; ++i) {
      q[i] = keys[i];
      q[n + i] = keys[i] | (K(1) << (sizeof(K) * 8 - 2)); // never inserted
   }
   Map hmap(power);
   keys.optimizeGPU();
   vals.optimizeGPU();
   hmap.insert(keys.data(), vals.data(), n, 1.0f);
   CHECK(hmap.peek_status() == status::success);
   CHECK(hmap.size() == n);
   CHECK(hmap.ordered()); // a large insert is followed by the Robin Hood reorder pass
   uint8_t* hit = nullptr;
   SPLIT_CHECK_ERR(hipMallocManaged(&hit, 2 * n));
   for (size_t i = 0; i < 2 * n; ++i) out[i] = K(7);
   q.optimizeGPU();
   out.optimizeGPU();
   hmap.retrieve(q.data(), out.data(), 2 * n, 0, hit);
   SPLIT_CHECK_ERR(hipDeviceSynchronize());
   for (size_t i = 0; i < n; ++i) {
      CHECK(hit[i] == 1 && out[i] == vals[i]);
      CHECK(hit[n + i] == 0 && out[n + i] == K(7)); // a miss leaves the value untouched
   }
   SPLIT_CHECK_ERR(hipFree(hit));
   // host-side const find on the same buckets (recover_all_elements)
   hmap.optimizeCPU();
   for (size_t i = 0; i < n; i += 7) {
      auto it = hmap.find(keys[i]);
      CHECK(it != hmap.end() && it->second == vals[i]);
      CHECK(hmap.find(q[n + i]) == hmap.end());
   }
   size_t iterated = 0;
   for (auto it = hmap.begin(); it != hmap.end(); ++it) ++iterated;
   CHECK(iterated == n);
   // a small insert drops the order flag; reorder() restores it; exposing the buckets drops it
   split::SplitVector<K> one(1, q[n]), oneval(1, K(5));
   one.optimizeGPU();
   oneval.optimizeGPU();
   hmap.insert(one.data(), oneval.data(), 1, 1.0f);
   CHECK(!hmap.ordered() && hmap.size() == n + 1);
   CHECK(hmap.reorder());
   (void)hmap.template expose_bucketdata<false>();
   CHECK(!hmap.ordered());
   return true;
}

template <typename K>
bool erase_clean_extract(int power) {
   using Map = Hashmap<K, K>;
   const size_t n = (size_t(1) << power) / 2;
   split::SplitVector<K> keys(n), vals(n);
   for (size_t i = 0; i < n; ++i) {
      keys[i] = static_cast<K>(i);
      vals[i] = static_cast<K>(i);
   }
   Map hmap(power);
   hmap.setAutoBuild(0.0f); // the probing insert kernel, whatever the batch size
   hmap.insert(keys.data(), vals.data(), n, 1.0f);
   // erase the even keys (test_hashmap_* "no even-valued survivors")
   split::SplitVector<K> even(n / 2);
   for (size_t i = 0; i < n / 2; ++i) even[i] = static_cast<K>(2 * i);
   hmap.erase(even.data(), n / 2);
   CHECK(hmap.size() == n - n / 2);
   CHECK(hmap.tombstone_count() == n / 2);
   hmap.clean_tombstones();
   CHECK(hmap.tombstone_count() == 0);
   CHECK(hmap.size() == n - n / 2);
   hmap.optimizeCPU();
   for (size_t i = 0; i < n; ++i) {
      const bool present = hmap.find(static_cast<K>(i)) != hmap.end();
      CHECK(present == (i % 2 == 1));
   }
   // extractPattern with a user rule == host filter of the bucket array, in bucket order
   split::SplitVector<hash_pair<K, K>> elements;
   const size_t m = hmap.extractPattern(elements, IsOdd<K>());
   CHECK(m == n - n / 2);
   std::vector<hash_pair<K, K>> expect, b(hmap.bucket_count());
   SPLIT_CHECK_ERR(hipMemcpy(b.data(), hmap.template expose_bucketdata<false>(), b.size() * sizeof(b[0]),
                             hipMemcpyDeviceToHost)); // the table lives in device memory
   for (size_t i = 0; i < hmap.bucket_count(); ++i)
      if (b[i].first != hmap.get_emptybucket() && b[i].first != hmap.get_tombstone() && (b[i].second & 1))
         expect.push_back(b[i]);
   CHECK(expect.size() == m);
   for (size_t i = 0; i < m; ++i) CHECK(elements[i] == expect[i]);
   // keys by pattern and all keys (extractKeysByPattern / extractAllKeys)
   split::SplitVector<K> ks;
   const size_t kb = hmap.extractKeysByPattern(ks, KeyBelow<K>{K(1000)});
   size_t eb = 0;
   for (auto& e : expect) eb += e.first < 1000;
   CHECK(kb == eb);
   for (size_t i = 0; i < kb; ++i) CHECK(ks[i] < 1000 && (ks[i] & 1));
   split::SplitVector<K> all;
   CHECK(hmap.extractAllKeys(all) == m);
   for (size_t i = 0; i < m; ++i) CHECK(all[i] == expect[i].first);
   // the *Loop variants (hashinator.h:1186-1259) give the same elements in the same order
   split::SplitVector<hash_pair<K, K>> el2;
   hmap.extractPatternLoop(el2, IsOdd<K>());
   CHECK(el2.size() == m);
   for (size_t i = 0; i < m; ++i) CHECK(el2[i] == expect[i]);
   split::SplitVector<hash_pair<K, K>> el3;
   hmap.extractLoop(el3);
   CHECK(el3.size() == hmap.size());
   split::SplitVector<K> ks2, all2;
   hmap.extractKeysByPatternLoop(ks2, KeyBelow<K>{K(1000)});
   CHECK(ks2.size() == kb);
   for (size_t i = 0; i < kb; ++i) CHECK(ks2[i] == ks[i]);
   hmap.extractAllKeysLoop(all2);
   CHECK(all2.size() == m);
   for (size_t i = 0; i < m; ++i) CHECK(all2[i] == all[i]);
   // split::tools::copy_if_loop / copy_if_keys_loop on SplitVectors (split_tools.h:994-1020)
   split::SplitVector<hash_pair<K, K>> in(hmap.bucket_count());
   SPLIT_CHECK_ERR(hipMemcpy(in.data(), b.data(), b.size() * sizeof(b[0]), hipMemcpyHostToDevice));
   split::SplitVector<hash_pair<K, K>> out;
   split::tools::copy_if_loop(in, out, IsOdd<K>());
   CHECK(out.size() == m);
   for (size_t i = 0; i < m; ++i) CHECK(out[i] == expect[i]);
   split::SplitVector<K> kout;
   split::tools::copy_if_keys_loop(in, kout, IsOdd<K>());
   CHECK(kout.size() == m);
   for (size_t i = 0; i < m; ++i) CHECK(kout[i] == expect[i].first);
   return true;
}

template <typename K>
bool resize_and_grow() {
   using Map = Hashmap<K, K>;
   Map hmap; // sizePower 5
   const size_t n = 5000;
   split::SplitVector<K> keys(n), vals(n);
   for (size_t i = 0; i < n; ++i) {
      keys[i] = static_cast<K>(3 * i + 1);
      vals[i] = static_cast<K>(i);
   }
   hmap.insert(keys.data(), vals.data(), n); // targetLF 0.5 grows to 2^14
   CHECK(hmap.getSizePower() == 14);
   CHECK(hmap.size() == n);
   hmap.resize(16, targets::device);
   CHECK(hmap.getSizePower() == 16 && hmap.peek_status() == status::success && hmap.size() == n);
   hmap.resize(17, targets::host);
   CHECK(hmap.getSizePower() >= 17 && hmap.size() == n);
   {
      // download() (hashinator.h:1457-1467) after device-side work: tombstones go, the set stays
      Map hd(14);
      hd.insert(keys.data(), vals.data(), n, 1.0f);
      hd.erase(keys.data(), n / 4);
      CHECK(hd.tombstone_count() == n / 4);
      hd.download();
      CHECK(hd.tombstone_count() == 0 && hd.size() == n - n / 4 && hd.getSizePower() == 14);
      hd.optimizeCPU();
      for (size_t i = 0; i < n; ++i) CHECK((hd.find(keys[i]) != hd.end()) == (i >= n / 4));
   }
   {
      // growing a map of many keys runs the rehash as one bulk build (the key threshold lowered here)
      Map hb(14);
      hb.setAutoBuild(1.0f / 16, 0);
      hb.insert(keys.data(), vals.data(), n, 1.0f);
      hb.resize(16, targets::device);
      CHECK(hb.getSizePower() == 16 && hb.peek_status() == status::success && hb.size() == n && hb.ordered());
      hb.optimizeCPU();
      for (size_t i = 0; i < n; ++i) {
         auto it = hb.find(keys[i]);
         CHECK(it != hb.end() && it->second == vals[i]);
      }
   }
   hmap.optimizeCPU();
   for (size_t i = 0; i < n; ++i) {
      auto it = hmap.find(keys[i]);
      CHECK(it != hmap.end() && it->second == vals[i]);
   }
   hmap.clear(targets::device);
   CHECK(hmap.size() == 0);
   hmap.optimizeCPU();
   CHECK(hmap.find(keys[0]) == hmap.end());
   // insertIndex: value = position
   Map h2(13);
   h2.insertIndex(keys.data(), n, 1.0f);
   h2.optimizeCPU();
   for (size_t i = 0; i < n; i += 13) CHECK(h2.find(keys[i])->second == static_cast<K>(i));
   // duplicate-heavy insert (hashmap_unit_test main.cu:1064-1085)
   Map h3(10);
   split::SplitVector<K> dk(1 << 12), dv(1 << 12);
   for (size_t i = 0; i < dk.size(); ++i) {
      dk[i] = static_cast<K>(i % 1024);
      dv[i] = static_cast<K>(i % 1024);
   }
   h3.insert(dk.data(), dv.data(), dk.size(), 1.0f);
   CHECK(h3.size() == 1024);
   return true;
}

/* User kernel of the device-side API: one wave per key, w_tid = lane. */
template <typename Map, typename K>
__global__ void k_warp_ops(Map* m, const K* keys, const K* vals, K* out, size_t n, int op) {
   const size_t wid = (static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x) / 64;
   const size_t w_tid = threadIdx.x % 64;
   if (wid >= n) return; // whole waves
   if (op == 0) {
      m->warpInsert(keys[wid], vals[wid], w_tid);
   } else if (op == 1) {
      K v = out[wid];
      m->warpFind(keys[wid], v, w_tid);
      if (w_tid == 0) out[wid] = v;
   } else if (op == 2) {
      m->warpErase(keys[wid], w_tid);
   } else {
      const bool added = m->warpInsert_V(keys[wid], vals[wid], w_tid);
      if (w_tid == 0) out[wid] = added ? K(1) : K(0);
   }
}
template <typename Map, typename K>
void warp_ops(Map* dm, split::SplitVector<K>& keys, split::SplitVector<K>& vals, split::SplitVector<K>& out, size_t n,
              int op) {
   const unsigned blocks = static_cast<unsigned>((n * 64 + 255) / 256);
   hipLaunchKernelGGL((k_warp_ops<Map, K>), dim3(blocks), dim3(256), 0, 0, dm, keys.data(), vals.data(), out.data(),
                      n, op);
   SPLIT_CHECK_ERR(hipGetLastError());
   SPLIT_CHECK_ERR(hipDeviceSynchronize());
}

/* upload() + warpInsert / warpInsert_V / warpFind / warpErase from a user
 * kernel (hashinator.h:853-1126, 1416-1467), then the host API on the same map. */
template <typename K>
bool device_api() {
   using Map = Hashmap<K, K>;
   const size_t n = 3000, m = 1000;
   split::SplitVector<K> keys(n + m), vals(n + m), out(n + m);
   for (size_t i = 0; i < n + m; ++i) {
      keys[i] = static_cast<K>(7 * i + 3);
      vals[i] = static_cast<K>(i ^ 0x55);
   }
   Map hmap(14);
   Map* dm = hmap.upload();
   CHECK(dm != nullptr && hmap.get_device_pointer() == dm);
   warp_ops(dm, keys, vals, out, n, 0); // insert the first n keys
   hmap.download();
   CHECK(hmap.size() == n && hmap.tombstone_count() == 0);
   {
      // linear probing's occupied-slot set does not depend on the insertion
      // order: it must equal a serial insert's on the host
      std::vector<hash_pair<K, K>> b(hmap.bucket_count());
      SPLIT_CHECK_ERR(hipMemcpy(b.data(), hmap.template expose_bucketdata<false>(), b.size() * sizeof(b[0]),
                                hipMemcpyDeviceToHost));
      const size_t N = b.size();
      std::vector<char> occ(N, 0);
      for (size_t i = 0; i < n; ++i) {
         size_t p = static_cast<size_t>(Hashinator::HashFunctions::Fibonacci<K>::_hash(keys[i], 14)) & (N - 1);
         while (occ[p]) p = (p + 1) & (N - 1);
         occ[p] = 1;
      }
      for (size_t j = 0; j < N; ++j) CHECK(occ[j] == (b[j].first != hmap.get_emptybucket()));
   }
   dm = hmap.upload(); // download() may have replaced the bucket array (a rehash)
   // host-side retrieve sees them
   split::SplitVector<K> got(n + m);
   for (size_t i = 0; i < n + m; ++i) got[i] = K(9);
   hmap.retrieve(keys.data(), got.data(), n + m);
   for (size_t i = 0; i < n + m; ++i) CHECK(got[i] == (i < n ? vals[i] : K(9)));
   // warpFind: values of present keys, absent keys untouched
   for (size_t i = 0; i < n + m; ++i) out[i] = K(5);
   warp_ops(dm, keys, vals, out, n + m, 1);
   for (size_t i = 0; i < n + m; ++i) CHECK(out[i] == (i < n ? vals[i] : K(5)));
   // warpInsert_V: false for present keys, true for new ones
   warp_ops(dm, keys, vals, out, n + m, 3);
   for (size_t i = 0; i < n + m; ++i) CHECK(out[i] == (i < n ? K(0) : K(1)));
   hmap.download();
   CHECK(hmap.size() == n + m);
   dm = hmap.upload();
   // warpErase of the first half; download() removes the tombstones
   warp_ops(dm, keys, vals, out, (n + m) / 2, 2);
   SPLIT_CHECK_ERR(hipDeviceSynchronize());
   hmap.download();
   CHECK(hmap.size() == (n + m) - (n + m) / 2 && hmap.tombstone_count() == 0);
   hmap.optimizeCPU();
   for (size_t i = 0; i < n + m; ++i) CHECK((hmap.find(keys[i]) != hmap.end()) == (i >= (n + m) / 2));
   // a Robin Hood ordered map (bulk build) takes new keys from the device and
   // the host lookups still find every key
   Map hb(13);
   hb.setAutoBuild(1.0f / 16, 0);
   hb.insert(keys.data(), vals.data(), n, 1.0f);
   CHECK(hb.ordered());
   Map* db = hb.upload();
   split::SplitVector<K> nk(m), nv(m), no(m);
   for (size_t i = 0; i < m; ++i) {
      nk[i] = keys[n + i];
      nv[i] = vals[n + i];
   }
   warp_ops(db, nk, nv, no, m, 0);
   hb.download();
   CHECK(hb.size() == n + m && !hb.ordered());
   for (size_t i = 0; i < n + m; ++i) got[i] = K(9);
   hb.retrieve(keys.data(), got.data(), n + m);
   for (size_t i = 0; i < n + m; ++i) CHECK(got[i] == vals[i]);
   return true;
}

/* User kernels of the single-thread device API: one thread per key. */
template <typename Map, typename K>
__global__ void k_thread_ops(Map* m, const K* keys, const K* vals, K* out, size_t n, int op) {
   const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
   if (i >= n) return;
   if (op == 0) {
      out[i] = m->set_element(keys[i], vals[i]) ? K(1) : K(0);
   } else if (op == 1) {
      auto it = m->device_find(keys[i]);
      out[i] = it != m->device_end() ? it->second : K(5);
   } else if (op == 2) {
      out[i] = static_cast<K>(m->device_erase(keys[i]));
   } else if (op == 3) {
      out[i] = static_cast<K>(m->device_count(keys[i]));
   } else if (op == 4) {
      auto r = m->device_insert(hash_pair<K, K>(keys[i], vals[i]));
      out[i] = (r.second ? K(1) : K(0)) + (r.first != m->device_end() && r.first->second == vals[i] ? K(2) : K(0));
   } else {
      out[i] = m->read_element(keys[i]);
   }
}
template <typename Map, typename K>
__global__ void k_iterate(Map* m, K* out) {
   if (threadIdx.x != 0 || blockIdx.x != 0) return;
   size_t c = 0;
   K sum = 0;
   for (auto it = m->device_begin(); it != m->device_end(); ++it) {
      ++c;
      sum += it->first;
   }
   out[0] = static_cast<K>(c);
   out[1] = sum;
}
template <typename Map, typename K>
void thread_ops(Map* dm, split::SplitVector<K>& keys, split::SplitVector<K>& vals, split::SplitVector<K>& out,
                size_t n, int op) {
   hipLaunchKernelGGL((k_thread_ops<Map, K>), dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, 0, dm,
                      keys.data(), vals.data(), out.data(), n, op);
   SPLIT_CHECK_ERR(hipGetLastError());
   SPLIT_CHECK_ERR(hipDeviceSynchronize());
}

/* set_element / device_insert / device_find / device_count / device_erase /
 * read_element and device iterators (hashinator.h:1470-1769). */
template <typename K>
bool thread_api() {
   using Map = Hashmap<K, K>;
   const size_t n = 3000, m = 1000;
   split::SplitVector<K> keys(n + m), vals(n + m), out(n + m), it(2);
   K ksum = 0;
   for (size_t i = 0; i < n + m; ++i) {
      keys[i] = static_cast<K>(11 * i + 5);
      vals[i] = static_cast<K>(i * 3 + 1);
      if (i < n) ksum += keys[i];
   }
   Map hmap(14);
   Map* dm = hmap.upload();
   thread_ops(dm, keys, vals, out, n, 0);
   for (size_t i = 0; i < n; ++i) CHECK(out[i] == K(1));
   thread_ops(dm, keys, vals, out, n, 0); // again: no new keys
   for (size_t i = 0; i < n; ++i) CHECK(out[i] == K(0));
   hmap.download();
   CHECK(hmap.size() == n);
   dm = hmap.upload();
   thread_ops(dm, keys, vals, out, n + m, 1);
   for (size_t i = 0; i < n + m; ++i) CHECK(out[i] == (i < n ? vals[i] : K(5)));
   thread_ops(dm, keys, vals, out, n + m, 3);
   for (size_t i = 0; i < n + m; ++i) CHECK(out[i] == (i < n ? K(1) : K(0)));
   thread_ops(dm, keys, vals, out, n, 5);
   for (size_t i = 0; i < n; ++i) CHECK(out[i] == vals[i]);
   hipLaunchKernelGGL((k_iterate<Map, K>), dim3(1), dim3(64), 0, 0, dm, it.data());
   SPLIT_CHECK_ERR(hipDeviceSynchronize());
   CHECK(it[0] == static_cast<K>(n) && it[1] == ksum);
   // host retrieve agrees with the device-side inserts
   split::SplitVector<K> got(n + m);
   for (size_t i = 0; i < n + m; ++i) got[i] = K(9);
   hmap.retrieve(keys.data(), got.data(), n + m);
   for (size_t i = 0; i < n + m; ++i) CHECK(got[i] == (i < n ? vals[i] : K(9)));
   // device_erase of the first half (1 each), then device_insert of the rest
   thread_ops(dm, keys, vals, out, n / 2, 2);
   for (size_t i = 0; i < n / 2; ++i) CHECK(out[i] == K(1));
   split::SplitVector<K> nk(m), nv(m), no(m);
   for (size_t i = 0; i < m; ++i) {
      nk[i] = keys[n + i];
      nv[i] = vals[n + i];
   }
   thread_ops(dm, nk, nv, no, m, 4);
   for (size_t i = 0; i < m; ++i) CHECK(no[i] == K(3));
   hmap.download();
   CHECK(hmap.size() == n - n / 2 + m && hmap.tombstone_count() == 0);
   hmap.optimizeCPU();
   for (size_t i = 0; i < n + m; ++i) CHECK((hmap.find(keys[i]) != hmap.end()) == (i >= n / 2));
   return true;
}

bool compaction_and_scan() {
   const size_t n = (1 << 20) + 333;
   split::SplitVector<uint32_t> in(n), out(n);
   std::mt19937 rng(3);
   for (size_t i = 0; i < n; ++i) in[i] = rng() % 4000;
   split::tools::copy_if(in, out, Below1000());
   std::vector<uint32_t> expect;
   for (size_t i = 0; i < n; ++i)
      if (in[i] < 1000) expect.push_back(in[i]);
   CHECK(out.size() == expect.size());
   for (size_t i = 0; i < expect.size(); ++i) CHECK(out[i] == expect[i]);
   // preallocated arena (unit_tests/stream_compaction/preallocated.cu)
   split::SplitVector<uint32_t> out2(n);
   const size_t need = split::tools::estimateMemoryForCompaction(in);
   void* stack = nullptr;
   SPLIT_CHECK_ERR(hipMalloc(&stack, need));
   split::tools::copy_if(in, out2, Below1000(), stack, need);
   SPLIT_CHECK_ERR(hipFree(stack));
   CHECK(out2 == out);
   // exclusive scan, any size, wrapping
   split::SplitVector<uint32_t> flags(n), sum(n);
   for (size_t i = 0; i < n; ++i) flags[i] = rng();
   split::tools::split_prefix_scan(flags, sum);
   uint32_t acc = 0;
   for (size_t i = 0; i < n; ++i) {
      CHECK(sum[i] == acc);
      acc += flags[i];
   }
   return true;
}

/* The raw-pointer copy_if over a caller's stack, driven through the
 * split_gpu* runtime names the reference's tests use
 * (unit_tests/stream_compaction/race.cu, preallocated.cu): ragged sizes
 * around the 8,192-element tile (0 included), a 12-byte element, and a rule
 * that writes its element, as the reference's clean_tombstones rule does
 * (hashinator.h:1285-1297): the edit stays in the input and the output holds
 * the edited element. */
struct Triple {
   uint32_t a, b, c;
};
struct TripleEven {
   __host__ __device__ bool operator()(Triple& t) const { return (t.b % 2) == 0; }
};
struct BumpOdd {
   __host__ __device__ bool operator()(hash_pair<uint64_t, uint64_t>& e) const {
      const bool odd = (e.second & 1) != 0;
      e.second += 10;
      return odd;
   }
};
bool raw_compaction_split_gpu_api() {
   split_gpuStream_t s;
   SPLIT_CHECK_ERR(split_gpuStreamCreate(&s));
   std::mt19937 rng(17);
   for (size_t n : {size_t(0), size_t(1), size_t(8191), size_t(8192), size_t(8193), size_t(3 * 8192 + 5), size_t(200001)}) {
      std::vector<Triple> h(n);
      for (size_t i = 0; i < n; ++i) h[i] = Triple{static_cast<uint32_t>(i), static_cast<uint32_t>(rng()), 7u};
      Triple *d_in = nullptr, *d_out = nullptr;
      void* stack = nullptr;
      const size_t need = split::tools::estimateMemoryForCompaction(n);
      SPLIT_CHECK_ERR(split_gpuMalloc((void**)&d_in, (n ? n : 1) * sizeof(Triple)));
      SPLIT_CHECK_ERR(split_gpuMalloc((void**)&d_out, (n ? n : 1) * sizeof(Triple)));
      SPLIT_CHECK_ERR(split_gpuMalloc(&stack, need));
      SPLIT_CHECK_ERR(split_gpuMemsetAsync(d_out, 0xff, (n ? n : 1) * sizeof(Triple), s));
      SPLIT_CHECK_ERR(split_gpuMemcpyAsync(d_in, h.data(), n * sizeof(Triple), split_gpuMemcpyHostToDevice, s));
      const size_t m = split::tools::copy_if(d_in, d_out, n, TripleEven(), stack, need, s);
      std::vector<Triple> got(m);
      SPLIT_CHECK_ERR(split_gpuMemcpyAsync(got.data(), d_out, m * sizeof(Triple), split_gpuMemcpyDeviceToHost, s));
      SPLIT_CHECK_ERR(split_gpuStreamSynchronize(s));
      size_t j = 0;
      for (size_t i = 0; i < n; ++i) {
         if (h[i].b % 2) continue;
         CHECK(j < m && got[j].a == h[i].a && got[j].b == h[i].b && got[j].c == 7u);
         ++j;
      }
      CHECK(j == m);
      SPLIT_CHECK_ERR(split_gpuFree(d_in));
      SPLIT_CHECK_ERR(split_gpuFree(d_out));
      SPLIT_CHECK_ERR(split_gpuFree(stack));
   }
   // a rule that writes its element
   const size_t n = 50000;
   using P = hash_pair<uint64_t, uint64_t>;
   std::vector<P> h(n);
   for (size_t i = 0; i < n; ++i) h[i] = P(i, (static_cast<uint64_t>(rng()) << 20) | (i % 3 == 0 ? 1u : 0u));
   split::SplitVector<P, split::split_device_allocator<P>> in(h), out(n);
   split::tools::copy_if(in, out, BumpOdd(), s);
   const auto got = out.to_host();
   const auto after = in.to_host();
   size_t j = 0;
   for (size_t i = 0; i < n; ++i) {
      CHECK(after[i].first == h[i].first && after[i].second == h[i].second + 10);
      if (!(h[i].second & 1)) continue;
      CHECK(j < got.size() && got[j].first == h[i].first && got[j].second == h[i].second + 10);
      ++j;
   }
   CHECK(j == got.size());
   SPLIT_CHECK_ERR(split_gpuStreamDestroy(s));
   SPLIT_CHECK_ERR(split_gpuDeviceSynchronize());
   return true;
}

bool splitvector_host_semantics() {
   split::SplitVector<int> v{1, 2, 3};
   v.push_back(4);
   CHECK(v.size() == 4 && v.back() == 4 && v.front() == 1);
   v.insert(v.begin() + 1, 2, 9);
   CHECK(v.size() == 6 && v[1] == 9 && v[2] == 9 && v[3] == 2);
   v.erase(v.begin(), v.begin() + 2);
   CHECK(v.size() == 4 && v[0] == 9 && v[1] == 2);
   v.resize(100);
   CHECK(v.size() == 100 && v.capacity() >= 100);
   v.resize(3);
   CHECK(v.size() == 3);
   split::SplitVector<int> w(v);
   CHECK(w == v);
   w[0] = -1;
   CHECK(w != v);
   v.clear();
   CHECK(v.empty());
   split::SplitVector<int> big(1 << 20, 5);
   big.optimizeCPU();
   CHECK(big[(1 << 20) - 1] == 5 && big[0] == 5);
   return true;
}

// ---------------------------------------------------------------- SplitVector on the device
// (splitvec.h:482-500 upload, 617-669 / 940-1040 HOSTDEVICE members,
// 1069-1097 device_push_back, 901-912 device_resize, 1361-1462 device_insert)

template <class Vec>
__global__ void push_kernel(Vec* v, uint32_t n, uint8_t* ok) {
   const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
   if (i < n) ok[i] = v->device_push_back(static_cast<typename Vec::value_type>(3ull * i + 1)) ? 1 : 0;
}

// Many threads push at once; against a serial replay of the accepted pushes:
// the same multiset after the initial elements, the size the reference's
// rule gives (a push succeeds while size + 1 < capacity), and the host's
// size() agreeing with the device's.
template <class Alloc>
bool splitvector_device_push_back() {
   using Vec = split::SplitVector<uint64_t, Alloc>;
   struct Case {
      size_t init, cap, n;
   } cases[] = {{0, 100000, 250000}, {0, 1 << 20, 300000}, {1000, 5000, 10000}, {0, 1, 64}, {0, 0, 64}};
   for (const Case& c : cases) {
      std::vector<uint64_t> first(c.init);
      for (size_t i = 0; i < c.init; ++i) first[i] = 7 * i + 5;
      Vec v(first);
      if (c.cap) v.reserve(c.cap, true);
      CHECK(v.capacity() == std::max(c.cap, c.init) && v.size() == c.init);
      Vec* dv = v.upload();
      uint8_t* ok = nullptr;
      SPLIT_CHECK_ERR(hipMalloc(&ok, c.n));
      hipLaunchKernelGGL(push_kernel<Vec>, dim3((c.n + 255) / 256), dim3(256), 0, 0, dv, static_cast<uint32_t>(c.n), ok);
      SPLIT_CHECK_ERR(hipGetLastError());
      SPLIT_CHECK_ERR(hipDeviceSynchronize());
      std::vector<uint8_t> hok(c.n);
      SPLIT_CHECK_ERR(hipMemcpy(hok.data(), ok, c.n, hipMemcpyDeviceToHost));
      SPLIT_CHECK_ERR(hipFree(ok));
      // serial replay of the accepted pushes
      std::vector<uint64_t> replay(first);
      size_t accepted = 0;
      for (size_t i = 0; i < c.n; ++i)
         if (hok[i]) {
            replay.push_back(3ull * i + 1);
            ++accepted;
         }
      const size_t expect = c.cap == 0 ? c.init : std::min(c.init + c.n, c.cap - 1);
      CHECK(accepted + c.init == expect);
      CHECK(v.size() == expect);
      std::vector<uint64_t> got = v.to_host();
      CHECK(got.size() == expect);
      CHECK(std::equal(first.begin(), first.end(), got.begin()));
      std::sort(got.begin() + c.init, got.end());
      std::sort(replay.begin() + c.init, replay.end());
      CHECK(got == replay);
   }
   return true;
}

template <class Vec>
__global__ void script_kernel(Vec* v, long long* st) {
   using T = typename Vec::value_type;
   auto it = v->device_insert(v->begin() + 3, T(100));
   st[0] = (it == v->begin() + 3) && (*it == T(100));
   it = v->device_insert(v->begin(), 2, T(7));
   st[1] = (it == v->begin()) && v->size() == 13;
   const T src[3] = {T(41), T(42), T(43)};
   it = v->device_insert(v->begin() + 5, src, src + 3);
   st[2] = (it == v->begin() + 5) && (*it == T(41));
   st[3] = v->device_resize(v->size() + 2) ? 1 : 0;
   v->erase(v->begin() + 1);
   v->erase(v->begin() + 2, v->begin() + 4);
   v->pop_back();
   const size_t before = v->size();
   st[4] = (!v->device_resize(v->capacity() + 1) && v->size() == before) ? 1 : 0;
   st[5] = (v->device_insert(v->begin(), v->capacity(), T(5)) == v->end() && v->size() == before) ? 1 : 0;
   st[6] = static_cast<long long>(v->size());
   // HOSTDEVICE accessors
   long long sum = 0;
   for (auto p = v->begin(); p != v->end(); ++p) sum += static_cast<long long>(*p);
   long long sum2 = 0;
   for (size_t i = 0; i < v->size(); ++i) sum2 += static_cast<long long>((*v)[i]) + static_cast<long long>(v->at(i));
   st[7] = sum;
   st[8] = sum2;
   st[9] = static_cast<long long>(v->front()) * 1000 + static_cast<long long>(v->back());
   st[10] = (v->data() == &(*v)[0] && !v->empty() && v->capacity() == 64) ? 1 : 0;
}

// One device thread runs a script of inserts, resizes and erases; a
// std::vector runs the same script on the host.
template <class Alloc>
bool splitvector_device_script() {
   using Vec = split::SplitVector<uint32_t, Alloc>;
   std::vector<uint32_t> init(10);
   for (uint32_t i = 0; i < 10; ++i) init[i] = i;
   Vec v(init);
   v.reserve(64, true);
   Vec* dv = v.upload();
   long long* st = nullptr;
   SPLIT_CHECK_ERR(hipMalloc(&st, 16 * sizeof(long long)));
   hipLaunchKernelGGL(script_kernel<Vec>, dim3(1), dim3(1), 0, 0, dv, st);
   SPLIT_CHECK_ERR(hipGetLastError());
   SPLIT_CHECK_ERR(hipDeviceSynchronize());
   long long h[16];
   SPLIT_CHECK_ERR(hipMemcpy(h, st, sizeof(h), hipMemcpyDeviceToHost));
   SPLIT_CHECK_ERR(hipFree(st));
   std::vector<uint32_t> e(init);
   e.insert(e.begin() + 3, 100u);
   e.insert(e.begin(), 2, 7u);
   const uint32_t src[3] = {41, 42, 43};
   e.insert(e.begin() + 5, src, src + 3);
   e.resize(e.size() + 2);
   e.erase(e.begin() + 1);
   e.erase(e.begin() + 2, e.begin() + 4);
   e.pop_back();
   for (int i = 0; i < 6; ++i) CHECK(h[i] == 1);
   CHECK(h[6] == static_cast<long long>(e.size()));
   CHECK(v.size() == e.size());
   CHECK(v.to_host() == e);
   long long sum = 0;
   for (uint32_t x : e) sum += x;
   CHECK(h[7] == sum && h[8] == 2 * sum);
   CHECK(h[9] == static_cast<long long>(e.front()) * 1000 + e.back());
   CHECK(h[10] == 1);
   // a host change after upload reaches the device header
   v.push_back(99);
   CHECK(v.size() == e.size() + 1);
   split::SplitInfo info{0, 0};
   v.copyMetadata(&info);
   SPLIT_CHECK_ERR(hipDeviceSynchronize());
   CHECK(info.size == e.size() + 1 && info.capacity == 64);
   return true;
}

// The drop-in compaction on device-resident SplitVectors (the stream
// compaction workload's path through split::tools::copy_if).
bool splitvector_device_copy_if() {
   using DVec = split::SplitVector<hash_pair<uint64_t, uint64_t>, split::split_device_allocator<hash_pair<uint64_t, uint64_t>>>;
   const size_t n = (1 << 21) + 77;
   std::vector<hash_pair<uint64_t, uint64_t>> h(n);
   std::mt19937_64 rng(11);
   for (size_t i = 0; i < n; ++i) h[i] = hash_pair<uint64_t, uint64_t>(i, rng());
   DVec in(h), out(n);
   split::tools::copy_if(in, out, IsOdd<uint64_t>());
   std::vector<hash_pair<uint64_t, uint64_t>> expect;
   for (auto& e : h)
      if (e.second & 1) expect.push_back(e);
   CHECK(out.size() == expect.size());
   auto got = out.to_host();
   CHECK(got.size() == expect.size());
   for (size_t i = 0; i < expect.size(); ++i) CHECK(got[i].first == expect[i].first && got[i].second == expect[i].second);
   return true;
}

/* setAutoMerge: a quarter of the table inserted into a half-full ordered
 * map (with keys already present and keys repeated in the batch) takes the
 * merge; every key then retrieves its last value, the fill is the union's
 * and the map stays ordered. */
template <typename K>
bool merge_insert() {
   using Map = Hashmap<K, K>;
   const int power = 16;
   const size_t N = size_t(1) << power, n_old = N / 2, n_new = N / 4;
   split::SplitVector<K> ok(n_old), ov(n_old), nk(n_new), nv(n_new);
   std::map<K, K> expect;
   for (size_t i = 0; i < n_old; ++i) {
      ok[i] = static_cast<K>(i * 2654435761ull + 17) & (std::numeric_limits<K>::max() >> 2);
      ov[i] = static_cast<K>(i);
      expect[ok[i]] = ov[i];
   }
   std::mt19937_64 rng(7);
   for (size_t i = 0; i < n_new; ++i) {
      K k;
      if (i % 10 == 0)
         k = ok[rng() % n_old]; // already present: overwritten
      else if (i % 17 == 0 && i > 0)
         k = nk[rng() % i]; // repeated in the batch: the last value stays
      else
         k = static_cast<K>((i * 40503ull + 5) | (K(1) << (sizeof(K) * 8 - 3)));
      nk[i] = k;
      nv[i] = static_cast<K>(1000000 + i);
      expect[k] = nv[i];
   }
   Map hmap(power);
   hmap.setAutoBuild(1.0f / 16, 1);
   hmap.setAutoMerge(0.25f);
   ok.optimizeGPU();
   ov.optimizeGPU();
   nk.optimizeGPU();
   nv.optimizeGPU();
   hmap.insert(ok.data(), ov.data(), n_old, 1.0f);
   CHECK(hmap.ordered() && hmap.size() == n_old);
   auto* before = hmap.template expose_bucketdata<false>();
   CHECK(hmap.reorder()); // expose_bucketdata dropped the flag; the reorder sets it again
   hmap.insert(nk.data(), nv.data(), n_new, 1.0f);
   CHECK(hmap.peek_status() == status::success);
   CHECK(hmap.template expose_bucketdata<false>() != before); // a fresh array: the merge ran
   CHECK(hmap.size() == expect.size());
   split::SplitVector<K> q(expect.size()), out(expect.size());
   size_t i = 0;
   for (auto& kv : expect) q[i++] = kv.first;
   q.optimizeGPU();
   out.optimizeGPU();
   uint8_t* hit = nullptr;
   SPLIT_CHECK_ERR(hipMallocManaged(&hit, expect.size()));
   hmap.retrieve(q.data(), out.data(), expect.size(), 0, hit);
   SPLIT_CHECK_ERR(hipDeviceSynchronize());
   i = 0;
   for (auto& kv : expect) {
      CHECK(hit[i] == 1 && out[i] == kv.second);
      ++i;
   }
   SPLIT_CHECK_ERR(hipFree(hit));
   return true;
}

int main() {
   struct Case {
      const char* name;
      bool (*fn)();
   } cases[] = {
       {"insert_retrieve_find<u32> sp=16 lf=0.9", [] { return insert_retrieve_find<uint32_t>(16, 0.9f); }},
       {"insert_retrieve_find<u64> sp=18 lf=0.9", [] { return insert_retrieve_find<uint64_t>(18, 0.9f); }},
       {"erase_clean_extract<u32> sp=14", [] { return erase_clean_extract<uint32_t>(14); }},
       {"erase_clean_extract<u64> sp=14", [] { return erase_clean_extract<uint64_t>(14); }},
       {"resize_and_grow<u32>", [] { return resize_and_grow<uint32_t>(); }},
       {"resize_and_grow<u64>", [] { return resize_and_grow<uint64_t>(); }},
       {"device_api<u32>", [] { return device_api<uint32_t>(); }},
       {"device_api<u64>", [] { return device_api<uint64_t>(); }},
       {"thread_api<u32>", [] { return thread_api<uint32_t>(); }},
       {"thread_api<u64>", [] { return thread_api<uint64_t>(); }},
       {"compaction_and_scan", compaction_and_scan},
       {"raw_compaction_split_gpu_api", raw_compaction_split_gpu_api},
       {"splitvector_host_semantics", splitvector_host_semantics},
       {"splitvector_device_push_back<unified>", splitvector_device_push_back<split::split_unified_allocator<uint64_t>>},
       {"splitvector_device_push_back<device>", splitvector_device_push_back<split::split_device_allocator<uint64_t>>},
       {"splitvector_device_script<unified>", splitvector_device_script<split::split_unified_allocator<uint32_t>>},
       {"splitvector_device_script<device>", splitvector_device_script<split::split_device_allocator<uint32_t>>},
       {"splitvector_device_copy_if", splitvector_device_copy_if},
       {"merge_insert<u32>", merge_insert<uint32_t>},
       {"merge_insert<u64>", merge_insert<uint64_t>},
   };
   for (auto& c : cases) {
      printf("%s ...\n", c.name);
      fflush(stdout);
      const bool ok = c.fn();
      printf("%s %s\n", ok ? "PASS" : "FAIL", c.name);
      fflush(stdout);
   }
   printf("%d failure(s)\n", failures);
   return failures ? 1 : 0;
}
