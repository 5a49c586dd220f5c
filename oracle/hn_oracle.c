/*
 * hn_oracle.c -- CPU restatement of Hashinator's map and compaction semantics.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path, and the CPU timing
 * baseline.  The shipped library (hashinator_amd/) never links or calls this.
 *
 * Reference citations are to kstppd/hashinator:
 *   include/hashinator/hashfunctions.h   Fibonacci hash
 *   include/hashinator/hashinator.h      host-only map (_at, find, erase, rehash,
 *                                        performCleanupTasks), accelerated wrappers
 *                                        (insert, erase, clean_tombstones, device_rehash)
 *   include/hashinator/kernels_AMD.h     per-key semantics of the accelerated kernels
 *   include/splitvector/split_tools.h    copy_if / copy_keys_if / split_prefix_scan
 *
 * Sentinels: EMPTYBUCKET = max(K), TOMBSTONE = EMPTYBUCKET - 1 (hashinator.h:59-60).
 */
#include "hn_oracle.h"

#include <stdlib.h>
#include <string.h>

#define OR_CPU_BUCKET_OVERFLOW 32u /* defaults.h:33, CPU-only build */
#define OR_GPU_BUCKET_OVERFLOW 64u /* defaults.h:28-31, wave64 HIP build */

/* hashfunctions.h:39-43.  The product is formed in 64 bits (the constant is
 * `ul`), shifted, then truncated to 32 bits; callers mask with 2^sp - 1. */
uint32_t or_fibhash32(uint32_t key, int size_power) {
   key ^= key >> (32 - size_power);
   uint64_t prod = (uint64_t)key * 2654435769ull;
   return (uint32_t)(prod >> (32 - size_power));
}

/* hashfunctions.h:52-56: product modulo 2^64, then shifted. */
uint64_t or_fibhash64(uint64_t key, int size_power) {
   key ^= key >> (64 - size_power);
   return (key * 11400714819323198485ull) >> (64 - size_power);
}

void or_fibhash32_n(const uint32_t* keys, const int32_t* sps, uint32_t* out, size_t n) {
   for (size_t i = 0; i < n; ++i) out[i] = or_fibhash32(keys[i], sps[i]);
}

void or_fibhash64_n(const uint64_t* keys, const int32_t* sps, uint64_t* out, size_t n) {
   for (size_t i = 0; i < n; ++i) out[i] = or_fibhash64(keys[i], sps[i]);
}

/* common.h:52-57 */
static uint64_t or_next_overflow(uint64_t cur, uint64_t w) {
   uint64_t r = cur % w;
   return r == 0 ? cur : cur + (w - r);
}

uint64_t or_splitmix64(uint64_t* state) {
   uint64_t z = (*state += 0x9E3779B97F4A7C15ull);
   z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
   z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
   return z ^ (z >> 31);
}

/* Hashmap::hash (hashinator.h:77-81) returns uint32_t for both key widths. */
static inline uint32_t or_hash_idx32(uint32_t k, int sp) { return or_fibhash32(k, sp); }
static inline uint32_t or_hash_idx64(uint64_t k, int sp) { return (uint32_t)or_fibhash64(k, sp); }

/* ------------------------------------------------------------------------ */
/* The key-width generic part is written once and instantiated twice.        */
/* ------------------------------------------------------------------------ */
#define OR_DEFINE(W, K)                                                                                         \
   struct or_map##W {                                                                                          \
      or_info info;                                                                                            \
      size_t n;                                                                                                \
      K* b; /* 2*n words: key, value */                                                                        \
   };                                                                                                          \
   static const K OR_EMPTY##W = (K)~(K)0;                                                                       \
   static const K OR_TOMB##W = (K)(~(K)0 - 1);                                                                 \
                                                                                                               \
   static K* or_alloc_table##W(size_t n) {                                                                     \
      K* b = (K*)malloc(2 * n * sizeof(K));                                                                    \
      for (size_t i = 0; i < n; ++i) {                                                                         \
         b[2 * i] = OR_EMPTY##W;                                                                               \
         b[2 * i + 1] = 0;                                                                                     \
      }                                                                                                        \
      return b;                                                                                                \
   }                                                                                                           \
                                                                                                               \
   /* Hashmap(int) + Info(int) (hashinator.h:108-116, defaults.h:46-48), CPU-only overflow 32 */               \
   or_map##W* or_map##W##_new(int sp) {                                                                        \
      or_map##W* m = (or_map##W*)calloc(1, sizeof(or_map##W));                                                 \
      m->info.size_power = sp;                                                                                 \
      m->info.fill = 0;                                                                                        \
      m->info.overflow = OR_CPU_BUCKET_OVERFLOW;                                                               \
      m->info.tombstones = 0;                                                                                  \
      m->info.err = OR_INVALID;                                                                                \
      m->n = (size_t)1 << sp;                                                                                  \
      m->b = or_alloc_table##W(m->n);                                                                          \
      return m;                                                                                                \
   }                                                                                                           \
   void or_map##W##_free(or_map##W* m) {                                                                       \
      if (!m) return;                                                                                          \
      free(m->b);                                                                                              \
      free(m);                                                                                                 \
   }                                                                                                           \
   const or_info* or_map##W##_info(const or_map##W* m) { return &m->info; }                                    \
   size_t or_map##W##_nbuckets(const or_map##W* m) { return m->n; }                                            \
   void or_map##W##_copy_buckets(const or_map##W* m, K* out) { memcpy(out, m->b, 2 * m->n * sizeof(K)); }      \
                                                                                                               \
   /* Host rehash (hashinator.h:236-279): 32-probe placement limit, grow and retry from the old table. */      \
   void or_map##W##_rehash(or_map##W* m, int nsp) {                                                            \
      for (;;) {                                                                                               \
         if (nsp > 32) abort(); /* reference throws std::out_of_range */                                       \
         size_t nn = (size_t)1 << nsp;                                                                         \
         K* nb = or_alloc_table##W(nn);                                                                        \
         m->info.size_power = nsp;                                                                             \
         size_t mask = nn - 1;                                                                                 \
         int ok = 1;                                                                                           \
         for (size_t e = 0; e < m->n && ok; ++e) {                                                             \
            K k = m->b[2 * e];                                                                                 \
            if (k == OR_EMPTY##W || k == OR_TOMB##W) continue;                                                 \
            uint32_t h = or_hash_idx##W(k, nsp);                                                               \
            int found = 0;                                                                                     \
            for (uint32_t i = 0; i < OR_CPU_BUCKET_OVERFLOW; ++i) {                                            \
               size_t s = (size_t)((uint32_t)(h + i)) & mask;                                                  \
               if (nb[2 * s] == OR_EMPTY##W) {                                                                 \
                  nb[2 * s] = k;                                                                               \
                  nb[2 * s + 1] = m->b[2 * e + 1];                                                             \
                  found = 1;                                                                                   \
                  break;                                                                                       \
               }                                                                                               \
            }                                                                                                  \
            if (!found) ok = 0;                                                                                \
         }                                                                                                     \
         if (!ok) {                                                                                            \
            free(nb);                                                                                          \
            ++nsp;                                                                                             \
            continue;                                                                                          \
         }                                                                                                     \
         free(m->b);                                                                                           \
         m->b = nb;                                                                                            \
         m->n = nn;                                                                                            \
         m->info.overflow = OR_CPU_BUCKET_OVERFLOW;                                                            \
         m->info.tombstones = 0;                                                                               \
         return;                                                                                               \
      }                                                                                                        \
   }                                                                                                           \
                                                                                                               \
   /* CPU performCleanupTasks (hashinator.h:631-639) */                                                        \
   static void or_map##W##_cleanup(or_map##W* m) {                                                             \
      while (m->info.overflow > OR_CPU_BUCKET_OVERFLOW) or_map##W##_rehash(m, m->info.size_power + 1);         \
      float ratio = m->info.tombstones == 0 ? 0.0f : (float)m->info.tombstones / (float)m->n;                  \
      if (ratio > 0.25) or_map##W##_rehash(m, m->info.size_power);                                             \
   }                                                                                                           \
                                                                                                               \
   /* _at (hashinator.h:340-403): returns the value slot index (word offset) for key. */                       \
   static size_t or_map##W##_at(or_map##W* m, K key) {                                                         \
      for (;;) {                                                                                               \
         size_t mask = m->n - 1;                                                                               \
         uint32_t h = or_hash_idx##W(key, m->info.size_power);                                                 \
         for (size_t i = 0; i < m->n; ++i) {                                                                   \
            size_t s = ((size_t)h + i) & mask;                                                                 \
            K c = m->b[2 * s];                                                                                 \
            if (c == key) return 2 * s + 1;                                                                    \
            if (c == OR_EMPTY##W) {                                                                            \
               m->b[2 * s] = key;                                                                              \
               m->info.fill++;                                                                                 \
               return 2 * s + 1;                                                                               \
            }                                                                                                  \
            if (c == OR_TOMB##W) {                                                                             \
               int exists = 0;                                                                                 \
               m->b[2 * s] = key;                                                                              \
               m->info.tombstones--;                                                                           \
               for (size_t j = i + 1; j < m->n; ++j) {                                                         \
                  size_t d = ((size_t)h + j) & mask;                                                           \
                  if (m->b[2 * d] == key) {                                                                    \
                     exists = 1;                                                                               \
                     m->b[2 * s + 1] = m->b[2 * d + 1];                                                        \
                     size_t nx = ((size_t)h + j + 1) & mask;                                                   \
                     if (m->b[2 * nx] == OR_EMPTY##W || j + 1 >= m->info.overflow) {                           \
                        m->b[2 * d] = OR_EMPTY##W;                                                             \
                     } else {                                                                                  \
                        m->b[2 * d] = OR_TOMB##W;                                                              \
                        m->info.tombstones++;                                                                  \
                     }                                                                                         \
                     break;                                                                                    \
                  }                                                                                            \
               }                                                                                               \
               if (!exists) m->info.fill++;                                                                    \
               return 2 * s + 1;                                                                               \
            }                                                                                                  \
         }                                                                                                     \
         /* full: rehash(sp+1) then at(key) = performCleanupTasks + _at */                                     \
         or_map##W##_rehash(m, m->info.size_power + 1);                                                        \
         or_map##W##_cleanup(m);                                                                               \
      }                                                                                                        \
   }                                                                                                           \
                                                                                                               \
   /* CPU-only batch insert (hashinator.h:1774-1778) */                                                        \
   void or_map##W##_insert(or_map##W* m, const K* keys, const K* vals, size_t n) {                             \
      for (size_t i = 0; i < n; ++i) {                                                                         \
         size_t slot = or_map##W##_at(m, keys[i]); /* may rehash: read m->b afterwards */                      \
         m->b[slot] = vals[i];                                                                                 \
      }                                                                                                        \
   }                                                                                                           \
                                                                                                               \
   /* const find (hashinator.h:742-768): skip TOMBSTONE, stop at EMPTY, whole-table bound. */                  \
   static int64_t or_map##W##_find1(const or_map##W* m, K key) {                                               \
      size_t mask = m->n - 1;                                                                                  \
      uint32_t h = or_hash_idx##W(key, m->info.size_power);                                                    \
      for (size_t i = 0; i < m->n; ++i) {                                                                      \
         size_t s = ((size_t)h + i) & mask;                                                                    \
         K c = m->b[2 * s];                                                                                    \
         if (c == OR_TOMB##W) continue;                                                                        \
         if (c == key) return (int64_t)s;                                                                      \
         if (c == OR_EMPTY##W) return -1;                                                                      \
      }                                                                                                        \
      return -1;                                                                                               \
   }                                                                                                           \
   void or_map##W##_find(const or_map##W* m, const K* keys, K* vals, uint8_t* hit, size_t n) {                 \
      for (size_t i = 0; i < n; ++i) {                                                                         \
         int64_t s = or_map##W##_find1(m, keys[i]);                                                            \
         if (s >= 0) vals[i] = m->b[2 * s + 1];                                                                \
         if (hit) hit[i] = s >= 0;                                                                             \
      }                                                                                                        \
   }                                                                                                           \
                                                                                                               \
   /* erase(key) (hashinator.h:842-850) -> non-const find (770-797, runs performCleanupTasks first)            \
    * -> erase(iterator) (822-832). */                                                                        \
   size_t or_map##W##_erase(or_map##W* m, const K* keys, size_t n) {                                           \
      size_t erased = 0;                                                                                       \
      for (size_t i = 0; i < n; ++i) {                                                                         \
         or_map##W##_cleanup(m);                                                                               \
         int64_t s = or_map##W##_find1(m, keys[i]);                                                            \
         if (s < 0) continue;                                                                                  \
         K c = m->b[2 * s];                                                                                    \
         if (c != OR_EMPTY##W && c != OR_TOMB##W) {                                                            \
            m->b[2 * s] = OR_TOMB##W;                                                                          \
            m->info.fill--;                                                                                    \
            m->info.tombstones++;                                                                              \
         }                                                                                                     \
         ++erased;                                                                                             \
      }                                                                                                        \
      return erased;                                                                                           \
   }                                                                                                           \
                                                                                                               \
   /* ---------------- accelerated-path semantics, one key at a time ---------------- */                       \
   void or_gpu##W##_init(K* b, or_info* info, int sp) {                                                        \
      size_t n = (size_t)1 << sp;                                                                              \
      for (size_t i = 0; i < n; ++i) {                                                                         \
         b[2 * i] = OR_EMPTY##W;                                                                               \
         b[2 * i + 1] = 0;                                                                                     \
      }                                                                                                        \
      info->size_power = sp;                                                                                   \
      info->_pad0 = 0;                                                                                         \
      info->fill = 0;                                                                                          \
      info->overflow = OR_GPU_BUCKET_OVERFLOW;                                                                 \
      info->tombstones = 0;                                                                                    \
      info->err = OR_INVALID;                                                                                  \
      info->_pad1 = 0;                                                                                         \
   }                                                                                                           \
                                                                                                               \
   /* insert_kernel (kernels_AMD.h:272-405): an existing key gets its value overwritten; otherwise the key   \
    * claims the first EMPTY in probe order (TOMBSTONEs count as occupied); fill += 1; overflow =            \
    * max(overflow, nextOverflow(offset + 1, 64)) as with atomicMax (kernels_NVIDIA.h:436). */               \
   static int or_gpu##W##_insert1(K* b, or_info* info, K key, K val) {                                         \
      size_t n = (size_t)1 << info->size_power, mask = n - 1;                                                  \
      uint64_t h = (uint64_t)or_hash_idx##W(key, info->size_power);                                            \
      for (size_t off = 0; off < n; ++off) {                                                                   \
         size_t s = (size_t)(h + off) & mask;                                                                  \
         K c = b[2 * s];                                                                                       \
         if (c == key) {                                                                                       \
            b[2 * s + 1] = val;                                                                                \
            return 1;                                                                                          \
         }                                                                                                     \
         if (c == OR_EMPTY##W) {                                                                               \
            b[2 * s] = key;                                                                                    \
            b[2 * s + 1] = val;                                                                                \
            info->fill++;                                                                                      \
            uint64_t o = or_next_overflow(off + 1, OR_GPU_BUCKET_OVERFLOW);                                    \
            if (o > info->overflow) info->overflow = o;                                                        \
            return 1;                                                                                          \
         }                                                                                                     \
      }                                                                                                        \
      return 0;                                                                                                \
   }                                                                                                           \
   /* Hasher::insert sets err = success before the launch (hashers.h:50); a key that finds no slot sets      \
    * fail (kernels_NVIDIA.h:278-281). */                                                                      \
   int or_gpu##W##_insert(K* b, or_info* info, const K* keys, const K* vals, size_t n) {                       \
      info->err = OR_SUCCESS;                                                                                  \
      for (size_t i = 0; i < n; ++i)                                                                           \
         if (!or_gpu##W##_insert1(b, info, keys[i], vals[i])) info->err = OR_FAIL;                             \
      return info->err;                                                                                        \
   }                                                                                                           \
   /* insert_index_kernel (kernels_AMD.h:507-641): value = input index. */                                   \
   int or_gpu##W##_insert_index(K* b, or_info* info, const K* keys, size_t n) {                                \
      info->err = OR_SUCCESS;                                                                                  \
      for (size_t i = 0; i < n; ++i)                                                                           \
         if (!or_gpu##W##_insert1(b, info, keys[i], (K)i)) info->err = OR_FAIL;                                \
      return info->err;                                                                                        \
   }                                                                                                           \
   /* insert_kernel over hash_pair input (kernels_AMD.h:136-270). */                                         \
   int or_gpu##W##_insert_pairs(K* b, or_info* info, const K* pairs, size_t n) {                               \
      info->err = OR_SUCCESS;                                                                                  \
      for (size_t i = 0; i < n; ++i)                                                                           \
         if (!or_gpu##W##_insert1(b, info, pairs[2 * i], pairs[2 * i + 1])) info->err = OR_FAIL;               \
      return info->err;                                                                                        \
   }                                                                                                           \
                                                                                                               \
   /* retrieve_kernel (kernels_AMD.h:647-697): probe offsets below currentMaxBucketOverflow; a match hits,   \
    * an EMPTY ends the probe as a miss, TOMBSTONEs are stepped over; a miss leaves vals[i] untouched. */      \
   void or_gpu##W##_retrieve(const K* b, const or_info* info, const K* keys, K* vals, uint8_t* hit,            \
                             size_t n) {                                                                       \
      size_t mask = ((size_t)1 << info->size_power) - 1;                                                       \
      for (size_t i = 0; i < n; ++i) {                                                                         \
         uint64_t h = (uint64_t)or_hash_idx##W(keys[i], info->size_power);                                     \
         int found = 0;                                                                                        \
         for (uint64_t off = 0; off < info->overflow; ++off) {                                                 \
            size_t s = (size_t)(h + off) & mask;                                                               \
            K c = b[2 * s];                                                                                    \
            if (c == keys[i]) {                                                                                \
               vals[i] = b[2 * s + 1];                                                                         \
               found = 1;                                                                                      \
               break;                                                                                          \
            }                                                                                                  \
            if (c == OR_EMPTY##W) break;                                                                       \
         }                                                                                                     \
         if (hit) hit[i] = (uint8_t)found;                                                                     \
      }                                                                                                        \
   }                                                                                                           \
                                                                                                               \
   /* delete_kernel (kernels_AMD.h:412-505) + Hashmap::erase fill bookkeeping (hashinator.h:1397-1409).      \
    * The key slot becomes TOMBSTONE, the value is kept. */                                                    \
   void or_gpu##W##_erase(K* b, or_info* info, const K* keys, size_t n) {                                      \
      size_t mask = ((size_t)1 << info->size_power) - 1;                                                       \
      for (size_t i = 0; i < n; ++i) {                                                                         \
         uint64_t h = (uint64_t)or_hash_idx##W(keys[i], info->size_power);                                     \
         for (uint64_t off = 0; off < info->overflow; ++off) {                                                 \
            size_t s = (size_t)(h + off) & mask;                                                               \
            K c = b[2 * s];                                                                                    \
            if (c == keys[i]) {                                                                                \
               b[2 * s] = OR_TOMB##W;                                                                          \
               info->tombstones++;                                                                             \
               info->fill--;                                                                                   \
               break;                                                                                          \
            }                                                                                                  \
            if (c == OR_EMPTY##W) break;                                                                       \
         }                                                                                                     \
      }                                                                                                        \
   }                                                                                                           \
                                                                                                               \
   /* reset_all_to_empty (kernels_AMD.h:26-43): keys of slots [0, len) -> EMPTY, fill = 0 (written by thread  \
    * 0, 38-40); values, tombstone count and overflow are left as they are.  Hasher::reset_all launches       \
    * ceil(len / MAX_BLOCKSIZE) blocks (hashers.h:155-163): len == 0 launches nothing and changes nothing. */ \
   void or_gpu##W##_reset_all(K* b, or_info* info, size_t len) {                                               \
      if (len == 0) return;                                                                                    \
      for (size_t i = 0; i < len; ++i) b[2 * i] = OR_EMPTY##W;                                                 \
      info->fill = 0;                                                                                          \
   }                                                                                                           \
                                                                                                               \
   /* reset_to_empty (kernels_AMD.h:65-134) through Hasher::reset (hashers.h:144-152): for each src pair in   \
    * turn, the first slot in probe order from its home that holds its key becomes EMPTY; the probe covers   \
    * the whole table (the loop bound is 1 << sizePower, 100) and EMPTY slots do not end it (only a match is \
    * voted on, 112); the slot's value is kept (only .first is written, 119).  fill -= len whether or not a  \
    * key was found (76-78).  len == 0 launches nothing. */                                                   \
   void or_gpu##W##_reset(K* b, or_info* info, const K* pairs, size_t len) {                                   \
      if (len == 0) return;                                                                                    \
      const int sp = info->size_power;                                                                         \
      const size_t n = (size_t)1 << sp, mask = n - 1;                                                          \
      for (size_t i = 0; i < len; ++i) {                                                                       \
         const K key = pairs[2 * i];                                                                           \
         const uint64_t h = (uint64_t)or_hash_idx##W(key, sp);                                                 \
         for (size_t off = 0; off < n; ++off) {                                                                \
            const size_t s = (size_t)(h + off) & mask;                                                         \
            if (b[2 * s] == key) {                                                                             \
               b[2 * s] = OR_EMPTY##W;                                                                         \
               break;                                                                                          \
            }                                                                                                  \
         }                                                                                                     \
      }                                                                                                        \
      info->fill -= len;                                                                                       \
   }                                                                                                           \
                                                                                                               \
   /* clean_tombstones (hashinator.h:1261-1315): TOMBSTONE -> EMPTY, extract the live elements that are not   \
    * in their home slot (bucket order), empty their slots (reset_to_empty, fill -= count), reset the        \
    * overflow to BUCKET_OVERFLOW, reinsert them (fill += count).  Returns the number reinserted. */          \
   size_t or_gpu##W##_clean_tombstones(K* b, or_info* info) {                                                  \
      if (info->tombstones == 0) return 0;                                                                     \
      info->tombstones = 0;                                                                                    \
      int sp = info->size_power;                                                                               \
      size_t n = (size_t)1 << sp, mask = n - 1, nd = 0;                                                        \
      K* d = (K*)malloc(2 * n * sizeof(K));                                                                    \
      for (size_t s = 0; s < n; ++s) {                                                                         \
         K c = b[2 * s];                                                                                       \
         if (c == OR_TOMB##W) {                                                                                \
            b[2 * s] = OR_EMPTY##W;                                                                            \
            continue;                                                                                          \
         }                                                                                                     \
         if (c == OR_EMPTY##W) continue;                                                                       \
         if (((size_t)or_hash_idx##W(c, sp) & mask) != s) {                                                    \
            d[2 * nd] = c;                                                                                     \
            d[2 * nd + 1] = b[2 * s + 1];                                                                      \
            ++nd;                                                                                              \
         }                                                                                                     \
      }                                                                                                        \
      info->overflow = OR_GPU_BUCKET_OVERFLOW;                                                                 \
      if (nd == 0) {                                                                                           \
         free(d);                                                                                              \
         return 0;                                                                                             \
      }                                                                                                        \
      for (size_t i = 0; i < n; ++i) {                                                                         \
         K c = b[2 * i];                                                                                       \
         if (c != OR_EMPTY##W && (((size_t)or_hash_idx##W(c, sp) & mask) != i)) b[2 * i] = OR_EMPTY##W;         \
      }                                                                                                        \
      info->fill -= nd;                                                                                        \
      or_gpu##W##_insert_pairs(b, info, d, nd);                                                                \
      free(d);                                                                                                 \
      return nd;                                                                                               \
   }                                                                                                           \
                                                                                                               \
   /* device_rehash (hashinator.h:284-336): every live element of the old table (bucket order) is inserted    \
    * into a fresh table of 2^new_sp slots; Info is reset to Info(new_sp) first; status = success iff the    \
    * fill is preserved.  new_buckets must not alias old_buckets. */                                           \
   int or_gpu##W##_rehash(const K* ob, int osp, K* nb, int nsp, or_info* info) {                               \
      uint64_t prior = info->fill;                                                                             \
      or_gpu##W##_init(nb, info, nsp);                                                                         \
      size_t on = (size_t)1 << osp;                                                                            \
      info->err = OR_SUCCESS;                                                                                  \
      for (size_t s = 0; s < on; ++s) {                                                                        \
         K c = ob[2 * s];                                                                                      \
         if (c == OR_EMPTY##W || c == OR_TOMB##W) continue;                                                    \
         if (!or_gpu##W##_insert1(nb, info, c, ob[2 * s + 1])) info->err = OR_FAIL;                            \
      }                                                                                                        \
      info->err = (prior == info->fill) ? OR_SUCCESS : OR_FAIL;                                                \
      return info->err;                                                                                        \
   }                                                                                                           \
                                                                                                               \
   /* Longest probe sequence (offset + 1) over the live elements of a table. */                               \
   uint64_t or_gpu##W##_max_probe(const K* b, int sp) {                                                        \
      size_t n = (size_t)1 << sp, mask = n - 1;                                                                \
      uint64_t mx = 0;                                                                                         \
      for (size_t s = 0; s < n; ++s) {                                                                         \
         K c = b[2 * s];                                                                                       \
         if (c == OR_EMPTY##W || c == OR_TOMB##W) continue;                                                    \
         uint64_t off = (uint64_t)((s - ((size_t)or_hash_idx##W(c, sp) & mask)) & mask) + 1;                    \
         if (off > mx) mx = off;                                                                               \
      }                                                                                                        \
      return mx;                                                                                               \
   }                                                                                                           \
                                                                                                               \
   /* Robin Hood order: every cluster (maximal run of non-EMPTY slots, taken cyclically) sorted by           \
    * (home - cluster start) mod N, then by key, written back over the cluster's own slots; overflow =        \
    * max(64, nextOverflow(longest probe of the new layout, 64)).  No-op (returns 0) when the table holds     \
    * TOMBSTONEs or has no EMPTY slot.  Returns 1 when the table is now ordered.  (An addition of this        \
    * library: the reference keeps arrival order; kernels_AMD.h insert_kernel.) */                           \
   int or_gpu##W##_reorder(K* b, or_info* info) {                                                              \
      const int sp = info->size_power;                                                                         \
      const size_t n = (size_t)1 << sp, mask = n - 1;                                                          \
      if (info->tombstones) return 0;                                                                          \
      size_t e0 = n;                                                                                           \
      for (size_t i = 0; i < n; ++i)                                                                           \
         if (b[2 * i] == OR_EMPTY##W) {                                                                        \
            e0 = i;                                                                                            \
            break;                                                                                             \
         }                                                                                                     \
      if (e0 == n) return 0;                                                                                   \
      K* tmp = (K*)malloc(2 * n * sizeof(K));                                                                  \
      uint64_t* rel = (uint64_t*)malloc(n * sizeof(uint64_t));                                                 \
      uint64_t mx = 0;                                                                                         \
      size_t u = e0 + 1; /* unwrapped position */                                                              \
      while (u < e0 + 1 + n) {                                                                                 \
         if (b[2 * (u & mask)] == OR_EMPTY##W) {                                                               \
            ++u;                                                                                               \
            continue;                                                                                          \
         }                                                                                                     \
         size_t start = u, len = 0;                                                                            \
         while (b[2 * ((start + len) & mask)] != OR_EMPTY##W) {                                                \
            size_t q = (start + len) & mask;                                                                   \
            tmp[2 * len] = b[2 * q];                                                                           \
            tmp[2 * len + 1] = b[2 * q + 1];                                                                   \
            rel[len] = ((size_t)or_hash_idx##W(b[2 * q], sp) - start) & mask;                                  \
            ++len;                                                                                             \
         }                                                                                                     \
         /* insertion sort by (rel, key): clusters are short */                                                \
         for (size_t i = 1; i < len; ++i) {                                                                    \
            uint64_t r = rel[i];                                                                               \
            K k = tmp[2 * i], v = tmp[2 * i + 1];                                                              \
            size_t j = i;                                                                                      \
            while (j > 0 && (rel[j - 1] > r || (rel[j - 1] == r && tmp[2 * (j - 1)] > k))) {                   \
               rel[j] = rel[j - 1];                                                                            \
               tmp[2 * j] = tmp[2 * (j - 1)];                                                                  \
               tmp[2 * j + 1] = tmp[2 * (j - 1) + 1];                                                          \
               --j;                                                                                            \
            }                                                                                                  \
            rel[j] = r;                                                                                        \
            tmp[2 * j] = k;                                                                                    \
            tmp[2 * j + 1] = v;                                                                                \
         }                                                                                                     \
         for (size_t i = 0; i < len; ++i) {                                                                    \
            size_t q = (start + i) & mask;                                                                     \
            b[2 * q] = tmp[2 * i];                                                                             \
            b[2 * q + 1] = tmp[2 * i + 1];                                                                     \
            if (i - rel[i] + 1 > mx) mx = i - rel[i] + 1;                                                      \
         }                                                                                                     \
         u = start + len;                                                                                      \
      }                                                                                                        \
      free(tmp);                                                                                               \
      free(rel);                                                                                               \
      info->overflow = mx > OR_GPU_BUCKET_OVERFLOW ? or_next_overflow(mx, OR_GPU_BUCKET_OVERFLOW)             \
                                                   : OR_GPU_BUCKET_OVERFLOW;                                   \
      return 1;                                                                                                \
   }                                                                                                           \
                                                                                                               \
   static int or_pred_ok##W(const or_pred* p, K key, K val) {                                                  \
      K f = p->field == 0 ? key : val;                                                                         \
      switch (p->kind) {                                                                                       \
      case 0: return 1;                                                                                        \
      case 1: return key != (K)p->empty && key != (K)p->tombstone;                                             \
      case 2: return (f & (K)p->mask) == (K)p->value;                                                          \
      case 3: return f < (K)p->value;                                                                          \
      case 4: return f >= (K)p->value;                                                                         \
      default: return 0;                                                                                       \
      }                                                                                                        \
   }                                                                                                           \
   /* copy_if (split_tools.h:1128-1144): stable; out = [x for x in in if rule(x)]. */                          \
   size_t or_copy_if_pairs##W(const K* in, size_t n, const or_pred* p, K* out) {                               \
      size_t c = 0;                                                                                            \
      for (size_t i = 0; i < n; ++i)                                                                           \
         if (or_pred_ok##W(p, in[2 * i], in[2 * i + 1])) {                                                     \
            out[2 * c] = in[2 * i];                                                                            \
            out[2 * c + 1] = in[2 * i + 1];                                                                    \
            ++c;                                                                                               \
         }                                                                                                     \
      return c;                                                                                                \
   }                                                                                                           \
   /* copy_keys_if (split_tools.h:1094-1111): as copy_if, emitting .first only. */                             \
   size_t or_copy_keys_if_pairs##W(const K* in, size_t n, const or_pred* p, K* out) {                          \
      size_t c = 0;                                                                                            \
      for (size_t i = 0; i < n; ++i)                                                                           \
         if (or_pred_ok##W(p, in[2 * i], in[2 * i + 1])) out[c++] = in[2 * i];                                 \
      return c;                                                                                                \
   }

OR_DEFINE(32, uint32_t)
OR_DEFINE(64, uint64_t)

size_t or_copy_if_flags(const void* in, const uint8_t* flags, size_t n, size_t elem_bytes, void* out) {
   const char* src = (const char*)in;
   char* dst = (char*)out;
   size_t c = 0;
   for (size_t i = 0; i < n; ++i)
      if (flags[i]) {
         memcpy(dst + c * elem_bytes, src + i * elem_bytes, elem_bytes);
         ++c;
      }
   return c;
}

/* split_prefix_scan (split_tools.h:363-418): exclusive, wrapping in the element type. */
void or_exclusive_scan_u32(const uint32_t* in, uint32_t* out, size_t n) {
   uint32_t acc = 0;
   for (size_t i = 0; i < n; ++i) {
      uint32_t x = in[i];
      out[i] = acc;
      acc += x;
   }
}

void or_exclusive_scan_u64(const uint64_t* in, uint64_t* out, size_t n) {
   uint64_t acc = 0;
   for (size_t i = 0; i < n; ++i) {
      uint64_t x = in[i];
      out[i] = acc;
      acc += x;
   }
}
