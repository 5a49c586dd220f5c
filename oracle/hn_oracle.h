/*
 * hn_oracle.h -- CPU restatement of Hashinator's map and compaction semantics.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  Nothing in hashinator_amd/ links or calls it.
 *
 * Two families of functions live here:
 *
 *  or_map{32,64}_*   the reference's host-only map (HASHINATOR_CPU_ONLY_MODE):
 *                    serial _at insertion with tombstone reuse, const find,
 *                    per-key erase with performCleanupTasks, host rehash with
 *                    the 32-probe placement limit.  Bucket images produced here
 *                    are pinned against the reference's own CPU-only build.
 *
 *  or_gpu{32,64}_*   a sequential statement of the accelerated path's
 *                    per-key semantics (kernels_AMD.h / kernels_NVIDIA.h):
 *                    tombstones count as occupied on insert, the insert claims
 *                    the first EMPTY in probe order, retrieve/erase are bounded
 *                    by currentMaxBucketOverflow, clean_tombstones extracts the
 *                    displaced elements and reinserts them, device rehash
 *                    reinserts every live element into a fresh table.
 *
 * plus stable compaction and exclusive scan (split_tools.h semantics).
 */
#ifndef HN_ORACLE_H
#define HN_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Byte-for-byte the layout of Hashinator::Info (defaults.h:44-54). */
typedef struct or_info {
   int32_t size_power;
   int32_t _pad0;
   uint64_t fill;
   uint64_t overflow;   /* currentMaxBucketOverflow */
   uint64_t tombstones; /* tombstoneCounter */
   int32_t err;         /* status: 0 success, 1 fail, 2 invalid */
   int32_t _pad1;
} or_info;

enum { OR_SUCCESS = 0, OR_FAIL = 1, OR_INVALID = 2 };

/* Predicate descriptor for compaction (same meaning as hn_pred in the C-ABI). */
typedef struct or_pred {
   int32_t kind;  /* 0 all, 1 valid key, 2 (field & mask) == value, 3 field < value, 4 field >= value */
   int32_t field; /* 0 key (first), 1 value (second) */
   uint64_t mask;
   uint64_t value;
   uint64_t empty;     /* EMPTYBUCKET sentinel of the element type */
   uint64_t tombstone; /* TOMBSTONE sentinel */
} or_pred;

/* ---- hash (hashfunctions.h:39-43, 52-56) ---- */
uint32_t or_fibhash32(uint32_t key, int size_power);
uint64_t or_fibhash64(uint64_t key, int size_power);
void or_fibhash32_n(const uint32_t* keys, const int32_t* sps, uint32_t* out, size_t n);
void or_fibhash64_n(const uint64_t* keys, const int32_t* sps, uint64_t* out, size_t n);

/* ---- host-only map (hashinator.h CPU-only mode) ---- */
typedef struct or_map32 or_map32;
typedef struct or_map64 or_map64;

or_map32* or_map32_new(int size_power);
void or_map32_free(or_map32* m);
const or_info* or_map32_info(const or_map32* m);
size_t or_map32_nbuckets(const or_map32* m);
void or_map32_copy_buckets(const or_map32* m, uint32_t* out /* 2*nbuckets */);
void or_map32_insert(or_map32* m, const uint32_t* keys, const uint32_t* vals, size_t n);
void or_map32_find(const or_map32* m, const uint32_t* keys, uint32_t* vals, uint8_t* hit, size_t n);
size_t or_map32_erase(or_map32* m, const uint32_t* keys, size_t n);
void or_map32_rehash(or_map32* m, int new_size_power);

or_map64* or_map64_new(int size_power);
void or_map64_free(or_map64* m);
const or_info* or_map64_info(const or_map64* m);
size_t or_map64_nbuckets(const or_map64* m);
void or_map64_copy_buckets(const or_map64* m, uint64_t* out);
void or_map64_insert(or_map64* m, const uint64_t* keys, const uint64_t* vals, size_t n);
void or_map64_find(const or_map64* m, const uint64_t* keys, uint64_t* vals, uint8_t* hit, size_t n);
size_t or_map64_erase(or_map64* m, const uint64_t* keys, size_t n);
void or_map64_rehash(or_map64* m, int new_size_power);

/* ---- accelerated-path semantics, sequential (caller owns buckets) ----
 * buckets are interleaved (key, value) pairs, 2*N words. */
void or_gpu32_init(uint32_t* buckets, or_info* info, int size_power);
int or_gpu32_insert(uint32_t* buckets, or_info* info, const uint32_t* keys, const uint32_t* vals, size_t n);
int or_gpu32_insert_index(uint32_t* buckets, or_info* info, const uint32_t* keys, size_t n);
int or_gpu32_insert_pairs(uint32_t* buckets, or_info* info, const uint32_t* pairs, size_t n);
void or_gpu32_retrieve(const uint32_t* buckets, const or_info* info, const uint32_t* keys, uint32_t* vals,
                       uint8_t* hit, size_t n);
void or_gpu32_erase(uint32_t* buckets, or_info* info, const uint32_t* keys, size_t n);
void or_gpu32_reset_all(uint32_t* buckets, or_info* info, size_t len);
void or_gpu32_reset(uint32_t* buckets, or_info* info, const uint32_t* pairs, size_t len);
size_t or_gpu32_clean_tombstones(uint32_t* buckets, or_info* info);
int or_gpu32_rehash(const uint32_t* old_buckets, int old_sp, uint32_t* new_buckets, int new_sp, or_info* info);
uint64_t or_gpu32_max_probe(const uint32_t* buckets, int size_power);
int or_gpu32_reorder(uint32_t* buckets, or_info* info);

void or_gpu64_init(uint64_t* buckets, or_info* info, int size_power);
int or_gpu64_insert(uint64_t* buckets, or_info* info, const uint64_t* keys, const uint64_t* vals, size_t n);
int or_gpu64_insert_index(uint64_t* buckets, or_info* info, const uint64_t* keys, size_t n);
int or_gpu64_insert_pairs(uint64_t* buckets, or_info* info, const uint64_t* pairs, size_t n);
void or_gpu64_retrieve(const uint64_t* buckets, const or_info* info, const uint64_t* keys, uint64_t* vals,
                       uint8_t* hit, size_t n);
void or_gpu64_erase(uint64_t* buckets, or_info* info, const uint64_t* keys, size_t n);
void or_gpu64_reset_all(uint64_t* buckets, or_info* info, size_t len);
void or_gpu64_reset(uint64_t* buckets, or_info* info, const uint64_t* pairs, size_t len);
size_t or_gpu64_clean_tombstones(uint64_t* buckets, or_info* info);
int or_gpu64_rehash(const uint64_t* old_buckets, int old_sp, uint64_t* new_buckets, int new_sp, or_info* info);
uint64_t or_gpu64_max_probe(const uint64_t* buckets, int size_power);
int or_gpu64_reorder(uint64_t* buckets, or_info* info);

/* ---- split::tools semantics ---- */
size_t or_copy_if_pairs32(const uint32_t* in, size_t n, const or_pred* p, uint32_t* out);
size_t or_copy_if_pairs64(const uint64_t* in, size_t n, const or_pred* p, uint64_t* out);
size_t or_copy_keys_if_pairs32(const uint32_t* in, size_t n, const or_pred* p, uint32_t* out);
size_t or_copy_keys_if_pairs64(const uint64_t* in, size_t n, const or_pred* p, uint64_t* out);
size_t or_copy_if_flags(const void* in, const uint8_t* flags, size_t n, size_t elem_bytes, void* out);
void or_exclusive_scan_u32(const uint32_t* in, uint32_t* out, size_t n);
void or_exclusive_scan_u64(const uint64_t* in, uint64_t* out, size_t n);

/* ---- input generation shared with the GPU side (SplitMix64) ---- */
uint64_t or_splitmix64(uint64_t* state);

#ifdef __cplusplus
}
#endif
#endif
