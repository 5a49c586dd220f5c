/*
 * hashinator_amd.h -- C-ABI of the MI355X (gfx950) accelerated hash-map path.
 *
 * Every entry point is asynchronous on the given HIP stream, takes plain
 * device-accessible pointers (hipMalloc or hipMallocManaged) and returns an
 * int: 0 on success, a hipError_t value on a HIP failure, or one of the
 * negative HN_E* codes below for invalid arguments.  Entry points are
 * reentrant across maps; calls on the same map must be ordered by the caller
 * (same stream, or events).
 *
 * Key/value instantiations: (uint32_t, uint32_t) as *_u32 and
 * (uint64_t, uint64_t) as *_u64.  Sentinels follow the reference:
 * EMPTYBUCKET = all ones, TOMBSTONE = EMPTYBUCKET - 1 (hashinator.h:59-60);
 * callers must not insert either as a key.
 *
 * Bucket layout is the reference's hash_pair<K,V> array (hash_pair.h:29-75):
 * 2^size_power slots of {K first; V second;}, 8 B (u32) or 16 B (u64) each.
 *
 * The reference interface each entry point replaces is cited as
 * file:line in kstppd/hashinator include/.
 */
#ifndef HASHINATOR_AMD_H
#define HASHINATOR_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t* hn_stream_t; /* == hipStream_t */

#if defined(__GNUC__)
#define HN_API __attribute__((visibility("default")))
#else
#define HN_API
#endif

/* Map metadata; byte-for-byte Hashinator::Info (hashinator/defaults.h:44-54).
 * reserved0 is zero between calls (hn_info_init sets it).  reserved1
 * holds flags: bit 0 is transient (zero between calls), bit 1
 * (HN_INFO_ORDERED) says the table is in Robin Hood order -- set by
 * hn_reorder_*, the bulk build and a complete clean, cleared by every
 * probing insert, reset and rehash; a caller that writes bucket memory
 * itself must clear it (lookups stop early on its word, and the clean
 * skips its sort for segments whose homes it finds in order). */
typedef struct hn_info {
   int32_t size_power;
   int32_t reserved0;
   uint64_t fill;
   uint64_t current_max_bucket_overflow;
   uint64_t tombstone_counter;
   int32_t err; /* HN_STATUS_* == Hashinator::status (common.h:85) */
   int32_t reserved1;
} hn_info;

enum { HN_STATUS_SUCCESS = 0, HN_STATUS_FAIL = 1, HN_STATUS_INVALID = 2 };
#define HN_INFO_ORDERED 2

/* BUCKET_OVERFLOW of the wave64 build (defaults.h:28-31): the initial and
 * post-clean value of current_max_bucket_overflow, and the unit it grows in. */
#define HN_BUCKET_OVERFLOW 64

/* Error codes (negative, so they never collide with hipError_t). */
#define HN_EINVAL (-1)   /* bad pointer, size or size_power */
#define HN_ESCRATCH (-2) /* scratch buffer too small */
#define HN_ENOLIB (-3)   /* no gfx950 code object for the current device */

/* Launch tuning.  NULL means "use the tuned defaults".  probe_width is the
 * number of lanes that probe one key together (1, 2, 4, 8, 16, 32 or 64; a
 * window of probe_width consecutive, aligned slots is read per step).
 * probe_width 0 (and NULL) selects the default: for inserts 4 lanes; for
 * reset 8 (u64) or 16 (u32); for retrieve and erase the lane kernel, one key per lane, whose
 * 128-byte windows are loaded by 8 lanes together and handed to their lanes
 * through LDS (it runs one wave per block and ignores block_size).
 * block_size is threads per workgroup (64..1024, multiple of 64); grid_blocks
 * 0 means derived from len and the device's CU count. */
typedef struct hn_tune {
   int32_t probe_width;
   int32_t block_size;
   int32_t grid_blocks;
   int32_t reserved;
} hn_tune;

/* Built-in compaction predicates (the C-ABI cannot carry a functor: a rule
 * evaluated elsewhere comes in as flags or bits, hn_copy_if_flags /
 * hn_copy_if_bits, and C++ callers compile the kernel with their rule from
 * <splitvector/split_compact.h>). */
typedef struct hn_pred {
   int32_t kind;  /* HN_PRED_* */
   int32_t field; /* 0: key (first), 1: value (second) */
   uint64_t mask;
   uint64_t value;
} hn_pred;

enum {
   HN_PRED_ALL = 0,    /* every element */
   HN_PRED_VALID = 1,  /* key != EMPTYBUCKET && key != TOMBSTONE */
   HN_PRED_MASKEQ = 2, /* (field & mask) == value */
   HN_PRED_LT = 3,     /* field < value */
   HN_PRED_GE = 4      /* field >= value */
};

/* ---------------------------------------------------------------- map setup */

/* Bucket arrays are n_slots hash_pair<K,K> in device memory, expected to be
 * aligned to the slot size (8 bytes for u32, 16 for u64), as hipMalloc and
 * the library's allocators give.  A 32-bit-key insert claims an EMPTY slot
 * with one 64-bit CAS of key and value when the slot is 8-byte aligned, and
 * with a CAS of the key followed by a value store otherwise. */

/* Info(int) (defaults.h:46-48) written to device memory: size_power, fill 0,
 * overflow HN_BUCKET_OVERFLOW, tombstones 0, err HN_STATUS_INVALID. */
HN_API int hn_info_init(hn_info* info, int size_power, hn_stream_t s);

/* SplitVector<hash_pair>(1<<sp, {EMPTYBUCKET, V()}) construction
 * (hashinator.h:111-114, splitvec.h:233-238), done on the device. */
HN_API int hn_table_init_u32(void* buckets, size_t n_slots, hn_stream_t s);
HN_API int hn_table_init_u64(void* buckets, size_t n_slots, hn_stream_t s);

/* ---------------------------------------------------------------- insert */

/* Hasher::insert(keys, vals, buckets, info, len, s) (hashers.h:46-65) ->
 * insert_kernel (kernels_AMD.h:272-405).  An existing key has its value
 * overwritten; a new key claims the first EMPTY in probe order (TOMBSTONEs are
 * not reused).  fill += new keys; overflow = max(overflow,
 * nextOverflow(longest probe, 64)); err = SUCCESS, or FAIL if a key found no
 * EMPTY slot in the whole table. */
HN_API int hn_insert_u32(const uint32_t* keys, const uint32_t* vals, void* buckets, hn_info* info, size_t len,
                  hn_stream_t s, const hn_tune* t);
HN_API int hn_insert_u64(const uint64_t* keys, const uint64_t* vals, void* buckets, hn_info* info, size_t len,
                  hn_stream_t s, const hn_tune* t);

/* Hasher::insertIndex (hashers.h:68-87) -> insert_index_kernel
 * (kernels_AMD.h:507-641): value = input index. */
HN_API int hn_insert_index_u32(const uint32_t* keys, void* buckets, hn_info* info, size_t len, hn_stream_t s,
                        const hn_tune* t);
HN_API int hn_insert_index_u64(const uint64_t* keys, void* buckets, hn_info* info, size_t len, hn_stream_t s,
                        const hn_tune* t);

/* Hasher::insert(src pairs, ...) (hashers.h:90-108) -> insert_kernel over
 * hash_pair input (kernels_AMD.h:136-270). */
HN_API int hn_insert_pairs_u32(const void* pairs, void* buckets, hn_info* info, size_t len, hn_stream_t s,
                        const hn_tune* t);
HN_API int hn_insert_pairs_u64(const void* pairs, void* buckets, hn_info* info, size_t len, hn_stream_t s,
                        const hn_tune* t);

/* ---------------------------------------------------------------- retrieve */

/* Hasher::retrieve(keys, vals, ...) (hashers.h:111-119) -> retrieve_kernel
 * (kernels_AMD.h:647-697).  Probe offsets below
 * current_max_bucket_overflow; on a hit vals[i] = value, on a miss vals[i] is
 * left untouched.  `hit` (nullable) receives 1/0 per key -- an addition, the
 * reference has no hit output. */
HN_API int hn_retrieve_u32(const uint32_t* keys, uint32_t* vals, uint8_t* hit, const void* buckets,
                    const hn_info* info, size_t len, hn_stream_t s, const hn_tune* t);
HN_API int hn_retrieve_u64(const uint64_t* keys, uint64_t* vals, uint8_t* hit, const void* buckets,
                    const hn_info* info, size_t len, hn_stream_t s, const hn_tune* t);

/* Hasher::retrieve(src pairs, ...) (hashers.h:121-129) -> retrieve_kernel over
 * hash_pair (kernels_AMD.h:703-752): src[i].second is written on a hit. */
HN_API int hn_retrieve_pairs_u32(void* pairs, uint8_t* hit, const void* buckets, const hn_info* info, size_t len,
                          hn_stream_t s, const hn_tune* t);
HN_API int hn_retrieve_pairs_u64(void* pairs, uint8_t* hit, const void* buckets, const hn_info* info, size_t len,
                          hn_stream_t s, const hn_tune* t);

/* ---------------------------------------------------------------- erase */

/* Hashmap::erase(keys, len) (hashinator.h:1397-1409) -> Hasher::erase
 * (hashers.h:132-142) -> delete_kernel (kernels_AMD.h:412-505).  A found key
 * becomes TOMBSTONE (compare-and-swap, so a key repeated in one batch counts
 * once); tombstone_counter += erased, fill -= erased (done on the device). */
HN_API int hn_erase_u32(const uint32_t* keys, void* buckets, hn_info* info, size_t len, hn_stream_t s,
                 const hn_tune* t);
HN_API int hn_erase_u64(const uint64_t* keys, void* buckets, hn_info* info, size_t len, hn_stream_t s,
                 const hn_tune* t);

/* ---------------------------------------------------------------- clear */

/* clear(targets::device) (hashinator.h:491-518) -> Hasher::reset_all
 * (hashers.h:155-163) -> reset_all_to_empty (kernels_AMD.h:26-43).  Keys of
 * the first `len` slots become EMPTYBUCKET, fill = 0; values, tombstone
 * counter and overflow are left as they are.  len == 0 changes nothing (the
 * reference launches no block then). */
HN_API int hn_reset_all_u32(void* buckets, hn_info* info, size_t len, hn_stream_t s);
HN_API int hn_reset_all_u64(void* buckets, hn_info* info, size_t len, hn_stream_t s);

/* Hasher::reset(src, dst, info, len, s) (hashers.h:144-152) -> reset_to_empty
 * (kernels_AMD.h:65-134): for each src pair, the first slot in probe order
 * holding its key (probing the whole table; EMPTY does not end the probe)
 * becomes EMPTY, its value kept; fill -= len, found or not.  The EMPTY holes
 * this leaves inside probe paths are what clean_tombstones' reinsert fills;
 * an insert of a key that still has a copy behind a hole updates that copy
 * when it lies in the same probe window as the hole (as the reference's
 * 64-slot windows do) and otherwise claims the hole. */
HN_API int hn_reset_u32(const void* src_pairs, void* buckets, hn_info* info, size_t len, hn_stream_t s);
HN_API int hn_reset_u64(const void* src_pairs, void* buckets, hn_info* info, size_t len, hn_stream_t s);

/* ---------------------------------------------------------------- clean / rehash */

/* Scratch for hn_clean_tombstones on a table of 2^size_power slots when the
 * caller can spare `budget_bytes`.  0: the default, the segment pass's fixed
 * part (a few MiB at 2^28 slots) plus a pool for segments of up to 1/32 of
 * the table (at least 1 MiB); 294 MiB at 2^28 u64 slots.  SIZE_MAX: enough
 * for any table, one with no EMPTY slot or one segment spanning it included
 * (36 bytes per slot for u64 keys, 28 for u32).  A budget of at least the fixed part plus 64 KiB
 * is returned as it is: the pool takes the rest.  Below that, the region
 * clean's size within the budget, at least hn_clean_tombstones_min_scratch()
 * bytes. */
HN_API size_t hn_clean_tombstones_scratch_bytes_u32(int size_power, size_t budget_bytes);
HN_API size_t hn_clean_tombstones_scratch_bytes_u64(int size_power, size_t budget_bytes);
HN_API size_t hn_clean_tombstones_min_scratch(void);

/* Hashmap::clean_tombstones (hashinator.h:1261-1315).  No-op when
 * tombstone_counter == 0 (checked on the device).  Otherwise TOMBSTONEs
 * become EMPTY and the live elements are placed again: tombstone_counter = 0,
 * fill is unchanged, overflow = max(64, nextOverflow(longest probe)), and the
 * occupied-slot set is the one a fresh insert of the live set gives.
 * Segment pass (scratch of at least the fixed part plus 64 KiB): the table
 * splits at its EMPTY slots into segments, which no live element's probe
 * crosses; each is rewritten in place in Robin Hood order (home slot, ties by
 * key), in LDS up to 8192 slots and through the pool above that, and the
 * table ends HN_INFO_ORDERED.  A segment longer than the pool holds is left
 * as it was: its tombstones stay counted in tombstone_counter, err = FAIL
 * (when it held any), the others are cleaned and HN_INFO_ORDERED is not
 * set.  A table with no EMPTY slot is cleaned by the region clean over the
 * whole table and then reordered when the pool holds the whole table
 * (SIZE_MAX), and left unchanged with err = FAIL otherwise.  Region clean
 * (less scratch): region by region from an EMPTY slot, every live element
 * not in its home slot is lifted out and reinserted; bounded by the region
 * size, not ordered; a table with no EMPTY slot gets err = FAIL unless one
 * region holds it. */
HN_API int hn_clean_tombstones_u32(void* buckets, hn_info* info, int size_power, void* scratch, size_t scratch_bytes,
                            hn_stream_t s);
HN_API int hn_clean_tombstones_u64(void* buckets, hn_info* info, int size_power, void* scratch, size_t scratch_bytes,
                            hn_stream_t s);

/* Hashmap::device_rehash (hashinator.h:284-336): every live element of
 * old_buckets (2^old_sp slots) is inserted into new_buckets (2^new_sp slots,
 * initialised here to {EMPTYBUCKET, 0}); *info becomes Info(new_sp) plus the
 * inserted elements; err = SUCCESS iff the fill is preserved.  The buffers
 * must not overlap.  `scratch` must hold hn_rehash_scratch_bytes() bytes. */
HN_API size_t hn_rehash_scratch_bytes(void);
HN_API int hn_rehash_u32(const void* old_buckets, int old_sp, void* new_buckets, int new_sp, hn_info* info, void* scratch,
                  hn_stream_t s);
HN_API int hn_rehash_u64(const void* old_buckets, int old_sp, void* new_buckets, int new_sp, hn_info* info, void* scratch,
                  hn_stream_t s);

/* Robin Hood reorder (no counterpart in the reference: the reference's
 * clusters stay in arrival order).  Every cluster (maximal run of non-EMPTY
 * slots) is rewritten in order of home slot, ties by key -- the layout that
 * inserting the same keys in sorted-home order gives.  Live set, occupancy,
 * fill and tombstone_counter are unchanged; current_max_bucket_overflow is
 * recomputed from the new layout (max(64, nextOverflow(longest probe)); it
 * never grows).  Then sets HN_INFO_ORDERED in info->reserved1, and retrieve and
 * erase stop a probe at the first occupant displaced less than the probe
 * offset (an absent key costs about as much as a present one).  No-op when
 * tombstone_counter != 0 or the table has no EMPTY slot; a cluster longer
 * than 8192 - 1024 slots is left as it is, the flag is not set and the
 * overflow is only raised (never lowered) by the reordered clusters.
 * `scratch` must hold hn_reorder_scratch_bytes(size_power) bytes. */
HN_API size_t hn_reorder_scratch_bytes(int size_power);
HN_API int hn_reorder_u32(void* buckets, hn_info* info, int size_power, void* scratch, size_t scratch_bytes,
                          hn_stream_t s);
HN_API int hn_reorder_u64(void* buckets, hn_info* info, int size_power, void* scratch, size_t scratch_bytes,
                          hn_stream_t s);

/* Bulk insert into an EMPTY table (fill == 0, tombstone_counter == 0):
 * Hasher::insert (hashers.h:55-70, kernels_AMD.h insert_kernel) on a fresh
 * map, producing directly the image hn_insert_* followed by hn_reorder_*
 * gives -- every cluster in home-slot order, ties by key, HN_INFO_ORDERED set,
 * overflow = max(64, nextOverflow(longest probe)).  A key repeated in the
 * batch keeps the value of its last occurrence (what a serial insert leaves).
 * `vals` NULL stores each key's index in the batch (insertIndex,
 * hashers.h:72-86); the _pairs variant reads hash_pair elements.  No probing
 * and no atomics on the table: a stable radix partition by 2048-slot region,
 * an LDS sort per region and a carry scan (about 10 streaming passes' worth
 * of the pair array).  Needs len < 2^size_power and `scratch` of
 * hn_build_scratch_bytes_*(size_power, len) bytes, 256-byte aligned as
 * hipMalloc gives (else HN_EINVAL; the same holds for the scratch of
 * hn_rehash_build_* and hn_merge_*); the table's previous slot
 * contents are overwritten, every slot: the bucket memory of a fresh table
 * need not be initialised (hn_table_init_*) first, only its Info
 * (hn_info_init).  When the table is not empty the call sets
 * err = INVALID and leaves it untouched (insert with hn_insert_* instead).
 * When one 2048-slot region receives more than 2560 keys (strongly skewed
 * keys) the table is filled by the probing insert instead, on the device:
 * same live set, arrival order, HN_INFO_ORDERED not set.  tune->reserved
 * (1..9, 0 = 9) caps the partition digit width in bits; probe_width and
 * block_size are not used. */
HN_API size_t hn_build_scratch_bytes_u32(int size_power, size_t len);
HN_API size_t hn_build_scratch_bytes_u64(int size_power, size_t len);
HN_API int hn_build_u32(const void* keys, const void* vals, void* buckets, hn_info* info, size_t len, int size_power,
                        void* scratch, size_t scratch_bytes, hn_stream_t s, const hn_tune* t);
HN_API int hn_build_u64(const void* keys, const void* vals, void* buckets, hn_info* info, size_t len, int size_power,
                        void* scratch, size_t scratch_bytes, hn_stream_t s, const hn_tune* t);
HN_API int hn_build_pairs_u32(const void* pairs, void* buckets, hn_info* info, size_t len, int size_power,
                              void* scratch, size_t scratch_bytes, hn_stream_t s, const hn_tune* t);
HN_API int hn_build_pairs_u64(const void* pairs, void* buckets, hn_info* info, size_t len, int size_power,
                              void* scratch, size_t scratch_bytes, hn_stream_t s, const hn_tune* t);

/* Insert into a table that already holds keys (Hasher::insert,
 * hashers.h:55-70, kernels_AMD.h insert_kernel 136-405), written to a
 * second bucket array of the same size: new_buckets receives the image
 * hn_insert_* followed by hn_reorder_* gives -- the union of the table's
 * live set and the batch, every cluster in home-slot order, ties by key; a
 * key already present takes the batch's value, a key repeated in the batch
 * its last one -- and *info becomes that table's (fill, overflow =
 * max(64, nextOverflow(longest probe)), HN_INFO_ORDERED, err = SUCCESS).
 * old_buckets is only read.  No probing and no atomics on the table: the
 * batch takes the bulk build's partition passes, then each 2048-slot region
 * merges its run of the old table (in Robin Hood order the keys of a region
 * stand together) with its share of the batch in LDS and places the union,
 * regions chained by a decoupled look-back.  Needs size_power <= 30 (the
 * table's; otherwise err = INVALID and nothing else changes), len <
 * 2^size_power, buffers that do not overlap, and `scratch` of
 * hn_merge_scratch_bytes_*(size_power, len) bytes.  The merge runs on a
 * table with HN_INFO_ORDERED and no TOMBSTONEs; when the table is not such,
 * a region holds more than 2,560 old slots or 1,536 batch pairs (skewed
 * keys), the wrap past the table's end would move region 0's spill, or the
 * union would fill the table, new_buckets receives instead, on the device,
 * the old table with the batch inserted by probing: hn_insert_*'s result
 * (its live set, arrival order, err = SUCCESS or FAIL, HN_INFO_ORDERED
 * cleared; of a key repeated in the batch one of its values, as there).  Either way the result is in new_buckets and
 * *info describes it.  len == 0 copies the table.  `vals` NULL stores each
 * key's index in the batch; the _pairs variant reads hash_pair elements. */
HN_API size_t hn_merge_scratch_bytes_u32(int size_power, size_t len);
HN_API size_t hn_merge_scratch_bytes_u64(int size_power, size_t len);
HN_API int hn_merge_u32(const void* keys, const void* vals, const void* old_buckets, void* new_buckets, hn_info* info,
                        size_t len, int size_power, void* scratch, size_t scratch_bytes, hn_stream_t s);
HN_API int hn_merge_u64(const void* keys, const void* vals, const void* old_buckets, void* new_buckets, hn_info* info,
                        size_t len, int size_power, void* scratch, size_t scratch_bytes, hn_stream_t s);
HN_API int hn_merge_pairs_u32(const void* pairs, const void* old_buckets, void* new_buckets, hn_info* info, size_t len,
                              int size_power, void* scratch, size_t scratch_bytes, hn_stream_t s);
HN_API int hn_merge_pairs_u64(const void* pairs, const void* old_buckets, void* new_buckets, hn_info* info, size_t len,
                              int size_power, void* scratch, size_t scratch_bytes, hn_stream_t s);

/* Hashmap::device_rehash into a LARGER table (hashinator.h:284-336: extract
 * the valid elements, Info(newSizePower), insert them) as one bulk build
 * whose first partition pass reads old_buckets' live slots: no intermediate
 * array, no probing, and the new table comes out Robin Hood ordered
 * (HN_INFO_ORDERED), the image hn_rehash_* followed by hn_reorder_* gives.
 * Needs old_sp < new_sp <= 32 (else HN_EINVAL: hn_rehash_* handles equal and
 * smaller sizes), buffers that do not overlap, and `scratch` of
 * hn_rehash_build_scratch_bytes_*(old_sp, new_sp) bytes (about 16 (u64) or
 * 8 (u32) bytes per OLD slot, plus bookkeeping).  new_buckets need not be
 * initialised.  *info becomes Info(new_sp) plus the elements; err = SUCCESS
 * iff the fill is preserved.  Strongly skewed keys (a region over capacity)
 * are placed by the probing insert instead, as in hn_build_*. */
HN_API size_t hn_rehash_build_scratch_bytes_u32(int old_sp, int new_sp);
HN_API size_t hn_rehash_build_scratch_bytes_u64(int old_sp, int new_sp);
HN_API int hn_rehash_build_u32(const void* old_buckets, int old_sp, void* new_buckets, int new_sp, hn_info* info,
                               void* scratch, size_t scratch_bytes, hn_stream_t s);
HN_API int hn_rehash_build_u64(const void* old_buckets, int old_sp, void* new_buckets, int new_sp, hn_info* info,
                               void* scratch, size_t scratch_bytes, hn_stream_t s);

/* ---------------------------------------------------------------- compaction and scan */

/* Scratch for one compaction or scan of n elements
 * (estimateMemoryForCompaction, split_tools.h:1064-1089, sized for the
 * single-pass kernel instead). */
HN_API size_t hn_compact_scratch_bytes(size_t n);

/* split::tools::copy_if over hash_pair<K,K> (split_tools.h:1128-1223) with a
 * built-in predicate: stable, out[0..count) = selected elements in input
 * order; *d_count (device pointer) = count. */
HN_API int hn_copy_if_pairs_u32(const void* in, size_t n, const hn_pred* pred, void* out, uint64_t* d_count, void* scratch,
                         size_t scratch_bytes, hn_stream_t s);
HN_API int hn_copy_if_pairs_u64(const void* in, size_t n, const hn_pred* pred, void* out, uint64_t* d_count, void* scratch,
                         size_t scratch_bytes, hn_stream_t s);

/* split::tools::copy_keys_if (split_tools.h:1094-1111, 1146-1159,
 * 1176-1191): as above, writing only .first of each selected element. */
HN_API int hn_copy_keys_if_pairs_u32(const void* in, size_t n, const hn_pred* pred, uint32_t* out, uint64_t* d_count,
                              void* scratch, size_t scratch_bytes, hn_stream_t s);
HN_API int hn_copy_keys_if_pairs_u64(const void* in, size_t n, const hn_pred* pred, uint64_t* out, uint64_t* d_count,
                              void* scratch, size_t scratch_bytes, hn_stream_t s);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  hn_copy_if_pairs_u64 without the scratch argument: the
 * look-back state lives in a static device array of the library (4 MiB per
 * device, held until the process ends; no hipMalloc, hipFree or allocating
 * library call).  Accepts n <= HN_NOALLOC_MAX_ELEMS (else HN_ESCRATCH, before
 * any launch).  Every call on a device shares that device's array: calls that
 * overlap on one device (other streams, other threads) race and give wrong
 * results; order them.  The state is re-initialised by each call. */
#define HN_NOALLOC_MAX_ELEMS (1ull << 32)
HN_API int hn_copy_if_pairs_noalloc_u64(const void* in, size_t n, const hn_pred* pred, void* out, uint64_t* d_count,
                                        hn_stream_t s);

/* copy_if with the rule already evaluated into flags (one byte per element,
 * nonzero = keep), for callers that evaluate their rule outside this library
 * (an FFI caller, or a rule computed by other kernels).  C++ callers with a
 * rule functor instantiate the same kernel with the rule in their own
 * translation unit instead (<splitvector/split_compact.h>, used by
 * split::tools::copy_if).  elem_bytes must be 4, 8 or 16. */
HN_API int hn_copy_if_flags(const void* in, const uint8_t* flags, size_t n, size_t elem_bytes, void* out,
                     uint64_t* d_count, void* scratch, size_t scratch_bytes, hn_stream_t s);

/* copy_keys_if with the rule already evaluated into flags: .first of each
 * flagged hash_pair, in input order. */
HN_API int hn_copy_keys_if_flags_u32(const void* in_pairs, const uint8_t* flags, size_t n, uint32_t* out,
                                     uint64_t* d_count, void* scratch, size_t scratch_bytes, hn_stream_t s);
HN_API int hn_copy_keys_if_flags_u64(const void* in_pairs, const uint8_t* flags, size_t n, uint64_t* out,
                                     uint64_t* d_count, void* scratch, size_t scratch_bytes, hn_stream_t s);

/* As hn_copy_if_flags / hn_copy_keys_if_flags with the rule packed one bit
 * per element (bit i is bit i%32 of word i/32) -- n/8 bytes of staging
 * instead of n. */
HN_API int hn_copy_if_bits(const void* in, const uint32_t* bits, size_t n, size_t elem_bytes, void* out,
                           uint64_t* d_count, void* scratch, size_t scratch_bytes, hn_stream_t s);
HN_API int hn_copy_keys_if_bits_u32(const void* in_pairs, const uint32_t* bits, size_t n, uint32_t* out,
                                    uint64_t* d_count, void* scratch, size_t scratch_bytes, hn_stream_t s);
HN_API int hn_copy_keys_if_bits_u64(const void* in_pairs, const uint32_t* bits, size_t n, uint64_t* out,
                                    uint64_t* d_count, void* scratch, size_t scratch_bytes, hn_stream_t s);

/* split::tools::split_prefix_scan (split_tools.h:363-418): exclusive prefix
 * sum with wraparound in the element type.  Any n (the reference requires a
 * power of two).  in == out is allowed. */
HN_API int hn_exclusive_scan_u32(const uint32_t* in, uint32_t* out, size_t n, void* scratch, size_t scratch_bytes,
                          hn_stream_t s);
HN_API int hn_exclusive_scan_u64(const uint64_t* in, uint64_t* out, size_t n, void* scratch, size_t scratch_bytes,
                          hn_stream_t s);

/* ---------------------------------------------------------------- many maps, one launch */

/* No counterpart in the reference (its API is one launch per map per
 * operation).  A batch is an array of independent maps, each with its own
 * bucket array and Info, and slices of shared input/output arrays.  For every
 * map the requested phases run in the reference's order -- insert
 * (hashers.h:46-65), retrieve (111-119), erase (132-142), clean_tombstones
 * (hashinator.h:1261-1315) -- with the single-map semantics above; one
 * workgroup owns one map at a time for insert and for erase + clean, maps are
 * taken in array order (list the largest first); retrieve runs in between,
 * one key per lane over units of every map's queries.  Up to three persistent
 * launches plus small set-up kernels (graph-capturable: no host
 * synchronisation, no allocation, no memset). */
typedef struct hn_batch_map {
   void* buckets;        /* 2^size_power slots; size_power is read from *info */
   hn_info* info;
   uint64_t ins_off, ins_len;       /* slice of ins_keys / ins_vals */
   uint64_t qry_off, qry_len;       /* slice of qry_keys / qry_vals / qry_hit */
   uint64_t ers_off, ers_len;       /* slice of ers_keys */
   uint64_t scratch_off, scratch_cap; /* slice of clean_scratch, in slots; clean needs cap >= fill */
} hn_batch_map;

enum { HN_BATCH_INSERT = 1, HN_BATCH_RETRIEVE = 2, HN_BATCH_ERASE = 4, HN_BATCH_CLEAN = 8 };

/* Device scratch of hn_batch_run for a batch of n_maps maps: the map and
 * unit tickets and each map's first retrieve unit (8 bytes per map). */
HN_API size_t hn_batch_scratch_bytes(uint32_t n_maps);

/* Fresh tables ({EMPTYBUCKET, 0}) and Info(sps[m]) for all maps, one launch.
 * `maps` and `sps` are device arrays. */
HN_API int hn_batch_init_u32(const hn_batch_map* maps, uint32_t n_maps, const int32_t* sps, hn_stream_t s);
HN_API int hn_batch_init_u64(const hn_batch_map* maps, uint32_t n_maps, const int32_t* sps, hn_stream_t s);

/* Run the phases in `ops` (HN_BATCH_* bits) on every map.  `maps` is a device
 * array.  Unused input pointers may be NULL; qry_hit is optional.  A map
 * whose clean would need more than scratch_cap slots gets err = FAIL and is
 * left as it is.  tune: probe_width, block_size (threads per map), grid_blocks
 * (resident workgroups); NULL = defaults. */
HN_API int hn_batch_run_u32(const hn_batch_map* maps, uint32_t n_maps, uint32_t ops, const uint32_t* ins_keys,
                            const uint32_t* ins_vals, const uint32_t* qry_keys, uint32_t* qry_vals, uint8_t* qry_hit,
                            const uint32_t* ers_keys, void* clean_scratch, void* scratch, hn_stream_t s,
                            const hn_tune* t);
HN_API int hn_batch_run_u64(const hn_batch_map* maps, uint32_t n_maps, uint32_t ops, const uint64_t* ins_keys,
                            const uint64_t* ins_vals, const uint64_t* qry_keys, uint64_t* qry_vals, uint8_t* qry_hit,
                            const uint64_t* ers_keys, void* clean_scratch, void* scratch, hn_stream_t s,
                            const hn_tune* t);

/* ---------------------------------------------------------------- misc */

/* Version string and the gfx target the library was built for. */
HN_API const char* hn_version(void);

#ifdef __cplusplus
}
#endif
#endif
