/*
 * hashinator_amd_calib.h -- measurement aids shipped with the library.
 *
 * No counterpart in the reference: these exist so that the random-access
 * HBM ceiling the hash-map kernels are held against can be measured on the
 * same device, in the same process, with the same launch machinery.
 */
#ifndef HASHINATOR_AMD_CALIB_H
#define HASHINATOR_AMD_CALIB_H

#include "hashinator_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Random gather: n_access reads of `gran` bytes (16, 32, 64, 128 or 256) at
 * uniformly random gran-aligned offsets of buf[0, bytes), each read issued by
 * gran/16 lanes as one coalesced request.  Every lane group keeps 4
 * independent address streams in flight; `chain` > 1 makes each stream issue
 * `chain` dependent reads (the next offset depends on the data read), which
 * models a probe sequence.  The reads are folded into *sink (device). */
HN_API int hn_calib_random_gather(const void* buf, size_t bytes, int gran, size_t n_access, int chain, uint64_t seed,
                                  uint64_t* sink, hn_stream_t s);

/* Per-lane random gather: n_access reads of `gran` bytes (16, 32, 64 or 128)
 * at uniformly random gran-aligned offsets of buf[0, bytes), each read by ONE
 * lane as gran/16 consecutive 16-byte loads; bytes/gran must be a power of
 * two.  Every lane keeps 2 chunks in flight. */
HN_API int hn_calib_random_gather_lane(const void* buf, size_t bytes, int gran, size_t n_access, uint64_t seed,
                                       uint64_t* sink, hn_stream_t s);

/* Streaming copy of `bytes` (a multiple of 16) from src to dst, 16-byte
 * vectors, grid-stride: the read + write rate a streaming kernel reaches on
 * this device, the ceiling compaction and scan are held against. */
HN_API int hn_calib_stream_copy(const void* src, void* dst, size_t bytes, hn_stream_t s);

#ifdef __cplusplus
}
#endif
#endif
