// hn_probe.hpp -- the probe loops shared by the single-map kernels
// (hn_map.hip) and the batched multi-map kernel (hn_batch.hip).
//
// Probing scheme.  A "probe group" of G lanes (G = 1..64, a power of two,
// several groups per wave64) owns one key at a time.  Step k of the probe
// reads one aligned window of W = min(G, 2^sp) consecutive slots: window k of
// a key with home slot h covers slots base + k*W + [0, W), base = h & ~(W-1),
// so every step is one aligned 64/128/.. byte read.  A lane's probe offset is
// k*W + lane - (h - base); lanes whose offset is negative (before h in the
// first window) or beyond the bound take no part.  Two wave ballots give the
// group's match and EMPTY masks; the first set bit in probe order decides,
// exactly as a serial linear probe would (hashinator.h:742-768 for lookups).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "hn_device.hpp"

namespace hn {

enum SrcKind { SRC_SOA = 0, SRC_INDEX = 1, SRC_PAIRS = 2, SRC_TABLE = 3 };

template <typename K>
struct InsertArgs {
   const K* keys;
   const K* vals;
   const Slot<K>* pairs;
   Slot<K>* buckets;
   hn_info* info;
   uint64_t len;
   const uint64_t* d_len;         // if set, the input length is read on the device
   const uint64_t* expected_fill; // if set, err = FAIL unless fill ends equal to it
   int32_t* fail_sink;            // if set, failures are OR-ed here and err is left alone
};

// Work distribution.  Inputs are consumed in blocks of G consecutive
// entries: lane gl of a group loads entry blk*G + gl with one coalesced read,
// the group then resolves the block's entries one after another (the current
// key is broadcast from its lane), and lookups keep each result in the lane
// that owns the entry, so the block's outputs leave in one coalesced write.
// Blocks are handed out grid-stride per group.

struct InsertCounters {
   uint64_t placed = 0;    // new keys placed by this lane
   uint64_t max_probe = 0; // longest probe (offset + 1) of those placements
   int failed = 0;         // a key of this lane's group found no EMPTY slot
};

// insert_kernel / insert_index_kernel semantics (kernels_AMD.h:136-405, 507-641)
// over input entries [0, len), starting at block `blk` and stepping `gstride`
// blocks.  Must be called by whole waves (every lane of the wave).
template <typename K, int G, int SRC>
__device__ __forceinline__ void insert_range(const InsertArgs<K>& a, const int sp, const uint64_t len, uint64_t blk,
                                             const uint64_t gstride, InsertCounters& cnt) {
   constexpr K kEmpty = empty_key<K>();
   constexpr K kTomb = tomb_key<K>();
   const int lane = __lane_id();
   const int gl = lane & (G - 1);
   const int gbase = lane & ~(G - 1);
   const uint64_t nsl = 1ull << sp;
   const uint64_t mask = nsl - 1;
   const uint64_t W = nsl < static_cast<uint64_t>(G) ? nsl : static_cast<uint64_t>(G);

   uint64_t placed = 0, max_probe = 0;
   int failed = 0;
   K my_key = kEmpty, my_val = 0; // this lane's entry of the current block
   K key = kEmpty, val = 0;       // the entry being inserted by the group
   uint64_t pending = 0, base = 0, kw = 0;
   int64_t d0 = 0;
   bool active = false;

   auto load_block = [&]() {
      active = false;
      while (blk * G < len) {
         const uint64_t q = blk * G + static_cast<uint64_t>(gl);
         bool ok = q < len;
         if (ok) {
            if constexpr (SRC == SRC_SOA) {
               my_key = a.keys[q];
               my_val = a.vals[q];
            } else if constexpr (SRC == SRC_INDEX) {
               my_key = a.keys[q];
               my_val = static_cast<K>(q);
            } else {
               const Slot<K> p = a.pairs[q];
               my_key = p.key;
               my_val = p.val;
            }
            if constexpr (SRC == SRC_TABLE) ok = my_key != kEmpty && my_key != kTomb;
         }
         pending = group_bits<G>(__ballot(ok), gbase);
         if (pending) {
            active = true;
            return;
         }
         blk += gstride;
      }
   };
   auto start_entry = [&]() {
      const int j = first_bit(pending);
      key = __shfl(my_key, gbase + j, kWave);
      val = __shfl(my_val, gbase + j, kWave);
      const uint64_t home = fib_home(key, sp) & mask;
      base = home & ~(W - 1);
      d0 = static_cast<int64_t>(home - base);
      kw = 0;
   };

   load_block();
   if (active) start_entry();
   while (__ballot(active)) {
      if (active) {
         const uint64_t pos = kw * W + static_cast<uint64_t>(gl);
         const int64_t off = static_cast<int64_t>(pos) - d0;
         const bool valid = static_cast<uint64_t>(gl) < W && off >= 0 && static_cast<uint64_t>(off) < nsl;
         Slot<K>* sl = a.buckets + ((base + pos) & mask);
         const Slot<K> s = *sl;
         const uint64_t bm = group_bits<G>(__ballot(valid && s.key == key), gbase);
         const uint64_t be = group_bits<G>(__ballot(valid && s.key == kEmpty), gbase);
         const int fm = bm ? first_bit(bm) : G;
         bool done = false;
         // Only EMPTY slots ahead of an existing copy of the key may be claimed.
         uint64_t em = fm >= 64 ? be : (be & ((1ull << fm) - 1ull));
         while (em) {
            const int j = first_bit(em);
            K old = kEmpty;
            if (gl == j) old = atomic_cas(&sl->key, kEmpty, key);
            old = __shfl(old, gbase + j, kWave);
            if (old == kEmpty) {
               if (gl == j) {
                  sl->val = val;
                  ++placed;
                  const uint64_t pr = static_cast<uint64_t>(off) + 1;
                  max_probe = pr > max_probe ? pr : max_probe;
               }
               done = true;
               break;
            }
            if (old == key) { // a copy of this key was placed concurrently
               if (gl == j) sl->val = val;
               done = true;
               break;
            }
            em &= em - 1;
         }
         if (!done && fm < G) {
            if (gl == fm) sl->val = val;
            done = true;
         }
         if (!done) {
            const uint64_t covered = (kw + 1) * W - static_cast<uint64_t>(d0); // offsets [0, covered) seen
            if (covered >= nsl) {
               failed = 1;
               done = true;
            } else {
               ++kw;
            }
         }
         if (done) {
            pending &= pending - 1;
            if (!pending) {
               blk += gstride;
               load_block();
            }
            if (active) start_entry();
         }
      }
   }

   cnt.placed += placed;
   cnt.max_probe = max_probe > cnt.max_probe ? max_probe : cnt.max_probe;
   cnt.failed |= failed;
}

template <typename K>
struct LookupArgs {
   const K* keys;   // SoA keys, or null for pairs
   K* vals;         // SoA values, or null for pairs
   Slot<K>* pairs;  // pairs variant: key in, value out
   uint8_t* hit;    // optional
   Slot<K>* buckets;
   hn_info* info;
   uint64_t len;
};

enum LookupMode { LK_RETRIEVE = 0, LK_ERASE = 1, LK_RESET = 2 };

// LK_RETRIEVE: retrieve_kernel (kernels_AMD.h:647-752).
// LK_ERASE:    delete_kernel (kernels_AMD.h:412-505), with CAS key -> TOMBSTONE.
// LK_RESET:    reset_to_empty (kernels_AMD.h:65-134): the first slot holding
//              the key (whole table, EMPTY does not stop the probe) -> EMPTY.
// Entries [0, len), starting at block `blk`, stepping `gstride` blocks; `ovf`
// is the map's currentMaxBucketOverflow.  Returns this lane's erase count.
// Must be called by whole waves.
//
// `ordered`: the table is in Robin Hood order (every cluster sorted by home,
// see hn_reorder_*).  Then an occupant whose displacement is smaller than the
// probe's offset has its home after the probed key's, so the key cannot lie
// further on and the probe stops as at an EMPTY slot.  TOMBSTONEs never stop.
//
// Index arithmetic is 32-bit where it can be (slot indices, masks, ballot
// slices, the occupant's home): the loop is issue-bound, not memory-bound.
template <int G>
using GroupMask = std::conditional_t<(G <= 32), uint32_t, uint64_t>;
template <int G>
__device__ __forceinline__ GroupMask<G> group_slice(uint64_t ballot, int gbase) {
   if constexpr (G <= 32)
      return group_bits32<G>(ballot, gbase);
   else
      return ballot;
}
template <typename M>
__device__ __forceinline__ int first_set(M m) {
   if constexpr (sizeof(M) == 4)
      return __ffs(m) - 1;
   else
      return __ffsll(static_cast<unsigned long long>(m)) - 1;
}

template <typename K, int G, bool PAIRS, int MODE>
__device__ __forceinline__ uint64_t lookup_range(const LookupArgs<K>& a, const int sp, const uint64_t ovf_in,
                                                 const uint64_t len, uint64_t blk, const uint64_t gstride,
                                                 const bool ordered_in = false) {
   using M = GroupMask<G>;
   constexpr bool ERASE = MODE == LK_ERASE;
   constexpr bool RESET = MODE == LK_RESET;
   constexpr K kEmpty = empty_key<K>();
   constexpr K kTomb = tomb_key<K>();
   constexpr int GLOG = G == 1 ? 0 : G == 2 ? 1 : G == 4 ? 2 : G == 8 ? 3 : G == 16 ? 4 : G == 32 ? 5 : 6;
   const int lane = __lane_id();
   const int gl = lane & (G - 1);
   const int gbase = lane & ~(G - 1);
   const uint32_t mask = slot_mask32(sp);
   const uint64_t nsl = static_cast<uint64_t>(mask) + 1;
   const int wlog = GLOG < sp ? GLOG : sp; // window of W = min(G, 2^sp) slots
   const uint32_t W = 1u << wlog;
   // Offsets past 2^sp only revisit slots already seen, so the bound is clamped there.
   const uint64_t ovf = RESET ? nsl : ovf_in;
   const uint64_t bound = ovf < nsl ? ovf : nsl;
   const bool ordered = ordered_in && !RESET;

   uint64_t erased = 0;
   K my_key = kEmpty, res_val = 0;
   bool res_hit = false;
   K key = kEmpty;
   M pending = 0;
   uint32_t base = 0;
   uint64_t kw = 0;
   uint32_t d0 = 0;
   bool active = false;

   auto load_block = [&]() {
      active = false;
      if (blk * G >= len) return;
      const uint64_t q = blk * G + static_cast<uint64_t>(gl);
      const bool ok = q < len;
      if (ok) {
         if constexpr (PAIRS) {
            my_key = a.pairs[q].key;
         } else {
            my_key = a.keys[q];
         }
      }
      res_hit = false;
      pending = group_slice<G>(__ballot(ok), gbase);
      active = pending != 0;
   };
   auto flush_block = [&]() {
      if constexpr (MODE == LK_RETRIEVE) {
         const uint64_t q = blk * G + static_cast<uint64_t>(gl);
         if (q < len) {
            if (res_hit) {
               if constexpr (PAIRS) {
                  a.pairs[q].val = res_val;
               } else {
                  a.vals[q] = res_val;
               }
            }
            if (a.hit) a.hit[q] = res_hit ? 1 : 0;
         }
      }
   };
   auto start_entry = [&]() {
      key = __shfl(my_key, gbase + first_set(pending), kWave);
      const uint32_t home = fib_home32(key, sp) & mask;
      base = home & ~(W - 1u);
      d0 = home - base;
      kw = 0;
   };

   load_block();
   if (active) start_entry();
   while (__ballot(active)) {
      if (active) {
         const uint64_t pos = (kw << wlog) + static_cast<uint64_t>(gl);
         const bool ahead = pos >= d0; // at or after the home slot
         const uint64_t off = pos - d0;
         const bool valid = static_cast<uint32_t>(gl) < W && ahead && off < bound;
         const uint32_t slot = (base + static_cast<uint32_t>(pos)) & mask;
         Slot<K>* sl = a.buckets + slot;
         const Slot<K> s = *sl;
         const M bm = group_slice<G>(__ballot(valid && s.key == key), gbase);
         bool stop = valid && s.key == kEmpty;
         if (ordered && valid && !stop && s.key != key && s.key != kTomb) {
            const uint32_t d_occ = (slot - fib_home32(s.key, sp)) & mask;
            stop = d_occ < off;
         }
         const M be = RESET ? M(0) : group_slice<G>(__ballot(stop), gbase);
         const int fm = bm ? first_set(bm) : G;
         const int fe = be ? first_set(be) : G;
         bool done = true;
         if (fm < fe) {
            if constexpr (ERASE) {
               if (gl == fm && atomic_cas(&sl->key, key, kTomb) == key) ++erased;
            } else if constexpr (RESET) {
               if (gl == fm) atomic_cas(&sl->key, key, kEmpty);
            } else {
               const K v = __shfl(s.val, gbase + fm, kWave);
               if (gl == first_set(pending)) {
                  res_val = v;
                  res_hit = true;
               }
            }
         } else {
            const uint64_t covered = ((kw + 1) << wlog) - d0;
            if (!be && covered < bound) {
               ++kw;
               done = false;
            }
         }
         if (done) {
            pending &= pending - 1;
            if (!pending) {
               flush_block();
               blk += gstride;
               load_block();
            }
            if (active) start_entry();
         }
      }
   }

   return erased;
}

} // namespace hn
