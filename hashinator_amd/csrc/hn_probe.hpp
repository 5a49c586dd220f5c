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
         // A copy of the key anywhere in the window wins over the window's
         // EMPTY slots, as in insert_kernel (kernels_AMD.h:339-349 before the
         // claim loop at 353): a table with EMPTY holes inside a probe path
         // (left by reset_to_empty) gets no second copy from this window.
         uint64_t em = fm < G ? 0 : be;
         while (em) {
            const int j = first_bit(em);
            K old = kEmpty;
            bool whole = false; // the claim stored the value too
            if (gl == j) {
               if constexpr (sizeof(Slot<K>) == 8) {
                  // 32-bit keys: one 64-bit CAS of the whole slot, (EMPTY, value
                  // seen by this window's load) -> (key, value), so a claim is
                  // one random store, not two (random stores are the insert's
                  // bound); if only the value bits differ, the key-only CAS.
                  // The reference's pair of 32-bit fields is only 4-byte
                  // aligned: a bucket array that is not 8-byte aligned takes
                  // the key-only CAS throughout.
                  if ((reinterpret_cast<uintptr_t>(sl) & 7u) == 0) {
                     const uint64_t exp = (static_cast<uint64_t>(s.val) << 32) | static_cast<uint64_t>(kEmpty);
                     const uint64_t des = (static_cast<uint64_t>(val) << 32) | static_cast<uint64_t>(key);
                     const uint64_t o = atomic_cas(reinterpret_cast<uint64_t*>(sl), exp, des);
                     whole = o == exp;
                     old = static_cast<K>(o);
                     if (!whole && old == kEmpty) old = atomic_cas(&sl->key, kEmpty, key);
                  } else {
                     old = atomic_cas(&sl->key, kEmpty, key);
                  }
               } else {
                  old = atomic_cas(&sl->key, kEmpty, key);
               }
            }
            old = __shfl(old, gbase + j, kWave);
            if (old == kEmpty) {
               if (gl == j) {
                  if (!whole) sl->val = val;
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

// ------------------------------------------------------------ lane windows
//
// Second lookup strategy: one key per LANE.  The probe loop above spends a
// whole group of G lanes on each key, so every per-key step (hash, window
// address, ballots, the decision) costs G lanes of issue; the loop is
// VALU-bound.  Here each lane owns its key for the whole probe.  The reads
// stay 128-byte coalesced: the 8 lanes of a group load the current window of
// each of the group's 8 keys in turn (8 loads of 16 bytes per lane; lane j
// reads chunk j of every window), stage them in LDS, and read back their own
// window: lane j gets the 8 chunks of key j.  The scan then runs in the lane
// with no cross-lane traffic.  A lane that finishes takes the next query of
// its wave's range at once, so lanes never wait for slow probes of others.
//
// Staging layout (per wave): chunk p of key k of group g at byte
// g*kLaneGroupBytes + k*kLaneRowBytes + p*16.  A row is padded to 144 bytes:
// the 8 lanes of a group write 128 contiguous bytes, and reading their rows
// they start 144 bytes apart, on 8 distinct bank groups; every address is a
// lane base plus a constant.
//
// The Robin Hood stop is tested on the window's LAST slot only: in an
// ordered table the stop predicate is monotone along a cluster (homes never
// decrease), so a window holds a stopping slot iff its last occupant stops
// (an EMPTY ends the probe anyway; a TOMBSTONE there only delays the stop by
// one window).  Results are identical, since a present key is never behind a
// stopping slot.  One hash per window instead of one per slot.
template <typename K>
struct LaneWin {
   static constexpr int kSlotBytes = static_cast<int>(sizeof(Slot<K>)); // 16 or 8
   static constexpr int kSlots = 128 / kSlotBytes;                       // 8 or 16 per window
   static constexpr int kLog = kSlots == 8 ? 3 : 4;
   static constexpr int kPerChunk = 16 / kSlotBytes; // slots per 16-byte chunk
};
constexpr int kLaneGroup = 8;                             // lanes that load one window together
constexpr int kLaneRowBytes = 144;                        // one staged window, padded
constexpr int kLaneGroupBytes = kLaneRowBytes * kLaneGroup; // 1152
constexpr int kLaneWaveBytes = kLaneGroupBytes * (kWave / kLaneGroup); // 9 KiB of staging per wave

__device__ __forceinline__ void wave_lds_sync() {
   __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
   __builtin_amdgcn_wave_barrier();
   __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Results of a wave's queries are written back in query order through a
// ring of kLaneRing entries in LDS: a finishing lane parks its result at ring
// entry q mod kLaneRing, and whenever the oldest 64 entries are complete the
// wave writes them with one coalesced store per array.  (Stored straight from
// the finishing lanes, the results of one store instruction scatter over a
// few hundred queries, and those partial-line writes cost a third of the
// kernel's time.)  A query is only handed out while its ring entry is free,
// so a long probe holds back at most kLaneRing queries of its wave.
constexpr int kLaneRing = 512;
constexpr uint8_t kRingPending = 2;

// Queries [q_next, q_end) by one wave (every lane must call).  `win`: this
// wave's kLaneWaveBytes of staging (16-byte aligned); `wbs`: 64 words; `ring_v`/`ring_s`:
// kLaneRing result values and states (retrieve only).  Requires
// sp >= LaneWin<K>::kLog (the table holds at least one window) and ovf >= 1.
// Returns this lane's erase count.
template <typename K, bool PAIRS, int MODE>
__device__ __forceinline__ uint64_t lane_lookup(const LookupArgs<K>& a, const int sp, const uint64_t ovf,
                                               const bool ordered, uint64_t q_next, const uint64_t q_end, char* win,
                                               uint32_t* wbs, K* ring_v, uint8_t* ring_s) {
   static_assert(MODE == LK_RETRIEVE || MODE == LK_ERASE, "lane windows serve retrieve and erase");
   constexpr bool RING = MODE == LK_RETRIEVE;
   using LW = LaneWin<K>;
   constexpr int WS = LW::kSlots;
   constexpr K kEmpty = empty_key<K>();
   constexpr K kTomb = tomb_key<K>();
   constexpr uint64_t kRingMask = kLaneRing - 1;
   const int lane = __lane_id();
   const int gl = lane & (kLaneGroup - 1);
   const int g = lane >> 3;
   const uint32_t mask = slot_mask32(sp);
   const uint64_t nsl = static_cast<uint64_t>(mask) + 1;
   const uint64_t bound = ovf < nsl ? ovf : nsl;

   // staging: I write chunk gl of the 8 windows of my group, and read my own window
   char* const wr = win + g * kLaneGroupBytes + gl * 16;
   const char* const rd = win + g * kLaneGroupBytes + gl * kLaneRowBytes;
   const char* const tbl = reinterpret_cast<const char*>(a.buckets) + gl * 16;

   K key = kEmpty;
   uint64_t q = 0;
   uint32_t base = 0, d0 = 0, kw = 0, nwin = 1, last_hi = WS;
   bool active = false;
   auto load_key = [&](uint64_t qq) -> K {
      if constexpr (PAIRS) {
         return a.pairs[qq].key;
      } else {
         return a.keys[qq];
      }
   };
   auto start = [&](uint64_t qq, K kk) {
      q = qq;
      key = kk;
      const uint32_t home = fib_home32(key, sp) & mask;
      base = home & ~static_cast<uint32_t>(WS - 1);
      d0 = home & static_cast<uint32_t>(WS - 1);
      kw = 0;
      // offsets [0, bound) are slots [d0, d0 + bound) counted from base
      const uint64_t span = static_cast<uint64_t>(d0) + bound;
      nwin = static_cast<uint32_t>((span + WS - 1) >> LW::kLog);
      last_hi = static_cast<uint32_t>(span - (static_cast<uint64_t>(nwin - 1) << LW::kLog));
      active = true;
      if constexpr (RING) ring_s[qq & kRingMask] = kRingPending;
   };
   uint64_t lo = q_next; // results of queries below lo are written out
   {
      const uint64_t qq = q_next + static_cast<uint64_t>(lane);
      if (qq < q_end) start(qq, load_key(qq));
      q_next = q_next + kWave < q_end ? q_next + kWave : q_end;
   }
   // Keys of queries q_next + lane, loaded one step ahead so a lane that
   // starts a query does not wait for memory.
   K pre = kEmpty;
   {
      const uint64_t qq = q_next + static_cast<uint64_t>(lane);
      if (qq < q_end) pre = load_key(qq);
   }

   // Write out complete 64-entry runs of results, oldest first.  A wave's
   // later loads wait for its earlier stores to complete, so the runs go out
   // in batches of kFlushRuns (or when the ring is nearly full, or at the end
   // of the range) rather than one at a time: one such wait per batch.
   constexpr int kFlushRuns = 4;
   auto flush = [&]() {
      if constexpr (RING) {
         wave_lds_sync();
         int ready = 0;
         uint64_t l = lo;
         for (int f = 0; f < kLaneRing / kWave; ++f) {
            const uint64_t run = q_end - l < static_cast<uint64_t>(kWave) ? q_end - l : kWave;
            if (run == 0) break;
            const uint64_t e = l + static_cast<uint64_t>(lane);
            const bool need = static_cast<uint64_t>(lane) < run;
            const bool pend = need && (e >= q_next || ring_s[e & kRingMask] == kRingPending);
            if (__ballot(pend)) break;
            ++ready;
            l += run;
         }
         if (ready > 0 && (ready >= kFlushRuns || q_next + kWave > lo + kLaneRing || l >= q_end)) {
            for (int f = 0; f < ready; ++f) {
               const uint64_t run = q_end - lo < static_cast<uint64_t>(kWave) ? q_end - lo : kWave;
               const uint64_t e = lo + static_cast<uint64_t>(lane);
               if (static_cast<uint64_t>(lane) < run) {
                  const uint8_t st = ring_s[e & kRingMask];
                  if (st == 1) {
                     const K v = ring_v[e & kRingMask];
                     if constexpr (PAIRS) {
                        a.pairs[e].val = v;
                     } else {
                        a.vals[e] = v;
                     }
                  }
                  if (a.hit) a.hit[e] = st;
               }
               lo += run;
            }
         }
         wave_lds_sync();
      }
   };

   uint64_t erased = 0;
   K cas_old = 0, cas_key = 0; // erase: last CAS of this lane, not yet counted
   bool cas_pending = false;
   for (;;) {
      const bool any = __ballot(active) != 0;
      if (!any && q_next >= q_end && (!RING || lo >= q_end)) break;
      if (!any) {
         flush();
      } else {
         const uint32_t wb = (base + (kw << LW::kLog)) & mask; // first slot of my window
         wbs[lane] = wb;
         wave_lds_sync();
         const uint4 w0 = *reinterpret_cast<const uint4*>(wbs + g * 8);
         const uint4 w1 = *reinterpret_cast<const uint4*>(wbs + g * 8 + 4);
         const uint32_t bw[8] = {w0.x, w0.y, w0.z, w0.w, w1.x, w1.y, w1.z, w1.w};
         auto ld = [&](int t) {
            return *reinterpret_cast<const uint4*>(tbl + static_cast<uint64_t>(bw[t]) * LW::kSlotBytes);
         };
         const uint4 c0 = ld(0), c1 = ld(1), c2 = ld(2), c3 = ld(3), c4 = ld(4), c5 = ld(5), c6 = ld(6), c7 = ld(7);
         if constexpr (MODE == LK_ERASE) {
            if (cas_pending) erased += cas_old == cas_key ? 1 : 0;
            cas_pending = false;
         }
         auto st = [&](int t, const uint4& v) { *reinterpret_cast<uint4*>(wr + t * kLaneRowBytes) = v; };
         st(0, c0), st(1, c1), st(2, c2), st(3, c3), st(4, c4), st(5, c5), st(6, c6), st(7, c7);
         // Result stores go out here, a whole scan ahead of the next window
         // loads: the wait for those loads' data also waits for every store
         // issued before them.
         flush();
         wave_lds_sync();

         // scan my window: match / EMPTY bits in slot order
         uint32_t mm = 0, em = 0;
#pragma unroll
         for (int t = 0; t < 8; ++t) {
            if constexpr (LW::kPerChunk == 1) {
               const K k0 = *reinterpret_cast<const K*>(rd + t * 16);
               mm |= static_cast<uint32_t>(k0 == key) << t;
               em |= static_cast<uint32_t>(k0 == kEmpty) << t;
            } else {
               const uint32_t* pc = reinterpret_cast<const uint32_t*>(rd + t * 16);
               const K k0 = static_cast<K>(pc[0]), k1 = static_cast<K>(pc[2]);
               mm |= (static_cast<uint32_t>(k0 == key) | (static_cast<uint32_t>(k1 == key) << 1)) << (2 * t);
               em |= (static_cast<uint32_t>(k0 == kEmpty) | (static_cast<uint32_t>(k1 == kEmpty) << 1)) << (2 * t);
            }
         }
         const uint32_t lo_s = kw == 0 ? d0 : 0u;
         const uint32_t hi_s = kw + 1 == nwin ? last_hi : static_cast<uint32_t>(WS);
         const uint32_t vm = ((1u << hi_s) - 1u) & ~((1u << lo_s) - 1u);
         mm &= vm;
         em &= vm;
         const uint32_t hm = mm & (em ? ((em & (0u - em)) - 1u) : ~0u); // matches before the first EMPTY
         bool fin = hm != 0 || em != 0 || kw + 1 >= nwin;
         if (ordered && !fin) {
            K kl;
            if constexpr (LW::kPerChunk == 1) {
               kl = *reinterpret_cast<const K*>(rd + 7 * 16);
            } else {
               kl = static_cast<K>(reinterpret_cast<const uint32_t*>(rd + 7 * 16)[2]);
            }
            if (kl != kTomb) {
               const uint32_t slot_l = wb + (WS - 1);
               const uint32_t d_occ = (slot_l - fib_home32(kl, sp)) & mask;
               const uint32_t off_l = (kw << LW::kLog) + (WS - 1) - d0;
               fin = d_occ < off_l;
            }
         }
         fin = fin && active;
         if (fin) {
            const bool hit = hm != 0;
            if (hit) {
               const uint32_t sl = static_cast<uint32_t>(__ffs(hm) - 1);
               if constexpr (MODE == LK_ERASE) {
                  // the CAS result is counted a step later, so the wave does
                  // not wait for the atomic here
                  cas_old = atomic_cas(&a.buckets[wb + sl].key, key, kTomb);
                  cas_key = key;
                  cas_pending = true;
               } else {
                  // the value sits next to the key in the staged window
                  const K v = *reinterpret_cast<const K*>(rd + sl * LW::kSlotBytes + sizeof(K));
                  ring_v[q & kRingMask] = v;
               }
            }
            if constexpr (RING) ring_s[q & kRingMask] = hit ? 1 : 0;
            active = false;
         } else if (active) {
            ++kw;
         }
      }
      wave_lds_sync();
      // idle lanes take the next queries of the wave's range, in lane order
      const uint64_t im = __ballot(!active);
      const uint64_t lim_r = RING ? lo + kLaneRing : q_end;
      const uint64_t lim = lim_r < q_end ? lim_r : q_end;
      if (im && q_next < lim) {
         const uint32_t rank =
             __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(im >> 32), __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(im), 0));
         const uint64_t room = lim - q_next;
         const uint32_t idle = static_cast<uint32_t>(__popcll(im));
         const uint32_t cnt = room < idle ? static_cast<uint32_t>(room) : idle;
         const K nk = __shfl(pre, static_cast<int>(rank & (kWave - 1)), kWave);
         const K shifted = __shfl(pre, static_cast<int>((static_cast<uint32_t>(lane) + cnt) & (kWave - 1)), kWave);
         if (!active && rank < cnt) start(q_next + rank, nk);
         q_next += cnt;
         if (static_cast<uint32_t>(lane) + cnt < static_cast<uint32_t>(kWave)) {
            pre = shifted;
         } else {
            const uint64_t pq = q_next + static_cast<uint64_t>(lane);
            if (pq < q_end) pre = load_key(pq);
         }
      }
      wave_lds_sync(); // the next window's staging overwrites this one
   }
   if (cas_pending) erased += cas_old == cas_key ? 1 : 0;
   return erased;
}

} // namespace hn
