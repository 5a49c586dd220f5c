/* split_compact.h -- single-pass stable stream compaction for gfx950, as a
 * template over the element type, the selector and the projection.
 *
 * One kernel over tiles of 8,192 elements (1,024 threads, 8 each) with
 * decoupled look-back: a tile takes a ticket (so every tile it waits on has
 * already started), counts its selected elements with wave ballots,
 * publishes that aggregate, folds its predecessors' published values 64 at
 * a time with one wave until it meets an inclusive prefix, publishes its
 * own inclusive prefix and writes its selected elements at their global
 * ranks.  The input is read once and the output written once, where the
 * reference runs a count pass, a Blelloch scan and a scatter pass
 * (split_tools.h:214-482, 920-989).  Ranks follow input order (element
 * j*1024 + t of a tile is item j of thread t), so the compaction is stable.
 *
 * libhashinator_amd instantiates it for its built-in predicates
 * (hn_copy_if_*, hn_copy_keys_if_*); split::tools::copy_if instantiates it in
 * the caller's translation unit with the caller's rule, so a user rule costs
 * no extra pass over the input.
 *
 * A selector receives a register copy of each element; through a writable
 * input pointer, an element the selector changed is stored back (the rule of
 * the reference's clean_tombstones, hashinator.h:1285-1297, turns tombstones
 * into empty buckets this way).
 *
 * Scratch: scratch_bytes(n) bytes of device memory, 256-byte aligned, whose
 * contents need not be initialised (launch() clears them on the stream).
 * Calls that share scratch must not overlap.
 */
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

namespace split {
namespace tools {
namespace compact {

// Large tiles because every tile pays a serial cost, its ticket and its step
// in the look-back chain: at 2^28 elements, 256-thread tiles ran copy_if in
// 2.0 ms; 1,024 threads with 8 elements each take 1.23 ms (4 or 16 elements
// per thread were slower).
constexpr int kThreads = 1024;
constexpr int kWaveSize = 64;
constexpr int kWaves = kThreads / kWaveSize;
constexpr int kPerThread = 8;
constexpr uint64_t kTile = static_cast<uint64_t>(kThreads) * kPerThread;

// A tile's look-back word: state in the top two bits (0 not yet published,
// 1 aggregate, 2 inclusive prefix), the element count below.  Words are
// written and read with relaxed device-scope atomics: a word carries its
// state and value together, so no release/acquire fence is needed (on a
// multi-XCD device those write back or invalidate the whole L2).
constexpr uint64_t kStateAgg = 1, kStateInc = 2;
constexpr uint64_t kValueMask = (1ull << 62) - 1;

struct TileState {
   uint32_t* ticket; // next tile to start
   uint64_t* word;   // [tiles]
};

constexpr uint64_t tiles_for(uint64_t n) { return (n + kTile - 1) / kTile; }
constexpr size_t scratch_bytes(uint64_t n) { return 256 + static_cast<size_t>(tiles_for(n)) * 8; }
inline TileState state_at(void* scratch) {
   TileState st;
   st.ticket = static_cast<uint32_t*>(scratch);
   st.word = reinterpret_cast<uint64_t*>(static_cast<char*>(scratch) + 256);
   return st;
}

__device__ __forceinline__ uint64_t first_set(uint64_t m) { return static_cast<uint64_t>(__ffsll(static_cast<unsigned long long>(m)) - 1); }

__device__ __forceinline__ uint64_t wave_add(uint64_t v) {
   for (int o = kWaveSize / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWaveSize);
   return v;
}

// Inclusive sum over the wave: DPP row shifts inside 16-lane rows, then row
// broadcasts across rows.
__device__ __forceinline__ uint32_t wave_scan_incl(uint32_t x) {
   int v = static_cast<int>(x);
   v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false); // row_shr:1
   v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false); // row_shr:2
   v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false); // row_shr:4
   v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false); // row_shr:8
   v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false); // row_bcast:15 -> rows 1, 3
   v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false); // row_bcast:31 -> rows 2, 3
   return static_cast<uint32_t>(v);
}

__device__ __forceinline__ void publish(const TileState& st, uint64_t tile, uint64_t state, uint64_t v) {
   __hip_atomic_store(&st.word[tile], (state << 62) | v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Exclusive prefix of tile `tile`, whose own count is `mine`; called by one
// full wave.  Publishes the tile's aggregate, then its inclusive prefix.
__device__ inline uint64_t tile_prefix(const TileState& st, uint32_t tile, uint64_t mine) {
   const int lane = __lane_id();
   if (tile == 0) {
      if (lane == 0) publish(st, 0, kStateInc, mine);
      return 0;
   }
   if (lane == 0) publish(st, tile, kStateAgg, mine);
   uint64_t prefix = 0;
   int64_t hi = static_cast<int64_t>(tile) - 1; // newest predecessor not yet folded in
   for (;;) {
      const int64_t p = hi - lane;
      uint64_t w = kStateInc << 62; // lanes before tile 0 read as an inclusive zero
      if (p >= 0) w = __hip_atomic_load(&st.word[p], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const uint64_t state = w >> 62;
      const uint64_t pending = __ballot(state == 0);
      const uint64_t inc = __ballot(state == kStateInc);
      // Lanes up to and including the first inclusive one, if none before it is still pending.
      const uint64_t first = inc ? first_set(inc) : 64;
      const uint64_t upto = first >= 63 ? ~0ull : ((2ull << first) - 1ull);
      if (pending & upto) {
         __builtin_amdgcn_s_sleep(1);
         continue;
      }
      prefix += wave_add(static_cast<uint64_t>(lane) <= first ? (w & kValueMask) : 0);
      if (first < 64) break;
      hi -= 64;
   }
   if (lane == 0) publish(st, tile, kStateInc, prefix + mine);
   return prefix;
}

// Whether two objects hold the same bytes (4-byte words where the size allows).
template <typename T>
__device__ __forceinline__ bool same_bytes(const T& a, const T& b) {
   constexpr int W = sizeof(T) % 4 == 0 ? 4 : 1;
   using Word = typename std::conditional<W == 4, uint32_t, unsigned char>::type;
   Word wa[sizeof(T) / W], wb[sizeof(T) / W];
   __builtin_memcpy(wa, &a, sizeof(T));
   __builtin_memcpy(wb, &b, sizeof(T));
   bool same = true;
#pragma unroll
   for (int k = 0; k < static_cast<int>(sizeof(T) / W); ++k) same = same && wa[k] == wb[k];
   return same;
}

// Clears the look-back words and the ticket; with no tiles, stores the count 0.
template <typename Count>
__global__ __launch_bounds__(256) void k_prepare(TileState st, uint64_t tiles, Count* count) {
   const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
   for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < tiles; i += stride)
      st.word[i] = 0;
   if (blockIdx.x == 0 && threadIdx.x == 0) {
      *st.ticket = 0;
      if (tiles == 0 && count) *count = 0;
   }
}

// out[rank] = proj(x) for every x = in[i] with sel(i, x), in input order;
// *count = the number written.  sel receives a register copy of the element
// and proj the copy as sel left it.  With a writable input (In not const), a
// copy that sel changed is stored back to in[i], so a rule that edits its
// element leaves the edit in the input, as a rule run on the element in
// place does; for a rule that edits nothing the check folds away at compile
// time.
template <typename In, typename Out, class Sel, class Proj>
__global__ __launch_bounds__(kThreads) void k_compact(In* __restrict__ in, uint64_t n, Out* __restrict__ out,
                                                      uint64_t* count, TileState st, uint64_t tiles, Sel sel,
                                                      Proj proj) {
   using T = typename std::remove_cv<In>::type;
   __shared__ uint32_t s_tile;
   __shared__ uint32_t s_cnt[kPerThread][kWaves];
   __shared__ uint64_t s_prefix;
   const int lane = __lane_id();
   const int wave = threadIdx.x / kWaveSize;
   if (threadIdx.x == 0) s_tile = __hip_atomic_fetch_add(st.ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
   __syncthreads();
   const uint32_t tile = s_tile;
   const uint64_t base = static_cast<uint64_t>(tile) * kTile;

   T v[kPerThread];
   uint64_t bal[kPerThread];
#pragma unroll
   for (int j = 0; j < kPerThread; ++j) {
      const uint64_t i = base + static_cast<uint64_t>(j) * kThreads + threadIdx.x;
      bool keep = false;
      if (i < n) {
         v[j] = in[i];
         if constexpr (std::is_const<In>::value) {
            keep = sel(i, v[j]);
         } else {
            const T before = v[j];
            keep = sel(i, v[j]);
            if (!same_bytes(before, v[j])) in[i] = v[j];
         }
      }
      bal[j] = __ballot(keep);
      if (lane == 0) s_cnt[j][wave] = static_cast<uint32_t>(__popcll(bal[j]));
   }
   __syncthreads();
   if (wave == 0) {
      // Exclusive prefix of the counts, j-major and wave-minor (= input
      // order), by one wave: each lane scans PER consecutive counts.
      constexpr int E = kPerThread * kWaves;
      constexpr int PER = (E + kWaveSize - 1) / kWaveSize;
      uint32_t* flat = &s_cnt[0][0];
      uint32_t c[PER];
      uint32_t sum = 0;
#pragma unroll
      for (int i = 0; i < PER; ++i) {
         const int idx = lane * PER + i;
         c[i] = idx < E ? flat[idx] : 0u;
         sum += c[i];
      }
      const uint32_t incl = wave_scan_incl(sum);
      uint32_t run = incl - sum;
#pragma unroll
      for (int i = 0; i < PER; ++i) {
         const int idx = lane * PER + i;
         if (idx < E) flat[idx] = run;
         run += c[i];
      }
      const uint64_t total = static_cast<uint32_t>(__shfl(static_cast<int>(incl), kWaveSize - 1, kWaveSize));
      const uint64_t pre = tile_prefix(st, tile, total);
      if (lane == 0) {
         s_prefix = pre;
         if (tile == tiles - 1 && count) *count = pre + total;
      }
   }
   __syncthreads();
   const uint64_t pre = s_prefix;
   const uint64_t below = (1ull << lane) - 1ull;
#pragma unroll
   for (int j = 0; j < kPerThread; ++j) {
      if ((bal[j] >> lane) & 1ull) {
         const uint64_t r = pre + s_cnt[j][wave] + static_cast<uint64_t>(__popcll(bal[j] & below));
         out[r] = proj(v[j]);
      }
   }
}

// The whole element, and a pair's first member (copy_keys_if).
struct Whole {
   template <typename T>
   __device__ T operator()(const T& x) const { return x; }
};
template <typename U>
struct FirstOf {
   template <typename T>
   __device__ U operator()(const T& x) const { return x.first; }
};

// A rule over elements (bool rule(T&)) as a selector.
template <class Rule>
struct RuleSel {
   Rule rule;
   template <typename T>
   __device__ bool operator()(uint64_t, T& x) const { return rule(x); }
};

// Enqueues the compaction of in[0, n) on stream s: out receives the selected
// elements' projections in input order and *count (device memory) their
// number.  scratch: scratch_bytes(n) bytes of device memory.
template <typename In, typename Out, class Sel, class Proj>
hipError_t launch(In* in, uint64_t n, Out* out, uint64_t* count, void* scratch, hipStream_t s, Sel sel, Proj proj) {
   const uint64_t tiles = tiles_for(n);
   const TileState st = state_at(scratch);
   const uint64_t want = (tiles + 255) / 256;
   const unsigned blocks = want == 0 ? 1u : static_cast<unsigned>(want < 2048 ? want : 2048);
   hipLaunchKernelGGL(k_prepare<uint64_t>, dim3(blocks), dim3(256), 0, s, st, tiles, count);
   if (tiles)
      hipLaunchKernelGGL((k_compact<In, Out, Sel, Proj>), dim3(static_cast<uint32_t>(tiles)), dim3(kThreads), 0, s,
                         in, n, out, count, st, tiles, sel, proj);
   return hipGetLastError();
}

} // namespace compact
} // namespace tools
} // namespace split
