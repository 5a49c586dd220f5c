// hn_compact.hip -- the C-ABI's stream compaction and exclusive scan (gfx950).
//
// Compaction is the single-pass kernel of <splitvector/split_compact.h>,
// instantiated here for the built-in predicates (hn_pred), flag bytes and
// bit masks.  The scan is one kernel over tiles with decoupled look-back: a
// tile takes a ticket (so every tile it waits on has already started),
// publishes its own aggregate, then walks back over its predecessors 64 at a
// time with one wave until it meets an inclusive prefix, and publishes its
// own inclusive prefix.  Each reads its input once and writes its output
// once, against the reference's count pass + Blelloch scan + scatter pass
// with a host sync between kernels (split_tools.h:920-989, 363-418).
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <mutex>

#include <splitvector/split_compact.h>

#include "hn_device.hpp"
#include "hn_host.hpp"

namespace hn {

namespace sc = split::tools::compact;

// Scan tiles are large because every tile pays a serial cost: its ticket (one
// atomic on one word) and its step in the look-back chain.  At 2^28 u32,
// 256-thread tiles ran the scan in 1.0 ms; 1024 threads with 256 bytes each
// take 0.41 ms.
constexpr int kBS = 1024;        // threads per tile
constexpr int kWaves = kBS / kWave;
constexpr int kScanBytes = 256;  // consecutive bytes per thread in a scan tile
constexpr uint64_t kFlagAgg = 1, kFlagInc = 2;

// Per-tile look-back state: the 64-bit value in two words, each carrying
// its own state (0 none, 1 aggregate, 2 inclusive) in the top bits and one
// 32-bit half below.  Words are written and read with relaxed device-scope
// atomics; a reader takes a tile's value only when both halves carry the
// same state, so no release/acquire fence is needed (on a multi-XCD device
// those write back or invalidate the whole L2).
struct Lookback {
   uint32_t* ticket;
   uint64_t* lo; // [tiles]
   uint64_t* hi; // [tiles]
};

constexpr size_t lookback_bytes(uint64_t tiles) { return 256 + tiles * 16 + 64; }

inline Lookback lookback_layout(void* scr, uint64_t tiles) {
   char* p = static_cast<char*>(scr);
   Lookback L;
   L.ticket = reinterpret_cast<uint32_t*>(p);
   L.lo = reinterpret_cast<uint64_t*>(p + 256);
   L.hi = L.lo + tiles;
   return L;
}

__global__ __launch_bounds__(256) void k_lookback_init(Lookback L, uint64_t tiles) {
   const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
   for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < tiles; i += stride) {
      L.lo[i] = 0;
      L.hi[i] = 0;
   }
   if (blockIdx.x == 0 && threadIdx.x == 0) *L.ticket = 0;
}

__device__ __forceinline__ void lb_publish(const Lookback& L, uint64_t tile, uint64_t flag, uint64_t v) {
   __hip_atomic_store(&L.hi[tile], (flag << 62) | (v >> 32), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
   __hip_atomic_store(&L.lo[tile], (flag << 62) | (v & 0xFFFFFFFFull), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// State of tile p (0 while its two halves disagree) and its value.
__device__ __forceinline__ uint64_t lb_read(const Lookback& L, uint64_t p, uint64_t* v) {
   const uint64_t a = __hip_atomic_load(&L.lo[p], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
   const uint64_t b = __hip_atomic_load(&L.hi[p], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
   if ((a >> 62) != (b >> 62)) return 0;
   *v = ((b & 0xFFFFFFFFull) << 32) | (a & 0xFFFFFFFFull);
   return a >> 62;
}

// Exclusive prefix of tile `tile` whose own aggregate is `mine`; called by
// one full wave.  Publishes the tile's aggregate and inclusive prefix.
__device__ uint64_t lookback_prefix(const Lookback& L, uint32_t tile, uint64_t mine) {
   const int lane = __lane_id();
   if (tile == 0) {
      if (lane == 0) lb_publish(L, 0, kFlagInc, mine);
      return 0;
   }
   if (lane == 0) lb_publish(L, tile, kFlagAgg, mine);
   uint64_t prefix = 0;
   int64_t hi = static_cast<int64_t>(tile) - 1; // newest predecessor not yet folded in
   for (;;) {
      const int64_t p = hi - lane;
      uint64_t f = kFlagInc; // lanes past tile 0 behave as an inclusive zero
      uint64_t v = 0;
      if (p >= 0) f = lb_read(L, static_cast<uint64_t>(p), &v);
      const uint64_t not_ready = __ballot(f == 0);
      const uint64_t inc_mask = __ballot(f == kFlagInc);
      // Lanes up to and including the first inclusive one, if none before it is still pending.
      const int first_inc = inc_mask ? first_bit(inc_mask) : 64;
      const uint64_t upto = first_inc >= 63 ? ~0ull : ((2ull << first_inc) - 1ull);
      if (not_ready & upto) {
         __builtin_amdgcn_s_sleep(1);
         continue;
      }
      const uint64_t take = (lane <= first_inc) ? v : 0;
      prefix += wave_sum(take);
      if (first_inc < 64) break;
      hi -= 64;
   }
   if (lane == 0) lb_publish(L, tile, kFlagInc, prefix + mine);
   return prefix;
}

template <typename T>
__device__ __forceinline__ T wave_incl_scan(T x) {
   for (int o = 1; o < kWave; o <<= 1) {
      T y = __shfl_up(x, o, kWave);
      if (__lane_id() >= o) x += y;
   }
   return x;
}
// u32: DPP row shifts inside 16-lane rows, then row broadcasts across rows.
template <>
__device__ __forceinline__ uint32_t wave_incl_scan<uint32_t>(uint32_t x) {
   int v = static_cast<int>(x);
   v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false); // row_shr:1
   v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false); // row_shr:2
   v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false); // row_shr:4
   v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false); // row_shr:8
   v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false); // row_bcast:15 -> rows 1, 3
   v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false); // row_bcast:31 -> rows 2, 3
   return static_cast<uint32_t>(v);
}

// ------------------------------------------------------------ compaction

template <typename K>
struct PairSel {
   hn_pred p;
   __device__ bool operator()(uint64_t, const Slot<K>& s) const {
      const K f = p.field == 0 ? s.key : s.val;
      switch (p.kind) {
      case HN_PRED_ALL: return true;
      case HN_PRED_VALID: return s.key != empty_key<K>() && s.key != tomb_key<K>();
      case HN_PRED_MASKEQ: return (f & static_cast<K>(p.mask)) == static_cast<K>(p.value);
      case HN_PRED_LT: return f < static_cast<K>(p.value);
      case HN_PRED_GE: return f >= static_cast<K>(p.value);
      default: return false;
      }
   }
};

template <typename T>
struct FlagSel {
   const uint8_t* flags;
   __device__ bool operator()(uint64_t i, const T&) const { return flags[i] != 0; }
};

// Rule already evaluated into a bit mask (bit i of word i/32, LSB first).
template <typename T>
struct BitSel {
   const uint32_t* bits;
   __device__ bool operator()(uint64_t i, const T&) const { return (bits[i >> 5] >> (i & 31)) & 1u; }
};

template <typename T>
struct Whole {
   __device__ T operator()(const T& x) const { return x; }
};
template <typename K>
struct KeyOf {
   __device__ K operator()(const Slot<K>& s) const { return s.key; }
};

// ------------------------------------------------------------ scan

// A tile is ROWS rows of kBS 16-byte vectors (kRows when both pointers are
// 16-byte aligned, kScanRowsScalar rows of element loads otherwise); thread
// t holds vector t of every row, so each load and store of a wave covers
// 1 KB of consecutive bytes.  Row sums are scanned across the wave, the (row, wave) totals in
// LDS, in row-major order.  (With 256 consecutive bytes per thread instead,
// a wave's 16-byte loads land 256 B apart and the 2^28-element u32 scan took
// 0.60 ms; striped, 0.41 ms.)
constexpr int kRows = kScanBytes / 16;
constexpr int kScanRowsScalar = 4; // rows of a tile read element by element (unaligned pointers)
template <typename T, bool VEC, int ROWS>
__global__ __launch_bounds__(kBS) void k_scan(const T* in, T* out, uint64_t n, Lookback L, uint64_t tiles) {
   constexpr int V = 16 / sizeof(T);
   constexpr int IPT = ROWS * V;
   constexpr uint64_t TILE = static_cast<uint64_t>(kBS) * IPT;
   __shared__ uint32_t s_tile;
   __shared__ T s_rw[ROWS * kWaves];
   __shared__ T s_part[ROWS * kWaves / kWave];
   __shared__ uint64_t s_prefix;
   const int lane = __lane_id();
   const int wave = threadIdx.x / kWave;
   if (threadIdx.x == 0) s_tile = __hip_atomic_fetch_add(L.ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
   __syncthreads();
   const uint32_t tile = s_tile;
   const uint64_t t0 = static_cast<uint64_t>(tile) * TILE;
   const bool full = VEC && t0 + TILE <= n;
   T x[IPT];
   if (full) {
      const uint4* src = reinterpret_cast<const uint4*>(in + t0);
#pragma unroll
      for (int c = 0; c < ROWS; ++c) reinterpret_cast<uint4*>(x)[c] = src[c * kBS + threadIdx.x];
   } else {
#pragma unroll
      for (int c = 0; c < ROWS; ++c)
#pragma unroll
         for (int k = 0; k < V; ++k) {
            const uint64_t i = t0 + (static_cast<uint64_t>(c) * kBS + threadIdx.x) * V + k;
            x[c * V + k] = i < n ? in[i] : T(0);
         }
   }
   T rs[ROWS], wi[ROWS];
#pragma unroll
   for (int c = 0; c < ROWS; ++c) {
      T a = 0;
#pragma unroll
      for (int k = 0; k < V; ++k) a += x[c * V + k];
      rs[c] = a;
      wi[c] = wave_incl_scan<T>(a);
      if (lane == kWave - 1) s_rw[c * kWaves + wave] = wi[c];
   }
   __syncthreads();
   // exclusive scan of the ROWS * kWaves (row, wave) totals, row-major
   constexpr int NE = ROWS * kWaves;
   static_assert(NE % kWave == 0 && NE / kWave <= kWaves, "one entry per thread of the first waves");
   T e = 0, ei = 0;
   if (threadIdx.x < NE) {
      e = s_rw[threadIdx.x];
      ei = wave_incl_scan<T>(e);
      if (lane == kWave - 1) s_part[wave] = ei;
   }
   __syncthreads();
   if (threadIdx.x < NE) {
      T before = 0;
      for (int w = 0; w < wave; ++w) before += s_part[w];
      s_rw[threadIdx.x] = before + ei - e;
   }
   T tile_total = 0;
#pragma unroll
   for (int w = 0; w < NE / kWave; ++w) tile_total += s_part[w];
   if (wave == 0) {
      const uint64_t pre = lookback_prefix(L, tile, static_cast<uint64_t>(tile_total));
      if (lane == 0) s_prefix = pre;
   }
   __syncthreads();
   const T pre = static_cast<T>(s_prefix);
#pragma unroll
   for (int c = 0; c < ROWS; ++c) {
      T run = pre + s_rw[c * kWaves + wave] + (wi[c] - rs[c]);
#pragma unroll
      for (int k = 0; k < V; ++k) {
         const T v = x[c * V + k];
         x[c * V + k] = run;
         run += v;
      }
   }
   if (full) {
      uint4* dst = reinterpret_cast<uint4*>(out + t0);
#pragma unroll
      for (int c = 0; c < ROWS; ++c) dst[c * kBS + threadIdx.x] = reinterpret_cast<uint4*>(x)[c];
   } else {
#pragma unroll
      for (int c = 0; c < ROWS; ++c)
#pragma unroll
         for (int k = 0; k < V; ++k) {
            const uint64_t i = t0 + (static_cast<uint64_t>(c) * kBS + threadIdx.x) * V + k;
            if (i < n) out[i] = x[c * V + k];
         }
   }
}

// ------------------------------------------------------------ host side

constexpr uint64_t compact_tiles(uint64_t n) { return sc::tiles_for(n); }

template <typename In, typename Out, class Sel, class Proj>
int run_compact(const In* in, uint64_t n, Out* out, uint64_t* d_count, void* scr, size_t bytes, hipStream_t s,
                Sel sel, Proj proj) {
   if ((n && (!in || !out)) || !d_count) return HN_EINVAL;
   if (!scr || bytes < sc::scratch_bytes(n)) return HN_ESCRATCH;
   return static_cast<int>(sc::launch(in, n, out, d_count, scr, s, sel, proj));
}

template <typename T>
int run_scan(const T* in, T* out, uint64_t n, void* scr, size_t bytes, hipStream_t s) {
   if (n && (!in || !out)) return HN_EINVAL;
   if (n == 0) return 0;
   // Pointers not 16-byte aligned take element loads and a tile of
   // kScanRowsScalar rows (16 rows of element loads ran out of registers);
   // either tile holds at least a compaction tile, so the scratch sized by
   // hn_compact_scratch_bytes covers its look-back words.
   const bool vec = (reinterpret_cast<uintptr_t>(in) % 16 == 0) && (reinterpret_cast<uintptr_t>(out) % 16 == 0);
   const uint64_t TILE = static_cast<uint64_t>(kBS) * (16 / sizeof(T)) * (vec ? kRows : kScanRowsScalar);
   const uint64_t tiles = (n + TILE - 1) / TILE;
   if (!scr || bytes < lookback_bytes(tiles)) return HN_ESCRATCH;
   Lookback L = lookback_layout(scr, tiles);
   hipLaunchKernelGGL(k_lookback_init, dim3(stream_blocks(tiles, 256)), dim3(256), 0, s, L, tiles);
   if (vec)
      hipLaunchKernelGGL((k_scan<T, true, kRows>), dim3(static_cast<uint32_t>(tiles)), dim3(kBS), 0, s, in, out, n, L,
                         tiles);
   else
      hipLaunchKernelGGL((k_scan<T, false, kScanRowsScalar>), dim3(static_cast<uint32_t>(tiles)), dim3(kBS), 0, s, in,
                         out, n, L, tiles);
   return static_cast<int>(hipGetLastError());
}

// ------------------------------------------------------------ compaction without allocation
//
// hn_copy_if_pairs_noalloc_u64 keeps its look-back state (ticket and tile
// words) in `scratch`, a file-scope device array sized for the largest
// input it accepts, HN_NOALLOC_MAX_ELEMS elements (a 2^32-slot table):
// the ticket and 2^19 one-word tile states, 4 MiB + 256 bytes, held by
// every device the library runs on until the process ends.  Host code never names the array
// in a launch or a copy: its host-side name is not a device address.  The
// device address is looked up once per device (with that device current)
// and kept in noalloc_ptr[].
constexpr uint64_t kNoAllocMaxElems = HN_NOALLOC_MAX_ELEMS;
constexpr size_t kNoAllocBytes = sc::scratch_bytes(kNoAllocMaxElems);
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
std::atomic<void*> noalloc_ptr[kMaxDevices];
std::mutex noalloc_mu;

// The current device's address of `scratch`, or a null pointer with *rc set.
void* noalloc_scratch(int* rc) {
   int dev = 0;
   hipError_t e = hipGetDevice(&dev);
   if (e != hipSuccess) {
      *rc = static_cast<int>(e);
      return nullptr;
   }
   if (dev < 0 || dev >= kMaxDevices) {
      *rc = HN_EINVAL;
      return nullptr;
   }
   void* p = noalloc_ptr[dev].load(std::memory_order_acquire);
   if (p) return p;
   std::lock_guard<std::mutex> lock(noalloc_mu);
   p = noalloc_ptr[dev].load(std::memory_order_acquire);
   if (!p) {
      e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !p) {
         *rc = e != hipSuccess ? static_cast<int>(e) : HN_ENOLIB;
         return nullptr;
      }
      noalloc_ptr[dev].store(p, std::memory_order_release);
   }
   return p;
}

} // namespace hn

using namespace hn;

extern "C" {

size_t hn_compact_scratch_bytes(size_t n) {
   // The compaction tile is the smaller one, so it bounds both kernels.
   return lookback_bytes(compact_tiles(n) + 1);
}

int hn_copy_if_pairs_u32(const void* in, size_t n, const hn_pred* p, void* out, uint64_t* c, void* sc, size_t sb,
                         hn_stream_t s) {
   if (!p) return HN_EINVAL;
   return run_compact(static_cast<const Slot<uint32_t>*>(in), n, static_cast<Slot<uint32_t>*>(out), c, sc, sb, s,
                      PairSel<uint32_t>{*p}, Whole<Slot<uint32_t>>{});
}
int hn_copy_if_pairs_u64(const void* in, size_t n, const hn_pred* p, void* out, uint64_t* c, void* sc, size_t sb,
                         hn_stream_t s) {
   if (!p) return HN_EINVAL;
   return run_compact(static_cast<const Slot<uint64_t>*>(in), n, static_cast<Slot<uint64_t>*>(out), c, sc, sb, s,
                      PairSel<uint64_t>{*p}, Whole<Slot<uint64_t>>{});
}
int hn_copy_if_pairs_noalloc_u64(const void* in, size_t n, const hn_pred* p, void* out, uint64_t* c, hn_stream_t s) {
   if (!p) return HN_EINVAL;
   if (n > kNoAllocMaxElems) return HN_ESCRATCH;
   int rc = 0;
   void* sc = noalloc_scratch(&rc);
   if (!sc) return rc;
   return run_compact(static_cast<const Slot<uint64_t>*>(in), n, static_cast<Slot<uint64_t>*>(out), c, sc,
                      kNoAllocBytes, s, PairSel<uint64_t>{*p}, Whole<Slot<uint64_t>>{});
}
int hn_copy_keys_if_pairs_u32(const void* in, size_t n, const hn_pred* p, uint32_t* out, uint64_t* c, void* sc,
                              size_t sb, hn_stream_t s) {
   if (!p) return HN_EINVAL;
   return run_compact(static_cast<const Slot<uint32_t>*>(in), n, out, c, sc, sb, s, PairSel<uint32_t>{*p},
                      KeyOf<uint32_t>{});
}
int hn_copy_keys_if_pairs_u64(const void* in, size_t n, const hn_pred* p, uint64_t* out, uint64_t* c, void* sc,
                              size_t sb, hn_stream_t s) {
   if (!p) return HN_EINVAL;
   return run_compact(static_cast<const Slot<uint64_t>*>(in), n, out, c, sc, sb, s, PairSel<uint64_t>{*p},
                      KeyOf<uint64_t>{});
}

int hn_copy_if_flags(const void* in, const uint8_t* flags, size_t n, size_t eb, void* out, uint64_t* c, void* sc,
                     size_t sb, hn_stream_t s) {
   if (n && !flags) return HN_EINVAL;
   switch (eb) {
   case 4:
      return run_compact(static_cast<const uint32_t*>(in), n, static_cast<uint32_t*>(out), c, sc, sb, s,
                         FlagSel<uint32_t>{flags}, Whole<uint32_t>{});
   case 8:
      return run_compact(static_cast<const uint64_t*>(in), n, static_cast<uint64_t*>(out), c, sc, sb, s,
                         FlagSel<uint64_t>{flags}, Whole<uint64_t>{});
   case 16:
      return run_compact(static_cast<const Slot<uint64_t>*>(in), n, static_cast<Slot<uint64_t>*>(out), c, sc, sb, s,
                         FlagSel<Slot<uint64_t>>{flags}, Whole<Slot<uint64_t>>{});
   default: return HN_EINVAL;
   }
}

int hn_copy_if_bits(const void* in, const uint32_t* bits, size_t n, size_t eb, void* out, uint64_t* c, void* sc,
                    size_t sb, hn_stream_t s) {
   if (n && !bits) return HN_EINVAL;
   switch (eb) {
   case 4:
      return run_compact(static_cast<const uint32_t*>(in), n, static_cast<uint32_t*>(out), c, sc, sb, s,
                         BitSel<uint32_t>{bits}, Whole<uint32_t>{});
   case 8:
      return run_compact(static_cast<const uint64_t*>(in), n, static_cast<uint64_t*>(out), c, sc, sb, s,
                         BitSel<uint64_t>{bits}, Whole<uint64_t>{});
   case 16:
      return run_compact(static_cast<const Slot<uint64_t>*>(in), n, static_cast<Slot<uint64_t>*>(out), c, sc, sb, s,
                         BitSel<Slot<uint64_t>>{bits}, Whole<Slot<uint64_t>>{});
   default: return HN_EINVAL;
   }
}
int hn_copy_keys_if_bits_u32(const void* in, const uint32_t* bits, size_t n, uint32_t* out, uint64_t* c, void* sc,
                             size_t sb, hn_stream_t s) {
   if (n && !bits) return HN_EINVAL;
   return run_compact(static_cast<const Slot<uint32_t>*>(in), n, out, c, sc, sb, s, BitSel<Slot<uint32_t>>{bits},
                      KeyOf<uint32_t>{});
}
int hn_copy_keys_if_bits_u64(const void* in, const uint32_t* bits, size_t n, uint64_t* out, uint64_t* c, void* sc,
                             size_t sb, hn_stream_t s) {
   if (n && !bits) return HN_EINVAL;
   return run_compact(static_cast<const Slot<uint64_t>*>(in), n, out, c, sc, sb, s, BitSel<Slot<uint64_t>>{bits},
                      KeyOf<uint64_t>{});
}

int hn_copy_keys_if_flags_u32(const void* in, const uint8_t* flags, size_t n, uint32_t* out, uint64_t* c, void* sc,
                              size_t sb, hn_stream_t s) {
   if (n && !flags) return HN_EINVAL;
   return run_compact(static_cast<const Slot<uint32_t>*>(in), n, out, c, sc, sb, s, FlagSel<Slot<uint32_t>>{flags},
                      KeyOf<uint32_t>{});
}
int hn_copy_keys_if_flags_u64(const void* in, const uint8_t* flags, size_t n, uint64_t* out, uint64_t* c, void* sc,
                              size_t sb, hn_stream_t s) {
   if (n && !flags) return HN_EINVAL;
   return run_compact(static_cast<const Slot<uint64_t>*>(in), n, out, c, sc, sb, s, FlagSel<Slot<uint64_t>>{flags},
                      KeyOf<uint64_t>{});
}

int hn_exclusive_scan_u32(const uint32_t* in, uint32_t* out, size_t n, void* sc, size_t sb, hn_stream_t s) {
   return run_scan<uint32_t>(in, out, n, sc, sb, s);
}
int hn_exclusive_scan_u64(const uint64_t* in, uint64_t* out, size_t n, void* sc, size_t sb, hn_stream_t s) {
   return run_scan<uint64_t>(in, out, n, sc, sb, s);
}

} // extern "C"
