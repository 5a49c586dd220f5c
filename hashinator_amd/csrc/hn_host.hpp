// hn_host.hpp -- host-side launch geometry shared by the C-ABI entry points.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <stdint.h>

#include "hashinator_amd.h"

namespace hn {

enum OpKind { OP_INSERT = 0, OP_RETRIEVE = 1, OP_ERASE = 2, OP_REINSERT = 3 };

struct Launch {
   int probe_width;
   int block_size;
   uint32_t blocks;
};

// Compute units of the current device (cached per device ordinal).
int cu_count();

// Default probe width of the group kernels.  Inserts: 4 lanes (windows of 64
// B for u64, 32 B for u32), measured fastest on MI355X for tables beyond the
// caches at LF 0.5-0.9 (u64 at 2^28 slots, LF 0.9: 21.8 ms against 22.9 ms
// with 8 lanes and 30.9 ms with 16; u32 at 2^26: 5.6 ms against 7.5 ms with
// 16).  Group lookups (reset_to_empty, and any caller that asks for the
// group path): one 128-byte window per step.  Retrieve and erase default to
// the lane kernel, which has no width.
template <typename K>
constexpr int default_probe_width(OpKind op) {
   if (op == OP_INSERT || op == OP_REINSERT) return 4;
   return sizeof(K) == 8 ? 8 : 16;
}

constexpr int kDefaultBlock = 256;
constexpr int kResidentBlocksPerCU = 8; // 8 x 256 threads = 32 waves = one full CU

template <typename K>
int resolve_launch(const hn_tune* t, OpKind op, uint64_t len, Launch* L) {
   int G = (t && t->probe_width) ? t->probe_width : default_probe_width<K>(op);
   int bs = (t && t->block_size) ? t->block_size : kDefaultBlock;
   if (G < 1 || G > 64 || (G & (G - 1))) return HN_EINVAL;
   if (bs < 64 || bs > 1024 || (bs % 64)) return HN_EINVAL;
   const uint64_t per_block = static_cast<uint64_t>(bs / G);
   uint64_t need = (len + per_block - 1) / per_block;
   if (need == 0) need = 1;
   uint64_t cap = (t && t->grid_blocks > 0) ? static_cast<uint64_t>(t->grid_blocks)
                                            : static_cast<uint64_t>(cu_count()) * kResidentBlocksPerCU * (1024 / bs) / 4;
   if (cap == 0) cap = 1;
   if (cap > (1u << 20)) cap = 1u << 20;
   L->probe_width = G;
   L->block_size = bs;
   L->blocks = static_cast<uint32_t>(need < cap ? need : cap);
   return 0;
}

// Resident blocks per CU of `kernel` at `block` threads (its occupancy),
// asked once per kernel and kept in `cache`: the lane kernels split their
// work statically over exactly that many waves per CU.
template <typename F>
int resident_blocks_per_cu(F kernel, int block, std::atomic<int>& cache) {
   int v = cache.load(std::memory_order_relaxed);
   if (v > 0) return v;
   int nb = 0;
   if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, block, 0) != hipSuccess || nb < 1) nb = 8;
   cache.store(nb, std::memory_order_relaxed);
   return nb;
}

// Grid for a grid-stride streaming kernel over n items.
inline uint32_t stream_blocks(uint64_t n, int bs) {
   uint64_t need = (n + bs - 1) / bs;
   const uint64_t cap = static_cast<uint64_t>(cu_count()) * 32;
   if (need == 0) need = 1;
   return static_cast<uint32_t>(need < cap ? need : cap);
}

} // namespace hn
