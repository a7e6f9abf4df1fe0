// Host-side helpers shared by the C-ABI entry points: typed constants,
// error reporting, workspace carving and launch geometry.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

#include "tsph_common.h"

namespace tsph {

void set_error(const char* fmt, ...);

#define TSPH_REQUIRE(cond, code, ...) \
  do {                                \
    if (!(cond)) {                    \
      ::tsph::set_error(__VA_ARGS__); \
      return (code);                  \
    }                                 \
  } while (0)

#define TSPH_LAUNCH_CHECK(what)                                                       \
  do {                                                                                \
    hipError_t e_ = hipGetLastError();                                                \
    if (e_ != hipSuccess) {                                                           \
      ::tsph::set_error("%s: %s", (what), hipGetErrorString(e_));                     \
      return TSPH_EHIP;                                                               \
    }                                                                                 \
  } while (0)

inline int blocks_for(long n) { return (int)((n + kBlock - 1) / kBlock); }

// Build the typed constants from the C parameter block.  Everything the
// reference stores in f32 fields arrives here already f32-rounded, so the
// conversion to T is exact; derived quotients are formed in T exactly as the
// kernels of the reference form them (sig/h, 32/pi/h^3, 0.01*h^2).
template <class T>
KP<T> make_kp(const TsphParams& p) {
  KP<T> k;
  std::memset(&k, 0, sizeof(k));
  k.dim = p.dim;
  k.P = p.phase_num;
  k.solver = p.solver_type;
  k.ntemp = p.ntemp;
  k.node_num = p.node_num;
  k.R = p.neighb_range;
  for (int d = 0; d < 3; ++d) {
    k.gs[d] = p.grid_size[d];
    k.coder[d] = p.coder[d];
    k.lb[d] = (T)p.lb[d];
    k.g[d] = (T)p.gravity[d];
  }
  for (int q = 0; q < kMaxP; ++q) k.rho[q] = (T)p.phase_rest_density[q];
  k.list_cap_f = p.list_cap_f;
  k.list_cap_b = p.list_cap_b;
  k.iter_min = p.iter_min;
  k.iter_max = p.iter_max;
  k.cs = p.cs;
  k.wc_gamma = p.wc_gamma;
  volatile T h = (T)p.kernel_h[1];
  k.h = h;
  k.h2 = (T)p.kernel_h[2];
  k.sig = (T)p.kernel_sig[p.dim];
  volatile T sig = k.sig;
  k.sig_h = sig / h;
  volatile T h3 = h * (h * h);
  k.c_coh = (T)(32.0 / M_PI) / h3;
  volatile T h2 = k.h2;
  k.eps = (T)0.01 * h2;
  k.r2max = (T)p.r2_max;
  k.inv_h = (T)1 / h;
  k.gps = (T)p.part_size[1];
  k.mu = (T)p.dynamic_viscosity;
  k.mu_pp = p.mu_per_phase ? 1 : 0;
  for (int q = 0; q < kMaxP; ++q) k.mu_k[q] = (T)p.phase_viscosity[q];
  k.gamma = (T)p.surface_tension_gamma;
  k.fbmD = (T)p.fbm_diffusion;
  k.fbmC = (T)p.fbm_convection;
  k.cfl = (T)p.cfl_factor;
  k.div_thr = (T)p.divergence_threshold;
  k.comp_thr = (T)p.compression_threshold;
  k.two_d2 = (T)(2 * (2 + p.dim));
  k.wg_a = (T)(6.0 * (double)k.sig / ((double)h * (double)h));
  k.wg_b = (T)(6.0 * (double)k.sig / (double)h);
  k.fastq = (sizeof(T) == 4 && (p.flags & TSPH_FLAG_FAST_QUOTIENT)) ? 1 : 0;
  k.st_rec = (p.flags & TSPH_FLAG_ST_RECORDS) ? 1 : 0;
  return k;
}

// Bump allocator over the caller's workspace (256-byte aligned segments).
struct Carve {
  char* base;
  size_t cap, off = 0;
  bool ok = true;
  Carve(void* b, size_t c) : base((char*)b), cap(c) {}
  template <class U>
  U* take(size_t count) {
    size_t bytes = (count * sizeof(U) + 255) & ~(size_t)255;
    if (off + bytes > cap) {
      ok = false;
      return nullptr;
    }
    U* r = (U*)(base + off);
    off += bytes;
    return r;
  }
};

__host__ __device__ inline size_t seg(size_t bytes) { return (bytes + 255) & ~(size_t)255; }

// Scan tile: counts per scan block.
constexpr int kScanTile = kBlock * 16;

__host__ __device__ inline long scan_blocks(long len) { return (len + kScanTile - 1) / kScanTile; }

// Workspace layout used by all calls for objects of up to `cap` slots and a
// grid of node_num cells (reference grid: node_num * 2 (cell,uid) buckets).
struct WsLayout {
  double* partial;     // [2 * maxblk] block partials for reductions
  int* scan_sums;      // [scan blocks]
  int* key;            // [cap]
  int* rank;           // [cap]
  int* tmp_idx;        // [cap]
  int* tmp_pid;        // [cap]
  int* tmp_key;        // [cap]
  int* cu_count;       // [2*node_num + 2] reference grid (cell, uid) counts
  int* cu_start;       // [2*node_num + 3]
  void* krec;          // [cap] V4 of the step's precision: position and kappa of
                       // each fluid particle in the current Jacobi iteration
  void* a1b;           // [cap] V4: boundary part of alpha_1 (sum_b X_j grad W_ij), pass A
};

inline size_t ws_bytes_needed(long cap, long node_num) {
  long maxblk = blocks_for(cap) + 8;
  long lenref = 2 * node_num + 2;
  return seg(sizeof(double) * 2 * maxblk) + seg(sizeof(int) * (scan_blocks(lenref + 1) + 8)) +
         5 * seg(sizeof(int) * cap) + seg(sizeof(int) * (lenref + 1)) +
         seg(sizeof(int) * (lenref + 2)) + 2 * seg(4 * sizeof(double) * cap) + 4096;
}

inline bool carve_ws(void* ws, size_t bytes, long cap, long node_num, WsLayout* L) {
  Carve c(ws, bytes);
  long maxblk = blocks_for(cap) + 8;
  long lenref = 2 * node_num + 2;
  L->partial = c.take<double>(2 * maxblk);
  L->scan_sums = c.take<int>(scan_blocks(lenref + 1) + 8);
  L->key = c.take<int>(cap);
  L->rank = c.take<int>(cap);
  L->tmp_idx = c.take<int>(cap);
  L->tmp_pid = c.take<int>(cap);
  L->tmp_key = c.take<int>(cap);
  L->cu_count = c.take<int>(lenref + 1);
  L->cu_start = c.take<int>(lenref + 2);
  L->krec = c.take<double>(4 * cap);
  L->a1b = c.take<double>(4 * cap);
  return c.ok;
}

// ---------------------------------------------------------------------------
// Batched scenes.  A device-resident table holds one entry per scene (its
// constants, object descriptors, control block and workspace layout),
// followed by per-kind tile prefixes: first[kind][s] is the first tile of
// scene s in a grouped launch of that kind.  Grouped kernels loop over tiles
// (grid-stride), find each tile's scene by binary search and run the same
// per-block body as the single-scene kernel.
// ---------------------------------------------------------------------------
template <class T>
struct SceneEntry {
  KP<T> k;
  TsphObject f, b;
  TsphControl* ctrl;
  WsLayout L;
};

enum TileKind { kTileFluid = 0, kTileBound = 1, kTileCell = 2, kTileScene = 3, kTileKinds = 4 };

template <class T>
__host__ __device__ inline size_t batch_entries_bytes(int ns) {
  return seg(sizeof(SceneEntry<T>) * (size_t)ns);
}
template <class T>
__host__ __device__ inline size_t batch_table_bytes(int ns) {
  return batch_entries_bytes<T>(ns) + sizeof(int) * kTileKinds * (size_t)(ns + 1);
}

// Cells of a scene's counting sort (cells + the trailing invalid bucket) and
// the scan tiles covering them plus the count array's extra slot.
__host__ __device__ inline long sort_len(long node_num) { return node_num + 1; }
__host__ __device__ inline long cell_tiles(long node_num) { return scan_blocks(sort_len(node_num) + 1); }

// The largest scene s with first[s] <= t, for a whole wave (all 64 lanes
// active, t wave-uniform): the lanes test 64 prefix entries at once and the
// vote counts those <= t, so the lookup costs one load per 64 scenes instead
// of a binary search's chain of dependent loads.  first[] is non-decreasing,
// so the entries <= t form a prefix.
__device__ __forceinline__ int scene_of_wave(const int* first, int ns, int t) {
  const int lane = (int)(threadIdx.x & 63);
  for (int c = 0; c < ns; c += 64) {
    const int s = c + lane;
    const int k = __popcll(__ballot(s < ns && first[s] <= t));
    if (k < 64) return c + k - 1;
  }
  return ns - 1;
}

// The tile prefix of one kind: first[s] is scene s's first tile, first[ns]
// the kind's tile count.
template <class T, int KIND>
__device__ __forceinline__ const int* tile_prefix(const char* __restrict__ tab, int ns) {
  return (const int*)(tab + batch_entries_bytes<T>(ns)) + KIND * (ns + 1);
}

// Runs f(entry, scene, blk) for tile t of one kind (t uniform per block).
template <class T, int KIND, class F>
__device__ __forceinline__ void on_tile(const char* __restrict__ tab, int ns, int t, F&& f) {
  // read-only here (the sort's pointer swap is a separate kernel), so the
  // scene entries can be scalar loads
  const SceneEntry<T>* __restrict__ E = (const SceneEntry<T>*)tab;
  const int* first = tile_prefix<T, KIND>(tab, ns);
  // uniform per block: readfirstlane lets the scene's constants live in
  // scalar registers
  const int s = __builtin_amdgcn_readfirstlane(scene_of_wave(first, ns, t));
  const int f0 = first[s];
  f(E[s], s, Blk{t - f0, first[s + 1] - f0});
}

// Grid-stride loop over the tiles of one kind; f(entry, scene, blk).
template <class T, int KIND, class F>
__device__ __forceinline__ void for_tiles(const char* __restrict__ tab, int ns, F&& f) {
  const int total = tile_prefix<T, KIND>(tab, ns)[ns];
  for (int t = blockIdx.x; t < total; t += gridDim.x) on_tile<T, KIND>(tab, ns, t, f);
}

inline int grid_for(int tiles, int grid) {
  if (tiles <= 0) return 0;
  return (grid <= 0 || grid > tiles) ? tiles : grid;
}

}  // namespace tsph
