// Neighbour iteration shared by the step and adjoint sweeps.
//
// Storage is in cell order.  The 27 (3D) / 9 (2D) cells of a particle's
// neighbourhood are visited as 9 / 3 contiguous row runs: cells (x-R .. x+R)
// of one (y, z) row have consecutive codes and therefore consecutive slots,
// in the reference's template order (x fastest, sph_config.py:219-224) and
// in-cell order.
#pragma once

#include "tsph_host.h"

namespace tsph {

// Visit the row runs [j0, j1) of `cstart`'s object in the cells around c[]:
// one run per (dy, dz) row, in template order.
template <class T, class F>
__device__ __forceinline__ void walk_runs(const KP<T>& k, const int* __restrict__ cstart,
                                          const int* c, F&& f) {
  const int R = k.R;
  const int zr = k.dim == 3 ? R : 0;
#pragma unroll 1
  for (int dz = -zr; dz <= zr; ++dz) {
#pragma unroll 1
    for (int dy = -R; dy <= R; ++dy) {
      int c0 = dim_encode(k, wadd(c[0], -R), wadd(c[1], dy), wadd(c[2], dz));
      long lo = c0 < 1 ? 1 : (long)c0;
      long hi = (long)c0 + 2 * R;
      if (hi > k.node_num - 1) hi = k.node_num - 1;
      if (lo > hi) continue;
      f(cstart[lo], cstart[hi + 1]);
    }
  }
}

// Visit every candidate slot of `cstart`'s object in the cells around c[].
template <class T, class F>
__device__ __forceinline__ void walk_cells(const KP<T>& k, const int* __restrict__ cstart,
                                           const int* c, F&& f) {
  walk_runs(k, cstart, c, [&](int j0, int j1) {
    for (int j = j0; j < j1; ++j) f(j);
  });
}

// Candidate walk with prefetch: the records load(j) of a run's candidates are
// gathered U at a time, all U in flight together, then f(j, record) runs for
// each in walk order.  Cell walks are bound by the latency of these gathers
// (a one-candidate-at-a-time loop waits a full round trip per candidate),
// not by their number.
template <int U, class T, class Load, class F>
__device__ __forceinline__ void walk_cells_prefetch(const KP<T>& k, const int* __restrict__ cstart,
                                                    const int* c, Load&& load, F&& f) {
  walk_runs(k, cstart, c, [&](int j0, int j1) {
    for (int j = j0; j < j1; j += U) {
      decltype(load(j0)) rec[U];
#pragma unroll
      for (int u = 0; u < U; ++u) rec[u] = load(j + u < j1 ? j + u : j);
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (j + u < j1) f(j + u, rec[u]);
    }
  });
}

// The pass A form: the prefetched record is the candidate's position.
template <int U, class T, class F>
__device__ __forceinline__ void walk_cells_batched(const KP<T>& k, const int* __restrict__ cstart,
                                                   const int* c, const V4<T>* __restrict__ pos,
                                                   F&& f) {
  walk_cells_prefetch<U>(k, cstart, c, [&](int j) { return pos[j]; }, f);
}
constexpr int kPassAU = 4;
constexpr int kStU = 4;  // prefetch depth of the surface-tension cell walks

template <class T>
__device__ __forceinline__ T rest_psi_of(const KP<T>& k, const T* rdens, int i) {
  return k.solver == TSPH_VFSPH ? T(1) : rdens[i];
}

// Iterate a particle's fluid neighbours: the recorded list, or, when the list
// overflowed, the cell walk with the same in-range test.
template <class T, class L, class F>
__device__ __forceinline__ void for_fluid_nbrs(const KP<T>& k, int i, int cap, int nf,
                                               const L& list,
                                               const int* __restrict__ cstart,
                                               const V4<T>* __restrict__ pos, V4<T> p, F&& f) {
  if (nf <= k.list_cap_f) {
    for (int s = 0; s < nf; ++s) f(list[(long)s * cap + i]);
  } else {
    int c[3];
    node_encode(k, p.x, p.y, p.z, c);
    walk_cells(k, cstart, c, [&](int j) {
      V4<T> q = pos[j];
      T r2 = s_r2(p.x - q.x, p.y - q.y, p.z - q.z);
      if (r2 > T(0) && r2 <= k.r2max) f(j);
    });
  }
}
template <class T, class F>
__device__ __forceinline__ void for_bound_nbrs(const KP<T>& k, int i, int cap, int nbb,
                                               const int* __restrict__ list,
                                               const int* __restrict__ cstart,
                                               const V4<T>* __restrict__ bpos, V4<T> p, F&& f) {
  if (nbb <= k.list_cap_b) {
    for (int s = 0; s < nbb; ++s) f(list[(long)s * cap + i]);
  } else {
    int c[3];
    node_encode(k, p.x, p.y, p.z, c);
    walk_cells(k, cstart, c, [&](int j) {
      V4<T> q = bpos[j];
      T r2 = s_r2(p.x - q.x, p.y - q.y, p.z - q.z);
      if (r2 > T(0) && r2 <= k.r2max) f(j);
    });
  }
}

// Neighbour iteration of the list sweeps for one lane, in the recorded
// (reference) order: fluid list, then boundary list.  The sweeps are bound
// by the latency of dependent loads (list entry -> neighbour data), so the
// lists are consumed in batches of kNb: all kNb indices are loaded, then all
// their neighbour records, then the pair terms are applied in order.  A lane
// whose list overflowed walks the cells with the same in-range test.
constexpr int kNb = 8;

// One list, software-pipelined: the next batch's list entries are loaded
// right after this batch's records are requested, so a batch costs one
// round trip to memory instead of two (index, then record).
template <class Rec, class Load, class Pair>
__device__ __forceinline__ void sweep_one_list(int cap, int i, const int* __restrict__ list, int n,
                                               Load&& load, Pair&& pair) {
  int j[kNb];
#pragma unroll
  for (int u = 0; u < kNb; ++u) j[u] = u < n ? list[(long)u * cap + i] : -1;
  for (int s0 = 0; s0 < n; s0 += kNb) {
    Rec d[kNb];
#pragma unroll
    for (int u = 0; u < kNb; ++u)
      if (j[u] >= 0) d[u] = load(j[u]);
    int jn[kNb];
#pragma unroll
    for (int u = 0; u < kNb; ++u) {
      const int s = s0 + kNb + u;
      jn[u] = s < n ? list[(long)s * cap + i] : -1;
    }
#pragma unroll
    for (int u = 0; u < kNb; ++u)
      if (j[u] >= 0) pair(d[u]);
#pragma unroll
    for (int u = 0; u < kNb; ++u) j[u] = jn[u];
  }
}

template <class T, class Rec, class LoadF, class LoadB, class PairF, class PairB>
__device__ __forceinline__ void sweep_lists(const KP<T>& k, int cap, int i, V4<T> p,
                                            const int* __restrict__ fstart,
                                            const int* __restrict__ bstart,
                                            const V4<T>* __restrict__ fpos,
                                            const V4<T>* __restrict__ bpos,
                                            const int* __restrict__ nbr_f,
                                            const int* __restrict__ nbr_b, int nf, int nbb,
                                            LoadF&& load_f, LoadB&& load_b, PairF&& pair_f,
                                            PairB&& pair_b) {
  if (nf <= k.list_cap_f && nbb <= k.list_cap_b) {
    sweep_one_list<Rec>(cap, i, nbr_f, nf, load_f, pair_f);
    if (nbb > 0) sweep_one_list<Rec>(cap, i, nbr_b, nbb, load_b, pair_b);
    return;
  }
  int c[3];
  node_encode(k, p.x, p.y, p.z, c);
  walk_cells(k, fstart, c, [&](int j) {
    V4<T> q = fpos[j];
    T r2 = s_r2(p.x - q.x, p.y - q.y, p.z - q.z);
    if (r2 > T(0) && r2 <= k.r2max) pair_f(load_f(j));
  });
  walk_cells(k, bstart, c, [&](int j) {
    V4<T> q = bpos[j];
    T r2 = s_r2(p.x - q.x, p.y - q.y, p.z - q.z);
    if (r2 > T(0) && r2 <= k.r2max) pair_b(load_b(j));
  });
}

template <class T>
struct Rec2 {
  V4<T> a, b;
};
template <class T>
struct Rec3 {
  V4<T> a, b, c;
};

// ---------------------------------------------------------------------------
// Cooperative list sweeps.  A group of kCoop adjacent lanes shares one
// particle; lane c of the group takes list slots c, c + kCoop, c + 2 kCoop,
// ...  At each step the group's lanes gather kCoop consecutive neighbours of
// the same particle, which lie in one or two cells, and the 64 lanes of a
// wave gather from about half as many cache sectors as with one lane per
// particle (whose lanes drift apart through their lists).  The sweeps are
// bound by the number of L1 sectors their gathers touch, not by the number
// of gathers.  The group's partial sums are combined by a butterfly, which
// leaves the same value in every lane of the group.
// ---------------------------------------------------------------------------
constexpr int kCoop = 4;    // lanes per particle of the gather-heavy sweeps
constexpr int kCoopU = 2;   // lanes per particle of the update sweep (one gather per pair)
constexpr int kCoopNb = 4;  // gathers in flight per lane

// A block covers its kBlock particles in C rounds of kBlock / C particles.
template <int C>
__device__ __forceinline__ int coop_lane() {
  return threadIdx.x & (C - 1);
}
template <int C>
__device__ __forceinline__ int coop_particle(int base, int round) {
  return base + round * (kBlock / C) + (int)(threadIdx.x / C);
}

template <int C, class T>
__device__ __forceinline__ T group_sum(T v) {
#pragma unroll
  for (int o = 1; o < C; o <<= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// One list, lane c's share, software-pipelined like sweep_one_list, NB
// gathers in flight per lane.
template <int C, class Rec, int NB = kCoopNb, class Load, class Pair>
__device__ __forceinline__ void sweep_one_list_coop(int cap, int i, int c,
                                                    const int* __restrict__ list, int n,
                                                    Load&& load, Pair&& pair) {
  int j[NB];
#pragma unroll
  for (int u = 0; u < NB; ++u) {
    const int s = c + C * u;
    j[u] = s < n ? list[(long)s * cap + i] : -1;
  }
  for (int s0 = c; s0 < n; s0 += C * NB) {
    Rec d[NB];
#pragma unroll
    for (int u = 0; u < NB; ++u)
      if (j[u] >= 0) d[u] = load(j[u]);
    int jn[NB];
#pragma unroll
    for (int u = 0; u < NB; ++u) {
      const int s = s0 + C * (NB + u);
      jn[u] = s < n ? list[(long)s * cap + i] : -1;
    }
#pragma unroll
    for (int u = 0; u < NB; ++u)
      if (j[u] >= 0) pair(d[u]);
#pragma unroll
    for (int u = 0; u < NB; ++u) j[u] = jn[u];
  }
}

// sweep_lists for lane c of a particle's group.  A particle whose list
// overflowed is swept by lane 0 alone, walking the cells.
template <int C, class T, class Rec, int NB = kCoopNb, class LoadF, class LoadB, class PairF,
          class PairB>
__device__ __forceinline__ void sweep_lists_coop(const KP<T>& k, int cap, int i, int c, V4<T> p,
                                                 const int* __restrict__ fstart,
                                                 const int* __restrict__ bstart,
                                                 const V4<T>* __restrict__ fpos,
                                                 const V4<T>* __restrict__ bpos,
                                                 const int* __restrict__ nbr_f,
                                                 const int* __restrict__ nbr_b, int nf, int nbb,
                                                 LoadF&& load_f, LoadB&& load_b, PairF&& pair_f,
                                                 PairB&& pair_b) {
  if (nf <= k.list_cap_f && nbb <= k.list_cap_b) {
    sweep_one_list_coop<C, Rec, NB>(cap, i, c, nbr_f, nf, load_f, pair_f);
    if (nbb > 0) sweep_one_list_coop<C, Rec, NB>(cap, i, c, nbr_b, nbb, load_b, pair_b);
    return;
  }
  if (c != 0) return;
  int cc[3];
  node_encode(k, p.x, p.y, p.z, cc);
  walk_cells(k, fstart, cc, [&](int j) {
    V4<T> q = fpos[j];
    T r2 = s_r2(p.x - q.x, p.y - q.y, p.z - q.z);
    if (r2 > T(0) && r2 <= k.r2max) pair_f(load_f(j));
  });
  if (nbr_b == nullptr) return;
  walk_cells(k, bstart, cc, [&](int j) {
    V4<T> q = bpos[j];
    T r2 = s_r2(p.x - q.x, p.y - q.y, p.z - q.z);
    if (r2 > T(0) && r2 <= k.r2max) pair_b(load_b(j));
  });
}

}  // namespace tsph
