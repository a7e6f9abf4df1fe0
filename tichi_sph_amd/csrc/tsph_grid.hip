// Neighbour grid kernels.
//
// Fused path (tsph_sort_object): a counting sort that moves one particle
// object into cell order every step.
//   keys      cell key per particle (node_encode + dim_encode,
//             sph_util.py:65-72), arrival rank via an atomic per cell
//             (encode, sph_obj.py:300-306); optional cfl_condition partials
//   scan      exclusive prefix sum of the per-cell counts (mem_shift,
//             sph_obj.py:308-314, in its single-thread meaning)
//   scatter   provisional slot = start[key] + arrival rank
//   permute   final slot = start[key] + rank of the particle's reference
//             index among its cell-mates (fill_node's single-thread order,
//             sph_obj.py:316-323), then the persistent state is gathered
//             into the back buffers.
// Reference layout (tsph_ngrid_build): the reference's shared Ngrid arrays
// over unsorted storage, same algorithm over (cell, uid) buckets.
#include <mutex>

#include "tsph_host.h"

namespace tsph {

// ---------------------------------------------------------------------------
// Scan: exclusive prefix of `cnt[0..len)` into `start[0..len]`.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void k_scan_reduce_body(Blk blk, const int* __restrict__ cnt, long len,
                                                        int* __restrict__ sums) {
  __shared__ double lds[kWaves];
  long base = (long)blk.b * kScanTile;
  long s = 0;
  for (int e = 0; e < 16; ++e) {
    long i = base + (long)e * kBlock + threadIdx.x;
    if (i < len) s += cnt[i];
  }
  double t = block_sum((double)s, lds);
  if (threadIdx.x == 0) sums[blk.b] = (int)t;
}

__global__ void __launch_bounds__(kBlock)
    k_scan_reduce(const int* __restrict__ cnt, long len,
                                                        int* __restrict__ sums) {
  k_scan_reduce_body(Blk{(int)blockIdx.x, (int)gridDim.x}, cnt, len, sums);
}

__device__ __forceinline__ void k_scan_apply_body(Blk blk, const int* __restrict__ cnt, long len,
                                                       const int* __restrict__ sums,
                                                       int* __restrict__ start) {
  __shared__ int wsum[kWaves];
  __shared__ int boff;
  // offset of this tile = sum of previous tile sums
  if (threadIdx.x < 64) {
    int acc = 0;
    for (int i = threadIdx.x; i < (int)blk.b; i += 64) acc += sums[i];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if (threadIdx.x == 0) boff = acc;
  }
  long base = (long)blk.b * kScanTile + (long)threadIdx.x * 16;
  int v[16];
  int loc = 0;
#pragma unroll
  for (int e = 0; e < 16; ++e) {
    long i = base + e;
    v[e] = i < len ? cnt[i] : 0;
    loc += v[e];
  }
  // inclusive scan of per-thread totals across the block
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int x = loc;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[w] = x;
  __syncthreads();
  int wo = 0;
  for (int i = 0; i < w; ++i) wo += wsum[i];
  int run = boff + wo + x - loc;  // exclusive prefix of this thread's first element
#pragma unroll
  for (int e = 0; e < 16; ++e) {
    long i = base + e;
    if (i < len) start[i] = run;
    run += v[e];
  }
  // start[len] = total, written by the thread owning the last element
  if (base <= len - 1 && len - 1 < base + 16) start[len] = run;
}

__global__ void __launch_bounds__(kBlock)
    k_scan_apply(const int* __restrict__ cnt, long len,
                                                       const int* __restrict__ sums,
                                                       int* __restrict__ start) {
  k_scan_apply_body(Blk{(int)blockIdx.x, (int)gridDim.x}, cnt, len, sums, start);
}

// The scan's first half with cfl_condition's finish (sph.py:41-47) in its
// block 0: the CFL partials of k_keys are reduced without a launch of their
// own (dt is first read by the solver, after the sort).
template <class T>
__global__ void __launch_bounds__(kBlock)
    k_scan_reduce_cfl(const int* __restrict__ cnt, long len, int* __restrict__ sums, KP<T> k,
                      const double* __restrict__ partial, int nblk, int n, TsphControl* ctrl) {
  const Blk blk{(int)blockIdx.x, (int)gridDim.x};
  k_scan_reduce_body(blk, cnt, len, sums);
  if (blk.b == 0) k_finalize_body<T>(blk, k, (int)kFinCfl, partial, nblk, n, ctrl);
}

static int launch_scan(const int* cnt, long len, int* sums, int* start, hipStream_t st) {
  long nb = scan_blocks(len);
  if (nb < 1) nb = 1;
  hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(kBlock), 0, st, cnt, len, sums);
  hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(kBlock), 0, st, cnt, len, sums, start);
  TSPH_LAUNCH_CHECK("scan");
  return 0;
}

__global__ void __launch_bounds__(kBlock) k_fill_i32(int* __restrict__ p, long n, int v) {
  long i = (long)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------
// Fused-path sort.
// ---------------------------------------------------------------------------
template <class T>
__device__ __forceinline__ void k_keys_body(Blk blk, KP<T> k, const V4<T>* __restrict__ pos, const V4<T>* __restrict__ vel, int n,
           int* __restrict__ key, int* __restrict__ rank, int* __restrict__ count, int cfl,
           double* __restrict__ partial, TsphControl* ctrl) {
  __shared__ double lds[kWaves];
  int nb = blk.nb;
  int b = xcd_block(blk.b, nb);
  int i = b * kBlock + threadIdx.x;
  double cand = 1e300;
  const int lane = threadIdx.x & 63;
  int kk = -1 - lane;  // inactive lanes: unique keys that never merge
  V4<T> p{};
  if (i < n) {
    p = pos[i];
    int c[3];
    node_encode(k, p.x, p.y, p.z, c);
    int code = dim_encode(k, c[0], c[1], c[2]);
    kk = (0 < code && code < k.node_num) ? code : k.node_num;  // invalid -> trailing bucket
    key[i] = kk;
  }
  // Arrival ranks: storage is in last step's cell order, so equal keys sit
  // in adjacent lanes; each run of equal keys takes one atomic for its length.
  {
    int left = __shfl_up(kk, 1, 64);
    bool lead = lane == 0 || left != kk;
    unsigned long long leaders = __ballot(lead);
    unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1);
    int start = 63 - __clzll(leaders & upto);
    unsigned long long after = start == 63 ? 0ull : (leaders & ~((2ull << start) - 1));
    int end = after ? __ffsll((long long)after) - 1 : 64;
    int base = 0;
    if (lead && i < n) base = atomicAdd(&count[kk], end - start);
    base = __shfl(base, start, 64);
    if (i < n) rank[i] = base + (lane - start);
  }
  if (i < n) {
    if (cfl) {  // cfl_condition, sph.py:44-47
      TSPH_STRICT
      V4<T> v = vel[i];
      T vn = s_sqrt(s_r2(v.x, v.y, v.z));
      if (vn > T(1e-4)) {
        T q = k.gps / vn;
        cand = (double)(q * k.cfl);
      }
    }
  }
  if (!cfl) return;
  double m = block_min(cand, lds);
  if (threadIdx.x == 0) partial[blk.b] = m;
}

template <class T>
__global__ void __launch_bounds__(kBlock)
    k_keys(KP<T> k, const V4<T>* __restrict__ pos, const V4<T>* __restrict__ vel, int n,
           int* __restrict__ key, int* __restrict__ rank, int* __restrict__ count, int cfl,
           double* __restrict__ partial, TsphControl* ctrl) {
  k_keys_body<T>(Blk{(int)blockIdx.x, (int)gridDim.x}, k, pos, vel, n, key, rank, count, cfl, partial, ctrl);
}

__device__ __forceinline__ void k_scatter_body(Blk blk, const int* __restrict__ key, const int* __restrict__ rank,
              const int* __restrict__ start, const int* __restrict__ pid, int n,
              int* __restrict__ tmp_idx, int* __restrict__ tmp_pid, int* __restrict__ tmp_key) {
  int i = xcd_block(blk.b, blk.nb) * kBlock + threadIdx.x;
  if (i >= n) return;
  int kk = key[i];
  int s = start[kk] + rank[i];
  tmp_idx[s] = i;
  tmp_pid[s] = pid[i];
  tmp_key[s] = kk;
}

__global__ void __launch_bounds__(kBlock)
    k_scatter(const int* __restrict__ key, const int* __restrict__ rank,
              const int* __restrict__ start, const int* __restrict__ pid, int n,
              int* __restrict__ tmp_idx, int* __restrict__ tmp_pid, int* __restrict__ tmp_key) {
  k_scatter_body(Blk{(int)blockIdx.x, (int)gridDim.x}, key, rank, start, pid, n, tmp_idx, tmp_pid, tmp_key);
}

template <class T>
__device__ __forceinline__ void k_permute_body(Blk blk, int P, int n, const int* __restrict__ start, const int* __restrict__ tmp_idx,
              const int* __restrict__ tmp_pid, const int* __restrict__ tmp_key,
              const V4<T>* __restrict__ pos, const V4<T>* __restrict__ vel,
              const T* __restrict__ mass, const T* __restrict__ rdens, const T* __restrict__ rvol,
              const T* __restrict__ vf, const int* __restrict__ pid, V4<T>* __restrict__ pos_b,
              V4<T>* __restrict__ vel_b, T* __restrict__ mass_b, T* __restrict__ rdens_b,
              T* __restrict__ rvol_b, T* __restrict__ vf_b, int* __restrict__ pid_b,
              int* __restrict__ cell) {
  int s = xcd_block(blk.b, blk.nb) * kBlock + threadIdx.x;
  if (s >= n) return;
  int c = tmp_key[s];
  int a = start[c], e = start[c + 1];
  int p = tmp_pid[s];
  int r = 0;
  for (int t = a; t < e; ++t) r += tmp_pid[t] < p;
  int dst = a + r;
  int src = tmp_idx[s];
  pos_b[dst] = pos[src];
  vel_b[dst] = vel[src];
  mass_b[dst] = mass[src];
  rdens_b[dst] = rdens[src];
  rvol_b[dst] = rvol[src];
  for (int q = 0; q < P; ++q) vf_b[(long)dst * P + q] = vf[(long)src * P + q];
  pid_b[dst] = p;
  cell[dst] = c;
}

template <class T>
__global__ void __launch_bounds__(kBlock)
    k_permute(int P, int n, const int* __restrict__ start, const int* __restrict__ tmp_idx,
              const int* __restrict__ tmp_pid, const int* __restrict__ tmp_key,
              const V4<T>* __restrict__ pos, const V4<T>* __restrict__ vel,
              const T* __restrict__ mass, const T* __restrict__ rdens, const T* __restrict__ rvol,
              const T* __restrict__ vf, const int* __restrict__ pid, V4<T>* __restrict__ pos_b,
              V4<T>* __restrict__ vel_b, T* __restrict__ mass_b, T* __restrict__ rdens_b,
              T* __restrict__ rvol_b, T* __restrict__ vf_b, int* __restrict__ pid_b,
              int* __restrict__ cell) {
  k_permute_body<T>(Blk{(int)blockIdx.x, (int)gridDim.x}, P, n, start, tmp_idx, tmp_pid, tmp_key, pos, vel, mass, rdens, rvol, vf, pid, pos_b, vel_b, mass_b, rdens_b, rvol_b, vf_b, pid_b, cell);
}

static void swap_state(TsphObject* o) {
  std::swap(o->pos, o->pos_b);
  std::swap(o->vel, o->vel_b);
  std::swap(o->mass, o->mass_b);
  std::swap(o->rest_density, o->rest_density_b);
  std::swap(o->rest_volume, o->rest_volume_b);
  std::swap(o->volume_frac, o->volume_frac_b);
  std::swap(o->pid, o->pid_b);
}

// The scratch of one counting sort: CFL block partials, scan tile sums, and
// per-particle keys, arrival ranks and the scatter's provisional slots.
// Every array is written before it is read within one sort, so nothing
// needs clearing between calls.
struct SortScratch {
  double* partial;  // [2 * (blocks_for(n) + 8)]
  int* scan_sums;   // [scan_blocks(node_num + 2) + 8]
  int *key, *rank, *tmp_idx, *tmp_pid, *tmp_key;  // [n]
};

template <class T>
static int sort_object_core(const TsphParams& prm, TsphObject* o, TsphControl* ctrl, int cfl,
                            const SortScratch& L, hipStream_t st) {
  KP<T> k = make_kp<T>(prm);
  int n = o->n;
  o->flags &= ~TSPH_OBJ_STATIC_SUMS;  // new storage order: the kept sums no longer apply
  long len = (long)prm.node_num + 1;  // cells + trailing invalid bucket
  hipLaunchKernelGGL(k_fill_i32, dim3(blocks_for(len + 1)), dim3(kBlock), 0, st, o->cell_count,
                     len + 1, 0);
  int nb = blocks_for(n);
  const int nbl = nb > 0 ? nb : 1;  // k_keys' blocks (one also for n = 0 with CFL)
  if (n > 0 || cfl)
    hipLaunchKernelGGL(k_keys<T>, dim3(nbl), dim3(kBlock), 0, st, k, (const V4<T>*)o->pos,
                       (const V4<T>*)o->vel, n, L.key, L.rank, o->cell_count, cfl, L.partial,
                       ctrl);
  TSPH_LAUNCH_CHECK("keys");
  if (cfl) {  // the scan, with the CFL partials reduced in its first kernel
    long nbs = scan_blocks(len);
    if (nbs < 1) nbs = 1;
    hipLaunchKernelGGL(k_scan_reduce_cfl<T>, dim3(nbs), dim3(kBlock), 0, st, o->cell_count, len,
                       L.scan_sums, k, L.partial, nbl, n, ctrl);
    hipLaunchKernelGGL(k_scan_apply, dim3(nbs), dim3(kBlock), 0, st, o->cell_count, len,
                       L.scan_sums, o->cell_start);
    TSPH_LAUNCH_CHECK("scan");
  } else if (int rc = launch_scan(o->cell_count, len, L.scan_sums, o->cell_start, st)) {
    return rc;
  }
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_scatter, dim3(nb), dim3(kBlock), 0, st, L.key, L.rank, o->cell_start,
                     o->pid, n, L.tmp_idx, L.tmp_pid, L.tmp_key);
  hipLaunchKernelGGL(k_permute<T>, dim3(nb), dim3(kBlock), 0, st, prm.phase_num, n,
                     o->cell_start, L.tmp_idx, L.tmp_pid, L.tmp_key, (const V4<T>*)o->pos,
                     (const V4<T>*)o->vel, (const T*)o->mass, (const T*)o->rest_density,
                     (const T*)o->rest_volume, (const T*)o->volume_frac, o->pid,
                     (V4<T>*)o->pos_b, (V4<T>*)o->vel_b, (T*)o->mass_b, (T*)o->rest_density_b,
                     (T*)o->rest_volume_b, (T*)o->volume_frac_b, o->pid_b, o->cell);
  TSPH_LAUNCH_CHECK("permute");
  swap_state(o);
  return 0;
}

template <class T>
int sort_object(const TsphParams& prm, TsphObject* o, TsphControl* ctrl, int cfl, void* ws,
                size_t wsb, hipStream_t st) {
  WsLayout W;
  TSPH_REQUIRE(carve_ws(ws, wsb, o->capacity, prm.node_num, &W), TSPH_ECAPACITY,
               "tsph_sort_object: workspace too small (%zu bytes)", wsb);
  const SortScratch L{W.partial, W.scan_sums, W.key, W.rank, W.tmp_idx, W.tmp_pid, W.tmp_key};
  return sort_object_core<T>(prm, o, ctrl, cfl, L, st);
}

// ---------------------------------------------------------------------------
// tsph_sort_object_static: the same sort with its scratch in a file-scope
// device array instead of the caller's workspace, for callers that cannot
// allocate.  Each device holds its own copy of `scratch` (the runtime places
// it when this code object is loaded on the device, for the life of the
// process); its device address is looked up once per device.  Calls that
// overlap on one device share the array and therefore race.
// ---------------------------------------------------------------------------
constexpr long kStaticSortMaxN = 1L << 21;      // particles of the object
constexpr long kStaticSortMaxCells = 1L << 24;  // node_num
constexpr size_t cseg(size_t b) { return (b + 255) & ~(size_t)255; }
constexpr size_t kStaticPartialBytes = cseg(sizeof(double) * 2 * (kStaticSortMaxN / kBlock + 8));
constexpr size_t kStaticScanBytes =
    cseg(sizeof(int) * ((kStaticSortMaxCells + 2 + kScanTile - 1) / kScanTile + 8));
constexpr size_t kStaticIdxBytes = cseg(sizeof(int) * kStaticSortMaxN);
constexpr size_t kStaticScratchBytes = kStaticPartialBytes + kStaticScanBytes + 5 * kStaticIdxBytes;

static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

constexpr int kStaticMaxDevices = 64;
static void* g_scratch_dev[kStaticMaxDevices];
static std::mutex g_scratch_mu;

// The device address of `scratch` on the current device (host code must not
// use the array's name as a device address: it is a host-side stand-in).
static int static_scratch(void** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  TSPH_REQUIRE(e == hipSuccess, TSPH_EHIP, "tsph_sort_object_static: hipGetDevice: %s",
               hipGetErrorString(e));
  TSPH_REQUIRE(dev >= 0 && dev < kStaticMaxDevices, TSPH_EUNSUPPORTED,
               "tsph_sort_object_static: device %d beyond the %d supported", dev, kStaticMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    TSPH_REQUIRE(e == hipSuccess && ptr != nullptr, TSPH_EHIP,
                 "tsph_sort_object_static: hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_scratch_dev[dev] = ptr;
  }
  *out = g_scratch_dev[dev];
  return 0;
}

size_t static_sort_scratch_bytes() { return kStaticScratchBytes; }
long static_sort_max_n() { return kStaticSortMaxN; }
long static_sort_max_cells() { return kStaticSortMaxCells; }

template <class T>
int sort_object_static(const TsphParams& prm, TsphObject* o, TsphControl* ctrl, int cfl,
                       hipStream_t st) {
  TSPH_REQUIRE(o->n <= kStaticSortMaxN, TSPH_ECAPACITY,
               "tsph_sort_object_static: %d particles exceed the static scratch's maximum of %ld",
               o->n, kStaticSortMaxN);
  TSPH_REQUIRE((long)prm.node_num <= kStaticSortMaxCells, TSPH_ECAPACITY,
               "tsph_sort_object_static: node_num %d exceeds the static scratch's maximum of %ld",
               prm.node_num, kStaticSortMaxCells);
  void* base = nullptr;
  if (int rc = static_scratch(&base)) return rc;
  char* p = (char*)base;
  SortScratch L;
  L.partial = (double*)p;
  p += kStaticPartialBytes;
  L.scan_sums = (int*)p;
  p += kStaticScanBytes;
  int** idx[5] = {&L.key, &L.rank, &L.tmp_idx, &L.tmp_pid, &L.tmp_key};
  for (int q = 0; q < 5; ++q, p += kStaticIdxBytes) *idx[q] = (int*)p;
  return sort_object_core<T>(prm, o, ctrl, cfl, L, st);
}

// ---------------------------------------------------------------------------
// Reference-layout Ngrid (shared by all objects, sph_obj.py:280-323).
// ---------------------------------------------------------------------------
template <class T>
__global__ void __launch_bounds__(kBlock)
    k_ref_encode(KP<T> k, const V4<T>* __restrict__ pos, int n, int u, int nobj,
                 int* __restrict__ cell, int* __restrict__ cell_structured,
                 int* __restrict__ rank, int* __restrict__ cu_count) {
  int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  V4<T> p = pos[i];
  int c[3];
  node_encode(k, p.x, p.y, p.z, c);
  int code = dim_encode(k, c[0], c[1], c[2]);
  cell[i] = code;
  cell_structured[4 * i + 0] = c[0];
  cell_structured[4 * i + 1] = c[1];
  cell_structured[4 * i + 2] = c[2];
  cell_structured[4 * i + 3] = 0;
  if (0 < code && code < k.node_num) rank[i] = atomicAdd(&cu_count[(long)code * nobj + u], 1);
}

__global__ void __launch_bounds__(kBlock)
    k_ref_cells(int node_num, int nobj, const int* __restrict__ cu_count,
                const int* __restrict__ cu_start, int* __restrict__ npc, int* __restrict__ nps,
                int* __restrict__ npsc) {
  int c = blockIdx.x * kBlock + threadIdx.x;
  if (c >= node_num) return;
  int s = 0;
  for (int u = 0; u < nobj; ++u) s += cu_count[(long)c * nobj + u];
  npc[c] = s;
  nps[c] = cu_start[(long)c * nobj];
  npsc[c] = cu_start[(long)c * nobj] + s;  // shift_count after fill_node
}

__global__ void __launch_bounds__(kBlock)
    k_ref_scatter(int node_num, const int* __restrict__ cell, const int* __restrict__ rank,
                  const int* __restrict__ pid, int n, int u, int nobj,
                  const int* __restrict__ cu_start, int* __restrict__ tmp_idx,
                  int* __restrict__ tmp_pid, int* __restrict__ tmp_key) {
  int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  int code = cell[i];
  if (!(0 < code && code < node_num)) return;
  long b = (long)code * nobj + u;
  int s = cu_start[b] + rank[i];
  tmp_idx[s] = i;
  tmp_pid[s] = pid[i];
  tmp_key[s] = (int)b;
}

__global__ void __launch_bounds__(kBlock)
    k_ref_fill(int cap, const int* __restrict__ total_ptr, const int* __restrict__ cu_start, const int* __restrict__ tmp_idx,
               const int* __restrict__ tmp_pid, const int* __restrict__ tmp_key, int nobj,
               int* __restrict__ pid_in_node, int* __restrict__ uid_in_node, int* in_cell0,
               int* in_cell1, int* in_cell2, int* in_cell3) {
  int s = blockIdx.x * kBlock + threadIdx.x;
  if (s >= cap || s >= *total_ptr) return;
  int b = tmp_key[s];
  int a = cu_start[b], e = cu_start[b + 1];
  int p = tmp_pid[s];
  int r = 0;
  for (int t = a; t < e; ++t) r += tmp_pid[t] < p;
  int dst = a + r;
  int u = b % nobj;
  int i = tmp_idx[s];
  pid_in_node[dst] = i;
  uid_in_node[dst] = u;
  int* ic = u == 0 ? in_cell0 : u == 1 ? in_cell1 : u == 2 ? in_cell2 : in_cell3;
  ic[i] = dst;
}

template <class T>
int ngrid_build(const TsphParams& prm, TsphObject* const* objs, int nobj, int* npc, int* nps,
                int* npsc, int* pin, int* uin, void* ws, size_t wsb, hipStream_t st) {
  TSPH_REQUIRE(nobj >= 1 && nobj <= 2, TSPH_EINVAL, "tsph_ngrid_build: 1 or 2 objects");
  KP<T> k = make_kp<T>(prm);
  long cap = 0, total_cap = 0;
  for (int u = 0; u < nobj; ++u) {
    TSPH_REQUIRE(objs[u]->uid == u, TSPH_EINVAL, "tsph_ngrid_build: object %d has uid %d", u,
                 objs[u]->uid);
    cap = std::max<long>(cap, objs[u]->capacity);
    total_cap += objs[u]->capacity;
  }
  WsLayout L;
  TSPH_REQUIRE(carve_ws(ws, wsb, total_cap, prm.node_num, &L), TSPH_ECAPACITY,
               "tsph_ngrid_build: workspace too small (%zu bytes)", wsb);
  long lenref = (long)prm.node_num * nobj;
  hipLaunchKernelGGL(k_fill_i32, dim3(blocks_for(lenref + 1)), dim3(kBlock), 0, st, L.cu_count,
                     lenref + 1, 0);
  // encode (sph_obj.py:300-306); rank temporaries are per object
  long roff = 0;
  for (int u = 0; u < nobj; ++u) {
    TsphObject* o = objs[u];
    if (o->n > 0)
      hipLaunchKernelGGL(k_ref_encode<T>, dim3(blocks_for(o->n)), dim3(kBlock), 0, st, k,
                         (const V4<T>*)o->pos, o->n, u, nobj, o->cell, o->cell_structured,
                         L.rank + roff, L.cu_count);
    roff += o->capacity;
  }
  TSPH_LAUNCH_CHECK("ngrid encode");
  // mem_shift (sph_obj.py:308-314)
  if (int rc = launch_scan(L.cu_count, lenref, L.scan_sums, L.cu_start, st)) return rc;
  hipLaunchKernelGGL(k_ref_cells, dim3(blocks_for(prm.node_num)), dim3(kBlock), 0, st,
                     prm.node_num, nobj, L.cu_count, L.cu_start, npc, nps, npsc);
  // fill_node (sph_obj.py:316-323)
  roff = 0;
  for (int u = 0; u < nobj; ++u) {
    TsphObject* o = objs[u];
    if (o->n > 0)
      hipLaunchKernelGGL(k_ref_scatter, dim3(blocks_for(o->n)), dim3(kBlock), 0, st,
                         prm.node_num, o->cell, L.rank + roff, o->pid, o->n, u, nobj, L.cu_start,
                         L.tmp_idx, L.tmp_pid, L.tmp_key);
    roff += o->capacity;
  }
  // total entries = cu_start[lenref]; launch over total capacity, extra threads exit
  long tc = total_cap;
  int* ic[4] = {nullptr, nullptr, nullptr, nullptr};
  for (int u = 0; u < nobj; ++u) ic[u] = objs[u]->in_cell;
  hipLaunchKernelGGL(k_ref_fill, dim3(blocks_for(tc)), dim3(kBlock), 0, st, (int)tc,
                     L.cu_start + lenref, L.cu_start, L.tmp_idx, L.tmp_pid, L.tmp_key, nobj, pin,
                     uin, ic[0], ic[1], ic[2], ic[3]);
  TSPH_LAUNCH_CHECK("ngrid fill");
  return 0;
}

// ---------------------------------------------------------------------------
// Grouped (batched) sort: the kernels above over every scene of a table.
// ---------------------------------------------------------------------------
template <class T>
__global__ void __launch_bounds__(kBlock) kb_fill_counts(const char* __restrict__ tab, int ns) {
  for_tiles<T, kTileCell>(tab, ns, [&](const SceneEntry<T>& e, int, Blk blk) {
    const long n = sort_len(e.k.node_num) + 1;  // cell_count[node_num + 2]
    const long base = (long)blk.b * kScanTile;
    for (int q = 0; q < 16; ++q) {
      long i = base + (long)q * kBlock + threadIdx.x;
      if (i < n) e.f.cell_count[i] = 0;
    }
  });
}

template <class T>
__global__ void __launch_bounds__(kBlock) kb_keys(const char* __restrict__ tab, int ns, int cfl) {
  for_tiles<T, kTileFluid>(tab, ns, [&](const SceneEntry<T>& e, int, Blk blk) {
    k_keys_body<T>(blk, e.k, (const V4<T>*)e.f.pos, (const V4<T>*)e.f.vel, e.f.n, e.L.key, e.L.rank,
                   e.f.cell_count, cfl, e.L.partial, e.ctrl);
  });
}

template <class T>
__global__ void __launch_bounds__(kBlock) kb_finalize_cfl(const char* __restrict__ tab, int ns) {
  const int* firstf = (const int*)(tab + batch_entries_bytes<T>(ns)) + kTileFluid * (ns + 1);
  for_tiles<T, kTileScene>(tab, ns, [&](const SceneEntry<T>& e, int s, Blk blk) {
    k_finalize_body<T>(blk, e.k, (int)kFinCfl, e.L.partial, firstf[s + 1] - firstf[s], e.f.n,
                       e.ctrl);
  });
}

template <class T>
__global__ void __launch_bounds__(kBlock) kb_scan_reduce(const char* __restrict__ tab, int ns) {
  for_tiles<T, kTileCell>(tab, ns, [&](const SceneEntry<T>& e, int, Blk blk) {
    k_scan_reduce_body(blk, e.f.cell_count, sort_len(e.k.node_num), e.L.scan_sums);
  });
}

template <class T>
__global__ void __launch_bounds__(kBlock) kb_scan_apply(const char* __restrict__ tab, int ns) {
  for_tiles<T, kTileCell>(tab, ns, [&](const SceneEntry<T>& e, int, Blk blk) {
    k_scan_apply_body(blk, e.f.cell_count, sort_len(e.k.node_num), e.L.scan_sums,
                      e.f.cell_start);
  });
}

template <class T>
__global__ void __launch_bounds__(kBlock) kb_scatter(const char* __restrict__ tab, int ns) {
  for_tiles<T, kTileFluid>(tab, ns, [&](const SceneEntry<T>& e, int, Blk blk) {
    k_scatter_body(blk, e.L.key, e.L.rank, e.f.cell_start, e.f.pid, e.f.n, e.L.tmp_idx,
                   e.L.tmp_pid, e.L.tmp_key);
  });
}

template <class T>
__global__ void __launch_bounds__(kBlock) kb_permute(const char* __restrict__ tab, int ns) {
  for_tiles<T, kTileFluid>(tab, ns, [&](const SceneEntry<T>& e, int, Blk blk) {
    const TsphObject& o = e.f;
    k_permute_body<T>(blk, e.k.P, o.n, o.cell_start, e.L.tmp_idx, e.L.tmp_pid, e.L.tmp_key,
                      (const V4<T>*)o.pos, (const V4<T>*)o.vel, (const T*)o.mass,
                      (const T*)o.rest_density, (const T*)o.rest_volume,
                      (const T*)o.volume_frac, o.pid, (V4<T>*)o.pos_b, (V4<T>*)o.vel_b,
                      (T*)o.mass_b, (T*)o.rest_density_b, (T*)o.rest_volume_b,
                      (T*)o.volume_frac_b, o.pid_b, o.cell);
  });
}

// Front/back swap of every scene's persistent state inside the table.
template <class T>
__global__ void kb_swap(char* tab, int ns) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ns) return;
  TsphObject& o = ((SceneEntry<T>*)tab)[s].f;
  void* t;
#define TSPH_SWAP(a, b) t = (void*)o.a; o.a = decltype(o.a)(o.b); o.b = decltype(o.b)(t);
  TSPH_SWAP(pos, pos_b)
  TSPH_SWAP(vel, vel_b)
  TSPH_SWAP(mass, mass_b)
  TSPH_SWAP(rest_density, rest_density_b)
  TSPH_SWAP(rest_volume, rest_volume_b)
  TSPH_SWAP(volume_frac, volume_frac_b)
  TSPH_SWAP(pid, pid_b)
#undef TSPH_SWAP
}

template <class T>
int batch_sort(const TsphBatchInfo& bi, char* tab, TsphObject* hostf, int cfl, int grid,
               hipStream_t st) {
  const int ns = bi.nscenes;
  const int gf = grid_for(bi.tiles[kTileFluid], grid), gc = grid_for(bi.tiles[kTileCell], grid);
  hipLaunchKernelGGL(kb_fill_counts<T>, dim3(gc), dim3(kBlock), 0, st, tab, ns);
  if (gf) hipLaunchKernelGGL(kb_keys<T>, dim3(gf), dim3(kBlock), 0, st, tab, ns, cfl);
  if (cfl)
    hipLaunchKernelGGL(kb_finalize_cfl<T>, dim3(ns), dim3(kBlock), 0, st, tab, ns);
  hipLaunchKernelGGL(kb_scan_reduce<T>, dim3(gc), dim3(kBlock), 0, st, tab, ns);
  hipLaunchKernelGGL(kb_scan_apply<T>, dim3(gc), dim3(kBlock), 0, st, tab, ns);
  if (gf) {
    hipLaunchKernelGGL(kb_scatter<T>, dim3(gf), dim3(kBlock), 0, st, tab, ns);
    hipLaunchKernelGGL(kb_permute<T>, dim3(gf), dim3(kBlock), 0, st, tab, ns);
  }
  hipLaunchKernelGGL(kb_swap<T>, dim3((ns + 63) / 64), dim3(64), 0, st, tab, ns);
  TSPH_LAUNCH_CHECK("batch_sort");
  if (hostf)
    for (int s = 0; s < ns; ++s) swap_state(&hostf[s]);
  return 0;
}

template int batch_sort<float>(const TsphBatchInfo&, char*, TsphObject*, int, int, hipStream_t);
template int batch_sort<double>(const TsphBatchInfo&, char*, TsphObject*, int, int, hipStream_t);

template int sort_object<float>(const TsphParams&, TsphObject*, TsphControl*, int, void*, size_t, hipStream_t);
template int sort_object<double>(const TsphParams&, TsphObject*, TsphControl*, int, void*, size_t, hipStream_t);
template int sort_object_static<float>(const TsphParams&, TsphObject*, TsphControl*, int, hipStream_t);
template int sort_object_static<double>(const TsphParams&, TsphObject*, TsphControl*, int, hipStream_t);
template int ngrid_build<float>(const TsphParams&, TsphObject* const*, int, int*, int*, int*, int*, int*, void*, size_t, hipStream_t);
template int ngrid_build<double>(const TsphParams&, TsphObject* const*, int, int*, int*, int*, int*, int*, void*, size_t, hipStream_t);

}  // namespace tsph
