// extern "C" entry points declared in include/tichi_sph.h.
#include <cstdarg>

#include <cstring>

#include "tsph_host.h"

namespace tsph {

static thread_local std::string g_err;

void set_error(const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
}

template <class T>
int sort_object(const TsphParams&, TsphObject*, TsphControl*, int, void*, size_t, hipStream_t);
template <class T>
int sort_object_static(const TsphParams&, TsphObject*, TsphControl*, int, hipStream_t);
size_t static_sort_scratch_bytes();
long static_sort_max_n();
long static_sort_max_cells();
template <class T>
int ngrid_build(const TsphParams&, TsphObject* const*, int, int*, int*, int*, int*, int*, void*,
                size_t, hipStream_t);
template <class T>
int density_alpha(const TsphParams&, TsphObject*, TsphObject*, TsphControl*, void*, size_t,
                  hipStream_t);
template <class T>
int solver_begin(const TsphParams&, TsphControl*, int, int*, int, hipStream_t);
int verify_fast_quotient(const TsphParams&, int*, hipStream_t);
template <class T>
int solver_half(const TsphParams&, TsphObject*, TsphObject*, TsphControl*, int, int, void*, void*,
                const void*, void*, size_t, hipStream_t);
template <class T>
int relist(const TsphParams&, TsphObject*, TsphObject*, TsphControl*, void*, size_t, hipStream_t);
template <class T>
int map_velocity(const TsphParams&, const TsphObject*, const int*, double, void*, hipStream_t);
template <class T>
int wcsph(const TsphParams&, TsphObject*, TsphObject*, TsphControl*, hipStream_t);
template <class T>
int batch_sort(const TsphBatchInfo&, char*, TsphObject*, int, int, hipStream_t);
template <class T>
int batch_density_alpha(const TsphBatchInfo&, char*, int, hipStream_t);
template <class T>
int batch_solver_begin(const TsphBatchInfo&, char*, int, uint64_t*, int, hipStream_t);
template <class T>
int batch_solver_half(const TsphBatchInfo&, char*, int, int, int, hipStream_t);
template <class T>
int batch_advect(const TsphBatchInfo&, char*, int, hipStream_t);
template <class T>
int batch_tail(const TsphBatchInfo&, char*, int, hipStream_t);
template <class T>
int batch_stop_state(const TsphBatchInfo&, char*, int*, hipStream_t);

template <class T>
int batch_prepare(const TsphParams* p, const TsphObject* f, const TsphObject* b,
                  TsphControl* const* ctrl, void* const* ws, const size_t* wsb, int ns,
                  void* host, void* table, size_t bytes, TsphBatchInfo* info, hipStream_t st) {
  TSPH_REQUIRE(bytes >= batch_table_bytes<T>(ns), TSPH_ECAPACITY,
               "tsph_batch_prepare: table needs %zu bytes", batch_table_bytes<T>(ns));
  std::memset(host, 0, batch_table_bytes<T>(ns));
  SceneEntry<T>* E = (SceneEntry<T>*)host;
  int* first = (int*)((char*)host + batch_entries_bytes<T>(ns));
  int tot[kTileKinds] = {0, 0, 0, 0};
  for (int s = 0; s < ns; ++s) {
    TSPH_REQUIRE(p[s].precision == p[0].precision, TSPH_EINVAL,
                 "tsph_batch_prepare: scene %d has another precision", s);
    TSPH_REQUIRE(f[s].n > 0, TSPH_EINVAL, "tsph_batch_prepare: scene %d has no fluid", s);
    TSPH_REQUIRE(p[s].list_cap_f >= 1 && p[s].list_cap_b >= 1, TSPH_EINVAL,
                 "tsph_batch_prepare: list capacities must be positive");
    E[s].k = make_kp<T>(p[s]);
    E[s].f = f[s];
    E[s].b = b[s];
    E[s].ctrl = ctrl[s];
    TSPH_REQUIRE(carve_ws(ws[s], wsb[s], f[s].capacity, p[s].node_num, &E[s].L), TSPH_ECAPACITY,
                 "tsph_batch_prepare: workspace of scene %d too small", s);
    const int t[kTileKinds] = {blocks_for(f[s].n), blocks_for(b[s].n),
                               (int)cell_tiles(p[s].node_num), 1};
    for (int q = 0; q < kTileKinds; ++q) {
      first[q * (ns + 1) + s] = tot[q];
      tot[q] += t[q];
    }
  }
  for (int q = 0; q < kTileKinds; ++q) {
    first[q * (ns + 1) + ns] = tot[q];
    info->tiles[q] = tot[q];
  }
  info->nscenes = ns;
  info->precision = p[0].precision;
  hipError_t e = hipMemcpyAsync(table, host, batch_table_bytes<T>(ns), hipMemcpyHostToDevice, st);
  TSPH_REQUIRE(e == hipSuccess, TSPH_EHIP, "tsph_batch_prepare: %s", hipGetErrorString(e));
  return 0;
}
template <class T>
int adj_begin(const TsphParams&, TsphObject*, TsphControl*, TsphAdjoint*, const int*, const void*,
              const void*, hipStream_t);
template <class T>
int adj_end(const TsphParams&, TsphObject*, TsphAdjoint*, const int*, void*, void*, hipStream_t);
template <class T>
int adj_solver_iteration(const TsphParams&, TsphObject*, TsphObject*, TsphControl*, TsphAdjoint*,
                         void*, size_t, hipStream_t);
template <class T>
int adj_advect(const TsphParams&, TsphObject*, TsphControl*, TsphAdjoint*, void*, size_t,
               hipStream_t);
template <class T>
int adj_density_alpha(const TsphParams&, TsphObject*, TsphObject*, TsphAdjoint*, hipStream_t);
template <class T>
int adj_bound_sums(const TsphParams&, TsphObject*, TsphObject*, TsphAdjoint*, hipStream_t);
template <class T>
int advect(const TsphParams&, TsphObject*, TsphObject*, TsphControl*, void*, size_t, hipStream_t);
template <class T>
int fbm_round(const TsphParams&, TsphObject*, TsphControl*, void*, size_t, hipStream_t);
template <class T>
int tail(const TsphParams&, TsphObject*, TsphControl*, int, void*, size_t, hipStream_t);
template <class T>
int ref_kernel(const TsphParams&, int, TsphObject*, const TsphObject*, const int*, const int*,
               const int*, const int*, TsphControl*, void*, size_t, hipStream_t);

// NeighborSearch.fill_neighbs, NeighborSearch.py:84-101, over the shared grid.
template <class T>
__global__ void __launch_bounds__(kBlock)
    k_neighbor_lists(KP<T> k, int n, const int* __restrict__ cs, int n_uids,
                     const int* __restrict__ npc, const int* __restrict__ nps,
                     const int* __restrict__ pin, const int* __restrict__ uin,
                     const int* __restrict__ row_start, int* __restrict__ neighbs,
                     int* __restrict__ shift, int* __restrict__ count) {
  int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const int side = 2 * k.R + 1;
  const int* c = cs + 4 * i;
  int cnt = 0;
  int base = neighbs ? row_start[i] : 0;
  for (int u = 0; u < n_uids; ++u) {
    shift[(long)i * n_uids + u] = cnt;
    for (int t = 0; t < k.ntemp; ++t) {
      int ox = t % side - k.R, oy = (t / side) % side - k.R;
      int oz = k.dim == 3 ? t / (side * side) - k.R : 0;
      int code = dim_encode(k, wadd(c[0], ox), wadd(c[1], oy), wadd(c[2], oz));
      if (0 < code && code < k.node_num) {
        int m = npc[code];
        for (int j = 0; j < m; ++j) {
          int s = nps[code] + j;
          if (uin[s] == u) {
            if (neighbs) neighbs[base + cnt] = pin[s];
            ++cnt;
          }
        }
      }
    }
    count[(long)i * n_uids + u] = cnt - shift[(long)i * n_uids + u];
  }
}

template <class T>
int neighbor_lists(const TsphParams& prm, const TsphObject* o, int n_uids, const int* npc,
                   const int* nps, const int* pin, const int* uin, const int* row_start,
                   int* neighbs, int* shift, int* count, hipStream_t st) {
  TSPH_REQUIRE(n_uids >= 1, TSPH_EINVAL, "tsph_neighbor_lists: n_uids >= 1");
  TSPH_REQUIRE(!neighbs || row_start, TSPH_EINVAL, "tsph_neighbor_lists: row_start required");
  if (o->n == 0) return 0;
  KP<T> k = make_kp<T>(prm);
  hipLaunchKernelGGL(k_neighbor_lists<T>, dim3(blocks_for(o->n)), dim3(kBlock), 0, st, k, o->n,
                     o->cell_structured, n_uids, npc, nps, pin, uin, row_start, neighbs, shift,
                     count);
  TSPH_LAUNCH_CHECK("neighbor_lists");
  return 0;
}

static int check_params(const TsphParams* p) {
  TSPH_REQUIRE(p, TSPH_EINVAL, "null params");
  TSPH_REQUIRE(p->dim == 2 || p->dim == 3, TSPH_EINVAL, "dim must be 2 or 3 (got %d)", p->dim);
  TSPH_REQUIRE(p->precision == 4 || p->precision == 8, TSPH_EINVAL,
               "precision must be 4 or 8 (got %d)", p->precision);
  TSPH_REQUIRE(p->phase_num >= 1 && p->phase_num <= TSPH_MAX_PHASE, TSPH_EINVAL,
               "phase_num must be 1..%d", TSPH_MAX_PHASE);
  TSPH_REQUIRE(p->neighb_range >= 1 && p->node_num >= 1, TSPH_EINVAL, "bad grid parameters");
  return 0;
}

// The adjoint differentiates one-phase DFSPH only.
static int check_adj(const TsphParams* p, const TsphAdjoint* a, const char* what) {
  TSPH_REQUIRE(a, TSPH_EINVAL, "%s: null adjoint", what);
  TSPH_REQUIRE(p->solver_type == TSPH_DFSPH && p->phase_num == 1, TSPH_EINVAL,
               "%s: the adjoint supports one-phase DFSPH only", what);
  return 0;
}

// The entries that read or write the boundary sums (tsph_adj_bound_sums).
static int check_bsums(const TsphAdjoint* a, const char* what) {
  TSPH_REQUIRE(a->bmom && a->bflag, TSPH_EINVAL, "%s: null boundary sums (bmom, bflag)", what);
  return 0;
}

static int check_obj(const TsphObject* o, const char* what) {
  TSPH_REQUIRE(o, TSPH_EINVAL, "%s: null object", what);
  TSPH_REQUIRE(o->n >= 0 && o->n <= o->capacity, TSPH_ECAPACITY,
               "%s: part_num %d exceeds capacity %d", what, o->n, o->capacity);
  return 0;
}

}  // namespace tsph

using namespace tsph;

#define DISPATCH(p, fn, ...) ((p)->precision == 8 ? fn<double>(__VA_ARGS__) : fn<float>(__VA_ARGS__))

extern "C" {

const char* tsph_last_error(void) { return g_err.c_str(); }
int tsph_abi_version(void) { return TSPH_ABI_VERSION; }
size_t tsph_control_bytes(void) { return sizeof(TsphControl); }

int tsph_verify_fast_quotient(const TsphParams* p, int32_t* mismatches, void* stream) {
  if (int rc = check_params(p)) return rc;
  TSPH_REQUIRE(mismatches, TSPH_EINVAL, "tsph_verify_fast_quotient: null output");
  return tsph::verify_fast_quotient(*p, (int*)mismatches, (hipStream_t)stream);
}

int tsph_signal_alloc(size_t bytes, void** host, void** device) {
  TSPH_REQUIRE(host && device && bytes > 0, TSPH_EINVAL, "tsph_signal_alloc: bad arguments");
  *host = nullptr;
  *device = nullptr;
  void* h = nullptr;
  hipError_t e = hipHostMalloc(&h, bytes, hipHostMallocMapped | hipHostMallocCoherent);
  TSPH_REQUIRE(e == hipSuccess, TSPH_EHIP, "tsph_signal_alloc: hipHostMalloc: %s",
               hipGetErrorString(e));
  memset(h, 0, bytes);
  void* d = nullptr;
  e = hipHostGetDevicePointer(&d, h, 0);
  if (e != hipSuccess) {
    (void)hipHostFree(h);
    TSPH_REQUIRE(false, TSPH_EHIP, "tsph_signal_alloc: hipHostGetDevicePointer: %s",
                 hipGetErrorString(e));
  }
  *host = h;
  *device = d;
  return 0;
}

int tsph_signal_free(void* host) {
  if (!host) return 0;
  hipError_t e = hipHostFree(host);
  TSPH_REQUIRE(e == hipSuccess, TSPH_EHIP, "tsph_signal_free: %s", hipGetErrorString(e));
  return 0;
}

size_t tsph_workspace_bytes(const TsphParams* p, int32_t capacity) {
  if (!p) return 0;
  return ws_bytes_needed(capacity, p->node_num);
}

int tsph_sort_object(const TsphParams* p, TsphObject* obj, TsphControl* ctrl, int32_t cfl,
                     void* ws, size_t wsb, void* stream) {
  if (int rc = check_params(p)) return rc;
  if (int rc = check_obj(obj, "tsph_sort_object")) return rc;
  return DISPATCH(p, sort_object, *p, obj, ctrl, cfl, ws, wsb, (hipStream_t)stream);
}

int tsph_sort_object_static(const TsphParams* p, TsphObject* obj, TsphControl* ctrl, int32_t cfl,
                            void* stream) {
  if (int rc = check_params(p)) return rc;
  if (int rc = check_obj(obj, "tsph_sort_object_static")) return rc;
  return DISPATCH(p, sort_object_static, *p, obj, ctrl, cfl, (hipStream_t)stream);
}

size_t tsph_sort_static_scratch_bytes(void) { return tsph::static_sort_scratch_bytes(); }
int32_t tsph_sort_static_max_particles(void) { return (int32_t)tsph::static_sort_max_n(); }
int32_t tsph_sort_static_max_cells(void) { return (int32_t)tsph::static_sort_max_cells(); }

int tsph_ngrid_build(const TsphParams* p, TsphObject* const* objs, int32_t nobj, int32_t* npc,
                     int32_t* nps, int32_t* npsc, int32_t* pin, int32_t* uin, void* ws, size_t wsb,
                     void* stream) {
  if (int rc = check_params(p)) return rc;
  for (int u = 0; u < nobj; ++u)
    if (int rc = check_obj(objs[u], "tsph_ngrid_build")) return rc;
  return DISPATCH(p, ngrid_build, *p, objs, nobj, npc, nps, npsc, pin, uin, ws, wsb,
                  (hipStream_t)stream);
}

int tsph_neighbor_lists(const TsphParams* p, const TsphObject* obj, int32_t n_uids,
                        const int32_t* npc, const int32_t* nps, const int32_t* pin,
                        const int32_t* uin, const int32_t* row_start, int32_t* neighbs,
                        int32_t* shift, int32_t* count, void* stream) {
  if (int rc = check_params(p)) return rc;
  if (int rc = check_obj(obj, "tsph_neighbor_lists")) return rc;
  return DISPATCH(p, neighbor_lists, *p, obj, n_uids, npc, nps, pin, uin, row_start, neighbs,
                  shift, count, (hipStream_t)stream);
}

int tsph_density_alpha(const TsphParams* p, TsphObject* f, TsphObject* b, TsphControl* ctrl,
                       void* ws, size_t wsb, void* stream) {
  if (int rc = check_params(p)) return rc;
  if (int rc = check_obj(f, "tsph_density_alpha")) return rc;
  if (int rc = check_obj(b, "tsph_density_alpha")) return rc;
  TSPH_REQUIRE(p->list_cap_f >= 1 && p->list_cap_b >= 1, TSPH_EINVAL,
               "tsph_density_alpha: list capacities must be positive");
  return DISPATCH(p, density_alpha, *p, f, b, ctrl, ws, wsb, (hipStream_t)stream);
}

int tsph_solver_begin(const TsphParams* p, TsphControl* ctrl, int32_t solve, void* stream) {
  if (int rc = check_params(p)) return rc;
  return DISPATCH(p, solver_begin, *p, ctrl, solve, (int*)nullptr, 0, (hipStream_t)stream);
}

int tsph_solver_begin_signal(const TsphParams* p, TsphControl* ctrl, int32_t solve,
                             uint64_t* signal, int32_t token, void* stream) {
  if (int rc = check_params(p)) return rc;
  return DISPATCH(p, solver_begin, *p, ctrl, solve, (int*)signal, (int)token,
                  (hipStream_t)stream);
}

static int solver_entry(const TsphParams* p, TsphObject* f, TsphObject* b, TsphControl* ctrl,
                        int32_t k, int half, void* ws, size_t wsb, void* stream,
                        void* tape_v = nullptr, void* tape_k = nullptr,
                        const void* clamp_ref = nullptr) {
  if (int rc = check_params(p)) return rc;
  if (int rc = check_obj(f, "tsph_solver_iteration")) return rc;
  if (int rc = check_obj(b, "tsph_solver_iteration")) return rc;
  TSPH_REQUIRE(k >= 0 && k < TSPH_MAX_ITER_LOG, TSPH_EINVAL,
               "tsph_solver_iteration: iteration %d outside [0, %d)", k, TSPH_MAX_ITER_LOG);
  return DISPATCH(p, solver_half, *p, f, b, ctrl, k, half, tape_v, tape_k, clamp_ref, ws, wsb,
                  (hipStream_t)stream);
}

int tsph_solver_iteration(const TsphParams* p, TsphObject* f, TsphObject* b, TsphControl* ctrl,
                          int32_t k, void* ws, size_t wsb, void* stream) {
  return solver_entry(p, f, b, ctrl, k, 3, ws, wsb, stream);
}
int tsph_solver_psi(const TsphParams* p, TsphObject* f, TsphObject* b, TsphControl* ctrl,
                    int32_t k, void* ws, size_t wsb, void* stream) {
  return solver_entry(p, f, b, ctrl, k, 1, ws, wsb, stream);
}
int tsph_solver_update(const TsphParams* p, TsphObject* f, TsphObject* b, TsphControl* ctrl,
                       int32_t k, void* ws, size_t wsb, void* stream) {
  return solver_entry(p, f, b, ctrl, k, 2, ws, wsb, stream);
}
int tsph_solver_psi_taped(const TsphParams* p, TsphObject* f, TsphObject* b, TsphControl* ctrl,
                          int32_t k, void* tape_vel_adv, void* tape_kappa, void* ws, size_t wsb,
                          void* stream) {
  TSPH_REQUIRE(tape_vel_adv && tape_kappa, TSPH_EINVAL, "tsph_solver_psi_taped: null tape");
  return solver_entry(p, f, b, ctrl, k, 1, ws, wsb, stream, tape_vel_adv, tape_kappa);
}

int tsph_solver_psi_frozen(const TsphParams* p, TsphObject* f, TsphObject* b, TsphControl* ctrl,
                           int32_t k, const void* kappa_ref, void* ws, size_t wsb, void* stream) {
  TSPH_REQUIRE(kappa_ref, TSPH_EINVAL, "tsph_solver_psi_frozen: null kappa_ref");
  return solver_entry(p, f, b, ctrl, k, 1, ws, wsb, stream, nullptr, nullptr, kappa_ref);
}

int tsph_relist(const TsphParams* p, TsphObject* f, TsphObject* b, TsphControl* ctrl, void* ws,
                size_t wsb, void* stream) {
  if (int rc = check_params(p)) return rc;
  if (int rc = check_obj(f, "tsph_relist")) return rc;
  if (int rc = check_obj(b, "tsph_relist")) return rc;
  TSPH_REQUIRE(p->list_cap_f >= 1 && p->list_cap_b >= 1, TSPH_EINVAL,
               "tsph_relist: list capacities must be positive");
  return DISPATCH(p, relist, *p, f, b, ctrl, ws, wsb, (hipStream_t)stream);
}

int tsph_advect(const TsphParams* p, TsphObject* f, TsphObject* b, TsphControl* ctrl, void* ws,
                size_t wsb, void* stream) {
  if (int rc = check_params(p)) return rc;
  if (int rc = check_obj(f, "tsph_advect")) return rc;
  return DISPATCH(p, advect, *p, f, b, ctrl, ws, wsb, (hipStream_t)stream);
}

int tsph_fbm_round(const TsphParams* p, TsphObject* f, TsphControl* ctrl, void* ws, size_t wsb,
                   void* stream) {
  if (int rc = check_params(p)) return rc;
  if (int rc = check_obj(f, "tsph_fbm_round")) return rc;
  return DISPATCH(p, fbm_round, *p, f, ctrl, ws, wsb, (hipStream_t)stream);
}

int tsph_tail(const TsphParams* p, TsphObject* f, TsphControl* ctrl, int32_t fbm, void* ws,
              size_t wsb, void* stream) {
  if (int rc = check_params(p)) return rc;
  if (int rc = check_obj(f, "tsph_tail")) return rc;
  return DISPATCH(p, tail, *p, f, ctrl, fbm, ws, wsb, (hipStream_t)stream);
}

int tsph_adj_begin(const TsphParams* p, TsphObject* f, TsphControl* ctrl, TsphAdjoint* adj,
                   const int32_t* pid, const void* xbar_ref, const void* vbar_ref, void* stream) {
  if (int rc = check_params(p)) return rc;
  if (int rc = check_obj(f, "tsph_adj_begin")) return rc;
  if (int rc = check_adj(p, adj, "tsph_adj_begin")) return rc;
  TSPH_REQUIRE(pid && xbar_ref && vbar_ref, TSPH_EINVAL, "tsph_adj_begin: null array");
  return DISPATCH(p, adj_begin, *p, f, ctrl, adj, pid, xbar_ref, vbar_ref, (hipStream_t)stream);
}

int tsph_adj_end(const TsphParams* p, TsphObject* f, TsphAdjoint* adj, const int32_t* pid,
                 void* xbar_ref, void* vbar_ref, void* stream) {
  if (int rc = check_params(p)) return rc;
  if (int rc = check_obj(f, "tsph_adj_end")) return rc;
  if (int rc = check_adj(p, adj, "tsph_adj_end")) return rc;
  TSPH_REQUIRE(pid && xbar_ref && vbar_ref, TSPH_EINVAL, "tsph_adj_end: null array");
  return DISPATCH(p, adj_end, *p, f, adj, pid, xbar_ref, vbar_ref, (hipStream_t)stream);
}

int tsph_adj_solver_iteration(const TsphParams* p, TsphObject* f, TsphObject* b,
                              TsphControl* ctrl, TsphAdjoint* adj, void* ws, size_t wsb,
                              void* stream) {
  if (int rc = check_params(p)) return rc;
  if (int rc = check_obj(f, "tsph_adj_solver_iteration")) return rc;
  if (int rc = check_obj(b, "tsph_adj_solver_iteration")) return rc;
  if (int rc = check_adj(p, adj, "tsph_adj_solver_iteration")) return rc;
  if (int rc = check_bsums(adj, "tsph_adj_solver_iteration")) return rc;
  TSPH_REQUIRE(adj->tape_vel_adv, TSPH_EINVAL, "tsph_adj_solver_iteration: null tape");
  return DISPATCH(p, adj_solver_iteration, *p, f, b, ctrl, adj, ws, wsb, (hipStream_t)stream);
}

int tsph_adj_advect(const TsphParams* p, TsphObject* f, TsphControl* ctrl, TsphAdjoint* adj,
                    void* ws, size_t wsb, void* stream) {
  if (int rc = check_params(p)) return rc;
  if (int rc = check_obj(f, "tsph_adj_advect")) return rc;
  if (int rc = check_adj(p, adj, "tsph_adj_advect")) return rc;
  TSPH_REQUIRE(adj->tape_vel && adj->grad && (adj->tape_normal || !adj->nbar), TSPH_EINVAL,
               "tsph_adj_advect: null tape or grad");
  TSPH_REQUIRE(adj->nbar || p->surface_tension_gamma == 0.0, TSPH_EINVAL,
               "tsph_adj_advect: nbar may be NULL only when surface_tension_gamma == 0");
  return DISPATCH(p, adj_advect, *p, f, ctrl, adj, ws, wsb, (hipStream_t)stream);
}

int tsph_adj_bound_sums(const TsphParams* p, TsphObject* f, TsphObject* b, TsphAdjoint* adj,
                        void* stream) {
  if (int rc = check_params(p)) return rc;
  if (int rc = check_obj(f, "tsph_adj_bound_sums")) return rc;
  if (int rc = check_obj(b, "tsph_adj_bound_sums")) return rc;
  if (int rc = check_adj(p, adj, "tsph_adj_bound_sums")) return rc;
  if (int rc = check_bsums(adj, "tsph_adj_bound_sums")) return rc;
  return DISPATCH(p, adj_bound_sums, *p, f, b, adj, (hipStream_t)stream);
}

int tsph_adj_density_alpha(const TsphParams* p, TsphObject* f, TsphObject* b, TsphAdjoint* adj,
                           void* stream) {
  if (int rc = check_params(p)) return rc;
  if (int rc = check_obj(f, "tsph_adj_density_alpha")) return rc;
  if (int rc = check_obj(b, "tsph_adj_density_alpha")) return rc;
  if (int rc = check_adj(p, adj, "tsph_adj_density_alpha")) return rc;
  if (int rc = check_bsums(adj, "tsph_adj_density_alpha")) return rc;
  return DISPATCH(p, adj_density_alpha, *p, f, b, adj, (hipStream_t)stream);
}

int tsph_wcsph(const TsphParams* p, TsphObject* f, TsphObject* b, TsphControl* ctrl,
               void* stream) {
  if (int rc = check_params(p)) return rc;
  if (int rc = check_obj(f, "tsph_wcsph")) return rc;
  if (int rc = check_obj(b, "tsph_wcsph")) return rc;
  TSPH_REQUIRE(p->wc_gamma != 0, TSPH_EINVAL, "tsph_wcsph: solver_wc_gamma must be non-zero");
  return DISPATCH(p, wcsph, *p, f, b, ctrl, (hipStream_t)stream);
}

int tsph_map_velocity(const TsphParams* p, const TsphObject* obj, const int32_t* grid_shape,
                      double grid_dist, void* grid_vel, void* stream) {
  if (int rc = check_params(p)) return rc;
  if (int rc = check_obj(obj, "tsph_map_velocity")) return rc;
  TSPH_REQUIRE(grid_shape && grid_vel, TSPH_EINVAL, "tsph_map_velocity: null grid");
  for (int d = 0; d < p->dim; ++d)
    TSPH_REQUIRE(grid_shape[d] >= 0, TSPH_EINVAL, "tsph_map_velocity: negative grid shape");
  return DISPATCH(p, map_velocity, *p, obj, grid_shape, grid_dist, grid_vel, (hipStream_t)stream);
}

size_t tsph_batch_table_bytes(int32_t precision, int32_t nscenes) {
  if (nscenes < 1) return 0;
  return precision == 8 ? batch_table_bytes<double>(nscenes) : batch_table_bytes<float>(nscenes);
}

int tsph_batch_prepare(const TsphParams* params, const TsphObject* fluid, const TsphObject* bound,
                       TsphControl* const* ctrl, void* const* workspace, const size_t* ws_bytes,
                       int32_t nscenes, void* host_staging, void* table, size_t table_bytes,
                       TsphBatchInfo* info, void* stream) {
  TSPH_REQUIRE(nscenes >= 1 && params && fluid && bound && ctrl && workspace && ws_bytes &&
                   host_staging && table && info,
               TSPH_EINVAL, "tsph_batch_prepare: null argument or no scenes");
  for (int s = 0; s < nscenes; ++s) {
    if (int rc = check_params(&params[s])) return rc;
    if (int rc = check_obj(&fluid[s], "tsph_batch_prepare")) return rc;
    if (int rc = check_obj(&bound[s], "tsph_batch_prepare")) return rc;
  }
  return DISPATCH(&params[0], batch_prepare, params, fluid, bound, ctrl, workspace, ws_bytes,
                  nscenes, host_staging, table, table_bytes, info, (hipStream_t)stream);
}

static int check_batch(const TsphBatchInfo* bi, void* table, const char* what) {
  TSPH_REQUIRE(bi && table && bi->nscenes >= 1, TSPH_EINVAL, "%s: bad batch", what);
  TSPH_REQUIRE(bi->precision == 4 || bi->precision == 8, TSPH_EINVAL, "%s: bad precision", what);
  return 0;
}

#define BDISPATCH(bi, fn, ...) \
  ((bi)->precision == 8 ? fn<double>(__VA_ARGS__) : fn<float>(__VA_ARGS__))

int tsph_batch_sort(const TsphBatchInfo* bi, void* table, TsphObject* fluid, int32_t cfl,
                    int32_t grid, void* stream) {
  if (int rc = check_batch(bi, table, "tsph_batch_sort")) return rc;
  return BDISPATCH(bi, batch_sort, *bi, (char*)table, fluid, cfl, grid, (hipStream_t)stream);
}
int tsph_batch_density_alpha(const TsphBatchInfo* bi, void* table, int32_t grid, void* stream) {
  if (int rc = check_batch(bi, table, "tsph_batch_density_alpha")) return rc;
  return BDISPATCH(bi, batch_density_alpha, *bi, (char*)table, grid, (hipStream_t)stream);
}
int tsph_batch_solver_begin(const TsphBatchInfo* bi, void* table, int32_t solve, void* stream) {
  if (int rc = check_batch(bi, table, "tsph_batch_solver_begin")) return rc;
  return BDISPATCH(bi, batch_solver_begin, *bi, (char*)table, solve, (uint64_t*)nullptr, 0,
                   (hipStream_t)stream);
}
int tsph_batch_solver_begin_signal(const TsphBatchInfo* bi, void* table, int32_t solve,
                                   uint64_t* signal, int32_t token, void* stream) {
  if (int rc = check_batch(bi, table, "tsph_batch_solver_begin_signal")) return rc;
  return BDISPATCH(bi, batch_solver_begin, *bi, (char*)table, solve, signal, (int)token,
                   (hipStream_t)stream);
}
int tsph_batch_solver_psi(const TsphBatchInfo* bi, void* table, int32_t k, int32_t grid,
                          void* stream) {
  if (int rc = check_batch(bi, table, "tsph_batch_solver_psi")) return rc;
  TSPH_REQUIRE(k >= 0 && k < TSPH_MAX_ITER_LOG, TSPH_EINVAL, "tsph_batch_solver_psi: bad k");
  return BDISPATCH(bi, batch_solver_half, *bi, (char*)table, k, 1, grid, (hipStream_t)stream);
}
int tsph_batch_solver_update(const TsphBatchInfo* bi, void* table, int32_t k, int32_t grid,
                             void* stream) {
  if (int rc = check_batch(bi, table, "tsph_batch_solver_update")) return rc;
  TSPH_REQUIRE(k >= 0 && k < TSPH_MAX_ITER_LOG, TSPH_EINVAL, "tsph_batch_solver_update: bad k");
  return BDISPATCH(bi, batch_solver_half, *bi, (char*)table, k, 2, grid, (hipStream_t)stream);
}
int tsph_batch_advect(const TsphBatchInfo* bi, void* table, int32_t grid, void* stream) {
  if (int rc = check_batch(bi, table, "tsph_batch_advect")) return rc;
  return BDISPATCH(bi, batch_advect, *bi, (char*)table, grid, (hipStream_t)stream);
}
int tsph_batch_tail(const TsphBatchInfo* bi, void* table, int32_t grid, void* stream) {
  if (int rc = check_batch(bi, table, "tsph_batch_tail")) return rc;
  return BDISPATCH(bi, batch_tail, *bi, (char*)table, grid, (hipStream_t)stream);
}
int tsph_batch_stop_state(const TsphBatchInfo* bi, void* table, int32_t* out, void* stream) {
  if (int rc = check_batch(bi, table, "tsph_batch_stop_state")) return rc;
  TSPH_REQUIRE(out, TSPH_EINVAL, "tsph_batch_stop_state: null out");
  return BDISPATCH(bi, batch_stop_state, *bi, (char*)table, out, (hipStream_t)stream);
}

int tsph_ref_kernel(const TsphParams* p, int32_t kernel, TsphObject* obj, const TsphObject* nobj,
                    const int32_t* npc, const int32_t* nps, const int32_t* pin, const int32_t* uin,
                    TsphControl* ctrl, void* ws, size_t wsb, void* stream) {
  if (int rc = check_params(p)) return rc;
  if (int rc = check_obj(obj, "tsph_ref_kernel")) return rc;
  return DISPATCH(p, ref_kernel, *p, kernel, obj, nobj, npc, nps, pin, uin, ctrl, ws, wsb,
                  (hipStream_t)stream);
}

}  // extern "C"
