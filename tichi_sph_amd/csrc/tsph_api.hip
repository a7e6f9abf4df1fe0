// extern "C" entry points declared in include/tichi_sph.h.
#include <cstdarg>

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
int ngrid_build(const TsphParams&, TsphObject* const*, int, int*, int*, int*, int*, int*, void*,
                size_t, hipStream_t);
template <class T>
int density_alpha(const TsphParams&, TsphObject*, TsphObject*, TsphControl*, void*, size_t,
                  hipStream_t);
template <class T>
int solver_begin(const TsphParams&, TsphControl*, int, hipStream_t);
template <class T>
int solver_half(const TsphParams&, TsphObject*, TsphObject*, TsphControl*, int, int, void*, void*,
                void*, size_t, hipStream_t);
template <class T>
int relist(const TsphParams&, TsphObject*, TsphObject*, TsphControl*, void*, size_t, hipStream_t);
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
  return DISPATCH(p, solver_begin, *p, ctrl, solve, (hipStream_t)stream);
}

static int solver_entry(const TsphParams* p, TsphObject* f, TsphObject* b, TsphControl* ctrl,
                        int32_t k, int half, void* ws, size_t wsb, void* stream,
                        void* tape_v = nullptr, void* tape_k = nullptr) {
  if (int rc = check_params(p)) return rc;
  if (int rc = check_obj(f, "tsph_solver_iteration")) return rc;
  if (int rc = check_obj(b, "tsph_solver_iteration")) return rc;
  TSPH_REQUIRE(k >= 0 && k < TSPH_MAX_ITER_LOG, TSPH_EINVAL,
               "tsph_solver_iteration: iteration %d outside [0, %d)", k, TSPH_MAX_ITER_LOG);
  return DISPATCH(p, solver_half, *p, f, b, ctrl, k, half, tape_v, tape_k, ws, wsb,
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

int tsph_adj_density_alpha(const TsphParams* p, TsphObject* f, TsphObject* b, TsphAdjoint* adj,
                           void* stream) {
  if (int rc = check_params(p)) return rc;
  if (int rc = check_obj(f, "tsph_adj_density_alpha")) return rc;
  if (int rc = check_obj(b, "tsph_adj_density_alpha")) return rc;
  if (int rc = check_adj(p, adj, "tsph_adj_density_alpha")) return rc;
  return DISPATCH(p, adj_density_alpha, *p, f, b, adj, (hipStream_t)stream);
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
