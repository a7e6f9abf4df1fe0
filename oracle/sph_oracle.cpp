// CPU oracle for the Tichi_SPH per-step particle hot path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP product is
// compared against; only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it.  Nothing in tichi_sph_amd/ links or calls it.
//
// It restates, line by line and in the reference's own schedule, the Taichi
// kernels of the reference (sph_config.py, sph_util.py, sph_obj.py, sph.py,
// NeighborSearch.py, main.py).  Every function cites the file:line it follows.
// It is templated on the arithmetic type: T=float is the FP32 oracle,
// T=double the FP64 oracle of the same step (same inputs, same f32-valued
// configuration constants, f64 arithmetic).
//
// Evaluation order.  Taichi evaluates Python expressions left to right with
// the usual precedence, vector ops component-wise; integer powers are
// demoted to multiplications by squaring, so q**2 = q*q and q**3 = q*(q*q).
// `norm()` is sqrt(((x0*x0)+(x1*x1))+(x2*x2)), `dot` is a left-to-right sum
// of products, `normalized()` is (1/norm)*x.  This file is compiled with
// -ffp-contract=off and without fast-math, so each operation rounds once
// (IEEE round-to-nearest), in that order.  The last-ulp behaviour of the
// reference's own Taichi build (fast-math codegen) is not knowable without
// Taichi and is documented as unpinned.
//
// Deliberate, documented deviations (all order-of-reduction only):
//  * The 0-D reductions that Taichi performs with float atomics in a parallel
//    loop (compression sph.py:221-226, energies sph.py:274-279) have no fixed
//    order in the reference.  Here the sum is accumulated in double and then
//    rounded to T, i.e. the order-independent value every f32 order
//    approximates.  compression = T(sum) / T(part_num) as in sph.py:226.
//  * The FBM round loop (main.py:244-248) has no bound in the reference; here
//    it is capped at kFbmMaxRounds rounds.
//  * Out-of-range float->int cell conversions (non-finite or |v| >= 2^31) give
//    INT_MIN, the x86 "integer indefinite" value; integer cell-code arithmetic
//    wraps modulo 2^32.
//
// Threading.  With one thread (oracle_set_threads(1)) every parallel loop
// runs in index order, which is the reference's single-thread semantics and
// defines the canonical bucket order (sph_obj.py:308-323).  With more threads
// the grid fill uses atomics exactly like the reference (non-deterministic
// in-cell order); that mode is only used to time the CPU baseline.

#include <algorithm>
#include <type_traits>
#include <atomic>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr int kMaxPhase = 8;
constexpr int kFbmMaxRounds = 1000;

// Raw configuration as read from config.json / scenario.json
// (sph_config.py:145-185).  Mirrored by a ctypes.Structure in oracle/oracle.py.
struct RawInit {
  int dim, phase_num, solver_type;  // solver_type: 0 = DFSPH, 1 = VFSPH
  int neighb_range;
  double global_part_size, dynamic_viscosity;
  double lb[3], rt[3], gravity[3];
  double phase_rest_density[kMaxPhase];
  int fluid_max, bound_max;
  int wc_gamma;
  double divergence_threshold, compression_threshold;
  int iter_min, iter_max;
  int cs;
  double cfl_factor;
  double fbm_diffusion, fbm_convection, surface_tension_gamma;
  // extension: per-phase viscosities (mu_per_phase != 0), see Consts
  int mu_per_phase;
  double phase_viscosity[kMaxPhase];
};

inline double f32(double x) { return (double)(float)x; }

// Derived constants, sph_config.py:145-232.  Values that the reference keeps
// in f32 fields are stored here already rounded to f32 (held in double).
struct Consts {
  int dim = 3, P = 1, solver = 0, R = 1;
  double part_size[5] = {0}, kernel_h[5] = {0}, kernel_sig[4] = {0};
  double lb[3] = {0}, rt[3] = {0}, g[3] = {0}, rho[kMaxPhase] = {0};
  double mu = 0;
  // Per-phase viscosity, an extension the reference lacks (one global mu,
  // sph_config.py:148): mu_i = sum_k phi_ik mu_k replaces mu in
  // SPH_advection_viscosity_acc when mu_pp is set.
  int mu_pp = 0;
  double muk[kMaxPhase] = {0};
  double grid_size[3] = {0};
  int coder[3] = {0, 0, 0};
  int node_num = 0;
  int ntemp = 0;
  int tmpl[125][3];
  double div_thr = 0, comp_thr = 0;
  int iter_min = 0, iter_max = 0, cs = 1, wc_gamma = 7;
  double cfl_factor = 0, fbm_D = 0, fbm_C = 0, gamma = 0;
};

Consts make_consts(const RawInit& in) {
  Consts c;
  c.dim = in.dim;
  c.P = in.phase_num;
  c.solver = in.solver_type;
  c.R = in.neighb_range;
  // init_sim_env, sph_config.py:145-163
  c.part_size[1] = f32(in.global_part_size);
  c.mu = f32(in.dynamic_viscosity);
  c.mu_pp = in.mu_per_phase;
  for (int k = 0; k < c.P; ++k) c.muk[k] = f32(in.phase_viscosity[k]);
  for (int d = 0; d < 3; ++d) {
    c.lb[d] = d < c.dim ? f32(in.lb[d]) : 0.0;
    c.rt[d] = d < c.dim ? f32(in.rt[d]) : 0.0;
    c.g[d] = d < c.dim ? f32(in.gravity[d]) : 0.0;
  }
  for (int k = 0; k < c.P; ++k) c.rho[k] = f32(in.phase_rest_density[k]);
  // math.pow on the f32 value read back from the field (sph_config.py:161-163)
  c.part_size[2] = f32(std::pow(c.part_size[1], 2));
  c.part_size[3] = f32(std::pow(c.part_size[1], 3));
  c.part_size[4] = f32(std::pow(c.part_size[1], 4));
  // init_solver, sph_config.py:165-185
  c.wc_gamma = in.wc_gamma;
  c.div_thr = f32(in.divergence_threshold);
  c.comp_thr = f32(in.compression_threshold);
  c.iter_min = in.iter_min;
  c.iter_max = in.iter_max;
  c.cs = in.cs;
  c.cfl_factor = f32(in.cfl_factor);
  c.fbm_D = f32(in.fbm_diffusion);
  c.fbm_C = f32(in.fbm_convection);
  c.gamma = f32(in.surface_tension_gamma);
  // calculate_kernel_param, sph_config.py:196-202 (f64 on the host, f32 store)
  c.kernel_h[1] = f32(c.part_size[1] * 2);
  c.kernel_h[2] = f32(std::pow(c.kernel_h[1], 2));
  c.kernel_h[3] = f32(std::pow(c.kernel_h[1], 3));
  c.kernel_h[4] = f32(std::pow(c.kernel_h[1], 4));
  c.kernel_sig[3] = f32(8 / M_PI / std::pow(c.kernel_h[1], 3));
  c.kernel_sig[2] = f32(40.0 / 7 / M_PI / std::pow(c.kernel_h[1], 2));
  // init_neighb_param, sph_config.py:204-208: ceil in f32 inside a kernel
  for (int d = 0; d < c.dim; ++d) {
    float ext = (float)c.rt[d] - (float)c.lb[d];
    c.grid_size[d] = (double)std::ceil(ext / (float)c.kernel_h[1]);
  }
  // calculate_neighb_param, sph_config.py:210-224
  c.node_num = 1;
  for (int d = 0; d < c.dim; ++d) {
    c.coder[d] = c.node_num;
    c.node_num *= (int)c.grid_size[d];
  }
  int side = 2 * c.R + 1;
  c.ntemp = 1;
  for (int d = 0; d < c.dim; ++d) c.ntemp *= side;
  for (int j = 0; j < c.ntemp; ++j) {
    int tmp = j;
    for (int d = 0; d < c.dim; ++d) {
      int p = 1;
      for (int e = 0; e < c.dim - d - 1; ++e) p *= side;
      int digit = tmp / p;
      tmp = tmp % p;
      c.tmpl[j][c.dim - d - 1] = -c.R + digit;
    }
    for (int d = c.dim; d < 3; ++d) c.tmpl[j][d] = 0;
  }
  return c;
}

// Taichi float->int cast of an integral-valued float, with the documented
// out-of-range convention.
template <class T>
inline int to_i32(T v) {
  if (!(v >= T(-2147483648.0) && v < T(2147483648.0))) return INT_MIN;
  return (int)v;
}
inline int wrap_add(int a, int b) { return (int)((uint32_t)a + (uint32_t)b); }
inline int wrap_mul(int a, int b) { return (int)((uint32_t)a * (uint32_t)b); }

template <class T>
inline T pow2(T x) { return x * x; }
template <class T>
inline T pow3(T x) { return x * (x * x); }

// Kernel functions, sph_util.py:4-43.
template <class T>
struct KernelMath {
  T h, h2, sig;
  int dim;
  // C, sph_util.py:4-13 (Akinci et al. 2013 cohesion spline)
  T C(T r) const {
    T q = r / h;
    T tmp = T(0);
    if (q <= T(0.5))
      tmp = (T(2) * pow3(T(1) - q)) * pow3(q) - T(1.0 / 64);
    else if (q > T(0.5) && q < T(1))
      tmp = pow3(T(1) - q) * pow3(q);
    tmp = tmp * (T(32 / M_PI) / pow3(h));
    return tmp;
  }
  // W, sph_util.py:16-25 (cubic spline)
  T W(T r) const {
    T q = r / h;
    T tmp = T(0);
    if (q <= T(0.5))
      tmp = T(6) * (pow3(q) - pow2(q)) + T(1);
    else if (q > T(0.5) && q < T(1))
      tmp = T(2) * pow3(T(1) - q);
    tmp = tmp * sig;
    return tmp;
  }
  // W_grad, sph_util.py:28-37 (scalar dW/dr)
  T W_grad(T r) const {
    T q = r / h;
    T tmp = T(0);
    if (q <= T(0.5))
      tmp = T(6) * (T(3) * pow2(q) - T(2) * q);
    else if (q > T(0.5) && q < T(1))
      tmp = T(-6) * pow2(T(1) - q);
    tmp = tmp * (sig / h);
    return tmp;
  }
};

template <class T>
struct Obj {
  int max = 0, n = 0, uid = 0;
  std::vector<T> pos, vel, vel_adv, acce, acce_adv, mass, rest_density, rest_volume;
  std::vector<T> volume_frac, volume_frac_tmp, W, W_grad, sph_compression, sph_density;
  std::vector<T> psi_adv, alpha, alpha_1, alpha_2, normal, drift_vel, fbm_zeta;
  std::vector<T> pressure, pressure_force;
  std::vector<int> flag, neighb_cell_seq, neighb_in_cell_seq, neighb_cell_structured_seq;
  T compression = T(0);
  int general_flag = 0;
  double ke = 0, pe = 0;

  void alloc(int m, int dim, int P) {
    max = m;
    auto v = [&](std::vector<T>& a, int k) { a.assign((size_t)m * k, T(0)); };
    v(pos, dim); v(vel, dim); v(vel_adv, dim); v(acce, dim); v(acce_adv, dim);
    v(mass, 1); v(rest_density, 1); v(rest_volume, 1); v(volume_frac, P);
    v(volume_frac_tmp, P); v(W, 1); v(W_grad, dim); v(sph_compression, 1);
    v(sph_density, 1); v(psi_adv, 1); v(alpha, 1); v(alpha_1, dim); v(alpha_2, 1);
    v(normal, dim); v(drift_vel, (size_t)P * dim); v(fbm_zeta, 1); v(pressure, 1);
    v(pressure_force, dim);
    flag.assign(m, 0);
    neighb_cell_seq.assign(m, 0);
    neighb_in_cell_seq.assign(m, 0);
    neighb_cell_structured_seq.assign((size_t)m * dim, 0);
  }
};

struct Grid {  // Ngrid, sph_obj.py:280-293
  std::vector<int> node_part_count, node_part_shift, node_part_shift_count;
  std::vector<int> part_pid_in_node, part_uid_in_node;
};

struct NSearch {  // NeighborSearch, NeighborSearch.py:5-29 (per-uid layout)
  int uids = 0;
  std::vector<int> node_part_count, node_part_shift, node_part_shift_count, part_pid_in_node;
  // per object: candidate lists, shift/count per uid (PartNeighbs, 5-9)
  std::vector<std::vector<int>> neighbs;        // flattened per particle, variable length
  std::vector<std::vector<int>> neighbs_start;  // prefix offsets into neighbs per particle (n+1)
  std::vector<std::vector<int>> neighbs_shift;  // [n*uids]
  std::vector<std::vector<int>> neighbs_count;  // [n*uids]
};

struct StepStats {  // mirrored in oracle/oracle.py
  int div_iter, incom_iter, fbm_rounds, pad;
  double dt, ke, pe;
  double div_err[128];
  double incom_err[128];
};

enum StepFlags {
  kStepCfl = 1,      // run cfl_condition first (main.py:348)
  kStepFbm = 2,      // enable the FBM block (main.py:243-250)
  kStepWcsph = 4,    // enable the WCSPH block (main.py:238-242)
};

int g_threads = 1;

template <class T>
struct Sim {
  Consts c;
  KernelMath<T> km;
  Obj<T> obj[2];  // 0 = fluid (uid 0), 1 = bound (uid 1), main.py:9-10
  Grid grid;
  NSearch ns;
  T dt = T(0);
  std::string err;

  explicit Sim(const RawInit& in) {
    c = make_consts(in);
    km.h = T(c.kernel_h[1]);
    km.h2 = T(c.kernel_h[2]);
    km.sig = T(c.kernel_sig[c.dim]);
    km.dim = c.dim;
    obj[0].alloc(in.fluid_max, c.dim, c.P);
    obj[1].alloc(in.bound_max, c.dim, c.P);
    obj[0].uid = 0;
    obj[1].uid = 1;
    int maxp = in.fluid_max + in.bound_max;  // sph_obj.py:279
    grid.node_part_count.assign(c.node_num, 0);
    grid.node_part_shift.assign(c.node_num, 0);
    grid.node_part_shift_count.assign(c.node_num, 0);
    grid.part_pid_in_node.assign(maxp, 0);
    grid.part_uid_in_node.assign(maxp, 0);
  }

  int D() const { return c.dim; }
  // DFSPH/VFSPH aliases, sph_obj.py:65-73
  T X(const Obj<T>& o, int i) const { return c.solver == 1 ? o.rest_volume[i] : o.mass[i]; }
  T sph_psi(const Obj<T>& o, int i) const {
    return c.solver == 1 ? o.sph_compression[i] : o.sph_density[i];
  }
  // rest_psi is the int field `ones` for VFSPH (value 1), rest_density for DFSPH
  T rest_psi(const Obj<T>& o, int i) const { return c.solver == 1 ? T(1) : o.rest_density[i]; }

  T norm_of(const T* x) const {
    T s = x[0] * x[0];
    for (int d = 1; d < D(); ++d) s = s + x[d] * x[d];
    return std::sqrt(s);
  }

  // node_encode, sph_util.py:65-67: int((pos - lb) // h), floor division
  void node_encode(const T* p, int* cell) const {
    for (int d = 0; d < D(); ++d) {
      T q = (p[d] - T(c.lb[d])) / km.h;
      cell[d] = to_i32(std::floor(q));
    }
  }
  // dim_encode, sph_util.py:70-72: coder . cell (int32, wrapping)
  int dim_encode(const int* cell) const {
    int s = wrap_mul(c.coder[0], cell[0]);
    for (int d = 1; d < D(); ++d) s = wrap_add(s, wrap_mul(c.coder[d], cell[d]));
    return s;
  }

  // The neighbour-walk idiom of every (ngrid, obj, nobj) kernel, sph.py:9-19.
  template <class F>
  void walk(const Obj<T>& o, int i, const Obj<T>& no, F&& f) const {
    const int* ci = &o.neighb_cell_structured_seq[(size_t)i * D()];
    for (int t = 0; t < c.ntemp; ++t) {
      int cell[3] = {0, 0, 0};
      for (int d = 0; d < D(); ++d) cell[d] = wrap_add(ci[d], c.tmpl[t][d]);
      int code = dim_encode(cell);
      if (0 < code && code < c.node_num) {
        int cnt = grid.node_part_count[code];
        for (int j = 0; j < cnt; ++j) {
          int shift = grid.node_part_shift[code] + j;
          if (grid.part_uid_in_node[shift] == no.uid) f(grid.part_pid_in_node[shift]);
        }
      }
    }
  }

  // map_velocity, sph.py:382-397, over a Grid (sph_obj.py:328-344): node I
  // sits at lb + I * dist (the field's float arithmetic); its velocity is the
  // sum of X_j / psi_j * vel_j * W(|x_I - x_j|) over the candidates of object
  // oi in the cells around the node's own cell, in template order.  Nodes
  // are row-major (last axis fastest); out is [nodes][dim].
  void map_velocity(int oi, const int* shape, T dist, T* out) {
    const Obj<T>& no = obj[oi];
    const int d = D();
    long total = 1;
    for (int a = 0; a < d; ++a) total *= shape[a];
    par_for((int)total, [&](int idx) {
      int I[3] = {0, 0, 0};
      long r = idx;
      for (int a = d - 1; a >= 0; --a) {
        I[a] = (int)(r % shape[a]);
        r /= shape[a];
      }
      T gp[3] = {T(0), T(0), T(0)};
      for (int a = 0; a < d; ++a) gp[a] = T(c.lb[a]) + T(I[a]) * dist;
      int node[3] = {0, 0, 0};
      node_encode(gp, node);
      T v[3] = {T(0), T(0), T(0)};
      for (int t = 0; t < c.ntemp; ++t) {
        int cell[3] = {0, 0, 0};
        for (int a = 0; a < d; ++a) cell[a] = wrap_add(node[a], c.tmpl[t][a]);
        int code = dim_encode(cell);
        if (!(0 < code && code < c.node_num)) continue;
        int cnt = grid.node_part_count[code];
        for (int q = 0; q < cnt; ++q) {
          int shift = grid.node_part_shift[code] + q;
          if (grid.part_uid_in_node[shift] != no.uid) continue;
          int j = grid.part_pid_in_node[shift];
          T x[3] = {T(0), T(0), T(0)};
          for (int a = 0; a < d; ++a) x[a] = gp[a] - no.pos[(size_t)j * d + a];
          T w = km.W(norm_of(x));
          T sc = X(no, j) / sph_psi(no, j);
          for (int a = 0; a < d; ++a) v[a] = v[a] + (sc * no.vel[(size_t)j * d + a]) * w;
        }
      }
      for (int a = 0; a < d; ++a) out[(size_t)idx * d + a] = v[a];
    });
  }

  template <class F>
  void par_for(int n, F&& f) {
#ifdef _OPENMP
    if (g_threads > 1) {
#pragma omp parallel for schedule(static) num_threads(g_threads)
      for (int i = 0; i < n; ++i) f(i);
      return;
    }
#endif
    for (int i = 0; i < n; ++i) f(i);
  }

  // ---------------- Ngrid, sph_obj.py:295-323 ----------------
  void clear_node() {  // 295-298
    std::fill(grid.node_part_count.begin(), grid.node_part_count.end(), 0);
  }
  void encode(Obj<T>& o) {  // 300-306
    par_for(o.n, [&](int i) {
      int* cs = &o.neighb_cell_structured_seq[(size_t)i * D()];
      node_encode(&o.pos[(size_t)i * D()], cs);
      o.neighb_cell_seq[i] = dim_encode(cs);
      int code = o.neighb_cell_seq[i];
      if (0 < code && code < c.node_num) {
        std::atomic_ref<int>(grid.node_part_count[code]).fetch_add(1, std::memory_order_relaxed);
      }
    });
  }
  void mem_shift() {  // 308-314: atomic reservation; a prefix sum when serial
    std::atomic<int> sum{0};
    par_for(c.node_num, [&](int i) {
      grid.node_part_shift[i] = sum.fetch_add(grid.node_part_count[i], std::memory_order_relaxed);
      grid.node_part_shift_count[i] = grid.node_part_shift[i];
    });
  }
  void fill_node(Obj<T>& o) {  // 316-323
    par_for(o.n, [&](int i) {
      int code = o.neighb_cell_seq[i];
      if (0 < code && code < c.node_num) {
        int s = std::atomic_ref<int>(grid.node_part_shift_count[code])
                    .fetch_add(1, std::memory_order_relaxed);
        o.neighb_in_cell_seq[i] = s;
        grid.part_pid_in_node[s] = i;
        grid.part_uid_in_node[s] = o.uid;
      }
    });
  }

  // ---------------- NeighborSearch.py:31-132 ----------------
  void ns_establish() {
    int U = 2;
    ns.uids = U;
    ns.node_part_count.assign((size_t)c.node_num * U, 0);
    ns.node_part_shift.assign((size_t)c.node_num * U, 0);
    ns.node_part_shift_count.assign((size_t)c.node_num * U, 0);
    ns.part_pid_in_node.assign((size_t)obj[0].max + obj[1].max, 0);
    // encode (54-60)
    for (int u = 0; u < U; ++u) {
      Obj<T>& o = obj[u];
      for (int i = 0; i < o.n; ++i) {
        int cell[3] = {0, 0, 0};
        node_encode(&o.pos[(size_t)i * D()], cell);
        int code = dim_encode(cell);
        if (0 < code && code < c.node_num) ns.node_part_count[(size_t)code * U + o.uid] += 1;
      }
    }
    // mem_shift (63-70), serial
    int sum = 0;
    for (int i = 0; i < c.node_num; ++i)
      for (int j = 0; j < U; ++j) {
        ns.node_part_shift[(size_t)i * U + j] = sum;
        sum += ns.node_part_count[(size_t)i * U + j];
        ns.node_part_shift_count[(size_t)i * U + j] = ns.node_part_shift[(size_t)i * U + j];
      }
    // fill_node (73-81)
    for (int u = 0; u < U; ++u) {
      Obj<T>& o = obj[u];
      for (int i = 0; i < o.n; ++i) {
        int cell[3] = {0, 0, 0};
        node_encode(&o.pos[(size_t)i * D()], cell);
        int code = dim_encode(cell);
        if (0 < code && code < c.node_num) {
          int s = ns.node_part_shift_count[(size_t)code * U + o.uid]++;
          ns.part_pid_in_node[s] = i;
        }
      }
    }
    // fill_neighbs (84-101)
    ns.neighbs.assign(U, {});
    ns.neighbs_start.assign(U, {});
    ns.neighbs_shift.assign(U, {});
    ns.neighbs_count.assign(U, {});
    for (int u = 0; u < U; ++u) {
      Obj<T>& o = obj[u];
      auto& nb = ns.neighbs[u];
      auto& st = ns.neighbs_start[u];
      auto& sh = ns.neighbs_shift[u];
      auto& ct = ns.neighbs_count[u];
      st.assign(o.n + 1, 0);
      sh.assign((size_t)o.n * U, 0);
      ct.assign((size_t)o.n * U, 0);
      for (int i = 0; i < o.n; ++i) {
        st[i] = (int)nb.size();
        int node[3] = {0, 0, 0};
        node_encode(&o.pos[(size_t)i * D()], node);
        int count = 0;
        for (int k = 0; k < U; ++k) {
          sh[(size_t)i * U + k] = count;
          for (int t = 0; t < c.ntemp; ++t) {
            int cell[3] = {0, 0, 0};
            for (int d = 0; d < D(); ++d) cell[d] = wrap_add(node[d], c.tmpl[t][d]);
            int code = dim_encode(cell);
            if (0 < code && code < c.node_num) {
              int cnt = ns.node_part_count[(size_t)code * U + k];
              for (int j = 0; j < cnt; ++j) {
                int shift = ns.node_part_shift[(size_t)code * U + k] + j;
                nb.push_back(ns.part_pid_in_node[shift]);
                count += 1;
              }
            }
          }
          ct[(size_t)i * U + k] = count - sh[(size_t)i * U + k];
        }
      }
      st[o.n] = (int)nb.size();
    }
  }

  // ---------------- sph.py kernels ----------------
  void SPH_clean_value(Obj<T>& o) {  // sph.py:22-38
    o.general_flag = 1;
    int d = D(), P = c.P;
    par_for(o.n, [&](int i) {
      o.W[i] = 0; o.sph_compression[i] = 0; o.sph_density[i] = 0; o.alpha_2[i] = 0; o.flag[i] = 0;
      for (int j = 0; j < d; ++j) {
        o.W_grad[(size_t)i * d + j] = 0; o.acce[(size_t)i * d + j] = 0;
        o.acce_adv[(size_t)i * d + j] = 0; o.alpha_1[(size_t)i * d + j] = 0;
        o.pressure_force[(size_t)i * d + j] = 0;
        for (int k = 0; k < P; ++k) o.drift_vel[((size_t)i * P + k) * d + j] = 0;
      }
    });
  }

  void cfl_condition(Obj<T>& o) {  // sph.py:41-47
    T dtv = T(c.part_size[1]) / T(c.cs);
    for (int i = 0; i < o.n; ++i) {
      T vn = norm_of(&o.vel[(size_t)i * D()]);
      if (vn > T(1e-4)) dtv = std::min(dtv, T(c.part_size[1]) / vn * T(c.cfl_factor));
    }
    dt = dtv;
  }

  void SPH_prepare_attr(Obj<T>& o, Obj<T>& no) {  // sph.py:50-64
    int d = D();
    par_for(o.n, [&](int i) {
      walk(o, i, no, [&](int j) {
        T x[3];
        for (int k = 0; k < d; ++k) x[k] = o.pos[(size_t)i * d + k] - no.pos[(size_t)j * d + k];
        T Wr = km.W(norm_of(x));
        o.W[i] = o.W[i] + Wr;
        o.sph_compression[i] = o.sph_compression[i] + Wr * no.rest_volume[j];
        o.sph_density[i] = o.sph_density[i] + Wr * no.mass[j];
      });
    });
  }

  void SPH_prepare_alpha_1(Obj<T>& o, Obj<T>& no) {  // sph.py:67-81
    int d = D();
    par_for(o.n, [&](int i) {
      walk(o, i, no, [&](int j) {
        T x[3];
        for (int k = 0; k < d; ++k) x[k] = o.pos[(size_t)i * d + k] - no.pos[(size_t)j * d + k];
        T r = norm_of(x);
        if (r > T(0)) {
          T Xj = X(no, j), wg = km.W_grad(r);
          for (int k = 0; k < d; ++k)
            o.alpha_1[(size_t)i * d + k] = o.alpha_1[(size_t)i * d + k] + Xj * x[k] / r * wg;
        }
      });
    });
  }

  void SPH_prepare_alpha_2(Obj<T>& o, Obj<T>& no) {  // sph.py:84-97
    int d = D();
    par_for(o.n, [&](int i) {
      walk(o, i, no, [&](int j) {
        T x[3];
        for (int k = 0; k < d; ++k) x[k] = o.pos[(size_t)i * d + k] - no.pos[(size_t)j * d + k];
        T r = norm_of(x);
        if (r > T(0)) {
          T wg = km.W_grad(r), Xj = X(no, j);
          o.alpha_2[i] = o.alpha_2[i] + pow2(wg) * pow2(Xj) / no.mass[j];
        }
      });
    });
  }

  void SPH_prepare_alpha(Obj<T>& o) {  // sph.py:100-105
    int d = D();
    par_for(o.n, [&](int i) {
      const T* a1 = &o.alpha_1[(size_t)i * d];
      T dot = a1[0] * a1[0];
      for (int k = 1; k < d; ++k) dot = dot + a1[k] * a1[k];
      o.alpha[i] = dot / o.mass[i] + o.alpha_2[i];
      if (o.alpha[i] < T(1e-4)) o.alpha[i] = T(1e-4);
    });
  }

  void SPH_advection_gravity_acc(Obj<T>& o) {  // sph.py:108-111
    int d = D();
    par_for(o.n, [&](int i) {
      for (int k = 0; k < d; ++k) o.acce_adv[(size_t)i * d + k] = o.acce_adv[(size_t)i * d + k] + T(c.g[k]);
    });
  }

  void SPH_advection_viscosity_acc(Obj<T>& o, Obj<T>& no) {  // sph.py:114-129
    int d = D();
    T two_d2 = T(2 * (2 + d));
    T eps = T(0.01) * km.h2;
    par_for(o.n, [&](int i) {
      T mu = T(c.mu);
      if (c.mu_pp) {  // extension: mu_i = sum_k phi_ik mu_k, k ascending
        const T* phi = &o.volume_frac[(size_t)i * c.P];
        mu = phi[0] * T(c.muk[0]);
        for (int k = 1; k < c.P; ++k) mu = mu + phi[k] * T(c.muk[k]);
      }
      walk(o, i, no, [&](int j) {
        T x[3], A[3];
        for (int k = 0; k < d; ++k) {
          x[k] = o.pos[(size_t)i * d + k] - no.pos[(size_t)j * d + k];
          A[k] = o.vel[(size_t)i * d + k] - no.vel[(size_t)j * d + k];
        }
        T r = norm_of(x);
        if (r > T(0)) {
          T Vj = X(no, j) / sph_psi(no, j);
          // W_lap, sph_util.py:40-43
          T s = two_d2 * Vj * km.W_grad(r);
          T inv = T(1) / norm_of(x);  // normalized(): (1/norm) * x
          T Ax = A[0] * x[0];
          for (int k = 1; k < d; ++k) Ax = Ax + A[k] * x[k];
          T den = eps + r * r;
          for (int k = 0; k < d; ++k) {
            T lap = s * (inv * x[k]) * Ax / den;
            o.acce_adv[(size_t)i * d + k] =
                o.acce_adv[(size_t)i * d + k] + lap * mu / o.rest_density[i];
          }
        }
      });
    });
  }

  void SPH_advection_surface_tension_acc(Obj<T>& o, Obj<T>& no) {  // sph.py:132-166
    int d = D();
    int P = c.P;
    par_for(o.n, [&](int i) {  // loop 1: normals
      for (int k = 0; k < d; ++k) o.normal[(size_t)i * d + k] = 0;
      walk(o, i, no, [&](int j) {
        T x[3];
        for (int k = 0; k < d; ++k) x[k] = o.pos[(size_t)i * d + k] - no.pos[(size_t)j * d + k];
        T r = norm_of(x);
        if (r > T(0)) {
          T s = -X(no, j) / sph_psi(no, j) * km.W_grad(r);
          for (int k = 0; k < d; ++k)
            o.normal[(size_t)i * d + k] = o.normal[(size_t)i * d + k] + s * x[k] / r;
        }
      });
      for (int k = 0; k < d; ++k) o.normal[(size_t)i * d + k] = o.normal[(size_t)i * d + k] * km.h;
    });
    T gam = T(c.gamma);
    par_for(o.n, [&](int i) {  // loop 2: cohesion + curvature
      walk(o, i, no, [&](int j) {
        T x[3];
        for (int k = 0; k < d; ++k) x[k] = o.pos[(size_t)i * d + k] - no.pos[(size_t)j * d + k];
        T r = norm_of(x);
        if (r > T(0) && o.volume_frac[(size_t)i * P] > T(0.99) && no.volume_frac[(size_t)j * P] > T(0.99)) {
          T cs = -gam * no.mass[j] * km.C(r);
          T f = T(2) * rest_psi(o, i) / (sph_psi(o, i) + sph_psi(no, j));
          for (int k = 0; k < d; ++k) {
            T coh = cs * x[k] / r;
            T curv = gam * (o.normal[(size_t)i * d + k] - no.normal[(size_t)j * d + k]);
            o.acce_adv[(size_t)i * d + k] = o.acce_adv[(size_t)i * d + k] + f * (coh + curv);
          }
        }
      });
    });
  }

  void WC_pressure_val(Obj<T>& o) {  // sph.py:169-175
    par_for(o.n, [&](int i) {
      T cs2 = T(pow2(c.cs));  // cs is an int field: cs**2 is an int
      T ratio = o.sph_density[i] / o.rest_density[i];
      T r7 = ratio;  // ratio**7 by squaring: r * r^2 * r^4
      {
        T res = T(1), a = ratio;
        int e = 7;
        while (e > 0) {
          if (e & 1) res = res * a;
          a = a * a;
          e >>= 1;
        }
        r7 = res;
      }
      o.pressure[i] = (o.rest_density[i] * cs2 / T(c.wc_gamma)) * (r7 - T(1));
      if (o.pressure[i] < T(0)) o.pressure[i] = T(0);
    });
  }

  void WC_pressure_acce(Obj<T>& o, Obj<T>& no) {  // sph.py:178-193
    int d = D();
    par_for(o.n, [&](int i) {
      walk(o, i, no, [&](int j) {
        T x[3];
        for (int k = 0; k < d; ++k) x[k] = o.pos[(size_t)i * d + k] - no.pos[(size_t)j * d + k];
        T r = norm_of(x);
        T p_term = o.pressure[i] / pow2(o.sph_density[i]) + no.pressure[j] / pow2(no.sph_density[j]);
        if (r > T(0)) {
          T wg = km.W_grad(r);
          for (int k = 0; k < d; ++k)
            o.acce_adv[(size_t)i * d + k] =
                o.acce_adv[(size_t)i * d + k] + -p_term * no.mass[j] * x[k] / r * wg;
        }
      });
    });
  }

  void IPPE_adv_psi_init(Obj<T>& o) {  // sph.py:196-199
    par_for(o.n, [&](int i) { o.psi_adv[i] = sph_psi(o, i) - rest_psi(o, i); });
  }

  void IPPE_adv_psi(Obj<T>& o, Obj<T>& no) {  // sph.py:202-216
    int d = D();
    T dtv = dt;
    par_for(o.n, [&](int i) {
      walk(o, i, no, [&](int j) {
        T x[3];
        for (int k = 0; k < d; ++k) x[k] = o.pos[(size_t)i * d + k] - no.pos[(size_t)j * d + k];
        T r = norm_of(x);
        if (r > T(0)) {
          T wg = km.W_grad(r);
          T dot = T(0);
          for (int k = 0; k < d; ++k) {
            T g = x[k] / r * wg;
            T dv = o.vel_adv[(size_t)i * d + k] - no.vel_adv[(size_t)j * d + k];
            dot = k == 0 ? g * dv : dot + g * dv;
          }
          o.psi_adv[i] = o.psi_adv[i] + dot * X(no, j) * dtv;
        }
      });
    });
  }

  void IPPE_psi_adv_non_negative(Obj<T>& o) {  // sph.py:219-226
    double sum = 0;
    for (int i = 0; i < o.n; ++i) {
      if (o.psi_adv[i] < T(0)) o.psi_adv[i] = T(0);
      sum += (double)(o.psi_adv[i] / rest_psi(o, i));
    }
    o.compression = T(sum) / T(o.n);
  }

  void IPPE_update_vel_adv(Obj<T>& o, Obj<T>& no) {  // sph.py:229-244
    int d = D();
    T inv_dt = -(T(1) / dt);
    par_for(o.n, [&](int i) {
      walk(o, i, no, [&](int j) {
        T x[3];
        for (int k = 0; k < d; ++k) x[k] = o.pos[(size_t)i * d + k] - no.pos[(size_t)j * d + k];
        T r = norm_of(x);
        if (r > T(0)) {
          T a = o.psi_adv[i] * X(no, j) / o.alpha[i];
          T b = no.psi_adv[j] * X(o, i) / no.alpha[j];
          T s = inv_dt * (a + b);
          T wg = km.W_grad(r);
          for (int k = 0; k < d; ++k) {
            T g = x[k] / r * wg;
            o.vel_adv[(size_t)i * d + k] = o.vel_adv[(size_t)i * d + k] + s * g / o.mass[i];
          }
        }
      });
    });
  }

  void SPH_advection_update_vel_adv(Obj<T>& o) {  // sph.py:247-250
    int d = D();
    par_for(o.n, [&](int i) {
      for (int k = 0; k < d; ++k)
        o.vel_adv[(size_t)i * d + k] = o.vel[(size_t)i * d + k] + o.acce_adv[(size_t)i * d + k] * dt;
    });
  }
  void SPH_vel_2_vel_adv(Obj<T>& o) {  // sph.py:253-256
    par_for(o.n * D(), [&](int i) { o.vel_adv[i] = o.vel[i]; });
  }
  void SPH_vel_adv_2_vel(Obj<T>& o) {  // sph.py:259-262
    par_for(o.n * D(), [&](int i) { o.vel[i] = o.vel_adv[i]; });
  }
  void SPH_update_pos(Obj<T>& o) {  // sph.py:265-269
    int d = D();
    par_for(o.n, [&](int i) {
      for (int k = 0; k < d; ++k) {
        o.vel[(size_t)i * d + k] = o.vel_adv[(size_t)i * d + k];
        o.pos[(size_t)i * d + k] = o.pos[(size_t)i * d + k] + o.vel[(size_t)i * d + k] * dt;
      }
    });
  }
  void SPH_update_energy(Obj<T>& o) {  // sph.py:272-279
    int d = D();
    double ke = 0, pe = 0;
    for (int i = 0; i < o.n; ++i) {
      const T* v = &o.vel[(size_t)i * d];
      T n2 = v[0] * v[0];
      for (int k = 1; k < d; ++k) n2 = n2 + v[k] * v[k];
      ke += (double)(T(0.5) * o.mass[i] * n2);
      pe += (double)(-o.mass[i] * T(c.g[1]) * (o.pos[(size_t)i * d + 1] - T(c.lb[1])));
    }
    o.ke = ke;
    o.pe = pe;
  }
  T phase_dot(const Obj<T>& o, int i) const {
    const T* phi = &o.volume_frac[(size_t)i * c.P];
    T s = T(c.rho[0]) * phi[0];
    for (int k = 1; k < c.P; ++k) s = s + T(c.rho[k]) * phi[k];
    return s;
  }
  void SPH_update_mass(Obj<T>& o) {  // sph.py:282-286
    par_for(o.n, [&](int i) {
      o.rest_density[i] = phase_dot(o, i);
      o.mass[i] = o.rest_density[i] * o.rest_volume[i];
    });
  }
  void SPH_FBM_clean_tmp(Obj<T>& o) {  // sph.py:301-305
    par_for(o.n * c.P, [&](int i) { o.volume_frac_tmp[i] = 0; });
  }
  void SPH_FBM_diffuse(Obj<T>& o, Obj<T>& no) {  // sph.py:308-326
    int d = D(), P = c.P;
    T eps = T(0.01) * km.h2;
    T pre = dt * T(c.fbm_D);
    par_for(o.n, [&](int i) {
      if (o.flag[i] != 0) return;
      walk(o, i, no, [&](int j) {
        if (no.flag[j] != 0) return;
        T x[3];
        for (int k = 0; k < d; ++k) x[k] = o.pos[(size_t)i * d + k] - no.pos[(size_t)j * d + k];
        T r = norm_of(x);
        if (r > T(0)) {
          T wg = km.W_grad(r);
          T den = r * r + eps;
          for (int k = 0; k < P; ++k) {
            T dphi = o.volume_frac[(size_t)i * P + k] - no.volume_frac[(size_t)j * P + k];
            T tmp = pre * dphi * no.rest_volume[j] * r * wg / den;
            o.volume_frac_tmp[(size_t)i * P + k] = o.volume_frac_tmp[(size_t)i * P + k] + tmp;
          }
        }
      });
    });
  }
  void SPH_FBM_convect(Obj<T>& o, Obj<T>& no) {  // sph.py:329-363
    int d = D(), P = c.P;
    par_for(o.n, [&](int i) {  // loop 1: drift velocities
      T* aa = &o.acce_adv[(size_t)i * d];
      T* fa = &o.acce[(size_t)i * d];  // fbm_acce is an alias of acce (sph_obj.py:57)
      for (int k = 0; k < d; ++k)
        aa[k] = (o.vel_adv[(size_t)i * d + k] - o.vel[(size_t)i * d + k]) / dt;
      o.fbm_zeta[i] = 0;
      const T* phi = &o.volume_frac[(size_t)i * P];
      for (int k = 0; k < P; ++k)
        o.fbm_zeta[i] = o.fbm_zeta[i] + phi[k] * (T(c.rho[k]) - o.rest_density[i]) / T(c.rho[k]);
      T z = o.fbm_zeta[i];
      for (int k = 0; k < d; ++k) fa[k] = (aa[k] - z * T(c.g[k])) / (T(1) - z);
      for (int p = 0; p < P; ++p) {
        T* u = &o.drift_vel[((size_t)i * P + p) * d];
        T s = phi[p] * (T(c.rho[p]) - o.rest_density[i]);
        for (int k = 0; k < d; ++k) u[k] = s * (T(c.g[k]) - fa[k]);
        T w = phi[p] * T(c.rho[p]);
        if (w > T(1e-6)) {
          for (int k = 0; k < d; ++k) u[k] = u[k] / w;
        } else {
          for (int k = 0; k < d; ++k) u[k] = u[k] * T(0);
        }
        for (int k = 0; k < d; ++k) u[k] = u[k] + (fa[k] - aa[k]);
        for (int k = 0; k < d; ++k) u[k] = u[k] * dt;
      }
    });
    T pre = T(c.fbm_C) * dt;
    par_for(o.n, [&](int i) {  // loop 2: convective flux
      if (o.flag[i] != 0) return;
      walk(o, i, no, [&](int j) {
        if (no.flag[j] != 0) return;
        T x[3];
        for (int k = 0; k < d; ++k) x[k] = o.pos[(size_t)i * d + k] - no.pos[(size_t)j * d + k];
        T r = norm_of(x);
        if (r > T(0)) {
          T wg = km.W_grad(r);
          for (int p = 0; p < P; ++p) {
            const T* ui = &o.drift_vel[((size_t)i * P + p) * d];
            const T* uj = &no.drift_vel[((size_t)j * P + p) * d];
            T pi = o.volume_frac[(size_t)i * P + p], pj = no.volume_frac[(size_t)j * P + p];
            T dot = T(0);
            for (int k = 0; k < d; ++k) {
              T f = pi * ui[k] + pj * uj[k];
              T e = x[k] / r;
              dot = k == 0 ? f * e : dot + f * e;
            }
            T tmp = pre * no.rest_volume[j] * dot * wg;
            o.volume_frac_tmp[(size_t)i * P + p] = o.volume_frac_tmp[(size_t)i * P + p] - tmp;
          }
        }
      });
    });
  }
  void SPH_FBM_check_tmp(Obj<T>& o) {  // sph.py:366-372
    o.general_flag = 0;
    for (int i = 0; i < o.n; ++i) {
      bool neg = false;  // has_negative, sph_util.py:56-62
      for (int k = 0; k < c.P; ++k)
        if (o.volume_frac[(size_t)i * c.P + k] + o.volume_frac_tmp[(size_t)i * c.P + k] < T(0)) neg = true;
      if (neg) {
        o.flag[i] = 1;
        o.general_flag = 1;
      }
    }
  }
  void SPH_update_volume_frac(Obj<T>& o) {  // sph.py:375-379
    par_for(o.n, [&](int i) {
      if (!(o.flag[i] > 0))
        for (int k = 0; k < c.P; ++k)
          o.volume_frac[(size_t)i * c.P + k] =
              o.volume_frac[(size_t)i * c.P + k] + o.volume_frac_tmp[(size_t)i * c.P + k];
    });
  }

  // main.py:175-256 (sph_step), preceded by cfl_condition as run_step does (348).
  void sph_step(int flags, StepStats* st) {
    Obj<T>& fluid = obj[0];
    Obj<T>& bound = obj[1];
    std::memset(st, 0, sizeof(*st));
    if (flags & kStepCfl) cfl_condition(fluid);
    // neighbour search (178-183)
    clear_node();
    encode(fluid);
    encode(bound);
    mem_shift();
    fill_node(fluid);
    fill_node(bound);
    // clean (185-186)
    SPH_clean_value(fluid);
    SPH_clean_value(bound);
    // density and alpha (188-197)
    SPH_prepare_attr(fluid, fluid);
    SPH_prepare_attr(fluid, bound);
    SPH_prepare_attr(bound, bound);
    SPH_prepare_attr(bound, fluid);
    SPH_prepare_alpha_1(fluid, fluid);
    SPH_prepare_alpha_1(fluid, bound);
    SPH_prepare_alpha_2(fluid, fluid);
    SPH_prepare_alpha_2(bound, fluid);
    SPH_prepare_alpha(fluid);
    SPH_prepare_alpha(bound);
    // divergence solve (199-214)
    int it = 0;
    SPH_vel_2_vel_adv(fluid);
    while (it < c.iter_min || fluid.compression > T(c.div_thr)) {
      IPPE_adv_psi_init(fluid);
      IPPE_adv_psi(fluid, fluid);
      IPPE_adv_psi(fluid, bound);
      IPPE_psi_adv_non_negative(fluid);
      if (it < 128) st->div_err[it] = (double)fluid.compression;
      IPPE_update_vel_adv(fluid, fluid);
      IPPE_update_vel_adv(fluid, bound);
      it += 1;
      if (it > c.iter_max) break;
    }
    st->div_iter = it;
    SPH_vel_adv_2_vel(fluid);
    // advection (216-219)
    SPH_advection_gravity_acc(fluid);
    SPH_advection_viscosity_acc(fluid, fluid);
    SPH_advection_surface_tension_acc(fluid, fluid);
    SPH_advection_update_vel_adv(fluid);
    // density solve (221-234)
    it = 0;
    while (it < c.iter_min || fluid.compression > T(c.comp_thr)) {
      IPPE_adv_psi_init(fluid);
      IPPE_adv_psi(fluid, fluid);
      IPPE_adv_psi(fluid, bound);
      IPPE_psi_adv_non_negative(fluid);
      if (it < 128) st->incom_err[it] = (double)fluid.compression;
      IPPE_update_vel_adv(fluid, fluid);
      IPPE_update_vel_adv(fluid, bound);
      it += 1;
      if (it > c.iter_max) break;
    }
    st->incom_iter = it;
    if (flags & kStepWcsph) {  // 238-242
      WC_pressure_val(fluid);
      WC_pressure_acce(fluid, fluid);
      WC_pressure_acce(fluid, bound);
      SPH_advection_update_vel_adv(fluid);
    }
    if (flags & kStepFbm) {  // 243-250
      int rounds = 0;
      while (fluid.general_flag > 0 && rounds < kFbmMaxRounds) {
        SPH_FBM_clean_tmp(fluid);
        SPH_FBM_convect(fluid, fluid);
        SPH_FBM_diffuse(fluid, fluid);
        SPH_FBM_check_tmp(fluid);
        ++rounds;
      }
      st->fbm_rounds = rounds;
      SPH_update_volume_frac(fluid);
    }
    // update (251-253)
    SPH_update_mass(fluid);
    SPH_update_pos(fluid);
    SPH_update_energy(fluid);
    st->dt = (double)dt;
    st->ke = fluid.ke;
    st->pe = fluid.pe;
  }

  // Field access by the reference's attribute names (sph_obj.py:24-63).
  std::vector<T>* ffield(Obj<T>& o, const std::string& n) {
    if (n == "pos") return &o.pos;
    if (n == "vel") return &o.vel;
    if (n == "vel_adv") return &o.vel_adv;
    if (n == "acce") return &o.acce;
    if (n == "acce_adv") return &o.acce_adv;
    if (n == "mass") return &o.mass;
    if (n == "rest_density") return &o.rest_density;
    if (n == "rest_volume") return &o.rest_volume;
    if (n == "volume_frac") return &o.volume_frac;
    if (n == "volume_frac_tmp") return &o.volume_frac_tmp;
    if (n == "W") return &o.W;
    if (n == "W_grad") return &o.W_grad;
    if (n == "sph_compression") return &o.sph_compression;
    if (n == "sph_density") return &o.sph_density;
    if (n == "psi_adv") return &o.psi_adv;
    if (n == "alpha") return &o.alpha;
    if (n == "alpha_1") return &o.alpha_1;
    if (n == "alpha_2") return &o.alpha_2;
    if (n == "normal") return &o.normal;
    if (n == "drift_vel") return &o.drift_vel;
    if (n == "fbm_zeta") return &o.fbm_zeta;
    if (n == "pressure") return &o.pressure;
    if (n == "pressure_force") return &o.pressure_force;
    return nullptr;
  }
  std::vector<int>* ifield(Obj<T>& o, const std::string& n) {
    if (n == "flag") return &o.flag;
    if (n == "neighb_cell_seq") return &o.neighb_cell_seq;
    if (n == "neighb_in_cell_seq") return &o.neighb_in_cell_seq;
    if (n == "neighb_cell_structured_seq") return &o.neighb_cell_structured_seq;
    return nullptr;
  }
  std::vector<int>* gfield(const std::string& n) {
    if (n == "node_part_count") return &grid.node_part_count;
    if (n == "node_part_shift") return &grid.node_part_shift;
    if (n == "node_part_shift_count") return &grid.node_part_shift_count;
    if (n == "part_pid_in_node") return &grid.part_pid_in_node;
    if (n == "part_uid_in_node") return &grid.part_uid_in_node;
    return nullptr;
  }

  int kernel(const std::string& n, int oi, int ni) {
    Obj<T>& o = obj[oi];
    Obj<T>& no = obj[ni < 0 ? 0 : ni];
    if (n == "clear_node") clear_node();
    else if (n == "encode") encode(o);
    else if (n == "mem_shift") mem_shift();
    else if (n == "fill_node") fill_node(o);
    else if (n == "SPH_clean_value") SPH_clean_value(o);
    else if (n == "cfl_condition") cfl_condition(o);
    else if (n == "SPH_prepare_attr") SPH_prepare_attr(o, no);
    else if (n == "SPH_prepare_alpha_1") SPH_prepare_alpha_1(o, no);
    else if (n == "SPH_prepare_alpha_2") SPH_prepare_alpha_2(o, no);
    else if (n == "SPH_prepare_alpha") SPH_prepare_alpha(o);
    else if (n == "SPH_advection_gravity_acc") SPH_advection_gravity_acc(o);
    else if (n == "SPH_advection_viscosity_acc") SPH_advection_viscosity_acc(o, no);
    else if (n == "SPH_advection_surface_tension_acc") SPH_advection_surface_tension_acc(o, no);
    else if (n == "WC_pressure_val") WC_pressure_val(o);
    else if (n == "WC_pressure_acce") WC_pressure_acce(o, no);
    else if (n == "IPPE_adv_psi_init") IPPE_adv_psi_init(o);
    else if (n == "IPPE_adv_psi") IPPE_adv_psi(o, no);
    else if (n == "IPPE_psi_adv_non_negative") IPPE_psi_adv_non_negative(o);
    else if (n == "IPPE_update_vel_adv") IPPE_update_vel_adv(o, no);
    else if (n == "SPH_advection_update_vel_adv") SPH_advection_update_vel_adv(o);
    else if (n == "SPH_vel_2_vel_adv") SPH_vel_2_vel_adv(o);
    else if (n == "SPH_vel_adv_2_vel") SPH_vel_adv_2_vel(o);
    else if (n == "SPH_update_pos") SPH_update_pos(o);
    else if (n == "SPH_update_energy") SPH_update_energy(o);
    else if (n == "SPH_update_mass") SPH_update_mass(o);
    else if (n == "SPH_FBM_clean_tmp") SPH_FBM_clean_tmp(o);
    else if (n == "SPH_FBM_diffuse") SPH_FBM_diffuse(o, no);
    else if (n == "SPH_FBM_convect") SPH_FBM_convect(o, no);
    else if (n == "SPH_FBM_check_tmp") SPH_FBM_check_tmp(o);
    else if (n == "SPH_update_volume_frac") SPH_update_volume_frac(o);
    else if (n == "establish_neighbs") ns_establish();
    else return -1;
    return 0;
  }
};

struct Handle {
  int precision;  // 4 or 8
  void* sim;
  RawInit init;
};

template <class T>
Sim<T>* S(Handle* h) { return (Sim<T>*)h->sim; }

}  // namespace

namespace {
// data: float array (precision 4) / double array (precision 8) for float
// fields, int32 for int fields.  count is the element count copied.
template <class T>
int set_get(Sim<T>* s, int oi, const char* name, void* data, long count, bool set) {
  std::string n(name);
  if (oi == 2) {
    std::vector<int>* g = s->gfield(n);
    if (!g || count > (long)g->size()) return -1;
    if (set) std::memcpy(g->data(), data, count * sizeof(int));
    else std::memcpy(data, g->data(), count * sizeof(int));
    return 0;
  }
  Obj<T>& o = s->obj[oi];
  if (auto* f = s->ffield(o, n)) {
    if (count > (long)f->size()) return -1;
    if (set) std::memcpy(f->data(), data, count * sizeof(T));
    else std::memcpy(data, f->data(), count * sizeof(T));
    return 0;
  }
  if (auto* f = s->ifield(o, n)) {
    if (count > (long)f->size()) return -1;
    if (set) std::memcpy(f->data(), data, count * sizeof(int));
    else std::memcpy(data, f->data(), count * sizeof(int));
    return 0;
  }
  return -2;
}

}  // namespace

// Constants block returned to Python for the a1 checks.
struct OracleConsts {
  double part_size[5], kernel_h[5], kernel_sig[4];
  double grid_size[3];
  int coder[3];
  int node_num, ntemp;
  int tmpl[125][3];
};

extern "C" {

void oracle_set_threads(int n) { g_threads = n < 1 ? 1 : n; }

void oracle_consts(const RawInit* in, OracleConsts* out) {
  Consts c = make_consts(*in);
  std::memcpy(out->part_size, c.part_size, sizeof(c.part_size));
  std::memcpy(out->kernel_h, c.kernel_h, sizeof(c.kernel_h));
  std::memcpy(out->kernel_sig, c.kernel_sig, sizeof(c.kernel_sig));
  std::memcpy(out->grid_size, c.grid_size, sizeof(c.grid_size));
  std::memcpy(out->coder, c.coder, sizeof(c.coder));
  out->node_num = c.node_num;
  out->ntemp = c.ntemp;
  std::memcpy(out->tmpl, c.tmpl, sizeof(c.tmpl));
}

// Kernel-function probes for the closed-form checks (sph_util.py:4-37).
double oracle_kernel_fn(int which, int precision, double h, double sig, double r) {
  if (precision == 8) {
    KernelMath<double> k{h, h * h, sig, 3};
    return which == 0 ? k.W(r) : which == 1 ? k.W_grad(r) : k.C(r);
  }
  KernelMath<float> k{(float)h, (float)h * (float)h, (float)sig, 3};
  return which == 0 ? k.W((float)r) : which == 1 ? k.W_grad((float)r) : k.C((float)r);
}

void* oracle_new(const RawInit* in, int precision) {
  Handle* h = new Handle{precision, nullptr, *in};
  if (precision == 8)
    h->sim = new Sim<double>(*in);
  else
    h->sim = new Sim<float>(*in);
  return h;
}

void oracle_free(void* p) {
  Handle* h = (Handle*)p;
  if (!h) return;
  if (h->precision == 8)
    delete S<double>(h);
  else
    delete S<float>(h);
  delete h;
}

int oracle_set_part_num(void* p, int oi, int n) {
  Handle* h = (Handle*)p;
  if (h->precision == 8) {
    if (n > S<double>(h)->obj[oi].max) return -1;
    S<double>(h)->obj[oi].n = n;
  } else {
    if (n > S<float>(h)->obj[oi].max) return -1;
    S<float>(h)->obj[oi].n = n;
  }
  return 0;
}

int oracle_set(void* p, int oi, const char* name, const void* data, long count) {
  Handle* h = (Handle*)p;
  return h->precision == 8 ? set_get(S<double>(h), oi, name, (void*)data, count, true)
                           : set_get(S<float>(h), oi, name, (void*)data, count, true);
}
int oracle_get(void* p, int oi, const char* name, void* data, long count) {
  Handle* h = (Handle*)p;
  return h->precision == 8 ? set_get(S<double>(h), oi, name, data, count, false)
                           : set_get(S<float>(h), oi, name, data, count, false);
}

// scalars: 0 compression, 1 dt (set/get), 2 general_flag, 3 ke, 4 pe
double oracle_scalar(void* p, int oi, int which, int set, double value) {
  Handle* h = (Handle*)p;
  auto run = [&](auto* s) -> double {
    auto& o = s->obj[oi];
    using T = std::remove_reference_t<decltype(o.compression)>;
    switch (which) {
      case 0: if (set) o.compression = (T)value; return (double)o.compression;
      case 1: if (set) s->dt = (T)value; return (double)s->dt;
      case 2: if (set) o.general_flag = (int)value; return (double)o.general_flag;
      case 3: return o.ke;
      case 4: return o.pe;
    }
    return 0.0;
  };
  return h->precision == 8 ? run(S<double>(h)) : run(S<float>(h));
}

int oracle_kernel(void* p, const char* name, int oi, int ni) {
  Handle* h = (Handle*)p;
  return h->precision == 8 ? S<double>(h)->kernel(name, oi, ni) : S<float>(h)->kernel(name, oi, ni);
}

int oracle_map_velocity(void* p, int oi, const int* shape, double dist, void* out) {
  Handle* h = (Handle*)p;
  if (h->precision == 8)
    S<double>(h)->map_velocity(oi, shape, dist, (double*)out);
  else
    S<float>(h)->map_velocity(oi, shape, (float)dist, (float*)out);
  return 0;
}

int oracle_sph_step(void* p, int flags, StepStats* st) {
  Handle* h = (Handle*)p;
  if (h->precision == 8)
    S<double>(h)->sph_step(flags, st);
  else
    S<float>(h)->sph_step(flags, st);
  return 0;
}

// NeighborSearch lists after establish_neighbs: total candidate count for
// object oi, then the flattened lists / starts / shift / count arrays.
long oracle_ns_total(void* p, int oi) {
  Handle* h = (Handle*)p;
  if (h->precision == 8) return (long)S<double>(h)->ns.neighbs[oi].size();
  return (long)S<float>(h)->ns.neighbs[oi].size();
}
int oracle_ns_get(void* p, int oi, int* neighbs, int* starts, int* shift, int* count) {
  Handle* h = (Handle*)p;
  auto run = [&](auto* s) {
    auto& ns = s->ns;
    std::memcpy(neighbs, ns.neighbs[oi].data(), ns.neighbs[oi].size() * sizeof(int));
    std::memcpy(starts, ns.neighbs_start[oi].data(), ns.neighbs_start[oi].size() * sizeof(int));
    std::memcpy(shift, ns.neighbs_shift[oi].data(), ns.neighbs_shift[oi].size() * sizeof(int));
    std::memcpy(count, ns.neighbs_count[oi].data(), ns.neighbs_count[oi].size() * sizeof(int));
  };
  if (h->precision == 8) run(S<double>(h));
  else run(S<float>(h));
  return 0;
}

// Initial-state builders, sph_obj.py:102-114, 140-151, 158-180, restated with
// the numpy dtypes the reference uses: the lattice count is computed in f32,
// the padding and positions in f64, then stored to the f32 field.
// mode 0 = cube, 1 = hollow box with `layers`.  Returns the particle count and
// writes positions (as double, already rounded to f32 values) when out != 0.
long oracle_lattice(int dim, double part_size_f32, double relaxing, const double* start,
                    const double* end, int mode, int layers, double* out) {
  double spacing = part_size_f32 * relaxing;  // sph_obj.py:112 / 141 (f64 product)
  int shape[3] = {1, 1, 1};
  double padding[3] = {0, 0, 0};
  for (int d = 0; d < dim; ++d) {
    float e = (float)end[d], s = (float)start[d];
    float q = (e - s + (float)1e-7) / (float)spacing;  // f32 arithmetic (sph_obj.py:105)
    shape[d] = (int)q;
    padding[d] = ((double)(e - s) - (double)shape[d] * spacing) / 2;  // f64 (sph_obj.py:106)
  }
  long n = 0;
  // np.where enumerates in C order: the last axis varies fastest.
  int ix[3];
  for (ix[0] = 0; ix[0] < shape[0]; ++ix[0])
    for (ix[1] = 0; ix[1] < (dim > 1 ? shape[1] : 1); ++ix[1])
      for (ix[2] = 0; ix[2] < (dim > 2 ? shape[2] : 1); ++ix[2]) {
        bool keep = true;
        if (mode == 1) {  // sph_obj.py:143-148: interior cleared
          bool inside = true;
          for (int d = 0; d < dim; ++d)
            inside = inside && ix[d] >= layers && ix[d] < shape[d] - layers;
          keep = !inside;
        }
        if (!keep) continue;
        if (out)
          for (int d = 0; d < dim; ++d) {
            double sp = start[d] + padding[d];  // list + f64 array (sph_obj.py:114/151)
            out[n * dim + d] = (double)(float)((double)ix[d] * spacing + sp);  // f64, then f32 store
          }
        ++n;
      }
  return n;
}

}  // extern "C"
