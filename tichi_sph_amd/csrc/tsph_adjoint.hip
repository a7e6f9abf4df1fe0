// Reverse mode of the fused DFSPH step (one phase, fixed dt, no FBM).
//
// The forward step (tsph_step.hip) is a chain of neighbour sums
//   out_i = sum_j F(x_ij, a_i, a_j)
// over the symmetric in-range lists (j in N(i) <=> i in N(j) for fluid
// pairs).  The adjoint of such a sum with respect to particle l collects
// the pair (l, j) from l's own equation and the pair (j, l) from j's
// equation; both are visible from l's list, so every reverse sweep is a
// gather over l's neighbours: no scatter, no atomics, a fixed order.
//
// Notation.  x = x_l - x_j, r = |x|, g = W'(r)/r so that grad W = g x, and
// J = dG/dx for G(x) = g x:  J w = g w + hq (x.w) x with hq = (W'' - g)/r^2.
// For a term c . G(x_lj) in l's equation and c' . G(x_jl) in j's equation,
// d/dx_l = J (c - c').
//
// The forward schedule being reversed (sph.py line numbers):
//   pass A: rho, alpha_1, alpha_2, alpha (50-105), vel_adv = vel (253-256)
//   divergence solve iterations (196-244)
//   advection: viscosity, normals, surface tension, vel_adv = vel + dt a
//     (108-166, 247-250)
//   density solve iterations (196-244)
//   tail: vel = vel_adv, pos += dt vel (259-269)
#include "tsph_sweep.h"

namespace tsph {

// g, hq and W'' of the cubic spline at r^2 (sph_util.py:28-37 differentiated
// once more): q <= 1/2: W' = 6 sig/h (3q^2 - 2q), W'' = 6 sig/h^2 (6q - 2);
// 1/2 < q < 1: W' = -6 sig/h (1-q)^2, W'' = 12 sig/h^2 (1-q).
template <class T>
struct D2 {
  T g, hq, w2, r;
};
template <class T>
__device__ __forceinline__ D2<T> kern_d2(const KP<T>& k, T r2) {
  T rinv = f_rsqrt(r2);
  T r = r2 * rinv;
  T q = r * k.inv_h;
  T u = T(1) - q;
  D2<T> d;
  d.r = r;
  if (q <= T(0.5)) {
    d.g = k.wg_a * (T(3) * q - T(2));
    d.w2 = k.wg_a * (T(6) * q - T(2));
  } else if (q < T(1)) {
    d.g = -k.wg_b * (u * u) * rinv;
    d.w2 = T(2) * k.wg_a * u;
  } else {
    d.g = T(0);
    d.w2 = T(0);
  }
  d.hq = (d.w2 - d.g) * (rinv * rinv);
  return d;
}

// Akinci cohesion C(r) and C'(r) (sph_util.py:4-13).
template <class T>
__device__ __forceinline__ void coh_d1(const KP<T>& k, T r, T* C, T* Cp) {
  T q = r * k.inv_h;
  T u = T(1) - q;
  T c = T(0), dc = T(0);
  if (q <= T(0.5)) {
    c = T(2) * (u * u * u) * (q * q * q) - T(1.0 / 64);
    dc = T(6) * (u * u) * (q * q) * (T(1) - T(2) * q);
  } else if (q < T(1)) {
    c = (u * u * u) * (q * q * q);
    dc = T(3) * (u * u) * (q * q) * (T(1) - T(2) * q);
  }
  *C = c * k.c_coh;
  *Cp = dc * k.c_coh * k.inv_h;
}

template <class T>
__device__ __forceinline__ V4<T> v4(T x, T y, T z, T w = T(0)) {
  return V4<T>{x, y, z, w};
}
template <class T>
__device__ __forceinline__ T dot3(const V4<T>& a, const V4<T>& b) {
  return a.x * b.x + a.y * b.y + a.z * b.z;
}
// acc += J(x) c
template <class T>
__device__ __forceinline__ void add_J(T& ax, T& ay, T& az, const D2<T>& d, T dx, T dy, T dz,
                                      T cx, T cy, T cz) {
  T s = d.hq * (dx * cx + dy * cy + dz * cz);
  ax += d.g * cx + s * dx;
  ay += d.g * cy + s * dy;
  az += d.g * cz + s * dz;
}

// ---------------------------------------------------------------------------
// Boundary sums.  Boundary particles carry no velocity, kappa or adjoint, so
// in the reverse of the Jacobi iterations and of pass A every (f,b) pair of
// particle l enters as  X_b G(x_lb) . a  and  J(x_lb) (X_b c)  with a, c
// depending on l alone.  Summed over b these are  a . A_l  and  M_l c  with
//   A_l = sum_b X_b g x,   M_l = sum_b X_b J(x) = G_l I + H_l,
//   G_l = sum_b X_b g,     H_l = sum_b X_b hq x x^T,
// so the reverse sweeps read three records instead of walking the boundary.
// ---------------------------------------------------------------------------
constexpr int kBsumU = 8;  // boundary candidates in flight per lane
// Lanes per particle of the reverse Jacobi and viscosity sweeps.  Their pair
// terms (kernel second derivatives, Hessian products) make them VALU-bound
// about as much as gather-bound, so two lanes per particle, which leave fewer
// lanes idle at the end of uneven lists, beat the forward sweeps' four.
constexpr int kCoopR = 2;
// Gathers in flight per lane in the reverse viscosity sweep.  Its four-vector
// records take 152 VGPRs at four in flight (3 waves per SIMD); two keep it at
// 110 (4 waves) and run faster.  The Jacobi sweeps keep four: at two they
// reach 5 waves but lose more to latency than they gain.
constexpr int kNbVis = 2;

template <class T>
__global__ void __launch_bounds__(kBlock)
    k_adj_bound_sums(KP<T> k, int n, int cap, const V4<T>* __restrict__ fpos,
                     const int* __restrict__ nbr_nb, const V4<T>* __restrict__ bpos,
                     const int* __restrict__ bstart, V4<T>* __restrict__ bmom,
                     int* __restrict__ bflag) {
  const int i = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  if (nbr_nb[i] == 0) {  // the step found no boundary neighbour: nothing to walk
    bflag[i] = 0;
    return;
  }
  const V4<T> p = fpos[i];
  int c[3];
  node_encode(k, p.x, p.y, p.z, c);
  T ax = 0, ay = 0, az = 0, G = 0, hxx = 0, hyy = 0, hzz = 0, hxy = 0, hxz = 0, hyz = 0;
  int any = 0;
  walk_cells_batched<kBsumU>(k, bstart, c, bpos, [&](int j, const V4<T>& q) {
    const T dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
    const T r2 = s_r2(dx, dy, dz);
    if (!(r2 <= k.r2max && r2 > T(0))) return;  // the boundary lists' test
    const D2<T> e = kern_d2(k, r2);
    const T xg = q.w * e.g, xh = q.w * e.hq;
    ax += xg * dx;
    ay += xg * dy;
    az += xg * dz;
    G += xg;
    hxx += xh * dx * dx;
    hyy += xh * dy * dy;
    hzz += xh * dz * dz;
    hxy += xh * dx * dy;
    hxz += xh * dx * dz;
    hyz += xh * dy * dz;
    any = 1;
  });
  bflag[i] = any;
  if (any) {
    bmom[i] = v4(ax, ay, az, G);
    bmom[(long)cap + i] = v4(hxx, hyy, hzz, hxy);
    bmom[2L * cap + i] = v4(hxz, hyz, T(0), T(0));
  }
}

// The boundary sums of particle i: A (xyz of m0) and M c = G c + H c.
template <class T>
struct BoundSums {
  V4<T> m0, m1, m2;
  __device__ __forceinline__ T dot_A(T x, T y, T z) const { return m0.x * x + m0.y * y + m0.z * z; }
  __device__ __forceinline__ void add_M(T& ox, T& oy, T& oz, T cx, T cy, T cz) const {
    ox += m0.w * cx + (m1.x * cx + m1.w * cy + m2.x * cz);
    oy += m0.w * cy + (m1.w * cx + m1.y * cy + m2.y * cz);
    oz += m0.w * cz + (m2.x * cx + m2.y * cy + m1.z * cz);
  }
};
template <class T>
__device__ __forceinline__ BoundSums<T> load_bound_sums(const V4<T>* __restrict__ bmom, int cap,
                                                        int i) {
  return BoundSums<T>{bmom[i], bmom[(long)cap + i], bmom[2L * cap + i]};
}

// ---------------------------------------------------------------------------
// Tail reverse: pos' = pos + dt vel_adv, vel' = vel_adv (sph.py:259-269).
// ---------------------------------------------------------------------------
template <class T>
__global__ void __launch_bounds__(kBlock)
    k_adj_begin(int n, const int* __restrict__ pid, const V4<T>* __restrict__ xref,
                const V4<T>* __restrict__ vref, const T* __restrict__ kappa_next,
                V4<T>* __restrict__ xbar, V4<T>* __restrict__ vbar, T* __restrict__ rhobar,
                T* __restrict__ alphabar, const TsphControl* ctrl) {
  int i = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  const T dt = (T)ctrl->dt;
  int p = pid[i];
  V4<T> xb = xref[p], vb = vref[p];
  xbar[i] = v4(xb.x, xb.y, xb.z);
  vbar[i] = v4(vb.x + dt * xb.x, vb.y + dt * xb.y, vb.z + dt * xb.z,
               kappa_next ? kappa_next[i] : T(0));
  rhobar[i] = T(0);
  alphabar[i] = T(0);
}

template <class T>
__global__ void __launch_bounds__(kBlock)
    k_adj_end(int n, const int* __restrict__ pid, const V4<T>* __restrict__ xbar,
              const V4<T>* __restrict__ vbar, V4<T>* __restrict__ xref, V4<T>* __restrict__ vref) {
  int i = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  int p = pid[i];
  V4<T> xb = xbar[i], vb = vbar[i];
  xref[p] = v4(xb.x, xb.y, xb.z);
  vref[p] = v4(vb.x, vb.y, vb.z);
}

// ---------------------------------------------------------------------------
// Jacobi iteration reverse, half 1: IPPE_update_vel_adv (sph.py:229-244)
//   v'_l = v_l - (1/dt)/m_l sum_j (kappa_l X_j + kappa_j X_l) G(x_lj)
// gives, with u = d loss / d v' and w = u / m,
//   kappa_bar_l = -(1/dt) sum_j X_j G(x_lj) . (w_l - w_j)
//   x_bar_l    += -(1/dt) sum_j J (kappa_l X_j + kappa_j X_l) (w_l - w_j)
// (boundary neighbours: w_j = 0, kappa_j = 0).  Then kappa = psi / alpha and
// the clamp psi = max(psi_pre, 0) (sph.py:219-226) give psi_bar, alpha_bar
// and rho_bar (psi_pre = rho - rho_0 + ..., sph.py:196-199).
// DFSPH: m = X = pos.w.  The iteration's kappa rides in vbar.w (packed by
// the kernel that wrote vbar), so a neighbour costs two 16-byte gathers;
// psi_bar is written into the w lane of the iteration's vel_adv record,
// where the second half gathers it.
template <class T>
__global__ void __launch_bounds__(kBlock)
    k_adj_kappa(KP<T> k, int n, int cap, const V4<T>* __restrict__ fpos,
                const T* __restrict__ alpha, const V4<T>* __restrict__ vbar,
                const int* __restrict__ fstart, const int* __restrict__ nbr_f,
                const int* __restrict__ nbr_nf, const V4<T>* __restrict__ bmom,
                const int* __restrict__ bflag, V4<T>* __restrict__ tvel, T* __restrict__ rhobar,
                T* __restrict__ alphabar, V4<T>* __restrict__ xbar, const TsphControl* ctrl) {
  const int base = xcd_block(blockIdx.x, gridDim.x) * kBlock;
  const int cl = coop_lane<kCoopR>();
  const T dt = (T)ctrl->dt;
#pragma unroll 1
  for (int round = 0; round < kCoopR; ++round) {
  const int i = coop_particle<kCoopR>(base, round);
  if (i >= n) break;
  const V4<T> p = fpos[i];
  const T Xi = p.w;
  const V4<T> u = vbar[i];
  const T rmi = T(1) / Xi;
  const T wx = u.x * rmi, wy = u.y * rmi, wz = u.z * rmi;
  const T ki = u.w;
  // the particle's own fields, requested before the walk so that their
  // latency hides behind it (only this particle's group writes them)
  const T al = alpha[i];
  const T rb0 = rhobar[i], ab0 = alphabar[i];
  const V4<T> xb = xbar[i];
  const bool hb = bflag[i] != 0;
  BoundSums<T> bs{};
  if (hb) bs = load_bound_sums(bmom, cap, i);
  T kb = 0, ax = 0, ay = 0, az = 0;
  auto term = [&](const Rec2<T>& d) {  // a fluid neighbour j
    const V4<T>& q = d.a;   // q.w = X_j = m_j
    const V4<T>& uj = d.b;  // xyz = u_j, w = kappa_j
    T dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
    D2<T> e = kern_d2(k, dx * dx + dy * dy + dz * dz);
    T rmj = f_rcp(q.w);  // hardware reciprocal in FP32 (the reverse is not a bit-exact sum)
    T mx = wx - uj.x * rmj, my = wy - uj.y * rmj, mz = wz - uj.z * rmj;
    kb += q.w * e.g * (dx * mx + dy * my + dz * mz);
    T c = ki * q.w + uj.w * Xi;
    add_J(ax, ay, az, e, dx, dy, dz, c * mx, c * my, c * mz);
  };
  // (f,f) pairs; the (f,b) pairs (u_b = kappa_b = 0) are w_l . A_l and
  // kappa_l M_l w_l from the boundary sums
  sweep_lists_coop<kCoopR, T, Rec2<T>>(
      k, cap, i, cl, p, fstart, nullptr, fpos, nullptr, nbr_f, nullptr, nbr_nf[i], 0,
      [&](int j) { return Rec2<T>{fpos[j], vbar[j]}; }, [&](int) { return Rec2<T>{}; },
      term, [&](const Rec2<T>&) {});
  kb = group_sum<kCoopR>(kb);
  ax = group_sum<kCoopR>(ax);
  ay = group_sum<kCoopR>(ay);
  az = group_sum<kCoopR>(az);
  if (cl != 0) continue;
  if (hb) {
    kb += bs.dot_A(wx, wy, wz);
    bs.add_M(ax, ay, az, ki * wx, ki * wy, ki * wz);
  }
  const T s = -(T(1) / dt);
  kb *= s;
  const T pb = ki > T(0) ? kb / al : T(0);
  tvel[i].w = pb;
  rhobar[i] = rb0 + pb;
  alphabar[i] = ab0 + -kb * ki / al;
  xbar[i] = v4(xb.x + s * ax, xb.y + s * ay, xb.z + s * az);
  }
}

// Jacobi iteration reverse, half 2: IPPE_adv_psi (sph.py:202-216)
//   psi_l = ... + dt sum_j X_j G(x_lj) . (v_l - v_j)
// gives
//   v_bar_l = u_l + dt sum_j (psib_l X_j + psib_j X_l) G(x_lj)
//   x_bar_l += dt sum_j J (psib_l X_j + psib_j X_l) (v_l - v_j)
// with v the recorded vel_adv entering the iteration (boundary: v_j = 0,
// psib_j = 0).
template <class T>
__global__ void __launch_bounds__(kBlock)
    k_adj_vel(KP<T> k, int n, int cap, const V4<T>* __restrict__ fpos,
              const V4<T>* __restrict__ tvel, const V4<T>* __restrict__ vbar,
              const T* __restrict__ kappa_next, const int* __restrict__ fstart,
              const int* __restrict__ nbr_f, const int* __restrict__ nbr_nf,
              const V4<T>* __restrict__ bmom, const int* __restrict__ bflag,
              V4<T>* __restrict__ vbar_out, V4<T>* __restrict__ xbar, const TsphControl* ctrl) {
  const int base = xcd_block(blockIdx.x, gridDim.x) * kBlock;
  const int cl = coop_lane<kCoopR>();
  const T dt = (T)ctrl->dt;
#pragma unroll 1
  for (int round = 0; round < kCoopR; ++round) {
  const int i = coop_particle<kCoopR>(base, round);
  if (i >= n) break;
  const V4<T> p = fpos[i];
  const T Xi = p.w;
  const V4<T> vi = tvel[i];  // w = psi_bar of this particle
  const T pbi = vi.w;
  // the particle's own fields, requested before the walk (only this
  // particle's group writes xbar[i])
  const V4<T> u = vbar[i];
  const T kn = kappa_next ? kappa_next[i] : T(0);
  const V4<T> xb = xbar[i];
  const bool hb = bflag[i] != 0;
  BoundSums<T> bs{};
  if (hb) bs = load_bound_sums(bmom, cap, i);
  T vx = 0, vy = 0, vz = 0, ax = 0, ay = 0, az = 0;
  auto term = [&](const Rec2<T>& d) {
    const V4<T>& q = d.a;   // q.w = X_j
    const V4<T>& vj = d.b;  // xyz = vel_adv_j, w = psib_j
    T dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
    D2<T> e = kern_d2(k, dx * dx + dy * dy + dz * dz);
    T c = pbi * q.w + vj.w * Xi;
    T cg = c * e.g;
    vx += cg * dx;
    vy += cg * dy;
    vz += cg * dz;
    add_J(ax, ay, az, e, dx, dy, dz, c * (vi.x - vj.x), c * (vi.y - vj.y), c * (vi.z - vj.z));
  };
  // (f,f) pairs; the (f,b) pairs (v_b = psib_b = 0) are psib_l A_l and
  // psib_l M_l v_l from the boundary sums
  sweep_lists_coop<kCoopR, T, Rec2<T>>(
      k, cap, i, cl, p, fstart, nullptr, fpos, nullptr, nbr_f, nullptr, nbr_nf[i], 0,
      [&](int j) { return Rec2<T>{fpos[j], tvel[j]}; }, [&](int) { return Rec2<T>{}; }, term,
      [&](const Rec2<T>&) {});
  vx = group_sum<kCoopR>(vx);
  vy = group_sum<kCoopR>(vy);
  vz = group_sum<kCoopR>(vz);
  ax = group_sum<kCoopR>(ax);
  ay = group_sum<kCoopR>(ay);
  az = group_sum<kCoopR>(az);
  if (cl != 0) continue;
  if (hb) {
    vx += pbi * bs.m0.x;
    vy += pbi * bs.m0.y;
    vz += pbi * bs.m0.z;
    bs.add_M(ax, ay, az, pbi * vi.x, pbi * vi.y, pbi * vi.z);
  }
  vbar_out[i] = v4(u.x + dt * vx, u.y + dt * vy, u.z + dt * vz, kn);
  xbar[i] = v4(xb.x + dt * ax, xb.y + dt * ay, xb.z + dt * az);
  }
}

// ---------------------------------------------------------------------------
// Advection reverse, part 1: surface tension loop 2 (sph.py:151-166)
//   a_l += sum_j f_lj (-gamma m_j C(r) x/r + gamma (n_l - n_j)),
//   f_lj = 2 rho0_l / (rho_l + rho_j),
// over every candidate of the neighbouring cells with r > 0 (the curvature
// term carries no kernel weight) and both volume fractions > 0.99.
// With a_bar = dt u (vel_adv = vel + dt a, sph.py:247-250) and
// e_lj = f_lj a_l - f_jl a_j:
//   n_bar_l  = gamma sum_j e_lj
//   x_bar_l += -gamma sum_j Jc (f_lj m_j a_l - f_jl m_l a_j),
//              Jc w = (C/r) w + ((C' - C/r)/r^2)(x.w) x
//   rho_bar_l += -sum_j (f_lj s_lj + f_jl s_jl) / (rho_l + rho_j)
//   d loss / d gamma += a_l . sum_j f_lj (-m_j C x/r + n_l - n_j)
// The terms without x (the curvature's) are summed over the cells, gathering
// no position: r > 0 is j != l except for a particle at l's own position,
// which lies in l's own cell, so only that cell's candidates are tested.
// The cohesion terms (C = C' = 0 beyond h) are summed over the in-range list.
// A neighbour's rest density and velocity adjoint enter only as
// b_j = rho0_j a_j (f_jl a_j = 2 b_j / (rho_l + rho_j)), which k_adj_st_rec
// writes per particle, so that a candidate costs two 16-byte gathers.
// ---------------------------------------------------------------------------
template <class T>
__global__ void __launch_bounds__(kBlock)
    k_adj_st_rec(int n, const T* __restrict__ rdens, const V4<T>* __restrict__ vbar,
                 V4<T>* __restrict__ brec, const TsphControl* ctrl) {
  const int i = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  const T s = rdens[i] * (T)ctrl->dt;
  const V4<T> u = vbar[i];
  brec[i] = V4<T>{s * u.x, s * u.y, s * u.z, T(0)};
}

template <class T>
__global__ void __launch_bounds__(kBlock)
    k_adj_surface_tension(KP<T> k, int n, int cap, const V4<T>* __restrict__ fpos,
                          const T* __restrict__ fmass, const T* __restrict__ dens,
                          const T* __restrict__ rdens, const T* __restrict__ vf,
                          const int* __restrict__ fstart, const int* __restrict__ nbr_f,
                          const int* __restrict__ nbr_nf, const V4<T>* __restrict__ tnormal,
                          const V4<T>* __restrict__ vbar, const V4<T>* __restrict__ brec,
                          V4<T>* __restrict__ nbar, T* __restrict__ rhobar,
                          V4<T>* __restrict__ xbar, double* __restrict__ partial,
                          const TsphControl* ctrl) {
  __shared__ double lds[kWaves];
  const int i = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  double gsum = 0;
  if (i < n) {
    const T dt = (T)ctrl->dt;
    const T gam = k.gamma;
    T nbx = 0, nby = 0, nbz = 0, ax = 0, ay = 0, az = 0, rb = 0, gs = 0;
    if (vf[(long)i * k.P] > T(0.99)) {
      const V4<T> p = fpos[i];
      const V4<T> ni = tnormal[i];
      const V4<T> ui = vbar[i];
      const T ali_x = dt * ui.x, ali_y = dt * ui.y, ali_z = dt * ui.z;
      const T rhoi = dens[i], r0i = rdens[i], mi = fmass[i];
      int c[3];
      node_encode(k, p.x, p.y, p.z, c);
      // the particle's own cell; a code outside (0, node_num) is the sort's
      // trailing bucket node_num, whose members no cell walk visits
      const int code = dim_encode(k, c[0], c[1], c[2]);
      const long own = (0 < code && code < k.node_num) ? (long)code : (long)k.node_num;
      const int own0 = fstart[own], own1 = fstart[own + 1];
      // curvature, per candidate: the recorded normal with +-rho_j in its
      // spare lane (negative where j's 0.99 gate fails; the forward advection
      // writes it) and b_j
      auto loadn = [&](int j) { return Rec2<T>{tnormal[j], brec[j]}; };
      walk_cells_prefetch<kStU>(k, fstart, c, loadn, [&](int j, const Rec2<T>& d) {
        if (j == i || !(d.a.w > T(0))) return;
        if (j >= own0 && j < own1) {
          const V4<T> q = fpos[j];
          if (!(s_r2(p.x - q.x, p.y - q.y, p.z - q.z) > T(0))) return;
        }
        const V4<T>& nj = d.a;
        const V4<T>& bj = d.b;
        const T rden = f_rcp(rhoi + nj.w);
        const T fij = T(2) * r0i * rden, r2d = T(2) * rden;
        const T dnx = ni.x - nj.x, dny = ni.y - nj.y, dnz = ni.z - nj.z;
        // e_lj = f_lj a_l - f_jl a_j
        const T ex = fij * ali_x - r2d * bj.x;
        const T ey = fij * ali_y - r2d * bj.y;
        const T ez = fij * ali_z - r2d * bj.z;
        gs += fij * (ali_x * dnx + ali_y * dny + ali_z * dnz);
        nbx += gam * ex;
        nby += gam * ey;
        nbz += gam * ez;
        rb += -gam * rden * (dnx * ex + dny * ey + dnz * ez);
      });
      // cohesion, per in-range neighbour: (x_j, X_j = m_j) as well
      auto loadc = [&](int j) { return Rec3<T>{fpos[j], tnormal[j], brec[j]}; };
      auto coh = [&](const Rec3<T>& d) {
        if (!(d.b.w > T(0))) return;
        const V4<T>& q = d.a;
        const V4<T>& bj = d.c;
        const T dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
        const T r2 = dx * dx + dy * dy + dz * dz;
        const T rinv = f_rsqrt(r2);
        T C, Cp;
        coh_d1(k, r2 * rinv, &C, &Cp);
        const T cr = C * rinv;
        const T cq = (Cp - cr) * (rinv * rinv);
        const T mj = q.w;  // DFSPH: the alias X is the mass
        const T rden = f_rcp(rhoi + d.b.w);
        const T fij = T(2) * r0i * rden, r2d = T(2) * rden;
        const T ad = ali_x * dx + ali_y * dy + ali_z * dz;
        const T bd = bj.x * dx + bj.y * dy + bj.z * dz;
        gs += fij * (-mj * cr) * ad;
        // f_lj m_j a_l - f_jl m_l a_j
        const T wx = fij * mj * ali_x - r2d * mi * bj.x;
        const T wy = fij * mj * ali_y - r2d * mi * bj.y;
        const T wz = fij * mj * ali_z - r2d * mi * bj.z;
        const T s = cq * (dx * wx + dy * wy + dz * wz);
        ax += -gam * (cr * wx + s * dx);
        ay += -gam * (cr * wy + s * dy);
        az += -gam * (cr * wz + s * dz);
        rb += -gam * rden * (r2d * mi * cr * bd - fij * mj * cr * ad);
      };
      const int nf = nbr_nf[i];
      if (nf <= k.list_cap_f) {
        sweep_one_list_coop<1, Rec3<T>, 2>(cap, i, 0, nbr_f, nf, loadc, coh);
      } else {
        walk_cells(k, fstart, c, [&](int j) {
          const V4<T> q = fpos[j];
          const T r2 = s_r2(p.x - q.x, p.y - q.y, p.z - q.z);
          if (r2 > T(0) && r2 <= k.r2max) coh(loadc(j));
        });
      }
    }
    nbar[i] = v4(nbx, nby, nbz);
    rhobar[i] += rb;
    V4<T> xb = xbar[i];
    xbar[i] = v4(xb.x + ax, xb.y + ay, xb.z + az);
    gsum = (double)gs;
  }
  double s = block_sum(gsum, lds);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// Advection reverse, part 2: viscosity (sph.py:114-129, W_lap sph_util.py:40-43)
//   a_l += fv_l sum_j V_j w(r) (v_lj . x) x,  w = g / (0.01 h^2 + r^2),
//   fv = 2 (2 + d) mu / rho0, V = X / rho,
// and the normals (sph.py:141-149)  n_l = -h sum_j V_j G(x).  With a = dt u:
//   v_bar_l  = u_l + sum_j w (e . x) x,           e = fv_l V_j a_l - fv_j V_l a_j
//   x_bar_l += sum_j [w'/r (v.x)(e.x) x + w ((e.x) v + (v.x) e)]
//              - h sum_j J (V_j nb_l - V_l nb_j)
//   V_bar_l  = sum_j [-fv_j w (v.x)(a_j.x) + h g (nb_j . x)]
//   rho_bar_l += -V_bar_l V_l / rho_l
//   d loss / d mu += a_l . sum_j (fv_l / mu) V_j w (v.x) x
template <class T>
struct Rec4 {
  V4<T> a, b, c, d;
};

// Per fluid particle: a_bar = dt u and fv = 2 (2 + d) mu / rho0, the record
// the viscosity reverse gathers per neighbour (one gather instead of u and
// rho0 apart).
template <class T>
__global__ void __launch_bounds__(kBlock)
    k_adj_vis_rec(KP<T> k, int n, const T* __restrict__ rdens, const V4<T>* __restrict__ vbar,
                  V4<T>* __restrict__ arec, const TsphControl* ctrl) {
  const int i = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  const T dt = (T)ctrl->dt;
  const V4<T> u = vbar[i];
  arec[i] = v4(dt * u.x, dt * u.y, dt * u.z, k.two_d2 * k.mu * f_rcp(rdens[i]));
}

template <class T>
__global__ void __launch_bounds__(kBlock)
    k_adj_viscosity(KP<T> k, int n, int cap, const V4<T>* __restrict__ fpos,
                    const T* __restrict__ dens,
                    const T* __restrict__ rdens, const int* __restrict__ fstart,
                    const int* __restrict__ bstart, const V4<T>* __restrict__ bpos,
                    const int* __restrict__ nbr_f, const int* __restrict__ nbr_nf,
                    const V4<T>* __restrict__ tvel, const V4<T>* __restrict__ nbar,
                    const V4<T>* __restrict__ vbar, const V4<T>* __restrict__ arec,
                    const T* __restrict__ kappa_next,
                    V4<T>* __restrict__ vbar_out, T* __restrict__ rhobar,
                    V4<T>* __restrict__ xbar, double* __restrict__ partial,
                    const TsphControl* ctrl) {
  __shared__ double lds[kWaves];
  const int base = xcd_block(blockIdx.x, gridDim.x) * kBlock;
  const int cl = coop_lane<kCoopR>();
  double msum = 0;
#pragma unroll 1
  for (int round = 0; round < kCoopR; ++round) {
    const int i = coop_particle<kCoopR>(base, round);
    if (i >= n) break;
    const T dt = (T)ctrl->dt;
    const V4<T> p = fpos[i];
    const V4<T> vi = tvel[i];
    const V4<T> ui = vbar[i];
    const V4<T> nbi = nbar ? nbar[i] : V4<T>{};
    const T ax_ = dt * ui.x, ay_ = dt * ui.y, az_ = dt * ui.z;
    const T Vi = vi.w;  // the recorded velocity's spare lane holds V_i = X_i / psi_i
    const T fmu_i = k.two_d2 / rdens[i];  // fv_i / mu
    const T fvi = fmu_i * k.mu;
    // the particle's own outputs' inputs, requested before the walk (only
    // this particle's group writes rhobar[i] and xbar[i])
    const T kn = kappa_next ? kappa_next[i] : T(0);
    const T rb0 = rhobar[i], di = dens[i];
    const V4<T> xb = xbar[i];
    T vbx = 0, vby = 0, vbz = 0, xbx = 0, xby = 0, xbz = 0, Vb = 0, ms = 0;
    auto term = [&](const Rec4<T>& d) {
      const V4<T>& q = d.a;   // position
      const V4<T>& vj = d.b;  // velocity, w = V_j
      const V4<T>& aj = d.c;  // a_bar_j, w = fv_j
      const V4<T>& nj = d.d;  // n_bar_j
      T dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
      T r2 = dx * dx + dy * dy + dz * dz;
      D2<T> e = kern_d2(k, r2);
      T den = k.eps + r2;
      T rden = f_rcp(den);
      T w = e.g * rden;
      T wq = (e.hq - T(2) * w) * rden;  // w'(r) / r
      T Vj = vj.w, fvj = aj.w;
      T vdx = vi.x - vj.x, vdy = vi.y - vj.y, vdz = vi.z - vj.z;
      T vx = vdx * dx + vdy * dy + vdz * dz;
      T alx = ax_ * dx + ay_ * dy + az_ * dz;
      ms += fmu_i * Vj * w * vx * alx;
      T ex_ = fvi * Vj * ax_ - fvj * Vi * aj.x;
      T ey_ = fvi * Vj * ay_ - fvj * Vi * aj.y;
      T ez_ = fvi * Vj * az_ - fvj * Vi * aj.z;
      T ex = ex_ * dx + ey_ * dy + ez_ * dz;
      vbx += w * ex * dx;
      vby += w * ex * dy;
      vbz += w * ex * dz;
      T t = wq * vx * ex;
      xbx += t * dx + w * (ex * vdx + vx * ex_);
      xby += t * dy + w * (ex * vdy + vx * ey_);
      xbz += t * dz + w * (ex * vdz + vx * ez_);
      T ajx = aj.x * dx + aj.y * dy + aj.z * dz;
      Vb += -fvj * w * vx * ajx + k.h * e.g * (nj.x * dx + nj.y * dy + nj.z * dz);
      T cx = Vj * nbi.x - Vi * nj.x, cy = Vj * nbi.y - Vi * nj.y, cz = Vj * nbi.z - Vi * nj.z;
      T s = e.hq * (dx * cx + dy * cy + dz * cz);
      xbx += -k.h * (e.g * cx + s * dx);
      xby += -k.h * (e.g * cy + s * dy);
      xbz += -k.h * (e.g * cz + s * dz);
    };
    sweep_lists_coop<kCoopR, T, Rec4<T>, kNbVis>(
        k, cap, i, cl, p, fstart, bstart, fpos, bpos, nbr_f, nullptr, nbr_nf[i], 0,
        [&](int j) {
          // tvel.w = V_j; arec = (dt u_j, fv_j)
          return Rec4<T>{fpos[j], tvel[j], arec[j], nbar ? nbar[j] : V4<T>{}};
        },
        [&](int j) { return Rec4<T>{}; }, term, [&](const Rec4<T>&) {});
    vbx = group_sum<kCoopR>(vbx);
    vby = group_sum<kCoopR>(vby);
    vbz = group_sum<kCoopR>(vbz);
    xbx = group_sum<kCoopR>(xbx);
    xby = group_sum<kCoopR>(xby);
    xbz = group_sum<kCoopR>(xbz);
    Vb = group_sum<kCoopR>(Vb);
    ms = group_sum<kCoopR>(ms);
    if (cl != 0) continue;
    vbar_out[i] = v4(ui.x + vbx, ui.y + vby, ui.z + vbz, kn);
    // V_i = X_i / psi_i enters only pair terms; a particle outside the grid
    // (psi_i = 0, V_i infinite) has none, so there is nothing to carry back
    rhobar[i] = di > T(0) ? rb0 + -Vb * Vi / di : rb0;
    xbar[i] = v4(xb.x + xbx, xb.y + xby, xb.z + xbz);
    msum += (double)ms;
  }
  double s = block_sum(msum, lds);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// Adds the block partials of a parameter gradient into grad[slot], in a
// fixed order.
__global__ void __launch_bounds__(kBlock)
    k_adj_grad_add(const double* __restrict__ partial, int nblk, double* grad, int slot) {
  __shared__ double lds[kWaves];
  double a = sum_partials(partial, nblk, lds);
  if (threadIdx.x == 0) grad[slot] += a;
}

// ---------------------------------------------------------------------------
// Pass A reverse.  alpha = max(|alpha_1|^2 / m + alpha_2, 1e-4) (sph.py:100-105):
// the clamp passes no gradient; a1bar = (2 alpha_bar alpha_1 / m, alpha_bar).
// ---------------------------------------------------------------------------
template <class T>
__global__ void __launch_bounds__(kBlock)
    k_adj_alpha_prep(int n, const T* __restrict__ alpha, const V4<T>* __restrict__ alpha_1,
                     const T* __restrict__ fmass, const T* __restrict__ alphabar,
                     V4<T>* __restrict__ a1bar) {
  int i = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  T ab = alpha[i] > T(1e-4) ? alphabar[i] : T(0);
  V4<T> a1 = alpha_1[i];
  T s = T(2) * ab / fmass[i];
  a1bar[i] = v4(s * a1.x, s * a1.y, s * a1.z, ab);
}

// rho_l = sum_j m_j W(r)            -> x_bar_l += (rb_l m_j + rb_j m_l) g x
// alpha_1_l = sum_j X_j G(x)         -> x_bar_l += J (X_j a1b_l - X_l a1b_j)
// alpha_2_l = sum_j W'^2 X_j^2 / m_j -> x_bar_l += 2 g W'' (a2b_l X_j^2/m_j + a2b_j X_l^2/m_l) x
// (boundary neighbours: density and alpha_1 terms of l's equation only;
// alpha_2 of the fluid is (f,f) only, sph.py:194).
template <class T>
__global__ void __launch_bounds__(kBlock)
    k_adj_pass_a(KP<T> k, int n, int cap, const V4<T>* __restrict__ fpos,
                 const T* __restrict__ fmass, const T* __restrict__ rhobar,
                 const V4<T>* __restrict__ a1bar, const int* __restrict__ fstart,
                 const int* __restrict__ nbr_f, const int* __restrict__ nbr_nf,
                 const V4<T>* __restrict__ bmom, const int* __restrict__ bflag,
                 V4<T>* __restrict__ xbar) {
  const int base = xcd_block(blockIdx.x, gridDim.x) * kBlock;
  const int cl = coop_lane<kCoop>();
#pragma unroll 1
  for (int round = 0; round < kCoop; ++round) {
  const int i = coop_particle<kCoop>(base, round);
  if (i >= n) break;
  const V4<T> p = fpos[i];
  const T Xi = p.w, mi = fmass[i];
  const T rbi = rhobar[i];
  const V4<T> abi = a1bar[i];
  const T xm_i = Xi * Xi / mi;
  const V4<T> xb = xbar[i];  // requested before the walk; only this group writes it
  const bool hb = bflag[i] != 0;
  BoundSums<T> bs{};
  if (hb) bs = load_bound_sums(bmom, cap, i);
  T ax = 0, ay = 0, az = 0;
  // (f,f) pairs.  Record: a = position (w = X_j), b = a1bar_j (w = a2bar_j),
  // c = (rho_bar_j, m_j).  The (f,b) pairs (rho_bar_b = 0, no alpha_2 term,
  // DFSPH: m_b = X_b) are rho_bar_l A_l + M_l a1bar_l from the boundary sums.
  sweep_lists_coop<kCoop, T, Rec4<T>>(
      k, cap, i, cl, p, fstart, nullptr, fpos, nullptr, nbr_f, nullptr, nbr_nf[i], 0,
      [&](int j) {
        V4<T> q = fpos[j];
        V4<T> extra = v4(rhobar[j], q.w, T(0), T(0));  // DFSPH: m_j = X_j
        return Rec4<T>{q, a1bar[j], extra, V4<T>{}};
      },
      [&](int) { return Rec4<T>{}; },
      [&](const Rec4<T>& d) {
        const V4<T>& q = d.a;
        const V4<T>& abj = d.b;
        const T rbj = d.c.x, mj = d.c.y;
        T dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
        D2<T> e = kern_d2(k, dx * dx + dy * dy + dz * dz);
        T cd = (rbi * mj + rbj * mi) * e.g;
        T c2 = T(2) * e.g * e.w2 * (abi.w * q.w * q.w * f_rcp(mj) + abj.w * xm_i);
        T s1 = cd + c2;
        ax += s1 * dx;
        ay += s1 * dy;
        az += s1 * dz;
        add_J(ax, ay, az, e, dx, dy, dz, q.w * abi.x - Xi * abj.x, q.w * abi.y - Xi * abj.y,
              q.w * abi.z - Xi * abj.z);
      },
      [&](const Rec4<T>&) {});
  ax = group_sum<kCoop>(ax);
  ay = group_sum<kCoop>(ay);
  az = group_sum<kCoop>(az);
  if (cl != 0) continue;
  if (hb) {
    ax += rbi * bs.m0.x;
    ay += rbi * bs.m0.y;
    az += rbi * bs.m0.z;
    bs.add_M(ax, ay, az, abi.x, abi.y, abi.z);
  }
  xbar[i] = v4(xb.x + ax, xb.y + ay, xb.z + az);
  }
}

// ---------------------------------------------------------------------------
// Host launchers.
// ---------------------------------------------------------------------------
template <class T>
int adj_begin(const TsphParams& prm, TsphObject* f, TsphControl* ctrl, TsphAdjoint* a,
              const int* pid, const void* xref, const void* vref, hipStream_t st) {
  (void)prm;
  if (f->n == 0) return 0;
  hipLaunchKernelGGL(k_adj_begin<T>, dim3(blocks_for(f->n)), dim3(kBlock), 0, st, f->n, pid,
                     (const V4<T>*)xref, (const V4<T>*)vref, (const T*)a->kappa_next,
                     (V4<T>*)a->xbar, (V4<T>*)a->vbar, (T*)a->rhobar, (T*)a->alphabar, ctrl);
  TSPH_LAUNCH_CHECK("adj_begin");
  return 0;
}

template <class T>
int adj_end(const TsphParams& prm, TsphObject* f, TsphAdjoint* a, const int* pid, void* xref,
            void* vref, hipStream_t st) {
  (void)prm;
  if (f->n == 0) return 0;
  hipLaunchKernelGGL(k_adj_end<T>, dim3(blocks_for(f->n)), dim3(kBlock), 0, st, f->n, pid,
                     (const V4<T>*)a->xbar, (const V4<T>*)a->vbar, (V4<T>*)xref, (V4<T>*)vref);
  TSPH_LAUNCH_CHECK("adj_end");
  return 0;
}

template <class T>
int adj_solver_iteration(const TsphParams& prm, TsphObject* f, TsphObject* b, TsphControl* ctrl,
                         TsphAdjoint* a, void* ws, size_t wsb, hipStream_t st) {
  (void)ws;
  (void)wsb;
  KP<T> k = make_kp<T>(prm);
  if (f->n == 0) return 0;
  int nb = blocks_for(f->n);
  (void)b;  // the boundary enters through the boundary sums
  hipLaunchKernelGGL(k_adj_kappa<T>, dim3(nb), dim3(kBlock), 0, st, k, f->n, f->capacity,
                     (const V4<T>*)f->pos, (const T*)f->alpha, (const V4<T>*)a->vbar,
                     f->cell_start, f->nbr_f, f->nbr_nf, (const V4<T>*)a->bmom, a->bflag,
                     (V4<T>*)a->tape_vel_adv, (T*)a->rhobar, (T*)a->alphabar, (V4<T>*)a->xbar,
                     ctrl);
  hipLaunchKernelGGL(k_adj_vel<T>, dim3(nb), dim3(kBlock), 0, st, k, f->n, f->capacity,
                     (const V4<T>*)f->pos, (const V4<T>*)a->tape_vel_adv, (const V4<T>*)a->vbar,
                     (const T*)a->kappa_next, f->cell_start, f->nbr_f, f->nbr_nf,
                     (const V4<T>*)a->bmom, a->bflag, (V4<T>*)a->vbar_out, (V4<T>*)a->xbar,
                     ctrl);
  TSPH_LAUNCH_CHECK("adj_solver_iteration");
  return 0;
}

template <class T>
int adj_advect(const TsphParams& prm, TsphObject* f, TsphControl* ctrl, TsphAdjoint* a, void* ws,
               size_t wsb, hipStream_t st) {
  KP<T> k = make_kp<T>(prm);
  WsLayout L;
  TSPH_REQUIRE(carve_ws(ws, wsb, f->capacity, prm.node_num, &L), TSPH_ECAPACITY,
               "tsph_adj_advect: workspace too small");
  if (f->n == 0) return 0;
  int nb = blocks_for(f->n);
  // nbar == NULL: gamma == 0 and its gradient is not wanted, so every term
  // of the surface-tension reverse vanishes and it is skipped
  if (a->nbar) {
    // b_j = rho0_j dt vbar_j into the workspace's record buffer (free between
    // forward Jacobi iterations)
    hipLaunchKernelGGL(k_adj_st_rec<T>, dim3(nb), dim3(kBlock), 0, st, f->n,
                       (const T*)f->rest_density, (const V4<T>*)a->vbar, (V4<T>*)L.krec, ctrl);
    hipLaunchKernelGGL(k_adj_surface_tension<T>, dim3(nb), dim3(kBlock), 0, st, k, f->n,
                       f->capacity, (const V4<T>*)f->pos, (const T*)f->mass,
                       (const T*)f->sph_density, (const T*)f->rest_density,
                       (const T*)f->volume_frac, f->cell_start, f->nbr_f, f->nbr_nf,
                       (const V4<T>*)a->tape_normal, (const V4<T>*)a->vbar,
                       (const V4<T>*)L.krec, (V4<T>*)a->nbar,
                       (T*)a->rhobar, (V4<T>*)a->xbar, L.partial, ctrl);
    hipLaunchKernelGGL(k_adj_grad_add, dim3(1), dim3(kBlock), 0, st, L.partial, nb, a->grad, 1);
  }
  // (dt vbar, fv) per particle into the workspace's boundary alpha_1 buffer
  // (pass A's output for the forward psi sweeps, free during a reverse)
  hipLaunchKernelGGL(k_adj_vis_rec<T>, dim3(nb), dim3(kBlock), 0, st, k, f->n,
                     (const T*)f->rest_density, (const V4<T>*)a->vbar, (V4<T>*)L.a1b, ctrl);
  hipLaunchKernelGGL(k_adj_viscosity<T>, dim3(nb), dim3(kBlock), 0, st, k, f->n, f->capacity,
                     (const V4<T>*)f->pos, (const T*)f->sph_density,
                     (const T*)f->rest_density, f->cell_start, (const int*)nullptr,
                     (const V4<T>*)nullptr, f->nbr_f, f->nbr_nf, (const V4<T>*)a->tape_vel,
                     (const V4<T>*)a->nbar, (const V4<T>*)a->vbar, (const V4<T>*)L.a1b,
                     (const T*)a->kappa_next,
                     (V4<T>*)a->vbar_out, (T*)a->rhobar, (V4<T>*)a->xbar, L.partial, ctrl);
  hipLaunchKernelGGL(k_adj_grad_add, dim3(1), dim3(kBlock), 0, st, L.partial, nb, a->grad, 0);
  TSPH_LAUNCH_CHECK("adj_advect");
  return 0;
}

template <class T>
int adj_density_alpha(const TsphParams& prm, TsphObject* f, TsphObject* b, TsphAdjoint* a,
                      hipStream_t st) {
  KP<T> k = make_kp<T>(prm);
  if (f->n == 0) return 0;
  int nb = blocks_for(f->n);
  hipLaunchKernelGGL(k_adj_alpha_prep<T>, dim3(nb), dim3(kBlock), 0, st, f->n, (const T*)f->alpha,
                     (const V4<T>*)f->alpha_1, (const T*)f->mass, (const T*)a->alphabar,
                     (V4<T>*)a->a1bar);
  hipLaunchKernelGGL(k_adj_pass_a<T>, dim3(nb), dim3(kBlock), 0, st, k, f->n, f->capacity,
                     (const V4<T>*)f->pos, (const T*)f->mass, (const T*)a->rhobar,
                     (const V4<T>*)a->a1bar, f->cell_start, f->nbr_f, f->nbr_nf,
                     (const V4<T>*)a->bmom, a->bflag, (V4<T>*)a->xbar);
  TSPH_LAUNCH_CHECK("adj_density_alpha");
  return 0;
}

template <class T>
int adj_bound_sums(const TsphParams& prm, TsphObject* f, TsphObject* b, TsphAdjoint* a,
                   hipStream_t st) {
  KP<T> k = make_kp<T>(prm);
  if (f->n == 0) return 0;
  hipLaunchKernelGGL(k_adj_bound_sums<T>, dim3(blocks_for(f->n)), dim3(kBlock), 0, st, k, f->n,
                     f->capacity, (const V4<T>*)f->pos, f->nbr_nb, (const V4<T>*)b->pos,
                     b->cell_start, (V4<T>*)a->bmom, a->bflag);
  TSPH_LAUNCH_CHECK("adj_bound_sums");
  return 0;
}

#define TSPH_INST(T)                                                                           \
  template int adj_begin<T>(const TsphParams&, TsphObject*, TsphControl*, TsphAdjoint*,         \
                            const int*, const void*, const void*, hipStream_t);                 \
  template int adj_end<T>(const TsphParams&, TsphObject*, TsphAdjoint*, const int*, void*,      \
                          void*, hipStream_t);                                                  \
  template int adj_solver_iteration<T>(const TsphParams&, TsphObject*, TsphObject*,             \
                                       TsphControl*, TsphAdjoint*, void*, size_t, hipStream_t); \
  template int adj_advect<T>(const TsphParams&, TsphObject*, TsphControl*, TsphAdjoint*, void*, \
                             size_t, hipStream_t);                                              \
  template int adj_density_alpha<T>(const TsphParams&, TsphObject*, TsphObject*, TsphAdjoint*,  \
                                    hipStream_t);                                               \
  template int adj_bound_sums<T>(const TsphParams&, TsphObject*, TsphObject*, TsphAdjoint*,     \
                                 hipStream_t);
TSPH_INST(float)
TSPH_INST(double)
#undef TSPH_INST

}  // namespace tsph
