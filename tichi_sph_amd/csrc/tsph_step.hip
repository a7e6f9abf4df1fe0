// Fused per-step kernels of the DFSPH / VFSPH step (main.py:175-256).
//
// Storage is in cell order (tsph_grid.hip).  The 27 (3D) / 9 (2D) cells of a
// particle's neighbourhood are visited as 9 / 3 contiguous row runs: cells
// (x-R .. x+R) of one (y, z) row have consecutive codes and therefore
// consecutive slots, in the reference's template order (x fastest,
// sph_config.py:219-224) and in-cell order.  Pass A walks the runs once per
// step, testing every candidate, and records the in-range neighbours
// (0 < r < h, exact f32 test) of each fluid particle in reference order;
// every later sweep iterates only those lists.  Terms beyond h are exact
// zeros in the reference, so skipping them changes no value.
#include <cstring>
#include <type_traits>

#include "tsph_sweep.h"

namespace tsph {

// Pair terms of pass A for one in-range candidate (strict arithmetic, the
// order of sph.py:61-64, 81, 97).
template <class T>
struct PassAAcc {
  T w = 0, cp = 0, dn = 0, ax = 0, ay = 0, az = 0, a2 = 0;
  T bx = 0, by = 0, bz = 0;  // the boundary neighbours' share of alpha_1
  // XM: the solver alias X is the mass (DFSPH), so X_j = m_j
  template <bool FQ, bool XM>
  __device__ __forceinline__ bool add(const KP<T>& k, const V4<T>& p, const V4<T>& q, T mj, T vj,
                                      bool fluid) {
    T dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
    T r2, r, qh;
    {
      TSPH_STRICT
      r2 = s_r2(dx, dy, dz);
      r = pa_sqrt<FQ>(r2);
      qh = pa_quot<FQ>(k, r);
      T Wr = s_W_q_sel(k, qh);
      w = w + Wr;
      cp = cp + Wr * vj;
      dn = dn + Wr * mj;
      if (!(r > T(0))) return false;
    }
    // W'(q) / r from the r and q just formed: 6 sig/h^2 (3q - 2) for q <= 1/2
    // (no division), -6 sig/h (1 - q)^2 / r above
    const T u = T(1) - qh;
    const T g = qh <= T(0.5) ? k.wg_a * (T(3) * qh - T(2)) : -k.wg_b * (u * u) * f_rcp(r);
    T Xj = q.w;
    T xg = Xj * g;
    ax += xg * dx;
    ay += xg * dy;
    az += xg * dz;
    if (fluid) {
      // W'^2 X_j^2 / m_j; for DFSPH X_j is m_j and the quotient is X_j
      const T x2m = XM ? Xj : (Xj * Xj) * f_rcp(mj);
      a2 += (g * g) * r2 * x2m;
    } else {
      bx += xg * dx;
      by += xg * dy;
      bz += xg * dz;
    }
    return true;
  }
};

// List slots per particle staged in LDS by pass A (kPassAK x kBlock ints per
// block; 22 KB keeps the kernel at its register-bound occupancy).
constexpr int kPassAK = 22;

// ---------------------------------------------------------------------------
// Pass A, fluid particles: SPH_prepare_attr (f,f)(f,b), SPH_prepare_alpha_1
// (f,f)(f,b), SPH_prepare_alpha_2 (f,f), SPH_prepare_alpha, SPH_vel_2_vel_adv,
// neighbour lists.  Strict arithmetic: bit-for-bit the reference's order.
// ---------------------------------------------------------------------------
template <class T>
__device__ __forceinline__ void pass_a_epilogue(const KP<T>& k, int i, const V4<T>& p, T mi,
                                                const PassAAcc<T>& a, int nf, int nbb,
                                                T* W, T* comp, T* dens, V4<T>* alpha_1,
                                                T* alpha_2, T* alpha, T* vol, const V4<T>* vel,
                                                V4<T>* vel_adv, int* nbr_nf, int* nbr_nb,
                                                TsphControl* ctrl) {
  TSPH_STRICT
  W[i] = a.w;
  comp[i] = a.cp;
  dens[i] = a.dn;
  alpha_1[i] = V4<T>{a.ax, a.ay, a.az, T(0)};
  alpha_2[i] = a.a2;
  T dot = s_r2(a.ax, a.ay, a.az);
  T al = dot / mi + a.a2;  // SPH_prepare_alpha, sph.py:103-105
  if (al < T(1e-4)) al = T(1e-4);
  alpha[i] = al;
  T psi = k.solver == TSPH_VFSPH ? a.cp : a.dn;
  const T Vi = p.w / psi;  // X / sph_psi, reused by the viscosity and normal sweeps
  vol[i] = Vi;
  // SPH_vel_2_vel_adv, sph.py:253-256.  The unused 4th lane carries V_i until
  // the advection sweep, so that sweep gathers one record less per pair; the
  // Jacobi sweeps only update x, y, z.
  const V4<T> v = vel[i];
  vel_adv[i] = V4<T>{v.x, v.y, v.z, Vi};
  if (nf > k.list_cap_f) {
    atomicAdd(&ctrl->overflow_f, 1);
    nf = k.list_cap_f + 1;
  }
  if (nbb > k.list_cap_b) {
    atomicAdd(&ctrl->overflow_b, 1);
    nbb = k.list_cap_b + 1;
  }
  nbr_nf[i] = nf;
  nbr_nb[i] = nbb;
}

template <class T, bool FQ, bool XM>
__device__ __forceinline__ void k_pass_a_fluid_body(Blk blk, const KP<T>& k, int n, int cap, const V4<T>* __restrict__ fpos,
                   const T* __restrict__ fmass, const T* __restrict__ frvol,
                   const int* __restrict__ fstart, int nb, const V4<T>* __restrict__ bpos,
                   const T* __restrict__ bmass, const T* __restrict__ brvol,
                   const int* __restrict__ bstart, T* __restrict__ W, T* __restrict__ comp,
                   T* __restrict__ dens, V4<T>* __restrict__ alpha_1, T* __restrict__ alpha_2,
                   T* __restrict__ alpha, T* __restrict__ vol, const V4<T>* __restrict__ vel,
                   V4<T>* __restrict__ vel_adv, int* __restrict__ nbr_f, int* __restrict__ nbr_b,
                   int* __restrict__ nbr_nf, int* __restrict__ nbr_nb, int* __restrict__ flag,
                   V4<T>* __restrict__ a1b, TsphControl* ctrl, int* __restrict__ stage) {
  const int i = xcd_block(blk.b, blk.nb) * kBlock + threadIdx.x;
  if (i >= n) return;
  const int capf = k.list_cap_f, capb = k.list_cap_b;
  // the first kPassAK slots of each list go to this lane's column of the
  // block's LDS stage; the rest are appended to the list directly
  const int nstage = capf < kPassAK ? capf : kPassAK;
  const int nstage_b = capb < kPassAK ? capb : kPassAK;
  flag[i] = 0;  // SPH_clean_value, sph.py:30
  const V4<T> p = fpos[i];
  const T mi = fmass[i];
  int c[3];
  node_encode(k, p.x, p.y, p.z, c);
  PassAAcc<T> acc;
  int nf = 0, nbb = 0;
  // pos.w is the solver alias X (mass for DFSPH, rest volume for VFSPH), so
  // only the other of the two fields is loaded
  const T* __restrict__ fo = XM ? frvol : fmass;
  const T* __restrict__ bo = XM ? brvol : bmass;
  walk_cells_batched<kPassAU>(k, fstart, c, fpos, [&](int j, const V4<T>& q) {
    if (s_r2(p.x - q.x, p.y - q.y, p.z - q.z) <= k.r2max) {
      const T o = fo[j];
      if (acc.template add<FQ, XM>(k, p, q, XM ? q.w : o, XM ? o : q.w, true)) {
        if (nf < nstage)
          stage[nf * kBlock + threadIdx.x] = j;
        else if (nf < capf)
          nbr_f[(long)nf * cap + i] = j;
        ++nf;
      }
    }
  });
  // the staged slots, row by row: the wave's lanes store consecutive
  // entries of one slot row together, whole lines at a time, instead of one
  // scattered entry per lane whenever a lane finds a neighbour
  const int ns = nf < nstage ? nf : nstage;
  for (int s = 0; s < ns; ++s) nbr_f[(long)s * cap + i] = stage[s * kBlock + threadIdx.x];
  walk_cells_batched<kPassAU>(k, bstart, c, bpos, [&](int j, const V4<T>& q) {
    if (s_r2(p.x - q.x, p.y - q.y, p.z - q.z) <= k.r2max) {
      const T o = bo[j];
      if (acc.template add<FQ, XM>(k, p, q, XM ? q.w : o, XM ? o : q.w, false)) {
        if (nbb < nstage_b)
          stage[nbb * kBlock + threadIdx.x] = j;
        else if (nbb < capb)
          nbr_b[(long)nbb * cap + i] = j;
        ++nbb;
      }
    }
  });
  const int nsb = nbb < nstage_b ? nbb : nstage_b;
  for (int s = 0; s < nsb; ++s) nbr_b[(long)s * cap + i] = stage[s * kBlock + threadIdx.x];
  a1b[i] = V4<T>{acc.bx, acc.by, acc.bz, T(0)};
  pass_a_epilogue(k, i, p, mi, acc, nf, nbb, W, comp, dens, alpha_1, alpha_2, alpha, vol, vel,
                  vel_adv, nbr_nf, nbr_nb, ctrl);
}

// Runs f(fq, xm) with the fluid body's compile-time variant for these
// constants: the verified fast quotient, and whether X is the mass.
template <class T, class F>
__device__ __forceinline__ void pass_a_fluid_variant(const KP<T>& k, F&& f) {
  using Y = std::true_type;
  using N = std::false_type;
  const bool xm = k.solver != TSPH_VFSPH;
  if (k.fastq) {
    if (xm) f(Y{}, Y{});
    else f(Y{}, N{});
  } else {
    if (xm) f(N{}, Y{});
    else f(N{}, N{});
  }
}

template <class T>
__global__ void __launch_bounds__(kBlock)
    k_pass_a_fluid(KP<T> k, int n, int cap, const V4<T>* __restrict__ fpos,
                   const T* __restrict__ fmass, const T* __restrict__ frvol,
                   const int* __restrict__ fstart, int nb, const V4<T>* __restrict__ bpos,
                   const T* __restrict__ bmass, const T* __restrict__ brvol,
                   const int* __restrict__ bstart, T* __restrict__ W, T* __restrict__ comp,
                   T* __restrict__ dens, V4<T>* __restrict__ alpha_1, T* __restrict__ alpha_2,
                   T* __restrict__ alpha, T* __restrict__ vol, const V4<T>* __restrict__ vel,
                   V4<T>* __restrict__ vel_adv, int* __restrict__ nbr_f, int* __restrict__ nbr_b,
                   int* __restrict__ nbr_nf, int* __restrict__ nbr_nb, int* __restrict__ flag,
                   V4<T>* __restrict__ a1b, TsphControl* ctrl) {
  __shared__ int stage[kPassAK * kBlock];
  const Blk blk{(int)blockIdx.x, (int)gridDim.x};
  pass_a_fluid_variant(k, [&](auto fq, auto xm) {
    k_pass_a_fluid_body<T, decltype(fq)::value, decltype(xm)::value>(blk, k, n, cap, fpos, fmass, frvol, fstart, nb, bpos, bmass, brvol, bstart, W, comp, dens, alpha_1, alpha_2, alpha, vol, vel, vel_adv, nbr_f, nbr_b, nbr_nf, nbr_nb, flag, a1b, ctrl, stage);
  });
}

// Pass A, boundary particles: SPH_prepare_attr (b,b)(b,f), SPH_prepare_alpha_2
// (b,f), SPH_prepare_alpha (alpha_1 of the boundary is never accumulated).
// The reference adds all (b,b) pairs before the (b,f) ones (main.py:189-190),
// and the boundary never moves, so the (b,b) partial sums are kept in ss
// (when given): with `kept` they are read instead of summed, otherwise summed
// and stored.
template <class T, bool FQ>
__device__ __forceinline__ void k_pass_a_bound_body(Blk blk, const KP<T>& k, int nb, const V4<T>* __restrict__ bpos, const T* __restrict__ bmass,
                   const T* __restrict__ brvol, const int* __restrict__ bstart,
                   const V4<T>* __restrict__ fpos, const T* __restrict__ fmass,
                   const T* __restrict__ frvol, const int* __restrict__ fstart,
                   T* __restrict__ W, T* __restrict__ comp, T* __restrict__ dens,
                   V4<T>* __restrict__ alpha_1, T* __restrict__ alpha_2, T* __restrict__ alpha,
                   T* __restrict__ vol, V4<T>* __restrict__ ss, bool kept) {
  TSPH_STRICT
  int i = xcd_block(blk.b, blk.nb) * kBlock + threadIdx.x;
  if (i >= nb) return;
  V4<T> p = bpos[i];
  int c[3];
  node_encode(k, p.x, p.y, p.z, c);
  T w = 0, cp = 0, dn = 0, a2 = 0;
  if (ss != nullptr && kept) {
    const V4<T> s0 = ss[i];
    w = s0.x;
    cp = s0.y;
    dn = s0.z;
  } else {
    walk_cells_batched<kPassAU>(k, bstart, c, bpos, [&](int j, const V4<T>& q) {
      T r2 = s_r2(p.x - q.x, p.y - q.y, p.z - q.z);
      if (r2 <= k.r2max) {
        T Wr = s_W_q(k, pa_quot<FQ>(k, pa_sqrt<FQ>(r2)));
        w = w + Wr;
        cp = cp + Wr * brvol[j];
        dn = dn + Wr * bmass[j];
      }
    });
    if (ss != nullptr) ss[i] = V4<T>{w, cp, dn, T(0)};
  }
  walk_cells_batched<kPassAU>(k, fstart, c, fpos, [&](int j, const V4<T>& q) {
    T r2 = s_r2(p.x - q.x, p.y - q.y, p.z - q.z);
    if (r2 <= k.r2max) {
      T r = pa_sqrt<FQ>(r2);
      T qh = pa_quot<FQ>(k, r);
      T Wr = s_W_q(k, qh);
      w = w + Wr;
      cp = cp + Wr * frvol[j];
      dn = dn + Wr * fmass[j];
      if (r > T(0)) {
        T wg = s_Wg_q(k, qh);
        T Xj = q.w;
        T wg2 = wg * wg;
        a2 = a2 + wg2 * (Xj * Xj) / fmass[j];
      }
    }
  });
  W[i] = w;
  comp[i] = cp;
  dens[i] = dn;
  alpha_1[i] = V4<T>{T(0), T(0), T(0), T(0)};
  alpha_2[i] = a2;
  T al = T(0) / bmass[i] + a2;
  if (al < T(1e-4)) al = T(1e-4);
  alpha[i] = al;
  T psi = k.solver == TSPH_VFSPH ? cp : dn;
  vol[i] = p.w / psi;
}

template <class T>
__global__ void __launch_bounds__(kBlock)
    k_pass_a_bound(KP<T> k, int nb, const V4<T>* __restrict__ bpos, const T* __restrict__ bmass,
                   const T* __restrict__ brvol, const int* __restrict__ bstart,
                   const V4<T>* __restrict__ fpos, const T* __restrict__ fmass,
                   const T* __restrict__ frvol, const int* __restrict__ fstart,
                   T* __restrict__ W, T* __restrict__ comp, T* __restrict__ dens,
                   V4<T>* __restrict__ alpha_1, T* __restrict__ alpha_2, T* __restrict__ alpha,
                   T* __restrict__ vol, V4<T>* __restrict__ ss, int kept) {
  k_pass_a_bound_body<T, false>(Blk{(int)blockIdx.x, (int)gridDim.x}, k, nb, bpos, bmass, brvol, bstart, fpos, fmass, frvol, fstart, W, comp, dens, alpha_1, alpha_2, alpha, vol, ss, kept != 0);
}

// Per-step reset of the control block's step-scoped state: general_flag = 1
// (SPH_clean_value, sph.py:24), FBM round count, list-overflow counters.
__device__ __forceinline__ void step_reset_body(TsphControl* ctrl) {
  if (threadIdx.x == 0) {
    ctrl->general_flag = 1;
    ctrl->fbm_rounds = 0;
    ctrl->overflow_f = 0;
    ctrl->overflow_b = 0;
  }
}
__global__ void k_step_reset(TsphControl* ctrl) { step_reset_body(ctrl); }

// Pass A of both objects in one launch: blocks [0, nbf) run the fluid body,
// the rest the boundary body.  The two halves are independent (each writes
// only its own object's fields), so the boundary blocks fill the tail of the
// fluid's grid instead of waiting for it.
template <class T>
__device__ __forceinline__ void pass_a_both_body(const KP<T>& k, const TsphObject& f,
                                                 const TsphObject& b, int nbf, V4<T>* a1b,
                                                 TsphControl* ctrl, int* stage) {
  if ((int)blockIdx.x < nbf) {
    pass_a_fluid_variant(k, [&](auto fq, auto xm) {
      k_pass_a_fluid_body<T, decltype(fq)::value, decltype(xm)::value>(
          Blk{(int)blockIdx.x, nbf}, k, f.n, f.capacity, (const V4<T>*)f.pos, (const T*)f.mass,
          (const T*)f.rest_volume, f.cell_start, b.n, (const V4<T>*)b.pos, (const T*)b.mass,
          (const T*)b.rest_volume, b.cell_start, (T*)f.W, (T*)f.sph_compression,
          (T*)f.sph_density, (V4<T>*)f.alpha_1, (T*)f.alpha_2, (T*)f.alpha, (T*)f.vol,
          (const V4<T>*)f.vel, (V4<T>*)f.vel_adv, f.nbr_f, f.nbr_b, f.nbr_nf, f.nbr_nb, f.flag,
          a1b, ctrl, stage);
    });
  } else {
    k_pass_a_bound_body<T, false>(Blk{(int)blockIdx.x - nbf, (int)gridDim.x - nbf}, k, b.n,
                               (const V4<T>*)b.pos, (const T*)b.mass, (const T*)b.rest_volume,
                               b.cell_start, (const V4<T>*)f.pos, (const T*)f.mass,
                               (const T*)f.rest_volume, f.cell_start, (T*)b.W,
                               (T*)b.sph_compression, (T*)b.sph_density, (V4<T>*)b.alpha_1,
                               (T*)b.alpha_2, (T*)b.alpha, (T*)b.vol, (V4<T>*)b.static_sums,
                               (b.flags & TSPH_OBJ_STATIC_SUMS) != 0);
  }
}
template <class T>
__global__ void __launch_bounds__(kBlock)
    k_pass_a_both(KP<T> k, TsphObject f, TsphObject b, int nbf, V4<T>* a1b, TsphControl* ctrl) {
  __shared__ int stage[kPassAK * kBlock];
  pass_a_both_body<T>(k, f, b, nbf, a1b, ctrl, stage);
}

// The check behind TSPH_FLAG_FAST_QUOTIENT: for every FP32 r^2 in
// [0, r2_max] (bit patterns 0 .. umax), the reduced and the IEEE sequences
// are compared on W(q) and r > 0 (out[0]: inputs that differ, out[2]: the
// largest such bit pattern) and on W'(q) (out[1]).  Each wave adds its
// counts once.
__global__ void __launch_bounds__(kBlock) k_verify_fast_quotient(KP<float> k, uint32_t umax, int* out) {
  int bad = 0, badg = 0;
  uint32_t last = 0;
  const uint32_t stride = gridDim.x * kBlock;
  for (uint32_t u = blockIdx.x * kBlock + threadIdx.x; u <= umax; u += stride) {
    const float x = __uint_as_float(u);
    float rs, ws, gs;
    {
      TSPH_STRICT
      rs = s_sqrt(x);
      const float qs = rs / k.h;
      ws = s_W_q(k, qs);
      gs = s_Wg_q(k, qs);
    }
    const float rf = f_sqrt_cr(x);
    const float qf = f_quot_h(rf, k.h, k.inv_h);
    const float wf = s_W_q(k, qf), gf = s_Wg_q(k, qf);
    const bool dw = (__float_as_uint(ws) != __float_as_uint(wf)) || ((rs > 0.0f) != (rf > 0.0f));
    bad += dw;
    badg += __float_as_uint(gs) != __float_as_uint(gf);
    if (dw) last = u;
  }
  for (int o = 32; o > 0; o >>= 1) {
    bad += __shfl_xor(bad, o, 64);
    badg += __shfl_xor(badg, o, 64);
    last = max(last, (uint32_t)__shfl_xor((int)last, o, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    if (bad) atomicAdd(out, bad);
    if (badg) atomicAdd(out + 1, badg);
    if (last) atomicMax((unsigned int*)(out + 2), last);
  }
}

__global__ void k_fill3_i32(int* p, int v0, int v1, int v2) {
  p[0] = v0;
  p[1] = v1;
  p[2] = v2;
}

int verify_fast_quotient(const TsphParams& prm, int* out, hipStream_t st) {
  if (prm.precision != 4) {
    hipLaunchKernelGGL(k_fill3_i32, dim3(1), dim3(1), 0, st, out, 1, 1, 0);
    TSPH_LAUNCH_CHECK("verify_fast_quotient");
    return 0;
  }
  KP<float> k = make_kp<float>(prm);
  uint32_t umax;
  std::memcpy(&umax, &k.r2max, sizeof(umax));
  TSPH_REQUIRE(k.r2max >= 0.0f && umax < 0x7f800000u, TSPH_EINVAL,
               "tsph_verify_fast_quotient: r2_max is not a finite non-negative value");
  hipLaunchKernelGGL(k_fill3_i32, dim3(1), dim3(1), 0, st, out, 0, 0, 0);
  hipLaunchKernelGGL(k_verify_fast_quotient, dim3(8192), dim3(kBlock), 0, st, k, umax, out);
  TSPH_LAUNCH_CHECK("verify_fast_quotient");
  return 0;
}

// ---------------------------------------------------------------------------
// Solver stop state (main.py:201-213 / 222-234).
// ---------------------------------------------------------------------------
template <class T>
__device__ __forceinline__ void solver_begin_body(const KP<T>& k, TsphControl* ctrl, int solve,
                                                  int* signal = nullptr, int token = 0) {
  int t = threadIdx.x;
  T thr = solve == 0 ? k.div_thr : k.comp_thr;
  T comp = (T)ctrl->compression;
  bool enter = (0 < k.iter_min) || (comp > thr);
  // run_u[q] = 1 iff iteration q of this solve runs; decided one iteration ahead
  for (int q = t; q < TSPH_MAX_ITER_LOG; q += blockDim.x) ctrl->run_u[q] = (q == 0 && enter) ? 1 : 0;
  if (t == 0) {
    ctrl->it = 0;
    ctrl->solve = solve;
    ctrl->stop = enter ? 0 : 1;
    if (solve == 0) ctrl->div_iter = 0;
    else ctrl->incom_iter = 0;
    ctrl->signal = (uint64_t)(uintptr_t)signal;
    ctrl->token = token;
    if (signal && !enter) {  // a solve that runs no iteration stops here: (it 0, stop)
      const uint64_t w = ((uint64_t)(uint32_t)token << 32) | 1u;
      __hip_atomic_store((uint64_t*)signal, w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
  }
}
template <class T>
__global__ void k_solver_begin(KP<T> k, TsphControl* ctrl, int solve, int* signal, int token) {
  solver_begin_body(k, ctrl, solve, signal, token);
}

// Jacobi half-step P: IPPE_adv_psi_init + IPPE_adv_psi (f,f)(f,b) +
// IPPE_psi_adv_non_negative (sph.py:196-226) for iteration kk of the solve.
template <class T>
__device__ __forceinline__ void k_solver_psi_body(Blk blk, const KP<T>& k, int n, int cap, int kk, const V4<T>* __restrict__ fpos,
                 const V4<T>* __restrict__ vadv, const T* __restrict__ psi_rho,
                 const T* __restrict__ rdens, const T* __restrict__ alpha,
                 const int* __restrict__ fstart, const V4<T>* __restrict__ bpos,
                 const int* __restrict__ bstart, const int* __restrict__ nbr_f,
                 const int* __restrict__ nbr_b, const int* __restrict__ nbr_nf,
                 const int* __restrict__ nbr_nb, T* __restrict__ psi_adv, T* __restrict__ kappa,
                 V4<T>* __restrict__ krec, const V4<T>* __restrict__ a1b,
                 V4<T>* __restrict__ tape_v, T* __restrict__ tape_k,
                 const T* __restrict__ clamp_ref,
                 double* __restrict__ partial, TsphControl* ctrl) {
  __shared__ double lds[kWaves];
  if (!ctrl->run_u[kk]) return;
  const int base = xcd_block(blk.b, blk.nb) * kBlock;
  const int c = coop_lane<kCoop>();
  const T dt = (T)ctrl->dt;
  double err = 0;
#pragma unroll 1
  for (int round = 0; round < kCoop; ++round) {
    const int i = coop_particle<kCoop>(base, round);
    if (i >= n) break;
    const V4<T> p = fpos[i];
    const V4<T> va = vadv[i];
    // the particle's own fields, requested before the walk so that their
    // latency hides behind it
    const V4<T> B = a1b[i];
    const T rp = rest_psi_of(k, rdens, i);
    const T pr = psi_rho[i];
    const T al = alpha[i];
    T sum = 0;  // this lane's share of sum_j X_j (v_i - v_j) . grad W_ij
    auto term = [&](const Rec2<T>& d) {
      const V4<T>& q = d.a;
      const V4<T>& vj = d.b;
      T dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
      T r2 = dx * dx + dy * dy + dz * dz, r;
      T g = f_grad(k, r2, &r);
      T dot = (dx * (va.x - vj.x) + dy * (va.y - vj.y) + dz * (va.z - vj.z)) * g;
      sum += dot * q.w;
    };
    // (f,f) pairs from the list; the (f,b) pairs have v_j = 0 (boundary
    // vel_adv is never written), so their sum is v_i . sum_b X_j grad W_ij,
    // the boundary share of alpha_1 that pass A recorded
    sweep_lists_coop<kCoop, T, Rec2<T>>(
        k, cap, i, c, p, fstart, bstart, fpos, bpos, nbr_f, nullptr, nbr_nf[i], 0,
        [&](int j) { return Rec2<T>{fpos[j], vadv[j]}; },
        [&](int) { return Rec2<T>{}; }, term, [&](const Rec2<T>&) {});
    sum = group_sum<kCoop>(sum);
    if (c != 0) continue;
    sum += va.x * B.x + va.y * B.y + va.z * B.z;
    T psi = (pr - rp) + sum * dt;
    // IPPE_psi_adv_non_negative (sph.py:219-226); with a clamp reference, the
    // clamp set of a recorded iteration instead (psi clamped exactly where the
    // recorded kappa is 0): the path whose derivative the adjoint takes
    if (clamp_ref ? clamp_ref[i] == T(0) : psi < T(0)) psi = T(0);
    psi_adv[i] = psi;
    const T kap = psi / al;
    kappa[i] = kap;
    krec[i] = V4<T>{p.x, p.y, p.z, kap};  // the record the update sweep gathers
    if (tape_v) {  // record the iteration for the adjoint
      tape_v[i] = va;
      tape_k[i] = kap;
    }
    err += (double)(psi / rp);
  }
  double s = block_sum(err, lds);
  if (threadIdx.x == 0) partial[blk.b] = s;
}

template <class T>
__global__ void __launch_bounds__(kBlock)
    k_solver_psi(KP<T> k, int n, int cap, int kk, const V4<T>* __restrict__ fpos,
                 const V4<T>* __restrict__ vadv, const T* __restrict__ psi_rho,
                 const T* __restrict__ rdens, const T* __restrict__ alpha,
                 const int* __restrict__ fstart, const V4<T>* __restrict__ bpos,
                 const int* __restrict__ bstart, const int* __restrict__ nbr_f,
                 const int* __restrict__ nbr_b, const int* __restrict__ nbr_nf,
                 const int* __restrict__ nbr_nb, T* __restrict__ psi_adv, T* __restrict__ kappa,
                 V4<T>* __restrict__ krec, const V4<T>* __restrict__ a1b,
                 V4<T>* __restrict__ tape_v, T* __restrict__ tape_k,
                 const T* __restrict__ clamp_ref,
                 double* __restrict__ partial, TsphControl* ctrl) {
  k_solver_psi_body<T>(Blk{(int)blockIdx.x, (int)gridDim.x}, k, n, cap, kk, fpos, vadv, psi_rho, rdens, alpha, fstart, bpos, bstart, nbr_f, nbr_b, nbr_nf, nbr_nb, psi_adv, kappa, krec, a1b, tape_v, tape_k, clamp_ref, partial, ctrl);
}

// Jacobi half-step U: IPPE_update_vel_adv (f,f)(f,b) (sph.py:229-244), with
// kappa = psi_adv / alpha formed once per particle.  Block 0 also finishes
// iteration kk's compression and takes the loop decision for iteration kk+1.
//
// The pair coefficient of a fluid neighbour is kappa_i X_j + kappa_j X_i, of
// a boundary neighbour kappa_i X_j (boundary kappa is never written: 0).
// Summed over all neighbours, the kappa_i X_j halves are kappa_i times
// alpha_1 = sum_j X_j grad W_ij of pass A (positions and X do not change
// within the step), so the sweep only visits the fluid list, gathering the
// one record (x_j, kappa_j) that the psi half-step wrote.
template <class T>
__device__ __forceinline__ void k_solver_update_body(Blk blk, const KP<T>& k, int n, int cap, int kk, const V4<T>* __restrict__ fpos,
                    const T* __restrict__ fmass, const T* __restrict__ kappa,
                    const V4<T>* __restrict__ krec, const V4<T>* __restrict__ alpha_1,
                    const int* __restrict__ fstart, const V4<T>* __restrict__ bpos,
                    const int* __restrict__ bstart, const int* __restrict__ nbr_f,
                    const int* __restrict__ nbr_b, const int* __restrict__ nbr_nf,
                    const int* __restrict__ nbr_nb, V4<T>* __restrict__ vadv,
                    const double* __restrict__ partial, TsphControl* ctrl) {
  __shared__ double lds[kWaves];
  if (!ctrl->run_u[kk]) return;
  if (blk.b == 0) {
    // IPPE_psi_adv_non_negative's 0-D reduction (sph.py:221-226) and the loop
    // test of main.py:201-213 / 222-234
    double tot = sum_partials(partial, blk.nb, lds);
    if (threadIdx.x == 0) {
      TSPH_STRICT
      T comp = (T)tot / (T)n;
      ctrl->compression = (double)comp;
      int solve = ctrl->solve;
      if (solve == 0) ctrl->err_div[kk] = (double)comp;
      else ctrl->err_incom[kk] = (double)comp;
      int it = kk + 1;
      ctrl->it = it;
      if (solve == 0) ctrl->div_iter = it;
      else ctrl->incom_iter = it;
      T thr = solve == 0 ? k.div_thr : k.comp_thr;
      bool cont = (it < k.iter_min) || (comp > thr);
      bool run = it <= k.iter_max && cont && it < TSPH_MAX_ITER_LOG;
      if (it < TSPH_MAX_ITER_LOG) ctrl->run_u[it] = run ? 1 : 0;
      ctrl->stop = run ? 0 : 1;
      if (ctrl->signal) {  // early stop signal to the host: one 64-bit system-scope store
        const uint64_t w = ((uint64_t)(uint32_t)ctrl->token << 32) | (uint32_t)(2 * it + (run ? 0 : 1));
        __hip_atomic_store((uint64_t*)(uintptr_t)ctrl->signal, w, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_SYSTEM);
      }
    }
  }
  const int base = xcd_block(blk.b, blk.nb) * kBlock;
  const int cl = coop_lane<kCoopU>();
  const T dt = (T)ctrl->dt;
#pragma unroll 1
  for (int round = 0; round < kCoopU; ++round) {
    const int i = coop_particle<kCoopU>(base, round);
    if (i >= n) break;
    const V4<T> p = fpos[i];
    const T Xi = p.w;
    T sx = 0, sy = 0, sz = 0;  // this lane's share of sum_j kappa_j grad W_ij
    auto term = [&](const V4<T>& q) {  // q = (x_j, kappa_j)
      T dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
      T r2 = dx * dx + dy * dy + dz * dz, r;
      T g = f_grad(k, r2, &r);
      T s = q.w * g;
      sx += s * dx;
      sy += s * dy;
      sz += s * dz;
    };
    sweep_lists_coop<kCoopU, T, V4<T>>(
        k, cap, i, cl, p, fstart, bstart, fpos, bpos, nbr_f, nullptr, nbr_nf[i], 0,
        [&](int j) { return krec[j]; }, [&](int) { return V4<T>{}; }, term,
        [&](const V4<T>&) {});
    sx = group_sum<kCoopU>(sx);
    sy = group_sum<kCoopU>(sy);
    sz = group_sum<kCoopU>(sz);
    if (cl != 0) continue;
    V4<T> va = vadv[i];
    const T ki = kappa[i];
    const T c = -(T(1) / dt) / fmass[i];
    const V4<T> A = alpha_1[i];
    const T cA = c * ki, cX = c * Xi;
    va.x += cA * A.x + cX * sx;
    va.y += cA * A.y + cX * sy;
    va.z += cA * A.z + cX * sz;
    vadv[i] = va;
  }
}

template <class T>
__global__ void __launch_bounds__(kBlock)
    k_solver_update(KP<T> k, int n, int cap, int kk, const V4<T>* __restrict__ fpos,
                    const T* __restrict__ fmass, const T* __restrict__ kappa,
                    const V4<T>* __restrict__ krec, const V4<T>* __restrict__ alpha_1,
                    const int* __restrict__ fstart, const V4<T>* __restrict__ bpos,
                    const int* __restrict__ bstart, const int* __restrict__ nbr_f,
                    const int* __restrict__ nbr_b, const int* __restrict__ nbr_nf,
                    const int* __restrict__ nbr_nb, V4<T>* __restrict__ vadv,
                    const double* __restrict__ partial, TsphControl* ctrl) {
  k_solver_update_body<T>(Blk{(int)blockIdx.x, (int)gridDim.x}, k, n, cap, kk, fpos, fmass, kappa, krec, alpha_1, fstart, bpos, bstart, nbr_f, nbr_b, nbr_nf, nbr_nb, vadv, partial, ctrl);
}

// ---------------------------------------------------------------------------
// Advection: SPH_vel_adv_2_vel, gravity, viscosity (f,f), surface-tension
// normals (f,f).  Neighbour velocities are read from vel_adv, so this sweep
// must not write vel_adv: the update vel_adv = vel + dt * acce_adv happens in
// the next kernel (k_surface_tension).
// ---------------------------------------------------------------------------
template <class T>
__device__ __forceinline__ void k_advect_body(Blk blk, const KP<T>& k, int n, int cap, const V4<T>* __restrict__ fpos, const T* __restrict__ vf,
             const T* __restrict__ rdens, const int* __restrict__ fstart,
             const int* __restrict__ bstart, const V4<T>* __restrict__ bpos,
             const int* __restrict__ nbr_f, const int* __restrict__ nbr_nf,
             V4<T>* __restrict__ vadv, V4<T>* __restrict__ vel, V4<T>* __restrict__ acce_adv,
             V4<T>* __restrict__ normal, const T* __restrict__ fmass,
             const T* __restrict__ psi_rho, V4<T>* __restrict__ strec) {
  const int base = xcd_block(blk.b, blk.nb) * kBlock;
  const int c = coop_lane<kCoop>();
#pragma unroll 1
  for (int round = 0; round < kCoop; ++round) {
    const int i = coop_particle<kCoop>(base, round);
    if (i >= n) break;
    const V4<T> p = fpos[i];
    const V4<T> vi = vadv[i];
    const T fvis = k.two_d2 * mu_of(k, vf, i) / rdens[i];
    T ax = 0, ay = 0, az = 0, nx = 0, ny = 0, nz = 0;  // this lane's share
    auto term = [&](const Rec2<T>& d) {
      const V4<T>& q = d.a;
      const V4<T>& vj = d.b;
      T Vj = vj.w;  // V_j = X_j / sph_psi_j rides in vel_adv's 4th lane (pass A)
      T dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
      T r2 = dx * dx + dy * dy + dz * dz, r;
      T g = f_grad(k, r2, &r);
      T Ax = (vi.x - vj.x) * dx + (vi.y - vj.y) * dy + (vi.z - vj.z) * dz;
      T s = fvis * Vj * g * Ax * f_rcp(k.eps + r2);  // W_lap, sph_util.py:40-43 (fast quotient)
      ax += s * dx;
      ay += s * dy;
      az += s * dz;
      T t = -Vj * g;  // normal, sph.py:149
      nx += t * dx;
      ny += t * dy;
      nz += t * dz;
    };
    // viscosity and normals are (f,f) only: the boundary list is not visited
    sweep_lists_coop<kCoop, T, Rec2<T>>(
        k, cap, i, c, p, fstart, bstart, fpos, bpos, nbr_f, nullptr, nbr_nf[i], 0,
        [&](int j) { return Rec2<T>{fpos[j], vadv[j]}; },
        [&](int j) { return Rec2<T>{}; }, term, [&](const Rec2<T>&) {});
    ax = group_sum<kCoop>(ax);
    ay = group_sum<kCoop>(ay);
    az = group_sum<kCoop>(az);
    nx = group_sum<kCoop>(nx);
    ny = group_sum<kCoop>(ny);
    nz = group_sum<kCoop>(nz);
    if (c != 0) continue;
    // SPH_vel_adv_2_vel; the spare lane keeps V_i until the tail writes 0:
    // the adjoint's tape of this velocity carries it to the reverse sweep
    vel[i] = vi;
    acce_adv[i] = V4<T>{k.g[0] + ax, k.g[1] + ay, k.g[2] + az, T(0)};
    if (k.gamma != T(0) || k.st_rec) {
      // the records the surface-tension walks gather per candidate j: the
      // normal's spare lane carries +-psi_j, negative where j's 0.99 gate
      // fails (psi_j > 0 for every particle in a cell: its own W(0) term),
      // and the forward walk's (x_j, m_j) goes to strec, so that it loads two
      // records instead of five
      const T psi = psi_rho[i];
      normal[i] = V4<T>{nx * k.h, ny * k.h, nz * k.h, vf[(long)i * k.P] > T(0.99) ? psi : -psi};
      strec[i] = V4<T>{p.x, p.y, p.z, fmass[i]};
    } else {
      normal[i] = V4<T>{nx * k.h, ny * k.h, nz * k.h, T(0)};
    }
  }
}

template <class T>
__global__ void __launch_bounds__(kBlock)
    k_advect(KP<T> k, int n, int cap, const V4<T>* __restrict__ fpos, const T* __restrict__ vf,
             const T* __restrict__ rdens, const int* __restrict__ fstart,
             const int* __restrict__ bstart, const V4<T>* __restrict__ bpos,
             const int* __restrict__ nbr_f, const int* __restrict__ nbr_nf,
             V4<T>* __restrict__ vadv, V4<T>* __restrict__ vel, V4<T>* __restrict__ acce_adv,
             V4<T>* __restrict__ normal, const T* __restrict__ fmass,
             const T* __restrict__ psi_rho, V4<T>* __restrict__ strec) {
  k_advect_body<T>(Blk{(int)blockIdx.x, (int)gridDim.x}, k, n, cap, fpos, vf, rdens, fstart, bstart, bpos, nbr_f, nbr_nf, vadv, vel, acce_adv, normal, fmass, psi_rho, strec);
}

// Surface tension loop 2 (sph.py:151-166) + SPH_advection_update_vel_adv
// (sph.py:247-250).  The curvature term carries no kernel weight, so every
// candidate of the neighbouring cells contributes to it: that part walks the
// cells, gathering only the normal record.  Its r > 0 test is j != i: two
// distinct particles at one position have bit-equal normals (their neighbour
// sums run over the same particles in the same order), so their curvature
// term is an exact zero either way.  The cohesion term's C(r) vanishes
// beyond h, so it is summed over the in-range list.  Both terms share the
// factor 2 psi_0 / (psi_i + psi_j) and gamma.  With gamma == 0 every term is
// an exact zero and only the update runs.
// Candidate records in flight per lane of the forward curvature walk.
constexpr int kStFwdU = 8;
template <class T>
__device__ __forceinline__ void k_surface_tension_body(Blk blk, const KP<T>& k, int n, int cap,
                      const V4<T>* __restrict__ fpos, const T* __restrict__ psi_rho,
                      const T* __restrict__ rdens, const T* __restrict__ vf,
                      const int* __restrict__ fstart, const V4<T>* __restrict__ normal,
                      const V4<T>* __restrict__ strec, const int* __restrict__ nbr_f,
                      const int* __restrict__ nbr_nf, const V4<T>* __restrict__ vel,
                      V4<T>* __restrict__ acce_adv, V4<T>* __restrict__ vadv,
                      const TsphControl* ctrl) {
  int i = xcd_block(blk.b, blk.nb) * kBlock + threadIdx.x;
  if (i >= n) return;
  const T dt = (T)ctrl->dt;
  V4<T> a = acce_adv[i];
  if (k.gamma != T(0) && vf[(long)i * k.P] > T(0.99)) {
    V4<T> p = fpos[i];
    V4<T> ni = normal[i];
    T rp = rest_psi_of(k, rdens, i);
    T psi_i = psi_rho[i];
    int c[3];
    node_encode(k, p.x, p.y, p.z, c);
    // the normal record's spare lane holds +-psi_j, negative where j's 0.99
    // gate fails (the advection sweep writes it).  Reciprocal square root
    // and reciprocals instead of the IEEE forms: the pair terms of these
    // sums are not part of the bit-exact sums.
    T ux = 0, uy = 0, uz = 0;  // sum of (n_i - n_j) / (psi_i + psi_j)
    walk_cells_prefetch<kStFwdU>(k, fstart, c, [&](int j) { return normal[j]; },
                                 [&](int j, const V4<T>& nj) {
      if (j != i && nj.w > T(0)) {
        const T w = f_rcp(psi_i + nj.w);
        ux += w * (ni.x - nj.x);
        uy += w * (ni.y - nj.y);
        uz += w * (ni.z - nj.z);
      }
    });
    // per in-range neighbour: (x_j, m_j) and the normal record
    auto load = [&](int j) { return Rec2<T>{strec[j], normal[j]}; };
    T hx = 0, hy = 0, hz = 0;  // sum of m_j C(r) x_ij / r / (psi_i + psi_j)
    auto coh = [&](const Rec2<T>& d) {
      if (d.b.w > T(0)) {
        const V4<T>& q = d.a;
        const T dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
        const T r2 = dx * dx + dy * dy + dz * dz;
        const T rinv = f_rsqrt(r2);
        const T s = q.w * f_coh(k, r2 * rinv * k.inv_h) * rinv * f_rcp(psi_i + d.b.w);
        hx += s * dx;
        hy += s * dy;
        hz += s * dz;
      }
    };
    const int nf = nbr_nf[i];
    if (nf <= k.list_cap_f) {
      sweep_one_list<Rec2<T>>(cap, i, nbr_f, nf, load, coh);
    } else {
      walk_cells(k, fstart, c, [&](int j) {
        const V4<T> q = fpos[j];
        const T r2 = s_r2(p.x - q.x, p.y - q.y, p.z - q.z);
        if (r2 > T(0) && r2 <= k.r2max) coh(load(j));
      });
    }
    // 2 psi_0 / (psi_i + psi_j) (cohesion + curvature), cohesion =
    // -gamma m_j C(r) x_ij / r, curvature = gamma (n_i - n_j)
    const T f = T(2) * rp * k.gamma;
    a.x += f * (ux - hx);
    a.y += f * (uy - hy);
    a.z += f * (uz - hz);
    acce_adv[i] = a;
  }
  {
    TSPH_STRICT  // SPH_advection_update_vel_adv, sph.py:247-250
    const V4<T> v = vel[i];
    vadv[i] = V4<T>{v.x + a.x * dt, v.y + a.y * dt, v.z + a.z * dt, T(0)};
  }
}

// SPH_advection_update_vel_adv alone (sph.py:247-250), for gamma = 0: a
// separate kernel, so that it runs at the occupancy of an elementwise kernel
// rather than at that of the surface-tension walk.
template <class T>
__global__ void __launch_bounds__(kBlock)
    k_vel_adv_update(int n, const V4<T>* __restrict__ vel, const V4<T>* __restrict__ acce_adv,
                     V4<T>* __restrict__ vadv, const TsphControl* ctrl) {
  TSPH_STRICT
  const int i = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  const T dt = (T)ctrl->dt;
  const V4<T> a = acce_adv[i], v = vel[i];
  vadv[i] = V4<T>{v.x + a.x * dt, v.y + a.y * dt, v.z + a.z * dt, T(0)};
}

template <class T>
__global__ void __launch_bounds__(kBlock)
    k_surface_tension(KP<T> k, int n, int cap, const V4<T>* __restrict__ fpos,
                      const T* __restrict__ psi_rho,
                      const T* __restrict__ rdens, const T* __restrict__ vf,
                      const int* __restrict__ fstart, const V4<T>* __restrict__ normal,
                      const V4<T>* __restrict__ strec, const int* __restrict__ nbr_f,
                      const int* __restrict__ nbr_nf, const V4<T>* __restrict__ vel,
                      V4<T>* __restrict__ acce_adv, V4<T>* __restrict__ vadv,
                      const TsphControl* ctrl) {
  k_surface_tension_body<T>(Blk{(int)blockIdx.x, (int)gridDim.x}, k, n, cap, fpos, psi_rho, rdens, vf, fstart, normal, strec, nbr_f, nbr_nf, vel, acce_adv, vadv, ctrl);
}

// ---------------------------------------------------------------------------
// FBM round (sph.py:301-372), predicated on ctrl->general_flag.
// ---------------------------------------------------------------------------
template <class T>
__global__ void __launch_bounds__(kBlock)
    k_fbm_drift(KP<T> k, int n, const V4<T>* __restrict__ vadv, const V4<T>* __restrict__ vel,
                const T* __restrict__ rdens, const T* __restrict__ vf, V4<T>* __restrict__ acce_adv,
                V4<T>* __restrict__ acce, T* __restrict__ zeta, V4<T>* __restrict__ drift,
                T* __restrict__ tmp, const int* __restrict__ flag, TsphControl* ctrl) {
  TSPH_STRICT
  if (!ctrl->general_flag) return;
  if (blockIdx.x == 0 && threadIdx.x == 0) ctrl->fbm_rounds += 1;
  int i = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  const T dt = (T)ctrl->dt;
  const int P = k.P;
  for (int q = 0; q < P; ++q) tmp[(long)i * P + q] = T(0);  // SPH_FBM_clean_tmp
  V4<T> va = vadv[i], v = vel[i];
  T aa[3] = {(va.x - v.x) / dt, (va.y - v.y) / dt, (va.z - v.z) / dt};
  T rd = rdens[i];
  T z = T(0);
  for (int q = 0; q < P; ++q) z = z + vf[(long)i * P + q] * (k.rho[q] - rd) / k.rho[q];
  T fa[3];
  for (int d = 0; d < 3; ++d) fa[d] = (aa[d] - z * k.g[d]) / (T(1) - z);
  acce_adv[i] = V4<T>{aa[0], aa[1], aa[2], T(0)};
  acce[i] = V4<T>{fa[0], fa[1], fa[2], T(0)};
  zeta[i] = z;
  for (int q = 0; q < P; ++q) {
    T phi = vf[(long)i * P + q];
    T s = phi * (k.rho[q] - rd);
    T u[3];
    for (int d = 0; d < 3; ++d) u[d] = s * (k.g[d] - fa[d]);
    T wgt = phi * k.rho[q];
    if (wgt > T(1e-6)) {
      for (int d = 0; d < 3; ++d) u[d] = u[d] / wgt;
    } else {
      for (int d = 0; d < 3; ++d) u[d] = u[d] * T(0);
    }
    for (int d = 0; d < 3; ++d) u[d] = (u[d] + (fa[d] - aa[d])) * dt;
    if (k.dim == 2) u[2] = T(0);
    // the spare lane carries what the cooperative flux sweep gathers of this
    // particle besides its drift: phi_q, or -1 once the particle is flagged
    // (sph.py:366-372; the flags of this round are those of the last check)
    drift[(long)i * P + q] = V4<T>{u[0], u[1], u[2], flag[i] != 0 ? T(-1) : phi};
  }
}

template <class T>
__global__ void __launch_bounds__(kBlock)
    k_fbm_flux(KP<T> k, int n, int cap, const V4<T>* __restrict__ fpos,
               const T* __restrict__ rvol, const T* __restrict__ vf,
               const V4<T>* __restrict__ drift, const int* __restrict__ flag,
               const int* __restrict__ fstart, const int* __restrict__ nbr_f_in,
               const int* __restrict__ nbr_nf, T* __restrict__ tmp, const TsphControl* ctrl) {
  TSPH_STRICT
  if (!ctrl->general_flag) return;
  int i = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  if (i >= n || flag[i] != 0) return;
  const int* nbr_f = nbr_f_in;
  const T dt = (T)ctrl->dt;
  const int P = k.P;
  V4<T> p = fpos[i];
  T acc[kMaxP];
  for (int q = 0; q < P; ++q) acc[q] = tmp[(long)i * P + q];
  const T pre_c = k.fbmC * dt;
  const T pre_d = dt * k.fbmD;
  int nf = nbr_nf[i];
  // SPH_FBM_convect loop 2 (sph.py:346-363)
  for_fluid_nbrs(k, i, cap, nf, nbr_f, fstart, fpos, p, [&](int j) {
    if (flag[j] != 0) return;
    V4<T> q = fpos[j];
    T dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
    T r = s_sqrt(s_r2(dx, dy, dz));
    T wg = s_Wg(k, r);
    T ex = dx / r, ey = dy / r, ez = dz / r;
    T Vj = rvol[j];
    for (int s = 0; s < P; ++s) {
      V4<T> ui = drift[(long)i * P + s], uj = drift[(long)j * P + s];
      T pi = vf[(long)i * P + s], pj = vf[(long)j * P + s];
      T fx = pi * ui.x + pj * uj.x, fy = pi * ui.y + pj * uj.y, fz = pi * ui.z + pj * uj.z;
      T dot = fx * ex;
      dot = dot + fy * ey;
      dot = dot + fz * ez;
      T t = pre_c * Vj * dot * wg;
      acc[s] = acc[s] - t;
    }
  });
  // SPH_FBM_diffuse (sph.py:308-326)
  for_fluid_nbrs(k, i, cap, nf, nbr_f, fstart, fpos, p, [&](int j) {
    if (flag[j] != 0) return;
    V4<T> q = fpos[j];
    T dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
    T r2 = s_r2(dx, dy, dz);
    T r = s_sqrt(r2);
    T wg = s_Wg(k, r);
    T den = r * r + k.eps;
    T Vj = rvol[j];
    for (int s = 0; s < P; ++s) {
      T dphi = vf[(long)i * P + s] - vf[(long)j * P + s];
      T t = pre_d * dphi * Vj * r * wg / den;
      acc[s] = acc[s] + t;
    }
  });
  for (int q = 0; q < P; ++q) tmp[(long)i * P + q] = acc[q];
}

// The same sweep for NP phases as one cooperative list sweep (kCoop lanes
// per particle, gathers in flight as in the solver sweeps).  The pair terms
// of SPH_FBM_convect and SPH_FBM_diffuse share the neighbour's record and
// W'/r (fast arithmetic); their sums are added to tmp in the reference's
// order, convection first.  With W'(r) e = (W'/r) x and r W'(r) = r^2 W'/r,
// the terms are pre_c V_j (f . x) W'/r and pre_d dphi V_j r^2 (W'/r) / (r^2 + eps).
template <class T, int NP>
struct FbmRec {
  V4<T> x;
  V4<T> u[NP];  // drift of each phase, (u, phi) or (u, -1) when flagged (k_fbm_drift)
  T V;
};

template <class T, int NP>
__global__ void __launch_bounds__(kBlock)
    k_fbm_flux_coop(KP<T> k, int n, int cap, const V4<T>* __restrict__ fpos,
                    const T* __restrict__ rvol, const T* __restrict__ vf,
                    const V4<T>* __restrict__ drift, const int* __restrict__ flag,
                    const int* __restrict__ fstart, const int* __restrict__ nbr_f,
                    const int* __restrict__ nbr_nf, T* __restrict__ tmp,
                    const TsphControl* ctrl) {
  if (!ctrl->general_flag) return;
  const int base = xcd_block(blockIdx.x, gridDim.x) * kBlock;
  const int c = coop_lane<kCoop>();
  const T dt = (T)ctrl->dt;
#pragma unroll 1
  for (int round = 0; round < kCoop; ++round) {
    const int i = coop_particle<kCoop>(base, round);
    if (i >= n) break;
    if (flag[i] != 0) continue;  // the whole group: one particle
    const V4<T> p = fpos[i];
    V4<T> ui[NP];
    T pi[NP], conv[NP], diff[NP];
#pragma unroll
    for (int s = 0; s < NP; ++s) {
      ui[s] = drift[(long)i * NP + s];
      pi[s] = vf[(long)i * NP + s];
      conv[s] = T(0);
      diff[s] = T(0);
    }
    // VFSPH: the position's alias lane is the rest volume (no extra gather)
    const bool vx = k.solver == TSPH_VFSPH;
    auto load = [&](int j) {
      FbmRec<T, NP> d;
      d.x = fpos[j];
#pragma unroll
      for (int s = 0; s < NP; ++s) d.u[s] = drift[(long)j * NP + s];
      d.V = vx ? d.x.w : rvol[j];
      return d;
    };
    auto pair = [&](const FbmRec<T, NP>& d) {
      if (d.u[0].w < T(0)) return;  // a flagged neighbour
      const T dx = p.x - d.x.x, dy = p.y - d.x.y, dz = p.z - d.x.z;
      const T r2 = dx * dx + dy * dy + dz * dz;
      T r;
      const T g = f_grad(k, r2, &r);
      const T gv = g * d.V;
      const T gd = gv * r2 * f_rcp(r2 + k.eps);
#pragma unroll
      for (int s = 0; s < NP; ++s) {
        const T pj = d.u[s].w;
        const T fx = pi[s] * ui[s].x + pj * d.u[s].x;
        const T fy = pi[s] * ui[s].y + pj * d.u[s].y;
        const T fz = pi[s] * ui[s].z + pj * d.u[s].z;
        conv[s] += (fx * dx + fy * dy + fz * dz) * gv;
        diff[s] += (pi[s] - pj) * gd;
      }
    };
    sweep_lists_coop<kCoop, T, FbmRec<T, NP>>(
        k, cap, i, c, p, fstart, fstart, fpos, fpos, nbr_f, nullptr, nbr_nf[i], 0, load, load,
        pair, [&](const FbmRec<T, NP>&) {});
#pragma unroll
    for (int s = 0; s < NP; ++s) {
      conv[s] = group_sum<kCoop>(conv[s]);
      diff[s] = group_sum<kCoop>(diff[s]);
    }
    if (c != 0) continue;
    const T pre_c = k.fbmC * dt, pre_d = dt * k.fbmD;
#pragma unroll
    for (int s = 0; s < NP; ++s) {
      T a = tmp[(long)i * NP + s];
      a = a - pre_c * conv[s];
      a = a + pre_d * diff[s];
      tmp[(long)i * NP + s] = a;
    }
  }
}

template <class T>
__global__ void __launch_bounds__(kBlock)
    k_fbm_check(KP<T> k, int n, const T* __restrict__ vf, const T* __restrict__ tmp,
                int* __restrict__ flag, double* __restrict__ partial, TsphControl* ctrl) {
  TSPH_STRICT
  __shared__ double lds[kWaves];
  if (!ctrl->general_flag) return;
  int nblk = gridDim.x;
  int i = xcd_block(blockIdx.x, nblk) * kBlock + threadIdx.x;
  double any = 0;
  if (i < n) {
    bool neg = false;  // has_negative, sph_util.py:56-62
    for (int q = 0; q < k.P; ++q)
      if (vf[(long)i * k.P + q] + tmp[(long)i * k.P + q] < T(0)) neg = true;
    if (neg) {
      flag[i] = 1;
      any = 1;
    }
  }
  double s = block_sum(any, lds);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// general_flag = any particle flagged this round (sph.py:368-372); skipped,
// like the round itself, once the flag is clear.
__global__ void __launch_bounds__(kBlock)
    k_fbm_check_finish(const double* __restrict__ partial, int nblk, TsphControl* ctrl) {
  __shared__ double lds[kWaves];
  if (!ctrl->general_flag) return;
  double a = sum_partials(partial, nblk, lds);
  if (threadIdx.x == 0) ctrl->general_flag = a > 0 ? 1 : 0;
}

// ---------------------------------------------------------------------------
// Tail: [SPH_update_volume_frac], SPH_update_mass, SPH_update_pos,
// SPH_update_energy (sph.py:265-286, 375-379).
// ---------------------------------------------------------------------------
template <class T>
__device__ __forceinline__ void k_tail_body(Blk blk, const KP<T>& k, int n, int fbm, V4<T>* __restrict__ pos, V4<T>* __restrict__ vel,
           const V4<T>* __restrict__ vadv, T* __restrict__ mass, T* __restrict__ rdens,
           const T* __restrict__ rvol, T* __restrict__ vf, const T* __restrict__ tmp,
           const int* __restrict__ flag, double* __restrict__ partial, TsphControl* ctrl) {
  TSPH_STRICT
  __shared__ double lds[kWaves];
  int nblk = blk.nb;
  int i = xcd_block(blk.b, nblk) * kBlock + threadIdx.x;
  const T dt = (T)ctrl->dt;
  double ke = 0, pe = 0;
  if (i < n) {
    const int P = k.P;
    if (fbm && !(flag[i] > 0))
      for (int q = 0; q < P; ++q) vf[(long)i * P + q] = vf[(long)i * P + q] + tmp[(long)i * P + q];
    T rd = k.rho[0] * vf[(long)i * P];
    for (int q = 1; q < P; ++q) rd = rd + k.rho[q] * vf[(long)i * P + q];
    T m = rd * rvol[i];
    rdens[i] = rd;
    mass[i] = m;
    V4<T> v = vadv[i];
    V4<T> p = pos[i];
    p.x = p.x + v.x * dt;
    p.y = p.y + v.y * dt;
    p.z = p.z + v.z * dt;
    p.w = k.solver == TSPH_VFSPH ? rvol[i] : m;
    pos[i] = p;
    vel[i] = V4<T>{v.x, v.y, v.z, T(0)};
    T n2 = s_r2(v.x, v.y, v.z);
    ke = (double)(T(0.5) * m * n2);
    pe = (double)(-m * k.g[1] * (p.y - k.lb[1]));
  }
  double s1 = block_sum(ke, lds);
  double s2 = block_sum(pe, lds);
  if (threadIdx.x == 0) {
    partial[blk.b] = s1;
    partial[nblk + blk.b] = s2;
  }
}

template <class T>
__global__ void __launch_bounds__(kBlock)
    k_tail(KP<T> k, int n, int fbm, V4<T>* __restrict__ pos, V4<T>* __restrict__ vel,
           const V4<T>* __restrict__ vadv, T* __restrict__ mass, T* __restrict__ rdens,
           const T* __restrict__ rvol, T* __restrict__ vf, const T* __restrict__ tmp,
           const int* __restrict__ flag, double* __restrict__ partial, TsphControl* ctrl) {
  k_tail_body<T>(Blk{(int)blockIdx.x, (int)gridDim.x}, k, n, fbm, pos, vel, vadv, mass, rdens, rvol, vf, tmp, flag, partial, ctrl);
}

// ---------------------------------------------------------------------------
// Host launchers.
// ---------------------------------------------------------------------------
template <class T>
static int pass_a(const TsphParams& prm, TsphObject* f, TsphObject* b, TsphControl* ctrl,
                  bool with_bound, const WsLayout& L, hipStream_t st) {
  KP<T> k = make_kp<T>(prm);
  hipLaunchKernelGGL(k_step_reset, dim3(1), dim3(64), 0, st, ctrl);
  if (with_bound && f->n > 0 && b->n > 0) {
    const int nbf = blocks_for(f->n);
    hipLaunchKernelGGL(k_pass_a_both<T>, dim3(nbf + blocks_for(b->n)), dim3(kBlock), 0, st, k, *f,
                       *b, nbf, (V4<T>*)L.a1b, ctrl);
    TSPH_LAUNCH_CHECK("density_alpha");
    if (b->static_sums) b->flags |= TSPH_OBJ_STATIC_SUMS;  // stream order: later launches read them
    return 0;
  }
  if (f->n > 0)
    hipLaunchKernelGGL(k_pass_a_fluid<T>, dim3(blocks_for(f->n)), dim3(kBlock), 0, st, k, f->n,
                       f->capacity, (const V4<T>*)f->pos, (const T*)f->mass,
                       (const T*)f->rest_volume, f->cell_start, b->n, (const V4<T>*)b->pos,
                       (const T*)b->mass, (const T*)b->rest_volume, b->cell_start, (T*)f->W,
                       (T*)f->sph_compression, (T*)f->sph_density, (V4<T>*)f->alpha_1,
                       (T*)f->alpha_2, (T*)f->alpha, (T*)f->vol, (const V4<T>*)f->vel,
                       (V4<T>*)f->vel_adv, f->nbr_f, f->nbr_b, f->nbr_nf, f->nbr_nb, f->flag,
                       (V4<T>*)L.a1b, ctrl);
  if (with_bound && b->n > 0)
    hipLaunchKernelGGL(k_pass_a_bound<T>, dim3(blocks_for(b->n)), dim3(kBlock), 0, st, k, b->n,
                       (const V4<T>*)b->pos, (const T*)b->mass, (const T*)b->rest_volume,
                       b->cell_start, (const V4<T>*)f->pos, (const T*)f->mass,
                       (const T*)f->rest_volume, f->cell_start, (T*)b->W, (T*)b->sph_compression,
                       (T*)b->sph_density, (V4<T>*)b->alpha_1, (T*)b->alpha_2, (T*)b->alpha,
                       (T*)b->vol, (V4<T>*)b->static_sums, b->flags & TSPH_OBJ_STATIC_SUMS);
  TSPH_LAUNCH_CHECK("density_alpha");
  if (with_bound && b->n > 0 && b->static_sums) b->flags |= TSPH_OBJ_STATIC_SUMS;
  return 0;
}

template <class T>
int density_alpha(const TsphParams& prm, TsphObject* f, TsphObject* b, TsphControl* ctrl,
                  void* ws, size_t wsb, hipStream_t st) {
  WsLayout L;
  TSPH_REQUIRE(carve_ws(ws, wsb, f->capacity, prm.node_num, &L), TSPH_ECAPACITY,
               "tsph_density_alpha: workspace too small");
  return pass_a<T>(prm, f, b, ctrl, true, L, st);
}

// The fluid's neighbour lists alone, as pass A builds them (8 candidates in
// flight: without pass A's pair math the walk is bound by the gathers): in-range
// candidates with r > 0 (r^2 > 0) in walk order.  The counters are updated
// without branches (per-branch increments get merged into an indexed
// increment of a stack array).
constexpr int kListsU = 8;
template <class T>
__global__ void __launch_bounds__(kBlock)
    k_lists(KP<T> k, int n, int cap, const V4<T>* __restrict__ fpos, const int* __restrict__ fstart,
            const V4<T>* __restrict__ bpos, const int* __restrict__ bstart, int* __restrict__ nbr_f,
            int* __restrict__ nbr_b, int* __restrict__ nbr_nf, int* __restrict__ nbr_nb,
            TsphControl* ctrl) {
  const int i = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  const int capf = k.list_cap_f, capb = k.list_cap_b;
  const V4<T> p = fpos[i];
  int c[3];
  node_encode(k, p.x, p.y, p.z, c);
  int nf = 0, nbb = 0;
  walk_cells_batched<kListsU>(k, fstart, c, fpos, [&](int j, const V4<T>& q) {
    const T r2 = s_r2(p.x - q.x, p.y - q.y, p.z - q.z);
    const bool take = r2 <= k.r2max && r2 > T(0);
    if (take && nf < capf) nbr_f[(long)nf * cap + i] = j;
    nf += take ? 1 : 0;
  });
  walk_cells_batched<kListsU>(k, bstart, c, bpos, [&](int j, const V4<T>& q) {
    const T r2 = s_r2(p.x - q.x, p.y - q.y, p.z - q.z);
    const bool take = r2 <= k.r2max && r2 > T(0);
    if (take && nbb < capb) nbr_b[(long)nbb * cap + i] = j;
    nbb += take ? 1 : 0;
  });
  if (nf > capf) {
    atomicAdd(&ctrl->overflow_f, 1);
    nf = capf + 1;
  }
  if (nbb > capb) {
    atomicAdd(&ctrl->overflow_b, 1);
    nbb = capb + 1;
  }
  nbr_nf[i] = nf;
  nbr_nb[i] = nbb;
}

// Rebuilds the fluid's neighbour lists from recorded sorted positions and
// cell table.  The adjoint recomputes per-step lists this way when it does
// not store them (it stores the step's densities and alpha factors).
template <class T>
int relist(const TsphParams& prm, TsphObject* f, TsphObject* b, TsphControl* ctrl, void* ws,
           size_t wsb, hipStream_t st) {
  (void)ws;
  (void)wsb;
  KP<T> k = make_kp<T>(prm);
  if (f->n == 0) return 0;
  hipLaunchKernelGGL(k_lists<T>, dim3(blocks_for(f->n)), dim3(kBlock), 0, st, k, f->n, f->capacity,
                     (const V4<T>*)f->pos, f->cell_start, (const V4<T>*)b->pos, b->cell_start,
                     f->nbr_f, f->nbr_b, f->nbr_nf, f->nbr_nb, ctrl);
  TSPH_LAUNCH_CHECK("relist");
  return 0;
}

template <class T>
int solver_begin(const TsphParams& prm, TsphControl* ctrl, int solve, int* signal, int token,
                 hipStream_t st) {
  KP<T> k = make_kp<T>(prm);
  hipLaunchKernelGGL(k_solver_begin<T>, dim3(1), dim3(kBlock), 0, st, k, ctrl, solve, signal,
                     token);
  TSPH_LAUNCH_CHECK("solver_begin");
  return 0;
}

template <class T>
int solver_half(const TsphParams& prm, TsphObject* f, TsphObject* b, TsphControl* ctrl, int kk,
                int half, void* tape_v, void* tape_k, const void* clamp_ref, void* ws, size_t wsb,
                hipStream_t st) {
  KP<T> k = make_kp<T>(prm);
  WsLayout L;
  TSPH_REQUIRE(carve_ws(ws, wsb, f->capacity, prm.node_num, &L), TSPH_ECAPACITY,
               "tsph_solver_iteration: workspace too small");
  // With no fluid particles one block still runs the stop rule: the mean of
  // an empty sum is 0/0 = NaN, as in the reference (sph.py:221-226), so the
  // solve runs its minimum iteration count.
  int nb = blocks_for(f->n > 0 ? f->n : 1);
  const T* psi = (const T*)(prm.solver_type == TSPH_VFSPH ? f->sph_compression : f->sph_density);
  if (half & 1)
    hipLaunchKernelGGL(k_solver_psi<T>, dim3(nb), dim3(kBlock), 0, st, k, f->n, f->capacity, kk,
                       (const V4<T>*)f->pos, (const V4<T>*)f->vel_adv, psi,
                       (const T*)f->rest_density, (const T*)f->alpha, f->cell_start,
                       (const V4<T>*)b->pos, b->cell_start, f->nbr_f, f->nbr_b, f->nbr_nf,
                       f->nbr_nb, (T*)f->psi_adv, (T*)f->kappa, (V4<T>*)L.krec,
                       (const V4<T>*)L.a1b, (V4<T>*)tape_v, (T*)tape_k, (const T*)clamp_ref,
                       L.partial, ctrl);
  if (half & 2)
    hipLaunchKernelGGL(k_solver_update<T>, dim3(nb), dim3(kBlock), 0, st, k, f->n, f->capacity,
                       kk, (const V4<T>*)f->pos, (const T*)f->mass, (const T*)f->kappa,
                       (const V4<T>*)L.krec, (const V4<T>*)f->alpha_1, f->cell_start, (const V4<T>*)b->pos, b->cell_start, f->nbr_f, f->nbr_b,
                       f->nbr_nf, f->nbr_nb, (V4<T>*)f->vel_adv, L.partial, ctrl);
  TSPH_LAUNCH_CHECK("solver_iteration");
  return 0;
}

template <class T>
int advect(const TsphParams& prm, TsphObject* f, TsphObject* b, TsphControl* ctrl, void* ws,
           size_t wsb, hipStream_t st) {
  KP<T> k = make_kp<T>(prm);
  WsLayout L;
  TSPH_REQUIRE(carve_ws(ws, wsb, f->capacity, prm.node_num, &L), TSPH_ECAPACITY,
               "tsph_advect: workspace too small");
  if (f->n == 0) return 0;
  int nb = blocks_for(f->n);
  const T* psi = (const T*)(prm.solver_type == TSPH_VFSPH ? f->sph_compression : f->sph_density);
  hipLaunchKernelGGL(k_advect<T>, dim3(nb), dim3(kBlock), 0, st, k, f->n, f->capacity,
                     (const V4<T>*)f->pos, (const T*)f->volume_frac, (const T*)f->rest_density,
                     f->cell_start, b->cell_start, (const V4<T>*)b->pos, f->nbr_f, f->nbr_nf,
                     (V4<T>*)f->vel_adv, (V4<T>*)f->vel, (V4<T>*)f->acce_adv,
                     (V4<T>*)f->normal, (const T*)f->mass, psi, (V4<T>*)L.krec);
  if (k.gamma == T(0))
    hipLaunchKernelGGL(k_vel_adv_update<T>, dim3(nb), dim3(kBlock), 0, st, f->n,
                       (const V4<T>*)f->vel, (const V4<T>*)f->acce_adv, (V4<T>*)f->vel_adv, ctrl);
  else
    hipLaunchKernelGGL(k_surface_tension<T>, dim3(nb), dim3(kBlock), 0, st, k, f->n,
                       f->capacity, (const V4<T>*)f->pos, psi, (const T*)f->rest_density,
                       (const T*)f->volume_frac, f->cell_start, (const V4<T>*)f->normal,
                       (const V4<T>*)L.krec, f->nbr_f, f->nbr_nf, (const V4<T>*)f->vel,
                       (V4<T>*)f->acce_adv, (V4<T>*)f->vel_adv, ctrl);
  TSPH_LAUNCH_CHECK("advect");
  return 0;
}

template <class T>
int fbm_round(const TsphParams& prm, TsphObject* f, TsphControl* ctrl, void* ws, size_t wsb,
              hipStream_t st) {
  KP<T> k = make_kp<T>(prm);
  WsLayout L;
  TSPH_REQUIRE(carve_ws(ws, wsb, f->capacity, prm.node_num, &L), TSPH_ECAPACITY,
               "tsph_fbm_round: workspace too small");
  if (f->n == 0) return 0;
  int nb = blocks_for(f->n);
  hipLaunchKernelGGL(k_fbm_drift<T>, dim3(nb), dim3(kBlock), 0, st, k, f->n,
                     (const V4<T>*)f->vel_adv, (const V4<T>*)f->vel, (const T*)f->rest_density,
                     (const T*)f->volume_frac, (V4<T>*)f->acce_adv, (V4<T>*)f->acce,
                     (T*)f->fbm_zeta, (V4<T>*)f->drift_vel, (T*)f->volume_frac_tmp, f->flag, ctrl);
  auto flux = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3(nb), dim3(kBlock), 0, st, k, f->n, f->capacity,
                       (const V4<T>*)f->pos, (const T*)f->rest_volume, (const T*)f->volume_frac,
                       (const V4<T>*)f->drift_vel, f->flag, f->cell_start, f->nbr_f, f->nbr_nf,
                       (T*)f->volume_frac_tmp, ctrl);
  };
  switch (prm.phase_num) {  // the cooperative sweep keeps NP phases' records in registers
    case 1: flux(k_fbm_flux_coop<T, 1>); break;
    case 2: flux(k_fbm_flux_coop<T, 2>); break;
    case 3: flux(k_fbm_flux_coop<T, 3>); break;
    default: flux(k_fbm_flux<T>); break;
  }
  hipLaunchKernelGGL(k_fbm_check<T>, dim3(nb), dim3(kBlock), 0, st, k, f->n,
                     (const T*)f->volume_frac, (const T*)f->volume_frac_tmp, f->flag, L.partial,
                     ctrl);
  hipLaunchKernelGGL(k_fbm_check_finish, dim3(1), dim3(kBlock), 0, st, L.partial, nb, ctrl);
  TSPH_LAUNCH_CHECK("fbm_round");
  return 0;
}

template <class T>
int tail(const TsphParams& prm, TsphObject* f, TsphControl* ctrl, int fbm, void* ws, size_t wsb,
         hipStream_t st) {
  KP<T> k = make_kp<T>(prm);
  WsLayout L;
  TSPH_REQUIRE(carve_ws(ws, wsb, f->capacity, prm.node_num, &L), TSPH_ECAPACITY,
               "tsph_tail: workspace too small");
  int nb = blocks_for(f->n > 0 ? f->n : 1);
  hipLaunchKernelGGL(k_tail<T>, dim3(nb), dim3(kBlock), 0, st, k, f->n, fbm, (V4<T>*)f->pos,
                     (V4<T>*)f->vel, (const V4<T>*)f->vel_adv, (T*)f->mass,
                     (T*)f->rest_density, (const T*)f->rest_volume, (T*)f->volume_frac,
                     (const T*)f->volume_frac_tmp, f->flag, L.partial, ctrl);
  hipLaunchKernelGGL(k_finalize<T>, dim3(1), dim3(kBlock), 0, st, k, (int)kFinEnergy, L.partial, nb,
                     f->n, ctrl);
  TSPH_LAUNCH_CHECK("tail");
  return 0;
}

// ---------------------------------------------------------------------------
// WCSPH block (main.py:236-240, commented in the reference's sph_step):
// WC_pressure_val (fluid), WC_pressure_acce (f,f)(f,b) and
// SPH_advection_update_vel_adv.  Boundary pressure is never set by the
// reference (WC_pressure_val runs for the fluid only), so it reads 0.
// Strict arithmetic, the reference's order.
// ---------------------------------------------------------------------------
template <class T>
__global__ void __launch_bounds__(kBlock)
    k_wc_pressure(KP<T> k, int n, const T* __restrict__ dens, const T* __restrict__ rdens,
                  T* __restrict__ pressure) {
  TSPH_STRICT
  int i = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  // (rho0 * cs^2 / gamma) * ((rho / rho0)^7 - 1), clamped at 0 (sph.py:169-175);
  // the integer power by repeated squaring
  T cs2 = (T)(k.cs * k.cs);
  T ratio = dens[i] / rdens[i];
  T res = T(1), b = ratio;
  int e = 7;
  while (e > 0) {
    if (e & 1) res = res * b;
    b = b * b;
    e >>= 1;
  }
  T pr = (rdens[i] * cs2 / (T)k.wc_gamma) * (res - T(1));
  if (pr < T(0)) pr = T(0);
  pressure[i] = pr;
}

template <class T>
__global__ void __launch_bounds__(kBlock)
    k_wc_acce(KP<T> k, int n, int cap, const V4<T>* __restrict__ fpos, const T* __restrict__ fmass,
              const T* __restrict__ fdens, const T* __restrict__ fpress, const int* __restrict__ fstart,
              const V4<T>* __restrict__ bpos, const T* __restrict__ bmass,
              const T* __restrict__ bdens, const T* __restrict__ bpress,
              const int* __restrict__ bstart, const int* __restrict__ nbr_f,
              const int* __restrict__ nbr_b, const int* __restrict__ nbr_nf,
              const int* __restrict__ nbr_nb, const V4<T>* __restrict__ vel,
              V4<T>* __restrict__ acce_adv, V4<T>* __restrict__ vel_adv,
              const TsphControl* ctrl) {
  TSPH_STRICT
  int i = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  const V4<T> p = fpos[i];
  V4<T> a = acce_adv[i];
  const T pi = fpress[i], di = fdens[i];
  auto term = [&](const V4<T>& q, T mj, T dj, T pj) {
    T dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
    T r = s_sqrt(s_r2(dx, dy, dz));
    T pt = pi / (di * di) + pj / (dj * dj);
    T wg = s_Wg(k, r);
    a.x = a.x + -pt * mj * dx / r * wg;
    a.y = a.y + -pt * mj * dy / r * wg;
    if (k.dim == 3) a.z = a.z + -pt * mj * dz / r * wg;
  };
  for_fluid_nbrs(k, i, cap, nbr_nf[i], nbr_f, fstart, fpos, p,
                 [&](int j) { term(fpos[j], fmass[j], fdens[j], fpress[j]); });
  for_bound_nbrs(k, i, cap, nbr_nb[i], nbr_b, bstart, bpos, p,
                 [&](int j) { term(bpos[j], bmass[j], bdens[j], bpress[j]); });
  acce_adv[i] = a;
  const T dt = (T)ctrl->dt;
  const V4<T> v = vel[i];
  vel_adv[i] = V4<T>{v.x + dt * a.x, v.y + dt * a.y, v.z + dt * a.z, T(0)};
}

template <class T>
int wcsph(const TsphParams& prm, TsphObject* f, TsphObject* b, TsphControl* ctrl,
          hipStream_t st) {
  KP<T> k = make_kp<T>(prm);
  if (f->n == 0) return 0;
  int nb = blocks_for(f->n);
  hipLaunchKernelGGL(k_wc_pressure<T>, dim3(nb), dim3(kBlock), 0, st, k, f->n,
                     (const T*)f->sph_density, (const T*)f->rest_density, (T*)f->pressure);
  hipLaunchKernelGGL(k_wc_acce<T>, dim3(nb), dim3(kBlock), 0, st, k, f->n, f->capacity,
                     (const V4<T>*)f->pos, (const T*)f->mass, (const T*)f->sph_density,
                     (const T*)f->pressure, f->cell_start, (const V4<T>*)b->pos,
                     (const T*)b->mass, (const T*)b->sph_density, (const T*)b->pressure,
                     b->cell_start, f->nbr_f, f->nbr_b, f->nbr_nf, f->nbr_nb,
                     (const V4<T>*)f->vel, (V4<T>*)f->acce_adv, (V4<T>*)f->vel_adv, ctrl);
  TSPH_LAUNCH_CHECK("wcsph");
  return 0;
}
template int wcsph<float>(const TsphParams&, TsphObject*, TsphObject*, TsphControl*, hipStream_t);
template int wcsph<double>(const TsphParams&, TsphObject*, TsphObject*, TsphControl*, hipStream_t);

// ---------------------------------------------------------------------------
// map_velocity (sph.py:382-397) over a Grid (sph_obj.py:328-344): one thread
// per grid node (row-major, last axis fastest) at lb + I * dist; the node's
// velocity sums X_j / psi_j * vel_j * W(|x_I - x_j|) over the object's
// candidates in the cells around the node's cell, in template and in-cell
// order (the object's cell table from this step's sort).  Strict arithmetic.
// ---------------------------------------------------------------------------
template <class T>
__global__ void __launch_bounds__(kBlock)
    k_map_velocity(KP<T> k, long total, int g0, int g1, int g2, T dist,
                   const V4<T>* __restrict__ pos, const V4<T>* __restrict__ vel,
                   const T* __restrict__ psi, const int* __restrict__ cstart,
                   V4<T>* __restrict__ out) {
  TSPH_STRICT
  const long idx = (long)blockIdx.x * kBlock + threadIdx.x;
  if (idx >= total) return;
  int I[3] = {0, 0, 0};
  if (k.dim == 3) {
    I[2] = (int)(idx % g2);
    long r = idx / g2;
    I[1] = (int)(r % g1);
    I[0] = (int)(r / g1);
  } else {
    I[1] = (int)(idx % g1);
    I[0] = (int)(idx / g1);
  }
  T gp[3] = {T(0), T(0), T(0)};
  for (int d = 0; d < k.dim; ++d) gp[d] = k.lb[d] + T(I[d]) * dist;
  int c[3];
  node_encode(k, gp[0], gp[1], gp[2], c);
  T vx = 0, vy = 0, vz = 0;
  walk_cells(k, cstart, c, [&](int j) {
    const V4<T> q = pos[j];
    const T w = s_W(k, s_sqrt(s_r2(gp[0] - q.x, gp[1] - q.y, gp[2] - q.z)));
    const T sc = q.w / psi[j];
    const V4<T> v = vel[j];
    vx = vx + (sc * v.x) * w;
    vy = vy + (sc * v.y) * w;
    vz = vz + (sc * v.z) * w;
  });
  out[idx] = V4<T>{vx, vy, vz, T(0)};
}

template <class T>
int map_velocity(const TsphParams& prm, const TsphObject* o, const int* shape, double dist,
                 void* out, hipStream_t st) {
  KP<T> k = make_kp<T>(prm);
  long total = 1;
  for (int d = 0; d < prm.dim; ++d) total *= shape[d];
  if (total <= 0) return 0;
  const T* psi = (const T*)(prm.solver_type == TSPH_VFSPH ? o->sph_compression : o->sph_density);
  hipLaunchKernelGGL(k_map_velocity<T>, dim3((unsigned)((total + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                     st, k, total, shape[0], shape[1], prm.dim == 3 ? shape[2] : 1, (T)dist,
                     (const V4<T>*)o->pos, (const V4<T>*)o->vel, psi, o->cell_start, (V4<T>*)out);
  TSPH_LAUNCH_CHECK("map_velocity");
  return 0;
}
template int map_velocity<float>(const TsphParams&, const TsphObject*, const int*, double, void*, hipStream_t);
template int map_velocity<double>(const TsphParams&, const TsphObject*, const int*, double, void*, hipStream_t);

// ---------------------------------------------------------------------------
// Grouped (batched) launches: the kernels above over every scene of a table
// (tsph_host.h:SceneEntry), with the single-scene launchers' arguments.
// ---------------------------------------------------------------------------
template <class T>
__device__ __forceinline__ const T* psi_of(const SceneEntry<T>& e) {
  return (const T*)(e.k.solver == TSPH_VFSPH ? e.f.sph_compression : e.f.sph_density);
}

template <class T>
__global__ void __launch_bounds__(kBlock) kb_step_reset(const char* __restrict__ tab, int ns) {
  for_tiles<T, kTileScene>(tab, ns, [&](const SceneEntry<T>& e, int, Blk) { step_reset_body(e.ctrl); });
}

// After the grouped boundary pass A: every scene whose boundary has a kept-sum
// buffer now holds its (b,b) sums, so later steps read them (the table's copy
// of TSPH_OBJ_STATIC_SUMS, as tsph_density_alpha sets it in the caller's
// descriptor).  One thread per scene, a vector store into the table; the
// grouped kernels read the table at their next launch.
template <class T>
__global__ void kb_mark_static_sums(char* tab, int ns) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ns) return;
  TsphObject& b = ((SceneEntry<T>*)tab)[s].b;
  if (b.static_sums != nullptr && b.n > 0) b.flags |= TSPH_OBJ_STATIC_SUMS;
}

// Pass A of one fluid tile / one boundary tile of a scene.
template <class T>
__device__ __forceinline__ void pass_a_fluid_tile(const SceneEntry<T>& e, Blk blk, int* stage) {
  const TsphObject& f = e.f;
  const TsphObject& b = e.b;
  pass_a_fluid_variant(e.k, [&](auto fq, auto xm) {
    k_pass_a_fluid_body<T, decltype(fq)::value, decltype(xm)::value>(
        blk, e.k, f.n, f.capacity, (const V4<T>*)f.pos, (const T*)f.mass,
        (const T*)f.rest_volume, f.cell_start, b.n, (const V4<T>*)b.pos, (const T*)b.mass,
        (const T*)b.rest_volume, b.cell_start, (T*)f.W, (T*)f.sph_compression,
        (T*)f.sph_density, (V4<T>*)f.alpha_1, (T*)f.alpha_2, (T*)f.alpha, (T*)f.vol,
        (const V4<T>*)f.vel, (V4<T>*)f.vel_adv, f.nbr_f, f.nbr_b, f.nbr_nf, f.nbr_nb, f.flag,
        (V4<T>*)e.L.a1b, e.ctrl, stage);
  });
}
template <class T>
__device__ __forceinline__ void pass_a_bound_tile(const SceneEntry<T>& e, Blk blk) {
  const TsphObject& f = e.f;
  const TsphObject& b = e.b;
  k_pass_a_bound_body<T, false>(
      blk, e.k, b.n, (const V4<T>*)b.pos, (const T*)b.mass, (const T*)b.rest_volume,
      b.cell_start, (const V4<T>*)f.pos, (const T*)f.mass, (const T*)f.rest_volume,
      f.cell_start, (T*)b.W, (T*)b.sph_compression, (T*)b.sph_density, (V4<T>*)b.alpha_1,
      (T*)b.alpha_2, (T*)b.alpha, (T*)b.vol, (V4<T>*)b.static_sums,
      (b.flags & TSPH_OBJ_STATIC_SUMS) != 0);
}

// Pass A of every scene's fluid and boundary in one grouped launch: tiles
// [0, fluid tiles) are fluid tiles, the rest boundary tiles.  The two
// halves write only their own object's fields, so the boundary tiles fill
// the tail of the fluid tiles' launch instead of waiting for it.
template <class T>
__global__ void __launch_bounds__(kBlock) kb_pass_a_both(const char* __restrict__ tab, int ns) {
  __shared__ int stage[kPassAK * kBlock];
  const int tf = tile_prefix<T, kTileFluid>(tab, ns)[ns];
  const int total = tf + tile_prefix<T, kTileBound>(tab, ns)[ns];
  for (int t = blockIdx.x; t < total; t += gridDim.x) {
    if (t < tf)
      on_tile<T, kTileFluid>(tab, ns, t, [&](const SceneEntry<T>& e, int, Blk blk) {
        pass_a_fluid_tile(e, blk, stage);
      });
    else
      on_tile<T, kTileBound>(tab, ns, t - tf, [&](const SceneEntry<T>& e, int, Blk blk) {
        pass_a_bound_tile(e, blk);
      });
  }
}

template <class T>
__global__ void __launch_bounds__(kBlock)
    kb_solver_begin(const char* __restrict__ tab, int ns, int solve, uint64_t* signal, int token) {
  // scene s's early stop signal goes to signal[s] (no signal when null)
  for_tiles<T, kTileScene>(tab, ns, [&](const SceneEntry<T>& e, int s, Blk) {
    solver_begin_body(e.k, e.ctrl, solve, signal ? (int*)(signal + s) : nullptr, token);
  });
}

template <class T>
__global__ void __launch_bounds__(kBlock) kb_solver_psi(const char* __restrict__ tab, int ns, int kk) {
  for_tiles<T, kTileFluid>(tab, ns, [&](const SceneEntry<T>& e, int, Blk blk) {
    const TsphObject& f = e.f;
    k_solver_psi_body<T>(blk, e.k, f.n, f.capacity, kk, (const V4<T>*)f.pos,
                         (const V4<T>*)f.vel_adv, psi_of(e), (const T*)f.rest_density,
                         (const T*)f.alpha, f.cell_start, (const V4<T>*)e.b.pos, e.b.cell_start,
                         f.nbr_f, f.nbr_b, f.nbr_nf, f.nbr_nb, (T*)f.psi_adv, (T*)f.kappa,
                         (V4<T>*)e.L.krec, (const V4<T>*)e.L.a1b, (V4<T>*)nullptr, (T*)nullptr,
                         (const T*)nullptr, e.L.partial, e.ctrl);
  });
}

template <class T>
__global__ void __launch_bounds__(kBlock) kb_solver_update(const char* __restrict__ tab, int ns, int kk) {
  for_tiles<T, kTileFluid>(tab, ns, [&](const SceneEntry<T>& e, int, Blk blk) {
    const TsphObject& f = e.f;
    k_solver_update_body<T>(blk, e.k, f.n, f.capacity, kk, (const V4<T>*)f.pos,
                            (const T*)f.mass, (const T*)f.kappa, (const V4<T>*)e.L.krec,
                            (const V4<T>*)f.alpha_1, f.cell_start,
                            (const V4<T>*)e.b.pos, e.b.cell_start, f.nbr_f, f.nbr_b, f.nbr_nf,
                            f.nbr_nb, (V4<T>*)f.vel_adv, e.L.partial, e.ctrl);
  });
}

template <class T>
__global__ void __launch_bounds__(kBlock) kb_advect(const char* __restrict__ tab, int ns) {
  for_tiles<T, kTileFluid>(tab, ns, [&](const SceneEntry<T>& e, int, Blk blk) {
    const TsphObject& f = e.f;
    k_advect_body<T>(blk, e.k, f.n, f.capacity, (const V4<T>*)f.pos, (const T*)f.volume_frac,
                     (const T*)f.rest_density, f.cell_start, e.b.cell_start,
                     (const V4<T>*)e.b.pos, f.nbr_f, f.nbr_nf, (V4<T>*)f.vel_adv, (V4<T>*)f.vel,
                     (V4<T>*)f.acce_adv, (V4<T>*)f.normal, (const T*)f.mass, psi_of(e),
                     (V4<T>*)e.L.krec);
  });
}

template <class T>
__global__ void __launch_bounds__(kBlock) kb_surface_tension(const char* __restrict__ tab, int ns) {
  for_tiles<T, kTileFluid>(tab, ns, [&](const SceneEntry<T>& e, int, Blk blk) {
    const TsphObject& f = e.f;
    k_surface_tension_body<T>(blk, e.k, f.n, f.capacity, (const V4<T>*)f.pos, psi_of(e),
                              (const T*)f.rest_density, (const T*)f.volume_frac, f.cell_start,
                              (const V4<T>*)f.normal, (const V4<T>*)e.L.krec, f.nbr_f, f.nbr_nf,
                              (const V4<T>*)f.vel, (V4<T>*)f.acce_adv, (V4<T>*)f.vel_adv, e.ctrl);
  });
}

template <class T>
__global__ void __launch_bounds__(kBlock) kb_tail(const char* __restrict__ tab, int ns) {
  for_tiles<T, kTileFluid>(tab, ns, [&](const SceneEntry<T>& e, int, Blk blk) {
    const TsphObject& f = e.f;
    k_tail_body<T>(blk, e.k, f.n, 0, (V4<T>*)f.pos, (V4<T>*)f.vel, (const V4<T>*)f.vel_adv,
                   (T*)f.mass, (T*)f.rest_density, (const T*)f.rest_volume, (T*)f.volume_frac,
                   (const T*)f.volume_frac_tmp, f.flag, e.L.partial, e.ctrl);
  });
}

template <class T>
__global__ void __launch_bounds__(kBlock) kb_finalize_energy(const char* __restrict__ tab, int ns) {
  const int* firstf = (const int*)(tab + batch_entries_bytes<T>(ns)) + kTileFluid * (ns + 1);
  for_tiles<T, kTileScene>(tab, ns, [&](const SceneEntry<T>& e, int s, Blk blk) {
    k_finalize_body<T>(blk, e.k, (int)kFinEnergy, e.L.partial, firstf[s + 1] - firstf[s], e.f.n,
                       e.ctrl);
  });
}

template <class T>
__global__ void kb_stop_state(const char* __restrict__ tab, int ns, int* out) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ns) return;
  const TsphControl* c = ((SceneEntry<T>*)tab)[s].ctrl;
  out[2 * s] = c->it;
  out[2 * s + 1] = c->stop;
}

template <class T>
int batch_density_alpha(const TsphBatchInfo& bi, char* tab, int grid, hipStream_t st) {
  const int ns = bi.nscenes;
  hipLaunchKernelGGL(kb_step_reset<T>, dim3(ns), dim3(kBlock), 0, st, tab, ns);
  if (int g = grid_for(bi.tiles[kTileFluid] + bi.tiles[kTileBound], grid))
    hipLaunchKernelGGL(kb_pass_a_both<T>, dim3(g), dim3(kBlock), 0, st, tab, ns);
  if (bi.tiles[kTileBound] > 0)
    hipLaunchKernelGGL(kb_mark_static_sums<T>, dim3((ns + 63) / 64), dim3(64), 0, st, tab, ns);
  TSPH_LAUNCH_CHECK("batch_density_alpha");
  return 0;
}

template <class T>
int batch_solver_begin(const TsphBatchInfo& bi, char* tab, int solve, uint64_t* signal, int token,
                       hipStream_t st) {
  hipLaunchKernelGGL(kb_solver_begin<T>, dim3(bi.nscenes), dim3(kBlock), 0, st, tab, bi.nscenes,
                     solve, signal, token);
  TSPH_LAUNCH_CHECK("batch_solver_begin");
  return 0;
}

template <class T>
int batch_solver_half(const TsphBatchInfo& bi, char* tab, int kk, int half, int grid,
                      hipStream_t st) {
  const int g = grid_for(bi.tiles[kTileFluid], grid);
  if (!g) return 0;
  if (half & 1) hipLaunchKernelGGL(kb_solver_psi<T>, dim3(g), dim3(kBlock), 0, st, tab, bi.nscenes, kk);
  if (half & 2)
    hipLaunchKernelGGL(kb_solver_update<T>, dim3(g), dim3(kBlock), 0, st, tab, bi.nscenes, kk);
  TSPH_LAUNCH_CHECK("batch_solver_iteration");
  return 0;
}

template <class T>
int batch_advect(const TsphBatchInfo& bi, char* tab, int grid, hipStream_t st) {
  const int g = grid_for(bi.tiles[kTileFluid], grid);
  if (!g) return 0;
  hipLaunchKernelGGL(kb_advect<T>, dim3(g), dim3(kBlock), 0, st, tab, bi.nscenes);
  hipLaunchKernelGGL(kb_surface_tension<T>, dim3(g), dim3(kBlock), 0, st, tab, bi.nscenes);
  TSPH_LAUNCH_CHECK("batch_advect");
  return 0;
}

template <class T>
int batch_tail(const TsphBatchInfo& bi, char* tab, int grid, hipStream_t st) {
  const int g = grid_for(bi.tiles[kTileFluid], grid);
  if (g) hipLaunchKernelGGL(kb_tail<T>, dim3(g), dim3(kBlock), 0, st, tab, bi.nscenes);
  hipLaunchKernelGGL(kb_finalize_energy<T>, dim3(bi.nscenes), dim3(kBlock), 0, st, tab, bi.nscenes);
  TSPH_LAUNCH_CHECK("batch_tail");
  return 0;
}

template <class T>
int batch_stop_state(const TsphBatchInfo& bi, char* tab, int* out, hipStream_t st) {
  hipLaunchKernelGGL(kb_stop_state<T>, dim3((bi.nscenes + 63) / 64), dim3(64), 0, st, tab,
                     bi.nscenes, out);
  TSPH_LAUNCH_CHECK("batch_stop_state");
  return 0;
}

#define TSPH_BATCH_INST(T)                                                                  \
  template int batch_density_alpha<T>(const TsphBatchInfo&, char*, int, hipStream_t);      \
  template int batch_solver_begin<T>(const TsphBatchInfo&, char*, int, uint64_t*, int, hipStream_t); \
  template int batch_solver_half<T>(const TsphBatchInfo&, char*, int, int, int, hipStream_t); \
  template int batch_advect<T>(const TsphBatchInfo&, char*, int, hipStream_t);             \
  template int batch_tail<T>(const TsphBatchInfo&, char*, int, hipStream_t);               \
  template int batch_stop_state<T>(const TsphBatchInfo&, char*, int*, hipStream_t);
TSPH_BATCH_INST(float)
TSPH_BATCH_INST(double)
#undef TSPH_BATCH_INST

template int relist<float>(const TsphParams&, TsphObject*, TsphObject*, TsphControl*, void*, size_t, hipStream_t);
template int relist<double>(const TsphParams&, TsphObject*, TsphObject*, TsphControl*, void*, size_t, hipStream_t);
template int density_alpha<float>(const TsphParams&, TsphObject*, TsphObject*, TsphControl*, void*, size_t, hipStream_t);
template int density_alpha<double>(const TsphParams&, TsphObject*, TsphObject*, TsphControl*, void*, size_t, hipStream_t);
template int solver_begin<float>(const TsphParams&, TsphControl*, int, int*, int, hipStream_t);
template int solver_begin<double>(const TsphParams&, TsphControl*, int, int*, int, hipStream_t);
template int solver_half<float>(const TsphParams&, TsphObject*, TsphObject*, TsphControl*, int, int, void*, void*, const void*, void*, size_t, hipStream_t);
template int solver_half<double>(const TsphParams&, TsphObject*, TsphObject*, TsphControl*, int, int, void*, void*, const void*, void*, size_t, hipStream_t);
template int advect<float>(const TsphParams&, TsphObject*, TsphObject*, TsphControl*, void*, size_t, hipStream_t);
template int advect<double>(const TsphParams&, TsphObject*, TsphObject*, TsphControl*, void*, size_t, hipStream_t);
template int fbm_round<float>(const TsphParams&, TsphObject*, TsphControl*, void*, size_t, hipStream_t);
template int fbm_round<double>(const TsphParams&, TsphObject*, TsphControl*, void*, size_t, hipStream_t);
template int tail<float>(const TsphParams&, TsphObject*, TsphControl*, int, void*, size_t, hipStream_t);
template int tail<double>(const TsphParams&, TsphObject*, TsphControl*, int, void*, size_t, hipStream_t);

}  // namespace tsph
