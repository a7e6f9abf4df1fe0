// Shared device-side definitions of the MI355X SPH hot path.
//
// Two flavours of pair arithmetic live here:
//  * strict: the reference's evaluation order (sph_util.py:4-43, sph.py) with
//    every operation rounded once (no contraction, correctly rounded divide
//    and sqrt).  Used where last-ulp agreement matters: cell keys, the
//    in-range test, the density sums (which the solvers difference against
//    the rest density) and the reference-surface kernels.
//  * fast: algebraically equal forms with one reciprocal square root per
//    pair.  Used by the Jacobi and advection sweeps.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/tichi_sph.h"

namespace tsph {

constexpr int kBlock = 256;     // threads per block (4 wave64)
constexpr int kWaves = kBlock / 64;
constexpr int kMaxP = TSPH_MAX_PHASE;

template <class T>
struct alignas(4 * sizeof(T)) V4 {
  T x, y, z, w;
};

// Typed per-scene constants, built on the host from TsphParams.  Values the
// reference keeps in f32 fields are exact in T.
template <class T>
struct KP {
  int dim, P, solver, ntemp, node_num, R;
  int gs[3], coder[3];
  int list_cap_f, list_cap_b;
  int iter_min, iter_max, cs, wc_gamma;
  int st_rec;    // advection writes the surface-tension records also at gamma = 0
  int fastq;     // FP32: q = r / h and sqrt by the verified reduced sequences (f_quot_h)
  T lb[3], g[3], rho[kMaxP];
  T h, h2, sig, sig_h, gps, c_coh, eps, r2max, inv_h;
  T mu, gamma, fbmD, fbmC, cfl;
  T div_thr, comp_thr;
  T two_d2;      // 2*(2+dim) of W_lap (sph_util.py:42)
  T wg_a, wg_b;  // fast gradient factors: 6*sig/h^2 and 6*sig/h
  int mu_pp;     // per-phase viscosity (TsphParams.mu_per_phase)
  T mu_k[kMaxP];
};

// A kernel body's block: index and count within its launch, or within its
// scene's share of a grouped (batched) launch.
struct Blk {
  int b, nb;
};

// ---------------------------------------------------------------------------
// Block mapping.  Workgroups are dispatched round-robin over the 8 XCDs
// (block b runs on XCD b % 8); remapping gives each XCD one contiguous range
// of particles, so a sweep's neighbour halo is shared inside one L2.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int xcd_block(int b, int nb) {
  int q = nb >> 3, rem = nb & 7, x = b & 7, idx = b >> 3;
  return x * q + (x < rem ? x : rem) + idx;
}

// ---------------------------------------------------------------------------
// Strict arithmetic (reference evaluation order, single rounding per op).
// ---------------------------------------------------------------------------
#define TSPH_STRICT _Pragma("clang fp contract(off)")

// The dynamic viscosity of particle i: the global mu (sph_config.py:148), or
// with the per-phase extension sum_k phi_ik mu_k, k ascending.
template <class T>
__device__ __forceinline__ T mu_of(const KP<T>& k, const T* __restrict__ vf, int i) {
  TSPH_STRICT
  if (!k.mu_pp) return k.mu;
  const T* phi = vf + (long)i * k.P;
  T m = phi[0] * k.mu_k[0];
  for (int q = 1; q < k.P; ++q) m = m + phi[q] * k.mu_k[q];
  return m;
}

template <class T>
__device__ __forceinline__ T s_sqrt(T x) {
  return __builtin_sqrt(x);
}
template <>
__device__ __forceinline__ float s_sqrt<float>(float x) {
  return __builtin_sqrtf(x);
}

// ((x0*x0) + (x1*x1)) + (x2*x2), the reference's norm_sqr (z = 0 in 2D adds +0)
template <class T>
__device__ __forceinline__ T s_r2(T dx, T dy, T dz) {
  TSPH_STRICT
  T a = dx * dx;
  T b = dy * dy;
  T c = dz * dz;
  T s = a + b;
  return s + c;
}

template <class T>
__device__ __forceinline__ T s_pow3(T x) {
  TSPH_STRICT
  T t = x * x;
  return x * t;
}

// W, sph_util.py:16-25, from q = r / h
template <class T>
__device__ __forceinline__ T s_W_q(const KP<T>& k, T q) {
  TSPH_STRICT
  T tmp = T(0);
  if (q <= T(0.5)) {
    T q3 = s_pow3(q);
    T q2 = q * q;
    tmp = T(6) * (q3 - q2) + T(1);
  } else if (q > T(0.5) && q < T(1)) {
    tmp = T(2) * s_pow3(T(1) - q);
  }
  return tmp * k.sig;
}
template <class T>
__device__ __forceinline__ T s_W(const KP<T>& k, T r) {
  TSPH_STRICT
  return s_W_q(k, r / k.h);
}
// s_W_q with both pieces evaluated and one selected: the same operations,
// so the same value, without a divergent branch per pair.  The empty asm
// statement only pins both pieces as computed (it emits no instruction), so
// that the compiler does not turn the select back into branches.
template <class T>
__device__ __forceinline__ T s_W_q_sel(const KP<T>& k, T q) {
  TSPH_STRICT
  T q3 = s_pow3(q);
  T q2 = q * q;
  T a = T(6) * (q3 - q2) + T(1);
  T b = T(2) * s_pow3(T(1) - q);
  __asm__("" : "+v"(a), "+v"(b));
  T tmp = q <= T(0.5) ? a : (q < T(1) ? b : T(0));
  return tmp * k.sig;
}

// W_grad, sph_util.py:28-37, from q = r / h
template <class T>
__device__ __forceinline__ T s_Wg_q(const KP<T>& k, T q) {
  TSPH_STRICT
  T tmp = T(0);
  if (q <= T(0.5)) {
    T q2 = q * q;
    T a = T(3) * q2;
    T b = T(2) * q;
    tmp = T(6) * (a - b);
  } else if (q > T(0.5) && q < T(1)) {
    T u = T(1) - q;
    tmp = T(-6) * (u * u);
  }
  return tmp * k.sig_h;
}
template <class T>
__device__ __forceinline__ T s_Wg(const KP<T>& k, T r) {
  TSPH_STRICT
  return s_Wg_q(k, r / k.h);
}

// ---------------------------------------------------------------------------
// Reduced sequences for pass A's r = sqrt(r^2) and q = r / h.
//   quotient: q0 = r * RN(1/h), then one residual correction
//             q = q0 + (r - q0 h) RN(1/h), the residual exact by FMA
//             (Markstein's correction step);
//   sqrt:     the hardware square root (within 1 ulp) followed by the
//             residual test of its two neighbours, without the IEEE
//             sequence's rescaling of tiny inputs.
// The fluid half of pass A reads r only through W(q) and the test r > 0.  A
// sweep over every FP32 r^2 in [0, r2_max] (k_verify_fast_quotient) compares
// those results of the reduced and the IEEE sequences bit for bit for this
// h; the host sets TSPH_FLAG_FAST_QUOTIENT only when none differs, so the
// strict sums stay bit-identical.  (The boundary half also reads W'(q),
// which differs for tiny r, and keeps the IEEE sequences.)
// ---------------------------------------------------------------------------
__device__ __forceinline__ float f_quot_h(float r, float h, float inv_h) {
  const float q0 = r * inv_h;
  const float e = __builtin_fmaf(-q0, h, r);
  return __builtin_fmaf(e, inv_h, q0);
}
__device__ __forceinline__ float f_sqrt_cr(float x) {
  float s = __builtin_amdgcn_sqrtf(x);
  const float sm = __int_as_float(__float_as_int(s) - 1);
  const float sp = __int_as_float(__float_as_int(s) + 1);
  const float rm = __builtin_fmaf(-sm, s, x);
  const float rp = __builtin_fmaf(-sp, s, x);
  s = rm <= 0.0f ? sm : s;
  s = rp > 0.0f ? sp : s;
  return s;
}

// Pass A's sqrt and q: the IEEE operations, or (FQ, FP32 only) the reduced
// sequences above.
template <bool FQ, class T>
__device__ __forceinline__ T pa_sqrt(T x) {
  return s_sqrt(x);
}
template <>
__device__ __forceinline__ float pa_sqrt<true, float>(float x) {
  return f_sqrt_cr(x);
}
template <bool FQ, class T>
__device__ __forceinline__ T pa_quot(const KP<T>& k, T r) {
  TSPH_STRICT
  return r / k.h;
}
template <>
__device__ __forceinline__ float pa_quot<true, float>(const KP<float>& k, float r) {
  return f_quot_h(r, k.h, k.inv_h);
}

// C, sph_util.py:4-13
template <class T>
__device__ __forceinline__ T s_C(const KP<T>& k, T r) {
  TSPH_STRICT
  T q = r / k.h;
  T tmp = T(0);
  if (q <= T(0.5)) {
    T a = T(2) * s_pow3(T(1) - q);
    tmp = a * s_pow3(q) - T(1.0 / 64);
  } else if (q > T(0.5) && q < T(1)) {
    tmp = s_pow3(T(1) - q) * s_pow3(q);
  }
  return tmp * k.c_coh;
}

// ---------------------------------------------------------------------------
// Fast arithmetic.
// ---------------------------------------------------------------------------
template <class T>
__device__ __forceinline__ T f_rsqrt(T x) {
  return rsqrt(x);
}
template <>
__device__ __forceinline__ float f_rsqrt<float>(float x) {
  return __builtin_amdgcn_rsqf(x);
}
template <class T>
__device__ __forceinline__ T f_rcp(T x) {
  return T(1) / x;
}
template <>
__device__ __forceinline__ float f_rcp<float>(float x) {
  return __builtin_amdgcn_rcpf(x);
}

// The cohesion kernel C (sph_util.py:4-13) of q = r / h, in fast arithmetic.
template <class T>
__device__ __forceinline__ T f_coh(const KP<T>& k, T q) {
  const T u = T(1) - q;
  const T uq = u * q;
  const T uq3 = uq * uq * uq;
  const T c = q <= T(0.5) ? T(2) * uq3 - T(1.0 / 64) : (q < T(1) ? uq3 : T(0));
  return c * k.c_coh;
}

// g = W_grad(r)/r so that grad W_ij = g * x_ij, for 0 < r2 < r2max.
// Returns r through *r_out.
template <class T>
__device__ __forceinline__ T f_grad(const KP<T>& k, T r2, T* r_out) {
  T rinv = f_rsqrt(r2);
  T r = r2 * rinv;
  T q = r * k.inv_h;
  T u = T(1) - q;
  // q <= 1/2: 6 sig/h (3q^2 - 2q) / r = 6 sig/h^2 (3q - 2)
  // q >  1/2: -6 sig/h (1 - q)^2 / r
  T ga = k.wg_a * (T(3) * q - T(2));
  T gb = -k.wg_b * (u * u) * rinv;
  *r_out = r;
  return q <= T(0.5) ? ga : gb;
}

// ---------------------------------------------------------------------------
// Cell coding, sph_util.py:65-72 (floor division in IEEE, int32 wrapping).
// ---------------------------------------------------------------------------
template <class T>
__device__ __forceinline__ int to_i32(T v) {
  if (!(v >= T(-2147483648.0) && v < T(2147483648.0))) return INT32_MIN;
  return (int)v;
}

template <class T>
__device__ __forceinline__ void node_encode(const KP<T>& k, T x, T y, T z, int* c) {
  TSPH_STRICT
  c[0] = to_i32(floor((x - k.lb[0]) / k.h));
  c[1] = to_i32(floor((y - k.lb[1]) / k.h));
  c[2] = k.dim == 3 ? to_i32(floor((z - k.lb[2]) / k.h)) : 0;
}

__device__ __forceinline__ int wadd(int a, int b) { return (int)((uint32_t)a + (uint32_t)b); }
__device__ __forceinline__ int wmul(int a, int b) { return (int)((uint32_t)a * (uint32_t)b); }

template <class T>
__device__ __forceinline__ int dim_encode(const KP<T>& k, int cx, int cy, int cz) {
  int s = wmul(k.coder[0], cx);
  s = wadd(s, wmul(k.coder[1], cy));
  if (k.dim == 3) s = wadd(s, wmul(k.coder[2], cz));
  return s;
}

// ---------------------------------------------------------------------------
// Reductions (wave64 shuffles, then LDS across the block's waves).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ int wave_or(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v |= __shfl_xor(v, o, 64);
  return v;
}

// Sum over the block; the result is valid in thread 0.  Fixed order for a
// fixed block size, so the value is deterministic.
__device__ __forceinline__ double block_sum(double v, double* lds) {
  v = wave_sum(v);
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) lds[w] = v;
  __syncthreads();
  double s = 0;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += lds[i];
  __syncthreads();
  return s;
}
__device__ __forceinline__ double block_min(double v, double* lds) {
  v = wave_min(v);
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) lds[w] = v;
  __syncthreads();
  double s = v;
  if (threadIdx.x == 0) {
    s = lds[0];
    for (int i = 1; i < (int)(blockDim.x >> 6); ++i) s = fmin(s, lds[i]);
  }
  __syncthreads();
  return s;
}

// Deterministic fixed-order sum of `n` partials by one block (result in
// thread 0).
// Block reductions of n partials: each thread loads eight at a time, all in
// flight together (one block over a few thousand partials is bound by the
// load latency), and adds them in index order, so the result is fixed.
constexpr int kPartU = 8;
__device__ __forceinline__ double sum_partials(const double* part, int n, double* lds) {
  double v = 0;
  const int s = blockDim.x;
  for (int base = threadIdx.x; base < n; base += kPartU * s) {
    double a[kPartU];
#pragma unroll
    for (int u = 0; u < kPartU; ++u)
      a[u] = base + u * s < n ? __builtin_nontemporal_load(part + base + u * s) : 0.0;
#pragma unroll
    for (int u = 0; u < kPartU; ++u) v += a[u];
  }
  return block_sum(v, lds);
}
__device__ __forceinline__ double min_partials(const double* part, int n, double* lds, double init) {
  double v = init;
  const int s = blockDim.x;
  for (int base = threadIdx.x; base < n; base += kPartU * s) {
    double a[kPartU];
#pragma unroll
    for (int u = 0; u < kPartU; ++u)
      a[u] = base + u * s < n ? __builtin_nontemporal_load(part + base + u * s) : init;
#pragma unroll
    for (int u = 0; u < kPartU; ++u) v = fmin(v, a[u]);
  }
  return block_min(v, lds);
}

// Reductions are finished by a later kernel in stream order (kernel
// boundaries order and publish the per-block partials); no block waits on
// another inside a launch.
enum FinalizeMode { kFinCfl = 0, kFinEnergy = 1, kFinFlag = 2, kFinCompression = 3 };

template <class T>
__device__ __forceinline__ void k_finalize_body(Blk blk, KP<T> k, int mode, const double* __restrict__ partial, int nblk, int n,
               TsphControl* ctrl) {
  __shared__ double lds[kWaves];
  if (mode == kFinCfl) {
    double all = min_partials(partial, nblk, lds, 1e300);
    if (threadIdx.x == 0) {
      TSPH_STRICT
      T d = k.gps / (T)k.cs;  // sph.py:43
      if (all < (double)d) d = (T)all;
      ctrl->dt = (double)d;
    }
  } else if (mode == kFinEnergy) {
    double a = sum_partials(partial, nblk, lds);
    double b = sum_partials(partial + nblk, nblk, lds);
    if (threadIdx.x == 0) {
      ctrl->ke = a;
      ctrl->pe = b;
    }
  } else if (mode == kFinFlag) {
    double a = sum_partials(partial, nblk, lds);
    if (threadIdx.x == 0) ctrl->general_flag = a > 0 ? 1 : 0;
  } else {
    double a = sum_partials(partial, nblk, lds);
    if (threadIdx.x == 0) {
      TSPH_STRICT
      ctrl->compression = (double)((T)a / (T)n);  // sph.py:226
    }
  }
}

template <class T>
__global__ void __launch_bounds__(kBlock)
    k_finalize(KP<T> k, int mode, const double* __restrict__ partial, int nblk, int n,
               TsphControl* ctrl) {
  k_finalize_body<T>(Blk{(int)blockIdx.x, (int)gridDim.x}, k, mode, partial, nblk, n, ctrl);
}

}  // namespace tsph
