// The reference's kernel surface on the GPU: each sph.py kernel with its
// (ngrid, obj, nobj) signature, walking the shared Ngrid
// (node_part_count / node_part_shift / part_pid_in_node / part_uid_in_node,
// sph_obj.py:280-323) with the uid filter exactly as the reference does
// (sph.py:9-19).  Strict arithmetic in the reference's evaluation order, one
// thread per particle of `obj`, accumulation into the existing field values.
// These kernels are the drop-in for callers that drive the solver kernel by
// kernel; the fused step (tsph_step.hip) is the fast path.
#include "tsph_host.h"

namespace tsph {

struct RefGrid {
  const int *npc, *nps, *pin, *uin;
};

template <class T>
struct RefObj {
  int n, uid;
  V4<T>*pos, *vel, *vel_adv, *acce_adv, *acce, *normal, *alpha_1;
  T *mass, *rest_density, *rest_volume, *vf, *vf_tmp;
  V4<T>* drift;
  T *W, *comp, *dens, *psi_adv, *alpha, *alpha_2, *zeta, *pressure;
  int *flag, *cs;
};

template <class T>
RefObj<T> ref_obj(const TsphObject* o) {
  RefObj<T> r;
  r.n = o->n;
  r.uid = o->uid;
  r.pos = (V4<T>*)o->pos;
  r.vel = (V4<T>*)o->vel;
  r.vel_adv = (V4<T>*)o->vel_adv;
  r.acce_adv = (V4<T>*)o->acce_adv;
  r.acce = (V4<T>*)o->acce;
  r.normal = (V4<T>*)o->normal;
  r.alpha_1 = (V4<T>*)o->alpha_1;
  r.mass = (T*)o->mass;
  r.rest_density = (T*)o->rest_density;
  r.rest_volume = (T*)o->rest_volume;
  r.vf = (T*)o->volume_frac;
  r.vf_tmp = (T*)o->volume_frac_tmp;
  r.drift = (V4<T>*)o->drift_vel;
  r.W = (T*)o->W;
  r.comp = (T*)o->sph_compression;
  r.dens = (T*)o->sph_density;
  r.psi_adv = (T*)o->psi_adv;
  r.alpha = (T*)o->alpha;
  r.alpha_2 = (T*)o->alpha_2;
  r.zeta = (T*)o->fbm_zeta;
  r.pressure = (T*)o->pressure;
  r.flag = o->flag;
  r.cs = o->cell_structured;
  return r;
}

// DFSPH/VFSPH aliases, sph_obj.py:65-73
template <class T>
__device__ __forceinline__ T Xof(const KP<T>& k, const RefObj<T>& o, int i) {
  return k.solver == TSPH_VFSPH ? o.rest_volume[i] : o.mass[i];
}
template <class T>
__device__ __forceinline__ T psiof(const KP<T>& k, const RefObj<T>& o, int i) {
  return k.solver == TSPH_VFSPH ? o.comp[i] : o.dens[i];
}
template <class T>
__device__ __forceinline__ T rpsiof(const KP<T>& k, const RefObj<T>& o, int i) {
  return k.solver == TSPH_VFSPH ? T(1) : o.rest_density[i];
}

// sph.py:11-19
template <class T, class F>
__device__ __forceinline__ void ref_walk(const KP<T>& k, const RefGrid& g, const int* cs, int nuid,
                                         F&& f) {
  const int side = 2 * k.R + 1;
  for (int t = 0; t < k.ntemp; ++t) {
    int ox = t % side - k.R;
    int oy = (t / side) % side - k.R;
    int oz = k.dim == 3 ? t / (side * side) - k.R : 0;
    int code = dim_encode(k, wadd(cs[0], ox), wadd(cs[1], oy), wadd(cs[2], oz));
    if (0 < code && code < k.node_num) {
      int cnt = g.npc[code];
      for (int j = 0; j < cnt; ++j) {
        int shift = g.nps[code] + j;
        if (g.uin[shift] == nuid) f(g.pin[shift]);
      }
    }
  }
}

template <class T>
__device__ __forceinline__ T ref_norm(T x, T y, T z) {
  return s_sqrt(s_r2(x, y, z));
}

template <class T>
__global__ void __launch_bounds__(kBlock)
    k_ref(int kid, KP<T> k, RefGrid g, RefObj<T> o, RefObj<T> no, double* partial,
          TsphControl* ctrl) {
  TSPH_STRICT
  __shared__ double lds[kWaves];
  const int nblk = gridDim.x;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  const bool act = i < o.n;
  const T dt = (T)ctrl->dt;
  const int P = k.P;
  double red0 = 0, red1 = 0;
  double redmin = 1e300;
  const int* cs = act ? o.cs + 4 * i : nullptr;
  switch (kid) {
    case TSPH_K_SPH_clean_value:  // sph.py:22-38
      if (blockIdx.x == 0 && threadIdx.x == 0) ctrl->general_flag = 1;
      if (act) {
        o.W[i] = 0; o.comp[i] = 0; o.dens[i] = 0; o.alpha_2[i] = 0; o.flag[i] = 0;
        V4<T> z{T(0), T(0), T(0), T(0)};
        o.acce[i] = z; o.acce_adv[i] = z; o.alpha_1[i] = z;
        for (int q = 0; q < P; ++q) o.drift[(long)i * P + q] = z;
      }
      return;
    case TSPH_K_cfl_condition:  // sph.py:41-47
      if (act) {
        V4<T> v = o.vel[i];
        T vn = ref_norm(v.x, v.y, v.z);
        if (vn > T(1e-4)) {
          T q = k.gps / vn;
          redmin = (double)(q * k.cfl);
        }
      }
      break;
    case TSPH_K_SPH_prepare_attr:  // sph.py:50-64
      if (act) {
        V4<T> p = o.pos[i];
        T w = o.W[i], cp = o.comp[i], dn = o.dens[i];
        ref_walk(k, g, cs, no.uid, [&](int j) {
          V4<T> q = no.pos[j];
          T Wr = s_W(k, ref_norm(p.x - q.x, p.y - q.y, p.z - q.z));
          w = w + Wr;
          cp = cp + Wr * no.rest_volume[j];
          dn = dn + Wr * no.mass[j];
        });
        o.W[i] = w; o.comp[i] = cp; o.dens[i] = dn;
      }
      return;
    case TSPH_K_SPH_prepare_alpha_1:  // sph.py:67-81
      if (act) {
        V4<T> p = o.pos[i];
        V4<T> a = o.alpha_1[i];
        ref_walk(k, g, cs, no.uid, [&](int j) {
          V4<T> q = no.pos[j];
          T dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
          T r = ref_norm(dx, dy, dz);
          if (r > T(0)) {
            T Xj = Xof(k, no, j), wg = s_Wg(k, r);
            a.x = a.x + Xj * dx / r * wg;
            a.y = a.y + Xj * dy / r * wg;
            a.z = a.z + Xj * dz / r * wg;
          }
        });
        o.alpha_1[i] = a;
      }
      return;
    case TSPH_K_SPH_prepare_alpha_2:  // sph.py:84-97
      if (act) {
        V4<T> p = o.pos[i];
        T a2 = o.alpha_2[i];
        ref_walk(k, g, cs, no.uid, [&](int j) {
          V4<T> q = no.pos[j];
          T r = ref_norm(p.x - q.x, p.y - q.y, p.z - q.z);
          if (r > T(0)) {
            T wg = s_Wg(k, r), Xj = Xof(k, no, j);
            T wg2 = wg * wg;
            a2 = a2 + wg2 * (Xj * Xj) / no.mass[j];
          }
        });
        o.alpha_2[i] = a2;
      }
      return;
    case TSPH_K_SPH_prepare_alpha:  // sph.py:100-105
      if (act) {
        V4<T> a = o.alpha_1[i];
        T al = s_r2(a.x, a.y, a.z) / o.mass[i] + o.alpha_2[i];
        if (al < T(1e-4)) al = T(1e-4);
        o.alpha[i] = al;
      }
      return;
    case TSPH_K_SPH_advection_gravity_acc:  // sph.py:108-111
      if (act) {
        V4<T> a = o.acce_adv[i];
        o.acce_adv[i] = V4<T>{a.x + k.g[0], a.y + k.g[1], a.z + k.g[2], a.w};
      }
      return;
    case TSPH_K_SPH_advection_viscosity_acc:  // sph.py:114-129
      if (act) {
        V4<T> p = o.pos[i], vi = o.vel[i], a = o.acce_adv[i];
        T rd = o.rest_density[i];
        const T mu = mu_of(k, o.vf, i);  // the global mu, or the per-phase extension
        ref_walk(k, g, cs, no.uid, [&](int j) {
          V4<T> q = no.pos[j], vj = no.vel[j];
          T dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
          T r = ref_norm(dx, dy, dz);
          if (r > T(0)) {
            T Vj = Xof(k, no, j) / psiof(k, no, j);
            T s = k.two_d2 * Vj * s_Wg(k, r);
            T inv = T(1) / ref_norm(dx, dy, dz);
            T Ax = (vi.x - vj.x) * dx;
            Ax = Ax + (vi.y - vj.y) * dy;
            if (k.dim == 3) Ax = Ax + (vi.z - vj.z) * dz;
            T den = k.eps + r * r;
            a.x = a.x + s * (inv * dx) * Ax / den * mu / rd;
            a.y = a.y + s * (inv * dy) * Ax / den * mu / rd;
            if (k.dim == 3) a.z = a.z + s * (inv * dz) * Ax / den * mu / rd;
          }
        });
        o.acce_adv[i] = a;
      }
      return;
    case TSPH_K_SPH_advection_surface_tension_acc:
    case TSPH_K_SPH_FBM_convect:
      return;  // two-loop kernels: launched as two passes by the host (see below)
    case TSPH_K_WC_pressure_val:  // sph.py:169-175
      if (act) {
        T cs2 = (T)(k.cs * k.cs);
        T ratio = o.dens[i] / o.rest_density[i];
        T res = T(1), b = ratio;
        int e = 7;
        while (e > 0) {
          if (e & 1) res = res * b;
          b = b * b;
          e >>= 1;
        }
        T pr = (o.rest_density[i] * cs2 / (T)k.wc_gamma) * (res - T(1));
        if (pr < T(0)) pr = T(0);
        o.pressure[i] = pr;
      }
      return;
    case TSPH_K_WC_pressure_acce:  // sph.py:178-193
      if (act) {
        V4<T> p = o.pos[i], a = o.acce_adv[i];
        T pi = o.pressure[i], di = o.dens[i];
        ref_walk(k, g, cs, no.uid, [&](int j) {
          V4<T> q = no.pos[j];
          T dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
          T r = ref_norm(dx, dy, dz);
          T dj = no.dens[j];
          T pt = pi / (di * di) + no.pressure[j] / (dj * dj);
          if (r > T(0)) {
            T wg = s_Wg(k, r);
            T mj = no.mass[j];
            a.x = a.x + -pt * mj * dx / r * wg;
            a.y = a.y + -pt * mj * dy / r * wg;
            if (k.dim == 3) a.z = a.z + -pt * mj * dz / r * wg;
          }
        });
        o.acce_adv[i] = a;
      }
      return;
    case TSPH_K_IPPE_adv_psi_init:  // sph.py:196-199
      if (act) o.psi_adv[i] = psiof(k, o, i) - rpsiof(k, o, i);
      return;
    case TSPH_K_IPPE_adv_psi:  // sph.py:202-216
      if (act) {
        V4<T> p = o.pos[i], va = o.vel_adv[i];
        T psi = o.psi_adv[i];
        ref_walk(k, g, cs, no.uid, [&](int j) {
          V4<T> q = no.pos[j], vj = no.vel_adv[j];
          T dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
          T r = ref_norm(dx, dy, dz);
          if (r > T(0)) {
            T wg = s_Wg(k, r);
            T dot = (dx / r * wg) * (va.x - vj.x);
            dot = dot + (dy / r * wg) * (va.y - vj.y);
            if (k.dim == 3) dot = dot + (dz / r * wg) * (va.z - vj.z);
            psi = psi + dot * Xof(k, no, j) * dt;
          }
        });
        o.psi_adv[i] = psi;
      }
      return;
    case TSPH_K_IPPE_psi_adv_non_negative:  // sph.py:219-226
      if (act) {
        T psi = o.psi_adv[i];
        if (psi < T(0)) psi = T(0);
        o.psi_adv[i] = psi;
        red0 = (double)(psi / rpsiof(k, o, i));
      }
      break;
    case TSPH_K_IPPE_update_vel_adv:  // sph.py:229-244
      if (act) {
        V4<T> p = o.pos[i], va = o.vel_adv[i];
        T pi = o.psi_adv[i], ai = o.alpha[i], Xi = Xof(k, o, i), mi = o.mass[i];
        T c = -(T(1) / dt);
        ref_walk(k, g, cs, no.uid, [&](int j) {
          V4<T> q = no.pos[j];
          T dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
          T r = ref_norm(dx, dy, dz);
          if (r > T(0)) {
            T a = pi * Xof(k, no, j) / ai;
            T b = no.psi_adv[j] * Xi / no.alpha[j];
            T s = c * (a + b);
            T wg = s_Wg(k, r);
            va.x = va.x + s * (dx / r * wg) / mi;
            va.y = va.y + s * (dy / r * wg) / mi;
            if (k.dim == 3) va.z = va.z + s * (dz / r * wg) / mi;
          }
        });
        o.vel_adv[i] = va;
      }
      return;
    case TSPH_K_SPH_advection_update_vel_adv:  // sph.py:247-250
      if (act) {
        V4<T> v = o.vel[i], a = o.acce_adv[i];
        o.vel_adv[i] = V4<T>{v.x + a.x * dt, v.y + a.y * dt, v.z + a.z * dt, T(0)};
      }
      return;
    case TSPH_K_SPH_vel_2_vel_adv:  // sph.py:253-256
      if (act) o.vel_adv[i] = o.vel[i];
      return;
    case TSPH_K_SPH_vel_adv_2_vel:  // sph.py:259-262
      if (act) o.vel[i] = o.vel_adv[i];
      return;
    case TSPH_K_SPH_update_pos:  // sph.py:265-269
      if (act) {
        V4<T> v = o.vel_adv[i], p = o.pos[i];
        o.vel[i] = v;
        o.pos[i] = V4<T>{p.x + v.x * dt, p.y + v.y * dt, p.z + v.z * dt, p.w};
      }
      return;
    case TSPH_K_SPH_update_energy:  // sph.py:272-279
      if (act) {
        V4<T> v = o.vel[i], p = o.pos[i];
        T m = o.mass[i];
        red0 = (double)(T(0.5) * m * s_r2(v.x, v.y, v.z));
        red1 = (double)(-m * k.g[1] * (p.y - k.lb[1]));
      }
      break;
    case TSPH_K_SPH_update_mass:  // sph.py:282-286
      if (act) {
        T rd = k.rho[0] * o.vf[(long)i * P];
        for (int q = 1; q < P; ++q) rd = rd + k.rho[q] * o.vf[(long)i * P + q];
        o.rest_density[i] = rd;
        T m = rd * o.rest_volume[i];
        o.mass[i] = m;
        V4<T> p = o.pos[i];
        p.w = k.solver == TSPH_VFSPH ? o.rest_volume[i] : m;
        o.pos[i] = p;
      }
      return;
    case TSPH_K_SPH_FBM_clean_tmp:  // sph.py:301-305
      if (act)
        for (int q = 0; q < P; ++q) o.vf_tmp[(long)i * P + q] = T(0);
      return;
    case TSPH_K_SPH_FBM_diffuse:  // sph.py:308-326
      if (act && o.flag[i] == 0) {
        V4<T> p = o.pos[i];
        T acc[kMaxP];
        for (int q = 0; q < P; ++q) acc[q] = o.vf_tmp[(long)i * P + q];
        T pre = dt * k.fbmD;
        ref_walk(k, g, cs, no.uid, [&](int j) {
          if (no.flag[j] != 0) return;
          V4<T> q = no.pos[j];
          T r = ref_norm(p.x - q.x, p.y - q.y, p.z - q.z);
          if (r > T(0)) {
            T wg = s_Wg(k, r);
            T den = r * r + k.eps;
            for (int s = 0; s < P; ++s) {
              T dphi = o.vf[(long)i * P + s] - no.vf[(long)j * P + s];
              acc[s] = acc[s] + pre * dphi * no.rest_volume[j] * r * wg / den;
            }
          }
        });
        for (int q = 0; q < P; ++q) o.vf_tmp[(long)i * P + q] = acc[q];
      }
      return;
    case TSPH_K_SPH_FBM_check_tmp:  // sph.py:366-372
      if (act) {
        bool neg = false;
        for (int q = 0; q < P; ++q)
          if (o.vf[(long)i * P + q] + o.vf_tmp[(long)i * P + q] < T(0)) neg = true;
        if (neg) {
          o.flag[i] = 1;
          red0 = 1;
        }
      }
      break;
    case TSPH_K_SPH_update_volume_frac:  // sph.py:375-379
      if (act && !(o.flag[i] > 0))
        for (int q = 0; q < P; ++q)
          o.vf[(long)i * P + q] = o.vf[(long)i * P + q] + o.vf_tmp[(long)i * P + q];
      return;
    default:
      return;
  }
  // reductions: block partials, finished by k_finalize in stream order
  double s0 = block_sum(red0, lds);
  double s1 = block_sum(red1, lds);
  double mn = block_min(redmin, lds);
  if (threadIdx.x == 0) {
    partial[blockIdx.x] = kid == TSPH_K_cfl_condition ? mn : s0;
    partial[nblk + blockIdx.x] = s1;
  }
}

// Surface tension: loop 1 (normals) and loop 2 (forces), sph.py:134-166.
template <class T>
__global__ void __launch_bounds__(kBlock)
    k_ref_st(int pass, KP<T> k, RefGrid g, RefObj<T> o, RefObj<T> no) {
  TSPH_STRICT
  int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= o.n) return;
  const int* cs = o.cs + 4 * i;
  V4<T> p = o.pos[i];
  if (pass == 0) {
    T nx = 0, ny = 0, nz = 0;
    ref_walk(k, g, cs, no.uid, [&](int j) {
      V4<T> q = no.pos[j];
      T dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
      T r = ref_norm(dx, dy, dz);
      if (r > T(0)) {
        T s = -Xof(k, no, j) / psiof(k, no, j) * s_Wg(k, r);
        nx = nx + s * dx / r;
        ny = ny + s * dy / r;
        nz = nz + s * dz / r;
      }
    });
    o.normal[i] = V4<T>{nx * k.h, ny * k.h, nz * k.h, T(0)};
    return;
  }
  V4<T> a = o.acce_adv[i];
  V4<T> ni = o.normal[i];
  T fi = o.vf[(long)i * k.P];
  ref_walk(k, g, cs, no.uid, [&](int j) {
    V4<T> q = no.pos[j];
    T dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
    T r = ref_norm(dx, dy, dz);
    if (r > T(0) && fi > T(0.99) && no.vf[(long)j * k.P] > T(0.99)) {
      T cs2 = -k.gamma * no.mass[j] * s_C(k, r);
      V4<T> nj = no.normal[j];
      T f = T(2) * rpsiof(k, o, i) / (psiof(k, o, i) + psiof(k, no, j));
      a.x = a.x + f * (cs2 * dx / r + k.gamma * (ni.x - nj.x));
      a.y = a.y + f * (cs2 * dy / r + k.gamma * (ni.y - nj.y));
      if (k.dim == 3) a.z = a.z + f * (cs2 * dz / r + k.gamma * (ni.z - nj.z));
    }
  });
  o.acce_adv[i] = a;
}

// SPH_FBM_convect loop 1 (drift velocities) and loop 2 (flux), sph.py:329-363.
template <class T>
__global__ void __launch_bounds__(kBlock)
    k_ref_convect(int pass, KP<T> k, RefGrid g, RefObj<T> o, RefObj<T> no,
                  const TsphControl* ctrl) {
  TSPH_STRICT
  int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= o.n) return;
  const T dt = (T)ctrl->dt;
  const int P = k.P;
  if (pass == 0) {
    V4<T> va = o.vel_adv[i], v = o.vel[i];
    T aa[3] = {(va.x - v.x) / dt, (va.y - v.y) / dt, (va.z - v.z) / dt};
    T rd = o.rest_density[i];
    T z = T(0);
    for (int q = 0; q < P; ++q) z = z + o.vf[(long)i * P + q] * (k.rho[q] - rd) / k.rho[q];
    T fa[3];
    for (int d = 0; d < 3; ++d) fa[d] = (aa[d] - z * k.g[d]) / (T(1) - z);
    o.acce_adv[i] = V4<T>{aa[0], aa[1], aa[2], T(0)};
    o.acce[i] = V4<T>{fa[0], fa[1], fa[2], T(0)};
    o.zeta[i] = z;
    for (int q = 0; q < P; ++q) {
      T phi = o.vf[(long)i * P + q];
      T s = phi * (k.rho[q] - rd);
      T u[3];
      for (int d = 0; d < 3; ++d) u[d] = s * (k.g[d] - fa[d]);
      T wgt = phi * k.rho[q];
      if (wgt > T(1e-6)) {
        for (int d = 0; d < 3; ++d) u[d] = u[d] / wgt;
      } else {
        for (int d = 0; d < 3; ++d) u[d] = u[d] * T(0);
      }
      for (int d = 0; d < 3; ++d) u[d] = u[d] + (fa[d] - aa[d]);
      for (int d = 0; d < 3; ++d) u[d] = u[d] * dt;
      if (k.dim == 2) u[2] = T(0);
      o.drift[(long)i * P + q] = V4<T>{u[0], u[1], u[2], T(0)};
    }
    return;
  }
  if (o.flag[i] != 0) return;
  const int* cs = o.cs + 4 * i;
  V4<T> p = o.pos[i];
  T acc[kMaxP];
  for (int q = 0; q < P; ++q) acc[q] = o.vf_tmp[(long)i * P + q];
  T pre = k.fbmC * dt;
  ref_walk(k, g, cs, no.uid, [&](int j) {
    if (no.flag[j] != 0) return;
    V4<T> q = no.pos[j];
    T dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
    T r = ref_norm(dx, dy, dz);
    if (r > T(0)) {
      T wg = s_Wg(k, r);
      for (int s = 0; s < P; ++s) {
        V4<T> ui = o.drift[(long)i * P + s], uj = no.drift[(long)j * P + s];
        T pi = o.vf[(long)i * P + s], pj = no.vf[(long)j * P + s];
        T dot = (pi * ui.x + pj * uj.x) * (dx / r);
        dot = dot + (pi * ui.y + pj * uj.y) * (dy / r);
        if (k.dim == 3) dot = dot + (pi * ui.z + pj * uj.z) * (dz / r);
        acc[s] = acc[s] - pre * no.rest_volume[j] * dot * wg;
      }
    }
  });
  for (int q = 0; q < P; ++q) o.vf_tmp[(long)i * P + q] = acc[q];
}

template <class T>
int ref_kernel(const TsphParams& prm, int kid, TsphObject* obj, const TsphObject* nobj,
               const int* npc, const int* nps, const int* pin, const int* uin, TsphControl* ctrl,
               void* ws, size_t wsb, hipStream_t st) {
  TSPH_REQUIRE(kid >= 0 && kid < TSPH_K_COUNT, TSPH_EINVAL, "tsph_ref_kernel: bad kernel id %d",
               kid);
  KP<T> k = make_kp<T>(prm);
  WsLayout L;
  TSPH_REQUIRE(carve_ws(ws, wsb, obj->capacity, prm.node_num, &L), TSPH_ECAPACITY,
               "tsph_ref_kernel: workspace too small");
  RefGrid g{npc, nps, pin, uin};
  RefObj<T> o = ref_obj<T>(obj);
  RefObj<T> no = ref_obj<T>(nobj ? nobj : obj);
  int nb = blocks_for(obj->n > 0 ? obj->n : 1);
  if (kid == TSPH_K_SPH_advection_surface_tension_acc) {
    hipLaunchKernelGGL(k_ref_st<T>, dim3(nb), dim3(kBlock), 0, st, 0, k, g, o, no);
    hipLaunchKernelGGL(k_ref_st<T>, dim3(nb), dim3(kBlock), 0, st, 1, k, g, o, no);
  } else if (kid == TSPH_K_SPH_FBM_convect) {
    hipLaunchKernelGGL(k_ref_convect<T>, dim3(nb), dim3(kBlock), 0, st, 0, k, g, o, no, ctrl);
    hipLaunchKernelGGL(k_ref_convect<T>, dim3(nb), dim3(kBlock), 0, st, 1, k, g, o, no, ctrl);
  } else {
    hipLaunchKernelGGL(k_ref<T>, dim3(nb), dim3(kBlock), 0, st, kid, k, g, o, no, L.partial,
                       ctrl);
    int mode = kid == TSPH_K_cfl_condition ? kFinCfl
               : kid == TSPH_K_SPH_update_energy ? kFinEnergy
               : kid == TSPH_K_SPH_FBM_check_tmp ? kFinFlag
               : kid == TSPH_K_IPPE_psi_adv_non_negative ? kFinCompression : -1;
    if (mode >= 0)
      hipLaunchKernelGGL(k_finalize<T>, dim3(1), dim3(kBlock), 0, st, k, mode, L.partial, nb,
                         obj->n, ctrl);
  }
  TSPH_LAUNCH_CHECK("ref_kernel");
  return 0;
}

template int ref_kernel<float>(const TsphParams&, int, TsphObject*, const TsphObject*, const int*, const int*, const int*, const int*, TsphControl*, void*, size_t, hipStream_t);
template int ref_kernel<double>(const TsphParams&, int, TsphObject*, const TsphObject*, const int*, const int*, const int*, const int*, TsphControl*, void*, size_t, hipStream_t);

}  // namespace tsph
