// kspace.hip -- pointwise k-space and real-space kernels of the NS step.
//
// k-space kernels run one thread per complex value of the half spectrum
// (the reference loops over the NZ_ reals of a row with identical real and
// imaginary coefficients, so both parts are handled by one thread).  The
// wavevector is computed from the index (Calc_K*_Ooura, fft_wrapper.h:
// 761-773) instead of being read from the KX/KY/KZ/K2/IK2 tables of Init_K
// (fft_wrapper.cxx:47-87): the same values, five fewer streams from HBM.
#include "kspace_ops.hpp"

namespace kapsel {

namespace {

template <typename T>
struct P5 {
  T *p[5];
};

template <typename T>
__global__ void truncate_kernel(P5<T> fp, Grid g) {
  size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.nc) return;
  KIdx k = kidx(g, c);
  if (is_cut(g, k)) st2(fp.p[blockIdx.y] + 2 * c, V2<T>{0, 0});
}

// Zeta_k2u_k (fft_wrapper.cxx:614-711); with trunc, Truncate_two_third_rule
// on zeta first (operate_omega.cxx:104-106, sp_3d_ns.cxx:38-40).  fix_mask
// bit d zeroes uk_dc[d] first (FIX_CELL, sp_3d_ns.cxx:30-36).
template <typename T>
__global__ void zeta_k2u_k_kernel(T *__restrict__ z0, T *__restrict__ z1, T *__restrict__ u0,
                                  T *__restrict__ u1, T *__restrict__ u2, double *__restrict__ dc, Grid g,
                                  int trunc, int fix_mask) {
  size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.nc) return;
  KIdx k = kidx(g, c);
  V2<T> d1 = ld2(z0 + 2 * c), d2 = ld2(z1 + 2 * c);
  if (trunc && is_cut(g, k)) {
    d1 = V2<T>{0, 0};
    d2 = V2<T>{0, 0};
    st2(z0 + 2 * c, d1);
    st2(z1 + 2 * c, d2);
  }
  V2<T> a, b, e;
  zeta2u_mode(g, k, d1, d2, a, b, e);
  if (c == 0) {
    double dd[3];
    for (int d = 0; d < 3; d++) {
      dd[d] = ((fix_mask >> d) & 1) ? 0.0 : dc[d];
      if ((fix_mask >> d) & 1) dc[d] = 0.0;
    }
    a.re = (T)dd[0];
    b.re = (T)dd[1];
    e.re = (T)dd[2];
  }
  st2(u0 + 2 * c, a);
  st2(u1 + 2 * c, b);
  st2(u2 + 2 * c, e);
}

// U_k2zeta_k (fft_wrapper.cxx:428-497)
template <typename T>
__global__ void u_k2zeta_k_kernel(const T *__restrict__ u0, const T *__restrict__ u1, const T *__restrict__ u2,
                                  T *__restrict__ z0, T *__restrict__ z1, double *__restrict__ dc, Grid g) {
  size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.nc) return;
  KIdx k = kidx(g, c);
  V2<T> a = ld2(u0 + 2 * c), b = ld2(u1 + 2 * c), e = ld2(u2 + 2 * c);
  if (c == 0) {
    dc[0] = (double)a.re;
    dc[1] = (double)b.re;
    dc[2] = (double)e.re;
  }
  V2<T> za, zb;
  u2zeta_mode(g, k, a, b, e, za, zb);
  st2(z0 + 2 * c, za);
  st2(z1 + 2 * c, zb);
}

// Solenoidal_uk after the 2/3 cut of the spectrum f, then U_k2zeta_k of
// u = a + P f, which is stored in o (a and f are left unchanged): the
// solenoidal_uk and u_k2zeta_k kernels of the sum in one pass
template <typename T>
__global__ void sole_u2zeta_sum_kernel(P5<T> a3, P5<T> f3, P5<T> o3, T *__restrict__ z0, T *__restrict__ z1,
                                       double *__restrict__ dc, Grid g) {
  size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.nc) return;
  KIdx k = kidx(g, c);
  V2<T> fa{0, 0}, fb{0, 0}, fe{0, 0};
  if (!is_cut(g, k)) {
    fa = ld2(f3.p[0] + 2 * c);
    fb = ld2(f3.p[1] + 2 * c);
    fe = ld2(f3.p[2] + 2 * c);
    solenoidal_mode(g, k, fa, fb, fe);
  }
  const V2<T> a = ld2(a3.p[0] + 2 * c) + fa, b = ld2(a3.p[1] + 2 * c) + fb, e = ld2(a3.p[2] + 2 * c) + fe;
  st2(o3.p[0] + 2 * c, a);
  st2(o3.p[1] + 2 * c, b);
  st2(o3.p[2] + 2 * c, e);
  if (c == 0) {
    dc[0] = (double)a.re;
    dc[1] = (double)b.re;
    dc[2] = (double)e.re;
  }
  V2<T> za, zb;
  u2zeta_mode(g, k, a, b, e, za, zb);
  st2(z0 + 2 * c, za);
  st2(z1 + 2 * c, zb);
}

// U2advection_k, real-space part (operate_omega.cxx:30-47)
template <typename T>
__global__ void advection_products_kernel(T *__restrict__ u0, T *__restrict__ u1, T *__restrict__ u2,
                                          T *__restrict__ a0, T *__restrict__ a1, Grid g) {
  size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= g.nr) return;
  size_t row = r / (unsigned)g.nz;
  size_t im = row * g.nz_ + (r - row * g.nz);
  T x1 = u0[im], x2 = u1[im], x3 = u2[im];
  u0[im] = x2 * x3;
  u1[im] = x3 * x1;
  u2[im] = x1 * x2;
  a0[im] = (x2 - x3) * (x2 + x3);
  a1[im] = (x3 - x1) * (x3 + x1);
}

// U2advection_k k-space combination + Omega_k2zeta_k (operate_omega.cxx:56-97,
// fft_wrapper.cxx:248-284) fused with the update of zeta on the kept modes:
// EULER 0: the slaved-Euler step of NS_solver_slavedEuler
//          (fluid_solver.cxx:346-365), dmy0 = -NU dt, dmy1 = 1/NU;
// EULER 1: Add_zeta_viscous_term (operate_omega.cxx:170-185) and the forward
//          Euler step Field_solver_Euler (fluid_solver.cxx:121-136) of
//          NSsolute_solver_Euler, dmy0 = NU, dmy1 = dt.
// The advection term is left in a0/a1 as the reference leaves it in its
// global f_ns0 (slaved; store_a 0: not stored); the Euler variant leaves the
// full right-hand side there, as Rhs_NS_solute leaves it in f_ns1.
template <typename T, int EULER>
__global__ void advection_combine_slaved_kernel(const T *__restrict__ u0, const T *__restrict__ u1,
                                                const T *__restrict__ u2, T *__restrict__ a0,
                                                T *__restrict__ a1, T *__restrict__ z0, T *__restrict__ z1,
                                                Grid g, double dmy0, double dmy1, int store_a = 1) {
  size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.nc) return;
  KIdx k = kidx(g, c);
  V2<T> w[3];
  advection_omega(AdvCoef<T>(g, k), ld2(u0 + 2 * c), ld2(u1 + 2 * c), ld2(u2 + 2 * c), ld2(a0 + 2 * c),
                  ld2(a1 + 2 * c), w);
  V2<T> f0, f1;
  omega_pick(k, w, f0, f1);
  if constexpr (EULER) {
    if (!is_cut(g, k)) {
      V2<T> za = ld2(z0 + 2 * c), zb = ld2(z1 + 2 * c);
      euler_update(g, k, dmy0, dmy1, f0, f1, za, zb);
      st2(z0 + 2 * c, za);
      st2(z1 + 2 * c, zb);
    } else {
      // Rhs_NS_solute cut zeta before (fluid_solver.cxx:201-203)
      st2(z0 + 2 * c, V2<T>{0, 0});
      st2(z1 + 2 * c, V2<T>{0, 0});
    }
    st2(a0 + 2 * c, f0);
    st2(a1 + 2 * c, f1);
    return;
  }
  if (store_a) {
    st2(a0 + 2 * c, f0);
    st2(a1 + 2 * c, f1);
  }
  if (!is_cut(g, k)) {
    V2<T> za = ld2(z0 + 2 * c), zb = ld2(z1 + 2 * c);
    slaved_update(g, k, dmy0, dmy1, f0, f1, za, zb);
    st2(z0 + 2 * c, za);
    st2(z1 + 2 * c, zb);
  }
}

// The slaved step's combination and zeta update (advection_combine_slaved
// kernel, EULER 0, the advection term not stored) followed by Zeta_k2u_k
// after the 2/3 cut with the FIX_CELL mask (zeta_k2u_k kernel): one pass
// over the modes.  A cut mode's zeta is zeroed as the cut zeroes it (its
// products are not read); the velocity spectrum goes to o.
// (o may be the product fields themselves: each mode is read before it is
// written, by the same thread)
template <typename T>
__global__ void slaved_zeta2u_kernel(const T *u0, const T *u1, const T *u2, const T *a0, const T *a1,
                                     T *__restrict__ z0, T *__restrict__ z1, P5<T> o, double *__restrict__ dc,
                                     Grid g, double dmy0, double dmy1, int fix_mask) {
  size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.nc) return;
  KIdx k = kidx(g, c);
  V2<T> za{0, 0}, zb{0, 0};
  if (!is_cut(g, k)) {
    V2<T> w[3];
    advection_omega(AdvCoef<T>(g, k), ld2(u0 + 2 * c), ld2(u1 + 2 * c), ld2(u2 + 2 * c), ld2(a0 + 2 * c),
                    ld2(a1 + 2 * c), w);
    V2<T> f0, f1;
    omega_pick(k, w, f0, f1);
    za = ld2(z0 + 2 * c);
    zb = ld2(z1 + 2 * c);
    slaved_update(g, k, dmy0, dmy1, f0, f1, za, zb);
  }
  st2(z0 + 2 * c, za);
  st2(z1 + 2 * c, zb);
  V2<T> a, b, e;
  zeta2u_mode(g, k, za, zb, a, b, e);
  if (c == 0) {
    double dd[3];
    for (int d = 0; d < 3; d++) {
      dd[d] = ((fix_mask >> d) & 1) ? 0.0 : dc[d];
      if ((fix_mask >> d) & 1) dc[d] = 0.0;
    }
    a.re = (T)dd[0];
    b.re = (T)dd[1];
    e.re = (T)dd[2];
  }
  st2(o.p[0] + 2 * c, a);
  st2(o.p[1] + 2 * c, b);
  st2(o.p[2] + 2 * c, e);
}

// o = a + s P f: the spectrum a plus s times the projection of f after the
// 2/3 cut (Solenoidal_uk with the cut), a and f left unchanged
template <typename T>
__global__ void sole_add_kernel(P5<T> a3, P5<T> f3, P5<T> o3, T s, Grid g) {
  size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.nc) return;
  KIdx k = kidx(g, c);
  V2<T> fa{0, 0}, fb{0, 0}, fe{0, 0};
  if (!is_cut(g, k)) {
    fa = ld2(f3.p[0] + 2 * c);
    fb = ld2(f3.p[1] + 2 * c);
    fe = ld2(f3.p[2] + 2 * c);
    solenoidal_mode(g, k, fa, fb, fe);
  }
  st2(o3.p[0] + 2 * c, ld2(a3.p[0] + 2 * c) + scal(s, fa));
  st2(o3.p[1] + 2 * c, ld2(a3.p[1] + 2 * c) + scal(s, fb));
  st2(o3.p[2] + 2 * c, ld2(a3.p[2] + 2 * c) + scal(s, fe));
}

// Solenoidal_uk (operate_omega.cxx:187-220), optionally after the 2/3 cut
// (Solenoidal_u, operate_omega.h:109-116); in place when o = u (each point
// is read before it is written)
template <typename T>
__global__ void solenoidal_uk_kernel(P5<T> u, P5<T> o, Grid g, int trunc) {
  size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.nc) return;
  KIdx k = kidx(g, c);
  if (trunc && is_cut(g, k)) {
#pragma unroll
    for (int d = 0; d < 3; d++) st2(o.p[d] + 2 * c, V2<T>{0, 0});
    return;
  }
  V2<T> a = ld2(u.p[0] + 2 * c), b = ld2(u.p[1] + 2 * c), e = ld2(u.p[2] + 2 * c);
  solenoidal_mode(g, k, a, b, e);
  st2(o.p[0] + 2 * c, a);
  st2(o.p[1] + 2 * c, b);
  st2(o.p[2] + 2 * c, e);
}

// Make_f_particle_dt_nonsole (f_particle.h:37-60): f = up - phi u
template <typename T>
__global__ void f_nonsole_kernel(P5<T> u, P5<T> up, P5<T> f, const T *__restrict__ phi, Grid g) {
  size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= g.nr) return;
  size_t row = r / (unsigned)g.nz;
  size_t im = row * g.nz_ + (r - row * g.nz);
  T ph = phi[im];
#pragma unroll
  for (int d = 0; d < 3; d++) f.p[d][im] = up.p[d][im] - ph * u.p[d][im];
}

// Add_f_particle (f_particle.h:79-95): u += f
template <typename T>
__global__ void add_f_kernel(P5<T> u, P5<T> f, P5<T> keep, Grid g) {
  size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= g.nr) return;
  size_t row = r / (unsigned)g.nz;
  size_t im = row * g.nz_ + (r - row * g.nz);
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const T v = u.p[d][im] + f.p[d][im];
    u.p[d][im] = v;
    if (keep.p[0]) keep.p[d][im] = v;
  }
}

template <typename D, typename S>
__global__ void convert_kernel(D *__restrict__ dst, const S *__restrict__ src, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = (D)src[i];
}

constexpr unsigned TPB = 256;

// complex values kept by the 2/3 rule, and cut
double kept_count(const Grid &g) {
  return (double)(2 * g.trn_x - 1) * (2 * g.trn_y - 1) * g.trn_z;
}
double cut_count(const Grid &g) { return (double)g.nc - kept_count(g); }

template <typename T>
P5<T> p5(void *const *f, int n) {
  P5<T> r{};
  for (int i = 0; i < n; i++) r.p[i] = (T *)f[i];
  return r;
}

template <typename T>
T *F(kapsel_ctx *ctx, int f) {
  return (T *)ctx->fields[f];
}

}  // namespace

#define KAPSEL_DISPATCH(ctx, FN, ...) \
  ((ctx)->prec == KAPSEL_FP64 ? FN<double>(__VA_ARGS__) : FN<float>(__VA_ARGS__))

template <typename T>
static void truncate_t(kapsel_ctx *ctx, void *const *fields, int nf) {
  LaunchTimer lt(ctx, KAPSEL_K_POINTWISE, nf * cut_count(ctx->g) * 2 * sizeof(T));
  hipLaunchKernelGGL(truncate_kernel<T>, dim3(blocks_for(ctx->g.nc, TPB), nf), dim3(TPB), 0, ctx->stream,
                     p5<T>(fields, nf), ctx->g);
  KAPSEL_LAUNCH_CHECK();
}
void truncate(kapsel_ctx *ctx, void *const *fields, int nf) { KAPSEL_DISPATCH(ctx, truncate_t, ctx, fields, nf); }

template <typename T>
static void zeta_k2u_k_t(kapsel_ctx *ctx, int trunc, int fix_mask, void *const *out3) {
  LaunchTimer lt(ctx, KAPSEL_K_POINTWISE,
                 (5.0 * ctx->g.f + (trunc ? 4.0 * cut_count(ctx->g) : 0.0)) * sizeof(T));
  void *const *u = out3 ? out3 : ctx->fields + KAPSEL_U0;
  hipLaunchKernelGGL(zeta_k2u_k_kernel<T>, dim3(blocks_for(ctx->g.nc, TPB)), dim3(TPB), 0, ctx->stream,
                     F<T>(ctx, KAPSEL_ZETA0), F<T>(ctx, KAPSEL_ZETA1), (T *)u[0], (T *)u[1], (T *)u[2],
                     ctx->d_uk_dc, ctx->g, trunc, fix_mask);
  KAPSEL_LAUNCH_CHECK();
}
// fix_dc < 0: plain Zeta_k2u_k; otherwise truncate zeta first and apply the
// FIX_CELL mask fix_dc.
void zeta_k2u_k(kapsel_ctx *ctx, int fix_dc, void *const *out3) {
  int trunc = fix_dc >= 0;
  int mask = fix_dc >= 0 ? fix_dc : 0;
  KAPSEL_DISPATCH(ctx, zeta_k2u_k_t, ctx, trunc, mask, out3);
}

template <typename T>
static void u_k2zeta_k_t(kapsel_ctx *ctx) {
  LaunchTimer lt(ctx, KAPSEL_K_POINTWISE, 5.0 * ctx->g.f * sizeof(T));
  hipLaunchKernelGGL(u_k2zeta_k_kernel<T>, dim3(blocks_for(ctx->g.nc, TPB)), dim3(TPB), 0, ctx->stream,
                     F<T>(ctx, KAPSEL_U0), F<T>(ctx, KAPSEL_U1), F<T>(ctx, KAPSEL_U2), F<T>(ctx, KAPSEL_ZETA0),
                     F<T>(ctx, KAPSEL_ZETA1), ctx->d_uk_dc, ctx->g);
  KAPSEL_LAUNCH_CHECK();
}
void u_k2zeta_k(kapsel_ctx *ctx) { KAPSEL_DISPATCH(ctx, u_k2zeta_k_t, ctx); }

template <typename T>
static void sole_u2zeta_sum_t(kapsel_ctx *ctx, void *const *a3, void *const *f3) {
  LaunchTimer lt(ctx, KAPSEL_K_POINTWISE, (8.0 * ctx->g.f + 3.0 * 2 * kept_count(ctx->g)) * sizeof(T));
  hipLaunchKernelGGL(sole_u2zeta_sum_kernel<T>, dim3(blocks_for(ctx->g.nc, TPB)), dim3(TPB), 0, ctx->stream,
                     p5<T>(a3, 3), p5<T>(f3, 3), p5<T>(ctx->fields + KAPSEL_U0, 3), F<T>(ctx, KAPSEL_ZETA0),
                     F<T>(ctx, KAPSEL_ZETA1), ctx->d_uk_dc, ctx->g);
  KAPSEL_LAUNCH_CHECK();
}
void sole_u2zeta_sum(kapsel_ctx *ctx, void *const *a3, void *const *f3) {
  KAPSEL_DISPATCH(ctx, sole_u2zeta_sum_t, ctx, a3, f3);
}

template <typename T>
static void sole_add_t(kapsel_ctx *ctx, void *const *a3, void *const *f3, void *const *o3, double s) {
  LaunchTimer lt(ctx, KAPSEL_K_POINTWISE, (6.0 * ctx->g.f + 6.0 * kept_count(ctx->g)) * sizeof(T));
  hipLaunchKernelGGL(sole_add_kernel<T>, dim3(blocks_for(ctx->g.nc, TPB)), dim3(TPB), 0, ctx->stream,
                     p5<T>(a3, 3), p5<T>(f3, 3), p5<T>(o3, 3), (T)s, ctx->g);
  KAPSEL_LAUNCH_CHECK();
}
void sole_add(kapsel_ctx *ctx, void *const *a3, void *const *f3, void *const *o3, double s) {
  KAPSEL_DISPATCH(ctx, sole_add_t, ctx, a3, f3, o3, s);
}

template <typename T>
static void advection_products_t(kapsel_ctx *ctx) {
  LaunchTimer lt(ctx, KAPSEL_K_POINTWISE, 8.0 * ctx->g.nr * sizeof(T));
  hipLaunchKernelGGL(advection_products_kernel<T>, dim3(blocks_for(ctx->g.nr, TPB)), dim3(TPB), 0, ctx->stream,
                     F<T>(ctx, KAPSEL_U0), F<T>(ctx, KAPSEL_U1), F<T>(ctx, KAPSEL_U2), (T *)ctx->adv_w[0],
                     (T *)ctx->adv_w[1], ctx->g);
  KAPSEL_LAUNCH_CHECK();
}
void advection_products(kapsel_ctx *ctx) { KAPSEL_DISPATCH(ctx, advection_products_t, ctx); }

template <typename T>
static void advection_combine_slaved_t(kapsel_ctx *ctx, double dt_fluid, void *const *in5, int store_a) {
  double dmy0 = -ctx->nu * dt_fluid;
  double dmy1 = 1. / ctx->nu;
  void *const def[5] = {ctx->fields[KAPSEL_U0], ctx->fields[KAPSEL_U1], ctx->fields[KAPSEL_U2], ctx->adv_w[0],
                        ctx->adv_w[1]};
  void *const *p = in5 ? in5 : def;
  LaunchTimer lt(ctx, KAPSEL_K_POINTWISE, ((store_a ? 7.0 : 5.0) * ctx->g.f + 8.0 * kept_count(ctx->g)) * sizeof(T));
  hipLaunchKernelGGL((advection_combine_slaved_kernel<T, 0>), dim3(blocks_for(ctx->g.nc, TPB)), dim3(TPB), 0,
                     ctx->stream, (const T *)p[0], (const T *)p[1], (const T *)p[2], (T *)p[3], (T *)p[4],
                     F<T>(ctx, KAPSEL_ZETA0), F<T>(ctx, KAPSEL_ZETA1), ctx->g, dmy0, dmy1, store_a);
  KAPSEL_LAUNCH_CHECK();
}
void advection_combine_slaved(kapsel_ctx *ctx, double dt_fluid, void *const *in5, int store_a) {
  KAPSEL_DISPATCH(ctx, advection_combine_slaved_t, ctx, dt_fluid, in5, store_a);
}

template <typename T>
static void slaved_zeta2u_t(kapsel_ctx *ctx, double dt_fluid, void *const *in5, int fix_mask, void *const *out3) {
  void *const *u = out3 ? out3 : ctx->fields + KAPSEL_U0;
  LaunchTimer lt(ctx, KAPSEL_K_POINTWISE, (5.0 * ctx->g.f + 7.0 * 2.0 * kept_count(ctx->g)) * sizeof(T));
  hipLaunchKernelGGL(slaved_zeta2u_kernel<T>, dim3(blocks_for(ctx->g.nc, TPB)), dim3(TPB), 0, ctx->stream,
                     (const T *)in5[0], (const T *)in5[1], (const T *)in5[2], (const T *)in5[3], (const T *)in5[4],
                     F<T>(ctx, KAPSEL_ZETA0), F<T>(ctx, KAPSEL_ZETA1), p5<T>(u, 3), ctx->d_uk_dc, ctx->g,
                     -ctx->nu * dt_fluid, 1. / ctx->nu, fix_mask);
  KAPSEL_LAUNCH_CHECK();
}
void slaved_zeta2u(kapsel_ctx *ctx, double dt_fluid, void *const *in5, int fix_mask, void *const *out3) {
  KAPSEL_DISPATCH(ctx, slaved_zeta2u_t, ctx, dt_fluid, in5, fix_mask, out3);
}

template <typename T>
static void advection_combine_euler_t(kapsel_ctx *ctx, double dt_fluid) {
  LaunchTimer lt(ctx, KAPSEL_K_POINTWISE, (7.0 * ctx->g.f + 4.0 * kept_count(ctx->g)) * sizeof(T));
  hipLaunchKernelGGL((advection_combine_slaved_kernel<T, 1>), dim3(blocks_for(ctx->g.nc, TPB)), dim3(TPB), 0,
                     ctx->stream, F<T>(ctx, KAPSEL_U0), F<T>(ctx, KAPSEL_U1), F<T>(ctx, KAPSEL_U2),
                     (T *)ctx->adv_w[0], (T *)ctx->adv_w[1], F<T>(ctx, KAPSEL_ZETA0),
                     F<T>(ctx, KAPSEL_ZETA1), ctx->g, ctx->nu, dt_fluid);
  KAPSEL_LAUNCH_CHECK();
}
void advection_combine_euler(kapsel_ctx *ctx, double dt_fluid) {
  KAPSEL_DISPATCH(ctx, advection_combine_euler_t, ctx, dt_fluid);
}

template <typename T>
static void solenoidal_uk_t(kapsel_ctx *ctx, void *const *f3, void *const *out3) {
  LaunchTimer lt(ctx, KAPSEL_K_POINTWISE, (3.0 * ctx->g.f + 6.0 * kept_count(ctx->g)) * sizeof(T));
  hipLaunchKernelGGL(solenoidal_uk_kernel<T>, dim3(blocks_for(ctx->g.nc, TPB)), dim3(TPB), 0, ctx->stream,
                     p5<T>(f3, 3), p5<T>(out3 ? out3 : f3, 3), ctx->g, 1);
  KAPSEL_LAUNCH_CHECK();
}
// truncation followed by the projection, as Solenoidal_u does in k-space
void solenoidal_uk(kapsel_ctx *ctx, void *const *f3, void *const *out3) {
  KAPSEL_DISPATCH(ctx, solenoidal_uk_t, ctx, f3, out3);
}

template <typename T>
static void make_f_nonsole_t(kapsel_ctx *ctx) {
  LaunchTimer lt(ctx, KAPSEL_K_POINTWISE, 10.0 * ctx->g.nr * sizeof(T));
  hipLaunchKernelGGL(f_nonsole_kernel<T>, dim3(blocks_for(ctx->g.nr, TPB)), dim3(TPB), 0, ctx->stream,
                     p5<T>(ctx->fields + KAPSEL_U0, 3), p5<T>(ctx->fields + KAPSEL_UP0, 3),
                     p5<T>(ctx->fields + KAPSEL_F0, 3), F<T>(ctx, KAPSEL_PHI), ctx->g);
  KAPSEL_LAUNCH_CHECK();
}
void make_f_nonsole(kapsel_ctx *ctx) { KAPSEL_DISPATCH(ctx, make_f_nonsole_t, ctx); }

template <typename T>
static void add_f_particle_t(kapsel_ctx *ctx, void *const *keep3) {
  LaunchTimer lt(ctx, KAPSEL_K_POINTWISE, (keep3 ? 12.0 : 9.0) * ctx->g.nr * sizeof(T));
  hipLaunchKernelGGL(add_f_kernel<T>, dim3(blocks_for(ctx->g.nr, TPB)), dim3(TPB), 0, ctx->stream,
                     p5<T>(ctx->fields + KAPSEL_U0, 3), p5<T>(ctx->fields + KAPSEL_F0, 3),
                     keep3 ? p5<T>(keep3, 3) : P5<T>{}, ctx->g);
  KAPSEL_LAUNCH_CHECK();
}
void add_f_particle(kapsel_ctx *ctx, void *const *keep3) { KAPSEL_DISPATCH(ctx, add_f_particle_t, ctx, keep3); }

template <typename T>
__global__ void copy_kernel(T *__restrict__ dst, const T *__restrict__ src, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

// Copy_v1_k (aux_field.h:82): the whole padded field, a plain streaming copy
void copy_field(kapsel_ctx *ctx, void *dst, const void *src) {
  LaunchTimer lt(ctx, KAPSEL_K_POINTWISE, 2.0 * ctx->g.f * ctx->elem);
  if (ctx->prec == KAPSEL_FP64)
    hipLaunchKernelGGL(copy_kernel<double>, dim3(blocks_for(ctx->g.f, TPB)), dim3(TPB), 0, ctx->stream,
                       (double *)dst, (const double *)src, ctx->g.f);
  else
    hipLaunchKernelGGL(copy_kernel<float>, dim3(blocks_for(ctx->g.f, TPB)), dim3(TPB), 0, ctx->stream,
                       (float *)dst, (const float *)src, ctx->g.f);
  KAPSEL_LAUNCH_CHECK();
}

void copy_convert_in(kapsel_ctx *ctx, void *dst, const double *src_dev) {
  hipLaunchKernelGGL((convert_kernel<float, double>), dim3(blocks_for(ctx->g.f, TPB)), dim3(TPB), 0, ctx->stream,
                     (float *)dst, src_dev, ctx->g.f);
  KAPSEL_LAUNCH_CHECK();
}
void copy_convert_out(kapsel_ctx *ctx, double *dst_dev, const void *src) {
  hipLaunchKernelGGL((convert_kernel<double, float>), dim3(blocks_for(ctx->g.f, TPB)), dim3(TPB), 0, ctx->stream,
                     dst_dev, (const float *)src, ctx->g.f);
  KAPSEL_LAUNCH_CHECK();
}

}  // namespace kapsel
