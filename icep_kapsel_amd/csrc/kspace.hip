// kspace.hip -- pointwise k-space and real-space kernels of the NS step.
//
// k-space kernels run one thread per complex value of the half spectrum
// (the reference loops over the NZ_ reals of a row with identical real and
// imaginary coefficients, so both parts are handled by one thread).  The
// wavevector is computed from the index (Calc_K*_Ooura, fft_wrapper.h:
// 761-773) instead of being read from the KX/KY/KZ/K2/IK2 tables of Init_K
// (fft_wrapper.cxx:47-87): the same values, five fewer streams from HBM.
#include "common.hpp"

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
  T kx = (T)(g.wave_x * k.kx), ky = (T)(g.wave_y * k.ky), kz = (T)(g.wave_z * k.kz);
  T k2 = kx * kx + ky * ky + kz * kz;
  T ik2 = k2 > (T)0 ? (T)1 / k2 : (T)0;
  V2<T> o0, o1, o2;
  if (k.kx != 0) {
    T ikx = (T)1 / kx;
    o0 = V2<T>{-ikx * (ky * d1.re + kz * d2.re), -ikx * (ky * d1.im + kz * d2.im)};
    o1 = d1;
    o2 = d2;
  } else if (k.ky != 0) {
    T dmy = -(kz / ky);
    o0 = d2;
    o1 = scal(dmy, d1);
    o2 = d1;
  } else {
    o0 = d1;
    o1 = d2;
    o2 = V2<T>{0, 0};
  }
  kx *= ik2;
  ky *= ik2;
  kz *= ik2;
  V2<T> a = {-(ky * o2.im - kz * o1.im), (ky * o2.re - kz * o1.re)};
  V2<T> b = {-(kz * o0.im - kx * o2.im), (kz * o0.re - kx * o2.re)};
  V2<T> e = {-(kx * o1.im - ky * o0.im), (kx * o1.re - ky * o0.re)};
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

// U_k2zeta_k (fft_wrapper.cxx:428-497); with p, u is the sum of the two
// spectra u + p, which is also stored in o
template <typename T>
__global__ void u_k2zeta_k_kernel(const T *__restrict__ u0, const T *__restrict__ u1, const T *__restrict__ u2,
                                  T *__restrict__ z0, T *__restrict__ z1, double *__restrict__ dc, Grid g,
                                  const T *__restrict__ p0 = nullptr, const T *__restrict__ p1 = nullptr,
                                  const T *__restrict__ p2 = nullptr, T *__restrict__ o0 = nullptr,
                                  T *__restrict__ o1 = nullptr, T *__restrict__ o2 = nullptr) {
  size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.nc) return;
  KIdx k = kidx(g, c);
  V2<T> a = ld2(u0 + 2 * c), b = ld2(u1 + 2 * c), e = ld2(u2 + 2 * c);
  if (p0) {
    a = a + ld2(p0 + 2 * c);
    b = b + ld2(p1 + 2 * c);
    e = e + ld2(p2 + 2 * c);
    st2(o0 + 2 * c, a);
    st2(o1 + 2 * c, b);
    st2(o2 + 2 * c, e);
  }
  if (c == 0) {
    dc[0] = (double)a.re;
    dc[1] = (double)b.re;
    dc[2] = (double)e.re;
  }
  T ks0 = (T)(k.kx * g.wave_x), ks1 = (T)(k.ky * g.wave_y), ks2 = (T)(k.kz * g.wave_z);
  V2<T> w0 = {-(ks1 * e.im - ks2 * b.im), (ks1 * e.re - ks2 * b.re)};
  V2<T> w1 = {-(ks2 * a.im - ks0 * e.im), (ks2 * a.re - ks0 * e.re)};
  V2<T> w2 = {-(ks0 * b.im - ks1 * a.im), (ks0 * b.re - ks1 * a.re)};
  V2<T> za, zb;
  if (k.kx != 0) {
    za = w1;
    zb = w2;
  } else if (k.ky != 0) {
    za = w2;
    zb = w0;
  } else {
    za = w0;
    zb = w1;
  }
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
// global f_ns0 (slaved); the Euler variant leaves the full right-hand side
// there, as Rhs_NS_solute leaves it in f_ns1.
template <typename T, int EULER>
__global__ void advection_combine_slaved_kernel(const T *__restrict__ u0, const T *__restrict__ u1,
                                                const T *__restrict__ u2, T *__restrict__ a0,
                                                T *__restrict__ a1, T *__restrict__ z0, T *__restrict__ z1,
                                                Grid g, double dmy0, double dmy1) {
  size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.nc) return;
  KIdx k = kidx(g, c);
  const T k1 = (T)(k.kx * g.wave_x), k2 = (T)(k.ky * g.wave_y), k3 = (T)(k.kz * g.wave_z);
  V2<T> u2u3 = ld2(u0 + 2 * c), u3u1 = ld2(u1 + 2 * c), u1u2 = ld2(u2 + 2 * c);
  V2<T> u22_u32 = ld2(a0 + 2 * c), u32_u12 = ld2(a1 + 2 * c);
  V2<T> u12_u22 = V2<T>{-u32_u12.re - u22_u32.re, -u32_u12.im - u22_u32.im};
  const T k1k2 = k1 * k2, k2k3 = k2 * k3, k3k1 = k3 * k1;
  const T c0 = (k3 - k2) * (k3 + k2), c1 = (k1 - k3) * (k1 + k3), c2 = (k2 - k1) * (k2 + k1);
  V2<T> w[3];
#define KAPSEL_OMEGA(out, p)                                                             \
  out[0].p = -(k2k3 * u22_u32.p + c0 * u2u3.p + k3k1 * u1u2.p - k1k2 * u3u1.p);          \
  out[1].p = -(k3k1 * u32_u12.p + c1 * u3u1.p + k1k2 * u2u3.p - k2k3 * u1u2.p);          \
  out[2].p = -(k1k2 * u12_u22.p + c2 * u1u2.p + k2k3 * u3u1.p - k3k1 * u2u3.p);
  KAPSEL_OMEGA(w, re)
  KAPSEL_OMEGA(w, im)
#undef KAPSEL_OMEGA
  V2<T> f0, f1;
  if (k.kx != 0) {
    f0 = w[1];
    f1 = w[2];
  } else if (k.ky != 0) {
    f0 = w[2];
    f1 = w[0];
  } else {
    f0 = w[0];
    f1 = w[1];
  }
  if constexpr (EULER) {
    if (!is_cut(g, k)) {
      const double q1 = k.kx * g.wave_x, q2 = k.ky * g.wave_y, q3 = k.kz * g.wave_z;
      const T nk2 = (T)(dmy0 * (q1 * q1 + q2 * q2 + q3 * q3));  // NU * K2[im]
      const T dt = (T)dmy1;
      V2<T> za = ld2(z0 + 2 * c), zb = ld2(z1 + 2 * c);
      f0 = V2<T>{f0.re + -(nk2 * za.re), f0.im + -(nk2 * za.im)};
      f1 = V2<T>{f1.re + -(nk2 * zb.re), f1.im + -(nk2 * zb.im)};
      st2(z0 + 2 * c, V2<T>{za.re + dt * f0.re, za.im + dt * f0.im});
      st2(z1 + 2 * c, V2<T>{zb.re + dt * f1.re, zb.im + dt * f1.im});
    } else {
      // Rhs_NS_solute cut zeta before (fluid_solver.cxx:201-203)
      st2(z0 + 2 * c, V2<T>{0, 0});
      st2(z1 + 2 * c, V2<T>{0, 0});
    }
    st2(a0 + 2 * c, f0);
    st2(a1 + 2 * c, f1);
    return;
  }
  st2(a0 + 2 * c, f0);
  st2(a1 + 2 * c, f1);
  if (!is_cut(g, k)) {
    const double q1 = k.kx * g.wave_x, q2 = k.ky * g.wave_y, q3 = k.kz * g.wave_z;
    const double K2 = q1 * q1 + q2 * q2 + q3 * q3;  // K2 of Init_K (fft_wrapper.cxx:71-73)
    const double ik2 = K2 > 0.0 ? 1. / K2 : 0.0;
    // the reference exponent and factors are double; the update is done in T
    const T dmy = (T)(exp(dmy0 * K2) - 1.);
    const T s = (T)(dmy1 * ik2);
    V2<T> za = ld2(z0 + 2 * c), zb = ld2(z1 + 2 * c);
    za = V2<T>{za.re + dmy * (za.re - s * f0.re), za.im + dmy * (za.im - s * f0.im)};
    zb = V2<T>{zb.re + dmy * (zb.re - s * f1.re), zb.im + dmy * (zb.im - s * f1.im)};
    st2(z0 + 2 * c, za);
    st2(z1 + 2 * c, zb);
  }
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
  const T kx = (T)(k.kx * g.wave_x), ky = (T)(k.ky * g.wave_y), kz = (T)(k.kz * g.wave_z);
  const T k2 = kx * kx + ky * ky + kz * kz;
  const T ik2 = k2 > (T)0 ? (T)1 / k2 : (T)0;
  V2<T> a = ld2(u.p[0] + 2 * c), b = ld2(u.p[1] + 2 * c), e = ld2(u.p[2] + 2 * c);
  V2<T> dmy = {ik2 * (a.re * kx + b.re * ky + e.re * kz), ik2 * (a.im * kx + b.im * ky + e.im * kz)};
  st2(o.p[0] + 2 * c, a - scal(kx, dmy));
  st2(o.p[1] + 2 * c, b - scal(ky, dmy));
  st2(o.p[2] + 2 * c, e - scal(kz, dmy));
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
static void u_k2zeta_sum_k_t(kapsel_ctx *ctx, void *const *a3, void *const *b3) {
  LaunchTimer lt(ctx, KAPSEL_K_POINTWISE, 11.0 * ctx->g.f * sizeof(T));
  hipLaunchKernelGGL(u_k2zeta_k_kernel<T>, dim3(blocks_for(ctx->g.nc, TPB)), dim3(TPB), 0, ctx->stream,
                     (const T *)a3[0], (const T *)a3[1], (const T *)a3[2], F<T>(ctx, KAPSEL_ZETA0),
                     F<T>(ctx, KAPSEL_ZETA1), ctx->d_uk_dc, ctx->g, (const T *)b3[0], (const T *)b3[1],
                     (const T *)b3[2], F<T>(ctx, KAPSEL_U0), F<T>(ctx, KAPSEL_U1), F<T>(ctx, KAPSEL_U2));
  KAPSEL_LAUNCH_CHECK();
}
void u_k2zeta_sum_k(kapsel_ctx *ctx, void *const *a3, void *const *b3) {
  KAPSEL_DISPATCH(ctx, u_k2zeta_sum_k_t, ctx, a3, b3);
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
static void advection_combine_slaved_t(kapsel_ctx *ctx, double dt_fluid) {
  double dmy0 = -ctx->nu * dt_fluid;
  double dmy1 = 1. / ctx->nu;
  LaunchTimer lt(ctx, KAPSEL_K_POINTWISE, (7.0 * ctx->g.f + 8.0 * kept_count(ctx->g)) * sizeof(T));
  hipLaunchKernelGGL((advection_combine_slaved_kernel<T, 0>), dim3(blocks_for(ctx->g.nc, TPB)), dim3(TPB), 0,
                     ctx->stream, F<T>(ctx, KAPSEL_U0), F<T>(ctx, KAPSEL_U1), F<T>(ctx, KAPSEL_U2),
                     (T *)ctx->adv_w[0], (T *)ctx->adv_w[1], F<T>(ctx, KAPSEL_ZETA0),
                     F<T>(ctx, KAPSEL_ZETA1), ctx->g, dmy0, dmy1);
  KAPSEL_LAUNCH_CHECK();
}
void advection_combine_slaved(kapsel_ctx *ctx, double dt_fluid) {
  KAPSEL_DISPATCH(ctx, advection_combine_slaved_t, ctx, dt_fluid);
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
