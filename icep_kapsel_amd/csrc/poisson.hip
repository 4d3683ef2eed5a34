// poisson.hip -- variable-permittivity Poisson equation of the ICEP model.
//
// Operator (Ax, operate_dielectric.cxx:485-522): A x = D E G x with
//   G x = IFFT(i k e^{+i k dx/2} x_k)   staggered gradient (A_k2da_k +
//                                        Shift_vec_fw_imag, fft_wrapper.h:751-858)
//   E   = face permittivity 0.5 (eps_i + eps_{i+e_d})
//   D v = IFFT(i k . e^{-i k dx/2} v_k)  staggered divergence
//                                        (U_k2divergence_k_shift, fft_wrapper.h:879-900)
// 8 real 3-D transforms per application.  A is symmetric negative
// semidefinite with the constants as null space, so conjugate gradients run
// on S = -A; the reference runs Eigen's BiCGSTAB with the identity
// preconditioner (operate_dielectric.cxx:556-562) and stops on
// ||b - A x|| / ||b|| < tol, the criterion kept here.
//
// Preconditioner (optional): M^{-1} r = IFFT(r_k / (eps_mean |k|^2)), the
// inverse of S for a uniform permittivity; it bounds the condition number of
// M^{-1} S by eps_max / eps_min independently of the grid.
//
// Precision: an FP64 context solves in FP64.  An FP32 context refines in
// FP64: the residual b - A x and the potential are FP64 (shadow fields),
// each correction is an FP32 (preconditioned) CG solve, so the reported
// residual is an FP64 residual.  All reductions accumulate in FP64.
#include <cmath>
#include <functional>
#include <utility>

#include "common.hpp"

namespace kapsel {

namespace {

enum { S_BB = 0, S_RZ, S_PAP, S_RR, S_RZN, S_BETA, S_TMP, S_INVEPS, S_M1, S_N };  // S_M1: -1
constexpr int NPART = 1024;  // blocks of the partial-sum reductions (fixed: deterministic sums)
constexpr unsigned TPB = 256;

// extract_id (periodic_boundary.h:21-27) for +1 neighbours
__device__ __forceinline__ size_t nbr(const Grid &g, int i, int j, int k) {
  i = i == g.nx ? 0 : i;
  j = j == g.ny ? 0 : j;
  k = k == g.nz ? 0 : k;
  return ((size_t)i * g.ny + j) * g.nz_ + k;
}

// extract_id for -1 neighbours
__device__ __forceinline__ size_t nbm(const Grid &g, int i, int j, int k) {
  i = i < 0 ? g.nx - 1 : i;
  j = j < 0 ? g.ny - 1 : j;
  k = k < 0 ? g.nz - 1 : k;
  return ((size_t)i * g.ny + j) * g.nz_ + k;
}

// E^2 at the cell centres from the face values E_d = -dpsi_d + E0_d
// (operate_dielectric.cxx:676-702)
template <typename T>
__global__ void efield2_kernel(const T *__restrict__ g0, const T *__restrict__ g1, const T *__restrict__ g2,
                               double e0x, double e0y, double e0z, T *__restrict__ e2, Grid g) {
  size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= g.nr) return;
  const RI p = ridx(g, r);
  const T fx = -g0[p.im] + (T)e0x, fy = -g1[p.im] + (T)e0y, fz = -g2[p.im] + (T)e0z;
  const T bx = -g0[nbm(g, p.i - 1, p.j, p.k)] + (T)e0x;
  const T by = -g1[nbm(g, p.i, p.j - 1, p.k)] + (T)e0y;
  const T bz = -g2[nbm(g, p.i, p.j, p.k - 1)] + (T)e0z;
  const T ax = (fx + bx) / (T)2, ay = (fy + by) / (T)2, az = (fz + bz) / (T)2;
  e2[p.im] = ax * ax + ay * ay + az * az;
}

// face force rho_f E - E2_f grad(eps) / 2 (operate_dielectric.cxx:704-757)
// averaged back to the cell centre (Interpolate_vec_on_normal_grid, :32-55)
template <typename T>
__device__ __forceinline__ T face_force(const T *gpsi, const T *geps, const T *rho, const T *e2, double e0d,
                                        size_t at, size_t up) {
  const T efw = -gpsi[at] + (T)e0d;
  const T rho_f = (T)0.5 * rho[up] + (T)0.5 * rho[at];
  const T e2f = (T)0.5 * e2[at] + (T)0.5 * e2[up];
  return rho_f * efw - e2f * geps[at] / (T)2;
}

template <typename T>
__global__ void maxwell_kernel(const T *__restrict__ p0, const T *__restrict__ p1, const T *__restrict__ p2,
                               const T *__restrict__ q0, const T *__restrict__ q1, const T *__restrict__ q2,
                               const T *__restrict__ rho, const T *__restrict__ e2, double e0x, double e0y,
                               double e0z, T *__restrict__ f0, T *__restrict__ f1, T *__restrict__ f2, Grid g) {
  size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= g.nr) return;
  const RI p = ridx(g, r);
  const size_t xm = nbm(g, p.i - 1, p.j, p.k), ym = nbm(g, p.i, p.j - 1, p.k), zm = nbm(g, p.i, p.j, p.k - 1);
  const size_t xp = nbr(g, p.i + 1, p.j, p.k), yp = nbr(g, p.i, p.j + 1, p.k), zp = nbr(g, p.i, p.j, p.k + 1);
  f0[p.im] = (T)0.5 * face_force(p0, q0, rho, e2, e0x, p.im, xp) + (T)0.5 * face_force(p0, q0, rho, e2, e0x, xm, p.im);
  f1[p.im] = (T)0.5 * face_force(p1, q1, rho, e2, e0y, p.im, yp) + (T)0.5 * face_force(p1, q1, rho, e2, e0y, ym, p.im);
  f2[p.im] = (T)0.5 * face_force(p2, q2, rho, e2, e0z, p.im, zp) + (T)0.5 * face_force(p2, q2, rho, e2, e0z, zm, p.im);
}

// u += f s (s = dt / rho, sp_3d_ns.cxx:83-93)
template <typename T>
__global__ void add_force_kernel(T *__restrict__ u0, T *__restrict__ u1, T *__restrict__ u2,
                                 const T *__restrict__ f0, const T *__restrict__ f1, const T *__restrict__ f2,
                                 double s, Grid g) {
  size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= g.nr) return;
  const RI p = ridx(g, r);
  u0[p.im] += f0[p.im] * (T)s;
  u1[p.im] += f1[p.im] * (T)s;
  u2[p.im] += f2[p.im] * (T)s;
}

template <typename T>
__global__ void copy_real_kernel(const T *__restrict__ x, T *__restrict__ y, Grid g) {
  size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= g.f) return;
  int k = (int)(e % (unsigned)g.nz_);
  y[e] = k < g.nz ? x[e] : (T)0;
}

// g_d = i k_d x_k e^{i k_d dx/2}
template <typename T>
__global__ void grad_shift_kernel(const T *__restrict__ xk, T *__restrict__ g0, T *__restrict__ g1,
                                  T *__restrict__ g2, const V2<T> *__restrict__ shx, const V2<T> *__restrict__ shy,
                                  const V2<T> *__restrict__ shz, Grid g) {
  size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.nc) return;
  KIdx k = kidx(g, c);
  V2<T> x = ld2(xk + 2 * c);
  const T w0 = (T)(k.kx * g.wave_x), w1 = (T)(k.ky * g.wave_y), w2 = (T)(k.kz * g.wave_z);
  st2(g0 + 2 * c, cmul(V2<T>{-w0 * x.im, w0 * x.re}, shx[k.i]));
  st2(g1 + 2 * c, cmul(V2<T>{-w1 * x.im, w1 * x.re}, shy[k.j]));
  st2(g2 + 2 * c, cmul(V2<T>{-w2 * x.im, w2 * x.re}, shz[k.kz]));
}

// g_d = i k_d e^{i k_d dx/2} x_k for two FP64 half spectra a, b into six
// FP32 spectra (the staggered gradients of eps and of the potential from the
// spectra the FP64 refinement holds)
__global__ void grad_shift2_f64_kernel(const double *__restrict__ a, const double *__restrict__ b,
                                       float *__restrict__ a0, float *__restrict__ a1, float *__restrict__ a2,
                                       float *__restrict__ b0, float *__restrict__ b1, float *__restrict__ b2,
                                       const V2<double> *__restrict__ shx, const V2<double> *__restrict__ shy,
                                       const V2<double> *__restrict__ shz, Grid g) {
  size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.nc) return;
  KIdx k = kidx(g, c);
  const double w[3] = {k.kx * g.wave_x, k.ky * g.wave_y, k.kz * g.wave_z};
  const V2<double> s[3] = {shx[k.i], shy[k.j], shz[k.kz]};
  const V2<double> x = ld2(a + 2 * c), y = ld2(b + 2 * c);
  float *oa[3] = {a0, a1, a2}, *ob[3] = {b0, b1, b2};
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const V2<double> ga = cmul(V2<double>{-w[d] * x.im, w[d] * x.re}, s[d]);
    const V2<double> gb = cmul(V2<double>{-w[d] * y.im, w[d] * y.re}, s[d]);
    st2(oa[d] + 2 * c, V2<float>{(float)ga.re, (float)ga.im});
    st2(ob[d] + 2 * c, V2<float>{(float)gb.re, (float)gb.im});
  }
}

// f_d = 0.5 (eps_i + eps_{i+e_d}) at the real points, 0 in the row padding:
// the face permittivities of the operator (operate_dielectric.cxx:492-508),
// which the plane pass multiplies in
template <typename T>
__global__ void face_fields_kernel(const T *__restrict__ eps, T *__restrict__ f0, T *__restrict__ f1,
                                   T *__restrict__ f2, Grid g) {
  size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= g.f) return;
  const int k = (int)(e % (unsigned)g.nz_);
  if (k >= g.nz) {
    f0[e] = f1[e] = f2[e] = (T)0;
    return;
  }
  const size_t row = e / (unsigned)g.nz_;
  const int j = (int)(row % (unsigned)g.ny), i = (int)(row / (unsigned)g.ny);
  const T en = eps[e];
  f0[e] = (T)0.5 * (eps[nbr(g, i + 1, j, k)] + en);
  f1[e] = (T)0.5 * (eps[nbr(g, i, j + 1, k)] + en);
  f2[e] = (T)0.5 * (eps[nbr(g, i, j, k + 1)] + en);
}

// out = i sum_d k_d conj(shift_d) w_d
template <typename T>
__global__ void div_shift_kernel(const T *__restrict__ w0, const T *__restrict__ w1, const T *__restrict__ w2,
                                 T *__restrict__ out, const V2<T> *__restrict__ shx,
                                 const V2<T> *__restrict__ shy, const V2<T> *__restrict__ shz, Grid g) {
  size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.nc) return;
  KIdx k = kidx(g, c);
  const T q0 = (T)(k.kx * g.wave_x), q1 = (T)(k.ky * g.wave_y), q2 = (T)(k.kz * g.wave_z);
  V2<T> a = cmul(ld2(w0 + 2 * c), scal(q0, conjv(shx[k.i])));
  a = a + cmul(ld2(w1 + 2 * c), scal(q1, conjv(shy[k.j])));
  a = a + cmul(ld2(w2 + 2 * c), scal(q2, conjv(shz[k.kz])));
  st2(out + 2 * c, V2<T>{-a.im, a.re});
}

// Interpolate_scalar_on_staggered_grid (operate_dielectric.cxx:57-76)
template <typename T>
__global__ void interp_staggered_kernel(T *__restrict__ v0, T *__restrict__ v1, T *__restrict__ v2,
                                        const T *__restrict__ eps, double f0, double f1, double f2, Grid g) {
  size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= g.nr) return;
  RI p = ridx(g, r);
  const T sn = (T)0.5 * eps[p.im];
  v0[p.im] = ((T)0.5 * eps[nbr(g, p.i + 1, p.j, p.k)] + sn) * (T)f0;
  v1[p.im] = ((T)0.5 * eps[nbr(g, p.i, p.j + 1, p.k)] + sn) * (T)f1;
  v2[p.im] = ((T)0.5 * eps[nbr(g, p.i, p.j, p.k + 1)] + sn) * (T)f2;
}

// a -= b over the real points; with bg, a -= b - bg[0] (a uniform background
// charge that makes the periodic problem solvable)
template <typename T>
__global__ void sub_kernel(T *__restrict__ a, const T *__restrict__ b, const double *__restrict__ bg, Grid g) {
  size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= g.nr) return;
  RI p = ridx(g, r);
  if (bg) a[p.im] -= b[p.im] - (T)bg[0];
  else a[p.im] -= b[p.im];
}

__device__ __forceinline__ double block_sum(double v) {
  __shared__ double red[TPB / 64];
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
  const int lane = threadIdx.x % 64, wid = threadIdx.x / 64;
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
    for (int w = 0; w < (int)(TPB / 64); w++) s += red[w];
  }
  __syncthreads();
  return s;
}

// The refinement's set-up in one pass (FP32 context): FP64 copies of eps
// and rho (rho NULL: known zero, not copied), the face permittivities 0.5 (eps_i + eps_{i+e_d}) of the FP32
// correction operator and of the FP64 residual operator (0 in the row
// padding, which the plane pass multiplies too), and per-block sums of eps,
// whose mean scales the preconditioner
__global__ __launch_bounds__(TPB) void refine_setup_kernel(const float *__restrict__ eps, const float *__restrict__ rho,
                                                           double *__restrict__ eps64, double *__restrict__ rho64,
                                                           float *__restrict__ a0, float *__restrict__ a1,
                                                           float *__restrict__ a2, double *__restrict__ b0,
                                                           double *__restrict__ b1, double *__restrict__ b2,
                                                           double *__restrict__ partial, Grid g) {
  double s = 0.0;
  for (size_t e = (size_t)blockIdx.x * TPB + threadIdx.x; e < g.f; e += (size_t)gridDim.x * TPB) {
    const int k = (int)(e % (unsigned)g.nz_);
    if (k >= g.nz) {
      a0[e] = a1[e] = a2[e] = 0.f;
      b0[e] = b1[e] = b2[e] = 0.0;
      continue;
    }
    const size_t row = e / (unsigned)g.nz_;
    const int j = (int)(row % (unsigned)g.ny), i = (int)(row / (unsigned)g.ny);
    const float en = eps[e];
    const float ex = eps[nbr(g, i + 1, j, k)], ey = eps[nbr(g, i, j + 1, k)], ez = eps[nbr(g, i, j, k + 1)];
    eps64[e] = (double)en;
    if (rho) rho64[e] = (double)rho[e];
    a0[e] = 0.5f * (ex + en);
    a1[e] = 0.5f * (ey + en);
    a2[e] = 0.5f * (ez + en);
    b0[e] = 0.5 * ((double)ex + (double)en);
    b1[e] = 0.5 * ((double)ey + (double)en);
    b2[e] = 0.5 * ((double)ez + (double)en);
    s += (double)en;
  }
  s = block_sum(s);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// partial[b] = sum over block b's real points of a*b (FP64)
template <typename T, typename U>
__global__ __launch_bounds__(TPB) void dot_partial_kernel(const T *__restrict__ a, const U *__restrict__ b,
                                                          double *__restrict__ partial, Grid g) {
  double s = 0.0;
  for (size_t r = (size_t)blockIdx.x * TPB + threadIdx.x; r < g.nr; r += (size_t)gridDim.x * TPB) {
    RI p = ridx(g, r);
    s += (double)a[p.im] * (double)b[p.im];
  }
  s = block_sum(s);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// Second stage of every reduction (a separate launch: a device-wide fence
// per block to let the last block finish costs more than the launch):
// scal[slot] = sum(partial) in block order; mode 1 also sets
// beta = new / scal[S_RZ] and scal[S_RZ] = new (the rz update of CG)
// with host (mapped pinned host memory): the sum goes to host[hslot] as
// well; with cnt, the store is counted in cnt[1] and the count stored in
// host[5] after it (the refinement polls that)
__global__ __launch_bounds__(TPB) void finish_kernel(const double *__restrict__ partial, int n,
                                                     double *__restrict__ scal, int slot, int mode,
                                                     double *host = nullptr, int hslot = 0,
                                                     unsigned long long *cnt = nullptr, double count = 0.0) {
  // in rounds of four partials per thread, their loads issued together (the
  // sum order is that of one partial per thread and round)
  double s = 0.0;
  for (int i0 = 0; i0 < n; i0 += 4 * TPB) {
    double v[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int i = i0 + threadIdx.x + k * TPB;
      v[k] = i < n ? partial[i] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < 4; k++) s += v[k];
  }
  s = block_sum(s);
  if (threadIdx.x == 0) {
    if (mode == 1) {
      const double old = scal[S_RZ];
      scal[S_BETA] = old != 0.0 ? s / old : 0.0;
      scal[S_RZ] = s;
    }
    if (mode == 2) {
      scal[S_INVEPS] = s > 0.0 ? count / s : 1.0;  // 1 / mean of a sum over count points
      scal[S_M1] = -1.0;                            // the refinement's extrapolation factor
    }
    scal[slot] = s;
    if (host) {
      host[hslot] = s;
      __threadfence_system();
      if (cnt) {
        const unsigned long long k = cnt[1] + 1ull;
        cnt[1] = k;
        host[5] = (double)k;
        __threadfence_system();
      }
    }
  }
}

// two sums at once: scal[S_RR] = <r,r> (partial[NPART..]); scal[S_RZ] =
// <r,z>; mode 1 also beta = <r,z> / old scal[S_RZ] (mode 0: beta = 0).
// With host (pinned host memory, mapped), mode 1 counts the iteration in
// cnt[0] and stores <r,r> in host[count % 2], then the count in host[4],
// which the pipelined convergence test polls.
__global__ __launch_bounds__(TPB) void finish2_kernel(const double *__restrict__ partial, double *__restrict__ scal,
                                                      int mode, double *host = nullptr,
                                                      unsigned long long *cnt = nullptr) {
  static_assert(NPART % TPB == 0, "whole rounds");
  double pa[NPART / TPB], pb[NPART / TPB];
#pragma unroll
  for (int k = 0; k < NPART / TPB; k++) {  // all loads issued before the sums
    pa[k] = partial[threadIdx.x + k * TPB];
    pb[k] = partial[NPART + threadIdx.x + k * TPB];
  }
  double a = 0.0, b = 0.0;
#pragma unroll
  for (int k = 0; k < NPART / TPB; k++) {
    a += pa[k];
    b += pb[k];
  }
  a = block_sum(a);
  b = block_sum(b);
  if (threadIdx.x == 0) {
    const double old = scal[S_RZ];
    scal[S_BETA] = (mode == 1 && old != 0.0) ? a / old : 0.0;
    scal[S_RZ] = a;
    scal[S_RR] = b;
    if (host && mode == 1) {
      const unsigned long long k = cnt[0] + 1ull;
      cnt[0] = k;
      host[k & 1ull] = b;
      __threadfence_system();
      host[4] = (double)k;
      __threadfence_system();
    }
  }
}

template <typename D, typename S>
__global__ void convert_real_kernel(D *__restrict__ d, const S *__restrict__ s, Grid g) {
  size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= g.f) return;
  d[e] = (D)s[e];
}

// x64 -= (double) d32   (r = A x - b, so the correction enters with a minus sign)
template <typename T>
__global__ void add_correction_kernel(double *__restrict__ x, const T *__restrict__ d, Grid g) {
  size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= g.nr) return;
  RI pt = ridx(g, q);
  x[pt.im] -= (double)d[pt.im];
}

// ---------------------------------------------------------- k-space PCG
// The FP32 correction solve keeps every CG vector as the half spectrum of a
// real field.  A maps spectra to spectra with 6 real 3-D transforms (the
// gradient's 3 inverse and the flux's 3 forward transforms); the
// preconditioner is a pointwise scale; dot products follow from Parseval:
//   sum_x a(x) b(x) = N^-3 sum_c w(kz) Re(conj(a_c) b_c),
// w = 1 on the kz = 0 and kz = NZ/2 planes (their conjugate partners are
// stored) and 2 elsewhere.  The sums below omit the N^-3 (it cancels in
// alpha and beta; the host applies it to ||r||^2).
__device__ __forceinline__ double kweight(const Grid &g, int kz) {
  return (kz == 0 || 2 * kz == g.nz) ? 1.0 : 2.0;
}

template <typename T>
__device__ __forceinline__ T precond_factor(const Grid &g, const KIdx &k, double inv_eps) {
  const double q0 = k.kx * g.wave_x, q1 = k.ky * g.wave_y, q2 = k.kz * g.wave_z;
  const double k2 = q0 * q0 + q1 * q1 + q2 * q2;
  return k2 > 0 ? (T)(inv_eps / k2) : (T)0;
}

// r = A 0 - b = -b, x = 0, z = M^{-1} r (z may alias r: no preconditioner),
// p = z; partial[b] = <r, z>, partial[NPART + b] = <r, r>.  The
// preconditioner scale 1 / eps_mean is read from sc[S_INVEPS].
template <typename T>
__global__ __launch_bounds__(TPB) void kcg_init_kernel(const T *__restrict__ b, T *__restrict__ x, T *r, T *z,
                                                       T *__restrict__ p, const double *__restrict__ sc,
                                                       int precond, double *__restrict__ partial, Grid g) {
  const double inv_eps = sc[S_INVEPS];
  double s_rz = 0.0, s_rr = 0.0;
  for (size_t c = (size_t)blockIdx.x * TPB + threadIdx.x; c < g.nc; c += (size_t)gridDim.x * TPB) {
    const KIdx k = kidx(g, c);
    const V2<T> bv = ld2(b + 2 * c);
    const V2<T> rv{-bv.re, -bv.im};
    st2(r + 2 * c, rv);
    st2(x + 2 * c, V2<T>{(T)0, (T)0});
    V2<T> zv = rv;
    if (precond) {
      zv = scal(precond_factor<T>(g, k, inv_eps), rv);
      st2(z + 2 * c, zv);
    }
    st2(p + 2 * c, zv);
    const double w = kweight(g, k.kz);
    s_rz += w * ((double)rv.re * zv.re + (double)rv.im * zv.im);
    s_rr += w * ((double)rv.re * rv.re + (double)rv.im * rv.im);
  }
  s_rz = block_sum(s_rz);
  s_rr = block_sum(s_rr);
  if (threadIdx.x == 0) {
    partial[blockIdx.x] = s_rz;
    partial[NPART + blockIdx.x] = s_rr;
  }
}

// p = z + beta p (beta = 0 before the first update: p = z), then
// g_d = i k_d e^{i k_d dx/2} p
template <typename T>
__global__ void kcg_dir_grad_kernel(const T *z, T *p, const double *__restrict__ sc, int first,
                                    T *__restrict__ g0, T *__restrict__ g1, T *__restrict__ g2,
                                    const V2<T> *__restrict__ shx, const V2<T> *__restrict__ shy,
                                    const V2<T> *__restrict__ shz, Grid g) {
  size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.nc) return;
  const KIdx k = kidx(g, c);
  V2<T> pv = ld2(z + 2 * c);
  if (!first) pv = pv + scal((T)sc[S_BETA], ld2(p + 2 * c));
  st2(p + 2 * c, pv);
  const T w0 = (T)(k.kx * g.wave_x), w1 = (T)(k.ky * g.wave_y), w2 = (T)(k.kz * g.wave_z);
  st2(g0 + 2 * c, cmul(V2<T>{-w0 * pv.im, w0 * pv.re}, shx[k.i]));
  st2(g1 + 2 * c, cmul(V2<T>{-w1 * pv.im, w1 * pv.re}, shy[k.j]));
  st2(g2 + 2 * c, cmul(V2<T>{-w2 * pv.im, w2 * pv.re}, shz[k.kz]));
}

// ap = i sum_d k_d e^{-i k_d dx/2} w_d (the staggered divergence of the
// fluxes); partial[b] = <p, ap>
template <typename T>
__global__ __launch_bounds__(TPB) void kcg_div_dot_kernel(const T *__restrict__ w0, const T *__restrict__ w1,
                                                          const T *__restrict__ w2, const T *__restrict__ p,
                                                          T *__restrict__ ap, const V2<T> *__restrict__ shx,
                                                          const V2<T> *__restrict__ shy,
                                                          const V2<T> *__restrict__ shz,
                                                          double *__restrict__ partial, Grid g) {
  double s = 0.0;
  for (size_t c = (size_t)blockIdx.x * TPB + threadIdx.x; c < g.nc; c += (size_t)gridDim.x * TPB) {
    const KIdx k = kidx(g, c);
    const T q0 = (T)(k.kx * g.wave_x), q1 = (T)(k.ky * g.wave_y), q2 = (T)(k.kz * g.wave_z);
    V2<T> a = cmul(ld2(w0 + 2 * c), scal(q0, conjv(shx[k.i])));
    a = a + cmul(ld2(w1 + 2 * c), scal(q1, conjv(shy[k.j])));
    a = a + cmul(ld2(w2 + 2 * c), scal(q2, conjv(shz[k.kz])));
    const V2<T> av{-a.im, a.re};
    st2(ap + 2 * c, av);
    const V2<T> pv = ld2(p + 2 * c);
    s += kweight(g, k.kz) * ((double)pv.re * av.re + (double)pv.im * av.im);
  }
  s = block_sum(s);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// alpha = <r,z> / <p, S p> with S p = -A p;  x += alpha p, r += alpha ap,
// z = M^{-1} r; partial[b] = <r, z>, partial[NPART + b] = <r, r>.  With
// npap > 0, <p, A p> is the sum of the npap partials in pap_part, which every
// block forms itself in the same order (no separate reduction launch).
template <typename T>
__global__ __launch_bounds__(TPB) void kcg_update_kernel(T *__restrict__ x, T *r, T *z, const T *__restrict__ p,
                                                         const T *__restrict__ ap, const double *__restrict__ sc,
                                                         int precond, double *__restrict__ partial,
                                                         const double *__restrict__ pap_part, int npap, Grid g) {
  const double inv_eps = sc[S_INVEPS];
  double pap;
  if (npap > 0) {
    __shared__ double pap_sh;
    // rounds of four partials per thread, their loads issued together (the
    // order of one partial per thread and round)
    double s = 0.0;
    for (int i0 = 0; i0 < npap; i0 += 4 * TPB) {
      double v[4];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int i = i0 + threadIdx.x + q * TPB;
        v[q] = i < npap ? pap_part[i] : 0.0;
      }
#pragma unroll
      for (int q = 0; q < 4; q++) s += v[q];
    }
    s = block_sum(s);
    if (threadIdx.x == 0) pap_sh = s;
    __syncthreads();
    pap = -pap_sh;
  } else {
    pap = -sc[S_PAP];
  }
  const T alpha = (T)(pap != 0.0 ? sc[S_RZ] / pap : 0.0);
  double s_rz = 0.0, s_rr = 0.0;
  for (size_t c = (size_t)blockIdx.x * TPB + threadIdx.x; c < g.nc; c += (size_t)gridDim.x * TPB) {
    const KIdx k = kidx(g, c);
    const V2<T> pv = ld2(p + 2 * c);
    st2(x + 2 * c, ld2(x + 2 * c) + scal(alpha, pv));
    const V2<T> rv = ld2(r + 2 * c) + scal(alpha, ld2(ap + 2 * c));
    st2(r + 2 * c, rv);
    V2<T> zv = rv;
    if (precond) {
      zv = scal(precond_factor<T>(g, k, inv_eps), rv);
      st2(z + 2 * c, zv);
    }
    const double w = kweight(g, k.kz);
    s_rz += w * ((double)rv.re * zv.re + (double)rv.im * zv.im);
    s_rr += w * ((double)rv.re * rv.re + (double)rv.im * rv.im);
  }
  s_rz = block_sum(s_rz);
  s_rr = block_sum(s_rr);
  if (threadIdx.x == 0) {
    partial[blockIdx.x] = s_rz;
    partial[NPART + blockIdx.x] = s_rr;
  }
}

// r = i sum_d k_d e^{-i k_d dx/2} w_d - b: the k-space residual A x - b of
// the refinement; partial[blk] = Parseval-weighted <r, r> (without N^-3)
template <typename T>
__global__ __launch_bounds__(TPB) void kres_kernel(const T *__restrict__ w0, const T *__restrict__ w1,
                                                   const T *__restrict__ w2, const T *__restrict__ b,
                                                   T *__restrict__ r, const V2<T> *__restrict__ shx,
                                                   const V2<T> *__restrict__ shy, const V2<T> *__restrict__ shz,
                                                   double *__restrict__ partial, Grid g) {
  double s = 0.0;
  for (size_t c = (size_t)blockIdx.x * TPB + threadIdx.x; c < g.nc; c += (size_t)gridDim.x * TPB) {
    const KIdx k = kidx(g, c);
    const T q0 = (T)(k.kx * g.wave_x), q1 = (T)(k.ky * g.wave_y), q2 = (T)(k.kz * g.wave_z);
    V2<T> a = cmul(ld2(w0 + 2 * c), scal(q0, conjv(shx[k.i])));
    a = a + cmul(ld2(w1 + 2 * c), scal(q1, conjv(shy[k.j])));
    a = a + cmul(ld2(w2 + 2 * c), scal(q2, conjv(shz[k.kz])));
    const V2<T> rv = V2<T>{-a.im, a.re} - ld2(b + 2 * c);
    st2(r + 2 * c, rv);
    s += kweight(g, k.kz) * ((double)rv.re * rv.re + (double)rv.im * rv.im);
  }
  s = block_sum(s);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// Build_rhs (operate_dielectric.cxx:524-540) as a half spectrum.  The face
// average E0_d (eps_i + eps_{i+e_d}) / 2 transforms to E0_d eps_k
// (1 + e^{i k_d dx}) / 2, and the shifted divergence multiplies it by
// i k_d e^{-i k_d dx/2}: b_k = i eps_k sum_d k_d E0_d cos(k_d dx / 2) - rho_k,
// with the background bg (when set) added to the k = 0 mode as nr bg
// (rhok NULL: rho = 0).
// Partial sums of ||b||^2 (Parseval weights) per block.
__global__ __launch_bounds__(TPB) void rhs_k_kernel(const double *__restrict__ epsk, const double *__restrict__ rhok,
                                                    double *__restrict__ b, const V2<double> *__restrict__ shx,
                                                    const V2<double> *__restrict__ shy,
                                                    const V2<double> *__restrict__ shz, double e0x, double e0y,
                                                    double e0z, const double *__restrict__ bg,
                                                    double *__restrict__ partial, Grid g) {
  double s = 0.0;
  for (size_t c = (size_t)blockIdx.x * TPB + threadIdx.x; c < g.nc; c += (size_t)gridDim.x * TPB) {
    const KIdx k = kidx(g, c);
    const double f = k.kx * g.wave_x * e0x * shx[k.i].re + k.ky * g.wave_y * e0y * shy[k.j].re +
                     k.kz * g.wave_z * e0z * shz[k.kz].re;
    const V2<double> e = ld2(epsk + 2 * c), q = rhok ? ld2(rhok + 2 * c) : V2<double>{0, 0};
    V2<double> v{-f * e.im - q.re, f * e.re - q.im};
    if (c == 0 && bg) v.re += (double)g.nr * bg[0];
    st2(b + 2 * c, v);
    s += kweight(g, k.kz) * (v.re * v.re + v.im * v.im);
  }
  s = block_sum(s);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// over every entry of a field (a half spectrum here): prev <- 2 x - prev (the
// caller then swaps the two buffers, so x holds the extrapolation and prev
// the old x without a copy)
__global__ void extrapolate_all_kernel(const double *__restrict__ x, double *__restrict__ prev, size_t n) {
  size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  prev[e] = 2.0 * x[e] - prev[e];
}

// over every entry: x64 -= (double) d   (r = A x - b: the correction enters with a minus sign)
__global__ void sub_f64_kernel(double *__restrict__ x, const double *__restrict__ d, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] -= d[i];
}

__global__ void sub_all_kernel(double *__restrict__ x, const float *__restrict__ d, size_t n) {
  size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) x[e] -= (double)d[e];
}


// x <- 2 x - prev, prev <- old x  (linear extrapolation of the potential in time)
__global__ void extrapolate_kernel(double *__restrict__ x, double *__restrict__ prev, Grid g) {
  size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= g.nr) return;
  RI pt = ridx(g, q);
  const double xn = x[pt.im];
  x[pt.im] = 2.0 * xn - prev[pt.im];
  prev[pt.im] = xn;
}

template <typename T>
__global__ void scale_real_kernel(T *__restrict__ x, double s, Grid g) {
  size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= g.nr) return;
  RI pt = ridx(g, q);
  x[pt.im] = (T)(s * (double)x[pt.im]);
}

// ------------------------------------------------------------------ host
template <typename T>
const V2<T> *sh(kapsel_ctx *ctx, int d) {
  const Twiddles &tw = ctx->tw[prec_of<T>()];
  return (const V2<T> *)(d == 0 ? tw.shx : d == 1 ? tw.shy : tw.shz);
}

double alg(kapsel_ctx *ctx, double fields, size_t elem) { return fields * ctx->g.f * elem; }

template <typename T>
void ax_t(kapsel_ctx *ctx, const T *x, T *out, const T *eps, T *const *s) {
  const Grid &g = ctx->g;
  {
    LaunchTimer lt(ctx, KAPSEL_K_POINTWISE, alg(ctx, 2, sizeof(T)));
    hipLaunchKernelGGL(copy_real_kernel<T>, dim3(blocks_for(g.f, TPB)), dim3(TPB), 0, ctx->stream, x, s[0], g);
    KAPSEL_LAUNCH_CHECK();
  }
  void *s0[1] = {s[0]};
  fft_r2c_t<T>(ctx, s0, 1);
  {
    LaunchTimer lt(ctx, KAPSEL_K_POINTWISE, alg(ctx, 4, sizeof(T)));
    hipLaunchKernelGGL(grad_shift_kernel<T>, dim3(blocks_for(g.nc, TPB)), dim3(TPB), 0, ctx->stream, s[0], s[1],
                       s[2], s[3], sh<T>(ctx, 0), sh<T>(ctx, 1), sh<T>(ctx, 2), g);
    KAPSEL_LAUNCH_CHECK();
  }
  // inverse transforms, face permittivity product, forward transforms: one
  // fused pass over the z (and y) axes (the face factors of the context
  // precision go to the operator scratch slots 5..7)
  void *s3[3] = {s[1], s[2], s[3]};
  void *f3[3] = {ctx->pscratch[5], ctx->pscratch[6], ctx->pscratch[7]};
  {
    LaunchTimer lt(ctx, KAPSEL_K_POINTWISE, alg(ctx, 4, sizeof(T)));
    hipLaunchKernelGGL(face_fields_kernel<T>, dim3(blocks_for(g.f, TPB)), dim3(TPB), 0, ctx->stream, eps,
                       (T *)f3[0], (T *)f3[1], (T *)f3[2], g);
    KAPSEL_LAUNCH_CHECK();
  }
  fft_face_t<T>(ctx, s3, f3);
  {
    LaunchTimer lt(ctx, KAPSEL_K_POINTWISE, alg(ctx, 4, sizeof(T)));
    hipLaunchKernelGGL(div_shift_kernel<T>, dim3(blocks_for(g.nc, TPB)), dim3(TPB), 0, ctx->stream, s[1], s[2],
                       s[3], out, sh<T>(ctx, 0), sh<T>(ctx, 1), sh<T>(ctx, 2), g);
    KAPSEL_LAUNCH_CHECK();
  }
  void *o1[1] = {out};
  fft_c2r_t<T>(ctx, o1, 1);
}

template <typename T>
void build_rhs_t(kapsel_ctx *ctx, T *rhs, const T *rho, const T *eps, const double e0[3], T *const *s) {
  const Grid &g = ctx->g;
  {
    LaunchTimer lt(ctx, KAPSEL_K_POINTWISE, alg(ctx, 4, sizeof(T)));
    hipLaunchKernelGGL(interp_staggered_kernel<T>, dim3(blocks_for(g.nr, TPB)), dim3(TPB), 0, ctx->stream, s[1],
                       s[2], s[3], eps, e0[0], e0[1], e0[2], g);
    KAPSEL_LAUNCH_CHECK();
  }
  void *s3[3] = {s[1], s[2], s[3]};
  fft_r2c_t<T>(ctx, s3, 3);
  {
    LaunchTimer lt(ctx, KAPSEL_K_POINTWISE, alg(ctx, 4, sizeof(T)));
    hipLaunchKernelGGL(div_shift_kernel<T>, dim3(blocks_for(g.nc, TPB)), dim3(TPB), 0, ctx->stream, s[1], s[2],
                       s[3], rhs, sh<T>(ctx, 0), sh<T>(ctx, 1), sh<T>(ctx, 2), g);
    KAPSEL_LAUNCH_CHECK();
  }
  void *o1[1] = {rhs};
  fft_c2r_t<T>(ctx, o1, 1);
  {
    LaunchTimer lt(ctx, KAPSEL_K_POINTWISE, alg(ctx, 3, sizeof(T)));
    hipLaunchKernelGGL(sub_kernel<T>, dim3(blocks_for(g.nr, TPB)), dim3(TPB), 0, ctx->stream, rhs, rho,
                       ctx->rho_bg, g);
    KAPSEL_LAUNCH_CHECK();
  }
}

void ensure_storage(kapsel_ctx *ctx) {
  if (!ctx->d_partial) {
    KAPSEL_HIP(hipMalloc(&ctx->d_partial, sizeof(double) * 3 * NPART));
    KAPSEL_HIP(hipMalloc(&ctx->d_count, sizeof(unsigned long long) * 2));
    KAPSEL_HIP(hipMemsetAsync(ctx->d_count, 0, sizeof(unsigned long long) * 2, ctx->stream));
    KAPSEL_HIP(hipMalloc(&ctx->d_scal, sizeof(double) * S_N));
    KAPSEL_HIP(hipMemsetAsync(ctx->d_scal, 0, sizeof(double) * S_N, ctx->stream));
    ctx->n_partial = NPART;
  }
  if (!ctx->h_slots) {
    KAPSEL_HIP(hipHostMalloc((void **)&ctx->h_slots, sizeof(double) * 8, hipHostMallocDefault));
    for (int i = 0; i < 8; i++) ctx->h_slots[i] = 0.0;
    ctx->seq_cg = ctx->seq_res = 0;
    KAPSEL_HIP(hipHostGetDevicePointer((void **)&ctx->h_slots_dev, ctx->h_slots, 0));
    for (hipEvent_t &e : ctx->slot_ev) KAPSEL_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  }
  if (!ctx->pscratch[0]) {
    for (void *&p : ctx->pscratch) {
      KAPSEL_HIP(hipMalloc(&p, ctx->g.f * ctx->elem));
      KAPSEL_HIP(hipMemsetAsync(p, 0, ctx->g.f * ctx->elem, ctx->stream));
    }
  }
  if (ctx->prec == KAPSEL_FP64 && !ctx->p64[2]) {
    // the residual and right-hand side spectra of the FP64 solve
    for (int i = 2; i < 4; i++) {
      KAPSEL_HIP(hipMalloc(&ctx->p64[i], ctx->g.f * sizeof(double)));
      KAPSEL_HIP(hipMemsetAsync(ctx->p64[i], 0, ctx->g.f * sizeof(double), ctx->stream));
    }
  }
  if (ctx->prec == KAPSEL_FP32 && !ctx->p64[0]) {
    for (double *&p : ctx->p64) {
      KAPSEL_HIP(hipMalloc(&p, ctx->g.f * sizeof(double)));
      KAPSEL_HIP(hipMemsetAsync(p, 0, ctx->g.f * sizeof(double), ctx->stream));
    }
    KAPSEL_HIP(hipMalloc(&ctx->x64_prev, ctx->g.f * sizeof(double)));
    KAPSEL_HIP(hipMemsetAsync(ctx->x64_prev, 0, ctx->g.f * sizeof(double), ctx->stream));
  }
}

template <typename T, typename U>
void dot_t(kapsel_ctx *ctx, const T *a, const U *b, int slot, int mode = 0) {
  const Grid &g = ctx->g;
  LaunchTimer lt(ctx, KAPSEL_K_POINTWISE, (double)g.nr * (sizeof(T) + sizeof(U)));
  hipLaunchKernelGGL((dot_partial_kernel<T, U>), dim3(NPART), dim3(TPB), 0, ctx->stream, a, b, ctx->d_partial, g);
  KAPSEL_LAUNCH_CHECK();
  hipLaunchKernelGGL(finish_kernel, dim3(1), dim3(TPB), 0, ctx->stream, ctx->d_partial, NPART, ctx->d_scal, slot,
                     mode);
  KAPSEL_LAUNCH_CHECK();
}

double read_scal(kapsel_ctx *ctx, int slot) {
  double v = 0.0;
  KAPSEL_HIP(hipMemcpyAsync(&v, ctx->d_scal + slot, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  KAPSEL_HIP(hipStreamSynchronize(ctx->stream));
  return v;
}

// mean of eps over the real points (preconditioner scale)
template <typename T>
double eps_mean(kapsel_ctx *ctx, const T *eps, T *ones) {
  dot_t<T, T>(ctx, eps, ones, S_TMP);
  return read_scal(ctx, S_TMP) / (double)ctx->g.nr;
}

template <typename T>
__global__ void fill_kernel(T *__restrict__ a, T v, size_t n) {
  size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) a[e] = v;
}

// Pipelined convergence test of the CG loops.  Each queued iteration copies
// its ||r||^2 into pinned slot it % 2 and records an event; the host keeps
// one more iteration queued while it waits for the older one, so the GPU
// does not drain between iterations.  It stops queueing ahead when the
// geometric trend of the last two values says the iteration in flight
// meets the bound, so normally no iteration runs past it.
// Waits until the sequence number a kernel stores in host slot `slot`
// reaches target.  The stores are system-scope, so polling sees them; after
// a bounded spin (seconds) it waits for the stream instead, which ends once
// the queued work has run.
void wait_seq(kapsel_ctx *ctx, int slot, unsigned long long target) {
  const volatile double *h = ctx->h_slots;
  const double want = (double)target;
  for (long spin = 0; spin < (1L << 26); spin++) {
    if (h[slot] >= want) return;
    __builtin_ia32_pause();
  }
  KAPSEL_HIP(hipStreamSynchronize(ctx->stream));
  if (h[slot] < want) {
    set_error("a kernel's host-side result (slot %d, number %llu) did not arrive", slot, target);
    throw Error{KAPSEL_ERR_HIP};
  }
}

// device_slots: the iteration's last kernel stores ||r||^2 and its sequence
// number into the host slots itself (finish2_kernel with host slots), and
// the host polls the number -- no marker enters the stream between two
// iterations; else a copy and an event are queued after each iteration.
struct ConvergenceWatch {
  kapsel_ctx *ctx;
  double scale;
  double rr, rr_prev = -1.0;
  int it = 0, known = 0;
  bool device_slots;
  unsigned long long seq[2] = {0, 0};
  ConvergenceWatch(kapsel_ctx *c, double s, double rr0, bool dev = false)
      : ctx(c), scale(s), rr(rr0), device_slots(dev) {}
  // back: how many iterations were queued after this one (0: it is the last)
  void queued(unsigned long long back = 0) {
    it++;
    if (device_slots) {
      seq[it & 1] = ctx->seq_cg - back;  // the iteration's sequence number
      return;
    }
    KAPSEL_HIP(hipMemcpyAsync(ctx->h_slots + (it & 1), ctx->d_scal + S_RR, sizeof(double),
                              hipMemcpyDeviceToHost, ctx->stream));
    KAPSEL_HIP(hipEventRecord(ctx->slot_ev[it & 1], ctx->stream));
  }
  void wait_one() {
    known++;
    rr_prev = rr;
    if (device_slots) {
      const unsigned long long k = seq[known & 1];
      wait_seq(ctx, 4, k);
      rr = static_cast<volatile double *>(ctx->h_slots)[k & 1] * scale;
      return;
    }
    KAPSEL_HIP(hipEventSynchronize(ctx->slot_ev[known & 1]));
    rr = static_cast<volatile double *>(ctx->h_slots)[known & 1] * scale;
  }
  // runs enqueue() until ||r||^2 < tol2 or maxit iterations.  With tail:
  // when the trend predicts that the iteration in flight is the last (or it
  // is the maxit-th), tail() is queued right behind it before the host reads
  // it, and the run ends with that iteration (tail_queued) whether or not it
  // met the bound -- the caller's next step then sees what it left.
  bool tail_queued = false;
  template <typename F>
  void run(double tol2, int maxit, F enqueue, const std::function<void()> *tail = nullptr) {
    for (;;) {
      if (known == it) {
        if (rr < tol2 || it >= maxit) return;
        enqueue();
        queued();
      } else {
        // before a trend exists, up to two iterations in flight while the
        // known residual is far above the bound (one iteration does not take
        // it 100x down), so the GPU does not wait for the host
        const bool trend = rr_prev > 0.0;
        const bool ahead =
            it < maxit && (trend ? rr * (rr / rr_prev) >= tol2 : (it - known < 2 && rr >= 1e4 * tol2));
        if (ahead) {
          enqueue();
          queued();
        } else if (tail && (trend || it >= maxit) && known + 1 == it) {
          (*tail)();
          tail_queued = true;
        }
        wait_one();
        if (tail_queued && known == it) return;
      }
    }
  }
};

// One k-space PCG iteration: p = z + beta p, ap = A p (3 inverse and 3
// forward transforms), alpha, x += alpha p, r += alpha ap, z = M^{-1} r,
// beta.  Every scalar lives in device memory: the host queues iterations
// ahead without reading any of them.
template <typename T>
void kcg_iteration(kapsel_ctx *ctx, T *x, const T *eps, int precond, T *const *w, T *const *s,
                   T *const *face) {
  const Grid &g = ctx->g;
  double *pap_part = ctx->d_partial + 2 * NPART;
  int npap = 0;
  T *r = w[0], *p = w[1], *ap = w[2], *z = precond ? w[3] : w[0];
  void *s3[3] = {s[0], s[1], s[2]};
  void *f3[3] = {face[0], face[1], face[2]};
  if (fft_x_grad_t<T>(ctx, z, p, ctx->d_scal, S_BETA, s3)) {
    // direction update and gradient in the first pass, divergence and
    // <p, Ap> in the last
    fft_face_mid_t<T>(ctx, s3, f3);
    // <p, A p> partials beside the update's own (summed in the update)
    npap = fft_x_div_t<T>(ctx, s3, p, ap, pap_part, NPART);
  } else {
    {
      LaunchTimer lt(ctx, KAPSEL_K_POINTWISE, alg(ctx, 6, sizeof(T)));
      hipLaunchKernelGGL(kcg_dir_grad_kernel<T>, dim3(blocks_for(g.nc, TPB)), dim3(TPB), 0, ctx->stream, z, p,
                         ctx->d_scal, 0, s[0], s[1], s[2], sh<T>(ctx, 0), sh<T>(ctx, 1), sh<T>(ctx, 2), g);
      KAPSEL_LAUNCH_CHECK();
    }
    fft_face_t<T>(ctx, s3, f3);
    {
      LaunchTimer lt(ctx, KAPSEL_K_POINTWISE, alg(ctx, 5, sizeof(T)));
      hipLaunchKernelGGL(kcg_div_dot_kernel<T>, dim3(NPART), dim3(TPB), 0, ctx->stream, s[0], s[1], s[2], p, ap,
                         sh<T>(ctx, 0), sh<T>(ctx, 1), sh<T>(ctx, 2), ctx->d_partial, g);
      KAPSEL_LAUNCH_CHECK();
      hipLaunchKernelGGL(finish_kernel, dim3(1), dim3(TPB), 0, ctx->stream, ctx->d_partial, NPART, ctx->d_scal,
                         S_PAP, 0);
      KAPSEL_LAUNCH_CHECK();
    }
  }
  {
    LaunchTimer lt(ctx, KAPSEL_K_POINTWISE, alg(ctx, precond ? 7 : 6, sizeof(T)));
    hipLaunchKernelGGL(kcg_update_kernel<T>, dim3(NPART), dim3(TPB), 0, ctx->stream, x, r, z, p, ap, ctx->d_scal,
                       precond, ctx->d_partial, (const double *)pap_part, npap, g);
    KAPSEL_LAUNCH_CHECK();
    hipLaunchKernelGGL(finish2_kernel, dim3(1), dim3(TPB), 0, ctx->stream, ctx->d_partial, ctx->d_scal, 1,
                       ctx->h_slots_dev, ctx->d_count);
    KAPSEL_LAUNCH_CHECK();
  }
}

// Queues one iteration, launched eagerly: replaying it as a captured HIP
// graph measured slower (908 against 936 steps/s for the ICEP step at
// 128^3 on one MI355X, runs alternated).
template <typename T>
void kcg_enqueue(kapsel_ctx *ctx, T *x, const T *eps, int precond, T *const *w, T *const *s, T *const *face) {
  ctx->seq_cg++;  // its finish2 stores this number in host slot 4
  kcg_iteration<T>(ctx, x, eps, precond, w, s, face);
}

// PCG on S = -A for A x = b with every vector in k-space: b and x are the
// half spectra (x written, b read); w = {r, p, ap, z} k-space work vectors,
// s = {g0, g1, g2} flux scratch; sc[S_INVEPS] holds 1 / eps_mean.  Stops
// when the real-space ||A x - b|| is below tol_abs; returns the iteration
// count and *rr_out = that ||A x - b||^2.  rr_known >= 0: ||b||^2 as the
// caller computed it (x starts at zero), so the start needs no read-back.
template <typename T>
void kcg_init_t(kapsel_ctx *ctx, T *x, const T *b, int precond, T *const *w) {
  const Grid &g = ctx->g;
  T *r = w[0], *p = w[1], *z = precond ? w[3] : w[0];
  LaunchTimer lt(ctx, KAPSEL_K_POINTWISE, alg(ctx, precond ? 5 : 4, sizeof(T)));
  hipLaunchKernelGGL(kcg_init_kernel<T>, dim3(NPART), dim3(TPB), 0, ctx->stream, b, x, r, z, p, ctx->d_scal, precond,
                     ctx->d_partial, g);
  KAPSEL_LAUNCH_CHECK();
  hipLaunchKernelGGL(finish2_kernel, dim3(1), dim3(TPB), 0, ctx->stream, ctx->d_partial, ctx->d_scal, 0,
                     ctx->h_slots_dev, ctx->d_count);
  KAPSEL_LAUNCH_CHECK();
}

// Queues the start of a correction and its first n iterations before the
// caller has read whether the correction is needed (kcg_t with begun = n then
// continues it; if it is not needed, they only wrote the correction's vectors)
template <typename T>
void kcg_begin(kapsel_ctx *ctx, T *x, const T *b, const T *eps, int precond, T *const *w, T *const *s,
               T *const *face, int n) {
  kcg_init_t<T>(ctx, x, b, precond, w);
  for (int i = 0; i < n; i++) kcg_enqueue<T>(ctx, x, eps, precond, w, s, face);
}

template <typename T>
int kcg_t(kapsel_ctx *ctx, T *x, const T *b, const T *eps, double tol_abs, int maxit, int precond, T *const *w,
          T *const *s, T *const *face, double *rr_out, double rr_known = -1.0, int begun = 0,
          const std::function<void()> *tail = nullptr, bool *tail_queued = nullptr) {
  const Grid &g = ctx->g;
  const double norm = 1.0 / (double)g.nr;  // Parseval factor of the k-space sums
  if (!begun) kcg_init_t<T>(ctx, x, b, precond, w);
  // ||b||^2 when the caller knows it (to rounding), else read back
  double rr = rr_known >= 0.0 ? rr_known : read_scal(ctx, S_RR) * norm;
  const double tol2 = tol_abs * tol_abs;
  ConvergenceWatch cw(ctx, norm, rr, true);
  for (int i = 0; i < begun; i++) cw.queued(begun - 1 - i);  // the iterations kcg_begin queued
  cw.run(tol2, maxit, [&]() { kcg_enqueue<T>(ctx, x, eps, precond, w, s, face); }, tail);
  if (tail_queued) *tail_queued = cw.tail_queued;
  rr = cw.rr;
  const int it = cw.it;
  *rr_out = rr;
  return it;
}

template <typename T>
T *fld(kapsel_ctx *ctx, int f) {
  return (T *)ctx->fields[f];
}

template <typename T>
double inv_eps_of(kapsel_ctx *ctx, const T *eps, T *scratch) {
  const Grid &g = ctx->g;
  hipLaunchKernelGGL(fill_kernel<T>, dim3(blocks_for(g.f, TPB)), dim3(TPB), 0, ctx->stream, scratch, (T)1, g.f);
  KAPSEL_LAUNCH_CHECK();
  double m = eps_mean<T>(ctx, eps, scratch);
  return m > 0 ? 1.0 / m : 1.0;
}

// FP64 context: the reference's CG (operate_dielectric.cxx:541-578: start
// from the current potential, stop at ||A x - b|| < tol ||b||, solve again
// from zero after maxit iterations) on the half spectra.  With x0 the
// start, r0 = A x0 - b once, then CG on A d = r0 from d = 0 -- the iterates
// of CG on A x = b from x0, x = x0 - d -- with the k-space iteration of the
// FP32 correction solves (fused x passes, one plane pass, one update kernel)
// in FP64.  relres: the residual of the final x, evaluated anew.
int solve_fp64(kapsel_ctx *ctx, const double e0[3], double tol, int maxit, int precond, double *relres) {
  const Grid &g = ctx->g;
  double *eps = fld<double>(ctx, KAPSEL_EPSILON), *x = fld<double>(ctx, KAPSEL_POTENTIAL);
  double *b = fld<double>(ctx, KAPSEL_RHS);
  double *s[4] = {(double *)ctx->pscratch[0], (double *)ctx->pscratch[1], (double *)ctx->pscratch[2],
                  (double *)ctx->pscratch[3]};
  double *w[4] = {fld<double>(ctx, KAPSEL_WORK0), fld<double>(ctx, KAPSEL_WORK1), fld<double>(ctx, KAPSEL_WORK2),
                  fld<double>(ctx, KAPSEL_WORK3)};
  double *face[3] = {(double *)ctx->pscratch[5], (double *)ctx->pscratch[6], (double *)ctx->pscratch[7]};
  double *xk = s[0], *d = (double *)ctx->pscratch[4], *rk = ctx->p64[2], *bk = ctx->p64[3];
  const unsigned fb = blocks_for(g.f, TPB);
  build_rhs_t<double>(ctx, b, fld<double>(ctx, KAPSEL_RHO_FREE), eps, e0, s);
  dot_t<double, double>(ctx, b, b, S_BB);
  const double bb = read_scal(ctx, S_BB);
  if (bb == 0.0) {
    KAPSEL_HIP(hipMemsetAsync(x, 0, g.f * sizeof(double), ctx->stream));
    *relres = 0.0;
    return 0;
  }
  const double inv_eps = precond ? inv_eps_of<double>(ctx, eps, w[3]) : 1.0;
  {
    LaunchTimer lt(ctx, KAPSEL_K_POINTWISE, alg(ctx, 4, sizeof(double)));
    hipLaunchKernelGGL(fill_kernel<double>, dim3(1), dim3(64), 0, ctx->stream, ctx->d_scal + S_INVEPS, inv_eps,
                       (size_t)1);
    hipLaunchKernelGGL(face_fields_kernel<double>, dim3(fb), dim3(TPB), 0, ctx->stream, eps, face[0], face[1], face[2],
                       g);
    KAPSEL_LAUNCH_CHECK();
  }
  // the start and the right-hand side as spectra
  {
    LaunchTimer lt(ctx, KAPSEL_K_POINTWISE, alg(ctx, 4, sizeof(double)));
    hipLaunchKernelGGL(copy_real_kernel<double>, dim3(fb), dim3(TPB), 0, ctx->stream, x, xk, g);
    hipLaunchKernelGGL(copy_real_kernel<double>, dim3(fb), dim3(TPB), 0, ctx->stream, b, bk, g);
    KAPSEL_LAUNCH_CHECK();
  }
  void *xb[2] = {xk, bk};
  fft_r2c_t<double>(ctx, xb, 2);
  void *g3[3] = {s[1], s[2], s[3]};
  void *f3[3] = {face[0], face[1], face[2]};
  // rk = A xk - bk and ||rk||^2 (Parseval) into S_RR
  auto residual = [&]() {
    if (fft_x_grad_t<double>(ctx, xk, xk, ctx->d_scal, -1, g3)) {
      fft_face_mid_t<double>(ctx, g3, f3);
      const int nb = fft_x_res_t<double>(ctx, g3, bk, rk, ctx->d_partial, NPART);
      hipLaunchKernelGGL(finish_kernel, dim3(1), dim3(TPB), 0, ctx->stream, ctx->d_partial, nb, ctx->d_scal, S_RR,
                         0);
    } else {
      {
        LaunchTimer lt(ctx, KAPSEL_K_POINTWISE, alg(ctx, 4, sizeof(double)));
        hipLaunchKernelGGL(grad_shift_kernel<double>, dim3(blocks_for(g.nc, TPB)), dim3(TPB), 0, ctx->stream, xk,
                           s[1], s[2], s[3], sh<double>(ctx, 0), sh<double>(ctx, 1), sh<double>(ctx, 2), g);
        KAPSEL_LAUNCH_CHECK();
      }
      fft_face_t<double>(ctx, g3, f3);
      LaunchTimer lt(ctx, KAPSEL_K_POINTWISE, alg(ctx, 5, sizeof(double)));
      hipLaunchKernelGGL(kres_kernel<double>, dim3(NPART), dim3(TPB), 0, ctx->stream, s[1], s[2], s[3], bk, rk,
                         sh<double>(ctx, 0), sh<double>(ctx, 1), sh<double>(ctx, 2), ctx->d_partial, g);
      hipLaunchKernelGGL(finish_kernel, dim3(1), dim3(TPB), 0, ctx->stream, ctx->d_partial, NPART, ctx->d_scal,
                         S_RR, 0);
    }
    KAPSEL_LAUNCH_CHECK();
    return read_scal(ctx, S_RR) / (double)g.nr;
  };
  const double tol_abs = tol * sqrt(bb);
  double rr0 = residual(), rr = 0.0;
  int it = kcg_t<double>(ctx, d, rk, eps, tol_abs, maxit, precond, w, s + 1, face, &rr, rr0);
  if (it >= maxit) {
    // operate_dielectric.cxx:563-570: solve again without the initial guess
    KAPSEL_HIP(hipMemsetAsync(xk, 0, g.f * sizeof(double), ctx->stream));
    rr0 = residual();
    it = kcg_t<double>(ctx, d, rk, eps, tol_abs, maxit, precond, w, s + 1, face, &rr, rr0);
  }
  {
    LaunchTimer lt(ctx, KAPSEL_K_POINTWISE, alg(ctx, 3, sizeof(double)));
    hipLaunchKernelGGL(sub_f64_kernel, dim3(fb), dim3(TPB), 0, ctx->stream, xk, d, g.f);
    KAPSEL_LAUNCH_CHECK();
  }
  *relres = sqrt(residual()) / sqrt(bb);
  void *x1[1] = {xk};
  fft_c2r_t<double>(ctx, x1, 1);
  {
    LaunchTimer lt(ctx, KAPSEL_K_POINTWISE, alg(ctx, 2, sizeof(double)));
    hipLaunchKernelGGL(copy_real_kernel<double>, dim3(fb), dim3(TPB), 0, ctx->stream, xk, x, g);
    KAPSEL_LAUNCH_CHECK();
  }
  return it;
}

// marks the FP64 part of an FP32 context's solve for the kernel timing
struct RefineScope {
  kapsel_ctx *c;
  explicit RefineScope(kapsel_ctx *x) : c(x) { c->refine64 = true; }
  ~RefineScope() { c->refine64 = false; }
};

// The FP32 outputs of an FP32-context solve: POTENTIAL and RHS as real
// fields, from the FP64 spectra of the solution (p64[1]) and of the
// right-hand side (p64[3]).
void solve_outputs_impl(kapsel_ctx *ctx) {
  const Grid &g = ctx->g;
  const unsigned fb = blocks_for(g.f, TPB);
  float *pot = fld<float>(ctx, KAPSEL_POTENTIAL), *b32 = fld<float>(ctx, KAPSEL_RHS);
  hipLaunchKernelGGL((convert_real_kernel<float, double>), dim3(fb), dim3(TPB), 0, ctx->stream, pot, ctx->p64[1], g);
  hipLaunchKernelGGL((convert_real_kernel<float, double>), dim3(fb), dim3(TPB), 0, ctx->stream, b32, ctx->p64[3], g);
  KAPSEL_LAUNCH_CHECK();
  void *p2[2] = {pot, b32};
  fft_c2r_t<float>(ctx, p2, 2);
}

// FP32 context: FP64 iterative refinement around FP32 (P)CG corrections.
// The host waits twice per solve besides the correction's own pipelined
// checks: once for ||b|| and the first residual, once for each later
// residual; everything that does not depend on a decision is queued first.
int solve_fp32(kapsel_ctx *ctx, const double e0[3], double tol, int maxit, int precond, double *relres) {
  const Grid &g = ctx->g;
  RefineScope refine(ctx);
  ctx->spec64_fresh = false;
  // FP64 shadow: eps (real, then its spectrum), the potential and the
  // residual as half spectra (rk holds rho until the RHS is built), the RHS
  // spectrum, operator scratch, face permittivities
  double *eps64 = ctx->p64[0], *xk = ctx->p64[1], *rk = ctx->p64[2], *bk = ctx->p64[3];
  double *s64[4] = {nullptr, ctx->p64[4], ctx->p64[5], ctx->p64[6]};  // [1..3]: the operator's three fields
  double *face64[3] = {ctx->p64[7], ctx->p64[8], ctx->p64[9]};
  float *s32[4] = {(float *)ctx->pscratch[0], (float *)ctx->pscratch[1], (float *)ctx->pscratch[2],
                   (float *)ctx->pscratch[3]};
  float *w32[4] = {fld<float>(ctx, KAPSEL_WORK0), fld<float>(ctx, KAPSEL_WORK1), fld<float>(ctx, KAPSEL_WORK2),
                   fld<float>(ctx, KAPSEL_WORK3)};
  float *d32 = (float *)ctx->pscratch[4], *eps32 = fld<float>(ctx, KAPSEL_EPSILON);
  const unsigned fb = blocks_for(g.f, TPB);
  float *face32[3] = {(float *)ctx->pscratch[5], (float *)ctx->pscratch[6], (float *)ctx->pscratch[7]};
  // a free charge never written is zero: no FP64 copy, no transform
  const bool rho_zero = !ctx->rho_free_set && !ctx->rho_bg;
  // FP64 copies of the FP32 inputs (eps and rho enter the residual in FP64),
  // the face factors of both operators, 1 / mean(eps) for the preconditioner
  {
    LaunchTimer lt(ctx, KAPSEL_K_POINTWISE, (4.0 + 4.0 + 8.0 + 8.0 + 12.0 + 24.0) * g.f);
    const float *rho32 = rho_zero ? nullptr : fld<float>(ctx, KAPSEL_RHO_FREE);
    hipLaunchKernelGGL(refine_setup_kernel, dim3(NPART), dim3(TPB), 0, ctx->stream, eps32, rho32, eps64, rk,
                       face32[0], face32[1], face32[2], face64[0], face64[1], face64[2], ctx->d_partial, g);
    KAPSEL_LAUNCH_CHECK();
    hipLaunchKernelGGL(finish_kernel, dim3(1), dim3(TPB), 0, ctx->stream, ctx->d_partial, NPART, ctx->d_scal, S_TMP,
                       2, nullptr, 0, nullptr, (double)g.nr);
    KAPSEL_LAUNCH_CHECK();
  }
  // start: the POTENTIAL field if it was written since the last solve; else
  // the last solution, extrapolated linearly in time once two solutions of
  // consecutive solves are known
  bool extrapolate_in_residual = false;
  if (!ctx->pot64_valid) {
    hipLaunchKernelGGL((convert_real_kernel<double, float>), dim3(fb), dim3(TPB), 0, ctx->stream, xk,
                       fld<float>(ctx, KAPSEL_POTENTIAL), g);
    KAPSEL_LAUNCH_CHECK();
    void *x1[1] = {xk};
    fft_r2c_t<double>(ctx, x1, 1);
    ctx->prev_valid = false;
  } else if (ctx->prev_valid) {
    // prev <- 2 x - prev, the buffers swapped: where the first residual's
    // first pass is fused, that pass forms it (reading x and prev, writing
    // the start it transforms)
    if (!(g.nx == 64 || g.nx == 128)) {
      LaunchTimer lt(ctx, KAPSEL_K_POINTWISE, alg(ctx, 3, sizeof(double)));
      hipLaunchKernelGGL(extrapolate_all_kernel, dim3(fb), dim3(TPB), 0, ctx->stream, xk, ctx->x64_prev, g.f);
      KAPSEL_LAUNCH_CHECK();
    } else {
      extrapolate_in_residual = true;
    }
    std::swap(ctx->p64[1], ctx->x64_prev);
    xk = ctx->p64[1];
  } else {
    KAPSEL_HIP(hipMemcpyAsync(ctx->x64_prev, xk, g.f * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    ctx->prev_valid = true;
  }
  // eps64 and rho as spectra
  void *er[2] = {eps64, rk};
  fft_r2c_t<double>(ctx, er, rho_zero ? 1 : 2);
  {
    LaunchTimer lt(ctx, KAPSEL_K_POINTWISE, alg(ctx, 3, sizeof(double)));
    const double *rho_k = rho_zero ? nullptr : rk;
    hipLaunchKernelGGL(rhs_k_kernel, dim3(NPART), dim3(TPB), 0, ctx->stream, eps64, rho_k, bk, sh<double>(ctx, 0),
                       sh<double>(ctx, 1), sh<double>(ctx, 2), e0[0], e0[1], e0[2], ctx->rho_bg, ctx->d_partial, g);
    KAPSEL_LAUNCH_CHECK();
    hipLaunchKernelGGL(finish_kernel, dim3(1), dim3(TPB), 0, ctx->stream, ctx->d_partial, NPART, ctx->d_scal, S_BB,
                       0, ctx->h_slots_dev, 2);
    KAPSEL_LAUNCH_CHECK();
  }
  void *g3[3] = {s64[1], s64[2], s64[3]};
  void *f3[3] = {face64[0], face64[1], face64[2]};
  const double norm = 1.0 / (double)g.nr;  // Parseval factor of the k-space sums
  float *b32 = fld<float>(ctx, KAPSEL_RHS);  // the correction's right-hand side spectrum meanwhile
  float *pot = fld<float>(ctx, KAPSEL_POTENTIAL);
  // outputs: the potential and the RHS b (FP32, real); with lazy_outputs
  // (inside kapsel_icep_step, whose later parts read the FP64 spectra) they
  // are only marked stale and made when the API next reads or uses them
  auto outputs = [&]() {
    if (ctx->lazy_outputs) {
      ctx->outputs_stale = true;
      return;
    }
    solve_outputs_impl(ctx);
  };
  ctx->outputs_stale = false;
  // rk = A xk - bk in k-space, ||r||^2 by Parseval (raw sum also into the
  // host slot 3); where NX has fused x passes, the gradient symbols ride in
  // the first and the divergence and the residual's partial sums in the
  // last.  sub: xk -= sub first (in the first pass where it is fused)
  auto residual = [&](const float *sub) {
    if (sub && !(ctx->g.nx == 64 || ctx->g.nx == 128)) {
      hipLaunchKernelGGL(sub_all_kernel, dim3(fb), dim3(TPB), 0, ctx->stream, xk, sub, g.f);
      KAPSEL_LAUNCH_CHECK();
      sub = nullptr;
    }
    // (xk = 2 x64_prev - xk first where the extrapolation was deferred here)
    const bool fused = extrapolate_in_residual
                           ? fft_x_grad_t<double>(ctx, ctx->x64_prev, xk, ctx->d_scal, S_M1, g3, nullptr, 2.0)
                           : fft_x_grad_t<double>(ctx, xk, xk, ctx->d_scal, -1, g3, sub);
    extrapolate_in_residual = false;
    if (fused) {
      fft_face_mid_t<double>(ctx, g3, f3);
      const int nb = fft_x_res_t<double>(ctx, g3, bk, rk, ctx->d_partial, NPART);
      hipLaunchKernelGGL(finish_kernel, dim3(1), dim3(TPB), 0, ctx->stream, ctx->d_partial, nb, ctx->d_scal, S_RR,
                         0, ctx->h_slots_dev, 3, ctx->d_count);
      KAPSEL_LAUNCH_CHECK();
    } else {
      {
        LaunchTimer lt(ctx, KAPSEL_K_POINTWISE, alg(ctx, 4, sizeof(double)));
        hipLaunchKernelGGL(grad_shift_kernel<double>, dim3(blocks_for(g.nc, TPB)), dim3(TPB), 0, ctx->stream, xk,
                           s64[1], s64[2], s64[3], sh<double>(ctx, 0), sh<double>(ctx, 1), sh<double>(ctx, 2), g);
        KAPSEL_LAUNCH_CHECK();
      }
      fft_face_t<double>(ctx, g3, f3);
      LaunchTimer lt(ctx, KAPSEL_K_POINTWISE, alg(ctx, 5, sizeof(double)));
      hipLaunchKernelGGL(kres_kernel<double>, dim3(NPART), dim3(TPB), 0, ctx->stream, s64[1], s64[2], s64[3], bk,
                         rk, sh<double>(ctx, 0), sh<double>(ctx, 1), sh<double>(ctx, 2), ctx->d_partial, g);
      KAPSEL_LAUNCH_CHECK();
      hipLaunchKernelGGL(finish_kernel, dim3(1), dim3(TPB), 0, ctx->stream, ctx->d_partial, NPART, ctx->d_scal,
                         S_RR, 0, ctx->h_slots_dev, 3, ctx->d_count);
      KAPSEL_LAUNCH_CHECK();
    }
    ctx->seq_res++;  // the finish stores this number in host slot 5
  };
  // x -= d, then the new residual: queued by the correction itself behind
  // the iteration its trend predicts to be the last (the next residual read
  // tells whether it was)
  bool hook_current = false;  // the solve hook is queued for the current xk
  ctx->solve_hook_done = false;
  const std::function<void()> apply_and_residual = [&]() {
    const bool was = ctx->refine64;
    ctx->refine64 = true;
    hook_current = false;
    residual(d32);
    ctx->refine64 = was;
  };
  // the caller's work on the potential about to be checked (it only writes
  // what a further correction rewrites before reading)
  auto hook = [&]() {
    if (!ctx->solve_hook) return;
    const bool was = ctx->refine64;
    ctx->refine64 = false;
    (*ctx->solve_hook)();
    ctx->refine64 = was;
    hook_current = true;
  };
  // While the host waits for a residual, the GPU runs what most likely
  // follows: after the first residual the correction's start and first
  // iteration, after a correction the outputs.  Whichever turns out not to
  // be needed only wrote scratch (or is redone before the solve returns).
  int total = 0;
  double rel = 1.0, bnorm = 0.0;
  const int nbegin = maxit < 2 ? maxit : 2;  // iterations queued before the first read-back
  bool outputs_current = false;
  residual(nullptr);
  for (int outer = 0; outer < 30; outer++) {
    const unsigned long long res_seq = ctx->seq_res;
    const bool begun = outer == 0;
    if (begun) {
      // the correction's right-hand side (the residual in FP32) and start
      hipLaunchKernelGGL((convert_real_kernel<float, double>), dim3(fb), dim3(TPB), 0, ctx->stream, b32, rk, g);
      KAPSEL_LAUNCH_CHECK();
      ctx->refine64 = false;
      kcg_begin<float>(ctx, d32, b32, eps32, precond, w32, s32 + 1, face32, nbegin);
      ctx->refine64 = true;
    } else {
      outputs();
      hook();
    }
    wait_seq(ctx, 5, res_seq);
    const volatile double *hs = ctx->h_slots;
    if (outer == 0) {
      const double bb = hs[2] * norm;
      if (bb == 0.0) {
        KAPSEL_HIP(hipMemsetAsync(pot, 0, g.f * sizeof(float), ctx->stream));
        KAPSEL_HIP(hipMemsetAsync(b32, 0, g.f * sizeof(float), ctx->stream));
        ctx->pot64_valid = false;
        *relres = 0.0;
        return 0;
      }
      bnorm = sqrt(bb);
    }
    const double rr = hs[3] * norm;
    rel = sqrt(rr) / bnorm;
    if (rel < tol || total >= maxit) {
      outputs_current = !begun;
      break;
    }
    // correction: A d = r in FP32 k-space, then x -= d.  Inner tolerance:
    // what the outer bound still needs (80 % of it: the FP32 operator's own
    // rounding leaves the new FP64 residual within a few percent of the
    // inner one), but never below 1e-4 of the current residual, the
    // reduction FP32 certifies
    if (!begun) {
      hipLaunchKernelGGL((convert_real_kernel<float, double>), dim3(fb), dim3(TPB), 0, ctx->stream, b32, rk, g);
      KAPSEL_LAUNCH_CHECK();
    }
    ctx->refine64 = false;  // the FP32 correction solve
    double rr_in = 0.0;
    const double tol_in = fmax(1e-4 * sqrt(rr), 0.8 * tol * bnorm);
    bool tail = false;
    int it = kcg_t<float>(ctx, d32, b32, eps32, tol_in, maxit - total, precond, w32, s32 + 1, face32, &rr_in, rr,
                          begun ? nbegin : 0, &apply_and_residual, &tail);
    total += it;
    ctx->refine64 = true;
    if (!tail) {
      if (it == 0) break;
      apply_and_residual();
    }
  }
  if (!outputs_current) outputs();
  ctx->refine64 = false;
  ctx->pot64_valid = true;
  ctx->spec64_fresh = true;
  ctx->solve_hook_done = hook_current;
  *relres = rel;
  return total;
}

// staggered gradient of a real field a into g[0..2] (s0 scratch)
template <typename T>
void staggered_grad_t(kapsel_ctx *ctx, const T *a, T *s0, T *const *gout) {
  const Grid &g = ctx->g;
  {
    LaunchTimer lt(ctx, KAPSEL_K_POINTWISE, alg(ctx, 2, sizeof(T)));
    hipLaunchKernelGGL(copy_real_kernel<T>, dim3(blocks_for(g.f, TPB)), dim3(TPB), 0, ctx->stream, a, s0, g);
    KAPSEL_LAUNCH_CHECK();
  }
  void *s1[1] = {s0};
  fft_r2c_t<T>(ctx, s1, 1);
  {
    LaunchTimer lt(ctx, KAPSEL_K_POINTWISE, alg(ctx, 4, sizeof(T)));
    hipLaunchKernelGGL(grad_shift_kernel<T>, dim3(blocks_for(g.nc, TPB)), dim3(TPB), 0, ctx->stream, s0, gout[0],
                       gout[1], gout[2], sh<T>(ctx, 0), sh<T>(ctx, 1), sh<T>(ctx, 2), g);
    KAPSEL_LAUNCH_CHECK();
  }
  void *g3[3] = {gout[0], gout[1], gout[2]};
  fft_c2r_t<T>(ctx, g3, 3);
}

// from_spectra (FP32 context, right after solve_fp32): the staggered
// gradients come from the FP64 spectra of eps and of the potential the
// refinement left in p64[0] and p64[1], which saves a copy and a forward
// transform of each field
template <typename T>
void maxwell_t(kapsel_ctx *ctx, const double e0[3], bool from_spectra) {
  const Grid &g = ctx->g;
  T *s0 = (T *)ctx->pscratch[0];
  T *ge[3] = {(T *)ctx->pscratch[1], (T *)ctx->pscratch[2], (T *)ctx->pscratch[3]};
  T *gp[3] = {fld<T>(ctx, KAPSEL_WORK1), fld<T>(ctx, KAPSEL_WORK2), fld<T>(ctx, KAPSEL_WORK3)};
  T *e2 = fld<T>(ctx, KAPSEL_WORK0);
  if constexpr (sizeof(T) == 4) {
    void *six[6] = {ge[0], ge[1], ge[2], gp[0], gp[1], gp[2]};
    if (from_spectra && !fft_grad2_f64_to_f32(ctx, ctx->p64[0], ctx->p64[1], six)) {
      {
        LaunchTimer lt(ctx, KAPSEL_K_POINTWISE, 2.0 * g.f * sizeof(double) + 6.0 * g.f * sizeof(float));
        hipLaunchKernelGGL(grad_shift2_f64_kernel, dim3(blocks_for(g.nc, TPB)), dim3(TPB), 0, ctx->stream,
                           ctx->p64[0], ctx->p64[1], ge[0], ge[1], ge[2], gp[0], gp[1], gp[2], sh<double>(ctx, 0),
                           sh<double>(ctx, 1), sh<double>(ctx, 2), g);
        KAPSEL_LAUNCH_CHECK();
      }
      fft_c2r_t<T>(ctx, six, 6);
    }
  }
  if (!from_spectra || sizeof(T) == 8) {
    staggered_grad_t<T>(ctx, fld<T>(ctx, KAPSEL_EPSILON), s0, ge);
    staggered_grad_t<T>(ctx, fld<T>(ctx, KAPSEL_POTENTIAL), s0, gp);
  }
  LaunchTimer lt(ctx, KAPSEL_K_POINTWISE, alg(ctx, 15, sizeof(T)));
  hipLaunchKernelGGL(efield2_kernel<T>, dim3(blocks_for(g.nr, TPB)), dim3(TPB), 0, ctx->stream, gp[0], gp[1], gp[2],
                     e0[0], e0[1], e0[2], e2, g);
  hipLaunchKernelGGL(maxwell_kernel<T>, dim3(blocks_for(g.nr, TPB)), dim3(TPB), 0, ctx->stream, gp[0], gp[1], gp[2],
                     ge[0], ge[1], ge[2], fld<T>(ctx, KAPSEL_RHO_FREE), e2, e0[0], e0[1], e0[2],
                     fld<T>(ctx, KAPSEL_F0), fld<T>(ctx, KAPSEL_F1), fld<T>(ctx, KAPSEL_F2), g);
  KAPSEL_LAUNCH_CHECK();
}

template <typename T>
void add_force_t(kapsel_ctx *ctx, double s) {
  LaunchTimer lt(ctx, KAPSEL_K_POINTWISE, alg(ctx, 9, sizeof(T)));
  hipLaunchKernelGGL(add_force_kernel<T>, dim3(blocks_for(ctx->g.nr, TPB)), dim3(TPB), 0, ctx->stream,
                     fld<T>(ctx, KAPSEL_U0), fld<T>(ctx, KAPSEL_U1), fld<T>(ctx, KAPSEL_U2), fld<T>(ctx, KAPSEL_F0),
                     fld<T>(ctx, KAPSEL_F1), fld<T>(ctx, KAPSEL_F2), s, ctx->g);
  KAPSEL_LAUNCH_CHECK();
}

}  // namespace

void maxwell_force(kapsel_ctx *ctx, const double e0[3], bool after_solve) {
  ensure_storage(ctx);
  if (ctx->prec == KAPSEL_FP64) maxwell_t<double>(ctx, e0, false);
  else maxwell_t<float>(ctx, e0, after_solve && ctx->spec64_fresh);
}

void maxwell_force_from_spectra(kapsel_ctx *ctx, const double e0[3]) {
  ensure_storage(ctx);
  maxwell_t<float>(ctx, e0, true);
}

void add_force_to_u(kapsel_ctx *ctx, double s) {
  if (ctx->prec == KAPSEL_FP64) add_force_t<double>(ctx, s);
  else add_force_t<float>(ctx, s);
}

void dielectric_field(kapsel_ctx *ctx) {
  if (ctx->prec == KAPSEL_FP64)
    make_dielectric_field_t<double>(ctx, fld<double>(ctx, KAPSEL_EPSILON), fld<double>(ctx, KAPSEL_WORK4));
  else
    make_dielectric_field_t<float>(ctx, fld<float>(ctx, KAPSEL_EPSILON), fld<float>(ctx, KAPSEL_WORK4));
}

void poisson_apply(kapsel_ctx *ctx, const void *x, void *out) {
  ensure_storage(ctx);
  if (ctx->prec == KAPSEL_FP64) {
    double *s[4] = {(double *)ctx->pscratch[0], (double *)ctx->pscratch[1], (double *)ctx->pscratch[2],
                    (double *)ctx->pscratch[3]};
    ax_t<double>(ctx, (const double *)x, (double *)out, fld<double>(ctx, KAPSEL_EPSILON), s);
  } else {
    float *s[4] = {(float *)ctx->pscratch[0], (float *)ctx->pscratch[1], (float *)ctx->pscratch[2],
                   (float *)ctx->pscratch[3]};
    ax_t<float>(ctx, (const float *)x, (float *)out, fld<float>(ctx, KAPSEL_EPSILON), s);
  }
}

void poisson_rhs(kapsel_ctx *ctx, const double e0[3]) {
  ensure_storage(ctx);
  if (ctx->prec == KAPSEL_FP64) {
    double *s[4] = {(double *)ctx->pscratch[0], (double *)ctx->pscratch[1], (double *)ctx->pscratch[2],
                    (double *)ctx->pscratch[3]};
    build_rhs_t<double>(ctx, fld<double>(ctx, KAPSEL_RHS), fld<double>(ctx, KAPSEL_RHO_FREE),
                        fld<double>(ctx, KAPSEL_EPSILON), e0, s);
  } else {
    float *s[4] = {(float *)ctx->pscratch[0], (float *)ctx->pscratch[1], (float *)ctx->pscratch[2],
                   (float *)ctx->pscratch[3]};
    build_rhs_t<float>(ctx, fld<float>(ctx, KAPSEL_RHS), fld<float>(ctx, KAPSEL_RHO_FREE),
                       fld<float>(ctx, KAPSEL_EPSILON), e0, s);
  }
}

// x <- gamma x, gamma = (x . b) / (x . A x): of all multiples of the warm
// start x, the one closest to the solution in the energy norm of S = -A, so
// the start is never worse than x = 0 in that norm.  Builds RHS for e0.
template <typename T>
double galerkin_scale_t(kapsel_ctx *ctx, const double e0[3]) {
  const Grid &g = ctx->g;
  T *s[4] = {(T *)ctx->pscratch[0], (T *)ctx->pscratch[1], (T *)ctx->pscratch[2], (T *)ctx->pscratch[3]};
  T *x = fld<T>(ctx, KAPSEL_POTENTIAL), *b = fld<T>(ctx, KAPSEL_RHS), *ax = fld<T>(ctx, KAPSEL_WORK0);
  const T *eps = fld<T>(ctx, KAPSEL_EPSILON);
  build_rhs_t<T>(ctx, b, fld<T>(ctx, KAPSEL_RHO_FREE), eps, e0, s);
  ax_t<T>(ctx, x, ax, eps, s);
  dot_t<T, T>(ctx, x, b, S_TMP);
  const double xb = read_scal(ctx, S_TMP);
  dot_t<T, T>(ctx, x, ax, S_TMP);
  const double xax = read_scal(ctx, S_TMP);
  const double gamma = xax != 0.0 ? xb / xax : 0.0;
  hipLaunchKernelGGL(scale_real_kernel<T>, dim3(blocks_for(g.nr, TPB)), dim3(TPB), 0, ctx->stream, x, gamma, g);
  KAPSEL_LAUNCH_CHECK();
  return gamma;
}

double galerkin_scale_potential(kapsel_ctx *ctx, const double e0[3]) {
  ensure_storage(ctx);
  ctx->pot64_valid = false;
  return ctx->prec == KAPSEL_FP64 ? galerkin_scale_t<double>(ctx, e0) : galerkin_scale_t<float>(ctx, e0);
}

void poisson_solve(kapsel_ctx *ctx, const double e0[3], double tol, int maxit, int precond, int *iters,
                   double *relres) {
  ensure_storage(ctx);
  if (ctx->prec == KAPSEL_FP64) *iters = solve_fp64(ctx, e0, tol, maxit, precond, relres);
  else *iters = solve_fp32(ctx, e0, tol, maxit, precond, relres);
}

void materialize_outputs(kapsel_ctx *ctx) {
  if (!ctx->outputs_stale) return;
  ctx->outputs_stale = false;
  solve_outputs_impl(ctx);
}

void poisson_free(kapsel_ctx *ctx) {
  for (void *&p : ctx->pscratch)
    if (p) (void)hipFree(p), p = nullptr;
  for (double *&p : ctx->p64)
    if (p) (void)hipFree(p), p = nullptr;
  if (ctx->x64_prev) (void)hipFree(ctx->x64_prev), ctx->x64_prev = nullptr;
  if (ctx->d_partial) (void)hipFree(ctx->d_partial), ctx->d_partial = nullptr;
  if (ctx->h_slots) (void)hipHostFree(ctx->h_slots), ctx->h_slots = nullptr;
  ctx->h_slots_dev = nullptr;
  if (ctx->d_count) (void)hipFree(ctx->d_count), ctx->d_count = nullptr;
  for (hipEvent_t &e : ctx->slot_ev)
    if (e) (void)hipEventDestroy(e), e = nullptr;
  if (ctx->d_scal) (void)hipFree(ctx->d_scal), ctx->d_scal = nullptr;
}

}  // namespace kapsel
