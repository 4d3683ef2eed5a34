// kspace_ops.hpp -- the per-mode k-space operators of the Navier-Stokes step,
// one definition for the pointwise kernels that apply them (kspace.hip),
// alone or several in one pass.
//
// A mode is one complex value of the half spectrum, given by its KIdx; the
// wavevector is formed from the index (Calc_K*_Ooura, fft_wrapper.h:761-773).
#pragma once

#include "common.hpp"

namespace kapsel {

// Zeta_k2u_k on one mode (fft_wrapper.cxx:614-711): the velocity spectrum
// (a, b, e) from the two stored vorticity components (d1, d2); the mode
// k = 0 is left to the caller (it carries uk_dc)
template <typename T>
__device__ __forceinline__ void zeta2u_mode(const Grid &g, const KIdx &k, V2<T> d1, V2<T> d2, V2<T> &a, V2<T> &b,
                                            V2<T> &e) {
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
  a = V2<T>{-(ky * o2.im - kz * o1.im), (ky * o2.re - kz * o1.re)};
  b = V2<T>{-(kz * o0.im - kx * o2.im), (kz * o0.re - kx * o2.re)};
  e = V2<T>{-(kx * o1.im - ky * o0.im), (kx * o1.re - ky * o0.re)};
}

// U_k2zeta_k on one mode (fft_wrapper.cxx:428-497): i k x u, of which the
// two components the case rule keeps
template <typename T>
__device__ __forceinline__ void u2zeta_mode(const Grid &g, const KIdx &k, V2<T> a, V2<T> b, V2<T> e, V2<T> &za,
                                            V2<T> &zb) {
  T ks0 = (T)(k.kx * g.wave_x), ks1 = (T)(k.ky * g.wave_y), ks2 = (T)(k.kz * g.wave_z);
  V2<T> w0 = {-(ks1 * e.im - ks2 * b.im), (ks1 * e.re - ks2 * b.re)};
  V2<T> w1 = {-(ks2 * a.im - ks0 * e.im), (ks2 * a.re - ks0 * e.re)};
  V2<T> w2 = {-(ks0 * b.im - ks1 * a.im), (ks0 * b.re - ks1 * a.re)};
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
}

// the two stored components of a vorticity vector w (the case rule of
// Omega_k2zeta_k, fft_wrapper.cxx:248-284)
template <typename T>
__device__ __forceinline__ void omega_pick(const KIdx &k, const V2<T> (&w)[3], V2<T> &f0, V2<T> &f1) {
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
}

// The wavevector products of U2advection_k's k-space part
// (operate_omega.cxx:56-97) on one mode
template <typename T>
struct AdvCoef {
  T k1k2, k2k3, k3k1, c0, c1, c2;
  __device__ __forceinline__ AdvCoef(const Grid &g, const KIdx &k) {
    const T k1 = (T)(k.kx * g.wave_x), k2 = (T)(k.ky * g.wave_y), k3 = (T)(k.kz * g.wave_z);
    k1k2 = k1 * k2;
    k2k3 = k2 * k3;
    k3k1 = k3 * k1;
    c0 = (k3 - k2) * (k3 + k2);
    c1 = (k1 - k3) * (k1 + k3);
    c2 = (k2 - k1) * (k2 + k1);
  }
};

// the vorticity of the advection term from the five product spectra
// u2u3, u3u1, u1u2, u2^2-u3^2, u3^2-u1^2
template <typename T>
__device__ __forceinline__ void advection_omega(const AdvCoef<T> &q, V2<T> u2u3, V2<T> u3u1, V2<T> u1u2,
                                                V2<T> u22_u32, V2<T> u32_u12, V2<T> (&w)[3]) {
  V2<T> u12_u22 = V2<T>{-u32_u12.re - u22_u32.re, -u32_u12.im - u22_u32.im};
#define KAPSEL_OMEGA(out, p)                                                                   \
  out[0].p = -(q.k2k3 * u22_u32.p + q.c0 * u2u3.p + q.k3k1 * u1u2.p - q.k1k2 * u3u1.p);        \
  out[1].p = -(q.k3k1 * u32_u12.p + q.c1 * u3u1.p + q.k1k2 * u2u3.p - q.k2k3 * u1u2.p);        \
  out[2].p = -(q.k1k2 * u12_u22.p + q.c2 * u1u2.p + q.k2k3 * u3u1.p - q.k3k1 * u2u3.p);
  KAPSEL_OMEGA(w, re)
  KAPSEL_OMEGA(w, im)
#undef KAPSEL_OMEGA
}

// NS_solver_slavedEuler's update of zeta on a kept mode (fluid_solver.cxx:
// 346-365), dmy0 = -NU dt, dmy1 = 1/NU: the exponent and factors in double,
// the update in T
template <typename T>
__device__ __forceinline__ void slaved_update(const Grid &g, const KIdx &k, double dmy0, double dmy1, V2<T> f0,
                                              V2<T> f1, V2<T> &za, V2<T> &zb) {
  const double q1 = k.kx * g.wave_x, q2 = k.ky * g.wave_y, q3 = k.kz * g.wave_z;
  const double K2 = q1 * q1 + q2 * q2 + q3 * q3;  // K2 of Init_K (fft_wrapper.cxx:71-73)
  const double ik2 = K2 > 0.0 ? 1. / K2 : 0.0;
  const T dmy = (T)(exp(dmy0 * K2) - 1.);
  const T s = (T)(dmy1 * ik2);
  za = V2<T>{za.re + dmy * (za.re - s * f0.re), za.im + dmy * (za.im - s * f0.im)};
  zb = V2<T>{zb.re + dmy * (zb.re - s * f1.re), zb.im + dmy * (zb.im - s * f1.im)};
}

// Add_zeta_viscous_term + Field_solver_Euler on a kept mode
// (operate_omega.cxx:170-185, fluid_solver.cxx:121-136), nu = NU: f += -NU
// K2 zeta, then zeta += dt f
template <typename T>
__device__ __forceinline__ void euler_update(const Grid &g, const KIdx &k, double nu, double dt_d, V2<T> &f0,
                                             V2<T> &f1, V2<T> &za, V2<T> &zb) {
  const double q1 = k.kx * g.wave_x, q2 = k.ky * g.wave_y, q3 = k.kz * g.wave_z;
  const T nk2 = (T)(nu * (q1 * q1 + q2 * q2 + q3 * q3));  // NU * K2[im]
  const T dt = (T)dt_d;
  f0 = V2<T>{f0.re + -(nk2 * za.re), f0.im + -(nk2 * za.im)};
  f1 = V2<T>{f1.re + -(nk2 * zb.re), f1.im + -(nk2 * zb.im)};
  za = V2<T>{za.re + dt * f0.re, za.im + dt * f0.im};
  zb = V2<T>{zb.re + dt * f1.re, zb.im + dt * f1.im};
}

// Solenoidal_uk on one mode (operate_omega.cxx:187-220): u - k (k.u) / k^2
template <typename T>
__device__ __forceinline__ void solenoidal_mode(const Grid &g, const KIdx &k, V2<T> &a, V2<T> &b, V2<T> &e) {
  const T kx = (T)(k.kx * g.wave_x), ky = (T)(k.ky * g.wave_y), kz = (T)(k.kz * g.wave_z);
  const T k2 = kx * kx + ky * ky + kz * kz;
  const T ik2 = k2 > (T)0 ? (T)1 / k2 : (T)0;
  V2<T> dmy = {ik2 * (a.re * kx + b.re * ky + e.re * kz), ik2 * (a.im * kx + b.im * ky + e.im * kz)};
  a = a - scal(kx, dmy);
  b = b - scal(ky, dmy);
  e = e - scal(kz, dmy);
}

}  // namespace kapsel
