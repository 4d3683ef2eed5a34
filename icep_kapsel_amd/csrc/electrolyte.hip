// electrolyte.hip -- two ion species: Nernst-Planck transport, the Coulomb
// field of particles and ions, and the Coulomb force on the fluid.
//
// The reference's electrolyte step (fluid_solver.cxx:184-338,
// operate_electrolyte.cxx:45-242, solute_rhs.cxx:42-233) is a chain of
// whole-field loops, each reading and writing full fields.  Here the chain is
// regrouped into a few passes, each reading its inputs once:
//   charge      rho = surface * s + sum_n z_n e c_n (1 - phi_qq)  (one pass,
//               s the surface normalisation computed on the device)
//   potential   psi = rho_k / (eps k^2) (optionally 2/3-cut), grad = i k psi
//               and the applied-field shift of the k = 0 mode, one pass
//               producing the three gradient spectra
//   flux (k)    per species: the 2/3 cut of c_k, kBT L i k c_k and c_k for the
//               inverse transforms, one pass
//   flux (x)    electromigration, impermeable band, advection, one pass
//   update      divergence of the flux spectrum and the Euler step on the
//               kept modes, one pass
// Reductions (surface normalisation, ion amounts) accumulate in FP64 with a
// fixed block count, so their sums are reproducible.
#include <cmath>

#include "common.hpp"
#include "fft_prims.hpp"

namespace kapsel {

namespace {

constexpr unsigned TPB = 256;
constexpr int NBR = 1024;  // blocks of a reduction (fixed: deterministic sums)

// device scalars
enum { S_SURF = 0, S_F0, S_F1, S_C0, S_C1, S_VOL, S_FE_IDEAL, S_FE_EL, S_PB0, S_PB1, S_RHO_MEAN, S_N };
// scratch fields
enum { E_N0 = 0, E_N1, E_N2, E_QQ, E_SURF, E_CX0, E_CX1, E_RHOK, E_RHO, E_UX0, E_UX1, E_UX2, E_NF };

__device__ __forceinline__ double block_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
  __shared__ double w[TPB / 64];
  const int lane = threadIdx.x % 64, wid = threadIdx.x / 64;
  __syncthreads();
  if (lane == 0) w[wid] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(TPB / 64); i++) s += w[i];
  return s;  // valid in thread 0
}

// OP 0: sum |a|                          (Make_phi_qq_particle, :218-229)
// OP 1: sums (1 - phi) a, (1 - phi) b    (Count_single_solute, solute_rhs.cxx:130-148)
// OP 2: sum (1 - phi)                    (Set_uniform_ion_charge_density_salt, :581-591)
// OP 3: sum_n (1 - phi) c_n (log c_n - 1), a = c_0, b = c_1 (Calc_free_energy_PB, :370-381)
// OP 4: sum rho (psi / 2 - E.(i,j,k) dx), a = rho, b = psi   (Calc_free_energy_PB, :388-411)
// OP 5: sums (1 - phi) exp(-e0 psi), (1 - phi) exp(-e1 psi), b = psi
//       (Set_Poisson_Boltzmann_ion_charge_density_salt, :652-662)
// OP 6: sum a                            (total of the free charge)
template <typename T, int OP>
__global__ __launch_bounds__(TPB) void reduce_kernel(const T *__restrict__ a, const T *__restrict__ b,
                                                     const T *__restrict__ phi, double *__restrict__ partial,
                                                     Grid g, double e0 = 0.0, double e1 = 0.0, double e2 = 0.0,
                                                     int external = 0) {
  double s0 = 0.0, s1 = 0.0;
  for (size_t r = (size_t)blockIdx.x * TPB + threadIdx.x; r < g.nr; r += (size_t)gridDim.x * TPB) {
    const RI p = ridx(g, r);
    if constexpr (OP == 0) {
      s0 += fabs((double)a[p.im]);
    } else if constexpr (OP == 1) {
      const double w = 1. - (double)phi[p.im];
      s0 += w * (double)a[p.im];
      s1 += w * (double)b[p.im];
    } else if constexpr (OP == 2) {
      s0 += 1. - (double)phi[p.im];
    } else if constexpr (OP == 3) {
      const double w = 1. - (double)phi[p.im];
      const double c0 = (double)a[p.im], c1 = (double)b[p.im];
      s0 += w * c0 * (log(c0) - 1.);
      s0 += w * c1 * (log(c1) - 1.);
    } else if constexpr (OP == 5) {
      const double w = 1. - (double)phi[p.im], v = (double)b[p.im];
      s0 += w * exp(-e0 * v);
      s1 += w * exp(-e1 * v);
    } else if constexpr (OP == 6) {
      s0 += (double)a[p.im];
    } else {
      double v = (double)b[p.im] * 0.5;
      if (external) v -= ((e0 * (double)p.i + e1 * (double)p.j + e2 * (double)p.k) * g.dx);
      s0 += (double)a[p.im] * v;
    }
  }
  s0 = block_sum(s0);
  if constexpr (OP == 1 || OP == 5) s1 = block_sum(s1);
  if (threadIdx.x == 0) {
    partial[blockIdx.x] = s0;
    partial[NBR + blockIdx.x] = s1;
  }
}

// MODE 0: scal[S_SURF] = abs_total / (sum dx^3)     (Make_phi_qq_particle, :230)
// MODE 1: counts c_n = sum_n dx^3, factors total_n / c_n   (Rescale_solute, :117-118)
// MODE 2: scal[S_VOL] = sum
// MODE 6: scal[S_RHO_MEAN] = sum * abs_total (abs_total carries 1 / points)
__global__ __launch_bounds__(TPB) void finish_kernel(const double *__restrict__ partial, double *__restrict__ scal,
                                                     int mode, double dx3, double abs_total, double t0,
                                                     double t1) {
  double s0 = 0.0, s1 = 0.0;
  for (int i = threadIdx.x; i < NBR; i += TPB) {
    s0 += partial[i];
    s1 += partial[NBR + i];
  }
  s0 = block_sum(s0);
  s1 = block_sum(s1);
  if (threadIdx.x != 0) return;
  if (mode == 0) {
    scal[S_SURF] = abs_total / (s0 * dx3);
  } else if (mode == 1) {
    const double c0 = s0 * dx3, c1 = s1 * dx3;
    scal[S_C0] = c0;
    scal[S_C1] = c1;
    scal[S_F0] = t0 / c0;
    scal[S_F1] = t1 / c1;
  } else if (mode == 2) {
    scal[S_VOL] = s0;
  } else if (mode == 3) {
    scal[S_FE_IDEAL] = s0 * abs_total;  // abs_total carries kBT DX3 here
  } else if (mode == 5) {
    scal[S_PB0] = s0;
    scal[S_PB1] = s1;
  } else if (mode == 6) {
    scal[S_RHO_MEAN] = s0 * abs_total;
  } else {
    scal[S_FE_EL] = s0 * dx3;
  }
}

// a *= scal[slot] over the real points (the surface normalisation, :231-240)
template <typename T>
__global__ void scale_real_kernel(T *__restrict__ a, const double *__restrict__ scal, int slot, Grid g) {
  size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= g.nr) return;
  const RI p = ridx(g, r);
  a[p.im] *= (T)scal[slot];
}

// c_k *= factor over the whole spectrum (Rescale_solute, :119-126); the real
// copy of the same species follows, so the Coulomb force can reuse it
template <typename T>
__global__ void rescale_kernel(T *__restrict__ c0k, T *__restrict__ c1k, const double *__restrict__ scal, Grid g) {
  size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= g.f) return;
  c0k[e] *= (T)scal[S_F0];
  c1k[e] *= (T)scal[S_F1];
}

// Conc_k2charge_field (operate_electrolyte.cxx:45-74): the normalised surface
// charge plus sum_n Valency_e[n] c_n (1 - phi_qq); without surf, the ions
// alone with qq the compact-sine profile (Conc_k2charge_field_no_surfase_charge,
// operate_dielectric.cxx:135-160); with scaled, the ion
// fields are multiplied by the rescale factors first (they hold the inverse
// transforms of the spectra before Rescale_solute scaled them).  The result
// goes to out (to be transformed) and, when keep is set, to keep as well.
template <typename T>
__global__ void charge_kernel(const T *__restrict__ qq, const T *__restrict__ surf, const T *__restrict__ c0,
                              const T *__restrict__ c1, const double *__restrict__ scal, double ze0, double ze1,
                              int scaled, T *__restrict__ out, T *__restrict__ keep, Grid g) {
  size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= g.nr) return;
  const RI p = ridx(g, r);
  const T one_phi = (T)1 - qq[p.im];
  T v0 = c0[p.im], v1 = c1[p.im];
  if (scaled) {
    v0 *= (T)scal[S_F0];
    v1 *= (T)scal[S_F1];
  }
  T rho = surf ? surf[p.im] * (T)scal[S_SURF] : (T)0;
  rho += (T)ze0 * v0 * one_phi;
  rho += (T)ze1 * v1 * one_phi;
  out[p.im] = rho;
  if (keep) keep[p.im] = rho;
}

// Charge_field_k2Coulomb_potential_k_PBC (operate_electrolyte.cxx:76-88),
// optionally Truncate_two_third_rule, A_k2da_k (fft_wrapper.h:755-759) and
// Add_constant_field_k (fluid_solver.cxx:231-244, dc = E_ext amp, 0 for the
// force): g_d = i k_d psi, g_d(k=0) -= dc_d
template <typename T>
__global__ void coulomb_grad_kernel(const T *__restrict__ rhok, T *__restrict__ g0, T *__restrict__ g1,
                                    T *__restrict__ g2, double idc, int trunc, double dc0, double dc1, double dc2,
                                    Grid g) {
  size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.nc) return;
  const KIdx k = kidx(g, c);
  const double wx = k.kx * g.wave_x, wy = k.ky * g.wave_y, wz = k.kz * g.wave_z;
  const double k2 = wx * wx + wy * wy + wz * wz;
  const double ik2 = k2 > 0.0 ? 1. / k2 : 0.0;  // IK2 of Init_K (fft_wrapper.cxx:71-80)
  V2<T> a = ld2(rhok + 2 * c);
  const T s = (T)(ik2 * idc);
  a = V2<T>{a.re * s, a.im * s};
  if (trunc && is_cut(g, k)) a = V2<T>{0, 0};
  V2<T> d0{-(T)wx * a.im, (T)wx * a.re}, d1{-(T)wy * a.im, (T)wy * a.re}, d2{-(T)wz * a.im, (T)wz * a.re};
  if (c == 0) {
    d0.re -= (T)dc0;
    d1.re -= (T)dc1;
    d2.re -= (T)dc2;
  }
  st2(g0 + 2 * c, d0);
  st2(g1 + 2 * c, d1);
  st2(g2 + 2 * c, d2);
}

// Per species, k-space half of the flux: Truncate_two_third_rule of c_k
// (Rhs_NS_solute, fluid_solver.cxx:209), Diffusion_flux_single's
// i k kBT L c_k (solute_rhs.cxx:212-233), and c_k for A_k2a_out (:211)
template <typename T>
__global__ void flux_k_kernel(T *__restrict__ ck, T *__restrict__ f0, T *__restrict__ f1, T *__restrict__ f2,
                              T *__restrict__ cx, double kl, Grid g) {
  size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.nc) return;
  const KIdx k = kidx(g, c);
  V2<T> a = ld2(ck + 2 * c);
  if (is_cut(g, k)) {
    // (a spectrum the previous Euler step left is zero there already)
    if (a.re != (T)0 || a.im != (T)0) st2(ck + 2 * c, V2<T>{0, 0});
    a = V2<T>{0, 0};
  }
  const V2<T> w{(T)kl * a.re, (T)kl * a.im};
  const T wx = (T)(k.kx * g.wave_x), wy = (T)(k.ky * g.wave_y), wz = (T)(k.kz * g.wave_z);
  st2(f0 + 2 * c, V2<T>{-wx * w.im, wx * w.re});
  st2(f1 + 2 * c, V2<T>{-wy * w.im, wy * w.re});
  st2(f2 + 2 * c, V2<T>{-wz * w.im, wz * w.re});
  if (cx) st2(cx + 2 * c, a);
}

// Real-space half of the flux of both species in one pass (the gradient,
// normals and velocity are read once): Solute_solver_rhs_nonlinear_x_single
// (solute_rhs.cxx:150-188), Solute_impermeability (:42-97) on every point
// whose surface normal is nonzero (exactly the points the reference's
// particle loop projects; the projection is idempotent, so points of two
// particles' cells need it once), Add_advection_flux (:190-210).  f: the
// fluxes of species 0 (p[0..2]) and 1 (p[3..5]), the diffusive part in,
// the whole flux out.  With fac, the real ion fields are the ones kept from
// before the last rescale and are multiplied by their factors fac[0..1].
template <typename T>
__global__ void flux_x2_kernel(FieldPtrs<T> f, const T *__restrict__ cx0, const T *__restrict__ cx1,
                               const double *__restrict__ fac, const T *__restrict__ g0, const T *__restrict__ g1,
                               const T *__restrict__ g2, const T *__restrict__ n0, const T *__restrict__ n1,
                               const T *__restrict__ n2, const T *__restrict__ u0, const T *__restrict__ u1,
                               const T *__restrict__ u2, double on0, double on1, double ze0, double ze1, int advect,
                               Grid g) {
  size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= g.nr) return;
  const size_t im = ridx(g, r).im;
  const T gr0 = g0[im], gr1 = g1[im], gr2 = g2[im];
  const T m0 = n0[im], m1 = n1[im], m2 = n2[im];
  const T norm2 = m0 * m0 + m1 * m1 + m2 * m2;
  T v0 = 0, v1 = 0, v2 = 0;
  if (advect) {
    v0 = u0[im];
    v1 = u1[im];
    v2 = u2[im];
  }
#pragma unroll
  for (int n = 0; n < 2; n++) {
    T *__restrict__ f0 = f.p[3 * n], *__restrict__ f1 = f.p[3 * n + 1], *__restrict__ f2 = f.p[3 * n + 2];
    T c = (n ? cx1 : cx0)[im];
    if (fac) c *= (T)fac[n];
    const T L = (T)(n ? on1 : on0), z = (T)(n ? ze1 : ze0);
    T a0 = f0[im] + c * (L * (gr0 * z));
    T a1 = f1[im] + c * (L * (gr1 * z));
    T a2 = f2[im] + c * (L * (gr2 * z));
    if (norm2 > (T)0) {
      const T dmy = (a0 * m0 + a1 * m1 + a2 * m2) / norm2;
      a0 -= m0 * dmy;
      a1 -= m1 * dmy;
      a2 -= m2 * dmy;
    }
    if (advect) {
      a0 += c * -v0;
      a1 += c * -v1;
      a2 += c * -v2;
    }
    f0[im] = a0;
    f1[im] = a1;
    f2[im] = a2;
  }
}

// The first (x, inverse) pass of the inverse transforms of the Coulomb
// potential's gradient (coulomb_grad_kernel), the symbol fused in: workgroup
// field d (grid.y) forms out_d = IFFT_x(i k_d b) - dc_d on the line of k = 0,
// b = a IK2 p0 (cut modes zeroed if trunc), as it reads the line of a (x line
// q = j*h + kz: elements q + i*plane), operation for operation as the
// pointwise kernel.  The fields of one tile run at consecutive ids
// (xcd_tile_by_tile), so a's lines come from HBM once.  FP32 only (the FP64
// context keeps the separate passes).  Against the separate kernel and pass
// (43 + 76 us at 256^3): 79 us with a just written, 120 us with a cold.
template <typename T, int L, int LPB>
__global__ __launch_bounds__((LPB * LineFFT<T, L>::S)) void el_x_grad_kernel(
    const T *__restrict__ a, FieldPtrs<T> out, long long Q, long long plane, Grid g, double p0, int trunc,
    double dc0, double dc1, double dc2, const V2<T> *__restrict__ tw) {
  using F = LineFFT<T, L>;
  constexpr int P = F::P;
  __shared__ V2<T> sm[LPB * F::LS];
  unsigned d;
  const unsigned tile = xcd_tile_by_tile(d);
  const int t = threadIdx.x;
  const int l = t % LPB, s = t / LPB;
  const long long q = (long long)tile * LPB + l;
  const bool valid = q < Q;
  const long long base = valid ? q : 0;
  KIdx k;
  k.j = (int)(base / g.h);
  k.kz = (int)(base - (long long)k.j * g.h);
  k.ky = k.j > g.hny ? k.j - g.ny : k.j;
  V2<T> v[P], twr[P];
#pragma unroll
  for (int n = 0; n < P; n++) {
    k.i = F::in_i(s, n);
    k.kx = k.i > g.hnx ? k.i - g.nx : k.i;
    V2<T> x = ld2(a + 2 * (base + (long long)k.i * plane));
    const double wx = k.kx * g.wave_x, wy = k.ky * g.wave_y, wz = k.kz * g.wave_z;
    const double k2 = wx * wx + wy * wy + wz * wz;
    const double ik2 = k2 > 0.0 ? 1. / k2 : 0.0;
    const T sc = (T)(ik2 * p0);
    x = V2<T>{x.re * sc, x.im * sc};
    if (trunc && is_cut(g, k)) x = V2<T>{0, 0};
    const T w = d == 0 ? (T)wx : d == 1 ? (T)wy : (T)wz;
    x = V2<T>{-w * x.im, w * x.re};
    if (base == 0 && k.i == 0) x.re -= (T)(d == 0 ? dc0 : d == 1 ? dc1 : dc2);
    v[n] = x;
  }
  line_twiddles<F>(twr, tw, s);
  F::template run<+1>(v, sm + l * F::LS, twr, s, LPB);
  if (valid) {
    T *__restrict__ od = out.p[d];
#pragma unroll
    for (int m = 0; m < F::P; m++) st2(od + 2 * (base + (long long)F::out_i(s, m) * plane), v[m]);
  }
}

// coulomb_grad_t's kernel and the first pass of the inverse transform of g3
template <typename T>
void el_x_grad(kapsel_ctx *ctx, const T *rhok, void *const *g3, int trunc, const double dc[3]) {
  static_assert(sizeof(T) == 4, "FP32 only");
  const Grid &g = ctx->g;
  constexpr int LPB = 32;
  const long long plane = (long long)g.ny * g.h;
  FieldPtrs<T> fo{};
  for (int i = 0; i < 3; i++) fo.p[i] = (T *)g3[i];
  const V2<T> *tw = (const V2<T> *)ctx->tw[prec_of<T>()].x;
  const dim3 grid((unsigned)((plane + LPB - 1) / LPB), 3);
  const double p0 = 1. / ctx->el.dielectric_cst;
  LaunchTimer lt(ctx, KAPSEL_K_FFT_LINES_INV, 4.0 * g.f * sizeof(T));
  switch (g.nx) {
#define KAPSEL_XGRAD(L)                                                                                     \
  case L:                                                                                                   \
    hipLaunchKernelGGL((el_x_grad_kernel<T, L, LPB>), grid, dim3(LPB * LineFFT<T, L>::S), 0, ctx->stream,     \
                       rhok, fo, plane, plane, g, p0, trunc, dc[0], dc[1], dc[2], tw);                      \
    break;
    KAPSEL_XGRAD(16)
    KAPSEL_XGRAD(32)
    KAPSEL_XGRAD(64)
    KAPSEL_XGRAD(128)
    KAPSEL_XGRAD(256)
#undef KAPSEL_XGRAD
    default: set_error("unsupported NX %d", g.nx); throw Error{KAPSEL_ERR_ARG};
  }
  KAPSEL_LAUNCH_CHECK();
}

// psi_k = rho_k IK2 / Dielectric_cst (Charge_field_k2Coulomb_potential_k_PBC,
// operate_electrolyte.cxx:76-88), in place
template <typename T>
__global__ void potential_k_kernel(T *__restrict__ a, double idc, Grid g) {
  size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.nc) return;
  const KIdx k = kidx(g, c);
  const double wx = k.kx * g.wave_x, wy = k.ky * g.wave_y, wz = k.kz * g.wave_z;
  const double k2 = wx * wx + wy * wy + wz * wz;
  const T s = (T)((k2 > 0.0 ? 1. / k2 : 0.0) * idc);
  V2<T> v = ld2(a + 2 * c);
  st2(a + 2 * c, V2<T>{v.re * s, v.im * s});
}

// U_k2divergence_k (fft_wrapper.cxx:896-920) of the flux spectrum and
// Field_solver_Euler (fluid_solver.cxx:121-136) of c_k on the kept modes
// (the cut modes were zeroed by flux_k_kernel)
template <typename T>
__global__ void div_euler_kernel(const T *__restrict__ f0, const T *__restrict__ f1, const T *__restrict__ f2,
                                 T *__restrict__ ck, double dt, Grid g) {
  size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.nc) return;
  const KIdx k = kidx(g, c);
  if (is_cut(g, k)) return;
  const T kx = (T)(k.kx * g.wave_x), ky = (T)(k.ky * g.wave_y), kz = (T)(k.kz * g.wave_z);
  const V2<T> a = ld2(f0 + 2 * c), b = ld2(f1 + 2 * c), e = ld2(f2 + 2 * c);
  const V2<T> div{-(kx * a.im + ky * b.im + kz * e.im), (kx * a.re + ky * b.re + kz * e.re)};
  V2<T> v = ld2(ck + 2 * c);
  v.re += (T)dt * div.re;
  v.im += (T)dt * div.im;
  st2(ck + 2 * c, v);
}

// Make_Coulomb_force_x_on_fluid, last loops (operate_electrolyte.cxx:143-176):
// f_d = rho (-d_d psi + E_d)
template <typename T>
__global__ void coulomb_force_kernel(T *__restrict__ f0, T *__restrict__ f1, T *__restrict__ f2,
                                     const T *__restrict__ rho, double e0, double e1, double e2, Grid g) {
  size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= g.nr) return;
  const size_t im = ridx(g, r).im;
  const T q = rho[im];
  T v = -f0[im];
  v += (T)e0;
  f0[im] = q * v;
  v = -f1[im];
  v += (T)e1;
  f1[im] = q * v;
  v = -f2[im];
  v += (T)e2;
  f2[im] = q * v;
}

// Rhs_NS_Nernst_Planck, Dielectric branch (fluid_solver.cxx:261-270): the
// potential's gradient on the staggered faces, i k_d e^{i k_d dx/2} psi_k
// (A_k2da_k, Shift_vec_fw_imag), the k = 0 mode shifted by the applied field
// (Add_constant_field_k, :231-244)
template <typename T>
__global__ void dielectric_grad_k_kernel(const T *__restrict__ psik, T *__restrict__ g0, T *__restrict__ g1,
                                         T *__restrict__ g2, const V2<T> *__restrict__ shx,
                                         const V2<T> *__restrict__ shy, const V2<T> *__restrict__ shz, double dc0,
                                         double dc1, double dc2, Grid g) {
  size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.nc) return;
  const KIdx k = kidx(g, c);
  const V2<T> x = ld2(psik + 2 * c);
  const T w0 = (T)(k.kx * g.wave_x), w1 = (T)(k.ky * g.wave_y), w2 = (T)(k.kz * g.wave_z);
  V2<T> d0 = cmul(V2<T>{-w0 * x.im, w0 * x.re}, shx[k.i]);
  V2<T> d1 = cmul(V2<T>{-w1 * x.im, w1 * x.re}, shy[k.j]);
  V2<T> d2 = cmul(V2<T>{-w2 * x.im, w2 * x.re}, shz[k.kz]);
  if (c == 0) {
    d0.re -= (T)dc0;
    d1.re -= (T)dc1;
    d2.re -= (T)dc2;
  }
  st2(g0 + 2 * c, d0);
  st2(g1 + 2 * c, d1);
  st2(g2 + 2 * c, d2);
}

// Interpolate_vec_on_normal_grid (operate_dielectric.cxx:32-55): the face
// values averaged onto the cells, f_d(i) = g_d(i) / 2 + g_d(i - e_d) / 2
template <typename T>
__global__ void normal_grid_kernel(const T *__restrict__ g0, const T *__restrict__ g1, const T *__restrict__ g2,
                                   T *__restrict__ f0, T *__restrict__ f1, T *__restrict__ f2, Grid g) {
  size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= g.nr) return;
  const RI p = ridx(g, r);
  const int im1 = p.i == 0 ? g.nx - 1 : p.i - 1, jm1 = p.j == 0 ? g.ny - 1 : p.j - 1;
  const int km1 = p.k == 0 ? g.nz - 1 : p.k - 1;
  const size_t xm = ((size_t)im1 * g.ny + p.j) * g.nz_ + p.k;
  const size_t ym = ((size_t)p.i * g.ny + jm1) * g.nz_ + p.k;
  const size_t zm = ((size_t)p.i * g.ny + p.j) * g.nz_ + km1;
  f0[p.im] = (T)0.5 * g0[p.im] + (T)0.5 * g0[xm];
  f1[p.im] = (T)0.5 * g1[p.im] + (T)0.5 * g1[ym];
  f2[p.im] = (T)0.5 * g2[p.im] + (T)0.5 * g2[zm];
}

// U2advection_k, real-space part (operate_omega.cxx:30-47), from the kept
// real velocity: products into U0..U2 and WORK0/WORK1
template <typename T>
__global__ void products_kernel(const T *__restrict__ x0, const T *__restrict__ x1, const T *__restrict__ x2,
                                T *__restrict__ u0, T *__restrict__ u1, T *__restrict__ u2, T *__restrict__ a0,
                                T *__restrict__ a1, Grid g) {
  size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= g.nr) return;
  const size_t im = ridx(g, r).im;
  const T v1 = x0[im], v2 = x1[im], v3 = x2[im];
  u0[im] = v2 * v3;
  u1[im] = v3 * v1;
  u2[im] = v1 * v2;
  a0[im] = (v2 - v3) * (v2 + v3);
  a1[im] = (v3 - v1) * (v3 + v1);
}

// Set_Poisson_Boltzmann_ion_charge_density_salt, one fixed-point iteration's
// pointwise parts (operate_electrolyte.cxx:670-695): the ion densities from
// the potential, and the charge density they make with the particle's
// surface charge
template <typename T>
__global__ void pb_conc_charge_kernel(const T *__restrict__ ep, T *__restrict__ c0, T *__restrict__ c1,
                                      const T *__restrict__ qq, const T *__restrict__ surf,
                                      const double *__restrict__ scal, double rp, double rn, double d0, double d1,
                                      double nu, double ze0, double ze1, T *__restrict__ rho, Grid g) {
  size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= g.nr) return;
  const size_t im = ridx(g, r).im;
  const double pot = (double)ep[im];
  const T v0 = (T)(rp * exp(-d0 * pot) * nu), v1 = (T)(rn * exp(-d1 * pot) / nu);
  c0[im] = v0;
  c1[im] = v1;
  const T one_phi = (T)1 - qq[im];
  T q = surf[im] * (T)scal[S_SURF];
  q += (T)ze0 * v0 * one_phi;
  q += (T)ze1 * v1 * one_phi;
  rho[im] = q;
}

// the damped update of the potential, with sum |psi| and sum |d psi| per block
// (operate_electrolyte.cxx:703-716)
template <typename T>
__global__ __launch_bounds__(TPB) void pb_update_kernel(T *__restrict__ ep, const T *__restrict__ pot, double alp,
                                                        double *__restrict__ partial, Grid g) {
  double s0 = 0.0, s1 = 0.0;
  for (size_t r = (size_t)blockIdx.x * TPB + threadIdx.x; r < g.nr; r += (size_t)gridDim.x * TPB) {
    const size_t im = ridx(g, r).im;
    const T dmy = (T)alp * (pot[im] - ep[im]);
    const T v = ep[im] + dmy;
    ep[im] = v;
    s0 += fabs((double)v);
    s1 += fabs((double)dmy);
  }
  s0 = block_sum(s0);
  s1 = block_sum(s1);
  if (threadIdx.x == 0) {
    partial[blockIdx.x] = s0;
    partial[NBR + blockIdx.x] = s1;
  }
}

// uniform ion densities on the real points, 0 in the row padding
// (Set_uniform_ion_charge_density_salt, :604-613)
template <typename T>
__global__ void fill_kernel(T *__restrict__ c0, T *__restrict__ c1, double v0, double v1, Grid g) {
  size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= g.f) return;
  const bool real = (int)(e % (unsigned)g.nz_) < g.nz;
  c0[e] = real ? (T)v0 : (T)0;
  c1[e] = real ? (T)v1 : (T)0;
}

double kept_count_el(const Grid &g) { return (double)(2 * g.trn_x - 1) * (2 * g.trn_y - 1) * g.trn_z; }
double cut_count_el(const Grid &g) { return (double)g.nc - kept_count_el(g); }

template <typename T>
T *fl(kapsel_ctx *ctx, int f) {
  return (T *)ctx->fields[f];
}
template <typename T>
T *ef(kapsel_ctx *ctx, int f) {
  return (T *)ctx->el_f[f];
}

double valency_e(const kapsel_ctx *ctx, int n) { return ctx->el.valency[n] * ctx->el.elementary_charge; }
double ac_factor(const kapsel_ctx *ctx, const kapsel_time &t) {
  return ctx->el.ac ? sin(ctx->el.angular_frequency * t.time) : 1.0;
}

void need_el(const kapsel_ctx *ctx) {
  if (!ctx->el_set) {
    set_error("electrolyte parameters not set (kapsel_set_electrolyte)");
    throw Error{KAPSEL_ERR_STATE};
  }
}

// the Dielectric electrolyte: one Dielectric_cst for the ions and the
// permittivity (the reference has one global)
void need_el_dielectric(const kapsel_ctx *ctx) {
  need_el(ctx);
  if (!ctx->el.dielectric) {
    set_error("the electrolyte was set without Dielectric");
    throw Error{KAPSEL_ERR_STATE};
  }
  if (ctx->el.dielectric_cst != ctx->P.dielectric_cst) {
    set_error("Dielectric: the electrolyte's Dielectric_cst (%g) differs from the permittivity's (%g, "
              "kapsel_set_dielectric)",
              ctx->el.dielectric_cst, ctx->P.dielectric_cst);
    throw Error{KAPSEL_ERR_ARG};
  }
}

// the branches that exist only without Dielectric
void need_el_plain(const kapsel_ctx *ctx, const char *what) {
  need_el(ctx);
  if (ctx->el.dielectric) {
    set_error("%s: not a Dielectric branch", what);
    throw Error{KAPSEL_ERR_STATE};
  }
}

template <typename T, int OP>
void reduce(kapsel_ctx *ctx, const void *a, const void *b, const void *phi) {
  const double fields = OP == 1 ? 3.0 : 1.0;
  LaunchTimer lt(ctx, KAPSEL_K_POINTWISE, fields * ctx->g.nr * sizeof(T));
  hipLaunchKernelGGL((reduce_kernel<T, OP>), dim3(NBR), dim3(TPB), 0, ctx->stream, (const T *)a, (const T *)b,
                     (const T *)phi, ctx->el_partial, ctx->g);
  KAPSEL_LAUNCH_CHECK();
}

void finish(kapsel_ctx *ctx, int mode, double abs_total = 0.0, double t0 = 0.0, double t1 = 0.0) {
  const double dx3 = ctx->g.dx * ctx->g.dx * ctx->g.dx;
  hipLaunchKernelGGL(finish_kernel, dim3(1), dim3(TPB), 0, ctx->stream, ctx->el_partial, ctx->el_scal, mode, dx3,
                     abs_total, t0, t1);
  KAPSEL_LAUNCH_CHECK();
}

double read_scal(kapsel_ctx *ctx, int slot) {
  double v = 0.0;
  KAPSEL_HIP(hipMemcpyAsync(&v, ctx->el_scal + slot, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  KAPSEL_HIP(hipStreamSynchronize(ctx->stream));
  return v;
}

// Make_phi_qq_particle (operate_electrolyte.cxx:179-242) without its final
// scaling pass: the profile sum and the raw surface charge, then the
// normalisation factor into scal[S_SURF]
template <typename T>
void phi_qq_t(kapsel_ctx *ctx, T *phi, T *surf) {
  const double qe = ctx->el.surface_charge * ctx->el.elementary_charge;
  make_phi_qq_raw(ctx, phi, surf, qe);
  reduce<T, 0>(ctx, surf, nullptr, nullptr);
  finish(ctx, 0, ctx->P.n_particles * fabs(qe));
}

// ion concentrations in real space: E_CX0/E_CX1 = c2r(CONC0/CONC1)
template <typename T>
void ions_real_t(kapsel_ctx *ctx) {
  void *two[2] = {ctx->el_f[E_CX0], ctx->el_f[E_CX1]};
  fft_c2r_t<T>(ctx, two, 2, ctx->fields + KAPSEL_CONC0);
}

// Conc_k2charge_field into out (and keep); the ion fields must hold the
// inverse transforms already (scaled: times the rescale factors)
template <typename T>
void charge_t(kapsel_ctx *ctx, T *out, T *keep, int scaled) {
  phi_qq_t<T>(ctx, ef<T>(ctx, E_QQ), ef<T>(ctx, E_SURF));
  LaunchTimer lt(ctx, KAPSEL_K_POINTWISE, (keep ? 6.0 : 5.0) * ctx->g.nr * sizeof(T));
  hipLaunchKernelGGL(charge_kernel<T>, dim3(blocks_for(ctx->g.nr, TPB)), dim3(TPB), 0, ctx->stream,
                     ef<T>(ctx, E_QQ), ef<T>(ctx, E_SURF), ef<T>(ctx, E_CX0), ef<T>(ctx, E_CX1), ctx->el_scal,
                     valency_e(ctx, 0), valency_e(ctx, 1), scaled, out, keep, ctx->g);
  KAPSEL_LAUNCH_CHECK();
}

template <typename T>
void coulomb_grad_t(kapsel_ctx *ctx, const T *rhok, T *const *g3, int trunc, const double dc[3]) {
  LaunchTimer lt(ctx, KAPSEL_K_POINTWISE, 4.0 * ctx->g.f * sizeof(T));
  hipLaunchKernelGGL(coulomb_grad_kernel<T>, dim3(blocks_for(ctx->g.nc, TPB)), dim3(TPB), 0, ctx->stream, rhok,
                     g3[0], g3[1], g3[2], 1. / ctx->el.dielectric_cst, trunc, dc[0], dc[1], dc[2], ctx->g);
  KAPSEL_LAUNCH_CHECK();
}

// Rhs_NS_Nernst_Planck + Rhs_NS_solute + Field_solver_Euler
// (fluid_solver.cxx:184-338), non-Dielectric branch.  Buffers: grad_potential
// in F0..F2 and solute_flux in UP0..UP2 as in the reference's call (:328),
// the second species' flux in WORK0..WORK2, the real c_n in E_CX0/E_CX1,
// the right-hand side of zeta in WORK0/WORK1 after the fluxes.
// reuse: the previous electrolyte step ended right before, so E_RHOK holds
// the charge spectrum of the current ions and particles (its Coulomb force),
// E_CX0/E_CX1 times the rescale factors are the real ion fields (the spectra
// are 2/3-cut already: the Euler step only touches kept modes and the cut
// modes were zeroed), and E_UX0..2 the real velocity whose spectrum zeta
// holds (solenoidal and cut, so Zeta_k2u gives it back).  ions_only: the
// Nernst-Planck part with the fluid at rest and zeta untouched, as
// Ion_diffusion_solver_Euler runs for Set_steadystate_ion_density
// (fluid_solver.cxx:293-316, operate_electrolyte.cxx:436-444; the solvent
// right-hand side it also computes is discarded there).
// Under Dielectric (never with reuse) the gradient is the one of the
// POTENTIAL field, the last solve's answer in the variable permittivity.
template <typename T>
void nssolute_t(kapsel_ctx *ctx, const kapsel_time &t, bool reuse, bool ions_only = false) {
  const Grid &g = ctx->g;
  const bool dielectric = ctx->el.dielectric != 0;
  T *rhok = ef<T>(ctx, E_RHOK);
  if (!reuse && !dielectric) {
    // charge density of the current state (concentrations not yet cut)
    ions_real_t<T>(ctx);
    charge_t<T>(ctx, rhok, nullptr, 0);
    void *one[1] = {rhok};
    fft_r2c_t<T>(ctx, one, 1);
  }
  const double amp = (double)(g.nx * g.ny * g.nz) * ac_factor(ctx, t);
  const double dc[3] = {ctx->el.e_ext[0] * amp, ctx->el.e_ext[1] * amp, ctx->el.e_ext[2] * amp};
  T *grad[3] = {fl<T>(ctx, KAPSEL_F0), fl<T>(ctx, KAPSEL_F1), fl<T>(ctx, KAPSEL_F2)};
  if (dielectric) {
    // E_RHOK: the potential's spectrum; E_UX0..2: the staggered gradient
    copy_field(ctx, rhok, ctx->fields[KAPSEL_POTENTIAL]);
    void *one[1] = {rhok};
    fft_r2c_t<T>(ctx, one, 1);
    T *face[3] = {ef<T>(ctx, E_UX0), ef<T>(ctx, E_UX1), ef<T>(ctx, E_UX2)};
    const Twiddles &tw = ctx->tw[prec_of<T>()];
    {
      LaunchTimer lt(ctx, KAPSEL_K_POINTWISE, 4.0 * g.f * sizeof(T));
      hipLaunchKernelGGL(dielectric_grad_k_kernel<T>, dim3(blocks_for(g.nc, TPB)), dim3(TPB), 0, ctx->stream, rhok,
                         face[0], face[1], face[2], (const V2<T> *)tw.shx, (const V2<T> *)tw.shy,
                         (const V2<T> *)tw.shz, dc[0], dc[1], dc[2], g);
      KAPSEL_LAUNCH_CHECK();
    }
    fft_c2r_t<T>(ctx, (void *const *)face, 3);
    {
      LaunchTimer lt(ctx, KAPSEL_K_POINTWISE, 6.0 * g.nr * sizeof(T));
      hipLaunchKernelGGL(normal_grid_kernel<T>, dim3(blocks_for(g.nr, TPB)), dim3(TPB), 0, ctx->stream, face[0],
                         face[1], face[2], grad[0], grad[1], grad[2], g);
      KAPSEL_LAUNCH_CHECK();
    }
  } else if constexpr (sizeof(T) == 4) {
    // coulomb_grad_t with its inverse transform, the first pass fused
    el_x_grad<T>(ctx, rhok, (void *const *)grad, 1, dc);
    fft_c2r_yz_t<T>(ctx, (void *const *)grad, 3);
  } else {
    coulomb_grad_t<T>(ctx, rhok, grad, 1, dc);
    fft_c2r_t<T>(ctx, (void *const *)grad, 3);
  }
  // u of the truncated zeta (Rhs_NS_solute :203-207)
  T *u[3] = {nullptr, nullptr, nullptr};
  if (ions_only) {
    // u = 0: the advective flux vanishes
  } else if (reuse) {
    for (int d = 0; d < 3; d++) u[d] = ef<T>(ctx, E_UX0 + d);
  } else {
    zeta_k2u_k(ctx, 0);
    fft_c2r_t<T>(ctx, ctx->fields + KAPSEL_U0, 3);
    for (int d = 0; d < 3; d++) u[d] = fl<T>(ctx, KAPSEL_U0 + d);
  }
  make_surface_normal(ctx, ctx->el_f + E_N0);
  // the fluxes of species 0 in UP0..UP2, of species 1 in WORK0..WORK2; the
  // real (cut) ion fields are formed in E_CX0/E_CX1 unless reused
  FieldPtrs<T> flux{{fl<T>(ctx, KAPSEL_UP0), fl<T>(ctx, KAPSEL_UP1), fl<T>(ctx, KAPSEL_UP2), fl<T>(ctx, KAPSEL_WORK0),
                     fl<T>(ctx, KAPSEL_WORK1), fl<T>(ctx, KAPSEL_WORK2)}};
  for (int n = 0; n < 2; n++) {
    T *ck = fl<T>(ctx, KAPSEL_CONC0 + n);
    T *cx = ef<T>(ctx, E_CX0 + n);
    void *four[4] = {flux.p[3 * n], flux.p[3 * n + 1], flux.p[3 * n + 2], cx};
    // (fused into the first inverse pass, with the line read once per output
    // field, this measured slower at 256^3 FP32: 150 us against 60 + 76 us)
    {
      LaunchTimer lt(ctx, KAPSEL_K_POINTWISE,
                     (2.0 * g.f + 2.0 * cut_count_el(g) + (reuse ? 3.0 : 4.0) * g.f) * sizeof(T));
      hipLaunchKernelGGL(flux_k_kernel<T>, dim3(blocks_for(g.nc, TPB)), dim3(TPB), 0, ctx->stream, ck,
                         flux.p[3 * n], flux.p[3 * n + 1], flux.p[3 * n + 2], reuse ? nullptr : cx,
                         ctx->el.kbt * ctx->el.onsager_coeff[n], g);
      KAPSEL_LAUNCH_CHECK();
    }
    fft_c2r_t<T>(ctx, four, reuse ? 3 : 4);
  }
  {
    LaunchTimer lt(ctx, KAPSEL_K_POINTWISE, (ions_only ? 20.0 : 23.0) * g.nr * sizeof(T));
    hipLaunchKernelGGL(flux_x2_kernel<T>, dim3(blocks_for(g.nr, TPB)), dim3(TPB), 0, ctx->stream, flux,
                       ef<T>(ctx, E_CX0), ef<T>(ctx, E_CX1), reuse ? ctx->el_scal + S_F0 : nullptr, grad[0],
                       grad[1], grad[2], ef<T>(ctx, E_N0), ef<T>(ctx, E_N1), ef<T>(ctx, E_N2), u[0], u[1], u[2],
                       ctx->el.onsager_coeff[0], ctx->el.onsager_coeff[1], valency_e(ctx, 0), valency_e(ctx, 1),
                       ions_only ? 0 : 1, g);
    KAPSEL_LAUNCH_CHECK();
  }
  for (int n = 0; n < 2; n++) {
    void *three[3] = {flux.p[3 * n], flux.p[3 * n + 1], flux.p[3 * n + 2]};
    fft_r2c_t<T>(ctx, three, 3);
    LaunchTimer lt(ctx, KAPSEL_K_POINTWISE, 5.0 * 2.0 * kept_count_el(g) * sizeof(T));
    hipLaunchKernelGGL(div_euler_kernel<T>, dim3(blocks_for(g.nc, TPB)), dim3(TPB), 0, ctx->stream, flux.p[3 * n],
                       flux.p[3 * n + 1], flux.p[3 * n + 2], fl<T>(ctx, KAPSEL_CONC0 + n), t.dt_fluid, g);
    KAPSEL_LAUNCH_CHECK();
  }
  if (ions_only) return;
  // Rhs_solvent (U2advection_k) + viscous term + Euler step of zeta; with
  // reuse in an FP32 context the five products of the real velocity are
  // formed inside the first pass of their forward transform (same
  // arithmetic as products_kernel)
  void *five[5] = {ctx->fields[KAPSEL_U0], ctx->fields[KAPSEL_U1], ctx->fields[KAPSEL_U2], ctx->adv_w[0],
                   ctx->adv_w[1]};
  void *b5[5] = {u[1], u[2], u[0], u[1], u[2]}, *c5[5] = {u[2], u[0], u[1], u[2], u[0]};
  if (!(reuse && sizeof(T) == 4 && fft_r2c_pro_t<T>(ctx, five, nullptr, b5, c5, 1.0, 5, 0x18u))) {
    if (reuse) {
      LaunchTimer lt(ctx, KAPSEL_K_POINTWISE, 8.0 * g.nr * sizeof(T));
      hipLaunchKernelGGL(products_kernel<T>, dim3(blocks_for(g.nr, TPB)), dim3(TPB), 0, ctx->stream, u[0], u[1],
                         u[2], fl<T>(ctx, KAPSEL_U0), fl<T>(ctx, KAPSEL_U1), fl<T>(ctx, KAPSEL_U2),
                         (T *)ctx->adv_w[0], (T *)ctx->adv_w[1], g);
      KAPSEL_LAUNCH_CHECK();
    } else {
      advection_products(ctx);
    }
    fft_r2c_t<T>(ctx, five, 5);
  }
  advection_combine_euler(ctx, t.dt_fluid);
}

// Rescale_solute (solute_rhs.cxx:99-128).  PHI becomes the unclamped profile
// sum, as the reference's Make_phi_particle leaves it; at the positions of
// the last profile scatter that is PHI_SUM exactly (same kernel, same sums)
// when phi_sum_current is set.  The ion fields keep the unscaled inverse
// transforms; charge_t(scaled = 1) applies the factors.
template <typename T>
void rescale_t(kapsel_ctx *ctx, bool phi_sum_current) {
  if (phi_sum_current) copy_field(ctx, ctx->fields[KAPSEL_PHI], ctx->fields[KAPSEL_PHI_SUM]);
  else make_phi_particle(ctx, ctx->fields[KAPSEL_PHI]);
  ions_real_t<T>(ctx);
  reduce<T, 1>(ctx, ctx->el_f[E_CX0], ctx->el_f[E_CX1], ctx->fields[KAPSEL_PHI]);
  finish(ctx, 1, 0.0, ctx->el_total[0], ctx->el_total[1]);
  LaunchTimer lt(ctx, KAPSEL_K_POINTWISE, 4.0 * ctx->g.f * sizeof(T));
  hipLaunchKernelGGL(rescale_kernel<T>, dim3(blocks_for(ctx->g.f, TPB)), dim3(TPB), 0, ctx->stream,
                     fl<T>(ctx, KAPSEL_CONC0), fl<T>(ctx, KAPSEL_CONC1), ctx->el_scal, ctx->g);
  KAPSEL_LAUNCH_CHECK();
}

// Make_Coulomb_force_x_on_fluid (operate_electrolyte.cxx:90-177), free
// particles; scaled: the ion fields hold the transforms from before the last
// rescale (see rescale_t), otherwise they are recomputed
// spectrum: F0..F2 receive the force's spectrum, formed inside the first pass
// of its forward transform instead of by the pointwise pass (same arithmetic)
template <typename T>
void coulomb_force_t(kapsel_ctx *ctx, const kapsel_time &t, int scaled, bool spectrum = false) {
  if (!scaled) ions_real_t<T>(ctx);
  T *rhok = ef<T>(ctx, E_RHOK), *rho = ef<T>(ctx, E_RHO);
  charge_t<T>(ctx, rhok, rho, scaled);
  void *one[1] = {rhok};
  fft_r2c_t<T>(ctx, one, 1);
  T *f[3] = {fl<T>(ctx, KAPSEL_F0), fl<T>(ctx, KAPSEL_F1), fl<T>(ctx, KAPSEL_F2)};
  const double zero[3] = {0.0, 0.0, 0.0};
  if constexpr (sizeof(T) == 4) {
    el_x_grad<T>(ctx, rhok, (void *const *)f, 0, zero);
    fft_c2r_yz_t<T>(ctx, (void *const *)f, 3);
  } else {
    coulomb_grad_t<T>(ctx, rhok, f, 0, zero);
    fft_c2r_t<T>(ctx, (void *const *)f, 3);
  }
  const double s = ac_factor(ctx, t);
  if (spectrum) {
    // f_d = -((grad_d psi - E_d) rho) = rho (E_d - grad_d psi)
    const double me[3] = {-ctx->el.e_ext[0] * s, -ctx->el.e_ext[1] * s, -ctx->el.e_ext[2] * s};
    void *r3[3] = {rho, rho, rho};
    if (fft_r2c_pro_t<T>(ctx, (void *const *)f, nullptr, (void *const *)f, r3, -1.0, 3, 0, me)) return;
  }
  {
    LaunchTimer lt(ctx, KAPSEL_K_POINTWISE, 7.0 * ctx->g.nr * sizeof(T));
    hipLaunchKernelGGL(coulomb_force_kernel<T>, dim3(blocks_for(ctx->g.nr, TPB)), dim3(TPB), 0, ctx->stream, f[0],
                       f[1], f[2], rho, ctx->el.e_ext[0] * s, ctx->el.e_ext[1] * s, ctx->el.e_ext[2] * s, ctx->g);
    KAPSEL_LAUNCH_CHECK();
  }
  if (spectrum) fft_r2c_t<T>(ctx, (void *const *)f, 3);
}

// Calc_free_energy_PB, non-Dielectric branch (operate_electrolyte.cxx:302-421):
// scal[S_FE_IDEAL], scal[S_FE_EL].  phi is the unclamped profile sum, which
// the phi_qq scatter of the charge computes as well.
template <typename T>
void free_energy_t(kapsel_ctx *ctx, const kapsel_time &t) {
  const Grid &g = ctx->g;
  ions_real_t<T>(ctx);
  T *rhok = ef<T>(ctx, E_RHOK), *rho = ef<T>(ctx, E_RHO);
  charge_t<T>(ctx, rhok, rho, 0);  // E_QQ: profile sum
  {
    LaunchTimer lt(ctx, KAPSEL_K_POINTWISE, 3.0 * g.nr * sizeof(T));
    hipLaunchKernelGGL((reduce_kernel<T, 3>), dim3(NBR), dim3(TPB), 0, ctx->stream, ef<T>(ctx, E_CX0),
                       ef<T>(ctx, E_CX1), ef<T>(ctx, E_QQ), ctx->el_partial, g, 0.0, 0.0, 0.0, 0);
    KAPSEL_LAUNCH_CHECK();
  }
  const double dx3 = g.dx * g.dx * g.dx;
  finish(ctx, 3, ctx->el.kbt * dx3);
  void *one[1] = {rhok};
  fft_r2c_t<T>(ctx, one, 1);
  {
    LaunchTimer lt(ctx, KAPSEL_K_POINTWISE, 2.0 * g.f * sizeof(T));
    hipLaunchKernelGGL(potential_k_kernel<T>, dim3(blocks_for(g.nc, TPB)), dim3(TPB), 0, ctx->stream, rhok,
                       1. / ctx->el.dielectric_cst, g);
    KAPSEL_LAUNCH_CHECK();
  }
  fft_c2r_t<T>(ctx, one, 1);
  const double s = ac_factor(ctx, t);
  const int external = ctx->el.e_ext[0] != 0. || ctx->el.e_ext[1] != 0. || ctx->el.e_ext[2] != 0.;
  {
    LaunchTimer lt(ctx, KAPSEL_K_POINTWISE, 2.0 * g.nr * sizeof(T));
    hipLaunchKernelGGL((reduce_kernel<T, 4>), dim3(NBR), dim3(TPB), 0, ctx->stream, rho, rhok, (const T *)nullptr,
                       ctx->el_partial, g, ctx->el.e_ext[0] * s, ctx->el.e_ext[1] * s, ctx->el.e_ext[2] * s,
                       external);
    KAPSEL_LAUNCH_CHECK();
  }
  finish(ctx, 4);
}

template <typename T>
void free_energy_host_t(kapsel_ctx *ctx, const kapsel_time &t, double *fe) {
  free_energy_t<T>(ctx, t);
  fe[1] = read_scal(ctx, S_FE_IDEAL);
  fe[2] = read_scal(ctx, S_FE_EL);
  fe[0] = fe[1] + fe[2];
}

// Set_steadystate_ion_density (operate_electrolyte.cxx:424-457): relax the
// ions with the fluid at rest until the free energy changes by less than
// TOL = 1e-4 (operate_electrolyte.h:22), at most max_iter - 1 steps
template <typename T>
int steadystate_t(kapsel_ctx *ctx, const kapsel_time &t, int max_iter, double *fe) {
  free_energy_host_t<T>(ctx, t, fe);
  double previous = fe[0];
  int steps = 0;
  for (int m = 1; m < max_iter; m++) {
    nssolute_t<T>(ctx, t, false, true);
    rescale_t<T>(ctx, false);
    steps++;
    free_energy_host_t<T>(ctx, t, fe);
    if (fabs(fe[0] - previous) < 1.e-4) break;
    previous = fe[0];
  }
  return steps;
}

// Set_Poisson_Boltzmann_ion_charge_density_salt (operate_electrolyte.cxx:
// 623-736): one particle, z:z salt; damped fixed point of the potential
// (alp = 0.01) until sum|d psi| / sum|psi| < TOL = 1e-4, at most max_iter - 1
// iterations; the ion spectra at the end.  The particle's profile sum and
// normalised surface charge are the same every iteration (the particle does
// not move), so they are made once.
template <typename T>
int pb_salt_single_t(kapsel_ctx *ctx, int max_iter) {
  const kapsel_electrolyte &e = ctx->el;
  const Grid &g = ctx->g;
  const double rho_inf = e.kbt * e.dielectric_cst / ((e.elementary_charge * e.debye_length) *
                                                      (e.elementary_charge * e.debye_length));
  const double rp = rho_inf / (e.valency[0] * (e.valency[0] - e.valency[1]));
  const double rn = -rho_inf / (e.valency[1] * (e.valency[0] - e.valency[1]));
  const double alp = 0.01;
  const double d0 = e.elementary_charge * e.valency[0] / e.kbt, d1 = e.elementary_charge * e.valency[1] / e.kbt;
  const double surface_ion_number = -e.surface_charge * ctx->P.n_particles;
  const double dx3 = g.dx * g.dx * g.dx;
  make_phi_particle(ctx, ctx->fields[KAPSEL_PHI]);
  phi_qq_t<T>(ctx, ef<T>(ctx, E_QQ), ef<T>(ctx, E_SURF));
  T *ep = ef<T>(ctx, E_UX0), *rhok = ef<T>(ctx, E_RHOK);
  T *c0 = fl<T>(ctx, KAPSEL_CONC0), *c1 = fl<T>(ctx, KAPSEL_CONC1);
  KAPSEL_HIP(hipMemsetAsync(ep, 0, g.f * sizeof(T), ctx->stream));
  int iters = 0;
  for (int m = 1; m < max_iter; m++) {
    iters++;
    {
      LaunchTimer lt(ctx, KAPSEL_K_POINTWISE, 2.0 * g.nr * sizeof(T));
      hipLaunchKernelGGL((reduce_kernel<T, 5>), dim3(NBR), dim3(TPB), 0, ctx->stream, (const T *)nullptr, ep,
                         fl<T>(ctx, KAPSEL_PHI), ctx->el_partial, g, d0, d1, 0.0, 0);
      KAPSEL_LAUNCH_CHECK();
    }
    finish(ctx, 5);
    const double sp = read_scal(ctx, S_PB0) * dx3, sn = read_scal(ctx, S_PB1) * dx3;
    const double a = e.valency[0] * rp * sp, b = -surface_ion_number, c = e.valency[1] * rn * sn;
    const double nu = (-b + sqrt(b * b - 4. * a * c)) / (2. * a);
    {
      LaunchTimer lt(ctx, KAPSEL_K_POINTWISE, 6.0 * g.nr * sizeof(T));
      hipLaunchKernelGGL(pb_conc_charge_kernel<T>, dim3(blocks_for(g.nr, TPB)), dim3(TPB), 0, ctx->stream, ep, c0,
                         c1, ef<T>(ctx, E_QQ), ef<T>(ctx, E_SURF), ctx->el_scal, rp, rn, d0, d1, nu,
                         valency_e(ctx, 0), valency_e(ctx, 1), rhok, g);
      KAPSEL_LAUNCH_CHECK();
    }
    void *one[1] = {rhok};
    fft_r2c_t<T>(ctx, one, 1);
    {
      LaunchTimer lt(ctx, KAPSEL_K_POINTWISE, 2.0 * g.f * sizeof(T));
      hipLaunchKernelGGL(potential_k_kernel<T>, dim3(blocks_for(g.nc, TPB)), dim3(TPB), 0, ctx->stream, rhok,
                         1. / e.dielectric_cst, g);
      KAPSEL_LAUNCH_CHECK();
    }
    fft_c2r_t<T>(ctx, one, 1);
    {
      LaunchTimer lt(ctx, KAPSEL_K_POINTWISE, 3.0 * g.nr * sizeof(T));
      hipLaunchKernelGGL(pb_update_kernel<T>, dim3(NBR), dim3(TPB), 0, ctx->stream, ep, rhok, alp, ctx->el_partial,
                         g);
      KAPSEL_LAUNCH_CHECK();
    }
    finish(ctx, 5);
    if (read_scal(ctx, S_PB1) / read_scal(ctx, S_PB0) < 1.e-4) break;
  }
  void *two[2] = {c0, c1};
  fft_r2c_t<T>(ctx, two, 2);
  return iters;
}

// Init_rho_ion, N_spec = 2, uniform salt (operate_electrolyte.cxx:559-621,
// 860); under Dielectric the bulk density carries eps_fluid (:597-600).
// no_surface: Set_uniform_ion_charge_density_salt_dielectric (:864-910), the
// bulk density kBT Dielectric_cst eps_fluid / (e lambda_D RADIUS)^2 and no
// counter-ions
template <typename T>
void init_rho_ion_t(kapsel_ctx *ctx, bool no_surface = false) {
  const kapsel_electrolyte &e = ctx->el;
  make_phi_particle(ctx, ctx->fields[KAPSEL_PHI]);
  reduce<T, 2>(ctx, nullptr, nullptr, ctx->fields[KAPSEL_PHI]);
  finish(ctx, 2);
  const double volume_phi = read_scal(ctx, S_VOL);
  double positive_ion_number = 0., negative_ion_number = 0.;
  const double q = e.surface_charge, npart = (double)ctx->P.n_particles;
  if (q < 0.) positive_ion_number += fabs(q * npart / e.valency[0]);
  else if (q > 0.) negative_ion_number += fabs(q * npart / e.valency[1]);
  const double dx3 = ctx->g.dx * ctx->g.dx * ctx->g.dx;
  const double pd = positive_ion_number / (volume_phi * dx3);
  const double nd = negative_ion_number / (volume_phi * dx3);
  double rho_inf = e.kbt * e.dielectric_cst / ((e.elementary_charge * e.debye_length) *
                                                (e.elementary_charge * e.debye_length));
  if (no_surface) {
    const double l = e.elementary_charge * e.debye_length * ctx->P.radius;
    rho_inf = e.kbt * e.dielectric_cst * ctx->P.eps_fluid / (l * l);
  } else if (e.dielectric) {
    rho_inf *= ctx->P.eps_fluid;
  }
  const double rp = rho_inf / (e.valency[0] * (e.valency[0] - e.valency[1]));
  const double rn = -rho_inf / (e.valency[1] * (e.valency[0] - e.valency[1]));
  {
    LaunchTimer lt(ctx, KAPSEL_K_POINTWISE, 2.0 * ctx->g.f * sizeof(T));
    hipLaunchKernelGGL(fill_kernel<T>, dim3(blocks_for(ctx->g.f, TPB)), dim3(TPB), 0, ctx->stream,
                       fl<T>(ctx, KAPSEL_CONC0), fl<T>(ctx, KAPSEL_CONC1), pd + rp, nd + rn, ctx->g);
    KAPSEL_LAUNCH_CHECK();
  }
  void *two[2] = {ctx->fields[KAPSEL_CONC0], ctx->fields[KAPSEL_CONC1]};
  fft_r2c_t<T>(ctx, two, 2);
  // Count_solute_each (solute_rhs.h:143-156): the totals are the amounts of
  // the transformed fields; PHI already holds Make_phi_particle
  ions_real_t<T>(ctx);
  reduce<T, 1>(ctx, ctx->el_f[E_CX0], ctx->el_f[E_CX1], ctx->fields[KAPSEL_PHI]);
  finish(ctx, 1, 0.0, 1.0, 1.0);
  ctx->el_total[0] = read_scal(ctx, S_C0);
  ctx->el_total[1] = read_scal(ctx, S_C1);
}

#define EL_DISPATCH(ctx, FN, ...) \
  ((ctx)->prec == KAPSEL_FP64 ? FN<double>(__VA_ARGS__) : FN<float>(__VA_ARGS__))

}  // namespace

void electrolyte_set(kapsel_ctx *ctx, const kapsel_electrolyte &el) {
  const bool surface = el.surface_charge != 0.0;
  if (surface && ctx->P.n_particles < 1) {
    set_error("the electrolyte needs particles: the surface charge is normalised by their total "
              "(Make_phi_qq_particle)");
    throw Error{KAPSEL_ERR_ARG};
  }
  if ((!surface && !el.dielectric) || !(el.valency[0] > 0.0) || !(el.valency[1] < 0.0) || !(el.kbt > 0.0) ||
      !(el.dielectric_cst > 0.0) || !(el.debye_length > 0.0) || !(el.elementary_charge > 0.0) ||
      !(el.onsager_coeff[0] >= 0.0) || !(el.onsager_coeff[1] >= 0.0)) {
    set_error("invalid electrolyte parameters (nonzero surface charge unless Dielectric, valencies +/-, "
              "positive kBT, Dielectric_cst, Debye length, elementary charge; Onsager coefficients >= 0)");
    throw Error{KAPSEL_ERR_ARG};
  }
  if (!ctx->el_f[0]) {
    for (int i = 0; i < E_NF; i++) {
      KAPSEL_HIP(hipMalloc(&ctx->el_f[i], ctx->g.f * ctx->elem));
      KAPSEL_HIP(hipMemsetAsync(ctx->el_f[i], 0, ctx->g.f * ctx->elem, ctx->stream));
    }
    KAPSEL_HIP(hipMalloc(&ctx->el_partial, sizeof(double) * 2 * NBR));
    KAPSEL_HIP(hipMalloc(&ctx->el_scal, sizeof(double) * S_N));
    KAPSEL_HIP(hipMemsetAsync(ctx->el_scal, 0, sizeof(double) * S_N, ctx->stream));
  }
  ctx->el = el;
  ctx->el_set = true;
}


void electrolyte_free(kapsel_ctx *ctx) {
  for (void *&p : ctx->el_f)
    if (p) {
      (void)hipFree(p);
      p = nullptr;
    }
  if (ctx->el_partial) (void)hipFree(ctx->el_partial);
  if (ctx->el_scal) (void)hipFree(ctx->el_scal);
  ctx->el_partial = ctx->el_scal = nullptr;
}

template <typename T>
static void recount_t(kapsel_ctx *ctx) {
  make_phi_particle(ctx, ctx->fields[KAPSEL_PHI]);
  ions_real_t<T>(ctx);
  reduce<T, 1>(ctx, ctx->el_f[E_CX0], ctx->el_f[E_CX1], ctx->fields[KAPSEL_PHI]);
  finish(ctx, 1, 0.0, 1.0, 1.0);
  ctx->el_total[0] = read_scal(ctx, S_C0);
  ctx->el_total[1] = read_scal(ctx, S_C1);
}
void recount_total_solute(kapsel_ctx *ctx) {
  need_el(ctx);
  EL_DISPATCH(ctx, recount_t, ctx);
}

void init_rho_ion(kapsel_ctx *ctx) {
  need_el(ctx);
  EL_DISPATCH(ctx, init_rho_ion_t, ctx);
}

template <typename T>
static int init_rho_ion_pb_t(kapsel_ctx *ctx, const kapsel_time &t, int max_iter, double *fe) {
  int it;
  if (ctx->P.n_particles == 1) {
    it = pb_salt_single_t<T>(ctx, max_iter > 0 ? max_iter : 100000);
    free_energy_host_t<T>(ctx, t, fe);
  } else {
    init_rho_ion_t<T>(ctx);
    it = steadystate_t<T>(ctx, t, max_iter > 0 ? max_iter : 10000, fe);
  }
  recount_t<T>(ctx);  // Count_solute_each (operate_electrolyte.cxx:860)
  return it;
}
int init_rho_ion_pb(kapsel_ctx *ctx, const kapsel_time &t, int max_iter, double *fe) {
  need_el_plain(ctx, "the Poisson-Boltzmann start");
  if (ctx->P.n_particles == 1 && ctx->el.valency[0] != -ctx->el.valency[1]) {
    // Set_Poisson_Boltzmann_ion_charge_density_salt (:629-633)
    set_error("the one-particle Poisson-Boltzmann start needs a z:z salt (valencies %g, %g)", ctx->el.valency[0],
              ctx->el.valency[1]);
    throw Error{KAPSEL_ERR_ARG};
  }
  return EL_DISPATCH(ctx, init_rho_ion_pb_t, ctx, t, max_iter, fe);
}

void free_energy_pb(kapsel_ctx *ctx, const kapsel_time &t, double *fe) {
  need_el_plain(ctx, "Calc_free_energy_PB");
  EL_DISPATCH(ctx, free_energy_host_t, ctx, t, fe);
}

template <typename T>
static void make_phi_qq_t(kapsel_ctx *ctx, void *phi, void *surface) {
  phi_qq_t<T>(ctx, (T *)phi, (T *)surface);
  LaunchTimer lt(ctx, KAPSEL_K_POINTWISE, 2.0 * ctx->g.nr * sizeof(T));
  hipLaunchKernelGGL(scale_real_kernel<T>, dim3(blocks_for(ctx->g.nr, TPB)), dim3(TPB), 0, ctx->stream,
                     (T *)surface, ctx->el_scal, (int)S_SURF, ctx->g);
  KAPSEL_LAUNCH_CHECK();
}
void make_phi_qq(kapsel_ctx *ctx, void *phi, void *surface) {
  need_el(ctx);
  EL_DISPATCH(ctx, make_phi_qq_t, ctx, phi, surface);
}

template <typename T>
static void charge_field_t(kapsel_ctx *ctx, void *out) {
  ions_real_t<T>(ctx);
  charge_t<T>(ctx, (T *)out, nullptr, 0);
}
void charge_field(kapsel_ctx *ctx, void *out) {
  need_el(ctx);
  EL_DISPATCH(ctx, charge_field_t, ctx, out);
}

void nssolute_euler(kapsel_ctx *ctx, const kapsel_time &t) {
  need_el(ctx);
  EL_DISPATCH(ctx, nssolute_t, ctx, t, false);
}

void nssolute_step(kapsel_ctx *ctx, const kapsel_time &t, bool reuse) {
  need_el(ctx);
  EL_DISPATCH(ctx, nssolute_t, ctx, t, reuse && !ctx->el.dielectric);
}

void *const *electrolyte_u_keep(kapsel_ctx *ctx) { return ctx->el_f + E_UX0; }

template <typename T>
static void rescale_solute_t(kapsel_ctx *ctx, double *factor_host) {
  rescale_t<T>(ctx, false);
  if (factor_host) {
    factor_host[0] = read_scal(ctx, S_F0);
    factor_host[1] = read_scal(ctx, S_F1);
  }
}
void rescale_solute(kapsel_ctx *ctx, double *factor_host) {
  need_el(ctx);
  EL_DISPATCH(ctx, rescale_solute_t, ctx, factor_host);
}

void coulomb_force(kapsel_ctx *ctx, const kapsel_time &t) {
  need_el_plain(ctx, "the Coulomb force");
  EL_DISPATCH(ctx, coulomb_force_t, ctx, t, 0);
}

template <typename T>
static void init_rho_ion_dielectric_t(kapsel_ctx *ctx) {
  init_rho_ion_t<T>(ctx, ctx->el.surface_charge == 0.0);
}
void init_rho_ion_dielectric(kapsel_ctx *ctx) {
  need_el_dielectric(ctx);
  EL_DISPATCH(ctx, init_rho_ion_dielectric_t, ctx);
}

// Make_Maxwell_force_x_on_fluid's free charge (operate_dielectric.cxx:
// 629-640): with surface charge Conc_k2charge_field, without it the ions
// weighted by the compact-sine profile in WORK4; then its mean
template <typename T>
static const double *dielectric_free_charge_t(kapsel_ctx *ctx) {
  const Grid &g = ctx->g;
  T *rho = fl<T>(ctx, KAPSEL_RHO_FREE);
  ions_real_t<T>(ctx);
  if (ctx->el.surface_charge != 0.0) {
    charge_t<T>(ctx, rho, nullptr, 0);
  } else {
    LaunchTimer lt(ctx, KAPSEL_K_POINTWISE, 4.0 * g.nr * sizeof(T));
    hipLaunchKernelGGL(charge_kernel<T>, dim3(blocks_for(g.nr, TPB)), dim3(TPB), 0, ctx->stream,
                       fl<T>(ctx, KAPSEL_WORK4), (const T *)nullptr, ef<T>(ctx, E_CX0), ef<T>(ctx, E_CX1),
                       ctx->el_scal, valency_e(ctx, 0), valency_e(ctx, 1), 0, rho, (T *)nullptr, g);
    KAPSEL_LAUNCH_CHECK();
  }
  reduce<T, 6>(ctx, rho, nullptr, nullptr);
  finish(ctx, 6, 1.0 / (double)g.nr);
  return ctx->el_scal + S_RHO_MEAN;
}
const double *dielectric_free_charge(kapsel_ctx *ctx) {
  need_el_dielectric(ctx);
  return EL_DISPATCH(ctx, dielectric_free_charge_t, ctx);
}

// the Electrolyte branch's middle part (sp_3d_ns.cxx:55-95): Rescale_solute
// right after the profile scatter (PHI_SUM current), then the Coulomb force
// from the rescaled ions
void rescale_and_coulomb_force(kapsel_ctx *ctx, const kapsel_time &t, bool spectrum) {
  need_el_plain(ctx, "the Coulomb force");
  if (ctx->prec == KAPSEL_FP64) {
    rescale_t<double>(ctx, true);
    coulomb_force_t<double>(ctx, t, 1, spectrum);
  } else {
    rescale_t<float>(ctx, true);
    coulomb_force_t<float>(ctx, t, 1, spectrum);
  }
}

}  // namespace kapsel
