// particles.hip -- particle <-> grid coupling and particle dynamics.
//
// Scatter (Make_phi_particle_sum, Make_u_particle_sum, Make_phi_particle,
// Make_phi_sin): the reference loops over particles and adds into the grid
// with `omp atomic` (make_phi.cxx:229-370).  Here each grid point sums the
// particles near it instead: a 4x4x16 tile of points first compacts, in
// particle order, the particles whose profile can reach the tile into LDS,
// then each thread adds their profile values for its own point.  No
// atomics, every field value is written once (no reset pass), and the sum
// order is fixed, so results are reproducible run to run.  The set of
// contributions is the reference's: a point of the Sekibun cell of particle
// i outside the profile support has phi = 0 exactly, and every point with
// phi > 0 lies in the cell (|o| <= r + |residue| < a + xi/2 + sqrt(3) dx).
//
// Gather (Calc_f_hydro_correct_precision, md_force.cxx:340-449): each
// particle's Sekibun table is cut into slices, one wavefront per (particle,
// slice) sums its slice with FP64 compensated (TwoSum) partial sums per lane
// and a wavefront reduction; a second kernel merges the slices in order.
#include <cfloat>
#include <mutex>
#include <type_traits>

#include "common.hpp"

namespace kapsel {

namespace {

struct Phys {
  double dx, idx, radius, hxi, xi;
  double eps_f, eps_p, dcst;  // eps_fluid, eps_particle_top, Dielectric_cst
  double qe;                  // Surface_charge_e of the particle species (Make_phi_qq_particle)
  double l[3];
  int n[3];
  int np;
  int rotation;
};

__device__ __forceinline__ int nint(double a) { return a >= 0.0 ? int(a + .5) : int(a - .5); }

// H and Phi (profile.h:29-49)
__device__ __forceinline__ double Hf(double x, double dx) {
  if (x > 0) {
    double q = dx / x;
    return exp(-(q * q));
  }
  return 0.0;
}
__device__ __forceinline__ double phi_smooth(double x, const Phys &p) {
  double dmy = Hf(p.radius + p.hxi - x, p.dx);
  return dmy / (dmy + Hf(x - p.radius + p.hxi, p.dx));
}
// Phi_compact_sin (profile.h:139-152)
__device__ __forceinline__ double phi_compact_sin(double x, const Phys &p) {
  double dmy_x = p.radius - x;
  if (fabs(dmy_x) <= p.hxi) return 0.5 * sin(M_PI * dmy_x / p.xi) + 0.5;
  return dmy_x > p.hxi ? 1. : 0.;
}

// Distance0 min image (interaction.h:122-135), returns squared norm
__device__ __forceinline__ double min_image(const double *x1, const double *x2, const Phys &p, double *x12) {
  double s = 0.0;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    double v = x2[d] - x1[d];
    v -= (double)nint(v / p.l[d]) * p.l[d];
    x12[d] = v;
    s += v * v;
  }
  return s;
}

__device__ __forceinline__ double pval(const double *ps, int f, int d, int n, int np) {
  return ps[((size_t)f * 3 + d) * np + n];
}
__device__ __forceinline__ double &pref(double *ps, int f, int d, int n, int np) {
  return ps[((size_t)f * 3 + d) * np + n];
}

constexpr int TP = 256;  // threads per scatter workgroup
constexpr int TXL = 8;   // points per thread along x

// profile values in the field precision M (float fields need no FP64 math)
template <typename M>
__device__ __forceinline__ M hm(M x, M dx) {
  if (x > (M)0) {
    M q = dx / x;
    return exp(-(q * q));
  }
  return (M)0;
}
template <typename M>
__device__ __forceinline__ M phi_m(M x, M a, M hxi, M dx) {
  M dmy = hm<M>(a + hxi - x, dx);
  return dmy / (dmy + hm<M>(x - a + hxi, dx));
}
template <typename M>
__device__ __forceinline__ M phi_sin_m(M x, M a, M hxi, M xi) {
  M dmy_x = a - x;
  if (fabs(dmy_x) <= hxi) return (M)0.5 * sin((M)M_PI * dmy_x / xi) + (M)0.5;
  return dmy_x > hxi ? (M)1 : (M)0;
}

// MODE 0: phi_sum and phi = min(phi_sum, 1) (Make_phi_particle_sum)
// MODE 1: unclamped smooth-profile sum (Make_phi_particle)
// MODE 2: unclamped compact-sine sum (Make_phi_sin)
// MODE 3: phi*u_p with phi_sum read back (Make_u_particle_sum)
// MODE 4: compact-sine sum and the permittivity built from it
//         (Make_dielectric_field non-Janus branch, operate_dielectric.cxx:298-307)
// MODE 5: unclamped smooth-profile sum and sum of Q_e DPhi_compact_sin, the
//         surface charge before its normalisation (Make_phi_qq_particle,
//         operate_electrolyte.cxx:179-224)
// MODE 6: sum of the unit vectors particle -> point over the band
//         |r - a| < xi/2 (Make_surface_normal, make_phi.cxx:17-61); computed
//         in FP64 for every field precision, since the band test is a
//         discontinuity a rounding difference could move
// A workgroup owns a tile of 8 x ty x tz points (ty * tz = 256); each
// thread walks 8 points along x.  The per-point arithmetic runs in the field
// precision T; displacements are formed in FP64 from the FP64 positions.
template <typename T, int MODE>
__global__ __launch_bounds__(TP) void points_kernel(const double *__restrict__ ps, T *__restrict__ out0,
                                                    T *__restrict__ out1, T *__restrict__ out2,
                                                    const T *__restrict__ phi_sum, Grid g, Phys p, int ty,
                                                    int tz) {
  using M = typename std::conditional<MODE == 6, double, T>::type;
  constexpr int NV = MODE == 3 ? 9 : 3;
  __shared__ double cand[NV][TP];
  __shared__ int wcount[TP / 64];
  const int t = threadIdx.x;
  const int ntx = g.nx / TXL, nty = g.ny / ty;
  int b = blockIdx.x;
  const int bx = b % ntx;
  b /= ntx;
  const int by = b % nty;
  const int bz = b / nty;
  const int j = by * ty + t / tz;
  const int k = bz * tz + t % tz;
  const int i0 = bx * TXL;
  const double il[3] = {1.0 / p.l[0], 1.0 / p.l[1], 1.0 / p.l[2]};
  // tile box centre and half widths for the conservative reach test
  const double cen[3] = {(i0 + 0.5 * (TXL - 1)) * p.dx, (by * ty + 0.5 * (ty - 1)) * p.dx,
                         (bz * tz + 0.5 * (tz - 1)) * p.dx};
  const double hw[3] = {0.5 * (TXL - 1) * p.dx, 0.5 * (ty - 1) * p.dx, 0.5 * (tz - 1) * p.dx};
  const double reach = p.radius + p.hxi + 0.5 * p.dx;
  const M a = (M)p.radius, hxi = (M)p.hxi, xi = (M)p.xi, dxm = (M)p.dx;
  const M rmax = a + hxi;
  M acc0[TXL], acc1[TXL], acc2[TXL];
#pragma unroll
  for (int x = 0; x < TXL; x++) acc0[x] = acc1[x] = acc2[x] = (M)0;
  const double yg = j * p.dx, zg = k * p.dx;
  for (int base = 0; base < p.np; base += TP) {
    const int n = base + t;
    bool c = false;
    double xp[3] = {0, 0, 0};
    if (n < p.np) {
      double s2 = 0.0;
#pragma unroll
      for (int d = 0; d < 3; d++) {
        xp[d] = pval(ps, KAPSEL_P_X, d, n, p.np);
        double dd = xp[d] - cen[d];
        dd -= p.l[d] * rint(dd * il[d]);
        dd = fabs(dd) - hw[d];
        if (dd > 0) s2 += dd * dd;
      }
      c = s2 < reach * reach;
    }
    const unsigned long long m = __ballot(c);
    const int wid = t / 64, lane = t % 64;
    if (lane == 0) wcount[wid] = __popcll(m);
    __syncthreads();
    int off = 0, total = 0;
#pragma unroll
    for (int w = 0; w < TP / 64; w++) {
      if (w < wid) off += wcount[w];
      total += wcount[w];
    }
    if (c) {
      const int slot = off + __popcll(m & ((1ull << lane) - 1ull));
#pragma unroll
      for (int d = 0; d < 3; d++) cand[d][slot] = xp[d];
      if constexpr (MODE == 3) {
#pragma unroll
        for (int d = 0; d < 3; d++) {
          cand[3 + d][slot] = pval(ps, KAPSEL_P_V, d, n, p.np);
          cand[6 + d][slot] = pval(ps, KAPSEL_P_OMEGA, d, n, p.np);
        }
      }
    }
    __syncthreads();
    for (int q = 0; q < total; q++) {
      // min-image displacement particle -> point (Distance0, interaction.h:122-135);
      // |x12| < a + xi/2 << L/2 wherever phi > 0, so Nint by multiplication is exact there
      double dy = yg - cand[1][q], dz = zg - cand[2][q];
      dy -= (double)nint(dy * il[1]) * p.l[1];
      dz -= (double)nint(dz * il[2]) * p.l[2];
      const M ry = (M)dy, rz = (M)dz;
      const M ryz2 = ry * ry + rz * rz;
      if (ryz2 >= rmax * rmax) continue;
#pragma unroll
      for (int x = 0; x < TXL; x++) {
        double dxv = (i0 + x) * p.dx - cand[0][q];
        dxv -= (double)nint(dxv * il[0]) * p.l[0];
        const M rx = (M)dxv;
        const M r2 = rx * rx + ryz2;
        if (r2 >= rmax * rmax) continue;  // H(a + xi/2 - r) = 0 there: phi = 0
        const M r = sqrt(r2);
        if constexpr (MODE == 2 || MODE == 4) {
          acc0[x] += phi_sin_m<M>(r, a, hxi, xi);
        } else if constexpr (MODE == 6) {
          if (fabs(r - a) < hxi) {
            const M ir = (M)1 / r;
            acc0[x] += rx * ir;
            acc1[x] += ry * ir;
            acc2[x] += rz * ir;
          }
        } else if constexpr (MODE == 5) {
          acc0[x] += phi_m<M>(r, a, hxi, dxm);
          // DPhi_compact_sin (profile.h:168-181)
          const M dmy_x = a - r;
          if (fabs(dmy_x) < hxi) {
            const M ixi = (M)1 / xi;
            acc1[x] += (M)p.qe * ((M)M_PI * (M)0.5 * ixi * cos((M)M_PI * dmy_x * ixi));
          }
        } else {
          const M ph = phi_m<M>(r, a, hxi, dxm);
          if constexpr (MODE == 3) {
            // r vector: particle centre -> point (Relative_coord, make_phi.h:162-179)
            M vr[3] = {0, 0, 0};
            if (p.rotation) {
              const M o0 = (M)cand[6][q], o1 = (M)cand[7][q], o2 = (M)cand[8][q];
              vr[0] = o1 * rz - o2 * ry;
              vr[1] = o2 * rx - o0 * rz;
              vr[2] = o0 * ry - o1 * rx;
            }
            const size_t im = ((size_t)(i0 + x) * g.ny + j) * g.nz_ + k;
            const M psum = (M)phi_sum[im];
            const M dmy_phi = ph / (psum > (M)1 ? psum : (M)1);
            acc0[x] += ((M)cand[3][q] + vr[0]) * dmy_phi;
            acc1[x] += ((M)cand[4][q] + vr[1]) * dmy_phi;
            acc2[x] += ((M)cand[5][q] + vr[2]) * dmy_phi;
          } else {
            acc0[x] += ph;
          }
        }
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int x = 0; x < TXL; x++) {
    const size_t im = ((size_t)(i0 + x) * g.ny + j) * g.nz_ + k;
    if constexpr (MODE == 0) {
      out0[im] = (T)acc0[x];                                    // phi_sum
      out1[im] = (T)(acc0[x] < (M)1 ? acc0[x] : (M)1);          // phi = MIN(phi_sum, 1)
    } else if constexpr (MODE == 3 || MODE == 6) {
      out0[im] = (T)acc0[x];
      out1[im] = (T)acc1[x];
      out2[im] = (T)acc2[x];
    } else if constexpr (MODE == 5) {
      out0[im] = (T)acc0[x];
      out1[im] = (T)acc1[x];
    } else if constexpr (MODE == 4) {
      out0[im] = (T)acc0[x];
      out1[im] = (T)((M)(p.eps_f * p.dcst) + (M)((p.eps_p - p.eps_f) * p.dcst) * acc0[x]);
    } else {
      out0[im] = (T)acc0[x];
    }
  }
}

// TwoSum compensated accumulation
struct CSum {
  double s, c;
};
__device__ __forceinline__ void cadd(CSum &a, double x) {
  double t = a.s + x;
  double bp = t - a.s;
  double e = (a.s - (t - bp)) + (x - bp);
  a.s = t;
  a.c += e;
}
__device__ __forceinline__ CSum cmerge(CSum a, CSum b) {
  CSum r = a;
  cadd(r, b.s);
  r.c += b.c;
  return r;
}

// Gather, first stage: unit u = (particle n, slice s) is one wavefront that
// sums the Sekibun points [s*chunk, min((s+1)*chunk, np_domain)) of particle
// n (lane-strided, TwoSum-compensated FP64), reduces across the wavefront and
// stores its 6 compensated sums (force, torque) as 12 doubles at
// partial[12*u].  Slicing the table puts nslices wavefronts on each particle,
// so a few dozen particles still fill the chip.
template <typename T>
__global__ __launch_bounds__(64) void hydro_gather_part_kernel(const double *__restrict__ ps,
                                                               const int *__restrict__ cells, int np_domain,
                                                               int nslices, int chunk, const T *__restrict__ u0,
                                                               const T *__restrict__ u1, const T *__restrict__ u2,
                                                               const T *__restrict__ phi_sum, Grid g, Phys p,
                                                               double *__restrict__ partial) {
  const int unit = blockIdx.x;
  const int n = unit / nslices, s = unit - n * nslices;
  const int t = threadIdx.x;
  double xp[3], vp[3], om[3];
  int x_int[3];
  double residue[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    xp[d] = pval(ps, KAPSEL_P_X, d, n, p.np);
    vp[d] = pval(ps, KAPSEL_P_V, d, n, p.np);
    om[d] = pval(ps, KAPSEL_P_OMEGA, d, n, p.np);
    // Particle_cell (make_phi.h:120-148); positions are validated to [0, L)
    // on upload and wrapped by md_position, the clamp only keeps indexing
    // in bounds
    double q = xp[d] * p.idx;
    if (!(q >= 0.0)) q = 0.0;
    if (q > (double)(p.n[d] - 1) + 0.999999999) q = (double)(p.n[d] - 1);
    x_int[d] = (int)q;
    residue[d] = (q - (double)x_int[d]) * p.dx;
  }
  CSum acc[6];
#pragma unroll
  for (int c = 0; c < 6; c++) acc[c] = CSum{0.0, 0.0};
  const int m0 = s * chunk;
  const int m1 = min(np_domain, m0 + chunk);
  for (int mesh = m0 + t; mesh < m1; mesh += 64) {
    int rm[3];
    double r[3], x[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
      const int cell = cells[3 * mesh + d];
      rm[d] = ((x_int[d] + cell) % p.n[d] + p.n[d]) % p.n[d];
      r[d] = (double)cell * p.dx - residue[d];
      x[d] = rm[d] * p.dx;
    }
    double x12[3];
    const double dmyR = sqrt(min_image(x, xp, p, x12));
    double vr[3] = {0, 0, 0};
    if (p.rotation) {
      vr[0] = om[1] * r[2] - om[2] * r[1];
      vr[1] = om[2] * r[0] - om[0] * r[2];
      vr[2] = om[0] * r[1] - om[1] * r[0];
    }
    const size_t im = ((size_t)rm[0] * g.ny + rm[1]) * g.nz_ + rm[2];
    const double psum = (double)phi_sum[im];
    const double dmy_phi = phi_smooth(dmyR, p) / (psum > 1.0 ? psum : 1.0);
    const double uu[3] = {(double)u0[im], (double)u1[im], (double)u2[im]};
    double fp[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
      fp[d] = ((vp[d] + vr[d]) - uu[d]) * dmy_phi;
      cadd(acc[d], fp[d]);
    }
    cadd(acc[3], r[1] * fp[2] - r[2] * fp[1]);
    cadd(acc[4], r[2] * fp[0] - r[0] * fp[2]);
    cadd(acc[5], r[0] * fp[1] - r[1] * fp[0]);
  }
#pragma unroll
  for (int c = 0; c < 6; c++) {
    for (int off = 32; off > 0; off >>= 1) {
      CSum o{__shfl_down(acc[c].s, off), __shfl_down(acc[c].c, off)};
      acc[c] = cmerge(acc[c], o);
    }
  }
  if (t == 0) {
    double *dst = partial + (size_t)12 * unit;
#pragma unroll
    for (int c = 0; c < 6; c++) {
      dst[2 * c] = acc[c].s;
      dst[2 * c + 1] = acc[c].c;
    }
  }
}

// Gather, second stage: one thread per (particle, component) merges the
// slices' compensated sums in slice order (a fixed order: the result does not
// depend on scheduling) and writes f_hydro / torque_hydro
// (md_force.cxx:440-447).
__global__ void hydro_gather_final_kernel(double *__restrict__ ps, const double *__restrict__ partial, int nslices,
                                          Phys p, double dmy) {
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= 6 * p.np) return;
  const int n = id / 6, c = id - 6 * n;
  const double *src = partial + (size_t)12 * n * nslices + 2 * c;
  CSum acc{src[0], src[1]};
  for (int s = 1; s < nslices; s++) acc = cmerge(acc, CSum{src[12 * s], src[12 * s + 1]});
  const double v = acc.s + acc.c;
  if (c < 3) pref(ps, KAPSEL_P_F_HYDRO, c, n, p.np) = dmy * v;
  else if (p.rotation) pref(ps, KAPSEL_P_TORQUE_HYDRO, c - 3, n, p.np) = dmy * v;
}

struct LJ {
  double cutoff, sigma, coeff;
  int powers;
  int lc[3];
  double lc_r;
  double body[3];
  int on;
};

// Lennard_Jones_f (interaction.h:175-200)
__device__ __forceinline__ double lj_f(double x, const LJ &lj) {
  double dmy = lj.sigma / x;
  double d2 = dmy * dmy;
  double d6 = d2 * d2 * d2;
  if (lj.powers == 0) dmy = d6;
  else if (lj.powers == 1) dmy = d6 * d6;
  else dmy = d6 * d6 * d6;
  return lj.coeff / (x * x) * (2.0 * dmy * dmy - dmy);
}

// number of times cell b appears among the three wrapped neighbours of
// cell a along one axis (md_force.cxx:68-81)
__device__ __forceinline__ int nb_count(int a, int b, int lc) {
  int cnt = 0;
#pragma unroll
  for (int o = -1; o <= 1; o++) {
    int v = a + o;
    if (v < 0) v = lc - 1;
    if (v > lc - 1) v = 0;
    cnt += (v == b);
  }
  return cnt;
}

// Force(): the cell-list LJ of md_force.cxx:16-171 (pairs i>j visited once
// per neighbour-cell match, each contributing +-f(r)(x_i - x_j)) plus the
// constant body force (Add_f_gravity, md_force.cxx:261-280).  One workgroup
// per particle i: its lanes take the partners j, a wavefront reduction sums
// them; every particle's force is written by one workgroup only.
__global__ __launch_bounds__(256) void force_kernel(double *__restrict__ ps, Phys p, LJ lj) {
  const int i = blockIdx.x;
  double fi[3] = {0, 0, 0};
  if (lj.on) {
    double xi[3];
    int ci[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
      xi[d] = pval(ps, KAPSEL_P_X, d, i, p.np);
      ci[d] = int(xi[d] / lj.lc_r);
    }
    for (int j = threadIdx.x; j < p.np; j += blockDim.x) {
      if (j == i) continue;
      double xj[3];
      int mult = 1;
#pragma unroll
      for (int d = 0; d < 3; d++) {
        xj[d] = pval(ps, KAPSEL_P_X, d, j, p.np);
        mult *= nb_count(ci[d], int(xj[d] / lj.lc_r), lj.lc[d]);
      }
      if (!mult) continue;
      double x12[3];
      const double r = sqrt(min_image(xi, xj, p, x12));  // x12 = x_j - x_i
      if (r < lj.cutoff) {
        const double f = lj_f(r, lj);
#pragma unroll
        for (int d = 0; d < 3; d++) fi[d] += mult * (f * (-x12[d]));
      }
    }
  }
  __shared__ double red[3][4];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    for (int off = 32; off > 0; off >>= 1) fi[d] += __shfl_down(fi[d], off);
    if ((threadIdx.x & 63) == 0) red[d][threadIdx.x >> 6] = fi[d];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int d = 0; d < 3; d++) {
      double s = 0.0;
      for (int w = 0; w < (int)(blockDim.x >> 6); w++) s += red[d][w];
      pref(ps, KAPSEL_P_FR, d, i, p.np) += s + lj.body[d];
    }
  }
}

// MD_solver_position_Euler/_AB2 (particle_solver.cxx:39-91), PBC (periodic_boundary.h:9-19)
// quaternion product alpha a.b (quaternion.h:173-183), (s, v) layout
__device__ __forceinline__ void qprod(double *q, const double *a, const double *b, double alpha) {
  q[0] = alpha * (a[0] * b[0] - (a[1] * b[1] + a[2] * b[2] + a[3] * b[3]));
  q[1] = alpha * (a[0] * b[1] + b[0] * a[1] + a[2] * b[3] - a[3] * b[2]);
  q[2] = alpha * (a[0] * b[2] + b[0] * a[2] + a[3] * b[1] - a[1] * b[3]);
  q[3] = alpha * (a[0] * b[3] + b[0] * a[3] + a[1] * b[2] - a[2] * b[1]);
}
// q.(0,v).q* (qtn_qvq_prod, quaternion.h:207-216)
__device__ __forceinline__ void qvq(double *out, const double *q, const double *v) {
  const double qs2 = q[0] * q[0] - (q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  const double qdotp = q[1] * v[0] + q[2] * v[1] + q[3] * v[2];
  out[0] = qs2 * v[0] + 2.0 * (qdotp * q[1] + q[0] * (q[2] * v[2] - q[3] * v[1]));
  out[1] = qs2 * v[1] + 2.0 * (qdotp * q[2] + q[0] * (q[3] * v[0] - q[1] * v[2]));
  out[2] = qs2 * v[2] + 2.0 * (qdotp * q[3] + q[0] * (q[1] * v[1] - q[2] * v[0]));
}

// MD_solver_position_Euler/_AB2 (particle_solver.cxx:39-91) and, with
// ROTATION, MD_solver_orientation_Euler (space-frame qdot) / _AB2 (body-frame
// qdot of 3 w_b - w_b,old) (particle_rotation_solver.h:15-39, rigid_body.cxx:
// 10-24, 151-165), each followed by qtn_normalize
__global__ void md_position_kernel(double *__restrict__ ps, double *__restrict__ qa, Phys p, int euler, double dt,
                                   double hdt) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= p.np) return;
  // every load before the first store (the state is one array, so a store
  // would hold back the loads after it)
  double x[3], v[3], vo[3], xn[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    x[d] = pval(ps, KAPSEL_P_X, d, n, p.np);
    v[d] = pval(ps, KAPSEL_P_V, d, n, p.np);
    vo[d] = euler ? 0.0 : pval(ps, KAPSEL_P_V_OLD, d, n, p.np);
    xn[d] = pval(ps, KAPSEL_P_X_NOPBC, d, n, p.np);
  }
  double q[4], qo[4], om[3], omo[3], dq[4];
  if (p.rotation) {
#pragma unroll
    for (int c = 0; c < 4; c++) {
      q[c] = qa[(size_t)c * p.np + n];
      qo[c] = qa[(size_t)(4 + c) * p.np + n];
    }
#pragma unroll
    for (int d = 0; d < 3; d++) {
      om[d] = pval(ps, KAPSEL_P_OMEGA, d, n, p.np);
      omo[d] = euler ? 0.0 : pval(ps, KAPSEL_P_OMEGA_OLD, d, n, p.np);
    }
  }
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const double dx = euler ? dt * v[d] : hdt * (3.0 * v[d] - vo[d]);
    pref(ps, KAPSEL_P_X_PREVIOUS, d, n, p.np) = x[d];
    pref(ps, KAPSEL_P_X_NOPBC, d, n, p.np) = xn[d] + dx;
    pref(ps, KAPSEL_P_X, d, n, p.np) = fmod(x[d] + dx + p.l[d], p.l[d]);
  }
  if (!p.rotation) return;
  if (euler) {
#pragma unroll
    for (int c = 0; c < 4; c++) qo[c] = q[c];
    const double qw[4] = {0.0, om[0], om[1], om[2]};
    qprod(dq, qw, q, 0.5);
#pragma unroll
    for (int c = 0; c < 4; c++) q[c] += dq[c] * dt;
  } else {
    double wb[3], wbo[3];
    const double qc[4] = {q[0], -q[1], -q[2], -q[3]}, qoc[4] = {qo[0], -qo[1], -qo[2], -qo[3]};
    qvq(wb, qc, om);
    qvq(wbo, qoc, omo);
#pragma unroll
    for (int d = 0; d < 3; d++) wb[d] = 3.0 * wb[d] - wbo[d];
#pragma unroll
    for (int c = 0; c < 4; c++) qo[c] = q[c];
    const double qw[4] = {0.0, wb[0], wb[1], wb[2]};
    qprod(dq, q, qw, 0.5);
#pragma unroll
    for (int c = 0; c < 4; c++) q[c] += dq[c] * hdt;
  }
  const double qnorm = 1.0 / sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
#pragma unroll
  for (int c = 0; c < 4; c++) {
    qa[(size_t)c * p.np + n] = q[c] * qnorm;
    qa[(size_t)(4 + c) * p.np + n] = qo[c];
  }
}

// Add_random_force_thermostat (fluct.cxx:22-35): the kicks of one step,
// deviates noise[k*np + n] (k < 6: velocity then angular velocity)
__global__ void thermostat_kernel(double *__restrict__ ps, const double *__restrict__ noise, Phys p, double nv,
                                  double imass, double no, double imoi) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= p.np) return;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    pref(ps, KAPSEL_P_V, d, n, p.np) += noise[(size_t)d * p.np + n] * nv * imass;
    if (p.rotation) pref(ps, KAPSEL_P_OMEGA, d, n, p.np) += noise[(size_t)(d + 3) * p.np + n] * no * imoi;
  }
}

// The velocity update of component d of particle n from the forces fh
// (hydrodynamic), fr (the others) and the hydrodynamic torque th, then
// reset_Forces: every load of the particle's state before the first store
// (the state is one array, so a store would hold back the loads after it).
__device__ __forceinline__ void velocity_update(double *__restrict__ ps, const Phys &p, int d, int n, int euler,
                                                double dmy, double dmy_rot, double fh, double fr, double th) {
  const double v = pval(ps, KAPSEL_P_V, d, n, p.np), om = pval(ps, KAPSEL_P_OMEGA, d, n, p.np);
  const double tr = pval(ps, KAPSEL_P_TORQUE_R, d, n, p.np);
  const double frp = euler ? 0.0 : pval(ps, KAPSEL_P_FR_PREVIOUS, d, n, p.np);
  const double trp = euler ? 0.0 : pval(ps, KAPSEL_P_TORQUE_R_PREVIOUS, d, n, p.np);
  const double vn = euler ? v + dmy * (fh + fr) : v + dmy * (2.0 * fh + fr + frp);
  const double omn = euler ? om + dmy_rot * (th + tr) : om + dmy_rot * (2.0 * th + tr + trp);
  pref(ps, KAPSEL_P_V_OLD, d, n, p.np) = v;
  pref(ps, KAPSEL_P_OMEGA_OLD, d, n, p.np) = om;
  pref(ps, KAPSEL_P_V, d, n, p.np) = vn;
  pref(ps, KAPSEL_P_OMEGA, d, n, p.np) = omn;
  pref(ps, KAPSEL_P_FR_PREVIOUS, d, n, p.np) = fr;
  pref(ps, KAPSEL_P_FR, d, n, p.np) = 0.0;
  pref(ps, KAPSEL_P_F_HYDRO_PREVIOUS, d, n, p.np) = fh;
  pref(ps, KAPSEL_P_F_HYDRO, d, n, p.np) = 0.0;
  pref(ps, KAPSEL_P_TORQUE_R_PREVIOUS, d, n, p.np) = tr;
  pref(ps, KAPSEL_P_TORQUE_R, d, n, p.np) = 0.0;
  pref(ps, KAPSEL_P_TORQUE_HYDRO_PREVIOUS, d, n, p.np) = th;
  pref(ps, KAPSEL_P_TORQUE_HYDRO, d, n, p.np) = 0.0;
}

// velocity update of MD_solver_velocity_Euler/_AB2_hydro (particle_solver.cxx:
// 109-150, 252-295) followed by reset_Forces (particle_solver.cxx:11-37)
__global__ void md_velocity_kernel(double *__restrict__ ps, Phys p, int euler, double dmy, double dmy_rot) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= p.np) return;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    velocity_update(ps, p, d, n, euler, dmy, dmy_rot, pval(ps, KAPSEL_P_F_HYDRO, d, n, p.np),
                    pval(ps, KAPSEL_P_FR, d, n, p.np), pval(ps, KAPSEL_P_TORQUE_HYDRO, d, n, p.np));
  }
}

// The end of Calc_f_hydro (the slices merged as hydro_gather_final_kernel
// merges them), Force() (as force_kernel) and the velocity update with
// reset_Forces (as md_velocity_kernel) of particle blockIdx.x in one
// workgroup: each reads and writes the particle's own state only (the LJ
// sum reads positions, which none of them writes), so the three launches
// of a step are one, with the same arithmetic.
__global__ __launch_bounds__(256) void hydro_force_velocity_kernel(double *__restrict__ ps,
                                                                   const double *__restrict__ partial, int nslices,
                                                                   Phys p, double dmy_f, LJ lj, int euler,
                                                                   double dmy, double dmy_rot) {
  const int i = blockIdx.x;
  __shared__ double fh[6], fr[3];
  __shared__ double red[3][4];
  // the slices' partials (12 per slice, contiguous) staged in LDS by the
  // whole workgroup, GS slices at a time, then merged in slice order by
  // threads 0-5 (the order of hydro_gather_final_kernel)
  constexpr int GS = 64;
  __shared__ double stage[12 * GS];
  const double *src = partial + (size_t)12 * i * nslices;
  CSum acc{0.0, 0.0};
  for (int s0 = 0; s0 < nslices; s0 += GS) {
    const int ns = min(GS, nslices - s0);
    for (int e = threadIdx.x; e < 12 * ns; e += blockDim.x) stage[e] = src[12 * s0 + e];
    __syncthreads();
    if (threadIdx.x < 6) {
      const int c = threadIdx.x;
      for (int sl = 0; sl < ns; sl++) {
        const CSum v{stage[12 * sl + 2 * c], stage[12 * sl + 2 * c + 1]};
        acc = s0 + sl == 0 ? v : cmerge(acc, v);
      }
    }
    __syncthreads();
  }
  if (threadIdx.x < 6) {
    const int c = threadIdx.x;
    const double v = acc.s + acc.c;
    fh[c] = c < 3 || p.rotation ? dmy_f * v : pval(ps, KAPSEL_P_TORQUE_HYDRO, c - 3, i, p.np);
  }
  double fi[3] = {0, 0, 0};
  if (lj.on) {
    double xi[3];
    int ci[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
      xi[d] = pval(ps, KAPSEL_P_X, d, i, p.np);
      ci[d] = int(xi[d] / lj.lc_r);
    }
    for (int j = threadIdx.x; j < p.np; j += blockDim.x) {
      if (j == i) continue;
      double xj[3];
      int mult = 1;
#pragma unroll
      for (int d = 0; d < 3; d++) {
        xj[d] = pval(ps, KAPSEL_P_X, d, j, p.np);
        mult *= nb_count(ci[d], int(xj[d] / lj.lc_r), lj.lc[d]);
      }
      if (!mult) continue;
      double x12[3];
      const double r = sqrt(min_image(xi, xj, p, x12));  // x12 = x_j - x_i
      if (r < lj.cutoff) {
        const double f = lj_f(r, lj);
#pragma unroll
        for (int d = 0; d < 3; d++) fi[d] += mult * (f * (-x12[d]));
      }
    }
  }
#pragma unroll
  for (int d = 0; d < 3; d++) {
    for (int off = 32; off > 0; off >>= 1) fi[d] += __shfl_down(fi[d], off);
    if ((threadIdx.x & 63) == 0) red[d][threadIdx.x >> 6] = fi[d];
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    const int d = threadIdx.x;
    double s = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) s += red[d][w];
    fr[d] = pval(ps, KAPSEL_P_FR, d, i, p.np) + (s + lj.body[d]);
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    const int d = threadIdx.x;
    velocity_update(ps, p, d, i, euler, dmy, dmy_rot, fh[d], fr[d], fh[3 + d]);
  }
}

Phys phys(kapsel_ctx *ctx) {
  Phys p;
  const Grid &g = ctx->g;
  p.dx = g.dx;
  p.idx = 1. / g.dx;
  p.radius = ctx->P.radius;
  p.xi = ctx->P.xi;
  p.hxi = ctx->hxi;
  p.l[0] = g.lx;
  p.l[1] = g.ly;
  p.l[2] = g.lz;
  p.n[0] = g.nx;
  p.n[1] = g.ny;
  p.n[2] = g.nz;
  p.np = ctx->P.n_particles;
  p.rotation = ctx->P.rotation;
  p.eps_f = ctx->P.eps_fluid;
  p.eps_p = ctx->P.eps_particle;
  p.dcst = ctx->P.dielectric_cst;
  p.qe = 0.0;
  return p;
}

template <typename T, int MODE>
void launch_points(kapsel_ctx *ctx, void *o0, void *o1, void *o2, const void *phi_sum, double qe = 0.0) {
  const Grid &g = ctx->g;
  const int tz = g.nz >= 32 ? 32 : 16, ty = TP / tz;
  unsigned nb = (unsigned)((g.nx / TXL) * (g.ny / ty) * (g.nz / tz));
  const double fields_moved = (MODE == 0 || MODE == 4 || MODE == 5) ? 2.0 : (MODE == 3 || MODE == 6) ? 4.0 : 1.0;
  LaunchTimer lt(ctx, KAPSEL_K_PARTICLE, fields_moved * g.nr * sizeof(T));
  Phys p = phys(ctx);
  p.qe = qe;
  hipLaunchKernelGGL((points_kernel<T, MODE>), dim3(nb), dim3(TP), 0, ctx->stream, ctx->d_ps, (T *)o0, (T *)o1,
                     (T *)o2, (const T *)phi_sum, g, p, ty, tz);
  KAPSEL_LAUNCH_CHECK();
}

// both stages of the gather on the given memory; partial holds 12 doubles per
// (particle, slice) unit
template <typename T>
void launch_gather(hipStream_t stream, double *ps, const int *cells, int np_domain, const T *u0, const T *u1,
                   const T *u2, const T *phi_sum, const Grid &g, const Phys &p, double dmy, double *partial) {
  const int ns = gather_slices(p.np, np_domain);
  const int chunk = (np_domain + ns - 1) / ns;
  hipLaunchKernelGGL(hydro_gather_part_kernel<T>, dim3((unsigned)(p.np * ns)), dim3(64), 0, stream, ps, cells,
                     np_domain, ns, chunk, u0, u1, u2, phi_sum, g, p, partial);
  KAPSEL_LAUNCH_CHECK();
  hipLaunchKernelGGL(hydro_gather_final_kernel, dim3(blocks_for((size_t)6 * p.np, 64)), dim3(64), 0, stream, ps,
                     partial, ns, p, dmy);
  KAPSEL_LAUNCH_CHECK();
}

template <typename T>
void gather_t(kapsel_ctx *ctx, double dmy) {
  LaunchTimer lt(ctx, KAPSEL_K_PARTICLE, 4.0 * ctx->P.n_particles * ctx->np_domain * sizeof(T));
  launch_gather<T>(ctx->stream, ctx->d_ps, ctx->d_sekibun, ctx->np_domain, (const T *)ctx->fields[KAPSEL_U0],
                   (const T *)ctx->fields[KAPSEL_U1], (const T *)ctx->fields[KAPSEL_U2],
                   (const T *)ctx->fields[KAPSEL_PHI_SUM], ctx->g, phys(ctx), dmy, ctx->d_gpart);
}

}  // namespace

void md_position(kapsel_ctx *ctx, const kapsel_time &t) {
  if (ctx->P.n_particles <= 0) return;
  LaunchTimer lt(ctx, KAPSEL_K_PARTICLE, 8.0 * KAPSEL_P_NFIELDS * 3 * ctx->P.n_particles);
  hipLaunchKernelGGL(md_position_kernel, dim3(blocks_for(ctx->P.n_particles, 64)), dim3(64), 0, ctx->stream,
                     ctx->d_ps, ctx->d_q, phys(ctx), t.ts == 0 ? 1 : 0, t.dt_md, t.hdt_md);
  KAPSEL_LAUNCH_CHECK();
}

void make_phi_particle_sum(kapsel_ctx *ctx) {
  if (ctx->prec == KAPSEL_FP64)
    launch_points<double, 0>(ctx, ctx->fields[KAPSEL_PHI_SUM], ctx->fields[KAPSEL_PHI], nullptr, nullptr);
  else
    launch_points<float, 0>(ctx, ctx->fields[KAPSEL_PHI_SUM], ctx->fields[KAPSEL_PHI], nullptr, nullptr);
}

void make_u_particle_sum(kapsel_ctx *ctx) {
  void **f = ctx->fields;
  if (ctx->prec == KAPSEL_FP64)
    launch_points<double, 3>(ctx, f[KAPSEL_UP0], f[KAPSEL_UP1], f[KAPSEL_UP2], f[KAPSEL_PHI_SUM]);
  else
    launch_points<float, 3>(ctx, f[KAPSEL_UP0], f[KAPSEL_UP1], f[KAPSEL_UP2], f[KAPSEL_PHI_SUM]);
}

void make_phi_particle(kapsel_ctx *ctx, void *field) {
  if (ctx->prec == KAPSEL_FP64) launch_points<double, 1>(ctx, field, nullptr, nullptr, nullptr);
  else launch_points<float, 1>(ctx, field, nullptr, nullptr, nullptr);
}

void make_phi_sin(kapsel_ctx *ctx, void *field) {
  if (ctx->prec == KAPSEL_FP64) launch_points<double, 2>(ctx, field, nullptr, nullptr, nullptr);
  else launch_points<float, 2>(ctx, field, nullptr, nullptr, nullptr);
}

void make_phi_qq_raw(kapsel_ctx *ctx, void *phi, void *surface, double q_e) {
  if (ctx->prec == KAPSEL_FP64) launch_points<double, 5>(ctx, phi, surface, nullptr, nullptr, q_e);
  else launch_points<float, 5>(ctx, phi, surface, nullptr, nullptr, q_e);
}

void make_surface_normal(kapsel_ctx *ctx, void *const *n3) {
  if (ctx->prec == KAPSEL_FP64) launch_points<double, 6>(ctx, n3[0], n3[1], n3[2], nullptr);
  else launch_points<float, 6>(ctx, n3[0], n3[1], n3[2], nullptr);
}

template <typename T>
void make_dielectric_field_t(kapsel_ctx *ctx, T *eps, T *phi_sin) {
  launch_points<T, 4>(ctx, phi_sin, eps, nullptr, nullptr);
}
template void make_dielectric_field_t<double>(kapsel_ctx *, double *, double *);
template void make_dielectric_field_t<float>(kapsel_ctx *, float *, float *);

void calc_f_hydro(kapsel_ctx *ctx, const kapsel_time &t) {
  if (ctx->P.n_particles <= 0) return;
  const double dmy = -(ctx->g.dx * ctx->g.dx * ctx->g.dx) * ctx->P.rho / t.dt_fluid;  // md_force.cxx:342-343
  if (ctx->prec == KAPSEL_FP64) gather_t<double>(ctx, dmy);
  else gather_t<float>(ctx, dmy);
}

namespace {
LJ lj_of(kapsel_ctx *ctx) {
  LJ lj;
  lj.on = ctx->P.lj_truncate >= 0;
  lj.cutoff = ctx->a_r_cutoff * ctx->P.lj_sigma;
  lj.sigma = ctx->P.lj_sigma;
  lj.powers = ctx->P.lj_powers;
  lj.coeff = (lj.powers == 0 ? 24. : lj.powers == 1 ? 48. : 72.) * ctx->P.lj_epsilon;
  const int cell_length = 16;  // md_force.cxx:14
  lj.lc[0] = ctx->g.nx / cell_length;
  lj.lc[1] = ctx->g.ny / cell_length;
  lj.lc[2] = ctx->g.nz / cell_length;
  lj.lc_r = cell_length * ctx->g.dx;
  for (int d = 0; d < 3; d++) lj.body[d] = ctx->P.body_force[d];
  return lj;
}
unsigned lj_threads(int np) { return np > 128 ? 256 : (np > 64 ? 128 : 64); }
}  // namespace

void force(kapsel_ctx *ctx) {
  if (ctx->P.n_particles <= 0) return;
  LaunchTimer lt(ctx, KAPSEL_K_PARTICLE, 8.0 * KAPSEL_P_NFIELDS * 3 * ctx->P.n_particles);
  hipLaunchKernelGGL(force_kernel, dim3(ctx->P.n_particles), dim3(lj_threads(ctx->P.n_particles)), 0, ctx->stream,
                     ctx->d_ps, phys(ctx), lj_of(ctx));
  KAPSEL_LAUNCH_CHECK();
}

void hydro_force_velocity(kapsel_ctx *ctx, const kapsel_time &t) {
  const int np = ctx->P.n_particles;
  if (np <= 0) return;
  const Grid &g = ctx->g;
  const Phys p = phys(ctx);
  const int ns = gather_slices(np, ctx->np_domain);
  const int chunk = (ctx->np_domain + ns - 1) / ns;
  const size_t elem = ctx->prec == KAPSEL_FP64 ? sizeof(double) : sizeof(float);
  {
    LaunchTimer lt(ctx, KAPSEL_K_PARTICLE, 4.0 * np * ctx->np_domain * elem);
    void *const *f = ctx->fields;
    if (ctx->prec == KAPSEL_FP64)
      hipLaunchKernelGGL(hydro_gather_part_kernel<double>, dim3((unsigned)(np * ns)), dim3(64), 0, ctx->stream,
                         ctx->d_ps, ctx->d_sekibun, ctx->np_domain, ns, chunk, (const double *)f[KAPSEL_U0],
                         (const double *)f[KAPSEL_U1], (const double *)f[KAPSEL_U2],
                         (const double *)f[KAPSEL_PHI_SUM], g, p, ctx->d_gpart);
    else
      hipLaunchKernelGGL(hydro_gather_part_kernel<float>, dim3((unsigned)(np * ns)), dim3(64), 0, ctx->stream,
                         ctx->d_ps, ctx->d_sekibun, ctx->np_domain, ns, chunk, (const float *)f[KAPSEL_U0],
                         (const float *)f[KAPSEL_U1], (const float *)f[KAPSEL_U2],
                         (const float *)f[KAPSEL_PHI_SUM], g, p, ctx->d_gpart);
    KAPSEL_LAUNCH_CHECK();
  }
  const double dmy_f = -(g.dx * g.dx * g.dx) * ctx->P.rho / t.dt_fluid;  // md_force.cxx:342-343
  const int euler = t.ts == 0;
  const double dtm = euler ? t.dt_md : t.hdt_md;
  LaunchTimer lt(ctx, KAPSEL_K_PARTICLE, 8.0 * KAPSEL_P_NFIELDS * 3 * np);
  hipLaunchKernelGGL(hydro_force_velocity_kernel, dim3(np), dim3(lj_threads(np)), 0, ctx->stream, ctx->d_ps,
                     (const double *)ctx->d_gpart, ns, p, dmy_f, lj_of(ctx), euler, dtm * ctx->imass,
                     dtm * ctx->imoi);
  KAPSEL_LAUNCH_CHECK();
}

namespace {
// genrand_real3 of mt19937ar.c: (int32 + 0.5) / 2^32, in (0, 1)
double real3(std::mt19937 &mt) { return ((double)mt() + 0.5) * (1.0 / 4294967296.0); }
// Gauss2 (fluct.h:52-63): polar method, two deviates
void gauss2(std::mt19937 &mt, double *r) {
  double x1, x2, rsq;
  do {
    x1 = 2. * real3(mt) - 1.;
    x2 = 2. * real3(mt) - 1.;
    rsq = x1 * x1 + x2 * x2;
  } while (rsq >= 1.0 || rsq == 0.0);
  const double factor = sqrt(-2. * log(rsq) / rsq);
  r[0] = factor * x1;
  r[1] = factor * x2;
}
}  // namespace

void thermostat(kapsel_ctx *ctx, const kapsel_time &t) {
  const int np = ctx->P.n_particles;
  if (np <= 0 || !(ctx->thermo_kbt > 0.0)) return;
  const size_t bytes = sizeof(double) * 6 * np;
  if (!ctx->d_noise) {
    KAPSEL_HIP(hipMalloc(&ctx->d_noise, bytes));
    for (int s = 0; s < 2; s++) {
      KAPSEL_HIP(hipHostMalloc(&ctx->h_noise[s], bytes, hipHostMallocDefault));
      KAPSEL_HIP(hipEventCreateWithFlags(&ctx->noise_ev[s], hipEventDisableTiming));
    }
  }
  // the copy that last read this host buffer has finished before it is refilled
  const int s = ctx->noise_slot;
  ctx->noise_slot ^= 1;
  KAPSEL_HIP(hipEventSynchronize(ctx->noise_ev[s]));
  double *h = ctx->h_noise[s];
  for (int n = 0; n < np; n++) {
    double dmy[6];
    gauss2(ctx->mt, dmy);
    gauss2(ctx->mt, dmy + 2);
    gauss2(ctx->mt, dmy + 4);
    for (int k = 0; k < 6; k++) h[(size_t)k * np + n] = dmy[k];
  }
  KAPSEL_HIP(hipMemcpyAsync(ctx->d_noise, h, bytes, hipMemcpyHostToDevice, ctx->stream));
  KAPSEL_HIP(hipEventRecord(ctx->noise_ev[s], ctx->stream));
  // Zeta_drag, Zeta_drag_rot, sdv_*, noise_intensity_* (fluct.cxx:10-15;
  // kT_snap_v/_o = 1.439/1.22, input.cxx:276-278)
  const double a = ctx->P.radius;
  const double zeta_drag = 6. * M_PI * ctx->P.eta * a;
  const double zeta_drag_rot = 8. * M_PI * ctx->P.eta * a * a * a;
  const double nv = 1.439 * sqrt(zeta_drag * t.dt_md * ctx->thermo_kbt * ctx->thermo_alpha_v);
  const double no = 1.22 * sqrt(zeta_drag_rot * t.dt_md * ctx->thermo_kbt * ctx->thermo_alpha_o);
  LaunchTimer lt(ctx, KAPSEL_K_PARTICLE, 8.0 * 12 * np);
  hipLaunchKernelGGL(thermostat_kernel, dim3(blocks_for(np, 64)), dim3(64), 0, ctx->stream, ctx->d_ps,
                     ctx->d_noise, phys(ctx), nv, ctx->imass, no, ctx->imoi);
  KAPSEL_LAUNCH_CHECK();
}

void thermostat_free(kapsel_ctx *ctx) {
  if (ctx->d_noise) (void)hipFree(ctx->d_noise);
  for (int s = 0; s < 2; s++) {
    if (ctx->h_noise[s]) (void)hipHostFree(ctx->h_noise[s]);
    if (ctx->noise_ev[s]) (void)hipEventDestroy(ctx->noise_ev[s]);
    ctx->h_noise[s] = nullptr;
    ctx->noise_ev[s] = nullptr;
  }
  ctx->d_noise = nullptr;
}

void md_velocity_integrate(kapsel_ctx *ctx, const kapsel_time &t) {
  if (ctx->P.n_particles <= 0) return;
  const int euler = t.ts == 0;
  const double dtm = euler ? t.dt_md : t.hdt_md;
  LaunchTimer lt(ctx, KAPSEL_K_PARTICLE, 8.0 * KAPSEL_P_NFIELDS * 3 * ctx->P.n_particles);
  hipLaunchKernelGGL(md_velocity_kernel, dim3(blocks_for(ctx->P.n_particles, 64)), dim3(64), 0, ctx->stream,
                     ctx->d_ps, phys(ctx), euler, dtm * ctx->imass, dtm * ctx->imoi);
  KAPSEL_LAUNCH_CHECK();
}

}  // namespace kapsel

// ---------------------------------------------------------------------------
// kapsel_calc_f_hydro_noalloc: the gather on caller-owned device memory with
// its scratch (the Sekibun table and the slices' partial sums) in a static
// device array, so that a call allocates nothing.
namespace kapsel {
namespace {

constexpr size_t kNoallocTableInts = (size_t)3 * KAPSEL_NOALLOC_MAX_CELL_POINTS;
constexpr size_t kNoallocTableDoubles = (kNoallocTableInts * sizeof(int) + sizeof(double) - 1) / sizeof(double);
constexpr size_t kNoallocDoubles = KAPSEL_NOALLOC_SCRATCH_BYTES / sizeof(double);
static_assert(kNoallocTableDoubles < kNoallocDoubles, "scratch too small for the Sekibun table");
constexpr size_t kNoallocUnits = (kNoallocDoubles - kNoallocTableDoubles) / 12;

// The array as each device holds it.  In host code `scratch` names a host-side
// stand-in, not device memory: host code only ever uses the per-device address
// that hipGetSymbolAddress returns (scratch_ptr below).
static __device__ __align__(256) double scratch[kNoallocDoubles];

// Particle_domain (profile.cxx:25-74) on the device, one workgroup: the cube
// of offsets in i-j-k order, 256 at a time; a kept offset goes to slot
// base + (kept offsets before it in this round), so the table has the
// reference order.  The test is sekibun_inside's, with each rounding explicit.
__global__ __launch_bounds__(256) void sekibun_build_kernel(int *__restrict__ cells, int nh, double dx,
                                                            double max_radius) {
  __shared__ int wave_count[4];
  const int side = 2 * nh + 1, total = side * side * side;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int base = 0;
  for (int c0 = 0; c0 < total; c0 += 256) {
    const int idx = c0 + (int)threadIdx.x;
    int i = 0, j = 0, k = 0;
    bool in = false;
    if (idx < total) {
      i = idx / (side * side) - nh;
      j = (idx / side) % side - nh;
      k = idx % side - nh;
      const double x0 = __dmul_rn((double)i, dx), x1 = __dmul_rn((double)j, dx), x2 = __dmul_rn((double)k, dx);
      const double r2 = __dadd_rn(__dadd_rn(__dmul_rn(x0, x0), __dmul_rn(x1, x1)), __dmul_rn(x2, x2));
      in = __dsqrt_rn(r2) < max_radius;
    }
    const unsigned long long m = __ballot(in);
    const int before = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) wave_count[w] = __popcll(m);
    __syncthreads();
    int off = base;
    for (int q = 0; q < w; q++) off += wave_count[q];
    if (in) {
      cells[3 * (off + before)] = i;
      cells[3 * (off + before) + 1] = j;
      cells[3 * (off + before) + 2] = k;
    }
    base += wave_count[0] + wave_count[1] + wave_count[2] + wave_count[3];
    __syncthreads();
  }
}

std::mutex g_scratch_mu;
void *g_scratch_ptr[KAPSEL_NOALLOC_MAX_DEVICES] = {};

// The device address of `scratch` on the current device, looked up once per
// device (with that device current) and kept.
double *scratch_ptr() {
  int dev = 0;
  KAPSEL_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= KAPSEL_NOALLOC_MAX_DEVICES) {
    set_error("device %d beyond the %d devices the allocation-free entry points serve", dev,
              KAPSEL_NOALLOC_MAX_DEVICES);
    throw Error{KAPSEL_ERR_ARG};
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *ptr = nullptr;
    KAPSEL_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) {
      set_error("hipGetSymbolAddress returned no address for the scratch array");
      throw Error{KAPSEL_ERR_HIP};
    }
    g_scratch_ptr[dev] = ptr;
  }
  return (double *)g_scratch_ptr[dev];
}

template <typename T>
void noalloc_gather(const kapsel_params &P, const Phys &p, const Grid &g, double dmy, double *ps, const void *u0,
                    const void *u1, const void *u2, const void *phi_sum, int np_domain, const SekibunGeom &sg,
                    hipStream_t stream, double *base) {
  int *cells = (int *)base;
  double *partial = base + kNoallocTableDoubles;
  hipLaunchKernelGGL(sekibun_build_kernel, dim3(1), dim3(256), 0, stream, cells, sg.nh, sg.dx, sg.max_radius);
  KAPSEL_LAUNCH_CHECK();
  launch_gather<T>(stream, ps, cells, np_domain, (const T *)u0, (const T *)u1, (const T *)u2, (const T *)phi_sum,
                   g, p, dmy, partial);
}

}  // namespace

}  // namespace kapsel

extern "C" int kapsel_calc_f_hydro_noalloc(const kapsel_params *params, const kapsel_time *jikan, double *ps,
                                           const void *u0, const void *u1, const void *u2, const void *phi_sum,
                                           void *stream) {
  using namespace kapsel;
  try {
    if (!params || !jikan || !ps || !u0 || !u1 || !u2 || !phi_sum) {
      set_error("kapsel_calc_f_hydro_noalloc: null argument");
      return KAPSEL_ERR_ARG;
    }
    const kapsel_params &P = *params;
    if (!(jikan->dt_fluid > 0) || !(P.dx > 0) || !(P.rho > 0) || P.n_particles < 0 || !(P.radius > 0) ||
        !(P.xi > 0) || (P.precision != KAPSEL_FP64 && P.precision != KAPSEL_FP32)) {
      set_error("kapsel_calc_f_hydro_noalloc: invalid parameters");
      return KAPSEL_ERR_ARG;
    }
    for (int n : {P.nx, P.ny, P.nz})
      if (n < 1) {
        set_error("kapsel_calc_f_hydro_noalloc: grid size %d", n);
        return KAPSEL_ERR_ARG;
      }
    if (P.n_particles == 0) return KAPSEL_OK;
    const SekibunGeom sg = sekibun_geom(P.radius, P.xi, P.dx);
    if (2 * sg.nh + 1 > P.nx || 2 * sg.nh + 1 > P.ny || 2 * sg.nh + 1 > P.nz) {
      set_error("kapsel_calc_f_hydro_noalloc: Sekibun cell (%d points across) wider than the grid", 2 * sg.nh + 1);
      return KAPSEL_ERR_ARG;
    }
    int np_domain = 0;
    for (int i = -sg.nh; i <= sg.nh; i++)
      for (int j = -sg.nh; j <= sg.nh; j++)
        for (int k = -sg.nh; k <= sg.nh; k++) np_domain += sekibun_inside(sg, i, j, k) ? 1 : 0;
    if (np_domain > KAPSEL_NOALLOC_MAX_CELL_POINTS) {
      set_error("kapsel_calc_f_hydro_noalloc: Sekibun cell of %d points exceeds the scratch's %d", np_domain,
                KAPSEL_NOALLOC_MAX_CELL_POINTS);
      return KAPSEL_ERR_ARG;
    }
    const size_t units = (size_t)P.n_particles * gather_slices(P.n_particles, np_domain);
    if (units > kNoallocUnits) {
      set_error("kapsel_calc_f_hydro_noalloc: %d particles need %zu partial-sum slots, the scratch holds %zu",
                P.n_particles, units, kNoallocUnits);
      return KAPSEL_ERR_ARG;
    }
    Grid g{};
    g.nx = P.nx;
    g.ny = P.ny;
    g.nz = P.nz;
    g.h = g.nz / 2 + 1;
    g.nz_ = 2 * g.h;
    g.dx = P.dx;
    Phys p{};
    p.dx = P.dx;
    p.idx = 1. / P.dx;
    p.radius = P.radius;
    p.xi = P.xi;
    p.hxi = 0.5 * P.xi;
    p.l[0] = P.nx * P.dx;
    p.l[1] = P.ny * P.dx;
    p.l[2] = P.nz * P.dx;
    p.n[0] = P.nx;
    p.n[1] = P.ny;
    p.n[2] = P.nz;
    p.np = P.n_particles;
    p.rotation = P.rotation;
    const double dmy = -(P.dx * P.dx * P.dx) * P.rho / jikan->dt_fluid;  // md_force.cxx:342-343
    double *base = scratch_ptr();
    if (P.precision == KAPSEL_FP64)
      noalloc_gather<double>(P, p, g, dmy, ps, u0, u1, u2, phi_sum, np_domain, sg, (hipStream_t)stream, base);
    else
      noalloc_gather<float>(P, p, g, dmy, ps, u0, u1, u2, phi_sum, np_domain, sg, (hipStream_t)stream, base);
    return KAPSEL_OK;
  } catch (const Error &e) {
    return e.code;
  } catch (const std::exception &e) {
    set_error("%s", e.what());
    return KAPSEL_ERR_STATE;
  }
}
