/*
 * kapsel_oracle.cpp -- CPU FP64 restatement of the KAPSEL SPM hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity checker for the HIP
 * library (tests/, __graft_entry__.smoke()) and the CPU baseline that
 * bench.py times beside it.  Nothing in the product links or calls it.
 *
 * It restates the reference loops one for one, in the reference's operation
 * order, FP64 throughout, OpenMP over the same loop nests (scatter with
 * `omp atomic` like make_phi.cxx:258).  The only part not taken from the
 * reference is the FFT: the reference calls FFTW/Ooura, this file carries
 * its own radix-2 transform with the same convention (forward e^{-ik.x}
 * unscaled, inverse scaled by 1/(NX NY NZ), fft_wrapper.h:681-723), checked
 * in tests against the reference's own Ooura rdft3d built from source, and
 * against numpy.fft.
 *
 * Pinning of each function (tests/test_oracle_*.py):
 *   fft_r2c/fft_c2r       Ooura rdft3d+rdft3dsort (oracle/_ref), numpy.fft, DFT closed forms
 *   truncate              index-set check against |k| >= TRN (closed form)
 *   zeta_k2u_k/u_k2zeta_k round trip on solenoidal truncated fields, curl/div identities
 *   u2advection_k         independent curl(u x omega) in numpy; Beltrami (ABC) flow gives 0
 *   ns_slaved_euler       shear wave decays by exp(-nu k^2 dt) exactly
 *   solenoidal_u          k.u_k = 0, idempotence, gradient fields removed
 *   phi / sekibun         closed-form profile values, lattice-point count, brute force O(Np N^3)
 *   scatter / gather      brute force numpy sums; rigid-motion fluid gives zero force/torque
 *   LJ                    all-pairs numpy sum (the reference's own brute-force form,
 *                         md_force.cxx:173-259)
 *   particle update       closed-form Euler/AB2 increments
 *   dielectric operator   constant eps: symbol -eps|k|^2; symmetry; residual of the solve
 *   init_rigidity, f_particle  closed-form centre velocity of one sphere on a grid node
 *   spm/icep/electrolyte step  exact per-step momentum budget (tests/test_oracle_step.py);
 *                         `plant` switches in known mistakes that the checks must catch
 */
#include <algorithm>
#include <cmath>
#include <random>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/kapsel_gpu.h"

typedef std::complex<double> cplx;

namespace {

inline double SQ(double x) { return x * x; }

/* Nint (macro.h:43-52) */
inline int Nint(double a) { return a >= 0.0 ? int(a + .5) : int(a - .5); }

/* ------------------------------------------------------------------------ */
/* 1-D radix-2 complex FFT on a contiguous line (own code, not the reference) */
struct Fft1d {
  int n = 0;
  std::vector<cplx> tw;  /* exp(-2 pi i k / n), k < n/2 */
  std::vector<int> rev;
  void init(int n_) {
    n = n_;
    tw.resize(std::max(1, n / 2));
    for (int k = 0; k < n / 2; k++) {
      long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)k / (long double)n;
      tw[k] = cplx((double)cosl(a), (double)sinl(a));
    }
    rev.resize(n);
    int lg = 0;
    while ((1 << lg) < n) lg++;
    for (int i = 0; i < n; i++) {
      int r = 0;
      for (int b = 0; b < lg; b++)
        if (i & (1 << b)) r |= 1 << (lg - 1 - b);
      rev[i] = r;
    }
  }
  /* sign = -1: forward e^{-i}; sign = +1: inverse e^{+i}, unscaled */
  void run(cplx *a, int sign) const {
    for (int i = 0; i < n; i++)
      if (i < rev[i]) std::swap(a[i], a[rev[i]]);
    for (int len = 2; len <= n; len <<= 1) {
      int half = len >> 1, step = n / len;
      for (int s = 0; s < n; s += len) {
        for (int k = 0; k < half; k++) {
          cplx w = tw[k * step];
          if (sign > 0) w = std::conj(w);
          cplx u = a[s + k], v = a[s + k + half] * w;
          a[s + k] = u + v;
          a[s + k + half] = u - v;
        }
      }
    }
  }
};

struct Oracle {
  kapsel_params P;
  int NX, NY, NZ, NZ_, HNZ_, HNX, HNY;
  int TRN_X, TRN_Y, TRN_Z;
  double DX, DX3, WAVE_X, WAVE_Y, WAVE_Z, NU, IRHO, HXI, RADIUS, XI;
  double L[3];
  double MASS, IMASS, MOI, IMOI;
  double A_R_cutoff;
  size_t F;
  Fft1d fx, fy, fzh;        /* lengths NX, NY, NZ/2 */
  std::vector<cplx> zrot;   /* exp(-2 pi i k / NZ), k <= NZ/2 */
  std::vector<double *> fields;
  double uk_dc[3];
  int Np;
  std::vector<double> ps;   /* particle state, SoA (kapsel_gpu.h) */
  std::vector<double> qs;   /* orientations q (c < 4) and q_old (c = 4..7) at [c*Np + n] */
  int NP_domain;
  std::vector<int> sekibun; /* 3 ints per cell point */
  /* Planted mistake (0 = none), for tests that show a check fails on it:
   * 1 Add_f_particle subtracts f; 2 the hydrodynamic gather runs after the
   * velocity update; 3 f is built from phi_sum instead of phi; 4 the
   * position update runs after the profile scatter; 5 the dielectric or
   * Coulomb force is subtracted from u. */
  int plant = 0;

  double *fld(int f) { return fields[f]; }
  double *pf(int f, int d) { return ps.data() + ((size_t)f * 3 + d) * Np; }

  /* k maps, Calc_K*_Ooura (fft_wrapper.h:761-773); k index is the double index */
  inline int KX(int i) const { return i > HNX ? i - NX : i; }
  inline int KY(int j) const { return j > HNY ? j - NY : j; }
  inline int KZ(int k) const { return k / 2; }
  inline double IK2(int i, int j, int k) const {
    double k2 = SQ(WAVE_X * KX(i)) + SQ(WAVE_Y * KY(j)) + SQ(WAVE_Z * KZ(k));
    return k2 > 0.0 ? 1. / k2 : 0.0;
  }
  inline double K2(int i, int j, int k) const {
    return SQ(WAVE_X * KX(i)) + SQ(WAVE_Y * KY(j)) + SQ(WAVE_Z * KZ(k));
  }

  ~Oracle() {
    for (double *p : fields) free(p);
  }

  /* ---------------------------------------------------------------- setup */
  int init(const kapsel_params *p) {
    P = *p;
    NX = P.nx; NY = P.ny; NZ = P.nz;
    HNZ_ = NZ / 2 + 1; NZ_ = 2 * HNZ_;   /* input.cxx:305-306 */
    HNX = NX / 2; HNY = NY / 2;
    TRN_X = (NX + 2) / 3; TRN_Y = (NY + 2) / 3; TRN_Z = (NZ + 2) / 3; /* input.cxx:310 */
    DX = P.dx; DX3 = DX * DX * DX;
    L[0] = NX * DX; L[1] = NY * DX; L[2] = NZ * DX;
    WAVE_X = 2 * M_PI / L[0]; WAVE_Y = 2 * M_PI / L[1]; WAVE_Z = 2 * M_PI / L[2];
    IRHO = 1. / P.rho; NU = P.eta * IRHO;
    RADIUS = P.radius; XI = P.xi; HXI = XI * 0.5;
    /* Set_global_parameters (input.cxx:287-296) */
    MASS = P.mass_ratio * P.rho * 4. / 3. * M_PI * RADIUS * RADIUS * RADIUS;
    IMASS = 1. / MASS;
    MOI = 2. / 5. * MASS * SQ(RADIUS);
    IMOI = 1. / MOI;
    /* A_R_cutoff (input.cxx:1530-1547) */
    if (P.lj_truncate > 0) {
      A_R_cutoff = P.lj_powers == 0 ? pow(2., 1. / 6.) : P.lj_powers == 1 ? pow(2., 1. / 12.) : pow(2., 1. / 18.);
    } else if (P.lj_truncate == 0) {
      int Nmin = std::min(NX, std::min(NY, NZ));
      A_R_cutoff = std::min(Nmin * DX * .5 / P.lj_sigma, 2.5);
    } else {
      A_R_cutoff = 0.;
    }
    F = (size_t)NX * NY * NZ_;
    fx.init(NX); fy.init(NY); fzh.init(NZ / 2);
    zrot.resize(NZ / 2 + 1);
    for (int k = 0; k <= NZ / 2; k++) {
      long double a = -2.0L * 3.14159265358979323846264338327950288L * k / (long double)NZ;
      zrot[k] = cplx((double)cosl(a), (double)sinl(a));
    }
    fields.resize(KAPSEL_NFIELDS);
    for (int f = 0; f < KAPSEL_NFIELDS; f++) {
      fields[f] = (double *)aligned_alloc(128, ((F * sizeof(double) + 127) / 128) * 128);
      memset(fields[f], 0, F * sizeof(double));
    }
    uk_dc[0] = uk_dc[1] = uk_dc[2] = 0;
    Np = P.n_particles;
    ps.assign((size_t)KAPSEL_P_NFIELDS * 3 * std::max(Np, 1), 0.0);
    qs.assign((size_t)8 * std::max(Np, 1), 0.0);
    for (int n = 0; n < Np; n++) qs[n] = qs[(size_t)4 * Np + n] = 1.0;
    particle_domain();
    return 0;
  }

  /* Particle_domain (profile.cxx:25-74) */
  void particle_domain() {
    /* A, A_XI are the UDF values with RADIUS = A*DX, XI = A_XI*DX (input.cxx:1396-1403) */
    int A = (int)(RADIUS / DX + 1e-9), A_XI = (int)(XI / DX + 1e-9);
    int nh = A + A_XI + 2;
    double MAX_RADIUS = RADIUS + HXI + 1.7321 * DX;
    sekibun.clear();
    for (int i = -nh; i <= nh; i++)
      for (int j = -nh; j <= nh; j++)
        for (int k = -nh; k <= nh; k++) {
          double x0 = (double)i * DX, x1 = (double)j * DX, x2 = (double)k * DX;
          double dmy = sqrt(SQ(x0) + SQ(x1) + SQ(x2));
          if (dmy < MAX_RADIUS) {
            sekibun.push_back(i); sekibun.push_back(j); sekibun.push_back(k);
          }
        }
    NP_domain = (int)sekibun.size() / 3;
  }

  /* ------------------------------------------------------------------ FFT */
  /* A2a_k: in-place r2c, forward e^{-i}, unscaled (fft_wrapper.h:681-701) */
  void fft_r2c(double *a) {
    const int M = NZ / 2;
    /* z: real rows -> NZ/2+1 complex, half-length complex FFT + split */
#pragma omp parallel
    {
      std::vector<cplx> z(M), X(M + 1);
#pragma omp for
      for (int ij = 0; ij < NX * NY; ij++) {
        double *row = a + (size_t)ij * NZ_;
        for (int n = 0; n < M; n++) z[n] = cplx(row[2 * n], row[2 * n + 1]);
        fzh.run(z.data(), -1);
        /* X[k] = E[k] + W^k O[k]; E = (Z[k]+conj Z[M-k])/2, O = (Z[k]-conj Z[M-k])/(2i) */
        for (int k = 0; k <= M; k++) {
          cplx zk = z[k % M], zm = std::conj(z[(M - k) % M]);
          cplx E = 0.5 * (zk + zm);
          cplx O = cplx(0, -0.5) * (zk - zm);
          X[k] = E + zrot[k] * O;
        }
        for (int k = 0; k <= M; k++) { row[2 * k] = X[k].real(); row[2 * k + 1] = X[k].imag(); }
      }
    }
    fft_yx(a, -1);
  }

  /* complex passes along y then x (forward) or x then y (inverse) */
  void fft_yx(double *a, int sign) {
    const int H = HNZ_;
    cplx *c = reinterpret_cast<cplx *>(a);
    auto ypass = [&]() {
#pragma omp parallel
      {
        std::vector<cplx> line(NY);
#pragma omp for
        for (int ik = 0; ik < NX * H; ik++) {
          int i = ik / H, kz = ik % H;
          cplx *base = c + (size_t)i * NY * H + kz;
          for (int j = 0; j < NY; j++) line[j] = base[(size_t)j * H];
          fy.run(line.data(), sign);
          for (int j = 0; j < NY; j++) base[(size_t)j * H] = line[j];
        }
      }
    };
    auto xpass = [&]() {
#pragma omp parallel
      {
        std::vector<cplx> line(NX);
#pragma omp for
        for (int jk = 0; jk < NY * H; jk++) {
          cplx *base = c + jk;
          for (int i = 0; i < NX; i++) line[i] = base[(size_t)i * NY * H];
          fx.run(line.data(), sign);
          for (int i = 0; i < NX; i++) base[(size_t)i * NY * H] = line[i];
        }
      }
    };
    if (sign < 0) { ypass(); xpass(); }
    else { xpass(); ypass(); }
  }

  /* A_k2a: in-place c2r, scaled by 1/(NX NY NZ) (fft_wrapper.h:704-723).
   * The imaginary parts of the kz = 0 and kz = NZ/2 entries are ignored
   * after the x and y passes (FFTW's c2r treats them as real). */
  void fft_c2r(double *a) {
    fft_yx(a, +1);
    const int M = NZ / 2;
    const double scale = 1.0 / ((double)NX * NY * NZ);
#pragma omp parallel
    {
      std::vector<cplx> z(M), X(M + 1);
#pragma omp for
      for (int ij = 0; ij < NX * NY; ij++) {
        double *row = a + (size_t)ij * NZ_;
        for (int k = 0; k <= M; k++) X[k] = cplx(row[2 * k], row[2 * k + 1]);
        X[0] = cplx(X[0].real(), 0.0);
        X[M] = cplx(X[M].real(), 0.0);
        for (int k = 0; k < M; k++) {
          cplx xk = X[k], xm = std::conj(X[M - k]);
          cplx E = xk + xm;                          /* 2 E[k] */
          cplx O = (xk - xm) * std::conj(zrot[k]);   /* 2 O[k] */
          z[k] = E + cplx(0, 1) * O;
        }
        fzh.run(z.data(), +1);
        for (int n = 0; n < M; n++) {
          row[2 * n] = z[n].real() * scale;
          row[2 * n + 1] = z[n].imag() * scale;
        }
        row[NZ] = 0.0; row[NZ + 1] = 0.0;
      }
    }
  }

  /* ------------------------------------------------------------ k-space */
  /* Truncate_two_third_rule_ooura (fft_wrapper.h:797-839): zero every mode
   * with |kx| >= TRN_X or |ky| >= TRN_Y or kz >= TRN_Z */
  void truncate(double *a) {
    const int trn_z2 = 2 * TRN_Z;
#pragma omp parallel for
    for (int i = 0; i < NX; i++)
      for (int j = 0; j < NY; j++)
        for (int k = 0; k < NZ_; k++) {
          bool cut = (k >= trn_z2) || (j >= TRN_Y && j <= NY - TRN_Y) ||
                     (i >= TRN_X && i <= NX - TRN_X);
          if (cut) a[((size_t)i * NY + j) * NZ_ + k] = 0.0;
        }
  }

  /* Zeta_k2u_k (fft_wrapper.cxx:614-711) */
  void zeta_k2u_k(double **zeta, const double dc[3], double **u) {
#pragma omp parallel for
    for (int i = 0; i < NX; i++)
      for (int j = 0; j < NY; j++)
        for (int k = 0; k < HNZ_; k++) {
          int k2 = 2 * k;
          size_t im = ((size_t)i * NY + j) * NZ_ + k2;
          double kx = WAVE_X * KX(i), ky = WAVE_Y * KY(j), kz = WAVE_Z * KZ(k2);
          double ik2 = IK2(i, j, k2);
          double d1r = zeta[0][im], d1i = zeta[0][im + 1];
          double d2r = zeta[1][im], d2i = zeta[1][im + 1];
          double orr[3], oi[3];
          if (KX(i) != 0) {
            orr[0] = -(1. / kx) * (ky * d1r + kz * d2r);
            oi[0] = -(1. / kx) * (ky * d1i + kz * d2i);
            orr[1] = d1r; oi[1] = d1i;
            orr[2] = d2r; oi[2] = d2i;
          } else if (KY(j) != 0) {
            double dmy = -(kz / ky);
            orr[0] = d2r; oi[0] = d2i;
            orr[1] = dmy * d1r; oi[1] = dmy * d1i;
            orr[2] = d1r; oi[2] = d1i;
          } else {
            orr[0] = d1r; oi[0] = d1i;
            orr[1] = d2r; oi[1] = d2i;
            orr[2] = 0.0; oi[2] = 0.0;
          }
          kx *= ik2; ky *= ik2; kz *= ik2;
          u[0][im] = -(ky * oi[2] - kz * oi[1]);
          u[0][im + 1] = (ky * orr[2] - kz * orr[1]);
          u[1][im] = -(kz * oi[0] - kx * oi[2]);
          u[1][im + 1] = (kz * orr[0] - kx * orr[2]);
          u[2][im] = -(kx * oi[1] - ky * oi[0]);
          u[2][im + 1] = (kx * orr[1] - ky * orr[0]);
        }
    u[0][0] = dc[0]; u[1][0] = dc[1]; u[2][0] = dc[2];
  }

  /* U_k2zeta_k (fft_wrapper.cxx:428-497) */
  void u_k2zeta_k(double **u, double **zeta, double dc[3]) {
    dc[0] = u[0][0]; dc[1] = u[1][0]; dc[2] = u[2][0];
#pragma omp parallel for
    for (int i = 0; i < NX; i++)
      for (int j = 0; j < NY; j++)
        for (int k = 0; k < HNZ_; k++) {
          int k2 = 2 * k;
          size_t im = ((size_t)i * NY + j) * NZ_ + k2;
          double ks0 = KX(i) * WAVE_X, ks1 = KY(j) * WAVE_Y, ks2 = KZ(k2) * WAVE_Z;
          double orr[3], oi[3];
          orr[0] = -(ks1 * u[2][im + 1] - ks2 * u[1][im + 1]);
          oi[0] = (ks1 * u[2][im] - ks2 * u[1][im]);
          orr[1] = -(ks2 * u[0][im + 1] - ks0 * u[2][im + 1]);
          oi[1] = (ks2 * u[0][im] - ks0 * u[2][im]);
          orr[2] = -(ks0 * u[1][im + 1] - ks1 * u[0][im + 1]);
          oi[2] = (ks0 * u[1][im] - ks1 * u[0][im]);
          int a0, a1;
          if (KX(i) != 0) { a0 = 1; a1 = 2; }
          else if (KY(j) != 0) { a0 = 2; a1 = 0; }
          else { a0 = 0; a1 = 1; }
          zeta[0][im] = orr[a0]; zeta[0][im + 1] = oi[a0];
          zeta[1][im] = orr[a1]; zeta[1][im + 1] = oi[a1];
        }
  }

  /* Omega_k2zeta_k (fft_wrapper.cxx:248-284) */
  void omega_k2zeta_k(double **omega, double **zetak) {
#pragma omp parallel for
    for (int i = 0; i < NX; i++)
      for (int j = 0; j < NY; j++)
        for (int k = 0; k < NZ_; k++) {
          size_t im = ((size_t)i * NY + j) * NZ_ + k;
          if (KX(i) != 0) { zetak[0][im] = omega[1][im]; zetak[1][im] = omega[2][im]; }
          else if (KY(j) != 0) { zetak[0][im] = omega[2][im]; zetak[1][im] = omega[0][im]; }
          else { zetak[0][im] = omega[0][im]; zetak[1][im] = omega[1][im]; }
        }
  }

  /* U2advection_k (operate_omega.cxx:12-98) */
  void u2advection_k(double **u, double **adv) {
#pragma omp parallel for
    for (int i = 0; i < NX; i++)
      for (int j = 0; j < NY; j++)
        for (int k = 0; k < NZ; k++) {
          size_t im = ((size_t)i * NY + j) * NZ_ + k;
          double u1 = u[0][im], u2 = u[1][im], u3 = u[2][im];
          u[0][im] = u2 * u3;
          u[1][im] = u3 * u1;
          u[2][im] = u1 * u2;
          adv[0][im] = (u2 - u3) * (u2 + u3);
          adv[1][im] = (u3 - u1) * (u3 + u1);
        }
    for (int d = 0; d < 3; d++) fft_r2c(u[d]);
    fft_r2c(adv[0]);
    fft_r2c(adv[1]);
#pragma omp parallel for
    for (int i = 0; i < NX; i++)
      for (int j = 0; j < NY; j++)
        for (int k = 0; k < NZ_; k++) {
          size_t im = ((size_t)i * NY + j) * NZ_ + k;
          double k1 = KX(i) * WAVE_X, k2 = KY(j) * WAVE_Y, k3 = KZ(k) * WAVE_Z;
          double u2u3 = u[0][im], u3u1 = u[1][im], u1u2 = u[2][im];
          double u22_u32 = adv[0][im], u32_u12 = adv[1][im];
          double u12_u22 = -u32_u12 - u22_u32;
          double k1k2 = k1 * k2, k2k3 = k2 * k3, k3k1 = k3 * k1;
          u[0][im] = -(k2k3 * u22_u32 + (k3 - k2) * (k3 + k2) * u2u3 + k3k1 * u1u2 - k1k2 * u3u1);
          u[1][im] = -(k3k1 * u32_u12 + (k1 - k3) * (k1 + k3) * u3u1 + k1k2 * u2u3 - k2k3 * u1u2);
          u[2][im] = -(k1k2 * u12_u22 + (k2 - k1) * (k2 + k1) * u1u2 + k2k3 * u3u1 - k3k1 * u2u3);
        }
    omega_k2zeta_k(u, adv);
  }

  double *U(int d) { return fld(KAPSEL_U0 + d); }
  double *UP(int d) { return fld(KAPSEL_UP0 + d); }
  double *FP(int d) { return fld(KAPSEL_F0 + d); }

  /* Zeta_k2advection_k (operate_omega.cxx:100-111); global u is scratch */
  void zeta_k2advection_k(double **zeta, double **adv) {
    truncate(zeta[0]);
    truncate(zeta[1]);
    double *u[3] = {U(0), U(1), U(2)};
    zeta_k2u_k(zeta, uk_dc, u);
    for (int d = 0; d < 3; d++) fft_c2r(u[d]);
    u2advection_k(u, adv);
  }

  /* NS_solver_slavedEuler (fluid_solver.cxx:340-366); f_ns0 is WORK0/WORK1 */
  void ns_slaved_euler(const kapsel_time &jikan) {
    double *zeta[2] = {fld(KAPSEL_ZETA0), fld(KAPSEL_ZETA1)};
    double *f_ns0[2] = {fld(KAPSEL_WORK0), fld(KAPSEL_WORK1)};
    zeta_k2advection_k(zeta, f_ns0);
    double dmy0 = -NU * jikan.dt_fluid;
    const double dmy1 = 1. / NU;
    /* the four kept boxes (fft_wrapper.cxx:185-197) */
    int rng[4][6] = {{0, TRN_X - 1, 0, TRN_Y - 1, 0, 2 * TRN_Z - 1},
                     {0, TRN_X - 1, NY - TRN_Y + 1, NY - 1, 0, 2 * TRN_Z - 1},
                     {NX - TRN_X + 1, NX - 1, 0, TRN_Y - 1, 0, 2 * TRN_Z - 1},
                     {NX - TRN_X + 1, NX - 1, NY - TRN_Y + 1, NY - 1, 0, 2 * TRN_Z - 1}};
    for (int n = 0; n < 4; n++) {
#pragma omp parallel for
      for (int i = rng[n][0]; i <= rng[n][1]; i++)
        for (int j = rng[n][2]; j <= rng[n][3]; j++)
          for (int k = rng[n][4]; k <= rng[n][5]; k++) {
            size_t im = ((size_t)i * NY + j) * NZ_ + k;
            double dmy = exp(dmy0 * K2(i, j, k)) - 1.;
            double ik2 = IK2(i, j, k);
            zeta[0][im] += dmy * (zeta[0][im] - dmy1 * ik2 * f_ns0[0][im]);
            zeta[1][im] += dmy * (zeta[1][im] - dmy1 * ik2 * f_ns0[1][im]);
          }
    }
  }

  /* Solenoidal_uk (operate_omega.cxx:187-220) */
  void solenoidal_uk(double **u) {
#pragma omp parallel for
    for (int i = 0; i < NX; i++)
      for (int j = 0; j < NY; j++)
        for (int k = 0; k < NZ_; k++) {
          size_t im = ((size_t)i * NY + j) * NZ_ + k;
          double kx = KX(i) * WAVE_X, ky = KY(j) * WAVE_Y, kz = KZ(k) * WAVE_Z;
          double dmy = IK2(i, j, k) * (u[0][im] * kx + u[1][im] * ky + u[2][im] * kz);
          u[0][im] -= dmy * kx;
          u[1][im] -= dmy * ky;
          u[2][im] -= dmy * kz;
        }
  }

  /* Solenoidal_u (operate_omega.h:109-116) */
  void solenoidal_u(double **u) {
    for (int d = 0; d < 3; d++) fft_r2c(u[d]);
    for (int d = 0; d < 3; d++) truncate(u[d]);
    solenoidal_uk(u);
    for (int d = 0; d < 3; d++) fft_c2r(u[d]);
  }

  /* Zeta_k2u (operate_omega.h:138-141) */
  void zeta_k2u() {
    double *zeta[2] = {fld(KAPSEL_ZETA0), fld(KAPSEL_ZETA1)};
    double *u[3] = {U(0), U(1), U(2)};
    zeta_k2u_k(zeta, uk_dc, u);
    for (int d = 0; d < 3; d++) fft_c2r(u[d]);
  }

  /* U2zeta_k (operate_omega.h:212-215) */
  void u2zeta_k() {
    double *zeta[2] = {fld(KAPSEL_ZETA0), fld(KAPSEL_ZETA1)};
    double *u[3] = {U(0), U(1), U(2)};
    for (int d = 0; d < 3; d++) fft_r2c(u[d]);
    u_k2zeta_k(u, zeta, uk_dc);
  }

  /* ------------------------------------------------------------ profile */
  /* H, Phi (profile.h:29-49) */
  inline double H(double x) const { return x > 0 ? exp(-SQ(DX / x)) : 0; }
  inline double Phi(double x, double radius) const {
    double dmy = H(radius + HXI - x);
    return dmy / (dmy + H(x - radius + HXI));
  }
  /* Phi_compact_sin (profile.h:139-152) */
  inline double Phi_compact_sin(double x, double radius) const {
    double dmy_x = radius - x;
    if (fabs(dmy_x) <= HXI) return 0.5 * sin(M_PI * dmy_x / XI) + 0.5;
    return dmy_x > HXI ? 1. : 0.;
  }
  /* Particle_cell (make_phi.h:120-148) */
  inline void Particle_cell(const double *xp, int *x_int, double *residue) const {
    const double idx = 1. / DX;
    for (int d = 0; d < 3; d++) {
      double dmy = xp[d] * idx;
      x_int[d] = (int)dmy;
      residue[d] = (dmy - (double)x_int[d]) * DX;
    }
  }
  /* Relative_coord (make_phi.h:162-179) */
  inline void Relative_coord(const int *cell, const int *x_int, const double *residue,
                             int *r_mesh, double *r) const {
    const int Nl[3] = {NX, NY, NZ};
    for (int d = 0; d < 3; d++) {
      r_mesh[d] = (x_int[d] + cell[d] + Nl[d]) % Nl[d];
      r[d] = (double)cell[d] * DX - residue[d];
    }
  }
  /* Distance0/Distance (interaction.h:122-147) */
  inline double Distance(const double *x1, const double *x2) const {
    double dmy = 0.0;
    for (int d = 0; d < 3; d++) {
      double x12 = x2[d] - x1[d];
      x12 -= (double)Nint(x12 / L[d]) * L[d];
      dmy += SQ(x12);
    }
    return sqrt(dmy);
  }
  inline void Distance0(const double *x1, const double *x2, double &r12, double *x12) const {
    double dmy = 0.0;
    for (int d = 0; d < 3; d++) {
      x12[d] = x2[d] - x1[d];
      x12[d] -= (double)Nint(x12[d] / L[d]) * L[d];
      dmy += SQ(x12[d]);
    }
    r12 = sqrt(dmy);
  }
  /* Angular2v_rot_on / _rot_off (make_phi.h:400-417) */
  inline void Angular2v(const double *omega, const double *r, double *v) const {
    if (P.rotation) {
      v[0] = omega[1] * r[2] - omega[2] * r[1];
      v[1] = omega[2] * r[0] - omega[0] * r[2];
      v[2] = omega[0] * r[1] - omega[1] * r[0];
    } else {
      v[0] = v[1] = v[2] = 0.0;
    }
  }
  inline void getx(int n, double *xp) { for (int d = 0; d < 3; d++) xp[d] = pf(KAPSEL_P_X, d)[n]; }

  /* Make_phi_particle_sum_primitive (make_phi.cxx:229-275); phi_sum is
   * accumulated (caller resets, sp_3d_ns.cxx:51-53) */
  void make_phi_particle_sum(double *phi, double *phi_sum) {
#pragma omp parallel for
    for (int n = 0; n < Np; n++) {
      double xp[3]; getx(n, xp);
      int x_int[3]; double residue[3];
      Particle_cell(xp, x_int, residue);
      int r_mesh[3]; double r[3], x[3];
      for (int mesh = 0; mesh < NP_domain; mesh++) {
        Relative_coord(&sekibun[3 * mesh], x_int, residue, r_mesh, r);
        for (int d = 0; d < 3; d++) x[d] = r_mesh[d] * DX;
        double dmy = Distance(x, xp);
        double dmy_phi = Phi(dmy, RADIUS);
        size_t im = ((size_t)r_mesh[0] * NY + r_mesh[1]) * NZ_ + r_mesh[2];
#pragma omp atomic
        phi_sum[im] += dmy_phi;
      }
    }
#pragma omp parallel for
    for (int i = 0; i < NX; i++)
      for (int j = 0; j < NY; j++)
        for (int k = 0; k < NZ; k++) {
          size_t im = ((size_t)i * NY + j) * NZ_ + k;
          phi[im] = std::min(phi_sum[im], 1.0);
        }
  }

  /* Make_phi_u_primitive with SW_UP = 0 (make_phi.cxx:146-227): unclamped sum */
  void make_phi_particle(double *phi) {
#pragma omp parallel for
    for (int n = 0; n < Np; n++) {
      double xp[3]; getx(n, xp);
      int x_int[3]; double residue[3];
      Particle_cell(xp, x_int, residue);
      int r_mesh[3]; double r[3], x[3];
      for (int mesh = 0; mesh < NP_domain; mesh++) {
        Relative_coord(&sekibun[3 * mesh], x_int, residue, r_mesh, r);
        for (int d = 0; d < 3; d++) x[d] = r_mesh[d] * DX;
        double dmy = Distance(x, xp);
        double dmy_phi = Phi(dmy, RADIUS);
        size_t im = ((size_t)r_mesh[0] * NY + r_mesh[1]) * NZ_ + r_mesh[2];
#pragma omp atomic
        phi[im] += dmy_phi;
      }
    }
  }

  /* Make_phi_sin_primitive (operate_dielectric.cxx:78-120) */
  void make_phi_sin(double *phi) {
#pragma omp parallel for
    for (int n = 0; n < Np; n++) {
      double xp[3]; getx(n, xp);
      int x_int[3]; double residue[3];
      Particle_cell(xp, x_int, residue);
      int r_mesh[3]; double r[3], x[3];
      for (int mesh = 0; mesh < NP_domain; mesh++) {
        Relative_coord(&sekibun[3 * mesh], x_int, residue, r_mesh, r);
        for (int d = 0; d < 3; d++) x[d] = r_mesh[d] * DX;
        double dmy = Distance(x, xp);
        double dmy_phi = Phi_compact_sin(dmy, RADIUS);
        size_t im = ((size_t)r_mesh[0] * NY + r_mesh[1]) * NZ_ + r_mesh[2];
#pragma omp atomic
        phi[im] += dmy_phi;
      }
    }
  }

  /* Make_u_particle_sum_primitive (make_phi.cxx:329-370); up accumulated */
  void make_u_particle_sum(double **up, const double *phi_sum) {
#pragma omp parallel for
    for (int n = 0; n < Np; n++) {
      double xp[3], vp[3], om[3];
      for (int d = 0; d < 3; d++) {
        xp[d] = pf(KAPSEL_P_X, d)[n]; vp[d] = pf(KAPSEL_P_V, d)[n]; om[d] = pf(KAPSEL_P_OMEGA, d)[n];
      }
      int x_int[3]; double residue[3];
      Particle_cell(xp, x_int, residue);
      int r_mesh[3]; double r[3], x[3], v_rot[3];
      for (int mesh = 0; mesh < NP_domain; mesh++) {
        Relative_coord(&sekibun[3 * mesh], x_int, residue, r_mesh, r);
        for (int d = 0; d < 3; d++) x[d] = r_mesh[d] * DX;
        size_t im = ((size_t)r_mesh[0] * NY + r_mesh[1]) * NZ_ + r_mesh[2];
        double dmy = Distance(x, xp);
        double dmy_phi = Phi(dmy, RADIUS) / std::max(phi_sum[im], 1.0);
        Angular2v(om, r, v_rot);
#pragma omp atomic
        up[0][im] += ((vp[0] + v_rot[0]) * dmy_phi);
#pragma omp atomic
        up[1][im] += ((vp[1] + v_rot[1]) * dmy_phi);
#pragma omp atomic
        up[2][im] += ((vp[2] + v_rot[2]) * dmy_phi);
      }
    }
  }

  /* Calc_f_hydro_correct_precision (md_force.cxx:340-449), free spheres */
  void calc_f_hydro(const double *phi_sum, double **u, const kapsel_time &jikan) {
    const double dmy0 = -DX3 * P.rho;
    double dmy = dmy0 / jikan.dt_fluid;
#pragma omp parallel for
    for (int n = 0; n < Np; n++) {
      double xp[3], vp[3], om[3], force[3] = {0, 0, 0}, torque[3] = {0, 0, 0};
      for (int d = 0; d < 3; d++) {
        xp[d] = pf(KAPSEL_P_X, d)[n]; vp[d] = pf(KAPSEL_P_V, d)[n]; om[d] = pf(KAPSEL_P_OMEGA, d)[n];
      }
      int x_int[3]; double residue[3];
      Particle_cell(xp, x_int, residue);
      int r_mesh[3]; double r[3], x[3], v_rot[3], dmy_fp[3];
      for (int mesh = 0; mesh < NP_domain; mesh++) {
        Relative_coord(&sekibun[3 * mesh], x_int, residue, r_mesh, r);
        for (int d = 0; d < 3; d++) x[d] = r_mesh[d] * DX;
        double dmyR = Distance(x, xp);
        Angular2v(om, r, v_rot);
        size_t im = ((size_t)r_mesh[0] * NY + r_mesh[1]) * NZ_ + r_mesh[2];
        double dmy_phi = Phi(dmyR, RADIUS) / std::max(phi_sum[im], 1.0);
        for (int d = 0; d < 3; d++) {
          dmy_fp[d] = ((vp[d] + v_rot[d]) - u[d][im]) * dmy_phi;
          force[d] += dmy_fp[d];
        }
        torque[0] += (r[1] * dmy_fp[2] - r[2] * dmy_fp[1]);
        torque[1] += (r[2] * dmy_fp[0] - r[0] * dmy_fp[2]);
        torque[2] += (r[0] * dmy_fp[1] - r[1] * dmy_fp[0]);
      }
      for (int d = 0; d < 3; d++) pf(KAPSEL_P_F_HYDRO, d)[n] = dmy * force[d];
      if (P.rotation)
        for (int d = 0; d < 3; d++) pf(KAPSEL_P_TORQUE_HYDRO, d)[n] = dmy * torque[d];
    }
  }

  /* Lennard_Jones_f (interaction.h:175-200) */
  inline double Lennard_Jones_f(double x, double sigma) const {
    int pw = P.lj_powers;
    double coeff = pw == 0 ? 24. * P.lj_epsilon : pw == 1 ? 48. * P.lj_epsilon : 72. * P.lj_epsilon;
    double dmy = sigma / x;
    double d2 = SQ(dmy);
    if (pw == 0) dmy = d2 * d2 * d2;
    else if (pw == 1) dmy = SQ(d2 * d2 * d2);
    else dmy = SQ(d2 * d2 * d2) * (d2 * d2 * d2);
    return coeff / SQ(x) * (2.0 * SQ(dmy) - dmy);
  }

  /* Calc_f_Lennard_Jones_shear_cap_primitive_lnk (md_force.cxx:16-171),
   * cap = DBL_MAX, free spheres */
  void lj_lnk() {
    const int Cell_length = 16;
    const double pair_cutoff = A_R_cutoff * P.lj_sigma;
    int lc[3] = {NX / Cell_length, NY / Cell_length, NZ / Cell_length};
    double lc_r = Cell_length * DX;
    int lcyz = lc[1] * lc[2], lcxyz = lc[0] * lcyz;
    std::vector<int> head(lcxyz, -1), lscl(Np);
    for (int n = 0; n < Np; n++) {
      int mc[3];
      for (int d = 0; d < 3; d++) mc[d] = int(pf(KAPSEL_P_X, d)[n] / lc_r);
      int cn = mc[0] * lcyz + mc[1] * lc[2] + mc[2];
      lscl[n] = head[cn];
      head[cn] = n;
    }
    int ic[3];
    for (ic[0] = 0; ic[0] < lc[0]; ic[0]++)
      for (ic[1] = 0; ic[1] < lc[1]; ic[1]++)
        for (ic[2] = 0; ic[2] < lc[2]; ic[2]++) {
          int cn = ic[0] * lcyz + ic[1] * lc[2] + ic[2];
          int icl[3], icls[3];
          for (icl[0] = ic[0] - 1; icl[0] <= ic[0] + 1; icl[0]++)
            for (icl[1] = ic[1] - 1; icl[1] <= ic[1] + 1; icl[1]++)
              for (icl[2] = ic[2] - 1; icl[2] <= ic[2] + 1; icl[2]++) {
                for (int d = 0; d < 3; d++) {
                  icls[d] = icl[d];
                  if (icls[d] < 0) icls[d] = lc[d] - 1;
                  if (icls[d] > lc[d] - 1) icls[d] = 0;
                }
                int cl = ((icls[0] + lc[0]) % lc[0]) * lcyz + ((icls[1] + lc[1]) % lc[1]) * lc[2] +
                         ((icls[2] + lc[2]) % lc[2]);
                int i = head[cn];
                while (i != -1) {
                  int j = head[cl];
                  while (j != -1) {
                    if (i > j) {
                      double xi[3], xj[3], rv[3], rij;
                      getx(i, xi); getx(j, xj);
                      Distance0(xi, xj, rij, rv);
                      if (rij < pair_cutoff) {
                        double dmy_r = Lennard_Jones_f(rij, P.lj_sigma);
                        for (int d = 0; d < 3; d++) {
                          double fi = dmy_r * (-rv[d]);
                          pf(KAPSEL_P_FR, d)[i] += fi;
                          pf(KAPSEL_P_FR, d)[j] -= fi;
                        }
                      }
                    }
                    j = lscl[j];
                  }
                  i = lscl[i];
                }
              }
        }
  }

  /* Force (particle_solver.h:92-104) */
  void force() {
    if (P.lj_truncate >= 0) lj_lnk();
    for (int n = 0; n < Np; n++)
      for (int d = 0; d < 3; d++) pf(KAPSEL_P_FR, d)[n] += P.body_force[d];
  }

  /* reset_Forces (particle_solver.cxx:11-37) */
  void reset_forces() {
    for (int n = 0; n < Np; n++)
      for (int d = 0; d < 3; d++) {
        pf(KAPSEL_P_FR_PREVIOUS, d)[n] = pf(KAPSEL_P_FR, d)[n];
        pf(KAPSEL_P_FR, d)[n] = 0.0;
        pf(KAPSEL_P_F_HYDRO_PREVIOUS, d)[n] = pf(KAPSEL_P_F_HYDRO, d)[n];
        pf(KAPSEL_P_F_HYDRO, d)[n] = 0.0;
        pf(KAPSEL_P_TORQUE_R_PREVIOUS, d)[n] = pf(KAPSEL_P_TORQUE_R, d)[n];
        pf(KAPSEL_P_TORQUE_R, d)[n] = 0.0;
        pf(KAPSEL_P_TORQUE_HYDRO_PREVIOUS, d)[n] = pf(KAPSEL_P_TORQUE_HYDRO, d)[n];
        pf(KAPSEL_P_TORQUE_HYDRO, d)[n] = 0.0;
      }
  }

  /* quaternion product q = alpha qa.qb (quaternion.h:173-183); q = (s, v) */
  static void qprod(double *q, const double *a, const double *b, double alpha) {
    q[0] = alpha * (a[0] * b[0] - (a[1] * b[1] + a[2] * b[2] + a[3] * b[3]));
    q[1] = alpha * (a[0] * b[1] + b[0] * a[1] + a[2] * b[3] - a[3] * b[2]);
    q[2] = alpha * (a[0] * b[2] + b[0] * a[2] + a[3] * b[1] - a[1] * b[3]);
    q[3] = alpha * (a[0] * b[3] + b[0] * a[3] + a[1] * b[2] - a[2] * b[1]);
  }
  /* v' = q.(0,v).q* (qtn_qvq_prod, quaternion.h:207-216) */
  static void qvq(double *out, const double *q, const double *v) {
    double qs2 = q[0] * q[0] - (q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    double qdotp = q[1] * v[0] + q[2] * v[1] + q[3] * v[2];
    out[0] = qs2 * v[0] + 2.0 * (qdotp * q[1] + q[0] * (q[2] * v[2] - q[3] * v[1]));
    out[1] = qs2 * v[1] + 2.0 * (qdotp * q[2] + q[0] * (q[3] * v[0] - q[1] * v[2]));
    out[2] = qs2 * v[2] + 2.0 * (qdotp * q[3] + q[0] * (q[1] * v[1] - q[2] * v[0]));
  }
  /* qtn_normalize (quaternion.h:283-292) */
  static void qnormalize(double *q) {
    double qnorm = 1.0 / sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    for (int c = 0; c < 4; c++) q[c] *= qnorm;
  }
  /* MD_solver_orientation_Euler / _AB2 (particle_rotation_solver.h:15-39)
   * with qdot (rigid_body.cxx:151-165) and rigid_body_rotation SPACE2BODY
   * (rigid_body.cxx:10-24) */
  void md_orientation(int n, bool euler, double dt, double hdt) {
    double q[4], qo[4], om[3], omo[3];
    for (int c = 0; c < 4; c++) { q[c] = qs[(size_t)c * Np + n]; qo[c] = qs[(size_t)(4 + c) * Np + n]; }
    for (int d = 0; d < 3; d++) { om[d] = pf(KAPSEL_P_OMEGA, d)[n]; omo[d] = pf(KAPSEL_P_OMEGA_OLD, d)[n]; }
    double dq[4];
    if (euler) {
      for (int c = 0; c < 4; c++) qo[c] = q[c];
      const double qw[4] = {0.0, om[0], om[1], om[2]};
      qprod(dq, qw, q, 0.5);                       /* SPACE_FRAME */
      for (int c = 0; c < 4; c++) q[c] += dq[c] * dt;
    } else {
      double wb[3], wbo[3];
      const double qc[4] = {q[0], -q[1], -q[2], -q[3]}, qoc[4] = {qo[0], -qo[1], -qo[2], -qo[3]};
      qvq(wb, qc, om);
      qvq(wbo, qoc, omo);
      for (int d = 0; d < 3; d++) wb[d] = 3.0 * wb[d] - wbo[d];
      for (int c = 0; c < 4; c++) qo[c] = q[c];
      const double qw[4] = {0.0, wb[0], wb[1], wb[2]};
      qprod(dq, q, qw, 0.5);                       /* BODY_FRAME */
      for (int c = 0; c < 4; c++) q[c] += dq[c] * hdt;
    }
    qnormalize(q);
    for (int c = 0; c < 4; c++) { qs[(size_t)c * Np + n] = q[c]; qs[(size_t)(4 + c) * Np + n] = qo[c]; }
  }

  /* MD_solver_position_Euler/_AB2 (particle_solver.cxx:39-91); PBC
   * (periodic_boundary.h:9-19); orientations with ROTATION */
  void md_position(const kapsel_time &jikan) {
    for (int n = 0; n < Np; n++) {
      for (int d = 0; d < 3; d++) {
        double &x = pf(KAPSEL_P_X, d)[n];
        pf(KAPSEL_P_X_PREVIOUS, d)[n] = x;
        double delta_x = jikan.ts == 0 ? jikan.dt_md * pf(KAPSEL_P_V, d)[n]
                                       : jikan.hdt_md * (3.0 * pf(KAPSEL_P_V, d)[n] - pf(KAPSEL_P_V_OLD, d)[n]);
        pf(KAPSEL_P_X_NOPBC, d)[n] += delta_x;
        x += delta_x;
        x = fmod(x + L[d], L[d]);
      }
      if (P.rotation) md_orientation(n, jikan.ts == 0, jikan.dt_md, jikan.hdt_md);
    }
  }

  /* MD_solver_velocity_Euler / _AB2_hydro (particle_solver.cxx:109-150, 252-295) */
  void md_velocity(const kapsel_time &jikan) {
    force();
    bool euler = jikan.ts == 0;
    double dmy = (euler ? jikan.dt_md : jikan.hdt_md) * IMASS;
    double dmy_rot = (euler ? jikan.dt_md : jikan.hdt_md) * IMOI;
    for (int n = 0; n < Np; n++)
      for (int d = 0; d < 3; d++) {
        double &v = pf(KAPSEL_P_V, d)[n];
        double &om = pf(KAPSEL_P_OMEGA, d)[n];
        pf(KAPSEL_P_V_OLD, d)[n] = v;
        pf(KAPSEL_P_OMEGA_OLD, d)[n] = om;
        if (euler) {
          v += dmy * (pf(KAPSEL_P_F_HYDRO, d)[n] + pf(KAPSEL_P_FR, d)[n]);
          om += dmy_rot * (pf(KAPSEL_P_TORQUE_HYDRO, d)[n] + pf(KAPSEL_P_TORQUE_R, d)[n]);
        } else {
          v += dmy * (2.0 * pf(KAPSEL_P_F_HYDRO, d)[n] + pf(KAPSEL_P_FR, d)[n] + pf(KAPSEL_P_FR_PREVIOUS, d)[n]);
          om += dmy_rot * (2.0 * pf(KAPSEL_P_TORQUE_HYDRO, d)[n] + pf(KAPSEL_P_TORQUE_R, d)[n] +
                           pf(KAPSEL_P_TORQUE_R_PREVIOUS, d)[n]);
        }
      }
    reset_forces();
  }

  void reset(double *a) { memset(a, 0, F * sizeof(double)); }

  /* Make_f_particle_dt_sole + Add_f_particle (f_particle.h:37-95) */
  void make_f_particle_dt_sole_add() {
    double *f[3] = {FP(0), FP(1), FP(2)};
    double *u[3] = {U(0), U(1), U(2)};
    double *up[3] = {UP(0), UP(1), UP(2)};
    double *phi = fld(plant == 3 ? KAPSEL_PHI_SUM : KAPSEL_PHI);
#pragma omp parallel for
    for (int i = 0; i < NX; i++)
      for (int j = 0; j < NY; j++)
        for (int k = 0; k < NZ; k++) {
          size_t im = ((size_t)i * NY + j) * NZ_ + k;
          for (int d = 0; d < 3; d++) f[d][im] = up[d][im] - phi[im] * u[d][im];
        }
    solenoidal_u(f);
    const double sgn = plant == 1 ? -1.0 : 1.0;
#pragma omp parallel for
    for (int i = 0; i < NX; i++)
      for (int j = 0; j < NY; j++)
        for (int k = 0; k < NZ; k++) {
          size_t im = ((size_t)i * NY + j) * NZ_ + k;
          for (int d = 0; d < 3; d++) u[d][im] += sgn * f[d][im];
        }
  }

  /* Time_evolution_hydro, Navier-Stokes branch (sp_3d_ns.cxx:24-186) */
  void spm_step(kapsel_time &jikan) {
    ns_slaved_euler(jikan);
    for (int d = 0; d < 3; d++)
      if (P.fix_cell[d]) uk_dc[d] = 0.0;
    truncate(fld(KAPSEL_ZETA0));
    truncate(fld(KAPSEL_ZETA1));
    zeta_k2u();
    if (plant != 4) md_position(jikan);
    reset(fld(KAPSEL_PHI));
    reset(fld(KAPSEL_PHI_SUM));
    make_phi_particle_sum(fld(KAPSEL_PHI), fld(KAPSEL_PHI_SUM));
    if (plant == 4) md_position(jikan);
    for (int d = 0; d < 3; d++) reset(UP(d));
    double *u[3] = {U(0), U(1), U(2)};
    if (plant == 2) md_velocity(jikan);
    calc_f_hydro(fld(KAPSEL_PHI_SUM), u, jikan);
    if (plant != 2) md_velocity(jikan);
    if (thermo_on && thermo_kbt > 0.) thermostat(jikan);  /* sp_3d_ns.cxx:155-157 */
    for (int d = 0; d < 3; d++) reset(UP(d));
    double *up[3] = {UP(0), UP(1), UP(2)};
    make_u_particle_sum(up, fld(KAPSEL_PHI_SUM));
    make_f_particle_dt_sole_add();
    u2zeta_k();
    jikan.time += jikan.dt_fluid;
    jikan.ts += 1;
  }

  /* ------------------------------------------------ dielectric Poisson */
  std::vector<double> pw[5];  /* padded scratch of the operator */
  double *scratch(int i) {
    if (pw[i].size() != F) pw[i].assign(F, 0.0);
    return pw[i].data();
  }
  /* extract_id (periodic_boundary.h:21-27) */
  inline size_t extract_id(int i, int j, int k) const {
    int ii = (i + NX) % NX, jj = (j + NY) % NY, kk = (k + NZ) % NZ;
    return ((size_t)ii * NY + jj) * NZ_ + kk;
  }
  /* Make_dielectric_field, non-Janus branch (operate_dielectric.cxx:270-309) */
  void make_dielectric_field(double *eps, double *phi_p) {
    reset(phi_p);
    reset(eps);
    make_phi_sin(phi_p);
    const double ef = P.eps_fluid, ep = P.eps_particle, dc = P.dielectric_cst;
#pragma omp parallel for
    for (int i = 0; i < NX; i++)
      for (int j = 0; j < NY; j++)
        for (int k = 0; k < NZ; k++) {
          size_t im = ((size_t)i * NY + j) * NZ_ + k;
          eps[im] = ef * dc + (ep - ef) * dc * phi_p[im];
        }
  }
  /* Interpolate_scalar_on_staggered_grid (operate_dielectric.cxx:57-76) */
  void interp_staggered(double **vec, const double *s, const double *factor) {
    for (int d = 0; d < 3; d++) reset(vec[d]);
#pragma omp parallel for
    for (int i = 0; i < NX; i++)
      for (int j = 0; j < NY; j++)
        for (int k = 0; k < NZ; k++) {
          size_t im = ((size_t)i * NY + j) * NZ_ + k;
          double sn = 0.5 * s[im];
          vec[0][im] = (0.5 * s[extract_id(i + 1, j, k)] + sn) * factor[0];
          vec[1][im] = (0.5 * s[extract_id(i, j + 1, k)] + sn) * factor[1];
          vec[2][im] = (0.5 * s[extract_id(i, j, k + 1)] + sn) * factor[2];
        }
  }
  /* Shift_d = exp(i DX WAVE_d k_d / 2) (fft_wrapper.cxx:83-85) */
  inline cplx shift(double wave, int kint) const { return std::exp(cplx(0, 0.5 * DX * wave * (double)kint)); }
  /* A_k2da_k (fft_wrapper.h:751-755) then Shift_vec_fw_imag (fft_wrapper.h:841-858) */
  void grad_shift_k(const double *a, double **da) {
#pragma omp parallel for
    for (int i = 0; i < NX; i++)
      for (int j = 0; j < NY; j++)
        for (int k = 0; k < HNZ_; k++) {
          size_t im = ((size_t)i * NY + j) * NZ_ + 2 * k;
          int kk[3] = {KX(i), KY(j), k};
          double wv[3] = {WAVE_X, WAVE_Y, WAVE_Z};
          for (int d = 0; d < 3; d++) {
            double wn = kk[d] * wv[d];
            cplx g(-wn * a[im + 1], wn * a[im]);
            g *= shift(wv[d], kk[d]);
            da[d][im] = g.real();
            da[d][im + 1] = g.imag();
          }
        }
  }
  /* U_k2divergence_k_shift (fft_wrapper.h:879-900) */
  void div_shift_k(double **uk, double *out) {
#pragma omp parallel for
    for (int i = 0; i < NX; i++)
      for (int j = 0; j < NY; j++)
        for (int k = 0; k < HNZ_; k++) {
          size_t im = ((size_t)i * NY + j) * NZ_ + 2 * k;
          int kk[3] = {KX(i), KY(j), k};
          double wv[3] = {WAVE_X, WAVE_Y, WAVE_Z};
          cplx acc(0, 0);
          for (int d = 0; d < 3; d++) {
            cplx ks = (double)kk[d] * wv[d] * std::conj(shift(wv[d], kk[d]));
            acc += cplx(uk[d][im], uk[d][im + 1]) * ks;
          }
          acc *= cplx(0, 1);
          out[im] = acc.real();
          out[im + 1] = acc.imag();
        }
  }
  /* Ax (operate_dielectric.cxx:485-522) on padded fields; reads EPSILON */
  void ax(const double *x, double *out) {
    double *w1 = scratch(0);
    double *w3[3] = {scratch(1), scratch(2), scratch(3)};
    reset(w1);
    for (int d = 0; d < 3; d++) reset(w3[d]);
#pragma omp parallel for
    for (int i = 0; i < NX; i++)
      for (int j = 0; j < NY; j++)
        for (int k = 0; k < NZ; k++) {
          size_t im = ((size_t)i * NY + j) * NZ_ + k;
          w1[im] = x[im];
        }
    fft_r2c(w1);
    grad_shift_k(w1, w3);
    for (int d = 0; d < 3; d++) fft_c2r(w3[d]);
    const double *eps = fld(KAPSEL_EPSILON);
#pragma omp parallel for
    for (int i = 0; i < NX; i++)
      for (int j = 0; j < NY; j++)
        for (int k = 0; k < NZ; k++) {
          size_t im = ((size_t)i * NY + j) * NZ_ + k;
          double en = eps[im];
          w3[0][im] *= 0.5 * (eps[extract_id(i + 1, j, k)] + en);
          w3[1][im] *= 0.5 * (eps[extract_id(i, j + 1, k)] + en);
          w3[2][im] *= 0.5 * (eps[extract_id(i, j, k + 1)] + en);
        }
    for (int d = 0; d < 3; d++) fft_r2c(w3[d]);
    div_shift_k(w3, w1);
    fft_c2r(w1);
#pragma omp parallel for
    for (int i = 0; i < NX; i++)
      for (int j = 0; j < NY; j++)
        for (int k = 0; k < NZ_; k++) {
          size_t im = ((size_t)i * NY + j) * NZ_ + k;
          out[im] = k < NZ ? w1[im] : 0.0;
        }
  }
  /* Build_rhs (operate_dielectric.cxx:524-540) */
  void build_rhs(double *rhs, const double *rho, const double *eps, const double *e0) {
    double *w3[3] = {scratch(1), scratch(2), scratch(3)};
    reset(rhs);
    interp_staggered(w3, eps, e0);
    for (int d = 0; d < 3; d++) fft_r2c(w3[d]);
    div_shift_k(w3, rhs);
    fft_c2r(rhs);
#pragma omp parallel for
    for (int i = 0; i < NX; i++)
      for (int j = 0; j < NY; j++)
        for (int k = 0; k < NZ; k++) {
          size_t im = ((size_t)i * NY + j) * NZ_ + k;
          rhs[im] -= rho[im];
        }
  }
  double dot(const double *a, const double *b) {
    double s = 0.0;
#pragma omp parallel for reduction(+ : s)
    for (int i = 0; i < NX; i++)
      for (int j = 0; j < NY; j++)
        for (int k = 0; k < NZ; k++) {
          size_t im = ((size_t)i * NY + j) * NZ_ + k;
          s += a[im] * b[im];
        }
    return s;
  }
  void axpy(double a, const double *x, double *y) {
#pragma omp parallel for
    for (size_t i = 0; i < F; i++) y[i] += a * x[i];
  }
  /* spectral preconditioner z = IFFT(FFT(r) / (eps_mean |k|^2)), mean removed */
  void precond_spectral(const double *r, double *z, double eps_mean) {
    memcpy(z, r, F * sizeof(double));
    fft_r2c(z);
#pragma omp parallel for
    for (int i = 0; i < NX; i++)
      for (int j = 0; j < NY; j++)
        for (int k = 0; k < NZ_; k++) {
          size_t im = ((size_t)i * NY + j) * NZ_ + k;
          double k2 = K2(i, j, k);
          z[im] = k2 > 0 ? z[im] / (eps_mean * k2) : 0.0;
        }
    fft_c2r(z);
  }
  /* Conjugate gradients on -A x = -b (A = D E G is symmetric negative
   * semidefinite with the constants as null space); stop when
   * ||b - A x|| / ||b|| < tol (Eigen's criterion, operate_dielectric.cxx:559-562) */
  int cg_solve(double *x, const double *b, double tol, int maxit, int precond, double *relres) {
    double *r = scratch(4);
    std::vector<double> zv(F), pv(F), apv(F);
    double *z = zv.data(), *p = pv.data(), *ap = apv.data();
    double bnorm = sqrt(dot(b, b));
    if (bnorm == 0.0) {
      reset(x);
      *relres = 0.0;
      return 0;
    }
    double eps_mean = 0.0;
    if (precond) eps_mean = dot(fld(KAPSEL_EPSILON), scratch_ones()) / ((double)NX * NY * NZ);
    ax(x, ap);
    for (size_t i = 0; i < F; i++) r[i] = ap[i] - b[i];  /* r = -b - (-A x) */
    auto apply_m = [&](const double *rr, double *zz) {
      if (precond) precond_spectral(rr, zz, eps_mean);
      else memcpy(zz, rr, F * sizeof(double));
    };
    apply_m(r, z);
    memcpy(p, z, F * sizeof(double));
    double rz = dot(r, z);
    double rn = sqrt(dot(r, r));
    int it = 0;
    while (rn / bnorm >= tol && it < maxit) {
      ax(p, ap);
      for (size_t i = 0; i < F; i++) ap[i] = -ap[i];
      double alpha = rz / dot(p, ap);
      axpy(alpha, p, x);
      axpy(-alpha, ap, r);
      apply_m(r, z);
      double rz_new = dot(r, z);
      double beta = rz_new / rz;
      rz = rz_new;
#pragma omp parallel for
      for (size_t i = 0; i < F; i++) p[i] = z[i] + beta * p[i];
      rn = sqrt(dot(r, r));
      it++;
    }
    /* true residual */
    ax(x, ap);
    for (size_t i = 0; i < F; i++) r[i] = b[i] - ap[i];
    *relres = sqrt(dot(r, r)) / bnorm;
    return it;
  }
  std::vector<double> ones;
  const double *scratch_ones() {
    if (ones.size() != F) ones.assign(F, 1.0);
    return ones.data();
  }
  /* BiCGSTAB with identity preconditioner, the reference's solver
   * (operate_dielectric.cxx:556-562): van der Vorst's algorithm with a
   * restart when r0.r vanishes, as Eigen implements it */
  int bicgstab_solve(double *x, const double *b, double tol, int maxit, double *relres) {
    std::vector<double> rv(F), r0v(F), pv(F, 0.0), vv(F, 0.0), sv(F), tv(F);
    double *r = rv.data(), *r0 = r0v.data(), *p = pv.data(), *v = vv.data(), *sq = sv.data(), *t = tv.data();
    ax(x, t);
    for (size_t i = 0; i < F; i++) r[i] = b[i] - t[i];
    memcpy(r0, r, F * sizeof(double));
    double rhs2 = dot(b, b);
    if (rhs2 == 0.0) {
      reset(x);
      *relres = 0.0;
      return 0;
    }
    double r0_sq = dot(r0, r0);
    double rho = 1, alpha = 1, w = 1;
    const double tol2 = tol * tol * rhs2, eps2 = 1e-32;
    int it = 0, restarts = 0;
    while (dot(r, r) > tol2 && it < maxit) {
      double rho_old = rho;
      rho = dot(r0, r);
      if (fabs(rho) < eps2 * r0_sq) {
        ax(x, t);
        for (size_t i = 0; i < F; i++) r[i] = b[i] - t[i];
        memcpy(r0, r, F * sizeof(double));
        rho = r0_sq = dot(r, r);
        if (restarts++ == 0) it = 0;
      }
      double beta = (rho / rho_old) * (alpha / w);
#pragma omp parallel for
      for (size_t i = 0; i < F; i++) p[i] = r[i] + beta * (p[i] - w * v[i]);
      ax(p, v);
      alpha = rho / dot(r0, v);
#pragma omp parallel for
      for (size_t i = 0; i < F; i++) sq[i] = r[i] - alpha * v[i];
      ax(sq, t);
      double tt = dot(t, t);
      w = tt > 0 ? dot(t, sq) / tt : 0.0;
#pragma omp parallel for
      for (size_t i = 0; i < F; i++) {
        x[i] += alpha * p[i] + w * sq[i];
        r[i] = sq[i] - w * t[i];
      }
      it++;
    }
    *relres = sqrt(dot(r, r) / rhs2);
    return it;
  }

  /* staggered gradient of a real field: IFFT(i k e^{i k dx/2} FFT(a))
   * (A2a_k, A_k2da_k, Shift_vec_fw_imag, U_k2u; operate_dielectric.cxx:649-662) */
  void staggered_grad(const double *a, double **g) {
    double *w = scratch(0);
    memcpy(w, a, F * sizeof(double));
    fft_r2c(w);
    grad_shift_k(w, g);
    for (int d = 0; d < 3; d++) fft_c2r(g[d]);
  }
  /* Make_Maxwell_force_x_on_fluid after the solve (operate_dielectric.cxx:646-763), AC off,
   * External_field on: cell-centred force in F0..F2; grad eps in WORK1..3, E^2 in WORK0 */
  void maxwell_force(const double *e0) {
    double *ge[3] = {fld(KAPSEL_WORK1), fld(KAPSEL_WORK2), fld(KAPSEL_WORK3)};
    double *fc[3] = {FP(0), FP(1), FP(2)};
    std::vector<double> ev[3];
    double *ef[3];
    for (int d = 0; d < 3; d++) {
      ev[d].assign(F, 0.0);
      ef[d] = ev[d].data();
    }
    staggered_grad(fld(KAPSEL_EPSILON), ge);
    staggered_grad(fld(KAPSEL_POTENTIAL), ef);  /* force[] holds grad(potential) at the faces */
    double *e2 = fld(KAPSEL_WORK0);
    reset(e2);
#pragma omp parallel for
    for (int i = 0; i < NX; i++)
      for (int j = 0; j < NY; j++)
        for (int k = 0; k < NZ; k++) {
          size_t im = ((size_t)i * NY + j) * NZ_ + k;
          double fx = -ef[0][im] + e0[0], fy = -ef[1][im] + e0[1], fz = -ef[2][im] + e0[2];
          double bx = -ef[0][extract_id(i - 1, j, k)] + e0[0];
          double by = -ef[1][extract_id(i, j - 1, k)] + e0[1];
          double bz = -ef[2][extract_id(i, j, k - 1)] + e0[2];
          e2[im] = SQ((fx + bx) / 2.0) + SQ((fy + by) / 2.0) + SQ((fz + bz) / 2.0);
        }
    const double *rho = fld(KAPSEL_RHO_FREE);
    std::vector<double> ff[3];
    for (int d = 0; d < 3; d++) ff[d].assign(F, 0.0);
#pragma omp parallel for
    for (int i = 0; i < NX; i++)
      for (int j = 0; j < NY; j++)
        for (int k = 0; k < NZ; k++) {
          size_t im = ((size_t)i * NY + j) * NZ_ + k;
          size_t nb[3] = {extract_id(i + 1, j, k), extract_id(i, j + 1, k), extract_id(i, j, k + 1)};
          for (int d = 0; d < 3; d++) {
            double efw = -ef[d][im] + e0[d];
            double rho_f = 0.5 * rho[nb[d]] + 0.5 * rho[im];
            double e2f = 0.5 * e2[im] + 0.5 * e2[nb[d]];
            ff[d][im] = rho_f * efw - e2f * ge[d][im] / 2.0;
          }
        }
    /* Interpolate_vec_on_normal_grid (operate_dielectric.cxx:32-55) */
    for (int d = 0; d < 3; d++) reset(fc[d]);
#pragma omp parallel for
    for (int i = 0; i < NX; i++)
      for (int j = 0; j < NY; j++)
        for (int k = 0; k < NZ; k++) {
          size_t im = ((size_t)i * NY + j) * NZ_ + k;
          fc[0][im] = 0.5 * ff[0][im] + 0.5 * ff[0][extract_id(i - 1, j, k)];
          fc[1][im] = 0.5 * ff[1][im] + 0.5 * ff[1][extract_id(i, j - 1, k)];
          fc[2][im] = 0.5 * ff[2][im] + 0.5 * ff[2][extract_id(i, j, k - 1)];
        }
  }
  /* u += f dt / rho, then Solenoidal_u(u) (sp_3d_ns.cxx:83-94) */
  void add_force_to_fluid(const kapsel_time &jikan) {
    double dmy = (plant == 5 ? -1.0 : 1.0) * jikan.dt_fluid * IRHO;
#pragma omp parallel for
    for (int i = 0; i < NX; i++)
      for (int j = 0; j < NY; j++)
        for (int k = 0; k < NZ; k++) {
          size_t im = ((size_t)i * NY + j) * NZ_ + k;
          for (int d = 0; d < 3; d++) U(d)[im] += FP(d)[im] * dmy;
        }
    double *u[3] = {U(0), U(1), U(2)};
    solenoidal_u(u);
  }
  /* ICEP step without ions: the Navier-Stokes step of Time_evolution_hydro with the
   * dielectric force of the Electrolyte+Dielectric branch inserted after the profile
   * scatter (sp_3d_ns.cxx:51-95): eps field, Poisson solve warm-started from the
   * previous potential, Maxwell force, u += f dt / rho, projection */
  int icep_step(kapsel_time &jikan, const double *e0, double tol, int maxit, int solver, double *relres) {
    ns_slaved_euler(jikan);
    for (int d = 0; d < 3; d++)
      if (P.fix_cell[d]) uk_dc[d] = 0.0;
    truncate(fld(KAPSEL_ZETA0));
    truncate(fld(KAPSEL_ZETA1));
    zeta_k2u();
    md_position(jikan);
    reset(fld(KAPSEL_PHI));
    reset(fld(KAPSEL_PHI_SUM));
    make_phi_particle_sum(fld(KAPSEL_PHI), fld(KAPSEL_PHI_SUM));
    make_dielectric_field(fld(KAPSEL_EPSILON), fld(KAPSEL_WORK4));
    build_rhs(fld(KAPSEL_RHS), fld(KAPSEL_RHO_FREE), fld(KAPSEL_EPSILON), e0);
    double *x = fld(KAPSEL_POTENTIAL);
    int it = solver == 2 ? bicgstab_solve(x, fld(KAPSEL_RHS), tol, maxit, relres)
                         : cg_solve(x, fld(KAPSEL_RHS), tol, maxit, solver, relres);
    maxwell_force(e0);
    add_force_to_fluid(jikan);
    double *u[3] = {U(0), U(1), U(2)};
    calc_f_hydro(fld(KAPSEL_PHI_SUM), u, jikan);
    md_velocity(jikan);
    for (int d = 0; d < 3; d++) reset(UP(d));
    double *up[3] = {UP(0), UP(1), UP(2)};
    make_u_particle_sum(up, fld(KAPSEL_PHI_SUM));
    make_f_particle_dt_sole_add();
    u2zeta_k();
    jikan.time += jikan.dt_fluid;
    jikan.ts += 1;
    return it;
  }

  /* -------------------------------------------------------- thermostat */
  /* kBT, alpha_v, alpha_o (input.cxx:280-281, 539-540) and the MT19937
   * stream of mt19937ar.c seeded by init_genrand (MT_seed, fluct.h:34-46);
   * std::mt19937 is the same generator with the same seeding */
  bool thermo_on = false;
  double thermo_kbt = 0.0, thermo_alpha_v = 0.0, thermo_alpha_o = 0.0;
  std::mt19937 mt{5489u};
  /* genrand_real3: (int32 + 0.5) / 2^32, in (0, 1) */
  double real3() { return ((double)mt() + 0.5) * (1.0 / 4294967296.0); }
  /* Gauss2 (fluct.h:52-63): polar method, two deviates */
  void gauss2(double *r) {
    double x1, x2, rsq;
    do {
      x1 = 2. * real3() - 1.;
      x2 = 2. * real3() - 1.;
      rsq = x1 * x1 + x2 * x2;
    } while (rsq >= 1.0 || rsq == 0.0);
    double factor = sqrt(-2. * log(rsq) / rsq);
    r[0] = factor * x1;
    r[1] = factor * x2;
  }
  /* Add_random_force_thermostat (fluct.cxx:9-40), free particles, in
   * particle order; kT_snap_v/_o of input.cxx:276-278 */
  void thermostat(const kapsel_time &jikan) {
    const double kT_snap_v = 1.439, kT_snap_o = 1.22;
    const double zeta_drag = 6. * M_PI * P.eta * RADIUS;
    const double zeta_drag_rot = 8. * M_PI * P.eta * RADIUS * RADIUS * RADIUS;
    const double sdv_v = sqrt(zeta_drag * jikan.dt_md * thermo_kbt * thermo_alpha_v);
    const double sdv_omega = sqrt(zeta_drag_rot * jikan.dt_md * thermo_kbt * thermo_alpha_o);
    const double noise_intensity_v = kT_snap_v * sdv_v;
    const double noise_intensity_o = kT_snap_o * sdv_omega;
    for (int n = 0; n < Np; n++) {
      double dmy[6];
      gauss2(dmy);
      gauss2(dmy + 2);
      gauss2(dmy + 4);
      for (int d = 0; d < 3; d++) {
        pf(KAPSEL_P_V, d)[n] += dmy[d] * noise_intensity_v * IMASS;
        if (P.rotation) pf(KAPSEL_P_OMEGA, d)[n] += dmy[d + 3] * noise_intensity_o * IMOI;
      }
    }
  }

  /* ------------------------------------------------------- electrolyte */
  /* Two ion species without Dielectric (N_spec = 2, "salt").  Scratch
   * buffers stand in for the reference's globals used as working memory
   * (u, up, f_particle, Surface_normal, Concentration_rhs0); none of their
   * contents leaks into a result, so the mapping is free. */
  kapsel_electrolyte EL{};
  bool el_set = false;
  double total_solute[2] = {0.0, 0.0};
  enum { EB_NORMAL = 0, EB_GRAD = 3, EB_FLUX = 6, EB_CHARGE = 9, EB_PHIQQ = 10, EB_WORK = 11,
         EB_RHS = 12, EB_POT = 14, EB_COUNT = 15 };
  std::vector<double> eb[EB_COUNT];
  double *EB(int i) {
    if (eb[i].size() != F) eb[i].assign(F, 0.0);
    return eb[i].data();
  }
  double *C(int n) { return fld(KAPSEL_CONC0 + n); }
  double valency_e(int n) const { return EL.valency[n] * EL.elementary_charge; }
  double ac_factor(const kapsel_time &t) const { return EL.ac ? sin(EL.angular_frequency * t.time) : 1.0; }

  /* DPhi_compact_sin (profile.h:168-181) */
  inline double DPhi_compact_sin(double x) const {
    const double ixi = 1. / XI;
    double dmy_x = RADIUS - x;
    if (fabs(dmy_x) < HXI) return M_PI * .5 * ixi * cos(M_PI * dmy_x * ixi);
    return 0.;
  }

  /* Make_phi_qq_particle (operate_electrolyte.cxx:179-242); phi and surface
   * accumulate (callers reset them); one particle species */
  void make_phi_qq_particle(double *phi, double *surface) {
    const double q = EL.surface_charge * EL.elementary_charge; /* Surface_charge_e (:740-746) */
    double abs_total_surface_charge = 0.;
    for (int n = 0; n < Np; n++) {
      double xp[3]; getx(n, xp);
      int x_int[3]; double residue[3];
      Particle_cell(xp, x_int, residue);
      int r_mesh[3]; double r[3], x[3];
      for (int mesh = 0; mesh < NP_domain; mesh++) {
        Relative_coord(&sekibun[3 * mesh], x_int, residue, r_mesh, r);
        for (int d = 0; d < 3; d++) x[d] = r_mesh[d] * DX;
        double dmy = Distance(x, xp);
        size_t im = ((size_t)r_mesh[0] * NY + r_mesh[1]) * NZ_ + r_mesh[2];
        phi[im] += Phi(dmy, RADIUS);
        surface[im] += q * DPhi_compact_sin(dmy);
      }
      abs_total_surface_charge += fabs(q);
    }
    double dmy = 0.0;
#pragma omp parallel for reduction(+ : dmy)
    for (int i = 0; i < NX; i++)
      for (int j = 0; j < NY; j++)
        for (int k = 0; k < NZ; k++) dmy += fabs(surface[((size_t)i * NY + j) * NZ_ + k]);
    double rescale = abs_total_surface_charge / (dmy * DX3);
#pragma omp parallel for
    for (int i = 0; i < NX; i++)
      for (int j = 0; j < NY; j++)
        for (int k = 0; k < NZ; k++) surface[((size_t)i * NY + j) * NZ_ + k] *= rescale;
  }

  /* Conc_k2charge_field (operate_electrolyte.cxx:45-74) */
  void conc_k2charge_field(double *charge, double *phi_p, double *work) {
    reset(phi_p);
    reset(charge);
    make_phi_qq_particle(phi_p, charge);
    for (int n = 0; n < 2; n++) {
      memcpy(work, C(n), F * sizeof(double)); /* A_k2a_out */
      fft_c2r(work);
      const double ze = valency_e(n);
#pragma omp parallel for
      for (int i = 0; i < NX; i++)
        for (int j = 0; j < NY; j++)
          for (int k = 0; k < NZ; k++) {
            size_t im = ((size_t)i * NY + j) * NZ_ + k;
            charge[im] += ze * work[im] * (1. - phi_p[im]);
          }
    }
  }

  /* Charge_field_k2Coulomb_potential_k_PBC (operate_electrolyte.cxx:76-88) */
  void coulomb_potential_k(double *a) {
    const double idc = 1. / EL.dielectric_cst;
#pragma omp parallel for
    for (int i = 0; i < NX; i++)
      for (int j = 0; j < NY; j++)
        for (int k = 0; k < NZ_; k++) a[((size_t)i * NY + j) * NZ_ + k] *= (IK2(i, j, k) * idc);
  }

  /* A_k2da_k (fft_wrapper.h:755-759, fft_wrapper.cxx:214-246): da_d = i k_d a */
  void grad_k(const double *a, double **da) {
#pragma omp parallel for
    for (int i = 0; i < NX; i++)
      for (int j = 0; j < NY; j++)
        for (int k = 0; k < HNZ_; k++) {
          int k2 = 2 * k;
          size_t im = ((size_t)i * NY + j) * NZ_ + k2;
          const double wn[3] = {KX(i) * WAVE_X, KY(j) * WAVE_Y, KZ(k2) * WAVE_Z};
          for (int d = 0; d < 3; d++) {
            da[d][im] = -wn[d] * a[im + 1];
            da[d][im + 1] = wn[d] * a[im];
          }
        }
  }

  /* U_k2divergence_k (fft_wrapper.cxx:896-920): div = i k . u */
  void divergence_k(double **u, double *div) {
#pragma omp parallel for
    for (int i = 0; i < NX; i++)
      for (int j = 0; j < NY; j++)
        for (int k = 0; k < HNZ_; k++) {
          int k2 = 2 * k;
          size_t im0 = ((size_t)i * NY + j) * NZ_ + k2, im1 = im0 + 1;
          const double ks[3] = {KX(i) * WAVE_X, KY(j) * WAVE_Y, KZ(k2) * WAVE_Z};
          div[im0] = -(ks[0] * u[0][im1] + ks[1] * u[1][im1] + ks[2] * u[2][im1]);
          div[im1] = (ks[0] * u[0][im0] + ks[1] * u[1][im0] + ks[2] * u[2][im0]);
        }
  }

  /* Make_surface_normal (make_phi.cxx:17-61) */
  void make_surface_normal(double **nrm) {
    for (int d = 0; d < 3; d++) reset(nrm[d]);
    for (int n = 0; n < Np; n++) {
      double xp[3]; getx(n, xp);
      int x_int[3]; double residue[3];
      Particle_cell(xp, x_int, residue);
      int r_mesh[3]; double r[3];
      for (int mesh = 0; mesh < NP_domain; mesh++) {
        Relative_coord(&sekibun[3 * mesh], x_int, residue, r_mesh, r);
        double dmy_r = sqrt(SQ(r[0]) + SQ(r[1]) + SQ(r[2]));
        double dmy = fabs(dmy_r - RADIUS);
        if (dmy < HXI) {
          double ir = 1. / dmy_r;
          size_t im = ((size_t)r_mesh[0] * NY + r_mesh[1]) * NZ_ + r_mesh[2];
          for (int d = 0; d < 3; d++) nrm[d][im] += r[d] * ir;
        }
      }
    }
  }

  /* Solute_impermeability (solute_rhs.cxx:42-97), particles in order */
  void solute_impermeability(double **flux, double **nrm) {
    for (int n = 0; n < Np; n++) {
      double xp[3]; getx(n, xp);
      int x_int[3]; double residue[3];
      Particle_cell(xp, x_int, residue);
      int r_mesh[3]; double r[3];
      for (int mesh = 0; mesh < NP_domain; mesh++) {
        Relative_coord(&sekibun[3 * mesh], x_int, residue, r_mesh, r);
        size_t im = ((size_t)r_mesh[0] * NY + r_mesh[1]) * NZ_ + r_mesh[2];
        double normal[3], norm2 = 0.;
        for (int d = 0; d < 3; d++) {
          normal[d] = nrm[d][im];
          norm2 += SQ(normal[d]);
        }
        if (norm2 > 0.) {
          double dmy = 0.;
          for (int d = 0; d < 3; d++) dmy += flux[d][im] * normal[d];
          dmy /= norm2;
          for (int d = 0; d < 3; d++) flux[d][im] -= normal[d] * dmy;
        }
      }
    }
  }

  /* Count_single_solute (solute_rhs.cxx:130-148) */
  double count_single_solute(const double *c, const double *phi) {
    double dmy = 0.;
#pragma omp parallel for reduction(+ : dmy)
    for (int i = 0; i < NX; i++)
      for (int j = 0; j < NY; j++)
        for (int k = 0; k < NZ; k++) {
          size_t im = ((size_t)i * NY + j) * NZ_ + k;
          dmy += (1. - phi[im]) * c[im];
        }
    return dmy * DX3;
  }

  /* Count_solute_each (solute_rhs.h:143-156) */
  void count_solute_each(double *n_solute, double *phi, double *work) {
    reset(phi);
    make_phi_particle(phi);
    for (int n = 0; n < 2; n++) {
      memcpy(work, C(n), F * sizeof(double));
      fft_c2r(work);
      n_solute[n] = count_single_solute(work, phi);
    }
  }

  /* Init_rho_ion (operate_electrolyte.cxx:738-861), N_spec = 2 without
   * Poisson_Boltzmann: Set_uniform_ion_charge_density_salt (:559-621); under
   * Dielectric the bulk density carries the fluid's permittivity */
  void init_rho_ion() {
    double *phi = fld(KAPSEL_PHI);
    reset(phi);
    make_phi_particle(phi);
    double positive_ion_number = 0., negative_ion_number = 0.;
    const double q = EL.surface_charge, npart = (double)Np;
    if (q < 0.) positive_ion_number += fabs(q * npart / EL.valency[0]);
    else if (q > 0.) negative_ion_number += fabs(q * npart / EL.valency[1]);
    double volume_phi = 0.;
#pragma omp parallel for reduction(+ : volume_phi)
    for (int i = 0; i < NX; i++)
      for (int j = 0; j < NY; j++)
        for (int k = 0; k < NZ; k++) volume_phi += (1. - phi[((size_t)i * NY + j) * NZ_ + k]);
    double pd = positive_ion_number / (volume_phi * DX3);
    double nd = negative_ion_number / (volume_phi * DX3);
    double rho_inf = EL.kbt * EL.dielectric_cst / SQ(EL.elementary_charge * EL.debye_length);
    if (EL.dielectric) rho_inf *= P.eps_fluid; /* :598-600 */
    double rp = rho_inf / (EL.valency[0] * (EL.valency[0] - EL.valency[1]));
    double rn = -rho_inf / (EL.valency[1] * (EL.valency[0] - EL.valency[1]));
    reset(C(0));
    reset(C(1));
#pragma omp parallel for
    for (int i = 0; i < NX; i++)
      for (int j = 0; j < NY; j++)
        for (int k = 0; k < NZ; k++) {
          size_t im = ((size_t)i * NY + j) * NZ_ + k;
          C(0)[im] = pd + rp;
          C(1)[im] = nd + rn;
        }
    total_solute[0] = count_single_solute(C(0), phi);
    total_solute[1] = count_single_solute(C(1), phi);
    fft_r2c(C(0));
    fft_r2c(C(1));
    count_solute_each(total_solute, phi, EB(EB_WORK));
  }

  /* Rescale_solute (solute_rhs.cxx:99-128) */
  void rescale_solute(double *factor) {
    double *phi = fld(KAPSEL_PHI), *work = EB(EB_WORK);
    reset(phi);
    make_phi_particle(phi);
    for (int n = 0; n < 2; n++) {
      memcpy(work, C(n), F * sizeof(double));
      fft_c2r(work);
      double dmy = count_single_solute(work, phi);
      factor[n] = total_solute[n] / dmy;
      double *c = C(n);
#pragma omp parallel for
      for (size_t im = 0; im < F; im++) c[im] *= factor[n];
    }
  }

  /* the four kept boxes of the 2/3 rule (fft_wrapper.cxx:185-197) */
  void kept_boxes(int rng[4][6]) const {
    int r[4][6] = {{0, TRN_X - 1, 0, TRN_Y - 1, 0, 2 * TRN_Z - 1},
                   {0, TRN_X - 1, NY - TRN_Y + 1, NY - 1, 0, 2 * TRN_Z - 1},
                   {NX - TRN_X + 1, NX - 1, 0, TRN_Y - 1, 0, 2 * TRN_Z - 1},
                   {NX - TRN_X + 1, NX - 1, NY - TRN_Y + 1, NY - 1, 0, 2 * TRN_Z - 1}};
    memcpy(rng, r, sizeof r);
  }

  /* Rhs_NS_solute (fluid_solver.cxx:184-229) with Diffusion_flux_single,
   * Solute_solver_rhs_nonlinear_x_single, Add_advection_flux
   * (solute_rhs.cxx:150-233) and Rhs_solvent (:172-182) */
  void rhs_ns_solute(double **grad, double **rhs_ns, double **rhs_solute) {
    truncate(fld(KAPSEL_ZETA0));
    truncate(fld(KAPSEL_ZETA1));
    zeta_k2u();
    double *u[3] = {U(0), U(1), U(2)};
    double *nrm[3] = {EB(EB_NORMAL), EB(EB_NORMAL + 1), EB(EB_NORMAL + 2)};
    make_surface_normal(nrm);
    double *flux[3] = {EB(EB_FLUX), EB(EB_FLUX + 1), EB(EB_FLUX + 2)};
    for (int n = 0; n < 2; n++) {
      double *ck = C(n);
      truncate(ck);
      /* Diffusion_flux_single: rhs_ns[0] is its working memory */
      double *w = rhs_ns[0];
      const double kl = EL.kbt * EL.onsager_coeff[n];
#pragma omp parallel for
      for (size_t im = 0; im < F; im++) w[im] = kl * ck[im];
      grad_k(w, flux);
      for (int d = 0; d < 3; d++) fft_c2r(flux[d]);
      double *c = rhs_solute[n];
      memcpy(c, ck, F * sizeof(double));
      fft_c2r(c);
      const double ze = valency_e(n), L = EL.onsager_coeff[n];
#pragma omp parallel for
      for (int i = 0; i < NX; i++)
        for (int j = 0; j < NY; j++)
          for (int k = 0; k < NZ; k++) {
            size_t im = ((size_t)i * NY + j) * NZ_ + k;
            for (int d = 0; d < 3; d++) flux[d][im] += c[im] * (L * (grad[d][im] * ze));
          }
      solute_impermeability(flux, nrm);
#pragma omp parallel for
      for (int i = 0; i < NX; i++)
        for (int j = 0; j < NY; j++)
          for (int k = 0; k < NZ; k++) {
            size_t im = ((size_t)i * NY + j) * NZ_ + k;
            for (int d = 0; d < 3; d++) flux[d][im] += c[im] * -u[d][im];
          }
      for (int d = 0; d < 3; d++) fft_r2c(flux[d]);
      divergence_k(flux, rhs_solute[n]);
    }
    u2advection_k(u, rhs_ns);
    int rng[4][6];
    kept_boxes(rng);
    double *zeta[2] = {fld(KAPSEL_ZETA0), fld(KAPSEL_ZETA1)};
    for (int b = 0; b < 4; b++) {
#pragma omp parallel for
      for (int i = rng[b][0]; i <= rng[b][1]; i++)
        for (int j = rng[b][2]; j <= rng[b][3]; j++)
          for (int k = rng[b][4]; k <= rng[b][5]; k++) {
            size_t im = ((size_t)i * NY + j) * NZ_ + k;
            double k2 = K2(i, j, k);
            rhs_ns[0][im] += -(NU * k2 * zeta[0][im]);
            rhs_ns[1][im] += -(NU * k2 * zeta[1][im]);
          }
    }
  }

  /* Rhs_NS_Nernst_Planck, Dielectric branch (fluid_solver.cxx:261-270): the
   * staggered gradient of the potential (A_k2da_k, Shift_vec_fw_imag), its
   * k = 0 shifted by the applied field (Add_constant_field_k, :231-244),
   * back on the cell centres (Interpolate_vec_on_normal_grid,
   * operate_dielectric.cxx:32-55) */
  void dielectric_grad(const kapsel_time &jikan, double **grad) {
    double *w = scratch(0);
    memcpy(w, fld(KAPSEL_POTENTIAL), F * sizeof(double));
    fft_r2c(w);
    grad_shift_k(w, grad);
    double amp = (double)(NX * NY * NZ) * ac_factor(jikan);
    for (int d = 0; d < 3; d++) grad[d][0] -= EL.e_ext[d] * amp;
    for (int d = 0; d < 3; d++) fft_c2r(grad[d]);
    std::vector<double> tmp[3];
    for (int d = 0; d < 3; d++) tmp[d].assign(grad[d], grad[d] + F);
#pragma omp parallel for
    for (int i = 0; i < NX; i++)
      for (int j = 0; j < NY; j++)
        for (int k = 0; k < NZ; k++) {
          size_t im = ((size_t)i * NY + j) * NZ_ + k;
          grad[0][im] = 0.5 * tmp[0][im] + 0.5 * tmp[0][extract_id(i - 1, j, k)];
          grad[1][im] = 0.5 * tmp[1][im] + 0.5 * tmp[1][extract_id(i, j - 1, k)];
          grad[2][im] = 0.5 * tmp[2][im] + 0.5 * tmp[2][extract_id(i, j, k - 1)];
        }
  }

  /* Rhs_NS_Nernst_Planck, non-Dielectric branch (fluid_solver.cxx:246-292)
   * with Add_constant_field_k (:231-244); the Dielectric branch when EL.dielectric */
  void rhs_ns_nernst_planck(const kapsel_time &jikan, double **rhs_ns, double **rhs_solute) {
    if (EL.dielectric) {
      double *grad[3] = {EB(EB_GRAD), EB(EB_GRAD + 1), EB(EB_GRAD + 2)};
      dielectric_grad(jikan, grad);
      rhs_ns_solute(grad, rhs_ns, rhs_solute);
      return;
    }
    double *pot = EB(EB_CHARGE);
    conc_k2charge_field(pot, EB(EB_PHIQQ), EB(EB_WORK));
    fft_r2c(pot);
    coulomb_potential_k(pot);
    truncate(pot);
    double *grad[3] = {EB(EB_GRAD), EB(EB_GRAD + 1), EB(EB_GRAD + 2)};
    grad_k(pot, grad);
    double amp = (double)(NX * NY * NZ) * ac_factor(jikan);
    for (int d = 0; d < 3; d++) grad[d][0] -= EL.e_ext[d] * amp;
    for (int d = 0; d < 3; d++) fft_c2r(grad[d]);
    rhs_ns_solute(grad, rhs_ns, rhs_solute);
  }

  /* NSsolute_solver_Euler (fluid_solver.cxx:318-338) with Field_solver_Euler
   * (:121-136); rhs_ns is WORK0/WORK1 */
  void nssolute_euler(const kapsel_time &jikan) {
    double *rhs_ns[2] = {fld(KAPSEL_WORK0), fld(KAPSEL_WORK1)};
    double *rhs_solute[2] = {EB(EB_RHS), EB(EB_RHS + 1)};
    rhs_ns_nernst_planck(jikan, rhs_ns, rhs_solute);
    double *zeta[2] = {fld(KAPSEL_ZETA0), fld(KAPSEL_ZETA1)};
    int rng[4][6];
    kept_boxes(rng);
    for (int b = 0; b < 4; b++) {
      for (int d = 0; d < 2; d++)
#pragma omp parallel for
        for (int i = rng[b][0]; i <= rng[b][1]; i++)
          for (int j = rng[b][2]; j <= rng[b][3]; j++)
            for (int k = rng[b][4]; k <= rng[b][5]; k++) {
              size_t im = ((size_t)i * NY + j) * NZ_ + k;
              zeta[d][im] += (jikan.dt_fluid * rhs_ns[d][im]);
            }
      for (int n = 0; n < 2; n++)
#pragma omp parallel for
        for (int i = rng[b][0]; i <= rng[b][1]; i++)
          for (int j = rng[b][2]; j <= rng[b][3]; j++)
            for (int k = rng[b][4]; k <= rng[b][5]; k++) {
              size_t im = ((size_t)i * NY + j) * NZ_ + k;
              C(n)[im] += (jikan.dt_fluid * rhs_solute[n][im]);
            }
    }
    /* uk_dc += dt * dc_rhs with dc_rhs = 0 (Rhs_solvent) */
  }

  /* Calc_free_energy_PB, non-Dielectric branch (operate_electrolyte.cxx:
   * 302-421): fe = {total, ideal, electrostatic} */
  void free_energy_pb(const kapsel_time &jikan, double *fe) {
    double *phi = EB(EB_PHIQQ), *work = EB(EB_WORK), *charge = EB(EB_CHARGE), *pot = EB(EB_POT);
    double ideal = 0., electrostatic = 0.;
    reset(phi);
    make_phi_particle(phi);
    for (int n = 0; n < 2; n++) {
      memcpy(work, C(n), F * sizeof(double));
      fft_c2r(work);
#pragma omp parallel for reduction(+ : ideal)
      for (int i = 0; i < NX; i++)
        for (int j = 0; j < NY; j++)
          for (int k = 0; k < NZ; k++) {
            size_t im = ((size_t)i * NY + j) * NZ_ + k;
            double c = work[im];
            ideal += (1. - phi[im]) * c * (log(c) - 1.);
          }
    }
    ideal *= (EL.kbt * DX3);
    conc_k2charge_field(charge, phi, work);
    memcpy(pot, charge, F * sizeof(double));
    fft_r2c(pot);
    coulomb_potential_k(pot);
    fft_c2r(pot);
    const double s = ac_factor(jikan);
    const double ext[3] = {EL.e_ext[0] * s, EL.e_ext[1] * s, EL.e_ext[2] * s};
    const bool external = EL.e_ext[0] != 0. || EL.e_ext[1] != 0. || EL.e_ext[2] != 0.;
#pragma omp parallel for reduction(+ : electrostatic)
    for (int i = 0; i < NX; i++)
      for (int j = 0; j < NY; j++)
        for (int k = 0; k < NZ; k++) {
          size_t im = ((size_t)i * NY + j) * NZ_ + k;
          double v = pot[im] * 0.5;
          if (external) v -= ((ext[0] * (double)i + ext[1] * (double)j + ext[2] * (double)k) * DX);
          electrostatic += charge[im] * v;
        }
    electrostatic *= DX3;
    fe[1] = ideal;
    fe[2] = electrostatic;
    fe[0] = fe[1] + fe[2];
  }

  /* Ion_diffusion_solver_Euler (fluid_solver.cxx:293-316) as Set_steadystate_ion_density
   * calls it (operate_electrolyte.cxx:436-444): zero velocity and vorticity, the
   * Nernst-Planck right-hand side, the Euler step of the ions, Rescale_solute.  The
   * solvent right-hand side the reference computes there is discarded, so it is not. */
  void ion_diffusion_euler(const kapsel_time &jikan) {
    for (int d = 0; d < 3; d++) reset(U(d));
    double *pot = EB(EB_CHARGE);
    conc_k2charge_field(pot, EB(EB_PHIQQ), EB(EB_WORK));
    fft_r2c(pot);
    coulomb_potential_k(pot);
    truncate(pot);
    double *grad[3] = {EB(EB_GRAD), EB(EB_GRAD + 1), EB(EB_GRAD + 2)};
    grad_k(pot, grad);
    double amp = (double)(NX * NY * NZ) * ac_factor(jikan);
    for (int d = 0; d < 3; d++) grad[d][0] -= EL.e_ext[d] * amp;
    for (int d = 0; d < 3; d++) fft_c2r(grad[d]);
    double *nrm[3] = {EB(EB_NORMAL), EB(EB_NORMAL + 1), EB(EB_NORMAL + 2)};
    make_surface_normal(nrm);
    double *flux[3] = {EB(EB_FLUX), EB(EB_FLUX + 1), EB(EB_FLUX + 2)};
    double *rhs_solute[2] = {EB(EB_RHS), EB(EB_RHS + 1)};
    for (int n = 0; n < 2; n++) {
      double *ck = C(n);
      truncate(ck);
      double *w = fld(KAPSEL_WORK0);
      const double kl = EL.kbt * EL.onsager_coeff[n];
#pragma omp parallel for
      for (size_t im = 0; im < F; im++) w[im] = kl * ck[im];
      grad_k(w, flux);
      for (int d = 0; d < 3; d++) fft_c2r(flux[d]);
      double *c = rhs_solute[n];
      memcpy(c, ck, F * sizeof(double));
      fft_c2r(c);
      const double ze = valency_e(n), L = EL.onsager_coeff[n];
#pragma omp parallel for
      for (int i = 0; i < NX; i++)
        for (int j = 0; j < NY; j++)
          for (int k = 0; k < NZ; k++) {
            size_t im = ((size_t)i * NY + j) * NZ_ + k;
            for (int d = 0; d < 3; d++) flux[d][im] += c[im] * (L * (grad[d][im] * ze));
          }
      solute_impermeability(flux, nrm);
      for (int d = 0; d < 3; d++) fft_r2c(flux[d]);
      divergence_k(flux, rhs_solute[n]);
    }
    int rng[4][6];
    kept_boxes(rng);
    for (int b = 0; b < 4; b++)
      for (int n = 0; n < 2; n++)
#pragma omp parallel for
        for (int i = rng[b][0]; i <= rng[b][1]; i++)
          for (int j = rng[b][2]; j <= rng[b][3]; j++)
            for (int k = rng[b][4]; k <= rng[b][5]; k++) {
              size_t im = ((size_t)i * NY + j) * NZ_ + k;
              C(n)[im] += (jikan.dt_fluid * rhs_solute[n][im]);
            }
    double f[2];
    rescale_solute(f);
  }

  /* Set_steadystate_ion_density (operate_electrolyte.cxx:424-457), TOL = 1e-4
   * (operate_electrolyte.h:22), at most max_iter - 1 relaxation steps (STEP = 10000) */
  int steadystate_ion_density(const kapsel_time &jikan, int max_iter, double *fe_out) {
    double fe[3];
    free_energy_pb(jikan, fe);
    double previous = fe[0];
    int steps = 0;
    for (int m = 1; m < max_iter; m++) {
      ion_diffusion_euler(jikan);
      steps++;
      free_energy_pb(jikan, fe);
      double dum = fabs(fe[0] - previous);
      if (dum < 1.e-4) break;
      previous = fe[0];
    }
    if (fe_out) for (int c = 0; c < 3; c++) fe_out[c] = fe[c];
    return steps;
  }

  /* Set_Poisson_Boltzmann_ion_charge_density_salt (operate_electrolyte.cxx:
   * 623-736): one particle, symmetric z:z salt; damped fixed point (alp = 0.01)
   * of the potential until sum|d psi| / sum|psi| < TOL = 1e-4, at most
   * max_iter - 1 iterations (STEP = 100000); returns the iterations taken */
  int pb_salt_single(int max_iter) {
    const double rho_inf = EL.kbt * EL.dielectric_cst / SQ(EL.elementary_charge * EL.debye_length);
    const double rp = rho_inf / (EL.valency[0] * (EL.valency[0] - EL.valency[1]));
    const double rn = -rho_inf / (EL.valency[1] * (EL.valency[0] - EL.valency[1]));
    const double alp = 0.01;
    const double dmy0 = EL.elementary_charge * EL.valency[0] / EL.kbt;
    const double dmy1 = EL.elementary_charge * EL.valency[1] / EL.kbt;
    const double surface_ion_number = -EL.surface_charge * Np;  /* Init_rho_ion :750-753 */
    double *phi = fld(KAPSEL_PHI);
    reset(phi);
    make_phi_particle(phi);
    double *ep0 = EB(EB_POT), *v0 = EB(EB_PHIQQ), *v1 = EB(EB_CHARGE);
    reset(ep0);
    int iters = 0;
    for (int m = 1; m < max_iter; m++) {
      iters++;
      double sp = 0., sn = 0.;
#pragma omp parallel for reduction(+ : sp, sn)
      for (int i = 0; i < NX; i++)
        for (int j = 0; j < NY; j++)
          for (int k = 0; k < NZ; k++) {
            size_t im = ((size_t)i * NY + j) * NZ_ + k;
            sp += (1. - phi[im]) * exp(-dmy0 * ep0[im]);
            sn += (1. - phi[im]) * exp(-dmy1 * ep0[im]);
          }
      sp *= DX3;
      sn *= DX3;
      const double a = EL.valency[0] * rp * sp, b = -surface_ion_number, c = EL.valency[1] * rn * sn;
      const double nu = (-b + sqrt(b * b - 4. * a * c)) / (2. * a);
#pragma omp parallel for
      for (int i = 0; i < NX; i++)
        for (int j = 0; j < NY; j++)
          for (int k = 0; k < NZ; k++) {
            size_t im = ((size_t)i * NY + j) * NZ_ + k;
            C(0)[im] = rp * exp(-dmy0 * ep0[im]) * nu;
            C(1)[im] = rn * exp(-dmy1 * ep0[im]) / nu;
          }
      reset(v0);
      reset(v1);
      make_phi_qq_particle(v0, v1);
      for (int n = 0; n < 2; n++) {
        const double ze = valency_e(n);
#pragma omp parallel for
        for (int i = 0; i < NX; i++)
          for (int j = 0; j < NY; j++)
            for (int k = 0; k < NZ; k++) {
              size_t im = ((size_t)i * NY + j) * NZ_ + k;
              v1[im] += ze * C(n)[im] * (1. - v0[im]);
            }
      }
      fft_r2c(v1);
      coulomb_potential_k(v1);
      fft_c2r(v1);
      double sum = 0., error = 0.;
#pragma omp parallel for reduction(+ : sum, error)
      for (int i = 0; i < NX; i++)
        for (int j = 0; j < NY; j++)
          for (int k = 0; k < NZ; k++) {
            size_t im = ((size_t)i * NY + j) * NZ_ + k;
            double dmy = alp * (v1[im] - ep0[im]);
            ep0[im] += dmy;
            sum += fabs(ep0[im]);
            error += fabs(dmy);
          }
      if (error / sum < 1.e-4) break;
    }
    fft_r2c(C(0));
    fft_r2c(C(1));
    return iters;
  }

  /* the free charge of the Maxwell force and the initial potential
   * (operate_dielectric.cxx:597-607, 631-643): with surface charge
   * Conc_k2charge_field (smooth profile and surface charge); without,
   * Conc_k2charge_field_no_surfase_charge (:135-160) with the compact-sine
   * profile Make_dielectric_field left in phi_sin */
  void dielectric_free_charge(double *rho, const double *phi_sin) {
    if (EL.surface_charge != 0.0) {
      conc_k2charge_field(rho, EB(EB_PHIQQ), EB(EB_WORK));
      return;
    }
    reset(rho);
    double *work = EB(EB_WORK);
    for (int n = 0; n < 2; n++) {
      memcpy(work, C(n), F * sizeof(double));
      fft_c2r(work);
      const double ze = valency_e(n);
#pragma omp parallel for
      for (int i = 0; i < NX; i++)
        for (int j = 0; j < NY; j++)
          for (int k = 0; k < NZ; k++) {
            size_t im = ((size_t)i * NY + j) * NZ_ + k;
            rho[im] += ze * work[im] * (1. - phi_sin[im]);
          }
    }
  }

  /* Init_rho_ion_dielectric (operate_electrolyte.cxx:912-979): with surface
   * charge Init_rho_ion; without, Set_uniform_ion_charge_density_salt_dielectric
   * (:864-910: bulk density kBT Dielectric_cst eps_fluid / (e lambda_D a)^2, no
   * counter-ions) and Count_solute_each */
  void init_rho_ion_dielectric() {
    if (EL.surface_charge != 0.0) {
      init_rho_ion();
      return;
    }
    double *phi = fld(KAPSEL_PHI);
    reset(phi);
    make_phi_particle(phi);
    double volume_phi = 0.;
#pragma omp parallel for reduction(+ : volume_phi)
    for (int i = 0; i < NX; i++)
      for (int j = 0; j < NY; j++)
        for (int k = 0; k < NZ; k++) volume_phi += (1. - phi[((size_t)i * NY + j) * NZ_ + k]);
    const double pd = 0.0 / (volume_phi * DX3), nd = 0.0 / (volume_phi * DX3);
    const double rho_inf = EL.kbt * EL.dielectric_cst * P.eps_fluid / SQ(EL.elementary_charge * EL.debye_length * RADIUS);
    const double rp = rho_inf / (EL.valency[0] * (EL.valency[0] - EL.valency[1]));
    const double rn = -rho_inf / (EL.valency[1] * (EL.valency[0] - EL.valency[1]));
    reset(C(0));
    reset(C(1));
#pragma omp parallel for
    for (int i = 0; i < NX; i++)
      for (int j = 0; j < NY; j++)
        for (int k = 0; k < NZ; k++) {
          size_t im = ((size_t)i * NY + j) * NZ_ + k;
          C(0)[im] = pd + rp;
          C(1)[im] = nd + rn;
        }
    total_solute[0] = count_single_solute(C(0), phi);
    total_solute[1] = count_single_solute(C(1), phi);
    fft_r2c(C(0));
    fft_r2c(C(1));
    count_solute_each(total_solute, phi, EB(EB_WORK));
  }

  /* the potential of the ions (and surface charge) in the variable
   * permittivity: Make_dielectric_field, the free charge, the solve warm-started
   * from POTENTIAL (Init_potential_dielectric, operate_dielectric.cxx:574-609;
   * Make_Maxwell_force_x_on_fluid :629-643) */
  int dielectric_potential(const kapsel_time &jikan, double tol, int maxit, int solver, double *relres, double *e0) {
    const double s = ac_factor(jikan);
    for (int d = 0; d < 3; d++) e0[d] = EL.e_ext[d] * s;
    make_dielectric_field(fld(KAPSEL_EPSILON), fld(KAPSEL_WORK4));
    dielectric_free_charge(fld(KAPSEL_RHO_FREE), fld(KAPSEL_WORK4));
    build_rhs(fld(KAPSEL_RHS), fld(KAPSEL_RHO_FREE), fld(KAPSEL_EPSILON), e0);
    /* the ions inside the particles (weight phi_sin) leave the free charge
     * with a nonzero total, which no periodic potential satisfies: a uniform
     * background cancels it, so the solve sees the consistent part of the
     * right-hand side (the reference's BiCGSTAB gets the raw total and stops
     * at MaxIter_potential) */
    const double mean = dot(fld(KAPSEL_RHO_FREE), scratch_ones()) / ((double)NX * NY * NZ);
    double *rhs = fld(KAPSEL_RHS);
#pragma omp parallel for
    for (int i = 0; i < NX; i++)
      for (int j = 0; j < NY; j++)
        for (int k = 0; k < NZ; k++) rhs[((size_t)i * NY + j) * NZ_ + k] += mean;
    double *x = fld(KAPSEL_POTENTIAL);
    return solver == 2 ? bicgstab_solve(x, fld(KAPSEL_RHS), tol, maxit, relres)
                       : cg_solve(x, fld(KAPSEL_RHS), tol, maxit, solver, relres);
  }

  /* Time_evolution_hydro, Electrolyte branch with Dielectric (sp_3d_ns.cxx:
   * 20, 24-186): NSsolute_solver_Euler with the potential's gradient, the
   * particle update, no Rescale_solute, Make_Maxwell_force_x_on_fluid with the
   * ions' charge, u += f dt / RHO, Solenoidal_u, the hydrodynamic force, the
   * velocity update, the rigidity constraint, U2zeta_k */
  int icep_electrolyte_step(kapsel_time &jikan, double tol, int maxit, int solver, double *relres) {
    nssolute_euler(jikan);
    for (int d = 0; d < 3; d++)
      if (P.fix_cell[d]) uk_dc[d] = 0.0;
    truncate(fld(KAPSEL_ZETA0));
    truncate(fld(KAPSEL_ZETA1));
    zeta_k2u();
    md_position(jikan);
    reset(fld(KAPSEL_PHI));
    reset(fld(KAPSEL_PHI_SUM));
    make_phi_particle_sum(fld(KAPSEL_PHI), fld(KAPSEL_PHI_SUM));
    double e0[3];
    int it = dielectric_potential(jikan, tol, maxit, solver, relres, e0);
    maxwell_force(e0);
    add_force_to_fluid(jikan);
    double *u[3] = {U(0), U(1), U(2)};
    calc_f_hydro(fld(KAPSEL_PHI_SUM), u, jikan);
    md_velocity(jikan);
    for (int d = 0; d < 3; d++) reset(UP(d));
    double *up[3] = {UP(0), UP(1), UP(2)};
    make_u_particle_sum(up, fld(KAPSEL_PHI_SUM));
    make_f_particle_dt_sole_add();
    u2zeta_k();
    jikan.time += jikan.dt_fluid;
    jikan.ts += 1;
    return it;
  }

  /* Make_Coulomb_force_x_on_fluid (operate_electrolyte.cxx:90-177), free particles */
  void coulomb_force(const kapsel_time &jikan) {
    double *charge = EB(EB_CHARGE), *pot = EB(EB_POT);
    double *force[3] = {FP(0), FP(1), FP(2)};
    conc_k2charge_field(charge, force[0], force[1]);
    memcpy(pot, charge, F * sizeof(double)); /* A2a_k_out */
    fft_r2c(pot);
    coulomb_potential_k(pot);
    grad_k(pot, force);
    for (int d = 0; d < 3; d++) fft_c2r(force[d]);
    double external[3];
    const double s = ac_factor(jikan);
    for (int d = 0; d < 3; d++) external[d] = EL.e_ext[d] * s;
    for (int d = 0; d < 3; d++)
#pragma omp parallel for
      for (int i = 0; i < NX; i++)
        for (int j = 0; j < NY; j++)
          for (int k = 0; k < NZ; k++) {
            size_t im = ((size_t)i * NY + j) * NZ_ + k;
            double electric_field = -force[d][im];
            electric_field += external[d];
            force[d][im] = charge[im] * electric_field;
          }
  }

  /* Time_evolution_hydro, Electrolyte branch without Dielectric (sp_3d_ns.cxx:24-186) */
  void electrolyte_step(kapsel_time &jikan) {
    nssolute_euler(jikan);
    for (int d = 0; d < 3; d++)
      if (P.fix_cell[d]) uk_dc[d] = 0.0;
    truncate(fld(KAPSEL_ZETA0));
    truncate(fld(KAPSEL_ZETA1));
    zeta_k2u();
    md_position(jikan);
    reset(fld(KAPSEL_PHI));
    reset(fld(KAPSEL_PHI_SUM));
    make_phi_particle_sum(fld(KAPSEL_PHI), fld(KAPSEL_PHI_SUM));
    double factor[2];
    rescale_solute(factor);
    coulomb_force(jikan);
    add_force_to_fluid(jikan);
    double *u[3] = {U(0), U(1), U(2)};
    calc_f_hydro(fld(KAPSEL_PHI_SUM), u, jikan);
    md_velocity(jikan);
    for (int d = 0; d < 3; d++) reset(UP(d));
    double *up[3] = {UP(0), UP(1), UP(2)};
    make_u_particle_sum(up, fld(KAPSEL_PHI_SUM));
    make_f_particle_dt_sole_add();
    u2zeta_k();
    jikan.time += jikan.dt_fluid;
    jikan.ts += 1;
  }

  /* main's initial rigidity projection (sp_3d_ns.cxx:431-447) */
  void init_rigidity() {
    reset(fld(KAPSEL_PHI));
    for (int d = 0; d < 3; d++) reset(UP(d));
    reset(fld(KAPSEL_PHI_SUM));
    make_phi_particle_sum(fld(KAPSEL_PHI), fld(KAPSEL_PHI_SUM));
    double *up[3] = {UP(0), UP(1), UP(2)};
    make_u_particle_sum(up, fld(KAPSEL_PHI_SUM));
    zeta_k2u();
    make_f_particle_dt_sole_add();
    u2zeta_k();
    uk_dc[0] = uk_dc[1] = uk_dc[2] = 0.0;
  }
};

thread_local char g_err[256] = "";

}  // namespace

/* ------------------------------------------------------------------------ */
extern "C" {

const char *oracle_last_error(void) { return g_err; }

void *oracle_create(const kapsel_params *p) {
  if (!p || p->nx < 16 || p->ny < 16 || p->nz < 16) {
    snprintf(g_err, sizeof g_err, "bad params");
    return nullptr;
  }
  for (int n : {p->nx, p->ny, p->nz})
    if (n & (n - 1)) {
      snprintf(g_err, sizeof g_err, "grid not a power of two");
      return nullptr;
    }
  Oracle *o = new Oracle();
  o->init(p);
  return o;
}
void oracle_destroy(void *h) { delete (Oracle *)h; }
int oracle_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
void oracle_set_num_threads(int n) {
#ifdef _OPENMP
  omp_set_num_threads(n);
#endif
}

#define O ((Oracle *)h)
int oracle_get_sekibun(void *h, int *np, int *cells) {
  *np = O->NP_domain;
  if (cells) memcpy(cells, O->sekibun.data(), sizeof(int) * O->sekibun.size());
  return 0;
}
int oracle_set_field(void *h, int f, const double *a) { memcpy(O->fld(f), a, O->F * sizeof(double)); return 0; }
int oracle_get_field(void *h, int f, double *a) { memcpy(a, O->fld(f), O->F * sizeof(double)); return 0; }
int oracle_set_uk_dc(void *h, const double *d) { for (int i = 0; i < 3; i++) O->uk_dc[i] = d[i]; return 0; }
int oracle_get_uk_dc(void *h, double *d) { for (int i = 0; i < 3; i++) d[i] = O->uk_dc[i]; return 0; }
int oracle_set_particles(void *h, const double *s) {
  memcpy(O->ps.data(), s, sizeof(double) * KAPSEL_P_NFIELDS * 3 * O->Np); return 0;
}
int oracle_get_particles(void *h, double *s) {
  memcpy(s, O->ps.data(), sizeof(double) * KAPSEL_P_NFIELDS * 3 * O->Np); return 0;
}
int oracle_fft_r2c(void *h, int f) { O->fft_r2c(O->fld(f)); return 0; }
int oracle_fft_c2r(void *h, int f) { O->fft_c2r(O->fld(f)); return 0; }
int oracle_truncate(void *h, int f) { O->truncate(O->fld(f)); return 0; }
int oracle_zeta_k2u_k(void *h) {
  double *z[2] = {O->fld(KAPSEL_ZETA0), O->fld(KAPSEL_ZETA1)};
  double *u[3] = {O->U(0), O->U(1), O->U(2)};
  O->zeta_k2u_k(z, O->uk_dc, u);
  return 0;
}
int oracle_u_k2zeta_k(void *h) {
  double *z[2] = {O->fld(KAPSEL_ZETA0), O->fld(KAPSEL_ZETA1)};
  double *u[3] = {O->U(0), O->U(1), O->U(2)};
  O->u_k2zeta_k(u, z, O->uk_dc);
  return 0;
}
/* U2advection_k on U0..U2 (real space); result in WORK0/WORK1 (k-space) */
int oracle_u2advection_k(void *h) {
  double *u[3] = {O->U(0), O->U(1), O->U(2)};
  double *adv[2] = {O->fld(KAPSEL_WORK0), O->fld(KAPSEL_WORK1)};
  O->u2advection_k(u, adv);
  return 0;
}
int oracle_solenoidal_u(void *h, int f0) {
  double *u[3] = {O->fld(f0), O->fld(f0 + 1), O->fld(f0 + 2)};
  O->solenoidal_u(u);
  return 0;
}
int oracle_ns_slaved_euler(void *h, const kapsel_time *t) { O->ns_slaved_euler(*t); return 0; }
int oracle_make_phi_particle_sum(void *h) {
  O->reset(O->fld(KAPSEL_PHI));
  O->reset(O->fld(KAPSEL_PHI_SUM));
  O->make_phi_particle_sum(O->fld(KAPSEL_PHI), O->fld(KAPSEL_PHI_SUM));
  return 0;
}
int oracle_make_u_particle_sum(void *h) {
  double *up[3] = {O->UP(0), O->UP(1), O->UP(2)};
  for (int d = 0; d < 3; d++) O->reset(up[d]);
  O->make_u_particle_sum(up, O->fld(KAPSEL_PHI_SUM));
  return 0;
}
int oracle_make_phi_particle(void *h, int f) {
  O->reset(O->fld(f));
  O->make_phi_particle(O->fld(f));
  return 0;
}
int oracle_make_phi_sin(void *h, int f) {
  O->reset(O->fld(f));
  O->make_phi_sin(O->fld(f));
  return 0;
}
int oracle_calc_f_hydro(void *h, const kapsel_time *t) {
  double *u[3] = {O->U(0), O->U(1), O->U(2)};
  O->calc_f_hydro(O->fld(KAPSEL_PHI_SUM), u, *t);
  return 0;
}
int oracle_force(void *h) { O->force(); return 0; }
int oracle_md_position(void *h, const kapsel_time *t) { O->md_position(*t); return 0; }
int oracle_md_velocity(void *h, const kapsel_time *t) { O->md_velocity(*t); return 0; }
int oracle_make_f_particle_dt_sole(void *h) { O->make_f_particle_dt_sole_add(); return 0; }
int oracle_make_dielectric_field(void *h) {
  O->make_dielectric_field(O->fld(KAPSEL_EPSILON), O->fld(KAPSEL_WORK4));
  return 0;
}
int oracle_poisson_apply(void *h, int fx, int fax) {
  O->ax(O->fld(fx), O->fld(fax));
  return 0;
}
/* Ax on unpadded NX*NY*NZ vectors, the reference's Eigen-side signature */
int oracle_ax_unpadded(void *h, const double *x, double *ax) {
  std::vector<double> xp(O->F, 0.0), yp(O->F, 0.0);
  for (int i = 0; i < O->NX; i++)
    for (int j = 0; j < O->NY; j++)
      for (int k = 0; k < O->NZ; k++) xp[((size_t)i * O->NY + j) * O->NZ_ + k] = x[((size_t)i * O->NY + j) * O->NZ + k];
  O->ax(xp.data(), yp.data());
  for (int i = 0; i < O->NX; i++)
    for (int j = 0; j < O->NY; j++)
      for (int k = 0; k < O->NZ; k++) ax[((size_t)i * O->NY + j) * O->NZ + k] = yp[((size_t)i * O->NY + j) * O->NZ_ + k];
  return 0;
}
int oracle_poisson_rhs(void *h, const double *e0) {
  O->build_rhs(O->fld(KAPSEL_RHS), O->fld(KAPSEL_RHO_FREE), O->fld(KAPSEL_EPSILON), e0);
  return 0;
}
/* solver: 0 = CG, 1 = spectral-preconditioned CG, 2 = BiCGSTAB (reference) */
int oracle_poisson_solve(void *h, const double *e0, double tol, int maxit, int solver, int *iters, double *relres) {
  O->build_rhs(O->fld(KAPSEL_RHS), O->fld(KAPSEL_RHO_FREE), O->fld(KAPSEL_EPSILON), e0);
  double *x = O->fld(KAPSEL_POTENTIAL);
  const double *b = O->fld(KAPSEL_RHS);
  int it = solver == 2 ? O->bicgstab_solve(x, b, tol, maxit, relres) : O->cg_solve(x, b, tol, maxit, solver, relres);
  if (it >= maxit) {
    /* the reference's fallback: solve again from a zero initial guess (operate_dielectric.cxx:563-570) */
    O->reset(x);
    it = solver == 2 ? O->bicgstab_solve(x, b, tol, maxit, relres) : O->cg_solve(x, b, tol, maxit, solver, relres);
  }
  *iters = it;
  return 0;
}
int oracle_maxwell_force(void *h, const double *e0) {
  O->maxwell_force(e0);
  return 0;
}
int oracle_add_force_to_fluid(void *h, const kapsel_time *t) {
  O->add_force_to_fluid(*t);
  return 0;
}
int oracle_icep_step(void *h, kapsel_time *t, const double *e0, double tol, int maxit, int solver, int *iters,
                     double *relres) {
  *iters = O->icep_step(*t, e0, tol, maxit, solver, relres);
  return 0;
}
int oracle_init_rigidity(void *h) { O->init_rigidity(); return 0; }
int oracle_set_plant(void *h, int code) { O->plant = code; return 0; }
int oracle_set_thermostat(void *h, double kbt, double alpha_v, double alpha_o, unsigned long seed) {
  O->thermo_on = true;
  O->thermo_kbt = kbt;
  O->thermo_alpha_v = alpha_v;
  O->thermo_alpha_o = alpha_o;
  O->mt.seed((std::mt19937::result_type)(seed & 0xffffffffUL));
  return 0;
}
int oracle_thermostat(void *h, const kapsel_time *t) { O->thermostat(*t); return 0; }
/* n values of the oracle's genrand_real3 stream (tests of the generator) */
int oracle_mt_real3(void *h, int n, double *out) {
  for (int i = 0; i < n; i++) out[i] = O->real3();
  return 0;
}
int oracle_set_orientations(void *h, const double *q) {
  memcpy(O->qs.data(), q, sizeof(double) * 8 * O->Np);
  return 0;
}
int oracle_get_orientations(void *h, double *q) {
  memcpy(q, O->qs.data(), sizeof(double) * 8 * O->Np);
  return 0;
}
int oracle_spm_step(void *h, kapsel_time *t) { O->spm_step(*t); return 0; }

/* electrolyte */
int oracle_set_electrolyte(void *h, const kapsel_electrolyte *el) {
  O->EL = *el;
  O->el_set = true;
  return 0;
}
#define NEED_EL                                                     \
  if (!O->el_set) {                                                 \
    snprintf(g_err, sizeof g_err, "electrolyte parameters not set"); \
    return -3;                                                      \
  }
int oracle_init_rho_ion(void *h) { NEED_EL O->init_rho_ion(); return 0; }
int oracle_get_total_solute(void *h, double *t) { t[0] = O->total_solute[0]; t[1] = O->total_solute[1]; return 0; }
int oracle_set_total_solute(void *h, const double *t) { O->total_solute[0] = t[0]; O->total_solute[1] = t[1]; return 0; }
int oracle_make_phi_qq(void *h, int fphi, int fsurf) {
  NEED_EL
  O->reset(O->fld(fphi));
  O->reset(O->fld(fsurf));
  O->make_phi_qq_particle(O->fld(fphi), O->fld(fsurf));
  return 0;
}
int oracle_make_surface_normal(void *h, int f0) {
  double *n[3] = {O->fld(f0), O->fld(f0 + 1), O->fld(f0 + 2)};
  O->make_surface_normal(n);
  return 0;
}
int oracle_charge_field(void *h, int f) {
  NEED_EL
  O->conc_k2charge_field(O->fld(f), O->EB(Oracle::EB_PHIQQ), O->EB(Oracle::EB_WORK));
  return 0;
}
int oracle_nssolute_euler(void *h, const kapsel_time *t) { NEED_EL O->nssolute_euler(*t); return 0; }
int oracle_rescale_solute(void *h, double *factor) {
  NEED_EL
  double f[2];
  O->rescale_solute(f);
  if (factor) { factor[0] = f[0]; factor[1] = f[1]; }
  return 0;
}
int oracle_coulomb_force(void *h, const kapsel_time *t) { NEED_EL O->coulomb_force(*t); return 0; }
int oracle_free_energy_pb(void *h, const kapsel_time *t, double *fe) { NEED_EL O->free_energy_pb(*t, fe); return 0; }
int oracle_init_rho_ion_pb(void *h, const kapsel_time *t, int max_iter, int *iters, double *fe) {
  NEED_EL
  /* Init_rho_ion with Poisson_Boltzmann, salt (operate_electrolyte.cxx:
   * 820-833): one particle -- the Poisson-Boltzmann fixed point; several --
   * uniform start and steady state; then Count_solute_each */
  if (O->Np == 1) {
    if (O->EL.valency[0] != -O->EL.valency[1]) {
      snprintf(g_err, sizeof g_err, "a z:z salt is needed (valencies %g %g)", O->EL.valency[0], O->EL.valency[1]);
      return -1;
    }
    *iters = O->pb_salt_single(max_iter > 0 ? max_iter : 100000);
    if (fe) O->free_energy_pb(*t, fe);
  } else {
    O->init_rho_ion();
    *iters = O->steadystate_ion_density(*t, max_iter > 0 ? max_iter : 10000, fe);
  }
  O->count_solute_each(O->total_solute, O->fld(KAPSEL_PHI), O->EB(Oracle::EB_WORK));
  return 0;
}
int oracle_electrolyte_step(void *h, kapsel_time *t) { NEED_EL O->electrolyte_step(*t); return 0; }
int oracle_init_rho_ion_dielectric(void *h) { NEED_EL O->init_rho_ion_dielectric(); return 0; }
/* the gradient the Dielectric branch of Rhs_NS_Nernst_Planck uses, into F0..F2 */
int oracle_dielectric_grad(void *h, const kapsel_time *t) {
  NEED_EL
  double *g[3] = {O->FP(0), O->FP(1), O->FP(2)};
  O->dielectric_grad(*t, g);
  return 0;
}
int oracle_dielectric_potential(void *h, const kapsel_time *t, double tol, int maxit, int solver, int *iters,
                                double *relres) {
  NEED_EL
  double e0[3];
  *iters = O->dielectric_potential(*t, tol, maxit, solver, relres, e0);
  return 0;
}
int oracle_icep_electrolyte_step(void *h, kapsel_time *t, double tol, int maxit, int solver, int *iters,
                                 double *relres) {
  NEED_EL
  *iters = O->icep_electrolyte_step(*t, tol, maxit, solver, relres);
  return 0;
}
#undef NEED_EL
#undef O

}  // extern "C"
