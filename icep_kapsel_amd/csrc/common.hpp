// common.hpp -- shared definitions of the gfx950 SPM library: context, error
// handling, complex helpers and launch instrumentation.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <functional>
#include <random>
#include <string>
#include <vector>

#include "../../include/kapsel_gpu.h"

namespace kapsel {

// ---------------------------------------------------------------- errors
void set_error(const char *fmt, ...);
struct Error {
  int code;
};

#define KAPSEL_HIP(call)                                                          \
  do {                                                                            \
    hipError_t e_ = (call);                                                       \
    if (e_ != hipSuccess) {                                                       \
      ::kapsel::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                          __FILE__, __LINE__);                                    \
      throw ::kapsel::Error{KAPSEL_ERR_HIP};                                      \
    }                                                                             \
  } while (0)

#define KAPSEL_LAUNCH_CHECK() KAPSEL_HIP(hipGetLastError())

// ---------------------------------------------------------------- complex
template <typename T>
struct V2 {
  T re, im;
};
template <typename T>
__device__ __forceinline__ V2<T> operator+(V2<T> a, V2<T> b) {
  return {a.re + b.re, a.im + b.im};
}
template <typename T>
__device__ __forceinline__ V2<T> operator-(V2<T> a, V2<T> b) {
  return {a.re - b.re, a.im - b.im};
}
template <typename T>
__device__ __forceinline__ V2<T> cmul(V2<T> a, V2<T> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename T>
__device__ __forceinline__ V2<T> conjv(V2<T> a) {
  return {a.re, -a.im};
}
template <typename T>
__device__ __forceinline__ V2<T> scal(T s, V2<T> a) {
  return {s * a.re, s * a.im};
}

// vector types for 2-element loads/stores
template <typename T>
struct VecT;
template <>
struct VecT<double> {
  using type = double2;
};
template <>
struct VecT<float> {
  using type = float2;
};
template <typename T>
__device__ __forceinline__ V2<T> ld2(const T *p) {
  auto v = *reinterpret_cast<const typename VecT<T>::type *>(p);
  return {v.x, v.y};
}
template <typename T>
__device__ __forceinline__ void st2(T *p, V2<T> a) {
  typename VecT<T>::type v;
  v.x = a.re;
  v.y = a.im;
  *reinterpret_cast<typename VecT<T>::type *>(p) = v;
}

// ---------------------------------------------------------------- grid
// Scalars every kernel needs, passed by value.
struct Grid {
  int nx, ny, nz;     // NX, NY, NZ
  int nz_;            // NZ_ = 2*(NZ/2+1), row pitch in reals
  int h;              // NZ/2+1, row pitch in complex values
  int hnx, hny;       // NX/2, NY/2
  int trn_x, trn_y, trn_z;
  double dx;
  double wave_x, wave_y, wave_z;
  double lx, ly, lz;  // box lengths L[d]
  size_t f;           // reals per padded field = NX*NY*NZ_
  size_t nc;          // complex values per field = NX*NY*h
  size_t nr;          // real grid points = NX*NY*NZ
};

// Index of complex value c of the half spectrum (one thread per value): the
// row (i, j), kz, and the signed wavenumbers of Calc_K*_Ooura
// (fft_wrapper.h:761-773), +N/2 kept for the Nyquist index
struct KIdx {
  int i, j, kz;
  int kx, ky;
};
__device__ __forceinline__ KIdx kidx(const Grid &g, size_t c) {
  KIdx r;
  const unsigned plane = (unsigned)g.ny * (unsigned)g.h;
  unsigned cc = (unsigned)c;
  r.i = cc / plane;
  unsigned rem = cc - r.i * plane;
  r.j = rem / (unsigned)g.h;
  r.kz = rem - r.j * g.h;
  r.kx = r.i > g.hnx ? r.i - g.nx : r.i;
  r.ky = r.j > g.hny ? r.j - g.ny : r.j;
  return r;
}
// Truncate_two_third_rule_ooura (fft_wrapper.h:797-839): the cut set
__device__ __forceinline__ bool is_cut(const Grid &g, const KIdx &k) {
  return (k.kz >= g.trn_z) || (k.j >= g.trn_y && k.j <= g.ny - g.trn_y) ||
         (k.i >= g.trn_x && k.i <= g.nx - g.trn_x);
}
// Real grid point r (k < NZ) -> its coordinates and padded index
struct RI {
  int i, j, k;
  size_t im;
};
__device__ __forceinline__ RI ridx(const Grid &g, size_t r) {
  RI o;
  size_t row = r / (unsigned)g.nz;
  o.k = (int)(r - row * g.nz);
  o.j = (int)(row % (unsigned)g.ny);
  o.i = (int)(row / (unsigned)g.ny);
  o.im = row * g.nz_ + o.k;
  return o;
}

// ---------------------------------------------------------------- timing
struct TimingRecord {
  int cls;
  double bytes;
  hipEvent_t start, stop;
};

// ---------------------------------------------------------------- context
struct Twiddles {
  void *x = nullptr, *y = nullptr, *zh = nullptr, *zrot = nullptr;  // V2<T> tables
  void *shx = nullptr, *shy = nullptr, *shz = nullptr;  // staggered shifts exp(i dx k/2)
};

struct Pnet;  // learned warm start (pnet.hip)

template <typename T>
constexpr int prec_of() {
  return sizeof(T) == 8 ? KAPSEL_FP64 : KAPSEL_FP32;
}

}  // namespace kapsel

struct kapsel_ctx {
  kapsel_params P;
  kapsel::Grid g;
  int prec;  // KAPSEL_FP64 or KAPSEL_FP32
  size_t elem;  // sizeof(T)
  hipStream_t stream = nullptr;
  // a second stream (default priority) for the ICEP step's Navier-Stokes
  // chain, which runs beside the Poisson chain on `stream`; forked and joined
  // with the two events
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  bool concurrent = true;         // kapsel_set_concurrency
  void *adv_w[2] = {};            // the advection's two extra product fields (WORK0, WORK1 by default)
  void *ns_work[2] = {};          // their stand-ins while the NS chain runs beside the solve
  void *adv_x[3] = {};            // the three other product fields where the products are formed
                                  // inside the forward transform (allocated on first use)
  bool lazy_outputs = false;      // an FP32 solve leaves POTENTIAL/RHS to materialize_outputs
  bool outputs_stale = false;     // POTENTIAL/RHS wait for materialize_outputs
  // queued by an FP32 solve with each potential it is about to check (while
  // the host waits for the check); solve_hook_done: queued for the potential
  // the solve returned
  const std::function<void()> *solve_hook = nullptr;
  bool solve_hook_done = false;
  int device = 0;
  void *fields[KAPSEL_NFIELDS] = {};
  double *d_uk_dc = nullptr;      // 3 doubles (device)
  double *d_ps = nullptr;         // particle state (device), SoA
  double *d_q = nullptr;          // orientations q, q_old (device): [c*np + n], c < 8
  int *d_sekibun = nullptr;       // np_domain * 3 ints
  int np_domain = 0;
  double *d_gpart = nullptr;      // gather partial sums: 12 doubles per (particle, slice)
  std::vector<int> sekibun;       // host copy
  kapsel::Twiddles tw[2];          // [KAPSEL_FP64], [KAPSEL_FP32]
  // Poisson solver storage, allocated on first use: operator scratch of the
  // context precision, and FP64 shadow fields for the iterative refinement
  // of an FP32 context
  void *pscratch[8] = {};          // [0..3] operator scratch, [4] refinement correction, [5..7] face eps
  double *p64[10] = {};           // eps (then its spectrum), x (k), r (k), b (k), 3 operator fields, 3 face eps
  // FP32 context: p64[1] holds the FP64 potential of the last solve; it seeds
  // the next solve unless the FP32 POTENTIAL field was written since.
  // x64_prev holds the solve before it (when prev_valid), so the next start
  // can be the linear extrapolation 2 x_n - x_{n-1} in time
  bool pot64_valid = false;
  // p64[0] and p64[1] hold the spectra of EPSILON and POTENTIAL as the last
  // FP32-context solve left them (cleared when a solve starts)
  bool spec64_fresh = false;
  bool refine64 = false;  // inside the FP64 part of an FP32 solve: launches time as KAPSEL_K_REFINE64
  bool prev_valid = false;
  // RHO_FREE has been written (through the API or by the dielectric
  // electrolyte's free charge) since the context was created zeroed; until
  // then an FP32-context solve takes rho = 0 without transforming it
  bool rho_free_set = false;
  double *x64_prev = nullptr;
  double *d_partial = nullptr;    // block partial sums (reductions)
  double *d_scal = nullptr;       // CG scalars
  // pinned host scalars that kernels store into (mapped): [0..1] ||r||^2 of
  // the k-space CG iterations by parity, [2] ||b||^2 and [3] ||r||^2 of the
  // refinement, [4] and [5] the sequence numbers (counts since the context's
  // start) of the last CG iteration and refinement residual stored
  double *h_slots = nullptr;
  double *h_slots_dev = nullptr;  // the same slots as the device sees them
  unsigned long long *d_count = nullptr;  // [0], [1]: the device side of the two counts
  unsigned long long seq_cg = 0, seq_res = 0;  // how many of each were queued
  hipEvent_t slot_ev[2] = {};     // real-space CG (FP64 context): copies and events
  int n_partial = 0;
  kapsel::Pnet *pnet = nullptr;   // learned warm-start network, on first use
  // electrolyte (electrolyte.hip): parameters, the conserved ion amounts, and
  // scratch fields allocated by kapsel_set_electrolyte: surface normal (3),
  // profile sum and raw surface charge of Make_phi_qq_particle (2), the two
  // ion concentrations in real space (2), the charge density to transform and
  // the real charge density kept for the force (2), the real velocity at the
  // end of a step (3)
  // thermostat (particles.hip): parameters, the MT19937 stream (host), pinned
  // double-buffered deviates and their device copy
  double thermo_kbt = 0.0, thermo_alpha_v = 0.0, thermo_alpha_o = 0.0;
  std::mt19937 mt{5489u};
  double *h_noise[2] = {};
  hipEvent_t noise_ev[2] = {};
  double *d_noise = nullptr;
  int noise_slot = 0;
  kapsel_electrolyte el{};
  bool el_set = false;
  double el_total[2] = {0.0, 0.0};
  void *el_f[12] = {};
  // Every entry point that can change state counts api_gen up.  An
  // electrolyte step sets el_reuse_gen to the count its successor will see
  // if nothing runs in between; that successor then reuses what the step
  // left: the charge spectrum, the real ion fields and the real velocity.
  unsigned long long api_gen = 0, el_reuse_gen = ~0ull;
  // An FP32 step that ends with U holding the cut, solenoidal velocity
  // spectrum whose curl it stored as zeta sets u_spec_gen the same way; a
  // slaved step that comes right after takes that spectrum for Zeta_k2u_k
  // of zeta (equal but for rounding) instead of forming it again.
  unsigned long long u_spec_gen = ~0ull;
  double *el_partial = nullptr;   // block partial sums of the electrolyte reductions
  double *el_scal = nullptr;      // device scalars (electrolyte.hip)
  // when set, the Poisson right-hand side is built from rho - rho_bg[0]
  // (a device scalar): the uniform background that neutralises the ions'
  // free charge in the variable permittivity
  const double *rho_bg = nullptr;
  // derived physics constants (Set_global_parameters, input.cxx:287-430)
  double nu, irho, hxi, mass, imass, moi, imoi, a_r_cutoff;
  // kernel timing
  int timing_mask = 0;
  int timing_every = 1;  // time one launch in timing_every of a selected class
  unsigned timing_seen[KAPSEL_K_NCLASSES] = {};
  std::vector<kapsel::TimingRecord> timing_pending;
  std::vector<hipEvent_t> event_pool;
  double timing_ms[KAPSEL_K_NCLASSES] = {};
  double timing_bytes[KAPSEL_K_NCLASSES] = {};
  int timing_count[KAPSEL_K_NCLASSES] = {};
};

namespace kapsel {

// Brackets one kernel launch with HIP events when its class is being timed;
// `bytes` is the launch's algorithmic HBM traffic (each input read once,
// each output written once).
// kernel_events: the scope launches one kernel with hipExtLaunchKernelGGL
// and passes it start and stop (null when this launch is not sampled), so
// that they hold the kernel's own start and end; otherwise the events are
// recorded on the stream around the scope's launches (their interval then
// holds the dispatch gap before the first kernel as well)
struct LaunchTimer {
  kapsel_ctx *ctx;
  int cls;
  double bytes;
  hipEvent_t start = nullptr, stop = nullptr;
  bool kernel_events = false;
  LaunchTimer(kapsel_ctx *c, int k, double b, bool kernel_events = false);
  ~LaunchTimer() noexcept(false);
};

inline unsigned blocks_for(size_t n, unsigned threads) {
  return (unsigned)((n + threads - 1) / threads);
}

// Particle_domain (profile.cxx:25-74): offsets (i, j, k), |i|,|j|,|k| <= nh,
// in i-j-k order, kept when sqrt((i dx)^2 + (j dx)^2 + (k dx)^2) < max_radius
// = RADIUS + XI/2 + 1.7321 DX, nh = A + A_XI + 2.
struct SekibunGeom {
  int nh;
  double dx, max_radius;
};
inline SekibunGeom sekibun_geom(double radius, double xi, double dx) {
  const int A = (int)(radius / dx + 1e-9), A_XI = (int)(xi / dx + 1e-9);
  return SekibunGeom{A + A_XI + 2, dx, radius + 0.5 * xi + 1.7321 * dx};
}
inline bool sekibun_inside(const SekibunGeom &s, int i, int j, int k) {
  const double x0 = (double)i * s.dx, x1 = (double)j * s.dx, x2 = (double)k * s.dx;
  return sqrt(x0 * x0 + x1 * x1 + x2 * x2) < s.max_radius;
}

// Slices of the Sekibun table per particle in the hydrodynamic gather: about
// 1024 wavefronts in all (four per CU), at least 128 table points each.
inline int gather_slices(int np, int np_domain) {
  if (np <= 0) return 1;
  int s = (1024 + np - 1) / np;
  const int smax = (np_domain + 127) / 128;
  if (s > smax) s = smax;
  return s < 1 ? 1 : s;
}

// ---- launchers (defined in the .hip files) ----------------------------------
// FFT (fft.hip)
void init_twiddles(kapsel_ctx *ctx);
void free_twiddles(kapsel_ctx *ctx);
// nf fields transformed together (1..6), of the context precision or of T;
// c2r with src: the spectra in src (left unchanged) go to fields
void fft_r2c(kapsel_ctx *ctx, void *const *fields, int nf);
void fft_c2r(kapsel_ctx *ctx, void *const *fields, int nf, void *const *src = nullptr);
template <typename T>
void fft_r2c_t(kapsel_ctx *ctx, void *const *fields, int nf);
template <typename T>
void fft_c2r_t(kapsel_ctx *ctx, void *const *fields, int nf, void *const *src = nullptr);
// fft_c2r_t after its first (x, inverse) pass
template <typename T>
void fft_c2r_yz_t(kapsel_ctx *ctx, void *const *fields, int nf);
// out_f = a_f + s (b_f + e_f) c_f transformed forward, the product formed in
// the first pass (a or c NULL: left out; bit f of dsq: (b_f - c_f)(b_f + c_f)
// instead of the product; e NULL: e_f = 0).  out_f may be one of field f's
// own inputs, no other field's.  false: nothing queued (no pass for NZ)
template <typename T>
bool fft_r2c_pro_t(kapsel_ctx *ctx, void *const *out, void *const *a, void *const *b, void *const *c, double s,
                   int nf, unsigned dsq = 0, const double *e = nullptr);
// whether the z and y transforms of the context's grid run as one plane pass
inline bool plane_ok_grid(const Grid &g) { return g.ny == g.nz && (g.nz == 32 || g.nz == 64 || g.nz == 128); }
// g_d <- r2c(face_d * c2r(g_d)) for the three staggered-gradient spectra
// (face_d real fields, 0 in the row padding): the inverse and forward z (and,
// where a plane fits in LDS, y) transforms and the product fused in one pass
template <typename T>
void fft_face_t(kapsel_ctx *ctx, void *const *g3, void *const *face3);
// fft_face_t without its first (x inverse) and last (x forward) passes
template <typename T>
void fft_face_mid_t(kapsel_ctx *ctx, void *const *g3, void *const *face3);
// the operator's first pass on a CG direction: p = zscale z + sc[beta_slot]
// p, then g_d = x-inverse of (i k_d e^{i k_d dx/2} p); beta_slot < 0: the
// gradient of z, p untouched, or with sub of p = z - sub (stored); false: NX
// has no fused pass
template <typename T>
bool fft_x_grad_t(kapsel_ctx *ctx, const T *z, T *p, const double *sc, int beta_slot, void *const *g3,
                  const float *sub = nullptr, double zscale = 1.0);
// the staggered gradients of two FP64 half spectra a, b as six real FP32
// fields (a's three, then b's); false where NX has no fused x pass (nothing
// queued)
bool fft_grad2_f64_to_f32(kapsel_ctx *ctx, const double *a, const double *b, void *const *out6);
// the operator's last pass: ap = staggered divergence of the x-forward of
// g_d; one Parseval-weighted <p, ap> partial per workgroup; returns the
// number of partials (0: NX has no fused pass)
template <typename T>
int fft_x_div_t(kapsel_ctx *ctx, void *const *g3, const T *p, T *ap, double *partial, int max_partials);
// fft_x_div_t's residual form: r = (divergence) - b, partials of |r|^2
template <typename T>
int fft_x_res_t(kapsel_ctx *ctx, void *const *g3, const T *b, T *r, double *partial, int max_partials);

// k-space and real-space pointwise (kspace.hip)
void truncate(kapsel_ctx *ctx, void *const *fields, int nf);
// out3: the velocity spectrum goes there instead of U0..U2
void zeta_k2u_k(kapsel_ctx *ctx, int fix_dc, void *const *out3 = nullptr);
void u_k2zeta_k(kapsel_ctx *ctx);
// the 2/3 cut and Solenoidal_uk of the spectrum f3, then U_k2zeta_k of
// a3 + P f3, which is stored in U0..U2 (a3 and f3 unchanged)
void sole_u2zeta_sum(kapsel_ctx *ctx, void *const *a3, void *const *f3);
// o3 = a3 + s P f3 (the projection after the 2/3 cut), a3 and f3 unchanged
void sole_add(kapsel_ctx *ctx, void *const *a3, void *const *f3, void *const *o3, double s);
void advection_products(kapsel_ctx *ctx);
// in5: the five product spectra (default U0..U2 and adv_w); the advection
// term is stored over the last two unless store_a is 0
void advection_combine_slaved(kapsel_ctx *ctx, double dt_fluid, void *const *in5 = nullptr, int store_a = 1);
// advection_combine_slaved (the term not stored) then zeta_k2u_k after the
// 2/3 cut with the FIX_CELL mask fix_mask, in one pass; out3: where the
// velocity spectrum goes (default U0..U2)
void slaved_zeta2u(kapsel_ctx *ctx, double dt_fluid, void *const *in5, int fix_mask, void *const *out3 = nullptr);
// out3: the projected spectrum goes there and f3 is left unchanged
void solenoidal_uk(kapsel_ctx *ctx, void *const *f3, void *const *out3 = nullptr);
void make_f_nonsole(kapsel_ctx *ctx);
// u += f over the real points; keep3: also store the sum there
void add_f_particle(kapsel_ctx *ctx, void *const *keep3 = nullptr);
void copy_field(kapsel_ctx *ctx, void *dst, const void *src);
void copy_convert_in(kapsel_ctx *ctx, void *dst, const double *src_dev);
void copy_convert_out(kapsel_ctx *ctx, double *dst_dev, const void *src);

// particles (particles.hip)
void md_position(kapsel_ctx *ctx, const kapsel_time &t);
void make_phi_particle_sum(kapsel_ctx *ctx);
void make_u_particle_sum(kapsel_ctx *ctx);
void make_phi_particle(kapsel_ctx *ctx, void *field);
void make_phi_sin(kapsel_ctx *ctx, void *field);
template <typename T>
void make_dielectric_field_t(kapsel_ctx *ctx, T *eps, T *phi_sin);

// dielectric Poisson (poisson.hip)
void dielectric_field(kapsel_ctx *ctx);
void poisson_apply(kapsel_ctx *ctx, const void *x, void *ax);
void poisson_rhs(kapsel_ctx *ctx, const double e0[3]);
void poisson_solve(kapsel_ctx *ctx, const double e0[3], double tol, int maxit, int precond, int *iters,
                   double *relres);
void poisson_free(kapsel_ctx *ctx);
void materialize_outputs(kapsel_ctx *ctx);
double galerkin_scale_potential(kapsel_ctx *ctx, const double e0[3]);
// after_solve: called right after poisson_solve, whose FP64 spectra of eps
// and the potential may then stand in for the transforms of the fields
void maxwell_force(kapsel_ctx *ctx, const double e0[3], bool after_solve = false);
// the Maxwell force of an FP32 context from the FP64 spectra its solve holds
// (the after_solve form), also while the solve runs
void maxwell_force_from_spectra(kapsel_ctx *ctx, const double e0[3]);
void add_force_to_u(kapsel_ctx *ctx, double s);

// learned warm start (pnet.hip)
void pnet_init(kapsel_ctx *ctx);
void pnet_free(kapsel_ctx *ctx);
long pnet_param_count(kapsel_ctx *ctx);
void pnet_set_params(kapsel_ctx *ctx, const float *host);
void pnet_get_params(kapsel_ctx *ctx, float *host);
double pnet_loss_grad(kapsel_ctx *ctx, double beta, const double e[3], double scale, float *grad_host);
double pnet_train_step(kapsel_ctx *ctx, double beta, const double e[3], double scale, double lr);
void pnet_warm_start(kapsel_ctx *ctx, double beta, const double e0[3], double scale);
void calc_f_hydro(kapsel_ctx *ctx, const kapsel_time &t);
void force(kapsel_ctx *ctx);
void md_velocity_integrate(kapsel_ctx *ctx, const kapsel_time &t);
// calc_f_hydro, force and md_velocity_integrate in two launches (the gather's
// slices, then one workgroup per particle for the rest)
void hydro_force_velocity(kapsel_ctx *ctx, const kapsel_time &t);
// Add_random_force_thermostat (fluct.cxx:9-40)
void thermostat(kapsel_ctx *ctx, const kapsel_time &t);
void thermostat_free(kapsel_ctx *ctx);
// Make_phi_qq_particle before its normalisation: unclamped profile sum into
// phi, sum of Q_e DPhi_compact_sin into surface (particles.hip)
void make_phi_qq_raw(kapsel_ctx *ctx, void *phi, void *surface, double q_e);
// Make_surface_normal into n3[0..2] (particles.hip)
void make_surface_normal(kapsel_ctx *ctx, void *const *n3);
// NSsolute_solver_Euler's vorticity update: U2advection_k combination and
// the forward Euler step with the viscous term on the kept modes (kspace.hip)
void advection_combine_euler(kapsel_ctx *ctx, double dt_fluid);

// restart files (restart.hip), host only
void restart_write(const char *path, const char *template_path, const kapsel_restart_data *d);
void restart_read(const char *path, kapsel_restart_data *d);

// electrolyte (electrolyte.hip)
void electrolyte_set(kapsel_ctx *ctx, const kapsel_electrolyte &el);
void electrolyte_free(kapsel_ctx *ctx);
void init_rho_ion(kapsel_ctx *ctx);
void make_phi_qq(kapsel_ctx *ctx, void *phi, void *surface);
void charge_field(kapsel_ctx *ctx, void *out);
void nssolute_euler(kapsel_ctx *ctx, const kapsel_time &t);
// NSsolute_solver_Euler inside an electrolyte step; reuse: the previous step
// ended right before (see api_gen), so its charge spectrum, real ion fields
// and real velocity are current
void nssolute_step(kapsel_ctx *ctx, const kapsel_time &t, bool reuse);
// where the electrolyte step keeps the real velocity (3 fields)
void *const *electrolyte_u_keep(kapsel_ctx *ctx);
void rescale_solute(kapsel_ctx *ctx, double *factor_host);
void coulomb_force(kapsel_ctx *ctx, const kapsel_time &t);
// Rescale_solute right after the profile scatter (PHI_SUM current) and the
// Coulomb force of the rescaled ions into F0..F2 (spectrum: its spectrum)
void rescale_and_coulomb_force(kapsel_ctx *ctx, const kapsel_time &t, bool spectrum = false);
// Count_solute_each (solute_rhs.h:143-156) into the conserved totals
void recount_total_solute(kapsel_ctx *ctx);
// Init_rho_ion with Poisson_Boltzmann, salt, several particles; returns the
// relaxation steps taken; fe: the final free energy {total, ideal, electrostatic}
int init_rho_ion_pb(kapsel_ctx *ctx, const kapsel_time &t, int max_iter, double *fe);
// Calc_free_energy_PB, non-Dielectric branch
void free_energy_pb(kapsel_ctx *ctx, const kapsel_time &t, double *fe);
// Init_rho_ion_dielectric (operate_electrolyte.cxx:912-979)
void init_rho_ion_dielectric(kapsel_ctx *ctx);
// the free charge of the Maxwell force into RHO_FREE (WORK4 holds the
// compact-sine profile of Make_dielectric_field) and its mean into a device
// scalar; returns that scalar's address
const double *dielectric_free_charge(kapsel_ctx *ctx);

}  // namespace kapsel
