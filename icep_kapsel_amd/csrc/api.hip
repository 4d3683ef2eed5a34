// api.hip -- the C ABI of include/kapsel_gpu.h: context lifetime, host <->
// device copies, and the reference's call sequences built from the kernels
// in fft.hip, kspace.hip and particles.hip.
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <exception>
#include <new>
#include <vector>

#include "common.hpp"

namespace kapsel {

namespace {
thread_local char g_err[512] = "";
}

void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
}

static hipEvent_t take_event(kapsel_ctx *ctx) {
  if (!ctx->event_pool.empty()) {
    hipEvent_t e = ctx->event_pool.back();
    ctx->event_pool.pop_back();
    return e;
  }
  hipEvent_t e;
  KAPSEL_HIP(hipEventCreate(&e));
  return e;
}

LaunchTimer::LaunchTimer(kapsel_ctx *c, int k, double b, bool kev)
    : ctx(c), cls(c->refine64 ? KAPSEL_K_REFINE64 : k), bytes(b), kernel_events(kev) {
  if ((ctx->timing_mask & (1 << cls)) && ctx->timing_seen[cls]++ % (unsigned)ctx->timing_every == 0) {
    start = take_event(ctx);
    stop = take_event(ctx);
    if (!kernel_events) KAPSEL_HIP(hipEventRecord(start, ctx->stream));
  }
}

LaunchTimer::~LaunchTimer() noexcept(false) {
  if (stop) {
    if (!kernel_events) KAPSEL_HIP(hipEventRecord(stop, ctx->stream));
    ctx->timing_pending.push_back(TimingRecord{cls, bytes, start, stop});
  }
}

static void harvest_timing(kapsel_ctx *ctx) {
  if (ctx->timing_pending.empty()) return;
  KAPSEL_HIP(hipStreamSynchronize(ctx->side));
  KAPSEL_HIP(hipStreamSynchronize(ctx->stream));
  for (const TimingRecord &r : ctx->timing_pending) {
    float ms = 0.f;
    KAPSEL_HIP(hipEventElapsedTime(&ms, r.start, r.stop));
    ctx->timing_ms[r.cls] += ms;
    ctx->timing_bytes[r.cls] += r.bytes;
    ctx->timing_count[r.cls] += 1;
    ctx->event_pool.push_back(r.start);
    ctx->event_pool.push_back(r.stop);
  }
  ctx->timing_pending.clear();
}

// Particle_domain (profile.cxx:25-74): every integer offset within
// RADIUS + XI/2 + 1.7321 DX, in the reference's i, j, k loop order
static void build_sekibun(kapsel_ctx *ctx) {
  const SekibunGeom sg = sekibun_geom(ctx->P.radius, ctx->P.xi, ctx->g.dx);
  const int nh = sg.nh;
  ctx->sekibun.clear();
  for (int i = -nh; i <= nh; i++)
    for (int j = -nh; j <= nh; j++)
      for (int k = -nh; k <= nh; k++)
        if (sekibun_inside(sg, i, j, k)) {
          ctx->sekibun.push_back(i);
          ctx->sekibun.push_back(j);
          ctx->sekibun.push_back(k);
        }
  ctx->np_domain = (int)ctx->sekibun.size() / 3;
  if (ctx->P.n_particles > 0 && (2 * nh + 1 > ctx->g.nx || 2 * nh + 1 > ctx->g.ny || 2 * nh + 1 > ctx->g.nz)) {
    set_error("Sekibun cell (%d points across) wider than the grid", 2 * nh + 1);
    throw Error{KAPSEL_ERR_ARG};
  }
  KAPSEL_HIP(hipMalloc(&ctx->d_sekibun, sizeof(int) * ctx->sekibun.size()));
  KAPSEL_HIP(hipMemcpy(ctx->d_sekibun, ctx->sekibun.data(), sizeof(int) * ctx->sekibun.size(),
                       hipMemcpyHostToDevice));
}

static bool pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

static void check_field(int f) {
  if (f < 0 || f >= KAPSEL_NFIELDS) {
    set_error("field id %d out of range", f);
    throw Error{KAPSEL_ERR_ARG};
  }
}

// a write to the FP32 potential through the API supersedes its FP64 shadow;
// the free charge is no longer known to be zero once written
static void written(kapsel_ctx *ctx, int f) {
  if (f == KAPSEL_POTENTIAL) ctx->pot64_valid = false;
  if (f == KAPSEL_RHO_FREE) ctx->rho_free_set = true;
}

static void destroy(kapsel_ctx *ctx) {
  if (!ctx) return;
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  if (ctx->side) (void)hipStreamSynchronize(ctx->side);
  for (void *&p : ctx->fields)
    if (p) (void)hipFree(p);
  for (void *&p : ctx->ns_work)
    if (p) (void)hipFree(p);
  for (void *&p : ctx->adv_x)
    if (p) (void)hipFree(p);
  if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
  if (ctx->ev_join) (void)hipEventDestroy(ctx->ev_join);
  if (ctx->side) (void)hipStreamDestroy(ctx->side);
  if (ctx->d_uk_dc) (void)hipFree(ctx->d_uk_dc);
  if (ctx->d_ps) (void)hipFree(ctx->d_ps);
  if (ctx->d_q) (void)hipFree(ctx->d_q);
  if (ctx->d_sekibun) (void)hipFree(ctx->d_sekibun);
  if (ctx->d_gpart) (void)hipFree(ctx->d_gpart);
  free_twiddles(ctx);
  poisson_free(ctx);
  pnet_free(ctx);
  electrolyte_free(ctx);
  thermostat_free(ctx);
  for (auto &r : ctx->timing_pending) {
    (void)hipEventDestroy(r.start);
    (void)hipEventDestroy(r.stop);
  }
  for (hipEvent_t e : ctx->event_pool) (void)hipEventDestroy(e);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

static void create(const kapsel_params *p, kapsel_ctx **out) {
  if (!p || !out) {
    set_error("null argument");
    throw Error{KAPSEL_ERR_ARG};
  }
  for (int n : {p->nx, p->ny, p->nz})
    if (!pow2(n) || n < 16 || n > 256) {
      set_error("grid size %d: must be a power of two in [16, 256]", n);
      throw Error{KAPSEL_ERR_ARG};
    }
  if (!(p->dx > 0) || !(p->rho > 0) || !(p->eta > 0) || p->n_particles < 0 ||
      (p->precision != KAPSEL_FP64 && p->precision != KAPSEL_FP32)) {
    set_error("invalid physical parameters or precision");
    throw Error{KAPSEL_ERR_ARG};
  }
  if (p->n_particles > 0 && (!(p->radius > 0) || !(p->xi > 0))) {
    set_error("particles need radius > 0 and xi > 0");
    throw Error{KAPSEL_ERR_ARG};
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    set_error("no HIP device");
    throw Error{KAPSEL_ERR_NODEV};
  }
  int dev = 0;
  KAPSEL_HIP(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  KAPSEL_HIP(hipGetDeviceProperties(&prop, dev));
  if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
    set_error("device %d is %s; this library is built for gfx950 only", dev, prop.gcnArchName);
    throw Error{KAPSEL_ERR_NODEV};
  }
  kapsel_ctx *ctx = new kapsel_ctx();
  try {
    ctx->P = *p;
    ctx->device = dev;
    ctx->prec = p->precision;
    ctx->elem = p->precision == KAPSEL_FP64 ? 8 : 4;
    Grid &g = ctx->g;
    g.nx = p->nx;
    g.ny = p->ny;
    g.nz = p->nz;
    g.h = g.nz / 2 + 1;  // HNZ_ (input.cxx:305)
    g.nz_ = 2 * g.h;     // NZ_ (input.cxx:306)
    g.hnx = g.nx / 2;
    g.hny = g.ny / 2;
    g.trn_x = (g.nx + 2) / 3;  // TRNs (input.cxx:310)
    g.trn_y = (g.ny + 2) / 3;
    g.trn_z = (g.nz + 2) / 3;
    g.dx = p->dx;
    g.lx = g.nx * g.dx;
    g.ly = g.ny * g.dx;
    g.lz = g.nz * g.dx;
    g.wave_x = 2 * M_PI / g.lx;  // WAVE_X = PI2/LX (input.cxx:326)
    g.wave_y = 2 * M_PI / g.ly;
    g.wave_z = 2 * M_PI / g.lz;
    g.f = (size_t)g.nx * g.ny * g.nz_;
    g.nc = (size_t)g.nx * g.ny * g.h;
    g.nr = (size_t)g.nx * g.ny * g.nz;
    // Set_global_parameters (input.cxx:287-296, 300-301, 404) and A_R_cutoff (input.cxx:1530-1547)
    ctx->irho = 1. / p->rho;
    ctx->nu = p->eta * ctx->irho;
    ctx->hxi = p->xi * 0.5;
    ctx->mass = p->mass_ratio * p->rho * 4. / 3. * M_PI * p->radius * p->radius * p->radius;
    ctx->imass = ctx->mass > 0 ? 1. / ctx->mass : 0.0;
    ctx->moi = 2. / 5. * ctx->mass * p->radius * p->radius;
    ctx->imoi = ctx->moi > 0 ? 1. / ctx->moi : 0.0;
    if (p->lj_truncate > 0) {
      ctx->a_r_cutoff = p->lj_powers == 0 ? pow(2., 1. / 6.) : p->lj_powers == 1 ? pow(2., 1. / 12.)
                                                                                  : pow(2., 1. / 18.);
    } else if (p->lj_truncate == 0) {
      int nmin = g.nx < g.ny ? g.nx : g.ny;
      nmin = nmin < g.nz ? nmin : g.nz;
      ctx->a_r_cutoff = fmin(nmin * g.dx * .5 / p->lj_sigma, 2.5);
    } else {
      ctx->a_r_cutoff = 0.;
    }
    KAPSEL_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    KAPSEL_HIP(hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking));
    KAPSEL_HIP(hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming));
    KAPSEL_HIP(hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming));
    for (int f = 0; f < KAPSEL_NFIELDS; f++) {
      KAPSEL_HIP(hipMalloc(&ctx->fields[f], g.f * ctx->elem));
      KAPSEL_HIP(hipMemsetAsync(ctx->fields[f], 0, g.f * ctx->elem, ctx->stream));
    }
    ctx->adv_w[0] = ctx->fields[KAPSEL_WORK0];
    ctx->adv_w[1] = ctx->fields[KAPSEL_WORK1];
    KAPSEL_HIP(hipMalloc(&ctx->d_uk_dc, 3 * sizeof(double)));
    KAPSEL_HIP(hipMemsetAsync(ctx->d_uk_dc, 0, 3 * sizeof(double), ctx->stream));
    size_t np = p->n_particles > 0 ? (size_t)p->n_particles : 1;
    KAPSEL_HIP(hipMalloc(&ctx->d_ps, sizeof(double) * KAPSEL_P_NFIELDS * 3 * np));
    KAPSEL_HIP(hipMemsetAsync(ctx->d_ps, 0, sizeof(double) * KAPSEL_P_NFIELDS * 3 * np, ctx->stream));
    {
      // orientations: q = q_old = identity (1, 0, 0, 0)
      std::vector<double> q(8 * np, 0.0);
      for (size_t i = 0; i < np; i++) q[i] = q[4 * np + i] = 1.0;
      KAPSEL_HIP(hipMalloc(&ctx->d_q, sizeof(double) * 8 * np));
      KAPSEL_HIP(hipMemcpyAsync(ctx->d_q, q.data(), sizeof(double) * 8 * np, hipMemcpyHostToDevice, ctx->stream));
      KAPSEL_HIP(hipStreamSynchronize(ctx->stream));
    }
    build_sekibun(ctx);
    KAPSEL_HIP(hipMalloc(&ctx->d_gpart, sizeof(double) * 12 * np *
                                            (size_t)gather_slices((int)np, ctx->np_domain)));
    init_twiddles(ctx);
    KAPSEL_HIP(hipStreamSynchronize(ctx->stream));
  } catch (...) {
    destroy(ctx);
    throw;
  }
  *out = ctx;
}

// entry points that only read state
static void check_ctx_ro(kapsel_ctx *ctx) {
  if (!ctx) {
    set_error("null context");
    throw Error{KAPSEL_ERR_ARG};
  }
}

// every other entry point: may change state
static void check_ctx(kapsel_ctx *ctx) {
  check_ctx_ro(ctx);
  ctx->api_gen++;
  // the FP32 POTENTIAL/RHS an ICEP step left to be made: every entry point
  // but the ICEP step itself may read or overwrite them
  materialize_outputs(ctx);
}

static void check_time(const kapsel_time *t) {
  if (!t || !(t->dt_fluid > 0)) {
    set_error("jikan must be non-null with dt_fluid > 0");
    throw Error{KAPSEL_ERR_ARG};
  }
}

// staging of a double host field into a device field of type T
static void upload(kapsel_ctx *ctx, void *dst, const double *host) {
  if (ctx->prec == KAPSEL_FP64) {
    KAPSEL_HIP(hipMemcpyAsync(dst, host, ctx->g.f * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  } else {
    double *tmp = nullptr;
    KAPSEL_HIP(hipMalloc(&tmp, ctx->g.f * sizeof(double)));
    KAPSEL_HIP(hipMemcpyAsync(tmp, host, ctx->g.f * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    copy_convert_in(ctx, dst, tmp);
    KAPSEL_HIP(hipStreamSynchronize(ctx->stream));
    KAPSEL_HIP(hipFree(tmp));
  }
  KAPSEL_HIP(hipStreamSynchronize(ctx->stream));
}

static void download(kapsel_ctx *ctx, double *host, const void *src) {
  if (ctx->prec == KAPSEL_FP64) {
    KAPSEL_HIP(hipMemcpyAsync(host, src, ctx->g.f * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  } else {
    double *tmp = nullptr;
    KAPSEL_HIP(hipMalloc(&tmp, ctx->g.f * sizeof(double)));
    copy_convert_out(ctx, tmp, src);
    KAPSEL_HIP(hipMemcpyAsync(host, tmp, ctx->g.f * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    KAPSEL_HIP(hipStreamSynchronize(ctx->stream));
    KAPSEL_HIP(hipFree(tmp));
  }
  KAPSEL_HIP(hipStreamSynchronize(ctx->stream));
}

static void *fld(kapsel_ctx *ctx, int f) { return ctx->fields[f]; }

// Zeta_k2u (operate_omega.h:138-141), optionally after the 2/3 cut and FIX_CELL
// (fix_dc as zeta_k2u_k); keep3: the velocity spectrum is kept there (the
// inverse transform reads it from there)
static void zeta_k2u(kapsel_ctx *ctx, int fix_dc, void *const *keep3 = nullptr) {
  zeta_k2u_k(ctx, fix_dc, keep3);
  fft_c2r(ctx, ctx->fields + KAPSEL_U0, 3, keep3);
}

// NS_solver_slavedEuler (fluid_solver.cxx:340-366) with Zeta_k2advection_k
// (operate_omega.cxx:100-111) and U2advection_k (operate_omega.cxx:12-98).
// Where a plane pass applies, the five products are formed inside it (into
// adv_x and adv_w, U left as it is) instead of by the products kernel.
// keep_adv false: the advection term (the reference's f_ns0) is not stored.
// fix_dc >= 0: followed by the step's Zeta_k2u after the 2/3 cut with the
// FIX_CELL mask fix_dc (sp_3d_ns.cxx:29-42), its Zeta_k2u_k in the same
// pass as the zeta update (keep3 as zeta_k2u; the advection term is then
// not stored; spectrum_only: keep3 receives the spectrum and U is not
// transformed).  In an FP32 context right after a step that left U as the
// spectrum zeta was formed from (u_spec_gen), the first Zeta_k2u transforms
// U back instead of forming the spectrum from zeta again.
template <typename T>
static void ensure_adv_x(kapsel_ctx *ctx) {
  for (int i = 0; i < 3; i++)
    if (!ctx->adv_x[i]) KAPSEL_HIP(hipMalloc(&ctx->adv_x[i], ctx->g.f * sizeof(T)));
}
template <typename T>
static void ns_slaved_euler_t(kapsel_ctx *ctx, const kapsel_time &t, bool keep_adv, int fix_dc,
                              void *const *keep3, bool spectrum_only) {
  void **u = ctx->fields + KAPSEL_U0;
  if (ctx->prec == KAPSEL_FP32 && ctx->u_spec_gen == ctx->api_gen) fft_c2r(ctx, u, 3);
  else zeta_k2u(ctx, 0);
  void *prod[5] = {ctx->adv_x[0], ctx->adv_x[1], ctx->adv_x[2], ctx->adv_w[0], ctx->adv_w[1]};
  void *five[5] = {u[0], u[1], u[2], ctx->adv_w[0], ctx->adv_w[1]};
  void *const *in5 = five;
  if (plane_ok_grid(ctx->g)) {
    ensure_adv_x<T>(ctx);
    for (int i = 0; i < 3; i++) prod[i] = ctx->adv_x[i];
    void *b[5] = {u[1], u[2], u[0], u[1], u[2]}, *c[5] = {u[2], u[0], u[1], u[2], u[0]};
    (void)fft_r2c_pro_t<T>(ctx, prod, nullptr, b, c, 1.0, 5, 0x18u);
    in5 = prod;
  } else {
    advection_products(ctx);
    fft_r2c(ctx, five, 5);
  }
  if (fix_dc >= 0) {
    slaved_zeta2u(ctx, t.dt_fluid, in5, fix_dc, keep3);
    if (!spectrum_only) fft_c2r(ctx, u, 3, keep3);
  } else {
    advection_combine_slaved(ctx, t.dt_fluid, in5, keep_adv);
  }
}
static void ns_slaved_euler(kapsel_ctx *ctx, const kapsel_time &t, bool keep_adv = true, int fix_dc = -1,
                            void *const *keep3 = nullptr, bool spectrum_only = false) {
  if (ctx->prec == KAPSEL_FP64) ns_slaved_euler_t<double>(ctx, t, keep_adv, fix_dc, keep3, spectrum_only);
  else ns_slaved_euler_t<float>(ctx, t, keep_adv, fix_dc, keep3, spectrum_only);
}

// Solenoidal_u (operate_omega.h:109-116)
static void solenoidal_u(kapsel_ctx *ctx, void *const *f3) {
  fft_r2c(ctx, f3, 3);
  solenoidal_uk(ctx, f3);
  fft_c2r(ctx, f3, 3);
}

// Make_f_particle_dt_sole + Add_f_particle (f_particle.h:37-95); keep3: also
// store the new real velocity there
static void f_particle_dt_sole_add(kapsel_ctx *ctx, void *const *keep3 = nullptr) {
  make_f_nonsole(ctx);
  solenoidal_u(ctx, ctx->fields + KAPSEL_F0);
  add_f_particle(ctx, keep3);
}

// U2zeta_k (operate_omega.h:212-215)
static void u2zeta_k(kapsel_ctx *ctx) {
  fft_r2c(ctx, ctx->fields + KAPSEL_U0, 3);
  u_k2zeta_k(ctx);
}

// The end of an ICEP step in an FP32 context, from Solenoidal_u(u) on.
// u = c2r(P u_k) and the particle force f = P(u_p - phi u) (f_particle.h)
// enter U2zeta_k only as their sum, whose spectrum is P u_k + P f_k: the
// first is kept when u is projected (keep3, three spare fields), the second
// is never transformed back.  U holds that spectrum afterwards and F the
// unprojected spectrum of f; the FP64 context keeps the reference's
// real-space order.  add_s > 0: u += f add_s first (the dielectric force,
// sp_3d_ns.cxx:77-84), formed inside the forward transform where a plane
// pass applies.
static void solenoidal_u_keep(kapsel_ctx *ctx, void *const *keep3, double add_s = 0.0) {
  void **u = ctx->fields + KAPSEL_U0;
  if (!(add_s != 0.0 && fft_r2c_pro_t<float>(ctx, u, u, ctx->fields + KAPSEL_F0, nullptr, add_s, 3))) {
    if (add_s != 0.0) add_force_to_u(ctx, add_s);
    fft_r2c(ctx, u, 3);
  }
  solenoidal_uk(ctx, u, keep3);
  fft_c2r(ctx, u, 3, keep3);
}
// The same end of the step from the velocity spectrum ns_keep the
// Navier-Stokes part left (cut and solenoidal, so that P(u + f s) = u + P f s):
// keep3 = ns_keep + s P f_k from the forward transform of the force alone,
// then U = c2r(keep3) -- the velocity is not transformed back after the
// Navier-Stokes part and forward again with the force.  f_spectrum: F holds
// the force's spectrum already
static void force_sole_keep(kapsel_ctx *ctx, void *const *ns_keep, void *const *keep3, double s,
                            bool f_spectrum = false) {
  void **f = ctx->fields + KAPSEL_F0;
  if (!f_spectrum) fft_r2c(ctx, f, 3);
  sole_add(ctx, ns_keep, f, keep3, s);
  fft_c2r(ctx, ctx->fields + KAPSEL_U0, 3, keep3);
}
// f = u_p - phi u (Make_f_particle_dt_nonsole) formed inside the forward
// transform of f where a plane pass applies; the projection of f, the sum
// and U_k2zeta_k in one pointwise pass
static void f_sole_u2zeta_k(kapsel_ctx *ctx, void *const *keep3) {
  void *phi3[3] = {fld(ctx, KAPSEL_PHI), fld(ctx, KAPSEL_PHI), fld(ctx, KAPSEL_PHI)};
  void **f = ctx->fields + KAPSEL_F0;
  if (!fft_r2c_pro_t<float>(ctx, f, ctx->fields + KAPSEL_UP0, phi3, ctx->fields + KAPSEL_U0, -1.0, 3)) {
    make_f_nonsole(ctx);
    fft_r2c(ctx, f, 3);
  }
  sole_u2zeta_sum(ctx, keep3, f);
  ctx->u_spec_gen = ctx->api_gen + 1;  // what the next entry point sees if it comes right after
}

static int fix_mask(const kapsel_ctx *ctx) {
  int m = 0;
  for (int d = 0; d < 3; d++)
    if (ctx->P.fix_cell[d]) m |= 1 << d;
  return m;
}

// Time_evolution_hydro, Navier-Stokes branch (sp_3d_ns.cxx:24-186)
// (FP32 context: the velocity spectrum of Zeta_k2u is kept in WORK0..WORK2
// and the step ends in k-space, as f_sole_u2zeta_k describes)
static void spm_step(kapsel_ctx *ctx, kapsel_time &t) {
  void *keep3[3] = {fld(ctx, KAPSEL_WORK0), fld(ctx, KAPSEL_WORK1), fld(ctx, KAPSEL_WORK2)};
  const bool fp32 = ctx->prec == KAPSEL_FP32;
  ns_slaved_euler(ctx, t, false, fix_mask(ctx), fp32 ? keep3 : nullptr);  // :27, 29-42
  md_position(ctx, t);                        // :44-50
  make_phi_particle_sum(ctx);                 // :51-53
  hydro_force_velocity(ctx, t);               // :99-110, Calc_f_hydro, Force, velocity
  thermostat(ctx, t);                         // :155-157, with kBT > 0
  make_u_particle_sum(ctx);                   // :166-167
  if (fp32) {
    f_sole_u2zeta_k(ctx, keep3);              // :168-169, 184
  } else {
    f_particle_dt_sole_add(ctx);              // :168-169
    u2zeta_k(ctx);                            // :184
  }
  t.time += t.dt_fluid;                       // :525
  t.ts += 1;
}

// ICEP step without ions: Time_evolution_hydro's Navier-Stokes step with the
// dielectric force of the Electrolyte+Dielectric branch inserted after the
// profile scatter (sp_3d_ns.cxx:51-95): eps, Poisson solve warm-started from
// the previous potential, Maxwell force, u += f dt / rho, projection
//
// The Navier-Stokes part (the slaved step and Zeta_k2u, which read and write
// zeta, uk_dc, U and the advection's two extra product fields) does not touch
// what the particle and Poisson part reads or writes (particles, PHI,
// EPSILON, the potential, the solver's vectors WORK0..WORK3 and scratch),
// and only u += f dt / rho joins them.  With concurrency on, the NS part is
// queued on the context's second stream with its own two product fields
// while the Poisson part, whose CG and refinement leave the device partly
// idle (single-workgroup reductions, plane passes on 1.5 planes per CU, host
// read-backs), runs on the context stream; an event joins them.  Same
// kernels, same inputs: the results are those of the one-stream order.
struct SideStream {
  kapsel_ctx *c;
  hipStream_t main;
  bool refine;
  explicit SideStream(kapsel_ctx *x) : c(x), main(x->stream), refine(x->refine64) {
    c->stream = c->side;
    c->refine64 = false;  // queued from inside the solve: not refinement work
    c->adv_w[0] = c->ns_work[0];
    c->adv_w[1] = c->ns_work[1];
  }
  ~SideStream() {
    c->stream = main;
    c->refine64 = refine;
    c->adv_w[0] = c->fields[KAPSEL_WORK0];
    c->adv_w[1] = c->fields[KAPSEL_WORK1];
  }
};

static void icep_step(kapsel_ctx *ctx, kapsel_time &t, const double e0[3], double tol, int maxit, int precond,
                      int *iters, double *relres) {
  // (kernel timing records each launch's events on the stream it goes to)
  const bool split = ctx->concurrent;
  // joins the context stream to the NS chain on every way out (an error in the
  // Poisson part included), so nothing later on the context stream races it
  struct Join {
    kapsel_ctx *c;
    bool armed = false;
    void now() {
      if (armed) {
        armed = false;
        KAPSEL_HIP(hipStreamWaitEvent(c->stream, c->ev_join, 0));
      }
    }
    ~Join() {
      if (armed) (void)hipStreamWaitEvent(c->stream, c->ev_join, 0);
    }
  } join{ctx};
  // the NS chain runs beside the start of the Poisson part (profile,
  // permittivity, the refinement's set-up and first FP64 residual, whose FP64
  // plane pass runs a second round on half the CUs) and is done before the
  // FP32 CG starts.  FP32 with a plane pass: it ends with its velocity
  // spectrum in adv_x, and the step's end adds the projected force to that
  const bool fp32 = ctx->prec == KAPSEL_FP32;
  const bool spec = fp32 && plane_ok_grid(ctx->g);
  if (spec) ensure_adv_x<float>(ctx);
  void *ns_keep[3] = {ctx->adv_x[0], ctx->adv_x[1], ctx->adv_x[2]};
  if (split) {
    for (void *&w : ctx->ns_work)
      if (!w) KAPSEL_HIP(hipMalloc(&w, ctx->g.f * ctx->elem));
    KAPSEL_HIP(hipEventRecord(ctx->ev_fork, ctx->stream));
    KAPSEL_HIP(hipStreamWaitEvent(ctx->side, ctx->ev_fork, 0));
    {
      SideStream on_side(ctx);
      ns_slaved_euler(ctx, t, false, fix_mask(ctx), spec ? ns_keep : nullptr, spec);
    }
    KAPSEL_HIP(hipEventRecord(ctx->ev_join, ctx->side));
    join.armed = true;
  } else {
    ns_slaved_euler(ctx, t, false, fix_mask(ctx), spec ? ns_keep : nullptr, spec);
  }
  md_position(ctx, t);
  make_phi_particle_sum(ctx);
  dielectric_field(ctx);
  {
    // FP32: the Maxwell force is queued with each potential the solve is
    // about to check, so the GPU computes it while the host waits for the
    // check (the solve says whether it was queued for its final potential)
    const std::function<void()> mx = [&]() { maxwell_force_from_spectra(ctx, e0); };
    struct Hook {
      kapsel_ctx *c;
      ~Hook() { c->solve_hook = nullptr; }
    } hook{ctx};
    if (fp32) ctx->solve_hook = &mx;
    poisson_solve(ctx, e0, tol, maxit, precond, iters, relres);
  }
  if (!(fp32 && ctx->solve_hook_done)) maxwell_force(ctx, e0, true);
  join.now();
  void *keep3[3] = {ctx->pscratch[5], ctx->pscratch[6], ctx->pscratch[7]};
  if (spec) {
    force_sole_keep(ctx, ns_keep, keep3, t.dt_fluid * ctx->irho);
  } else if (fp32) {
    solenoidal_u_keep(ctx, keep3, t.dt_fluid * ctx->irho);  // u += f dt / rho inside
  } else {
    add_force_to_u(ctx, t.dt_fluid * ctx->irho);
    solenoidal_u(ctx, ctx->fields + KAPSEL_U0);
  }
  hydro_force_velocity(ctx, t);
  make_u_particle_sum(ctx);
  if (fp32) {
    f_sole_u2zeta_k(ctx, keep3);
  } else {
    f_particle_dt_sole_add(ctx);
    u2zeta_k(ctx);
  }
  t.time += t.dt_fluid;
  t.ts += 1;
}

// Time_evolution_hydro, Electrolyte branch without Dielectric
// (sp_3d_ns.cxx:24-186): the Nernst-Planck Euler step replaces the slaved
// step, Rescale_solute follows the profile scatter, and the Coulomb force is
// applied to u before the hydrodynamic force.  The branch's first
// Calc_f_hydro (:60-68) only fills the diagnostic f_hydro1; it is skipped
// (the second call overwrites f_hydro and torque_hydro before any use).
// FP32 context: the velocity of Zeta_k2u stays a spectrum (cut and
// solenoidal) until the Coulomb force is added in k-space, and the step ends
// in k-space (force_sole_keep, f_sole_u2zeta_k), leaving U as the velocity
// spectrum; a step that reuses it transforms that back first.
static void electrolyte_step(kapsel_ctx *ctx, kapsel_time &t) {
  const bool reuse = ctx->el_reuse_gen == ctx->api_gen;
  const bool fp32 = ctx->prec == KAPSEL_FP32;
  void **u = ctx->fields + KAPSEL_U0;
  if (reuse && fp32) fft_c2r(ctx, electrolyte_u_keep(ctx), 3, u);
  nssolute_step(ctx, t, reuse);               // :20
  if (fp32) zeta_k2u_k(ctx, fix_mask(ctx));  // :29-42, U: the spectrum
  else zeta_k2u(ctx, fix_mask(ctx));
  md_position(ctx, t);                        // :44-50
  make_phi_particle_sum(ctx);                 // :51-53
  rescale_and_coulomb_force(ctx, t, fp32);    // :55-72 (FP32: F receives the spectrum)
  void *keep3[3] = {fld(ctx, KAPSEL_WORK0), fld(ctx, KAPSEL_WORK1), fld(ctx, KAPSEL_WORK2)};
  if (fp32) {
    force_sole_keep(ctx, u, keep3, t.dt_fluid * ctx->irho, true);  // :75-85
  } else {
    add_force_to_u(ctx, t.dt_fluid * ctx->irho);  // :75-84
    solenoidal_u(ctx, u);                         // :85
  }
  hydro_force_velocity(ctx, t);               // :99-110, Calc_f_hydro, Force, velocity
  make_u_particle_sum(ctx);                   // :166-167
  if (fp32) {
    f_sole_u2zeta_k(ctx, keep3);              // :168-169, 184
  } else {
    f_particle_dt_sole_add(ctx, electrolyte_u_keep(ctx));  // :168-169
    u2zeta_k(ctx);                            // :184
  }
  t.time += t.dt_fluid;
  t.ts += 1;
  // the next entry point sees api_gen + 1; if it is this step again, what
  // this step left (charge spectrum, real ion fields, real velocity) is current
  ctx->el_reuse_gen = ctx->api_gen + 1;
}

// the free charge of the ions (and surface charge) in the variable
// permittivity and its potential, with the uniform background that cancels
// the charge's total (the ions inside the particles, weight phi_sin, leave one
// that no periodic potential satisfies) in the right-hand side only
// (Make_Maxwell_force_x_on_fluid, operate_dielectric.cxx:629-640)
static void dielectric_potential(kapsel_ctx *ctx, const double e0[3], double tol, int maxit, int precond,
                                 int *iters, double *relres) {
  dielectric_field(ctx);
  struct Background {
    kapsel_ctx *c;
    ~Background() { c->rho_bg = nullptr; }
  } bg{ctx};
  ctx->rho_bg = dielectric_free_charge(ctx);  // writes RHO_FREE
  ctx->rho_free_set = true;
  poisson_solve(ctx, e0, tol, maxit, precond, iters, relres);
}

// Time_evolution_hydro, Electrolyte branch with Dielectric (sp_3d_ns.cxx:
// 20, 24-186): NSsolute_solver_Euler with the potential's gradient, no
// Rescale_solute, Make_Maxwell_force_x_on_fluid with the ions' charge
static void icep_electrolyte_step(kapsel_ctx *ctx, kapsel_time &t, double tol, int maxit, int precond, int *iters,
                                  double *relres) {
  const bool fp32 = ctx->prec == KAPSEL_FP32;
  void **u = ctx->fields + KAPSEL_U0;
  nssolute_step(ctx, t, false);               // :20
  if (fp32) zeta_k2u_k(ctx, fix_mask(ctx));  // :29-42, U: the spectrum (as electrolyte_step)
  else zeta_k2u(ctx, fix_mask(ctx));
  md_position(ctx, t);                        // :44-50
  make_phi_particle_sum(ctx);                 // :51-53
  const double s = ctx->el.ac ? sin(ctx->el.angular_frequency * t.time) : 1.0;
  const double e0[3] = {ctx->el.e_ext[0] * s, ctx->el.e_ext[1] * s, ctx->el.e_ext[2] * s};
  dielectric_potential(ctx, e0, tol, maxit, precond, iters, relres);  // :72-74
  maxwell_force(ctx, e0, true);
  void *keep3[3] = {ctx->pscratch[5], ctx->pscratch[6], ctx->pscratch[7]};
  if (fp32) {
    force_sole_keep(ctx, u, keep3, t.dt_fluid * ctx->irho);  // :75-85
  } else {
    add_force_to_u(ctx, t.dt_fluid * ctx->irho);  // :75-84
    solenoidal_u(ctx, ctx->fields + KAPSEL_U0);   // :85
  }
  hydro_force_velocity(ctx, t);               // :99-110, Calc_f_hydro, Force, velocity
  make_u_particle_sum(ctx);                   // :166-167
  if (fp32) {
    f_sole_u2zeta_k(ctx, keep3);              // :168-169, 184
  } else {
    f_particle_dt_sole_add(ctx);              // :168-169
    u2zeta_k(ctx);                            // :184
  }
  t.time += t.dt_fluid;
  t.ts += 1;
}

// main, initial rigidity projection (sp_3d_ns.cxx:431-447)
static void init_rigidity(kapsel_ctx *ctx) {
  make_phi_particle_sum(ctx);
  make_u_particle_sum(ctx);
  zeta_k2u(ctx, -1);
  f_particle_dt_sole_add(ctx);
  u2zeta_k(ctx);
  KAPSEL_HIP(hipMemsetAsync(ctx->d_uk_dc, 0, 3 * sizeof(double), ctx->stream));
}

}  // namespace kapsel

using namespace kapsel;

#define API_BEGIN try {
#define API_END                                   \
  }                                               \
  catch (const Error &e) {                        \
    return e.code;                                \
  }                                               \
  catch (const std::bad_alloc &) {                \
    set_error("host allocation failed");          \
    return KAPSEL_ERR_STATE;                      \
  }                                               \
  catch (const std::exception &e) {               \
    set_error("%s", e.what());                    \
    return KAPSEL_ERR_STATE;                      \
  }                                               \
  return KAPSEL_OK;

extern "C" {

const char *kapsel_last_error(void) { return g_err; }

int kapsel_create(const kapsel_params *params, kapsel_ctx **out) {
  API_BEGIN
  create(params, out);
  API_END
}

int kapsel_destroy(kapsel_ctx *ctx) {
  API_BEGIN
  destroy(ctx);
  API_END
}

int kapsel_synchronize(kapsel_ctx *ctx) {
  API_BEGIN
  check_ctx_ro(ctx);
  KAPSEL_HIP(hipStreamSynchronize(ctx->side));
  KAPSEL_HIP(hipStreamSynchronize(ctx->stream));
  API_END
}

int kapsel_set_concurrency(kapsel_ctx *ctx, int on) {
  API_BEGIN
  check_ctx(ctx);
  ctx->concurrent = on != 0;
  API_END
}

int kapsel_get_sekibun(kapsel_ctx *ctx, int *np_domain, int *cells) {
  API_BEGIN
  check_ctx_ro(ctx);
  if (!np_domain) {
    set_error("null np_domain");
    throw Error{KAPSEL_ERR_ARG};
  }
  *np_domain = ctx->np_domain;
  if (cells) memcpy(cells, ctx->sekibun.data(), sizeof(int) * ctx->sekibun.size());
  API_END
}

int kapsel_set_field(kapsel_ctx *ctx, int field, const double *host) {
  API_BEGIN
  check_ctx(ctx);
  check_field(field);
  written(ctx, field);
  if (!host) {
    set_error("null host buffer");
    throw Error{KAPSEL_ERR_ARG};
  }
  upload(ctx, ctx->fields[field], host);
  API_END
}

int kapsel_get_field(kapsel_ctx *ctx, int field, double *host) {
  API_BEGIN
  check_ctx_ro(ctx);
  check_field(field);
  if (!host) {
    set_error("null host buffer");
    throw Error{KAPSEL_ERR_ARG};
  }
  if (field == KAPSEL_POTENTIAL || field == KAPSEL_RHS) materialize_outputs(ctx);
  download(ctx, host, ctx->fields[field]);
  API_END
}

int kapsel_set_uk_dc(kapsel_ctx *ctx, const double uk_dc[3]) {
  API_BEGIN
  check_ctx(ctx);
  KAPSEL_HIP(hipMemcpyAsync(ctx->d_uk_dc, uk_dc, 3 * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  KAPSEL_HIP(hipStreamSynchronize(ctx->stream));
  API_END
}

int kapsel_get_uk_dc(kapsel_ctx *ctx, double uk_dc[3]) {
  API_BEGIN
  check_ctx_ro(ctx);
  KAPSEL_HIP(hipMemcpyAsync(uk_dc, ctx->d_uk_dc, 3 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  KAPSEL_HIP(hipStreamSynchronize(ctx->stream));
  API_END
}

int kapsel_set_particles(kapsel_ctx *ctx, const double *state) {
  API_BEGIN
  check_ctx(ctx);
  size_t n = sizeof(double) * KAPSEL_P_NFIELDS * 3 * ctx->P.n_particles;
  if (n && !state) {
    set_error("null particle state");
    throw Error{KAPSEL_ERR_ARG};
  }
  const double L[3] = {ctx->g.lx, ctx->g.ly, ctx->g.lz};
  for (int d = 0; d < 3; d++)
    for (int i = 0; i < ctx->P.n_particles; i++) {
      double x = state[((size_t)KAPSEL_P_X * 3 + d) * ctx->P.n_particles + i];
      if (!(x >= 0.0 && x < L[d])) {
        set_error("particle %d position[%d] = %g outside [0, %g)", i, d, x, L[d]);
        throw Error{KAPSEL_ERR_ARG};
      }
    }
  if (n) {
    KAPSEL_HIP(hipMemcpyAsync(ctx->d_ps, state, n, hipMemcpyHostToDevice, ctx->stream));
    KAPSEL_HIP(hipStreamSynchronize(ctx->stream));
  }
  API_END
}

int kapsel_get_particles(kapsel_ctx *ctx, double *state) {
  API_BEGIN
  check_ctx_ro(ctx);
  size_t n = sizeof(double) * KAPSEL_P_NFIELDS * 3 * ctx->P.n_particles;
  if (n) {
    KAPSEL_HIP(hipMemcpyAsync(state, ctx->d_ps, n, hipMemcpyDeviceToHost, ctx->stream));
    KAPSEL_HIP(hipStreamSynchronize(ctx->stream));
  }
  API_END
}

int kapsel_set_orientations(kapsel_ctx *ctx, const double *q) {
  API_BEGIN
  check_ctx(ctx);
  const int np = ctx->P.n_particles;
  if (!q && np > 0) {
    set_error("null orientations");
    throw Error{KAPSEL_ERR_ARG};
  }
  for (int i = 0; i < np; i++)
    for (int h = 0; h < 2; h++) {
      double s2 = 0.0;
      for (int c = 0; c < 4; c++) s2 += q[(size_t)(4 * h + c) * np + i] * q[(size_t)(4 * h + c) * np + i];
      if (!(fabs(sqrt(s2) - 1.0) < 1e-6)) {  // qtn_isnormal (quaternion.h:272-281)
        set_error("particle %d: %s is not a unit quaternion (norm %g)", i, h ? "q_old" : "q", sqrt(s2));
        throw Error{KAPSEL_ERR_ARG};
      }
    }
  if (np > 0) {
    KAPSEL_HIP(hipMemcpyAsync(ctx->d_q, q, sizeof(double) * 8 * np, hipMemcpyHostToDevice, ctx->stream));
    KAPSEL_HIP(hipStreamSynchronize(ctx->stream));
  }
  API_END
}

int kapsel_get_orientations(kapsel_ctx *ctx, double *q) {
  API_BEGIN
  check_ctx_ro(ctx);
  const int np = ctx->P.n_particles;
  if (!q && np > 0) {
    set_error("null output");
    throw Error{KAPSEL_ERR_ARG};
  }
  if (np > 0) {
    KAPSEL_HIP(hipMemcpyAsync(q, ctx->d_q, sizeof(double) * 8 * np, hipMemcpyDeviceToHost, ctx->stream));
    KAPSEL_HIP(hipStreamSynchronize(ctx->stream));
  }
  API_END
}

int kapsel_copy_field(kapsel_ctx *ctx, int dst, int src) {
  API_BEGIN
  check_ctx(ctx);
  check_field(dst);
  check_field(src);
  written(ctx, dst);
  copy_field(ctx, ctx->fields[dst], ctx->fields[src]);
  API_END
}

int kapsel_fft_r2c(kapsel_ctx *ctx, int field) {
  API_BEGIN
  check_ctx(ctx);
  check_field(field);
  written(ctx, field);
  fft_r2c(ctx, ctx->fields + field, 1);
  API_END
}

int kapsel_fft_c2r(kapsel_ctx *ctx, int field) {
  API_BEGIN
  check_ctx(ctx);
  check_field(field);
  written(ctx, field);
  fft_c2r(ctx, ctx->fields + field, 1);
  API_END
}

int kapsel_truncate(kapsel_ctx *ctx, int field) {
  API_BEGIN
  check_ctx(ctx);
  check_field(field);
  written(ctx, field);
  truncate(ctx, ctx->fields + field, 1);
  API_END
}

int kapsel_zeta_k2u_k(kapsel_ctx *ctx) {
  API_BEGIN
  check_ctx(ctx);
  zeta_k2u_k(ctx, -1);
  API_END
}

int kapsel_u_k2zeta_k(kapsel_ctx *ctx) {
  API_BEGIN
  check_ctx(ctx);
  u_k2zeta_k(ctx);
  API_END
}

int kapsel_solenoidal_u(kapsel_ctx *ctx, int field0) {
  API_BEGIN
  check_ctx(ctx);
  check_field(field0);
  check_field(field0 + 2);
  for (int d = 0; d < 3; d++) written(ctx, field0 + d);
  solenoidal_u(ctx, ctx->fields + field0);
  API_END
}

int kapsel_ns_slaved_euler(kapsel_ctx *ctx, const kapsel_time *jikan) {
  API_BEGIN
  check_ctx(ctx);
  check_time(jikan);
  ns_slaved_euler(ctx, *jikan);
  API_END
}

int kapsel_make_phi_particle_sum(kapsel_ctx *ctx) {
  API_BEGIN
  check_ctx(ctx);
  make_phi_particle_sum(ctx);
  API_END
}

int kapsel_make_u_particle_sum(kapsel_ctx *ctx) {
  API_BEGIN
  check_ctx(ctx);
  make_u_particle_sum(ctx);
  API_END
}

int kapsel_make_phi_particle(kapsel_ctx *ctx, int field) {
  API_BEGIN
  check_ctx(ctx);
  check_field(field);
  written(ctx, field);
  make_phi_particle(ctx, ctx->fields[field]);
  API_END
}

int kapsel_calc_f_hydro(kapsel_ctx *ctx, const kapsel_time *jikan) {
  API_BEGIN
  check_ctx(ctx);
  check_time(jikan);
  calc_f_hydro(ctx, *jikan);
  API_END
}

int kapsel_force(kapsel_ctx *ctx) {
  API_BEGIN
  check_ctx(ctx);
  force(ctx);
  API_END
}

int kapsel_md_position(kapsel_ctx *ctx, const kapsel_time *jikan) {
  API_BEGIN
  check_ctx(ctx);
  check_time(jikan);
  md_position(ctx, *jikan);
  API_END
}

int kapsel_md_velocity(kapsel_ctx *ctx, const kapsel_time *jikan) {
  API_BEGIN
  check_ctx(ctx);
  check_time(jikan);
  force(ctx);
  md_velocity_integrate(ctx, *jikan);
  API_END
}

int kapsel_make_f_particle_dt_sole(kapsel_ctx *ctx) {
  API_BEGIN
  check_ctx(ctx);
  f_particle_dt_sole_add(ctx);
  API_END
}

int kapsel_init_rigidity(kapsel_ctx *ctx) {
  API_BEGIN
  check_ctx(ctx);
  init_rigidity(ctx);
  API_END
}

int kapsel_spm_step(kapsel_ctx *ctx, kapsel_time *jikan) {
  API_BEGIN
  check_ctx(ctx);
  check_time(jikan);
  spm_step(ctx, *jikan);
  API_END
}

int kapsel_dielectric_field(kapsel_ctx *ctx) {
  API_BEGIN
  check_ctx(ctx);
  dielectric_field(ctx);
  API_END
}

int kapsel_poisson_apply(kapsel_ctx *ctx, int x_field, int ax_field) {
  API_BEGIN
  check_ctx(ctx);
  check_field(x_field);
  check_field(ax_field);
  if (x_field == ax_field) {
    set_error("poisson_apply needs distinct input and output fields");
    throw Error{KAPSEL_ERR_ARG};
  }
  written(ctx, ax_field);
  poisson_apply(ctx, ctx->fields[x_field], ctx->fields[ax_field]);
  API_END
}

int kapsel_poisson_apply_host(kapsel_ctx *ctx, const double *x, double *ax) {
  API_BEGIN
  check_ctx(ctx);
  if (!x || !ax) {
    set_error("null vector");
    throw Error{KAPSEL_ERR_ARG};
  }
  const Grid &g = ctx->g;
  std::vector<double> pad(g.f, 0.0);
  for (size_t row = 0; row < (size_t)g.nx * g.ny; row++)
    memcpy(&pad[row * g.nz_], x + row * g.nz, g.nz * sizeof(double));
  upload(ctx, ctx->fields[KAPSEL_WORK3], pad.data());
  poisson_apply(ctx, ctx->fields[KAPSEL_WORK3], ctx->fields[KAPSEL_WORK4]);
  download(ctx, pad.data(), ctx->fields[KAPSEL_WORK4]);
  for (size_t row = 0; row < (size_t)g.nx * g.ny; row++)
    memcpy(ax + row * g.nz, &pad[row * g.nz_], g.nz * sizeof(double));
  API_END
}

int kapsel_poisson_rhs(kapsel_ctx *ctx, const double e0[3]) {
  API_BEGIN
  check_ctx(ctx);
  if (!e0) {
    set_error("null external field");
    throw Error{KAPSEL_ERR_ARG};
  }
  poisson_rhs(ctx, e0);
  API_END
}

int kapsel_poisson_solve(kapsel_ctx *ctx, const double e0[3], double tol, int maxit, int precond, int *iters,
                         double *relres) {
  API_BEGIN
  check_ctx(ctx);
  if (!e0 || !iters || !relres || !(tol > 0) || maxit < 1 ||
      (precond != KAPSEL_PRECOND_NONE && precond != KAPSEL_PRECOND_SPECTRAL)) {
    set_error("invalid poisson_solve arguments");
    throw Error{KAPSEL_ERR_ARG};
  }
  poisson_solve(ctx, e0, tol, maxit, precond, iters, relres);
  if (!(*relres < tol)) {
    set_error("poisson_solve stopped at relative residual %.3e after %d iterations", *relres, *iters);
    return KAPSEL_ERR_NOCONV;
  }
  API_END
}

int kapsel_pnet_init(kapsel_ctx *ctx, long *n_params) {
  API_BEGIN
  check_ctx(ctx);
  pnet_init(ctx);
  if (n_params) *n_params = pnet_param_count(ctx);
  API_END
}

int kapsel_pnet_set_params(kapsel_ctx *ctx, const float *params) {
  API_BEGIN
  check_ctx(ctx);
  if (!params) {
    set_error("null parameters");
    throw Error{KAPSEL_ERR_ARG};
  }
  pnet_set_params(ctx, params);
  API_END
}

int kapsel_pnet_get_params(kapsel_ctx *ctx, float *params) {
  API_BEGIN
  check_ctx(ctx);
  if (!params) {
    set_error("null parameters");
    throw Error{KAPSEL_ERR_ARG};
  }
  pnet_get_params(ctx, params);
  API_END
}

static void check_unit(const double e[3]) {
  if (!e || !(fabs(e[0] * e[0] + e[1] * e[1] + e[2] * e[2] - 1.0) < 1e-6)) {
    set_error("field direction must be a unit vector");
    throw Error{KAPSEL_ERR_ARG};
  }
}

int kapsel_pnet_loss_grad(kapsel_ctx *ctx, double beta, const double e_hat[3], double scale, double *loss,
                          float *grad) {
  API_BEGIN
  check_ctx(ctx);
  check_unit(e_hat);
  if (!loss || !(scale > 0)) {
    set_error("null loss or non-positive scale");
    throw Error{KAPSEL_ERR_ARG};
  }
  *loss = pnet_loss_grad(ctx, beta, e_hat, scale, grad);
  API_END
}

int kapsel_pnet_train_step(kapsel_ctx *ctx, double beta, const double e_hat[3], double scale, double lr,
                           double *loss) {
  API_BEGIN
  check_ctx(ctx);
  check_unit(e_hat);
  if (!loss || !(scale > 0) || !(lr > 0)) {
    set_error("null loss, non-positive scale or learning rate");
    throw Error{KAPSEL_ERR_ARG};
  }
  *loss = pnet_train_step(ctx, beta, e_hat, scale, lr);
  API_END
}

int kapsel_pnet_warm_start(kapsel_ctx *ctx, double beta, const double e0[3], double scale) {
  API_BEGIN
  check_ctx(ctx);
  if (!e0 || !(scale > 0)) {
    set_error("null field or non-positive scale");
    throw Error{KAPSEL_ERR_ARG};
  }
  written(ctx, KAPSEL_POTENTIAL);
  pnet_warm_start(ctx, beta, e0, scale);
  API_END
}

int kapsel_poisson_galerkin_start(kapsel_ctx *ctx, const double e0[3], double *gamma) {
  API_BEGIN
  check_ctx(ctx);
  if (!e0) {
    set_error("null field");
    throw Error{KAPSEL_ERR_ARG};
  }
  const double g = galerkin_scale_potential(ctx, e0);
  if (gamma) *gamma = g;
  API_END
}

int kapsel_set_dielectric(kapsel_ctx *ctx, double eps_fluid, double eps_particle, double dielectric_cst) {
  API_BEGIN
  check_ctx(ctx);
  if (!(eps_fluid > 0) || !(eps_particle > 0) || !(dielectric_cst > 0)) {
    set_error("permittivities must be positive");
    throw Error{KAPSEL_ERR_ARG};
  }
  ctx->P.eps_fluid = eps_fluid;
  ctx->P.eps_particle = eps_particle;
  ctx->P.dielectric_cst = dielectric_cst;
  API_END
}

int kapsel_maxwell_force(kapsel_ctx *ctx, const double e0[3]) {
  API_BEGIN
  check_ctx(ctx);
  if (!e0) {
    set_error("null external field");
    throw Error{KAPSEL_ERR_ARG};
  }
  maxwell_force(ctx, e0);
  API_END
}

int kapsel_icep_step(kapsel_ctx *ctx, kapsel_time *jikan, const double e0[3], double tol, int maxit, int precond,
                     int *iters, double *relres) {
  API_BEGIN
  // no materialize_outputs: the step's solve reads the FP64 shadow, not the
  // FP32 POTENTIAL, and its outputs replace the pending ones
  check_ctx_ro(ctx);
  ctx->api_gen++;
  check_time(jikan);
  if (!e0 || !iters || !relres || !(tol > 0) || maxit < 1 ||
      (precond != KAPSEL_PRECOND_NONE && precond != KAPSEL_PRECOND_SPECTRAL)) {
    set_error("invalid icep_step arguments");
    throw Error{KAPSEL_ERR_ARG};
  }
  {
    struct Lazy {
      kapsel_ctx *c;
      explicit Lazy(kapsel_ctx *x) : c(x) { c->lazy_outputs = true; }
      ~Lazy() { c->lazy_outputs = false; }
    } lazy(ctx);
    icep_step(ctx, *jikan, e0, tol, maxit, precond, iters, relres);
  }
  if (!(*relres < tol)) {
    set_error("icep_step: Poisson solve stopped at relative residual %.3e", *relres);
    return KAPSEL_ERR_NOCONV;
  }
  API_END
}

int kapsel_set_electrolyte(kapsel_ctx *ctx, const kapsel_electrolyte *el) {
  API_BEGIN
  check_ctx(ctx);
  if (!el) {
    set_error("null electrolyte parameters");
    throw Error{KAPSEL_ERR_ARG};
  }
  electrolyte_set(ctx, *el);
  API_END
}

int kapsel_init_rho_ion(kapsel_ctx *ctx) {
  API_BEGIN
  check_ctx(ctx);
  init_rho_ion(ctx);
  written(ctx, KAPSEL_PHI);
  API_END
}

int kapsel_init_rho_ion_pb(kapsel_ctx *ctx, const kapsel_time *jikan, int max_iter, int *iters,
                           double free_energy[3]) {
  API_BEGIN
  check_ctx(ctx);
  check_time(jikan);
  if (!iters) {
    set_error("null iters");
    throw Error{KAPSEL_ERR_ARG};
  }
  double fe[3];
  *iters = init_rho_ion_pb(ctx, *jikan, max_iter, fe);
  if (free_energy)
    for (int c = 0; c < 3; c++) free_energy[c] = fe[c];
  written(ctx, KAPSEL_PHI);
  API_END
}

int kapsel_free_energy_pb(kapsel_ctx *ctx, const kapsel_time *jikan, double free_energy[3]) {
  API_BEGIN
  check_ctx(ctx);
  check_time(jikan);
  if (!free_energy) {
    set_error("null output");
    throw Error{KAPSEL_ERR_ARG};
  }
  free_energy_pb(ctx, *jikan, free_energy);
  API_END
}

int kapsel_get_total_solute(kapsel_ctx *ctx, double total[2]) {
  API_BEGIN
  check_ctx_ro(ctx);
  if (!total) {
    set_error("null output");
    throw Error{KAPSEL_ERR_ARG};
  }
  total[0] = ctx->el_total[0];
  total[1] = ctx->el_total[1];
  API_END
}

int kapsel_set_total_solute(kapsel_ctx *ctx, const double total[2]) {
  API_BEGIN
  check_ctx(ctx);
  if (!total || !(total[0] > 0) || !(total[1] > 0)) {
    set_error("totals must be positive");
    throw Error{KAPSEL_ERR_ARG};
  }
  ctx->el_total[0] = total[0];
  ctx->el_total[1] = total[1];
  API_END
}

int kapsel_make_phi_qq(kapsel_ctx *ctx, int phi_field, int surface_field) {
  API_BEGIN
  check_ctx(ctx);
  check_field(phi_field);
  check_field(surface_field);
  if (phi_field == surface_field) {
    set_error("phi and surface fields must differ");
    throw Error{KAPSEL_ERR_ARG};
  }
  make_phi_qq(ctx, fld(ctx, phi_field), fld(ctx, surface_field));
  written(ctx, phi_field);
  written(ctx, surface_field);
  API_END
}

int kapsel_make_surface_normal(kapsel_ctx *ctx, int field0) {
  API_BEGIN
  check_ctx(ctx);
  check_field(field0);
  check_field(field0 + 2);
  make_surface_normal(ctx, ctx->fields + field0);
  for (int d = 0; d < 3; d++) written(ctx, field0 + d);
  API_END
}

int kapsel_charge_field(kapsel_ctx *ctx, int field) {
  API_BEGIN
  check_ctx(ctx);
  check_field(field);
  charge_field(ctx, fld(ctx, field));
  written(ctx, field);
  API_END
}

int kapsel_nssolute_euler(kapsel_ctx *ctx, const kapsel_time *jikan) {
  API_BEGIN
  check_ctx(ctx);
  check_time(jikan);
  nssolute_euler(ctx, *jikan);
  API_END
}

int kapsel_rescale_solute(kapsel_ctx *ctx, double factor[2]) {
  API_BEGIN
  check_ctx(ctx);
  rescale_solute(ctx, factor);
  API_END
}

int kapsel_coulomb_force(kapsel_ctx *ctx, const kapsel_time *jikan) {
  API_BEGIN
  check_ctx(ctx);
  check_time(jikan);
  coulomb_force(ctx, *jikan);
  API_END
}

int kapsel_electrolyte_step(kapsel_ctx *ctx, kapsel_time *jikan) {
  API_BEGIN
  check_ctx(ctx);
  check_time(jikan);
  if (!ctx->el_set) {
    set_error("electrolyte parameters not set (kapsel_set_electrolyte)");
    throw Error{KAPSEL_ERR_STATE};
  }
  if (ctx->el.dielectric) {
    set_error("the electrolyte was set with Dielectric: kapsel_icep_electrolyte_step");
    throw Error{KAPSEL_ERR_STATE};
  }
  if (!(ctx->el_total[0] > 0) || !(ctx->el_total[1] > 0)) {
    set_error("ion totals not set (kapsel_init_rho_ion, kapsel_set_total_solute, or kapsel_load_restart "
              "after kapsel_set_electrolyte)");
    throw Error{KAPSEL_ERR_STATE};
  }
  electrolyte_step(ctx, *jikan);
  API_END
}

static void check_solve_args(double tol, int maxit, int precond, const int *iters, const double *relres) {
  if (!iters || !relres || !(tol > 0) || maxit < 1 ||
      (precond != KAPSEL_PRECOND_NONE && precond != KAPSEL_PRECOND_SPECTRAL)) {
    set_error("invalid solve arguments (tol > 0, maxit >= 1, a known preconditioner, non-null outputs)");
    throw Error{KAPSEL_ERR_ARG};
  }
}

int kapsel_init_rho_ion_dielectric(kapsel_ctx *ctx) {
  API_BEGIN
  check_ctx(ctx);
  init_rho_ion_dielectric(ctx);
  written(ctx, KAPSEL_PHI);
  API_END
}

int kapsel_dielectric_potential(kapsel_ctx *ctx, const kapsel_time *jikan, double tol, int maxit, int precond,
                                int *iters, double *relres) {
  API_BEGIN
  check_ctx(ctx);
  check_time(jikan);
  check_solve_args(tol, maxit, precond, iters, relres);
  if (!ctx->el_set || !ctx->el.dielectric) {
    set_error("kapsel_dielectric_potential needs the electrolyte set with Dielectric");
    throw Error{KAPSEL_ERR_STATE};
  }
  const double s = ctx->el.ac ? sin(ctx->el.angular_frequency * jikan->time) : 1.0;
  const double e0[3] = {ctx->el.e_ext[0] * s, ctx->el.e_ext[1] * s, ctx->el.e_ext[2] * s};
  dielectric_potential(ctx, e0, tol, maxit, precond, iters, relres);
  if (!(*relres < tol)) {
    set_error("dielectric_potential: Poisson solve stopped at relative residual %.3e", *relres);
    return KAPSEL_ERR_NOCONV;
  }
  API_END
}

int kapsel_icep_electrolyte_step(kapsel_ctx *ctx, kapsel_time *jikan, double tol, int maxit, int precond,
                                 int *iters, double *relres) {
  API_BEGIN
  check_ctx(ctx);
  check_time(jikan);
  check_solve_args(tol, maxit, precond, iters, relres);
  if (!ctx->el_set || !ctx->el.dielectric) {
    set_error("kapsel_icep_electrolyte_step needs the electrolyte set with Dielectric");
    throw Error{KAPSEL_ERR_STATE};
  }
  icep_electrolyte_step(ctx, *jikan, tol, maxit, precond, iters, relres);
  if (!(*relres < tol)) {
    set_error("icep_electrolyte_step: Poisson solve stopped at relative residual %.3e", *relres);
    return KAPSEL_ERR_NOCONV;
  }
  API_END
}

int kapsel_set_thermostat(kapsel_ctx *ctx, double kbt, double alpha_v, double alpha_o, unsigned long seed) {
  API_BEGIN
  check_ctx(ctx);
  if (!(kbt >= 0.0) || !(alpha_v >= 0.0) || !(alpha_o >= 0.0)) {
    set_error("kBT, alpha_v and alpha_o must be >= 0");
    throw Error{KAPSEL_ERR_ARG};
  }
  ctx->thermo_kbt = kbt;
  ctx->thermo_alpha_v = alpha_v;
  ctx->thermo_alpha_o = alpha_o;
  ctx->mt.seed((std::mt19937::result_type)(seed & 0xffffffffUL));  // init_genrand (32-bit seed)
  API_END
}

int kapsel_thermostat(kapsel_ctx *ctx, const kapsel_time *jikan) {
  API_BEGIN
  check_ctx(ctx);
  check_time(jikan);
  thermostat(ctx, *jikan);
  API_END
}

int kapsel_restart_write(const char *path, const char *template_path, const kapsel_restart_data *data) {
  API_BEGIN
  restart_write(path, template_path, data);
  API_END
}

int kapsel_restart_read(const char *path, kapsel_restart_data *data) {
  API_BEGIN
  restart_read(path, data);
  API_END
}

// host buffers of a context's restart state
struct RestartBuffers {
  std::vector<double> z0, z1, ps, q, c0, c1;
  kapsel_restart_data d{};
  explicit RestartBuffers(kapsel_ctx *ctx) {
    const size_t f = ctx->g.f;
    const int np = ctx->P.n_particles;
    z0.resize(f);
    z1.resize(f);
    ps.resize((size_t)KAPSEL_P_NFIELDS * 3 * (np > 0 ? np : 1));
    q.resize((size_t)8 * (np > 0 ? np : 1));
    d.nx = ctx->g.nx;
    d.ny = ctx->g.ny;
    d.nz = ctx->g.nz;
    d.n_particles = np;
    d.n_spec = ctx->el_set ? 2 : 0;
    d.zeta0 = z0.data();
    d.zeta1 = z1.data();
    d.particles = ps.data();
    d.orientations = q.data();
    if (ctx->el_set) {
      c0.resize(f);
      c1.resize(f);
      d.conc0 = c0.data();
      d.conc1 = c1.data();
    }
  }
};

int kapsel_save_restart(kapsel_ctx *ctx, const char *path, const kapsel_time *jikan, const char *template_path) {
  API_BEGIN
  check_ctx_ro(ctx);
  if (!jikan) {
    set_error("null jikan");
    throw Error{KAPSEL_ERR_ARG};
  }
  RestartBuffers b(ctx);
  download(ctx, b.d.zeta0, fld(ctx, KAPSEL_ZETA0));
  download(ctx, b.d.zeta1, fld(ctx, KAPSEL_ZETA1));
  KAPSEL_HIP(hipMemcpyAsync(b.d.uk_dc, ctx->d_uk_dc, 3 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  const int np = ctx->P.n_particles;
  if (np > 0) {
    KAPSEL_HIP(hipMemcpyAsync(b.d.particles, ctx->d_ps, sizeof(double) * KAPSEL_P_NFIELDS * 3 * np,
                              hipMemcpyDeviceToHost, ctx->stream));
    KAPSEL_HIP(hipMemcpyAsync(b.d.orientations, ctx->d_q, sizeof(double) * 8 * np, hipMemcpyDeviceToHost,
                              ctx->stream));
  }
  if (ctx->el_set) {
    download(ctx, b.d.conc0, fld(ctx, KAPSEL_CONC0));
    download(ctx, b.d.conc1, fld(ctx, KAPSEL_CONC1));
  }
  KAPSEL_HIP(hipStreamSynchronize(ctx->stream));
  b.d.ts = jikan->ts;
  b.d.time = jikan->time;
  restart_write(path, template_path, &b.d);
  API_END
}

int kapsel_load_restart(kapsel_ctx *ctx, const char *path, kapsel_time *jikan) {
  API_BEGIN
  check_ctx(ctx);
  if (!jikan) {
    set_error("null jikan");
    throw Error{KAPSEL_ERR_ARG};
  }
  RestartBuffers b(ctx);
  restart_read(path, &b.d);
  const double L[3] = {ctx->g.lx, ctx->g.ly, ctx->g.lz};
  const int np = ctx->P.n_particles;
  for (int d = 0; d < 3; d++)
    for (int i = 0; i < np; i++) {
      const double x = b.d.particles[((size_t)KAPSEL_P_X * 3 + d) * np + i];
      if (!(x >= 0.0 && x < L[d])) {
        set_error("%s: particle %d position[%d] = %g outside [0, %g)", path, i, d, x, L[d]);
        throw Error{KAPSEL_ERR_ARG};
      }
    }
  upload(ctx, fld(ctx, KAPSEL_ZETA0), b.d.zeta0);
  upload(ctx, fld(ctx, KAPSEL_ZETA1), b.d.zeta1);
  KAPSEL_HIP(hipMemcpyAsync(ctx->d_uk_dc, b.d.uk_dc, 3 * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  if (np > 0) {
    KAPSEL_HIP(hipMemcpyAsync(ctx->d_ps, b.d.particles, sizeof(double) * KAPSEL_P_NFIELDS * 3 * np,
                              hipMemcpyHostToDevice, ctx->stream));
    KAPSEL_HIP(hipMemcpyAsync(ctx->d_q, b.d.orientations, sizeof(double) * 8 * np, hipMemcpyHostToDevice,
                              ctx->stream));
  }
  if (ctx->el_set) {
    upload(ctx, fld(ctx, KAPSEL_CONC0), b.d.conc0);
    upload(ctx, fld(ctx, KAPSEL_CONC1), b.d.conc1);
    recount_total_solute(ctx);
  }
  KAPSEL_HIP(hipStreamSynchronize(ctx->stream));
  // the Poisson warm starts belong to the state before the load
  ctx->pot64_valid = false;
  ctx->prev_valid = false;
  jikan->ts = b.d.ts;
  jikan->time = b.d.time;
  API_END
}

int kapsel_kernel_timing(kapsel_ctx *ctx, int mask) {
  API_BEGIN
  check_ctx_ro(ctx);
  harvest_timing(ctx);
  ctx->timing_mask = mask;
  API_END
}

int kapsel_kernel_timing_every(kapsel_ctx *ctx, int every) {
  API_BEGIN
  check_ctx_ro(ctx);
  if (every < 1) {
    set_error("kernel timing: every must be >= 1");
    throw Error{KAPSEL_ERR_ARG};
  }
  harvest_timing(ctx);
  ctx->timing_every = every;
  for (unsigned &c : ctx->timing_seen) c = 0;
  API_END
}

int kapsel_kernel_timing_get(kapsel_ctx *ctx, int kernel_class, double *total_ms, int *launches,
                             double *total_bytes) {
  API_BEGIN
  check_ctx_ro(ctx);
  if (kernel_class < 0 || kernel_class >= KAPSEL_K_NCLASSES || !total_ms || !launches || !total_bytes) {
    set_error("bad kernel class or null output");
    throw Error{KAPSEL_ERR_ARG};
  }
  harvest_timing(ctx);
  *total_ms = ctx->timing_ms[kernel_class];
  *launches = ctx->timing_count[kernel_class];
  *total_bytes = ctx->timing_bytes[kernel_class];
  ctx->timing_ms[kernel_class] = 0.0;
  ctx->timing_bytes[kernel_class] = 0.0;
  ctx->timing_count[kernel_class] = 0;
  API_END
}

}  // extern "C"
