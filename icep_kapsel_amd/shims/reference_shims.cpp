// reference_shims.cpp -- the reference's operator signatures over the C ABI
// (declarations and conventions in reference_shims.hpp).
#include "reference_shims.hpp"

#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace {

struct ShimState {
  kapsel_ctx *ctx = nullptr;
  int n = 0;                       // Particle_Number
  double tol = 1e-6;               // Tol_potential
  int maxit = 1000;                // MaxIter_potential
  const double *epsilon = nullptr; // the reference's global permittivity field
  int rotation = 1;                // ROTATION
  double dielectric_cst = 1.0;     // Dielectric_cst
  int precond = KAPSEL_PRECOND_NONE;
  void (*on_error)(const char *) = nullptr;
};
ShimState S;

void fail(const char *what) {
  char msg[512];
  snprintf(msg, sizeof msg, "%s: %s", what, kapsel_last_error());
  if (S.on_error) {
    S.on_error(msg);
    return;
  }
  fprintf(stderr, "%s\n", msg);
  exit(EXIT_FAILURE);
}

#define KG(call)                       \
  do {                                 \
    if ((call) != KAPSEL_OK) fail(#call); \
  } while (0)

// the library's SoA particle fields and the Particle members they mirror
struct Member {
  int field;
  size_t offset;
};
const Member kMembers[] = {
    {KAPSEL_P_X, offsetof(Particle, x)},
    {KAPSEL_P_X_PREVIOUS, offsetof(Particle, x_previous)},
    {KAPSEL_P_X_NOPBC, offsetof(Particle, x_nopbc)},
    {KAPSEL_P_V, offsetof(Particle, v)},
    {KAPSEL_P_V_OLD, offsetof(Particle, v_old)},
    {KAPSEL_P_OMEGA, offsetof(Particle, omega)},
    {KAPSEL_P_OMEGA_OLD, offsetof(Particle, omega_old)},
    {KAPSEL_P_F_HYDRO, offsetof(Particle, f_hydro)},
    {KAPSEL_P_F_HYDRO_PREVIOUS, offsetof(Particle, f_hydro_previous)},
    {KAPSEL_P_FR, offsetof(Particle, fr)},
    {KAPSEL_P_FR_PREVIOUS, offsetof(Particle, fr_previous)},
    {KAPSEL_P_TORQUE_HYDRO, offsetof(Particle, torque_hydro)},
    {KAPSEL_P_TORQUE_HYDRO_PREVIOUS, offsetof(Particle, torque_hydro_previous)},
    {KAPSEL_P_TORQUE_R, offsetof(Particle, torque_r)},
    {KAPSEL_P_TORQUE_R_PREVIOUS, offsetof(Particle, torque_r_previous)},
};
static_assert(sizeof(kMembers) / sizeof(kMembers[0]) == KAPSEL_P_NFIELDS, "every library particle field");

const double *member(const Particle &p, size_t off) {
  return reinterpret_cast<const double *>(reinterpret_cast<const char *>(&p) + off);
}
double *member(Particle &p, size_t off) { return reinterpret_cast<double *>(reinterpret_cast<char *>(&p) + off); }

void upload_particles(const Particle *p) {
  std::vector<double> s((size_t)KAPSEL_P_NFIELDS * 3 * S.n);
  kapsel_shims_pack(p, S.n, s.data());
  KG(kapsel_set_particles(S.ctx, s.data()));
}

void download_particles(Particle *p) {
  std::vector<double> s((size_t)KAPSEL_P_NFIELDS * 3 * S.n);
  KG(kapsel_get_particles(S.ctx, s.data()));
  kapsel_shims_unpack(s.data(), S.n, p);
}

kapsel_time time_of(const CTime &t) {
  kapsel_time k;
  k.ts = t.ts;
  k.time = t.time;
  k.dt_fluid = t.dt_fluid;
  k.hdt_fluid = t.hdt_fluid;
  k.dt_md = t.dt_md;
  k.hdt_md = t.hdt_md;
  return k;
}

void check_attached(const char *who) {
  if (!S.ctx) {
    char msg[160];
    snprintf(msg, sizeof msg, "%s: kapsel_shims_attach has not run", who);
    if (S.on_error) S.on_error(msg);
    else {
      fprintf(stderr, "%s\n", msg);
      exit(EXIT_FAILURE);
    }
  }
}

}  // namespace

void kapsel_shims_attach(kapsel_ctx *ctx, int particle_number, int rotation, double dielectric_cst,
                         double tol_potential, int maxiter_potential, const double *epsilon, int precond) {
  S.ctx = ctx;
  S.n = particle_number;
  S.rotation = rotation;
  S.dielectric_cst = dielectric_cst;
  S.tol = tol_potential;
  S.maxit = maxiter_potential;
  S.epsilon = epsilon;
  S.precond = precond;
}

void kapsel_shims_set_error_handler(void (*handler)(const char *message)) { S.on_error = handler; }

void kapsel_shims_pack(const Particle *p, int n, double *state) {
  for (const Member &m : kMembers)
    for (int i = 0; i < n; i++) {
      const double *v = member(p[i], m.offset);
      for (int d = 0; d < DIM; d++) state[((size_t)m.field * 3 + d) * n + i] = v[d];
    }
}

void kapsel_shims_unpack(const double *state, int n, Particle *p) {
  for (const Member &m : kMembers)
    for (int i = 0; i < n; i++) {
      double *v = member(p[i], m.offset);
      for (int d = 0; d < DIM; d++) v[d] = state[((size_t)m.field * 3 + d) * n + i];
    }
}

void kapsel_shims_pack_orientations(const Particle *p, int n, double *q) {
  for (int i = 0; i < n; i++) {
    q[i] = p[i].q.s;
    q[(size_t)4 * n + i] = p[i].q_old.s;
    for (int c = 0; c < 3; c++) {
      q[(size_t)(1 + c) * n + i] = p[i].q.v[c];
      q[(size_t)(5 + c) * n + i] = p[i].q_old.v[c];
    }
  }
}

void kapsel_shims_unpack_orientations(const double *q, int n, Particle *p) {
  for (int i = 0; i < n; i++) {
    p[i].q.s = q[i];
    p[i].q_old.s = q[(size_t)4 * n + i];
    for (int c = 0; c < 3; c++) {
      p[i].q.v[c] = q[(size_t)(1 + c) * n + i];
      p[i].q_old.v[c] = q[(size_t)(5 + c) * n + i];
    }
  }
}

void A2a_k(double *a) {
  check_attached("A2a_k");
  KG(kapsel_set_field(S.ctx, KAPSEL_WORK0, a));
  KG(kapsel_fft_r2c(S.ctx, KAPSEL_WORK0));
  KG(kapsel_get_field(S.ctx, KAPSEL_WORK0, a));
}

void A_k2a(double *a) {
  check_attached("A_k2a");
  KG(kapsel_set_field(S.ctx, KAPSEL_WORK0, a));
  KG(kapsel_fft_c2r(S.ctx, KAPSEL_WORK0));
  KG(kapsel_get_field(S.ctx, KAPSEL_WORK0, a));
}

void Make_phi_particle_sum(double *phi, double *phi_sum, Particle *p, const double radius) {
  (void)radius;  // the context's RADIUS: the reference passes its default, RADIUS
  check_attached("Make_phi_particle_sum");
  upload_particles(p);
  KG(kapsel_make_phi_particle_sum(S.ctx));
  KG(kapsel_get_field(S.ctx, KAPSEL_PHI_SUM, phi_sum));
  KG(kapsel_get_field(S.ctx, KAPSEL_PHI, phi));
}

void Make_u_particle_sum(double **up, double const *phi_sum, Particle *p, const double radius) {
  (void)radius;
  check_attached("Make_u_particle_sum");
  upload_particles(p);
  KG(kapsel_set_field(S.ctx, KAPSEL_PHI_SUM, phi_sum));
  KG(kapsel_make_u_particle_sum(S.ctx));
  for (int d = 0; d < DIM; d++) KG(kapsel_get_field(S.ctx, KAPSEL_UP0 + d, up[d]));
}

void Calc_f_hydro_correct_precision(Particle *p, double const *phi_sum, double const *const *u, const CTime &jikan) {
  check_attached("Calc_f_hydro_correct_precision");
  upload_particles(p);
  KG(kapsel_set_field(S.ctx, KAPSEL_PHI_SUM, phi_sum));
  for (int d = 0; d < DIM; d++) KG(kapsel_set_field(S.ctx, KAPSEL_U0 + d, u[d]));
  const kapsel_time t = time_of(jikan);
  KG(kapsel_calc_f_hydro(S.ctx, &t));
  download_particles(p);
  // md_force.cxx:429-437: the slip force (and, with ROTATION, the slip
  // torque) is zeroed beside the hydrodynamic one; the library keeps no slip
  // members.  Without ROTATION torque_hydro comes back as it went in.
  for (int i = 0; i < S.n; i++)
    for (int d = 0; d < DIM; d++) {
      p[i].f_slip[d] = 0.0;
      if (S.rotation) p[i].torque_slip[d] = 0.0;
    }
}

void Make_dielectric_field(Particle *p, double *eps, double *phi_p, const double eps_p_top, const double eps_p_bottom,
                           const double eps_f) {
  (void)eps_p_bottom;  // Janus particles are not covered
  check_attached("Make_dielectric_field");
  upload_particles(p);
  KG(kapsel_set_dielectric(S.ctx, eps_f, eps_p_top, S.dielectric_cst));
  KG(kapsel_dielectric_field(S.ctx));
  KG(kapsel_get_field(S.ctx, KAPSEL_EPSILON, eps));
  KG(kapsel_get_field(S.ctx, KAPSEL_WORK4, phi_p));
}

void Ax(const double *x, double *ax) {
  check_attached("Ax");
  // Ax reads the global epsilon (operate_dielectric.cxx:499-502)
  if (S.epsilon) KG(kapsel_set_field(S.ctx, KAPSEL_EPSILON, S.epsilon));
  KG(kapsel_poisson_apply_host(S.ctx, x, ax));
}

void Build_rhs(double *rhs, double *free_charge_density, double *eps, double external_e_field[]) {
  check_attached("Build_rhs");
  KG(kapsel_set_field(S.ctx, KAPSEL_RHO_FREE, free_charge_density));
  KG(kapsel_set_field(S.ctx, KAPSEL_EPSILON, eps));
  KG(kapsel_poisson_rhs(S.ctx, external_e_field));
  KG(kapsel_get_field(S.ctx, KAPSEL_RHS, rhs));
}

void Charge_field2potential_dielectric_LO(double *free_charge_density, double *potential, double *eps, double *rhs,
                                          double external_e_field[]) {
  check_attached("Charge_field2potential_dielectric_LO");
  // the reference builds b from eps and applies A with the global epsilon;
  // its callers pass the global as eps (operate_dielectric.cxx:600, 604, 635,
  // 641), so both come from eps here
  KG(kapsel_set_field(S.ctx, KAPSEL_RHO_FREE, free_charge_density));
  KG(kapsel_set_field(S.ctx, KAPSEL_EPSILON, eps));
  KG(kapsel_set_field(S.ctx, KAPSEL_POTENTIAL, potential));
  int iters = 0;
  double relres = 0.0;
  const int rc = kapsel_poisson_solve(S.ctx, external_e_field, S.tol, S.maxit, S.precond, &iters, &relres);
  if (rc != KAPSEL_OK && rc != KAPSEL_ERR_NOCONV) fail("kapsel_poisson_solve");
  KG(kapsel_get_field(S.ctx, KAPSEL_POTENTIAL, potential));
  KG(kapsel_get_field(S.ctx, KAPSEL_RHS, rhs));
}
