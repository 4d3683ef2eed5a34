// reference_shims.hpp -- the library behind the reference's own operator
// signatures, for drop-in use at the reference's call sites.
//
// The reference (no FFI layer) calls free C++ functions on raw double arrays
// in the padded Ooura layout, with its parameters in globals.  Each function
// below keeps the reference's name and signature and forwards to the C ABI of
// kapsel_gpu.h on the context given to kapsel_shims_attach: the host arrays
// go to the device, the operator runs there, the results come back (one
// device round trip per call; the device-resident binding of the whole step
// is kapsel_spm_step / kapsel_icep_step).  A maintainer compiles this unit
// into the reference with the reference's own variable.h in place of the
// layout declarations below (same members, same order).
//
// Differences the caller sees:
//   - Make_phi_particle_sum and Make_u_particle_sum overwrite their outputs;
//     the reference adds into arrays its callers have just reset
//     (sp_3d_ns.cxx:50-53, 167-168), so the result is the same;
//   - a failing call reports through the error handler (default: message on
//     stderr, exit(EXIT_FAILURE), which is what the reference's exit_job does).
#pragma once

#include "../../include/kapsel_gpu.h"

#ifndef DIM
#define DIM 3
#endif

// CTime (variable.h:18-25)
typedef struct CTime {
  int ts;
  double time;
  double dt_fluid;
  double hdt_fluid;
  double dt_md;
  double hdt_md;
} CTime;

// quaternion (quaternion.h:23-26)
typedef struct quaternion {
  double s;
  double v[DIM];
} quaternion;

// Particle (variable.h:27-77): the layout the reference's arrays have
typedef struct Particle {
  int spec;
  double x[DIM];
  double x_previous[DIM];
  double x_nopbc[DIM];

  double v[DIM];
  double v_old[DIM];
  double v_slip[DIM];

  double f_hydro[DIM];
  double f_hydro_previous[DIM];
  double f_hydro1[DIM];
  double f_slip[DIM];
  double f_slip_previous[DIM];

  double fr[DIM];
  double fr_previous[DIM];
  double torque_r[DIM];
  double torque_r_previous[DIM];

  double omega[DIM];
  double omega_old[DIM];
  double omega_slip[DIM];

  double torque_hydro[DIM];
  double torque_hydro_previous[DIM];
  double torque_hydro1[DIM];
  double torque_slip[DIM];
  double torque_slip_previous[DIM];

  double momentum_depend_fr[DIM];

  double QR[DIM][DIM];
  double QR_old[DIM][DIM];

  quaternion q;
  quaternion q_old;

  double mass;
  double mass_center[DIM];
  double inertia[DIM][DIM];

  double surface_mass;
  double surface_mass_center[DIM];
  double surface_inertia[DIM][DIM];

  double surface_dv[DIM];
  double surface_dw[DIM];
} Particle;

// The reference globals the shims read, handed over once after Init_fft
// (sp_3d_ns.cxx:404) together with the context: Particle_Number, ROTATION,
// Dielectric_cst, Tol_potential and MaxIter_potential (input.h:294-295), the
// global permittivity field `epsilon` that Ax reads
// (operate_dielectric.cxx:499-502), and the preconditioner of the solve
// (KAPSEL_PRECOND_NONE: the reference's identity preconditioner).
void kapsel_shims_attach(kapsel_ctx *ctx, int particle_number, int rotation, double dielectric_cst,
                         double tol_potential, int maxiter_potential, const double *epsilon,
                         int precond = KAPSEL_PRECOND_NONE);
// Called with the library's message when a shim's call fails; the default
// prints it and calls exit(EXIT_FAILURE).
void kapsel_shims_set_error_handler(void (*handler)(const char *message));

// Particle AoS <-> the library's structure-of-arrays state (kapsel_gpu.h)
void kapsel_shims_pack(const Particle *p, int n, double *state);
void kapsel_shims_unpack(const double *state, int n, Particle *p);
// Particle.q / q_old <-> the library's orientations (q[c*n + i], q_old at c >= 4)
void kapsel_shims_pack_orientations(const Particle *p, int n, double *q);
void kapsel_shims_unpack_orientations(const double *q, int n, Particle *p);

// fft_wrapper.h:681 / :704
void A2a_k(double *a);
void A_k2a(double *a);
// make_phi.h:51, :56
void Make_phi_particle_sum(double *phi, double *phi_sum, Particle *p, const double radius);
void Make_u_particle_sum(double **up, double const *phi_sum, Particle *p, const double radius);
// md_force.h:55
void Calc_f_hydro_correct_precision(Particle *p, double const *phi_sum, double const *const *u, const CTime &jikan);
// operate_dielectric.h:126 (non-Janus branch: eps_p_bottom is not read)
void Make_dielectric_field(Particle *p, double *eps, double *phi_p, const double eps_p_top, const double eps_p_bottom,
                           const double eps_f);
// operate_dielectric.h:193 -- also the MatrixReplacement mv_function (LinearOperator.h:32)
void Ax(const double *x, double *Ax);
// operate_dielectric.h:207
void Build_rhs(double *rhs, double *free_charge_density, double *eps, double external_e_field[]);
// operate_dielectric.h:221
void Charge_field2potential_dielectric_LO(double *free_charge_density, double *potential, double *eps, double *rhs,
                                          double external_e_field[]);
