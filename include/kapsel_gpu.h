/*
 * kapsel_gpu.h -- C ABI of the MI355X (gfx950) implementation of the KAPSEL
 * smoothed-profile-method (SPM) step.
 *
 * Conventions (all entry points):
 *   - every function returns an int status: KAPSEL_OK (0) or a negative
 *     KAPSEL_ERR_* code; kapsel_last_error() returns the message of the last
 *     failure on the calling thread.  Nothing calls exit().
 *   - a context owns its device memory and one HIP stream (default priority);
 *     it is not thread-safe.  Host buffers stay owned by the caller.
 *   - host field buffers are always double precision and use the reference's
 *     padded Ooura layout a[(i*NY + j)*NZ_ + k], NZ_ = 2*(NZ/2+1): real space
 *     uses k < NZ, reciprocal space stores complex (re, im) pairs at
 *     k = 2*kz, 2*kz+1 for kz = 0..NZ/2 (fft_wrapper.h:681-723,
 *     fft_wrapper.cxx:47-87).  Device fields are double (KAPSEL_FP64) or float
 *     (KAPSEL_FP32) in the same layout.
 *   - particle state is structure-of-arrays doubles:
 *     state[(f*3 + d)*n_particles + n] for field f (enum kapsel_particle_field),
 *     component d, particle n.  It is the reference's Particle AoS
 *     (variable.h:27-77) restricted to what the free-sphere step touches.
 *
 * Every entry point cites the reference function it replaces.
 */
#ifndef KAPSEL_GPU_H
#define KAPSEL_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

#define KAPSEL_OK 0
#define KAPSEL_ERR_ARG (-1)      /* invalid argument */
#define KAPSEL_ERR_HIP (-2)      /* HIP runtime failure */
#define KAPSEL_ERR_STATE (-3)    /* call not valid in the current state */
#define KAPSEL_ERR_NOCONV (-4)   /* iterative solve did not reach the tolerance */
#define KAPSEL_ERR_NODEV (-5)    /* no usable gfx950 device */

enum kapsel_precision {
  KAPSEL_FP64 = 0, /* every field and transform in double */
  KAPSEL_FP32 = 1  /* fields/transforms in float; particle reductions and
                      particle update stay double */
};

/* Device-resident fields (host copies via kapsel_set_field/kapsel_get_field). */
enum kapsel_field {
  KAPSEL_ZETA0 = 0, KAPSEL_ZETA1,          /* reduced vorticity (k-space) */
  KAPSEL_U0, KAPSEL_U1, KAPSEL_U2,         /* fluid velocity */
  KAPSEL_UP0, KAPSEL_UP1, KAPSEL_UP2,      /* particle velocity field phi*u_p */
  KAPSEL_F0, KAPSEL_F1, KAPSEL_F2,         /* constraint force f_particle */
  KAPSEL_PHI, KAPSEL_PHI_SUM,              /* clamped / unclamped profile */
  KAPSEL_EPSILON, KAPSEL_POTENTIAL,        /* permittivity, potential */
  KAPSEL_RHO_FREE, KAPSEL_RHS,             /* free charge, Poisson RHS */
  KAPSEL_WORK0, KAPSEL_WORK1, KAPSEL_WORK2, KAPSEL_WORK3, KAPSEL_WORK4,
  KAPSEL_CONC0, KAPSEL_CONC1,              /* ion concentrations (k-space),
                                              Concentration[0..1] (solute_rhs.cxx:16) */
  KAPSEL_NFIELDS
};

enum kapsel_particle_field {
  KAPSEL_P_X = 0, KAPSEL_P_X_PREVIOUS, KAPSEL_P_X_NOPBC,
  KAPSEL_P_V, KAPSEL_P_V_OLD, KAPSEL_P_OMEGA, KAPSEL_P_OMEGA_OLD,
  KAPSEL_P_F_HYDRO, KAPSEL_P_F_HYDRO_PREVIOUS,
  KAPSEL_P_FR, KAPSEL_P_FR_PREVIOUS,
  KAPSEL_P_TORQUE_HYDRO, KAPSEL_P_TORQUE_HYDRO_PREVIOUS,
  KAPSEL_P_TORQUE_R, KAPSEL_P_TORQUE_R_PREVIOUS,
  KAPSEL_P_NFIELDS
};

/* Preconditioner choice for kapsel_poisson_solve. */
enum kapsel_precond {
  KAPSEL_PRECOND_NONE = 0,     /* plain CG (identity preconditioner) */
  KAPSEL_PRECOND_SPECTRAL = 1  /* inverse of the mean-permittivity spectral Laplacian */
};

/* Values of the reference's globals that the hot path reads (input.h). */
typedef struct kapsel_params {
  int nx, ny, nz;            /* NX, NY, NZ: powers of two, 16..256 */
  double dx;                 /* DX */
  double rho, eta;           /* RHO, ETA (NU = ETA/RHO) */
  int n_particles;           /* Particle_Number (one species) */
  double radius, xi;         /* RADIUS = A*DX, XI = A_XI*DX */
  double mass_ratio;         /* MASS_RATIOS[0] */
  int rotation;              /* ROTATION: 1 binds Angular2v_rot_on */
  int lj_truncate;           /* LJ_truncate: 1 cut at the minimum, 0 at 2.5 sigma, -1 no LJ */
  int lj_powers;             /* LJ_powers: 0 = 12:6, 1 = 24:12, 2 = 36:18 */
  double lj_epsilon;         /* EPSILON */
  double lj_sigma;           /* LJ_dia (SIGMA = 2 RADIUS unless sheared) */
  double body_force[3];      /* constant force added to fr by Force(); the
                                reference's gravity is body_force[G_direction]
                                = -G*RHO*(4/3)pi a^3 (MASS_RATIO-1) */
  int fix_cell[3];           /* FIX_CELLxyz: zero uk_dc[d] each step */
  int precision;             /* enum kapsel_precision */
  double eps_fluid;          /* eps_fluid */
  double eps_particle;       /* eps_particle_top (non-Janus) */
  double dielectric_cst;     /* Dielectric_cst */
} kapsel_params;

/* The reference's CTime (variable.h:18-25). */
typedef struct kapsel_time {
  int ts;
  double time;
  double dt_fluid, hdt_fluid, dt_md, hdt_md;
} kapsel_time;

typedef struct kapsel_ctx kapsel_ctx;

/* ---- lifetime ---------------------------------------------------------- */
/* Replaces Init_fft (fft_wrapper.cxx:166-199), Mem_alloc_var
 * (sp_3d_ns.cxx:295-340) and Particle_domain (profile.cxx:25-74). */
int kapsel_create(const kapsel_params *params, kapsel_ctx **out);
int kapsel_destroy(kapsel_ctx *ctx);
const char *kapsel_last_error(void);
/* Blocks until all work queued on the context's stream has finished. */
int kapsel_synchronize(kapsel_ctx *ctx);
/* Sekibun cell (profile.cxx:25-74): count, and the offsets (3 ints each,
 * reference order) when cells != NULL. */
int kapsel_get_sekibun(kapsel_ctx *ctx, int *np_domain, int *cells);

/* ---- host <-> device --------------------------------------------------- */
int kapsel_set_field(kapsel_ctx *ctx, int field, const double *host);
int kapsel_get_field(kapsel_ctx *ctx, int field, double *host);
int kapsel_set_uk_dc(kapsel_ctx *ctx, const double uk_dc[3]);
int kapsel_get_uk_dc(kapsel_ctx *ctx, double uk_dc[3]);
int kapsel_set_particles(kapsel_ctx *ctx, const double *state);
int kapsel_get_particles(kapsel_ctx *ctx, double *state);
/* Orientation quaternions Particle.q and q_old (variable.h:27-77), scalar
 * part first: q[c*n_particles + i] for c = 0..3, q_old at c = 4..7.
 * kapsel_md_position advances them when rotation is on
 * (MD_solver_orientation_Euler/_AB2, particle_rotation_solver.h:15-39);
 * kapsel_create sets both to the identity. */
int kapsel_set_orientations(kapsel_ctx *ctx, const double *q);
int kapsel_get_orientations(kapsel_ctx *ctx, double *q);

/* Copy_v1_k (aux_field.h:82): dst = src over the whole padded field. */
int kapsel_copy_field(kapsel_ctx *ctx, int dst, int src);

/* ---- FFT and k-space algebra ------------------------------------------- */
/* A2a_k (fft_wrapper.h:681-701): in-place forward r2c, e^{-ik.x}, unscaled. */
int kapsel_fft_r2c(kapsel_ctx *ctx, int field);
/* A_k2a (fft_wrapper.h:704-723): in-place inverse c2r, scaled by 1/(NX NY NZ). */
int kapsel_fft_c2r(kapsel_ctx *ctx, int field);
/* Truncate_two_third_rule_ooura (fft_wrapper.h:797-839). */
int kapsel_truncate(kapsel_ctx *ctx, int field);
/* Zeta_k2u_k (fft_wrapper.cxx:614-711): zeta, uk_dc -> u_k in U0..U2. */
int kapsel_zeta_k2u_k(kapsel_ctx *ctx);
/* U_k2zeta_k (fft_wrapper.cxx:428-497): u_k in U0..U2 -> zeta, uk_dc. */
int kapsel_u_k2zeta_k(kapsel_ctx *ctx);
/* Solenoidal_u (operate_omega.h:109-116) on the real vector field at
 * field0..field0+2. */
int kapsel_solenoidal_u(kapsel_ctx *ctx, int field0);
/* NS_solver_slavedEuler (fluid_solver.cxx:340-366): advances zeta in place. */
int kapsel_ns_slaved_euler(kapsel_ctx *ctx, const kapsel_time *jikan);

/* ---- particle <-> grid --------------------------------------------------- */
/* Make_phi_particle_sum (make_phi.cxx:229-275, 519-523): PHI_SUM, PHI. */
int kapsel_make_phi_particle_sum(kapsel_ctx *ctx);
/* Make_u_particle_sum (make_phi.cxx:329-370, 532-536): UP0..UP2 from PHI_SUM. */
int kapsel_make_u_particle_sum(kapsel_ctx *ctx);
/* Make_phi_particle (make_phi.cxx:146-227, 506-518 with SW_UP = 0): the
 * unclamped profile sum into `field`. */
int kapsel_make_phi_particle(kapsel_ctx *ctx, int field);
/* Calc_f_hydro_correct_precision (md_force.cxx:340-449): f_hydro and,
 * with rotation, torque_hydro from U0..U2 and PHI_SUM. */
int kapsel_calc_f_hydro(kapsel_ctx *ctx, const kapsel_time *jikan);
/* Calc_f_Lennard_Jones_shear_cap_primitive_lnk (md_force.cxx:16-171) plus
 * the body force of Add_f_gravity (md_force.cxx:261-280): Force()
 * (particle_solver.h:92-104), accumulated into fr. */
int kapsel_force(kapsel_ctx *ctx);
/* MD_solver_position_Euler/_AB2 (particle_solver.cxx:39-91), by jikan->ts. */
int kapsel_md_position(kapsel_ctx *ctx, const kapsel_time *jikan);
/* MD_solver_velocity_Euler/_AB2_hydro (particle_solver.cxx:109-150,
 * 252-295) including Force() and reset_Forces(). */
int kapsel_md_velocity(kapsel_ctx *ctx, const kapsel_time *jikan);
/* Make_f_particle_dt_sole + Add_f_particle (f_particle.h:37-95):
 * F = UP - PHI*U, projected, then U += F. */
int kapsel_make_f_particle_dt_sole(kapsel_ctx *ctx);

/* ---- whole step ---------------------------------------------------------- */
/* Initial rigidity projection of main (sp_3d_ns.cxx:431-447). */
int kapsel_init_rigidity(kapsel_ctx *ctx);
/* Time_evolution_hydro, Navier-Stokes branch (sp_3d_ns.cxx:24-186), then
 * jikan->time += dt_fluid and jikan->ts += 1 (sp_3d_ns.cxx:524-525).
 * Afterwards U0..U2 hold the spectrum of u (U2zeta_k transforms in place).
 * An FP32 context ends the step in k-space (the particle force's projected
 * spectrum is added to the kept velocity spectrum): F0..F2 then hold that
 * force spectrum and WORK0..WORK2 the velocity spectrum before it.
 * kapsel_icep_step and kapsel_icep_electrolyte_step of an FP32 context end
 * the same way (F0..F2 likewise; the kept spectrum in internal scratch). */
int kapsel_spm_step(kapsel_ctx *ctx, kapsel_time *jikan);

/* ---- dielectric Poisson (ICEP) ------------------------------------------- */
/* Make_dielectric_field, non-Janus branch (operate_dielectric.cxx:270-309,
 * Make_phi_sin_primitive :78-120): EPSILON = eps_f Dc + (eps_p - eps_f) Dc
 * sum_i phi_sin_i, the compact-sine sum left in WORK4. */
int kapsel_dielectric_field(kapsel_ctx *ctx);
/* Ax (operate_dielectric.cxx:485-522) = D E G x on device fields (padded
 * layout), E the face average of EPSILON; x_field != ax_field. */
int kapsel_poisson_apply(kapsel_ctx *ctx, int x_field, int ax_field);
/* Ax with the reference's Eigen-side signature: unpadded NX*NY*NZ host
 * vectors (MatrixReplacement mv_function, LinearOperator.h:32). */
int kapsel_poisson_apply_host(kapsel_ctx *ctx, const double *x, double *ax);
/* Build_rhs (operate_dielectric.cxx:524-540): RHS = D(eps_face E0) - RHO_FREE. */
int kapsel_poisson_rhs(kapsel_ctx *ctx, const double e0[3]);
/* Charge_field2potential_dielectric_LO (operate_dielectric.cxx:542-572):
 * builds RHS, solves A POTENTIAL = RHS warm-started from POTENTIAL until
 * ||RHS - A x|| / ||RHS|| < tol, at most maxit iterations, and re-solves
 * from zero when maxit is reached (:563-570).  Conjugate gradients on -A
 * (the reference uses BiCGSTAB; A is symmetric negative semidefinite).
 * precond: enum kapsel_precond.  An FP32 context refines in FP64: *relres
 * is always an FP64 residual.  Returns KAPSEL_ERR_NOCONV (with *iters and
 * *relres set) when tol is not reached. */
int kapsel_poisson_solve(kapsel_ctx *ctx, const double e0[3], double tol, int maxit, int precond, int *iters,
                         double *relres);

/* Sets eps_fluid, eps_particle_top and Dielectric_cst (input.cxx:688-736)
 * for the following kapsel_dielectric_field calls. */
int kapsel_set_dielectric(kapsel_ctx *ctx, double eps_fluid, double eps_particle, double dielectric_cst);
/* Make_Maxwell_force_x_on_fluid after the solve (operate_dielectric.cxx:646-763,
 * AC off, External_field on): staggered gradients of EPSILON and POTENTIAL,
 * face force rho E - E^2 grad(eps) / 2, averaged to the cell centres into
 * F0..F2 (scratch: WORK0..WORK3). */
int kapsel_maxwell_force(kapsel_ctx *ctx, const double e0[3]);
/* One ICEP step without ions: kapsel_spm_step's sequence with, after the
 * profile scatter, kapsel_dielectric_field, kapsel_poisson_solve (warm start:
 * the potential of the previous step), kapsel_maxwell_force, u += f dt/rho
 * and Solenoidal_u(u) -- the dielectric part of the Electrolyte+Dielectric
 * branch of Time_evolution_hydro (sp_3d_ns.cxx:55-95).  Returns
 * KAPSEL_ERR_NOCONV (step completed) when the solve missed tol. */
int kapsel_icep_step(kapsel_ctx *ctx, kapsel_time *jikan, const double e0[3], double tol, int maxit, int precond,
                     int *iters, double *relres);

/* Whether kapsel_icep_step queues its Navier-Stokes part (slaved step and
 * Zeta_k2u) on the context's second stream beside the particle and Poisson
 * part (1, the default) or runs everything in order on one stream (0).  The
 * results are the same; no reference counterpart (the reference runs its
 * OpenMP loops in order). */
int kapsel_set_concurrency(kapsel_ctx *ctx, int on);

/* ---- learned warm start (no counterpart in the reference) ----------------- */
/* A 3-level periodic convolutional encoder-decoder (182,083 parameters) maps
 * (phi_sin, beta phi_sin) -- the compact-sine sum left in WORK4 by
 * kapsel_dielectric_field and the Clausius-Mossotti factor
 * beta = (eps_p - eps_f) / (eps_p + 2 eps_f) -- to three potentials psi_d,
 * the response to a unit external field along d.  BF16 weights and
 * activations on MFMA, FP32 accumulation, FP32 master weights and Adam.
 * Needs a cubic grid with N divisible by 4. */
int kapsel_pnet_init(kapsel_ctx *ctx, long *n_params);
/* parameters: FP32, per layer W[tap][cin][cout] then b[cout] (layer order
 * e1a e1b e2a e2b e3a e3b d2 d1 out) */
int kapsel_pnet_set_params(kapsel_ctx *ctx, const float *params);
int kapsel_pnet_get_params(kapsel_ctx *ctx, float *params);
/* loss = mean_v (sum_d e_d psi_d - (POTENTIAL - mean) / scale)^2 and its
 * gradient (grad may be NULL), no update; e_hat a unit vector */
int kapsel_pnet_loss_grad(kapsel_ctx *ctx, double beta, const double e_hat[3], double scale, double *loss,
                          float *grad);
/* one Adam step on that loss (learning rate lr) */
int kapsel_pnet_train_step(kapsel_ctx *ctx, double beta, const double e_hat[3], double scale, double lr,
                           double *loss);
/* POTENTIAL = scale * sum_d e0_d psi_d: the warm start of kapsel_poisson_solve */
int kapsel_pnet_warm_start(kapsel_ctx *ctx, double beta, const double e0[3], double scale);
/* Rescales the warm start in POTENTIAL by gamma = (x . b) / (x . A x), b the
 * RHS for e0 (built into RHS; WORK0 scratch): the multiple of x nearest the
 * solution in the energy norm of -A, never worse than a zero start in that
 * norm.  No reference counterpart (the reference starts from the previous
 * potential, operate_dielectric.cxx:558-562).  *gamma may be NULL. */
int kapsel_poisson_galerkin_start(kapsel_ctx *ctx, const double e0[3], double *gamma);

/* ---- electrolyte: two ion species, Nernst-Planck, Coulomb force ----------- */
/* The globals the Electrolyte equation reads (input.cxx:249-265,
 * operate_electrolyte.cxx:10-14), for N_spec = 2 ("salt") without
 * Dielectric: species 0 is the positive ion, species 1 the negative ion. */
typedef struct kapsel_electrolyte {
  double kbt;                /* kBT */
  double elementary_charge;  /* Elementary_charge */
  double dielectric_cst;     /* Dielectric_cst */
  double surface_charge;     /* Surface_charge[0]: charge of one sphere, in units of
                                Elementary_charge (nonzero) */
  double valency[2];         /* Valency_positive_ion (> 0), Valency_negative_ion (< 0) */
  double onsager_coeff[2];   /* Onsager_coeff_positive_ion, Onsager_coeff_negative_ion */
  double debye_length;       /* Debye_length */
  double e_ext[3];           /* E_ext; all zero when External_field is OFF */
  int ac;                    /* AC: the applied field is E_ext sin(angular_frequency * time) */
  double angular_frequency;  /* Angular_Frequency (2 pi Frequency) */
  int dielectric;            /* Dielectric ON: the ions live in the variable permittivity of
                                kapsel_params (eps_fluid, eps_particle); surface_charge may be 0,
                                and then debye_length is in units of RADIUS (Init_rho_ion_dielectric,
                                operate_electrolyte.cxx:864-979) */
} kapsel_electrolyte;

/* Sets the electrolyte parameters; checks of Init_rho_ion
 * (operate_electrolyte.cxx:740-745, 820-827). */
int kapsel_set_electrolyte(kapsel_ctx *ctx, const kapsel_electrolyte *el);
/* Init_rho_ion, N_spec = 2 without Poisson_Boltzmann
 * (operate_electrolyte.cxx:738-861): Set_uniform_ion_charge_density_salt
 * (:559-621) fills CONC0/CONC1 (k-space) from the particle positions, then
 * Count_solute_each (solute_rhs.h:143-156) sets the conserved totals. */
int kapsel_init_rho_ion(kapsel_ctx *ctx);
/* Init_rho_ion with Poisson_Boltzmann, N_spec = 2 (operate_electrolyte.cxx:
 * 820-833, 860).  One particle: Set_Poisson_Boltzmann_ion_charge_density_salt
 * (:623-736), the damped fixed point of the potential (z:z salt only) until
 * its relative change falls below 1e-4 (max_iter <= 0: the reference's
 * 100000).  Several particles: the uniform start above, then
 * Set_steadystate_ion_density (:424-457) -- the ions relax with the fluid at
 * rest (Ion_diffusion_solver_Euler, fluid_solver.cxx:293-316, time step
 * jikan->dt_fluid) until the free energy changes by less than 1e-4
 * (operate_electrolyte.h:22) or max_iter - 1 steps were taken (max_iter <= 0:
 * the reference's 10000).  Then Count_solute_each.  *iters: the iterations
 * taken; free_energy (may be NULL): {total, ideal, electrostatic} at the end. */
int kapsel_init_rho_ion_pb(kapsel_ctx *ctx, const kapsel_time *jikan, int max_iter, int *iters,
                           double free_energy[3]);
/* Calc_free_energy_PB, non-Dielectric branch (operate_electrolyte.cxx:302-421):
 * {total, ideal, electrostatic} of the current ions and particles. */
int kapsel_free_energy_pb(kapsel_ctx *ctx, const kapsel_time *jikan, double free_energy[3]);
/* Total_solute[0..1] (solute_rhs.cxx:13): the amounts Rescale_solute keeps. */
int kapsel_get_total_solute(kapsel_ctx *ctx, double total[2]);
int kapsel_set_total_solute(kapsel_ctx *ctx, const double total[2]);
/* Make_phi_qq_particle (operate_electrolyte.cxx:179-242): unclamped profile
 * sum into phi_field, normalised surface charge density into surface_field. */
int kapsel_make_phi_qq(kapsel_ctx *ctx, int phi_field, int surface_field);
/* Make_surface_normal (make_phi.cxx:17-61) into field0..field0+2. */
int kapsel_make_surface_normal(kapsel_ctx *ctx, int field0);
/* Conc_k2charge_field (operate_electrolyte.cxx:45-74): real-space charge
 * density of the particles and the ions (CONC0/CONC1) into `field`. */
int kapsel_charge_field(kapsel_ctx *ctx, int field);
/* NSsolute_solver_Euler (fluid_solver.cxx:318-338) with Rhs_NS_Nernst_Planck
 * (:246-292), Rhs_NS_solute (:184-229) and the solute_rhs.cxx fluxes
 * (:42-233): one Euler step of zeta and CONC0/CONC1.  Under Dielectric the
 * electric field is the staggered gradient of the POTENTIAL field averaged
 * onto the cells (:261-270, Interpolate_vec_on_normal_grid,
 * operate_dielectric.cxx:32-55). */
int kapsel_nssolute_euler(kapsel_ctx *ctx, const kapsel_time *jikan);
/* Rescale_solute (solute_rhs.cxx:99-128): PHI <- unclamped profile sum,
 * CONC_n *= Total_solute[n] / Count_single_solute; factors may be NULL. */
int kapsel_rescale_solute(kapsel_ctx *ctx, double factor[2]);
/* Make_Coulomb_force_x_on_fluid (operate_electrolyte.cxx:90-177), free
 * particles: the force density into F0..F2. */
int kapsel_coulomb_force(kapsel_ctx *ctx, const kapsel_time *jikan);
/* Time_evolution_hydro, Electrolyte branch without Dielectric
 * (sp_3d_ns.cxx:24-186): NSsolute_solver_Euler, FIX_CELL, Zeta_k2u, particle
 * position update, profile scatter, Rescale_solute, Coulomb force,
 * u += f dt / RHO, Solenoidal_u, hydrodynamic force, particle velocity update
 * (no thermostat in this branch), rigidity constraint, U2zeta_k; then
 * time += dt_fluid, ts += 1.  The reference's first Calc_f_hydro of the
 * branch (:60-68) only fills the diagnostic f_hydro1 and is not repeated. */
int kapsel_electrolyte_step(kapsel_ctx *ctx, kapsel_time *jikan);

/* ---- electrolyte in the variable permittivity (Dielectric ON) ------------ */
/* Init_rho_ion_dielectric (operate_electrolyte.cxx:912-979): with surface
 * charge Init_rho_ion (bulk density times eps_fluid, :597-600); without,
 * Set_uniform_ion_charge_density_salt_dielectric (:864-910), the bulk density
 * kBT Dielectric_cst eps_fluid / (Elementary_charge Debye_length RADIUS)^2 of
 * both species and no counter-ions; then Count_solute_each. */
int kapsel_init_rho_ion_dielectric(kapsel_ctx *ctx);
/* The potential of the ions (and surface charge) in the permittivity:
 * Make_dielectric_field, the free charge (Conc_k2charge_field with surface
 * charge, Conc_k2charge_field_no_surfase_charge without,
 * operate_dielectric.cxx:135-160, 629-640) into RHO_FREE, the solve into
 * POTENTIAL warm-started from it, as Init_potential_dielectric
 * (operate_dielectric.cxx:574-609) and the Maxwell force do.  The applied
 * field is the electrolyte's E_ext (times sin(omega t) under AC).  The
 * right-hand side gets a uniform background that cancels the free charge's
 * total (nonzero once ions sit inside particles), without which the periodic
 * problem has no solution; RHO_FREE itself keeps the total.  Returns
 * KAPSEL_ERR_NOCONV when the relative residual is not below tol. */
int kapsel_dielectric_potential(kapsel_ctx *ctx, const kapsel_time *jikan, double tol, int maxit, int precond,
                                int *iters, double *relres);
/* Time_evolution_hydro, Electrolyte branch with Dielectric (sp_3d_ns.cxx:
 * 20, 24-186): NSsolute_solver_Euler (Dielectric gradient), FIX_CELL,
 * Zeta_k2u, particle position update, profile scatter, the potential as in
 * kapsel_dielectric_potential, Make_Maxwell_force_x_on_fluid
 * (operate_dielectric.cxx:610-760), u += f dt / RHO, Solenoidal_u,
 * hydrodynamic force, particle velocity update, rigidity constraint,
 * U2zeta_k; no Rescale_solute.  Needs Dielectric_cst equal in the electrolyte
 * and kapsel_set_dielectric. */
int kapsel_icep_electrolyte_step(kapsel_ctx *ctx, kapsel_time *jikan, double tol, int maxit, int precond,
                                 int *iters, double *relres);

/* ---- thermal fluctuations ------------------------------------------------ */
/* kBT, alpha_v, alpha_o of the Navier_Stokes equation (input.cxx:280-281,
 * 539-540) and the seed of the random stream (MT_seed, fluct.h:34-46:
 * init_genrand of MT19937; the reference seeds 0, sp_3d_ns.cxx:402).  With
 * kbt > 0, kapsel_spm_step applies Add_random_force_thermostat after the
 * velocity update (sp_3d_ns.cxx:155-157); the ICEP and electrolyte steps do
 * not (the reference's Electrolyte branch skips it).  kbt = 0 turns it off. */
int kapsel_set_thermostat(kapsel_ctx *ctx, double kbt, double alpha_v, double alpha_o, unsigned long seed);
/* Add_random_force_thermostat (fluct.cxx:9-40), free particles: Gauss2
 * deviates (fluct.h:52-63) of the context's MT19937 stream, three pairs per
 * particle in particle order, drawn on the host and applied on the device. */
int kapsel_thermostat(kapsel_ctx *ctx, const kapsel_time *jikan);

/* ---- restart files: the reference's UDF restart record -------------------- */
/* Saved state of Save_Restart_udf / Force_restore_parameters (resume.cxx:
 * 10-78, 308-398): zeta and the ion spectra in the padded k-space layout
 * (NX*NY*NZ_ doubles each), uk_dc, the particle state (layout of
 * kapsel_get_particles) and orientations (layout of kapsel_get_orientations),
 * jikan.  Members of the reference's sParticle this library does not keep
 * (f_hydro1, torque_hydro1, slip forces and torques) are written as zero and
 * skipped on reading; x_previous is set to x on reading, as the reference
 * does (resume.cxx:409-410).  Buffers stay owned by the caller. */
typedef struct kapsel_restart_data {
  int nx, ny, nz;              /* grid */
  int n_particles;
  int n_spec;                  /* 0: no ion concentrations, 2: conc0/conc1 */
  int ts;                      /* jikan.ts */
  double time;                 /* jikan.time */
  double uk_dc[3];
  double *zeta0, *zeta1;       /* NX*NY*NZ_ each */
  double *particles;           /* KAPSEL_P_NFIELDS*3*n_particles */
  double *orientations;        /* 8*n_particles */
  double *conc0, *conc1;       /* NX*NY*NZ_ each when n_spec == 2 */
} kapsel_restart_data;
/* Host only (no device needed).  Writes the record resume.CONTINUE.Saved_Data
 * in UDF data syntax (UDF/define.udf:317-338).  template_path == NULL: a
 * file holding only that record; otherwise the template UDF with its resume
 * record replaced (or added before \end{data}), e.g. the run's input UDF. */
int kapsel_restart_write(const char *path, const char *template_path, const kapsel_restart_data *data);
/* Host only.  Reads resume.CONTINUE.Saved_Data of a UDF file into the
 * caller's buffers; the grid, particle count and (for n_spec == 2) the
 * concentrations must match the file. */
int kapsel_restart_read(const char *path, kapsel_restart_data *data);
/* Save_Restart_udf of a context: its zeta, uk_dc, particles, orientations
 * and (once kapsel_set_electrolyte has run) ion concentrations, with jikan. */
int kapsel_save_restart(kapsel_ctx *ctx, const char *path, const kapsel_time *jikan, const char *template_path);
/* Force_restore_parameters into a context: sets zeta, uk_dc, particles,
 * orientations, jikan->ts and jikan->time (the time steps are the caller's)
 * and, for an electrolyte context, the ion concentrations, whose conserved
 * totals (not part of the record) are recounted from them
 * (Count_solute_each, solute_rhs.h:143-156). */
int kapsel_load_restart(kapsel_ctx *ctx, const char *path, kapsel_time *jikan);

/* ---- allocation-free hydrodynamic force ------------------------------------ */
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): Calc_f_hydro_correct_precision as kapsel_calc_f_hydro
 * computes it (same kernels, same summation order), on caller-owned device
 * memory, with no context.  Its scratch (the Sekibun table and the per-slice
 * partial sums) is a static device array of KAPSEL_NOALLOC_SCRATCH_BYTES that
 * each device holds until the process ends; a call allocates and frees
 * nothing.  Every call on one device shares that device's array: calls that
 * overlap on one device race.  Limits: a Sekibun cell of at most
 * KAPSEL_NOALLOC_MAX_CELL_POINTS points (a + xi/2 up to about 23 dx) and
 * 35,498 partial-sum slots, one per particle and slice of its table (the
 * library takes about 1024 / n_particles slices, at least one, so up to 35,498
 * particles); beyond them the call returns KAPSEL_ERR_ARG.  Devices 0 ..
 * KAPSEL_NOALLOC_MAX_DEVICES-1, the calling thread's current device.
 *   params: grid, dx, rho, radius, xi, rotation, n_particles; precision
 *           gives the field type (double for KAPSEL_FP64, float for FP32)
 *   ps:     device particle state, layout of kapsel_get_particles (doubles);
 *           reads X, V, OMEGA, writes F_HYDRO and (rotation) TORQUE_HYDRO
 *   u0..u2, phi_sum: device fields in the padded layout
 *   stream: a hipStream_t, NULL for the null stream; the call only queues work */
#define KAPSEL_NOALLOC_SCRATCH_BYTES (4u * 1024u * 1024u)
#define KAPSEL_NOALLOC_MAX_CELL_POINTS 65536
#define KAPSEL_NOALLOC_MAX_DEVICES 64
int kapsel_calc_f_hydro_noalloc(const kapsel_params *params, const kapsel_time *jikan, double *ps,
                                const void *u0, const void *u1, const void *u2, const void *phi_sum,
                                void *stream);

/* ---- kernel timing (HIP events on the context stream) --------------------- */
enum kapsel_kernel_class {
  KAPSEL_K_FFT_Z_R2C = 0, /* z-direction r2c pass (one kernel) */
  KAPSEL_K_FFT_Z_C2R,     /* z-direction c2r pass (one kernel) */
  KAPSEL_K_FFT_LINES_FWD, /* forward c2c line pass, y and x (one kernel on a cubic grid) */
  KAPSEL_K_FFT_LINES_INV, /* inverse c2c line pass, y and x (one kernel on a cubic grid) */
  KAPSEL_K_POINTWISE,     /* k-space and real-space pointwise kernels */
  KAPSEL_K_PARTICLE,      /* scatter, gather, LJ, particle update */
  KAPSEL_K_NETWORK,       /* warm-start network kernels */
  KAPSEL_K_REFINE64,      /* FP64 work of an FP32 context's iterative refinement */
  KAPSEL_K_FFT_PLANE_FACE, /* inverse z+y, face-permittivity product, forward z+y of the operator (one kernel) */
  KAPSEL_K_FFT_X_GRAD,     /* CG direction update + gradient symbols + x inverse pass (one kernel) */
  KAPSEL_K_FFT_X_DIV,      /* x forward pass + divergence symbols + <p, Ap> partials (one kernel) */
  KAPSEL_K_NCLASSES
};
/* Records a HIP event pair around every launch of the kernel classes whose
 * bit is set in `mask` (0 disables); kapsel_kernel_timing_get returns the
 * accumulated milliseconds, launch count and algorithmic bytes (each input
 * read once, each output written once) of one class and resets them.
 * No reference counterpart: the reference only prints wall-clock block
 * times (sp_3d_ns.cxx:506-518). */
int kapsel_kernel_timing(kapsel_ctx *ctx, int mask);
int kapsel_kernel_timing_get(kapsel_ctx *ctx, int kernel_class, double *total_ms, int *launches,
                             double *total_bytes);
/* Times one launch in `every` (>= 1; 1, the default, times all) of each class
 * selected by kapsel_kernel_timing: an event pair delays the launch it
 * brackets by the dispatch of its markers, so a sample perturbs the timed
 * run less.  No reference counterpart. */
int kapsel_kernel_timing_every(kapsel_ctx *ctx, int every);

#ifdef __cplusplus
}
#endif
#endif /* KAPSEL_GPU_H */
