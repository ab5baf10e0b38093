/*
 * fsibe_hip.h -- C ABI of libfsibe_hip.so, the MI355X (gfx950) implementation
 * of the FSIBE_2D fluid time step.
 *
 * Conventions shared by every entry point
 *  - Fields are IEEE binary64 arrays of nx*ny values in the reference's order
 *    real(8) f(0:nx-1,0:ny-1), x fastest (reference NavierStokes.f90:19-28).
 *  - Spectral fields in the public API use the reference's packed real layout
 *    produced by coftxy (cof_fftw33.f90:142-193): four reals per (kx,ky) block,
 *    both Nyquist modes dropped.
 *  - Every call returns 0 on success or an FSIBE_E* code; fsibe_last_error()
 *    gives the message of the last failure on the calling thread.  The
 *    reference instead prints and STOPs (e.g. Interpolation.f90:19-24); the
 *    Fortran binding (fsibe_2d_amd/fortran, fsibe_check) turns a nonzero
 *    code into write+stop.
 *  - Entry points without a _dev_ infix take host pointers, copy in, compute on
 *    the GPU and copy out before returning.  _dev_ entry points take device
 *    pointers and are asynchronous on the handle's stream.
 *  - A handle must not be used from two threads at once (the reference is not
 *    reentrant either: module globals and SAVE variables).
 */
#ifndef FSIBE_HIP_H
#define FSIBE_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  FSIBE_OK = 0,
  FSIBE_EARG = 1,         /* null pointer or invalid argument */
  FSIBE_ESIZE = 2,        /* grid size not supported (nx, ny even, n/2 and n of the form 2^a 3^b, b<=1) */
  FSIBE_EHIP = 3,         /* HIP runtime error */
  FSIBE_EUNSUPPORTED = 4, /* parameter combination outside the implemented path */
  FSIBE_EDOMAIN = 5,      /* interpolation target outside the field (Interpolation.f90:19-24) */
  FSIBE_ESTATE = 6        /* call sequence error */
};

/* Message of the last failed call on this thread ("" if none). */
const char *fsibe_last_error(void);

/* ========================================================================= */
/* Tier A -- the FFT / spectral-operator plugin.                              */
/* The reference selects its FFT backend at link time (makefile:40,75); every */
/* backend exports fft_initialize, fft_free, coftxy, cofitxy with grid sizes  */
/* taken from module share_vars.  This tier is the same plugin for a host    */
/* that binds it through iso_c_binding (fsibe_2d_amd/fortran/cof_hip.f90).   */
/* ========================================================================= */
typedef struct fsibe_spectral fsibe_spectral;

/* replaces fft_initialize (cof_fftw33.f90:18-66); xl, yl, nu are the
 * share_vars values the operators read (cofdx.f90:17, cal_vis.f90:20) */
int fsibe_spectral_create(int nx, int ny, double xl, double yl, double nu,
                          fsibe_spectral **out);
/* replaces fft_free (cof_fftw33.f90:119-137) */
int fsibe_spectral_destroy(fsibe_spectral *ctx);

/* replaces coftxy(f, fk) (cof_fftw33.f90:142-193) */
int fsibe_coftxy(fsibe_spectral *ctx, const double *f, double *fk);
/* replaces cofitxy(fk, f) (cof_fftw33.f90:197-243) */
int fsibe_cofitxy(fsibe_spectral *ctx, const double *fk, double *f);
/* replace cofdx (cofdx.f90:1-29), cofdy (cofdy.f90:1-30), cofdxdx
 * (cofdxdx.f90:1-27), cofdydy (cofdydy.f90:1-27), cofdxdy (cofdxdy.f90:1-40) */
int fsibe_cofdx(fsibe_spectral *ctx, const double *fk, double *out);
int fsibe_cofdy(fsibe_spectral *ctx, const double *fk, double *out);
int fsibe_cofdxdx(fsibe_spectral *ctx, const double *fk, double *out);
int fsibe_cofdydy(fsibe_spectral *ctx, const double *fk, double *out);
int fsibe_cofdxdy(fsibe_spectral *ctx, const double *fk, double *out);
/* replaces poisson(f, ans) (poisson.f90:1-30) */
int fsibe_poisson(fsibe_spectral *ctx, const double *f, double *ans);
/* replaces vis(dt, work) (cal_vis.f90:1-36) */
int fsibe_vis(fsibe_spectral *ctx, double dt, double *work);
/* replaces dealiase_mask (dealiase_mask.f90:1-33); writes the 0/1 mask */
int fsibe_dealiase_mask(fsibe_spectral *ctx, double *dealiase);

/* replace cofts(f, fk, L, n) and cofits(fk, f, L, n) (cof_fftw33.f90:248-340):
 * n real sequences of length L in columns, r2c scaled by 1/L and truncated
 * to L reals (Re0, Im0, Re1, Im1, ...), and the unnormalised inverse.  These
 * run on the host: their only caller is the host solid solver's curvature
 * filter (SolidSolver.f90:1388,1417) on a few hundred points. */
int fsibe_cofts(const double *f, double *fk, int L, int n);
int fsibe_cofits(const double *fk, double *f, int L, int n);

/* Device-pointer twins of the above, asynchronous on the context's stream
 * (stream = NULL) or on the given hipStream_t. */
int fsibe_dev_coftxy(fsibe_spectral *ctx, const double *d_f, double *d_fk, void *stream);
int fsibe_dev_cofitxy(fsibe_spectral *ctx, const double *d_fk, double *d_f, void *stream);
/* which: 0 cofdx, 1 cofdy, 2 cofdxdx, 3 cofdydy, 4 cofdxdy, 5 poisson */
int fsibe_dev_spectral_op(fsibe_spectral *ctx, int which, const double *d_in, double *d_out,
                          void *stream);
int fsibe_dev_vis(fsibe_spectral *ctx, double dt, double *d_work, void *stream);
int fsibe_spectral_sync(fsibe_spectral *ctx);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): coftxy (cof_fftw33.f90:142-193) of the device field d_f
 * (nx*ny reals) into d_fk (packed layout), bit-identical to
 * fsibe_dev_coftxy, with no context and no allocation.  The intermediate half
 * spectrum and the twiddle tables live in a static device buffer of the
 * library (32 MiB + 128 KiB on each GPU it runs on, held until the process
 * ends): nx*ny <= 2048*2048 and nx, ny <= 4096, else FSIBE_ESIZE.  The first
 * call of a grid size on a device (and the first after a call of another
 * size) uploads the twiddle tables and synchronises `stream` (NULL: the
 * default stream); later calls only launch the two kernels, asynchronously.
 * Calls that overlap on one GPU share the buffer and race. */
int fsibe_dev_coftxy_noalloc(int nx, int ny, const double *d_f, double *d_fk, void *stream);

/* ========================================================================= */
/* Tier B -- device-resident fluid step (the performance path).               */
/* Mirrors the reference's per-step calls (time_step.f90:187,193,199):       */
/*   create_mask(time, beams)                       create_mask.f90:2-218    */
/*   EvolveFluidAB2 / EvolveFluidExplicit           NavierStokes.f90:422-508 */
/*   GetForces(time, beams, press, u, "new")        BeamForces.f90:6-51      */
/* State (vortk, nlk, u, masks, vor_mean) lives in HBM between calls.         */
/* ========================================================================= */

/* Run parameters: the share_vars values the fluid step reads
 * (share_vars.f90:8-25, params.f90:9-81).  U_mean_true as set by
 * time_step.f90:96-100.  heave_* and motion_coef parameterise two motion
 * cases the reference does not have (600: y = heave_amp sin(2 pi heave_freq t);
 * 601: Fourier heave + pitch, see fsibe_mouvement). */
typedef struct fsibe_params {
  int32_t nx, ny, ns;
  int32_t iMeanVelocity, iSponge, iBeam, iWalls, iCylinder, iSharpTrailing;
  int32_t iFLUSI, iMotion, iViscous, iSaveBeam, idealis;
  double xl, yl, nu, eps, eps_sponge, cfl, dt_fixed, Time_end;
  double N_smooth, t_beam, length, SpongeSize, R_cylinder, h_channel;
  double x0, y0, AngleBeam, theta_inf, T_fullspeed, T_release, tau;
  double U_mean_true;
  double heave_amp, heave_freq;
  double motion_coef[8];
  /* the beam's material for the host solid solver (share_vars.f90:17-21):
   * mass ratio, stiffness, damping, gravity, time scheme (1 implicit Euler,
   * 2 Crank-Nicolson, 5 BDF2), gravity impulse case */
  double mue, eta, sigma, grav;
  int32_t TimeMethodSolid, iImpulse;
  /* 0: create_mask's cut-offs (segments and hinges drawn out to
   * t* = t_beam + 2 N_smooth max(dx,dy), where the smooth step is still
   * ~2e-7; cal_drag's window 1.5 t_beam around the beam, inside the mask's
   * smooth zone).  1 (no reference counterpart; for differentiation):
   * segments and hinges drawn out to t_beam + 7 delta, where the smooth
   * step is below the rounding of 1, and the cal_drag window grown to hold
   * every pixel within t_beam + 7 delta + max(dx,dy) of a node, so that the
   * mask and Force are continuous functions of the beam's position. */
  int32_t iSmoothCutoff;
} fsibe_params;

/* Per-step results copied back to the host (type Solid fields filled by
 * cal_drag NavierStokes.f90:567-589 and GetForcesInterp BeamForces.f90:254-270,
 * plus the diagnostics the reference appends to files every step). */
typedef struct fsibe_step_out {
  double dt1, u_max;
  double Force[2], Force_unst[2], Force_press[2];
  double energies[6]; /* vor_rms kinetic dissipation porous_diss inflow_solid inflow_mean (NavierStokes.f90:236-256) */
  double bsi[10];     /* drag lift drag_surf lift_surf Ft1 Fn1 Fp1 Ft2 Fn2 Fp2 (BeamForces.f90:485-496) */
  double fx_int, fy_int; /* BeamForces.f90:279 */
  int32_t status;    /* 1: the solid solver stopped the run (emergency brake) */
  int32_t ibes_iter; /* Newton iterations of the step's solid solve (iFLUSI=1) */
} fsibe_step_out;

typedef struct fsibe_sim fsibe_sim;

/* allocate a simulation (StartSimulation, dns.f90:183-201: dealiasing mask,
 * FFT plans; create_sponge_mask is run by the first fsibe_create_mask) */
int fsibe_sim_create(const fsibe_params *p, fsibe_sim **out);
int fsibe_sim_destroy(fsibe_sim *sim);

/* initial vorticity, packed spectral layout, written into time level n0
 * (init_fields.f90:38-290 decides the values; the host supplies them) */
int fsibe_set_vortk(fsibe_sim *sim, const double *vortk_packed);

/* create_mask(time, beams): upload the beam state (arrays of length ns, the
 * type Solid members x, y, vx, vy, theta) together with the leading-edge
 * kinematics mouvement(time) returns (alpha, LeadingEdge(1:6)); the masks are
 * rebuilt on the GPU.  The same beam state and kinematics are used by the
 * next fsibe_evolve_* (cal_drag) and fsibe_get_forces calls.  For iBeam <= 1;
 * with two beams FSIBE_EARG (use fsibe_create_mask_beams). */
int fsibe_create_mask(fsibe_sim *sim, double time, const double *x, const double *y,
                      const double *vx, const double *vy, const double *theta, double alpha,
                      const double LeadingEdge[6]);
/* create_mask(time, beams) for iBeam beams (create_mask.f90:74-119, a loop
 * over beams(ib) calling mouvement(time, ..., beams(ib)) for each): x .. theta
 * hold iBeam*ns values, beam ib's nodes from ib*ns; alpha[iBeam] and
 * LeadingEdge[6*iBeam] are each beam's mouvement(time).  The beams are drawn
 * in order, each primitive where its value is strictly larger, as the
 * reference draws them. */
int fsibe_create_mask_beams(fsibe_sim *sim, double time, const double *x, const double *y,
                            const double *vx, const double *vy, const double *theta,
                            const double *alpha, const double *LeadingEdge);

/* EvolveFluidExplicit(time, dt1, n0, n1, ...) -- NavierStokes.f90:422-459.
 * Returns dt1 chosen by the CFL logic of cal_nlk (NavierStokes.f90:59-78). */
int fsibe_evolve_explicit(fsibe_sim *sim, double time, double *dt1);
/* EvolveFluidAB2(time, dt0, dt1, n0, n1, ...) -- NavierStokes.f90:463-508 */
int fsibe_evolve_ab2(fsibe_sim *sim, double time, double dt0, double *dt1);
/* GetForces(time, beams, press, u, "new") -- BeamForces.f90:6-288 with the
 * hard-wired SPL=3 Delta interpolation, every beam (:24-47).  pressure_beam,
 * tau_beam: iBeam*ns values (beam ib from ib*ns); Force_press: 2*iBeam values
 * (2 without a beam). */
int fsibe_get_forces(fsibe_sim *sim, double time, double *pressure_beam, double *tau_beam,
                     double Force_press[2]);
/* results of the last evolve / get_forces (cal_drag forces, energies, ...);
 * the per-beam fields (Force, Force_unst, Force_press, bsi[2..9], fx_int,
 * fy_int, ibes_iter) are the first beam's */
int fsibe_get_step_out(fsibe_sim *sim, fsibe_step_out *out);
/* the same with the per-beam fields of beam ib (0 <= ib < iBeam): beams(ib+1)
 * of cal_drag (NavierStokes.f90:528-589) and GetForces (BeamForces.f90:24-47) */
int fsibe_get_step_out_beam(fsibe_sim *sim, int ib, fsibe_step_out *out);
/* swap time levels n0 <-> n1 (time_step.f90:152,222) */
int fsibe_swap_levels(fsibe_sim *sim);

/* Field download.  which: see FSIBE_FIELD_*.  Spectral fields come back in the
 * packed layout; physical fields as nx*ny reals. */
enum {
  FSIBE_FIELD_VORTK = 0, /* vortk(:,:,n0) packed */
  FSIBE_FIELD_NLK = 1,   /* nlk(:,:,n0) packed */
  FSIBE_FIELD_U1 = 3, FSIBE_FIELD_U2 = 4,
  FSIBE_FIELD_PRESS = 5, /* static pressure (computed on request over the whole grid) */
  FSIBE_FIELD_MASK = 6, FSIBE_FIELD_MASKVX = 7, FSIBE_FIELD_MASKVY = 8,
  FSIBE_FIELD_MASK_SPONGE = 9, FSIBE_FIELD_VOR_MEAN = 10,
  FSIBE_FIELD_VORT = 14 /* vorticity of vortk(:,:,n0) in physical space */
};
int fsibe_get_field(fsibe_sim *sim, int which, double *out);

/* Fields kept only on the rectangle of grid points the step needs (the
 * reference keeps them on the whole grid): which 0 static pressure, 1
 * stress_a, 2 stress_b (the Delta-interpolation stencils of GetForcesInterp,
 * BeamForces.f90:152-203), 3 mask, 4 maskvx, 5 maskvy (the support of the
 * beam's mask; zero elsewhere).  box receives x0, x1, y0, y1 (inclusive);
 * out (may be NULL to query the box) receives (x1-x0+1)*(y1-y0+1) values,
 * x fastest. */
int fsibe_get_box(fsibe_sim *sim, int which, double *out, int box[4]);

/* ---- host-side kinematics (mouvement.f90:5-241, integrate_position.f90:1-47,
 * init_beam.f90:48-169) for hosts that do not bring their own ------------- */
int fsibe_mouvement(const fsibe_params *p, int iMouvement, double x0, double y0,
                    double AngleBeam, double phase, double time, double *alpha,
                    double *alpha_t, double *alpha_tt, double LeadingEdge[6]);
int fsibe_integrate_position(const fsibe_params *p, int iMouvement, double time,
                             const double *theta, const double *theta_dot, double *x, double *y,
                             double *vx, double *vy);
/* placement of beam ib (init_beam.f90:44-76): ib 0 at (x0, y0), phase 0; ib 1
 * (iMotion 300..307 only, else FSIBE_EARG as the reference stops) at
 * (x0 + 1.25, y0), phase (iMotion - 300) pi / 4 */
int fsibe_beam_placement(const fsibe_params *p, int ib, double *x0, double *y0, double *phase);
/* integrate_position(time, beams(ib+1)) with iMotion and beam ib's placement */
int fsibe_integrate_position_beam(const fsibe_params *p, int ib, double time,
                                  const double *theta, const double *theta_dot, double *x,
                                  double *y, double *vx, double *vy);

/* ---- host solid solver (SolidSolver.f90) -----------------------------------
 * One beam: the Newton solver IBES_solver (SolidSolver.f90:99-414) for the
 * angle theta and tension T of the inextensible beam, time scheme
 * TimeMethodSolid (1 implicit Euler, 2 Crank-Nicolson, 5 BDF2 starting with
 * one Crank-Nicolson step), material mue, eta, sigma, grav of fsibe_params.
 * A host without the reference's solver (it links MKL) uses this one; the
 * Fortran module fsibe_2d_amd/fortran/solid_hip.f90 wraps it as
 * InitializeSolidSolver / SolidSolverWrapper. */
typedef struct fsibe_solid fsibe_solid;
/* init_beams for one beam (init_beam.f90:80-120: straight beam at the leading
 * edge of mouvement(0), integrate_position(0)) and InitializeSolidSolver
 * (SolidSolver.f90:54-67).  Needs ns >= 8. */
int fsibe_solid_create(const fsibe_params *p, fsibe_solid **out);
/* the same for beam ib at its placement (fsibe_beam_placement) */
int fsibe_solid_create_beam(const fsibe_params *p, int ib, fsibe_solid **out);
int fsibe_solid_destroy(fsibe_solid *b);
/* arrays of type(solid) (share_vars.f90:50-63), ns values each; OLDOLD is
 * beam_oldold(0:ns-1, 1:6), 6*ns values, column-major (x y vx vy theta
 * theta_dot) */
enum {
  FSIBE_SOLID_X = 0, FSIBE_SOLID_Y = 1, FSIBE_SOLID_VX = 2, FSIBE_SOLID_VY = 3, FSIBE_SOLID_THETA = 4,
  FSIBE_SOLID_THETA_DOT = 5, FSIBE_SOLID_AX = 6, FSIBE_SOLID_AY = 7, FSIBE_SOLID_PRESSURE_OLD = 8,
  FSIBE_SOLID_PRESSURE_NEW = 9, FSIBE_SOLID_TAU_OLD = 10, FSIBE_SOLID_TAU_NEW = 11, FSIBE_SOLID_OLDOLD = 12
};
int fsibe_solid_get(fsibe_solid *b, int which, double *out);
int fsibe_solid_set(fsibe_solid *b, int which, const double *in);
/* scalars of type(solid) and of the solver */
typedef struct fsibe_solid_info {
  double dt_old;             /* previous step, for BDF2 */
  double grav;               /* gravity after GravityImpulse (SolidSolver.f90:874-911) */
  double E_kinetic, E_pot, E_elastic; /* SolidEnergies (SolidSolver.f90:76-92) */
  double Inertial_Force[2];
  int32_t StartupStep;       /* 1: the next IBES step is the first */
  int32_t iter;              /* Newton iterations of the last step (file IBES_iter) */
  int32_t brake;             /* 1: the last step tripped the emergency brake
                              * max|theta_dot - theta_dot_old| > 100 (SolidSolver.f90:370-380) */
  int32_t method;            /* scheme of the next IBES step */
} fsibe_solid_info;
int fsibe_solid_get_info(fsibe_solid *b, fsibe_solid_info *out);
/* restore dt_old, grav and StartupStep (e.g. from a backup file) */
int fsibe_solid_set_info(fsibe_solid *b, const fsibe_solid_info *in);
/* GravityImpulse(time); SolidSolverWrapper(time, dt, beams): an IBES step if
 * time > T_release, else integrate_position(time + dt); SolidEnergies; then
 * pressure_old = pressure_new, tau_old = tau_new (time_step.f90:209-228).
 * The loads pressure_new, tau_new are those GetForces wrote for this step
 * (set them with fsibe_solid_set).  FSIBE_EDOMAIN if Newton does not
 * converge in 49 iterations (the reference stops, SolidSolver.f90:311-314). */
int fsibe_solid_step(fsibe_solid *b, double time, double dt);
/* SolidSolverWrapper(time, dt, beams) alone (SolidSolver.f90:13-44), for a
 * host that runs GravityImpulse and advances the loads itself (the Fortran
 * module solid_hip.f90): grav is the solver's (fsibe_solid_set_info). */
int fsibe_solid_wrapper(fsibe_solid *b, double time, double dt);
/* GravityImpulse(time) (SolidSolver.f90:874-911): *grav in the share_vars
 * variable's role, overwritten for iImpulse 1..3, unchanged otherwise */
int fsibe_gravity_impulse(const fsibe_params *p, double time, double *grav);
/* The Newton system of the next IBES step t -> t + dt, x of 2*ns+4 values
 * [theta(-1..ns+1), T(-1..ns-1)]: the predictor (EE1 and the ghost values,
 * SolidSolver.f90:167-209), F_nonlinear (:552-667) and the Jacobian
 * J[k*(2ns+4) + l] = dF_k/dx_l, which is transpose(J) of Jacobi (:671-863),
 * the matrix Solve_LGS factors. */
int fsibe_solid_guess(fsibe_solid *b, double time, double dt, double *x);
int fsibe_solid_residual(fsibe_solid *b, double time, double dt, const double *x, double *F);
int fsibe_solid_jacobian(fsibe_solid *b, double time, double dt, const double *x, double *J);
/* RHS_beameqn (SolidSolver.f90:918-994) at time: theta_t (= theta_dot with
 * theta_dot(0) = 0), theta_tt and the tension T (Tension, :1000-1081) */
int fsibe_solid_rhs(fsibe_solid *b, double time, const double *theta, const double *theta_dot,
                    const double *pressure, const double *tau, double *theta_t, double *theta_tt,
                    double *T);
/* the solver fsibe_run uses for iFLUSI=1 (created by its first call); of the
 * first beam, or of beam ib */
int fsibe_sim_solid(fsibe_sim *sim, fsibe_solid **out);
int fsibe_sim_solid_beam(fsibe_sim *sim, int ib, fsibe_solid **out);

/* ---- the time loop of time_step.f90:120-341 --------------------------------
 * The startup Euler step if not yet taken, then AB2 steps, each create_mask
 * -> evolve -> GetForces -> beam -> swap -> time += dt1, where the beam is
 * integrate_position for prescribed kinematics (iFLUSI=0) or, for iFLUSI=1,
 * GravityImpulse + SolidSolverWrapper of the library's solid solver
 * (fsibe_sim_solid).  Runs nsteps steps, or until Time_end, or until the
 * solid solver's emergency brake (then out->status = 1, as the reference
 * ends its loop with continue_timestep = .false.).  out (may be NULL)
 * receives the last step's results; out->ibes_iter the Newton iterations of
 * its solid step. */
int fsibe_run(fsibe_sim *sim, int nsteps, fsibe_step_out *out);
/* current time, dt1 and number of AB2 steps taken by fsibe_run */
int fsibe_get_time(fsibe_sim *sim, double *time, double *dt1, int *it);
/* beam state held by fsibe_run (x, y, vx, vy, theta, pressure_new, tau_new),
 * iBeam*ns values each (beam ib from ib*ns) */
int fsibe_get_beam(fsibe_sim *sim, double *x, double *y, double *vx, double *vy, double *theta,
                   double *pressure_new, double *tau_new);
/* ---- batches of independent simulations --------------------------------
 * No reference counterpart (the reference runs one simulation per process;
 * sequential runs at several resolutions, dns.f90:70-101).  n members with
 * their own parameters (grid sizes may differ; iSponge!=10; a member with
 * iFLUSI=1 has its beam solved by the library's solid solver; each size
 * group runs its own chain of steps, its members' host work overlapping the
 * other groups' steps),
 * vortk[i] the packed initial vorticity of member i (or NULL).  Every AB2
 * step of the batch is one launch per stage and grid size; each member's
 * results are those of fsibe_run on that member alone.  The members are
 * ordinary simulations: fsibe_batch_member returns one for fsibe_get_field,
 * fsibe_get_step_out, fsibe_get_time, fsibe_get_beam (after fsibe_batch_run
 * returns); they belong to the batch and are destroyed with it. */
typedef struct fsibe_batch fsibe_batch;
int fsibe_batch_create(const fsibe_params *params, int n, const double *const *vortk, fsibe_batch **out);
int fsibe_batch_destroy(fsibe_batch *batch);
/* advance every member nsteps steps (the first of a member is its startup
 * step); stops early after a step in which a member's solid solver tripped
 * its emergency brake (that member's step_out status = 1; members of
 * other size groups may be one step further) */
int fsibe_batch_run(fsibe_batch *batch, int nsteps);
int fsibe_batch_size(fsibe_batch *batch);
fsibe_sim *fsibe_batch_member(fsibe_batch *batch, int i);
/* Step mode and tile autotuning (no reference counterpart).  At creation a
 * batch times every valid launch choice of its column stages (compile-time
 * or runtime FFT plan, 1..16 columns per tile) per grid size on its own
 * members and keeps the fastest, and each size group advances by replaying
 * two captured HIP graphs per step on its own stream (the stages up to the
 * nonlinear term, then the rest).
 * use_graph: 1 graphs, 0 eager launches, < 0 unchanged; autotune: 1 measure
 * again, 0 the single-simulation path's choices, < 0 unchanged.  No setting
 * changes a result: every choice runs the same passes on every column.
 * Single simulations take the tuner's choices too (measured at the first
 * fsibe_run of a grid in the process).  The environment variable
 * FSIBE_AUTOTUNE=0 keeps the static choices everywhere, for profiles that
 * must launch the same kernels from run to run. */
int fsibe_batch_configure(fsibe_batch *batch, int use_graph, int autotune);
/* pin the launch choice of column stage `stage` (1, 4, 5 or 7: K1, K4, K5,
 * K7) of size group `group` (the order of fsibe_batch_describe): the
 * compile-time FFT plan (ct = 1) or the runtime plan (ct = 0) with tx
 * columns per tile; FSIBE_EARG for a choice that size cannot launch */
int fsibe_batch_set_tiles(fsibe_batch *batch, int group, int stage, int ct, int tx);
/* the groups, the chosen launch shapes with their measured times, and the
 * step mode, as a NUL-terminated JSON object (FSIBE_ESIZE if len is short) */
int fsibe_batch_describe(fsibe_batch *batch, char *buf, size_t len);

/* ---- discrete adjoint ---------------------------------------------------
 * No reference counterpart (the reference has no adjoint).  For a
 * prescribed-motion simulation (iFLUSI=0, one beam, motion case 601: two
 * harmonics of heave and of pitch, coefficients motion_coef[8]):
 *     J(c) = (1/T) sum_n (-Force(1)^n) dt1^n,  T = sum_n dt1^n
 * over nsteps steps from the initial vorticity vortk (packed layout), where
 * Force(1) is cal_drag's penalisation force (NavierStokes.f90:567-575).
 * fsibe_adjoint_gradient returns J and dJ/dc by the discrete adjoint of the
 * GPU step (reverse sweep over recomputed units of chk steps, each recorded
 * on a tape of chk steps just before its adjoint steps run);
 * fsibe_adjoint_objective runs the forward only.  The time step is taken as
 * fixed (dt_fixed, CFL not binding).
 * fsibe_adjoint_create keeps the state of every unit's start (one forward
 * sweep, ceil(nsteps/chk) - 1 snapshot slots of 7 fields each).
 * fsibe_adjoint_create_binomial bounds the slots to nsnap (nsnap < 0: one
 * per unit start, as fsibe_adjoint_create): the reverse sweep then runs the
 * binomial checkpointing schedule of Griewank's revolve, which advances the
 * fewest units any schedule with nsnap slots can (fsibe_adjoint_plan_cost);
 * results are bit-identical for every nsnap. */
typedef struct fsibe_adjoint fsibe_adjoint;
int fsibe_adjoint_create(const fsibe_params *p, const double *vortk, int nsteps, int chk, fsibe_adjoint **out);
int fsibe_adjoint_create_binomial(const fsibe_params *p, const double *vortk, int nsteps, int chk, int nsnap,
                                  fsibe_adjoint **out);
int fsibe_adjoint_destroy(fsibe_adjoint *adj);
int fsibe_adjoint_objective(fsibe_adjoint *adj, const double coef[8], double *J);
int fsibe_adjoint_gradient(fsibe_adjoint *adj, const double coef[8], double *J, double grad[8]);
/* wall times of the last objective / gradient call (the gradient's time
 * includes its forward run; reverse_ms is the part after it) */
int fsibe_adjoint_timings(fsibe_adjoint *adj, double *forward_ms, double *gradient_ms, double *reverse_ms);
/* the snapshot slots, and the steps the last objective / gradient call ran
 * forward without a tape (a gradient's forward sweep included) and with one */
int fsibe_adjoint_stats(fsibe_adjoint *adj, int *slots, long long *steps_advanced, long long *steps_recorded);
/* The reverse schedule for nunits units and nslots slots (host only, no
 * GPU): action i is ops[3i..3i+2] = (op, unit, arg) with op 0 = restore the
 * state of unit `unit` from slot arg (-1: the initial state), 1 = advance
 * arg units from unit `unit`, 2 = store the state of unit `unit` in slot
 * arg, 3 = record unit `unit` and run its adjoint steps.  *n_acts is the
 * action count (FSIBE_ESIZE when it exceeds cap; ops may be NULL to ask for
 * it), *advances the units it advances. */
int fsibe_adjoint_plan(int nunits, int nslots, int *ops, int cap, int *n_acts, long long *advances);
/* the least number of units advanced by any schedule reversing nunits units
 * with nslots slots and the initial state at hand: r l - C(nslots+1+r, r-1)
 * with r the least integer such that C(nslots+1+r, r) >= l (-1 for bad sizes) */
long long fsibe_adjoint_plan_cost(int nunits, int nslots);

/* ---- restart files --------------------------------------------------------
 * MakeRuntimeBackup (save_fields.f90:109-152): the runtime_backup<n>.in file
 * of the time loop's state after the last step, as the reference writes it
 * (Fortran unformatted sequential records, 4-byte markers, REAL 8 bytes,
 * INTEGER/LOGICAL 4): time; 123.0; n1, dt1, vortk, nlk, workvis, mask,
 * maskvx, maskvy, ivideo, u; ns; colorscale, colorscale_done; the beam's six
 * records; vor_mean.  ivideo, colorscale, colorscale_done are the host's
 * output counters, stored as given.  The spectral arrays omit the Nyquist
 * modes the library does not hold (zero after dealiasing; nlk's are read
 * by nothing).  Returns FSIBE_EARG if the file cannot be written. */
int fsibe_backup_write(fsibe_sim *sim, const char *path, int ivideo, double colorscale, int colorscale_done);
/* init_fields with inicond = 2 (init_fields.f90:44-97) for one file: the
 * state is restored so that the next fsibe_run / step continues the time
 * loop (AB2, no startup step; create_mask runs again from the stored beam).
 * The beam's StartupStep is not in the file and is set, so a BDF2 solid
 * restarts with a Crank-Nicolson step, as in the reference.  FSIBE_EDOMAIN
 * on a wrong checksum, ns or record size.  Choosing the newer of the two
 * backup files (init_fields.f90:47-52) is the host's part. */
int fsibe_backup_read(fsibe_sim *sim, const char *path, int *ivideo, double *colorscale, int *colorscale_done);
/* the final file of time_step.f90:350-355: nx, ny and cofitxy(vortk) of
 * Fortran level 0 (the whole vortk array passed as a 2-D argument) */
int fsibe_inicond_write(fsibe_sim *sim, const char *path);
/* init_fields with inicond = 111 (init_fields.f90:109-128): vortk(:,:,0) =
 * coftxy of the file's vorticity; only before the first step */
int fsibe_inicond_read(fsibe_sim *sim, const char *path);

/* ---- field outputs ----------------------------------------------------------
 * ComputePressureSnapshot(time, vortk(:,:,level), press) (NavierStokes.f90:
 * 598-665): the static pressure of Fortran level `level` (0/1; the time
 * loop passes n1) at `time`, with the current mask, into press[nx*ny].
 * Leaves the simulation's state alone.  Not a step kernel: built from the
 * packed-layout spectral operators on the device. */
int fsibe_pressure_snapshot(fsibe_sim *sim, double time, int level, double *press);
/* SaveField(filename, field, NSave, lx, ly, fieldtype) (FieldExport.f90:
 * 612-648) for a host field [ny][nx]: header = the PARAMS lines that
 * WriteHeaderToFile writes (BeamIO.f90:364-376), '\n'-separated, each
 * written blank-padded to 256 characters (NULL: none).  No GPU involved. */
int fsibe_save_field(const char *path, const double *field, int nx, int ny, int nsave, double xl, double yl,
                     const char *fieldtype, const char *header);
/* SaveGIF(field, filename, iColorTable, color_min, color_max)
 * (FieldExport.f90:305-384, written by gif_util.f90's writegif) for a host
 * field [ny][nx]: writes path + ".gif", a GIF89a of 256 colours with the
 * field's last row at the top.  palette: FargePalette's computed palettes
 * 11, 13 (SaveGIF's default) or 14 (FSIBE_EUNSUPPORTED for the tabulated
 * ones); have_range = 0 takes minval/maxval of the field as SaveGIF does
 * when color_min/color_max are absent.  No GPU involved. */
int fsibe_save_gif(const char *path, const double *field, int nx, int ny, int palette, int have_range,
                   double color_min, double color_max);

/* HIP stream of the simulation (hipStream_t) for event timing */
void *fsibe_sim_stream(fsibe_sim *sim);

/* Per-kernel timing with HIP events recorded on the simulation's stream
 * around each launch (no reference counterpart: the reference times whole
 * routines with system_clock, PerformanceMeasurement.f90:26-121).
 * fsibe_profile_enable resets the counters and times the kernel classes whose
 * bits are set in `classes` (-1: all, 0: off; timing synchronises once per
 * step); fsibe_profile_read fills up to n
 * totals (milliseconds) and launch counts and returns the number of kernel
 * classes: 0 create_mask, 1 startup velocity, 2 columns of vortk(n0),
 * 3 rows: nonlinear term (with the dt/AB2 scalars), 5 columns: AB2 update,
 * 6 columns of vortk(n1), 7 rows: velocity/pressure rhs and row sums,
 * 9 columns: pressure, 10 rows: pressure/stress box, 11 Delta interpolation
 * and per-node beam loads.  Classes 4, 8 and 12 time host phases with the
 * steady clock: 4 the solid solver's step (iFLUSI=1), 8 the mask geometry
 * and its upload, 12 the wait for a step's results (the host adds the row
 * sums and the per-node terms when it reads a step). */
int fsibe_profile_enable(fsibe_sim *sim, int classes);
int fsibe_profile_read(fsibe_sim *sim, double *total_ms, long long *count, int n);

#ifdef __cplusplus
}
#endif
#endif
