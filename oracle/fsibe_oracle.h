/*
 * fsibe_oracle.h -- CPU restatement of the FSIBE_2D fluid time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * product path and the timed CPU baseline.  Only tests/, the smoke check in
 * __graft_entry__.py and the cpu_baseline leg of bench.py may load it.  The
 * product library (fsibe_2d_amd/libfsibe_hip.so) never links or calls it.
 *
 * Every routine restates one reference routine (Fortran 90, compiled -r8, so
 * every REAL and every literal is IEEE binary64) and keeps its loop order and
 * arithmetic order, except the inside of the 2-D real FFT: the reference calls
 * FFTW3 (not vendored, no version pinned), whose internal algorithm is not
 * part of the contract -- only the DFT it computes is.
 *
 * Array convention (reference NavierStokes.f90:19-28): real(8) f(0:nx-1,0:ny-1)
 * column-major, x fastest, i.e. f[iy*nx + ix] in C.
 *
 * How each part is pinned (tests/test_oracle_*.py, tests/test_golden.py):
 * the FFT and packing by a brute-force DFT, numpy and single-mode closed
 * forms; the operators, vis, dealiase_mask, Mean_ux/uy, mouvement and
 * integrate_position bit for bit by the reference's own Fortran (oracle/_ref)
 * and its golden vectors; cal_velocity, cal_nlk (nonlinear, penalisation and
 * sponge terms), the CFL step, cal_energy, time_avg_vort and the integrator
 * by closed forms (Taylor-Green decay, AB2 order); SmoothStep, DrawADot, a
 * straight beam, the channel walls and the sponge masks of types 1, 2, 3
 * and 10 by erf closed forms and symmetry;
 * phi*, DeltaInterpolation and GetForcesInterp by moment conditions and
 * linear / sinusoidal fields.
 * cal_drag's Force by the integral of chi (u - u_s) over the grid, which
 * it equals with the window grown to the whole mask; the pressure part of
 * BeamSurfaceIntegral by a linear pressure field on a straight beam.
 * tests/test_oracle_pins.py pins the rest: DrawHinge's sectors (every
 * branch pair of beta1/beta2, both turn directions) by the distance to a
 * bent polyline computed by brute force; DrawSharpEnd by the product of its
 * two erf steps on a straight beam; the iSponge = 10 term by the closed form
 * of its trapezoid sum and its second-order convergence; cal_drag's window
 * by a sharp mask it must hold and a one-column shift of its truncated
 * bound, and Force_unst by a rigid heave's volume times acceleration; the
 * stress part of BeamSurfaceIntegral by the tractions S.n of linear stress
 * fields.  Each of these tests fails on a planted mistake in the routine it
 * checks (a wrong branch, half weight, off-by-one window, nint for int,
 * swapped sign or node, max instead of overwrite).
 * Parity unpinned: none of the fluid step's routines.
 */
#ifndef FSIBE_ORACLE_H
#define FSIBE_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Run parameters (reference share_vars.f90:8-25, params.f90:9-81).  The same
 * field order is used by the product's fsibe_params so one host record fills
 * both. */
typedef struct o_params {
  int32_t nx, ny, ns;
  int32_t iMeanVelocity, iSponge, iBeam, iWalls, iCylinder, iSharpTrailing;
  int32_t iFLUSI, iMotion, iViscous, iSaveBeam, idealis;
  double xl, yl, nu, eps, eps_sponge, cfl, dt_fixed, Time_end;
  double N_smooth, t_beam, length, SpongeSize, R_cylinder, h_channel;
  double x0, y0, AngleBeam, theta_inf, T_fullspeed, T_release, tau;
  double U_mean_true;
  /* motion case 600 (sinusoidal heave) and 601 (Fourier heave + pitch) */
  double heave_amp, heave_freq;
  double motion_coef[8];
  /* beam material (used by the host solid solver, not by the oracle) */
  double mue, eta, sigma, grav;
  int32_t TimeMethodSolid, iImpulse;
  /* 1: segments and hinges drawn out to t + 7 delta and the cal_drag window
   * grown to the beam's whole mask (a form the reference does not have) */
  int32_t iSmoothCutoff;
} o_params;

/* One beam (reference share_vars.f90:50-63, type Solid); arrays length ns. */
typedef struct o_beam {
  double *x, *y, *vx, *vy, *theta, *theta_dot;
  double *pressure_new, *tau_new;
  double Force[2], Force_unst[2], Force_press[2];
  double x0, y0, AngleBeam, phase;
  double drag_unst_new, drag_unst_old, lift_unst_new, lift_unst_old;
  int32_t UnsteadyCorrectionsReady;
  int32_t iMouvement;
} o_beam;

typedef struct o_sim o_sim;

/* ---- lifecycle (dns.f90:183-205, fft_initialize cof_fftw33.f90:18-66) -- */
o_sim *o_sim_create(const o_params *p);
void o_sim_destroy(o_sim *s);
void o_set_threads(int n);
double o_pi(const o_sim *s);

/* ---- FFT plugin (cof_fftw33.f90:142-243) ---------------------------------*/
void o_coftxy(o_sim *s, const double *f, double *fk);
void o_cofitxy(o_sim *s, const double *fk, double *f);
/* raw complex half spectrum helpers for pinning tests: ft is (nx/2+1) x ny
 * complex, kx fastest, interleaved re/im, unnormalised r2c */
void o_rfft2(o_sim *s, const double *f, double *ft);
void o_irfft2(o_sim *s, const double *ft, double *f);
/* 1-D complex FFT of length n (sign -1 forward, +1 inverse, unnormalised) */
void o_cfft1(int n, double *x, int sign);

/* ---- spectral operators (cofdx.f90 .. dealiase_mask.f90) -----------------*/
void o_cofdx(o_sim *s, const double *fk, double *out);
void o_cofdy(o_sim *s, const double *fk, double *out);
void o_cofdxdx(o_sim *s, const double *fk, double *out);
void o_cofdydy(o_sim *s, const double *fk, double *out);
void o_cofdxdy(o_sim *s, const double *fk, double *out);
void o_poisson(o_sim *s, const double *f, double *ans);
void o_vis(o_sim *s, double dt, double *work);
void o_dealiase_mask(o_sim *s, double *dealiase);

/* ---- kinematics (mouvement.f90, integrate_position.f90, init_beam.f90) ---*/
void o_mouvement(o_sim *s, double time, const o_beam *b, double *alpha,
                 double *alpha_t, double *alpha_tt, double LeadingEdge[6]);
void o_integrate_position(o_sim *s, double time, o_beam *b);
void o_init_beam(o_sim *s, o_beam *b); /* allocates arrays of length ns */
void o_init_beam_k(o_sim *s, o_beam *b, int ib); /* beam ib of init_beam.f90:48-76 */
void o_free_beam(o_beam *b);
double o_mean_ux(o_sim *s, double time);
double o_mean_uy(o_sim *s, double time);

/* ---- geometry (create_mask.f90) -----------------------------------------*/
void o_smoothstep(o_sim *s, double *f, double x, double t, double h);
/* the beam routines take b = beams(1:iBeam), o_nbeams(s) consecutive beams */
void o_create_mask(o_sim *s, double time, o_beam *b);
void o_create_sponge_mask(o_sim *s);
void o_frame_mask(o_sim *s, double *mask1);
void o_pressure_snapshot(o_sim *s, double time, int level, double *press);

/* ---- fluid solver (NavierStokes.f90) -------------------------------------*/
void o_cal_velocity(o_sim *s, double time, const double *vortk, double *u,
                    double *stream);
/* u is (nx*ny*2): u(:,:,1) then u(:,:,2) */
void o_sponge10_term(o_sim *s, const double *u1, double *pen1);
void o_cal_nlk(o_sim *s, double time, double *dt1, double *nlk,
               const double *vortk, const double *u);
void o_pressure(o_sim *s, double time, double dt1, const double *vortk,
                double *press, double *u, o_beam *b);
void o_evolve_explicit(o_sim *s, double time, double *dt1, int n0, int n1,
                       o_beam *b);
void o_evolve_ab2(o_sim *s, double time, double dt0, double *dt1, int n0,
                  int n1, o_beam *b);
void o_cal_drag(o_sim *s, double time, double dt1, const double *penal,
                o_beam *b);

/* ---- beam forces (BeamForces.f90, Interpolation.f90) ---------------------*/
double o_delta_phi4star(double x, double dx, double pi);
double o_delta_interp(o_sim *s, double x_target, double y_target,
                      const double *field, int *err);
void o_surface_integral(o_sim *s, double time, const double *u, const double *sa, const double *sb,
                        const double *press);
void o_get_forces(o_sim *s, double time, o_beam *b, const double *press,
                  const double *u);

/* ---- time loop (time_step.f90:120-341, prescribed kinematics) ------------*/
void o_set_vortk(o_sim *s, const double *vortk_P); /* into slab n0 */
void o_startup_step(o_sim *s);
void o_ab2_step(o_sim *s);
void o_run(o_sim *s, int nsteps); /* startup (if needed) + AB2 steps */
/* per-step calls for a host that moves the beam itself (iFLUSI=1) */
void o_host_create_mask(o_sim *s, double time);
double o_host_evolve(o_sim *s, double time, double dt0, int explicit_step);
void o_host_get_forces(o_sim *s, double time);
void o_host_swap(o_sim *s);

/* ---- state access --------------------------------------------------------*/
enum {
  O_F_VORTK = 0, O_F_NLK = 1, O_F_WORKVIS = 2, O_F_U1 = 3, O_F_U2 = 4,
  O_F_PRESS = 5, O_F_MASK = 6, O_F_MASKVX = 7, O_F_MASKVY = 8,
  O_F_MASK_SPONGE = 9, O_F_VOR_MEAN = 10, O_F_DEALIASE = 11,
  O_F_STRESS_A = 12, O_F_STRESS_B = 13, O_F_VORT = 14
};
void o_get_field(o_sim *s, int which, double *out); /* current n0 slab */
void o_set_field(o_sim *s, int which, const double *in);
o_beam *o_get_beam(o_sim *s);           /* the first beam (of o_nbeams consecutive ones) */
o_beam *o_get_beam_k(o_sim *s, int ib);
int o_nbeams(const o_sim *s);
void o_get_beam_bsi(o_sim *s, int ib, double *out);
/* scalars: time, dt0, dt1, u_max, energies[6], bsi[11], n0, it */
typedef struct o_diag {
  double time, dt0, dt1, u_max;
  double energies[6]; /* vor_rms kinetic dissipation porous_diss inflow_solid inflow_mean */
  double bsi[11];     /* drag lift drag_surf lift_surf Ft1 Fn1 Fp1 Ft2 Fn2 Fp2 (10 used) */
  double fx_int, fy_int;
  int32_t n0, it, started, pad;
  double u_mean[2];   /* iMeanVelocity = 11 */
} o_diag;
void o_get_diag(o_sim *s, o_diag *d);

/* ---- synthetic inputs ----------------------------------------------------*/
uint64_t o_splitmix64(uint64_t *state);
/* nx*ny normals in column-major order, SplitMix64(seed) + Box-Muller */
void o_normals(uint64_t seed, int n, double *out);
/* band-limited noise vorticity in packed layout: coftxy(amp*xi), keep blocks
 * with kx^2+ky^2 <= kmax2, times dealiase */
void o_init_noise(o_sim *s, uint64_t seed, double amp, int kmax2,
                  double *vortk_P);

#ifdef __cplusplus
}
#endif
#endif
