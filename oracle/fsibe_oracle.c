/*
 * fsibe_oracle.c -- CPU restatement of the FSIBE_2D fluid time step.
 *
 * TEST INFRASTRUCTURE ONLY (parity checker and timed CPU baseline); see the
 * header.  Each function cites the reference routine (file:line) it follows.
 * Arithmetic is written in the reference's evaluation order (Fortran evaluates
 * a+b+c as (a+b)+c, -a*b as -(a*b)) so that differences to the reference stem
 * only from the FFT's internal rounding and from libm.
 *
 * Parity pinning (what checks each part) is listed in tests/test_oracle_*.py.
 */
#include "fsibe_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXFAC 40

/* ========================================================================= */
/* 1-D complex FFT: Stockham autosort, radix 4/2/3/5 + generic odd factor.   */
/* Twiddles exp(-2 pi i m/n) are computed in long double and rounded once.   */
/* ========================================================================= */
typedef struct fplan {
  int n, nf, fac[MAXFAC];
  double *tw; /* 2n: re, im of exp(-2 pi i m / n) */
} fplan;

static fplan *fplan_create(int n) {
  fplan *p = (fplan *)calloc(1, sizeof(fplan));
  p->n = n;
  int m = n;
  while (m % 4 == 0) { p->fac[p->nf++] = 4; m /= 4; }
  while (m % 2 == 0) { p->fac[p->nf++] = 2; m /= 2; }
  while (m % 3 == 0) { p->fac[p->nf++] = 3; m /= 3; }
  while (m % 5 == 0) { p->fac[p->nf++] = 5; m /= 5; }
  for (int f = 7; m > 1; f += 2)
    while (m % f == 0) { p->fac[p->nf++] = f; m /= f; }
  p->tw = (double *)malloc(sizeof(double) * 2 * (size_t)n);
  const long double twopi = 6.283185307179586476925286766559005768L;
  for (int k = 0; k < n; k++) {
    long double a = twopi * (long double)k / (long double)n;
    p->tw[2 * k] = (double)cosl(a);
    p->tw[2 * k + 1] = (double)(-sinl(a));
  }
  return p;
}

static void fplan_destroy(fplan *p) {
  if (!p) return;
  free(p->tw);
  free(p);
}

/* in-place DFT of R points (vr, vi); sign -1 forward, +1 inverse */
static void small_dft(int R, double *vr, double *vi, int sign,
                      const fplan *p) {
  if (R == 2) {
    double ar = vr[0] + vr[1], ai = vi[0] + vi[1];
    double br = vr[0] - vr[1], bi = vi[0] - vi[1];
    vr[0] = ar; vi[0] = ai; vr[1] = br; vi[1] = bi;
  } else if (R == 4) {
    double a0r = vr[0] + vr[2], a0i = vi[0] + vi[2];
    double a1r = vr[0] - vr[2], a1i = vi[0] - vi[2];
    double a2r = vr[1] + vr[3], a2i = vi[1] + vi[3];
    double dr = vr[1] - vr[3], di = vi[1] - vi[3];
    /* a3 = d * (sign*i) */
    double a3r = -sign * di, a3i = sign * dr;
    vr[0] = a0r + a2r; vi[0] = a0i + a2i;
    vr[2] = a0r - a2r; vi[2] = a0i - a2i;
    vr[1] = a1r + a3r; vi[1] = a1i + a3i;
    vr[3] = a1r - a3r; vi[3] = a1i - a3i;
  } else if (R == 3) {
    const double s3 = 0.86602540378443864676; /* sin(2 pi/3) */
    double tr = vr[1] + vr[2], ti = vi[1] + vi[2];
    double dr = vr[1] - vr[2], di = vi[1] - vi[2];
    double mr = vr[0] - 0.5 * tr, mi = vi[0] - 0.5 * ti;
    /* (sign*i*s3) * d */
    double er = -sign * s3 * di, ei = sign * s3 * dr;
    vr[0] = vr[0] + tr; vi[0] = vi[0] + ti;
    vr[1] = mr + er; vi[1] = mi + ei;
    vr[2] = mr - er; vi[2] = mi - ei;
  } else {
    /* generic: exp(sign 2 pi i r q / R) = tw[(r q mod R) * n/R] (conj for +) */
    double tr[64], ti[64];
    int step = p->n / R;
    for (int q = 0; q < R; q++) {
      double sr = 0.0, si = 0.0;
      for (int r = 0; r < R; r++) {
        int idx = ((r * q) % R) * step;
        double wr = p->tw[2 * idx], wi = p->tw[2 * idx + 1] * (double)(-sign);
        sr += vr[r] * wr - vi[r] * wi;
        si += vr[r] * wi + vi[r] * wr;
      }
      tr[q] = sr; ti[q] = si;
    }
    for (int q = 0; q < R; q++) { vr[q] = tr[q]; vi[q] = ti[q]; }
  }
}

/* x: 2n doubles, work: 2n doubles; result in x */
static void fplan_exec(const fplan *p, double *x, double *work, int sign) {
  int n = p->n;
  if (n == 1) return;
  double *in = x, *out = work;
  int Ns = 1;
  for (int f = 0; f < p->nf; f++) {
    int R = p->fac[f];
    int m = n / R;
    int tstep = n / (Ns * R);
    double vr[64], vi[64];
    for (int j = 0; j < m; j++) {
      int k = j % Ns;
      for (int r = 0; r < R; r++) {
        vr[r] = in[2 * (j + r * m)];
        vi[r] = in[2 * (j + r * m) + 1];
      }
      if (Ns > 1) {
        for (int r = 1; r < R; r++) {
          int idx = r * k * tstep; /* < n */
          double wr = p->tw[2 * idx], wi = p->tw[2 * idx + 1];
          if (sign > 0) wi = -wi;
          double a = vr[r] * wr - vi[r] * wi;
          double b = vr[r] * wi + vi[r] * wr;
          vr[r] = a; vi[r] = b;
        }
      }
      small_dft(R, vr, vi, sign, p);
      int base = (j / Ns) * Ns * R + k;
      for (int r = 0; r < R; r++) {
        out[2 * (base + r * Ns)] = vr[r];
        out[2 * (base + r * Ns) + 1] = vi[r];
      }
    }
    Ns *= R;
    double *t = in; in = out; out = t;
  }
  if (in != x) memcpy(x, in, sizeof(double) * 2 * (size_t)n);
}

void o_cfft1(int n, double *x, int sign) {
  fplan *p = fplan_create(n);
  double *w = (double *)malloc(sizeof(double) * 2 * (size_t)n);
  fplan_exec(p, x, w, sign);
  free(w);
  fplan_destroy(p);
}

/* ========================================================================= */
/* simulation state (share_vars.f90 globals + time_step.f90 locals)          */
/* ========================================================================= */
struct o_sim {
  o_params p;
  int nx, ny, ns, H;
  double pi, scalex, scaley, dx, dy, ds;
  int sharp;
  fplan *fx; /* length nx/2 */
  fplan *fy; /* length ny   */
  double *twx; /* exp(-2 pi i k / nx), k in [0, nx/2] for the real post-process */
  double *dealiase, *mask, *maskvx, *maskvy, *mask_sponge, *mask_frame,
      *vor_mean;
  double *vortk[2], *nlk[2], *workvis[2];
  double *u;
  double *press, *stress_a, *stress_b, *vort;
  o_beam beams[2]; /* beams(1:iBeam), iBeam <= 2 (init_beam.f90:48-76) */
  int nb;
  double time, dt0, dt1, u_max;
  int n0, n1, it, started, startup_done;
  double energies[6], bsi[11], fx_int, fy_int;
  double bsi_b[2][11], fx_int_b[2], fy_int_b[2]; /* per beam; bsi, fx_int, fy_int: the first beam's */
  double u_mean[2];
  double fmean[2][2];
  double fm_new[2]; /* f_mean of the last cal_nlk (NavierStokes.f90:97-100) */
  int fm_set;       /* 1 when cal_nlk computed it (penalisation on) */
  int frame_done;   /* mask_frame set up (create_mask.f90:188-192) */
};

static int g_threads = 0;
#ifdef _OPENMP
#define NTHREADS (g_threads > 0 ? g_threads : omp_get_max_threads())
#else
#define NTHREADS 1
#endif
void o_set_threads(int n) { g_threads = n; }
double o_pi(const o_sim *s) { return s->pi; }

static double *dalloc(size_t n) {
  double *p = (double *)calloc(n ? n : 1, sizeof(double));
  if (!p) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
  return p;
}

/* dns.f90:183-201 and params.f90:139-141,190-208 */
o_sim *o_sim_create(const o_params *p) {
  o_sim *s = (o_sim *)calloc(1, sizeof(o_sim));
  s->p = *p;
  s->nx = p->nx; s->ny = p->ny; s->ns = p->ns; s->H = p->nx / 2;
  s->pi = 4.0 * atan(1.0);                      /* params.f90:139 */
  { double a = 2.0 * s->pi / p->xl; s->scalex = a * a; } /* params.f90:140 */
  { double a = 2.0 * s->pi / p->yl; s->scaley = a * a; } /* params.f90:141 */
  s->dx = p->xl / (double)p->nx;
  s->dy = p->yl / (double)p->ny;
  s->ds = (p->ns > 1) ? p->length / (double)(p->ns - 1) : 0.0; /* params.f90:190 */
  s->sharp = (fabs(p->N_smooth) < 1e-4);        /* params.f90:205-208 */
  if (s->p.U_mean_true == 0.0) s->p.U_mean_true = 1.0; /* time_step.f90:96 */
  s->fx = fplan_create(s->H);
  s->fy = fplan_create(s->ny);
  s->twx = dalloc(2 * (size_t)(s->H + 1));
  const long double twopi = 6.283185307179586476925286766559005768L;
  for (int k = 0; k <= s->H; k++) {
    long double a = twopi * (long double)k / (long double)s->nx;
    s->twx[2 * k] = (double)cosl(a);
    s->twx[2 * k + 1] = (double)(-sinl(a));
  }
  size_t N = (size_t)s->nx * s->ny;
  s->dealiase = dalloc(N);
  s->mask = dalloc(N); s->maskvx = dalloc(N); s->maskvy = dalloc(N);
  s->mask_sponge = dalloc(N); s->mask_frame = dalloc(N); s->vor_mean = dalloc(N);
  for (int i = 0; i < 2; i++) {
    s->vortk[i] = dalloc(N); s->nlk[i] = dalloc(N); s->workvis[i] = dalloc(N);
  }
  s->u = dalloc(2 * N);
  s->press = dalloc(N); s->stress_a = dalloc(N); s->stress_b = dalloc(N);
  s->vort = dalloc(N);
  /* dealiasing mask: dns.f90:199-200 */
  for (size_t i = 0; i < N; i++) s->dealiase[i] = 1.0;
  if (p->idealis == 1) o_dealiase_mask(s, s->dealiase);
  /* time_step.f90:11,43,79 */
  s->n1 = 1; s->n0 = 0; s->time = 0.0; s->dt1 = 0.0; s->dt0 = 0.0;
  s->nb = p->iBeam > 2 ? 2 : (p->iBeam > 0 ? p->iBeam : 0);
  if (p->iBeam > 2) { fprintf(stderr, "oracle: at most two beams\n"); abort(); }
  for (int ib = 0; ib < s->nb; ib++) o_init_beam_k(s, &s->beams[ib], ib);
  return s;
}

void o_sim_destroy(o_sim *s) {
  if (!s) return;
  fplan_destroy(s->fx); fplan_destroy(s->fy);
  free(s->twx);
  free(s->dealiase); free(s->mask); free(s->maskvx); free(s->maskvy);
  free(s->mask_sponge); free(s->mask_frame); free(s->vor_mean);
  for (int i = 0; i < 2; i++) { free(s->vortk[i]); free(s->nlk[i]); free(s->workvis[i]); }
  free(s->u); free(s->press); free(s->stress_a); free(s->stress_b); free(s->vort);
  for (int ib = 0; ib < s->nb; ib++) o_free_beam(&s->beams[ib]);
  free(s);
}

/* ========================================================================= */
/* 2-D real FFT (the DFT FFTW computes for dfftw_plan_dft_r2c_2d(nx,ny) and  */
/* its c2r inverse, cof_fftw33.f90:52,54).  ft is (H+1) x ny complex.        */
/* ========================================================================= */
void o_rfft2(o_sim *s, const double *f, double *ft) {
  const int nx = s->nx, ny = s->ny, H = s->H, W = H + 1;
#pragma omp parallel num_threads(NTHREADS)
  {
    double *z = dalloc(2 * (size_t)(H > 0 ? H : 1));
    double *w = dalloc(2 * (size_t)(ny > H ? ny : H));
#pragma omp for schedule(static)
    for (int iy = 0; iy < ny; iy++) {
      const double *row = f + (size_t)iy * nx;
      for (int n = 0; n < H; n++) { z[2 * n] = row[2 * n]; z[2 * n + 1] = row[2 * n + 1]; }
      fplan_exec(s->fx, z, w, -1);
      double *out = ft + 2 * (size_t)iy * W;
      /* X[k] = 1/2 (Z_k + conj Z_{H-k}) - i/2 e^{-2 pi i k/nx} (Z_k - conj Z_{H-k}) */
      out[0] = z[0] + z[1]; out[1] = 0.0;
      out[2 * H] = z[0] - z[1]; out[2 * H + 1] = 0.0;
      for (int k = 1; k < H; k++) {
        double ar = z[2 * k], ai = z[2 * k + 1];
        double br = z[2 * (H - k)], bi = -z[2 * (H - k) + 1]; /* conj */
        double er = 0.5 * (ar + br), ei = 0.5 * (ai + bi);
        double dr = 0.5 * (ar - br), di = 0.5 * (ai - bi);
        double wr = s->twx[2 * k], wi = s->twx[2 * k + 1];
        /* -i * w * d */
        double pr = wr * dr - wi * di, pi_ = wr * di + wi * dr;
        out[2 * k] = er + pi_;
        out[2 * k + 1] = ei - pr;
      }
    }
    free(z); free(w);
  }
#pragma omp parallel num_threads(NTHREADS)
  {
    double *c = dalloc(2 * (size_t)ny), *w = dalloc(2 * (size_t)ny);
#pragma omp for schedule(static)
    for (int kx = 0; kx < W; kx++) {
      for (int iy = 0; iy < ny; iy++) {
        c[2 * iy] = ft[2 * ((size_t)iy * W + kx)];
        c[2 * iy + 1] = ft[2 * ((size_t)iy * W + kx) + 1];
      }
      fplan_exec(s->fy, c, w, -1);
      for (int iy = 0; iy < ny; iy++) {
        ft[2 * ((size_t)iy * W + kx)] = c[2 * iy];
        ft[2 * ((size_t)iy * W + kx) + 1] = c[2 * iy + 1];
      }
    }
    free(c); free(w);
  }
}

/* unnormalised c2r; Im of the x-DC and x-Nyquist inputs are ignored (c2r
 * convention) */
void o_irfft2(o_sim *s, const double *ft_in, double *f) {
  const int nx = s->nx, ny = s->ny, H = s->H, W = H + 1;
  double *ft = dalloc(2 * (size_t)W * ny);
  memcpy(ft, ft_in, sizeof(double) * 2 * (size_t)W * ny);
#pragma omp parallel num_threads(NTHREADS)
  {
    double *c = dalloc(2 * (size_t)ny), *w = dalloc(2 * (size_t)ny);
#pragma omp for schedule(static)
    for (int kx = 0; kx < W; kx++) {
      for (int iy = 0; iy < ny; iy++) {
        c[2 * iy] = ft[2 * ((size_t)iy * W + kx)];
        c[2 * iy + 1] = ft[2 * ((size_t)iy * W + kx) + 1];
      }
      fplan_exec(s->fy, c, w, +1);
      for (int iy = 0; iy < ny; iy++) {
        ft[2 * ((size_t)iy * W + kx)] = c[2 * iy];
        ft[2 * ((size_t)iy * W + kx) + 1] = c[2 * iy + 1];
      }
    }
    free(c); free(w);
  }
#pragma omp parallel num_threads(NTHREADS)
  {
    double *z = dalloc(2 * (size_t)(H > 0 ? H : 1));
    double *w = dalloc(2 * (size_t)(H > 0 ? H : 1));
#pragma omp for schedule(static)
    for (int iy = 0; iy < ny; iy++) {
      const double *X = ft + 2 * (size_t)iy * W;
      /* Z_k = (X_k + conj X_{H-k}) + i e^{+2 pi i k/nx} (X_k - conj X_{H-k}) */
      for (int k = 0; k < H; k++) {
        double ar = X[2 * k], ai = (k == 0) ? 0.0 : X[2 * k + 1];
        double br = X[2 * (H - k)], bi = (k == 0) ? 0.0 : -X[2 * (H - k) + 1];
        double er = ar + br, ei = ai + bi;
        double dr = ar - br, di = ai - bi;
        double wr = s->twx[2 * k], wi = -s->twx[2 * k + 1]; /* e^{+} */
        double pr = wr * dr - wi * di, pi_ = wr * di + wi * dr;
        z[2 * k] = er - pi_;
        z[2 * k + 1] = ei + pr;
      }
      fplan_exec(s->fx, z, w, +1);
      double *row = f + (size_t)iy * nx;
      for (int n = 0; n < H; n++) { row[2 * n] = z[2 * n]; row[2 * n + 1] = z[2 * n + 1]; }
    }
    free(z); free(w);
  }
  free(ft);
}

/* coftxy: cof_fftw33.f90:142-193 */
void o_coftxy(o_sim *s, const double *f, double *fk) {
  const int nx = s->nx, ny = s->ny, W = s->H + 1;
  double *ft = dalloc(2 * (size_t)W * ny);
  o_rfft2(s, f, ft);
  double norm = 1.0 / ((double)nx * (double)ny); /* :162 */
#define FT_RE(ix, iy) ft[2 * ((size_t)(iy) * W + (ix))]
#define FT_IM(ix, iy) ft[2 * ((size_t)(iy) * W + (ix)) + 1]
#define FK(ix, iy) fk[(size_t)(iy) * nx + (ix)]
  for (int ix = 0; ix <= nx / 2 - 1; ix++) { /* :167-174 */
    FK(2 * ix, 0) = FT_RE(ix, 0) * norm;
    FK(2 * ix + 1, 0) = FT_IM(ix, 0) * norm;
    FK(2 * ix, 1) = 0.0;
    FK(2 * ix + 1, 1) = 0.0;
  }
  for (int iy = 1; iy <= ny / 2 - 1; iy++) { /* :177-189 */
    for (int ix = 0; ix <= nx / 2 - 1; ix++) {
      FK(2 * ix, 2 * iy) = 0.5 * (FT_RE(ix, iy) + FT_RE(ix, ny - iy)) * norm;
      FK(2 * ix, 2 * iy + 1) = 0.5 * (FT_IM(ix, iy) - FT_IM(ix, ny - iy)) * norm;
      FK(2 * ix + 1, 2 * iy) = 0.5 * (FT_IM(ix, iy) + FT_IM(ix, ny - iy)) * norm;
      FK(2 * ix + 1, 2 * iy + 1) = 0.5 * (-FT_RE(ix, iy) + FT_RE(ix, ny - iy)) * norm;
    }
  }
#undef FK
  free(ft);
}

/* cofitxy: cof_fftw33.f90:197-243 */
void o_cofitxy(o_sim *s, const double *fk, double *f) {
  const int nx = s->nx, ny = s->ny, W = s->H + 1;
  double *ft = dalloc(2 * (size_t)W * ny);
#define FK(ix, iy) fk[(size_t)(iy) * nx + (ix)]
  for (int ix = 0; ix <= nx / 2 - 1; ix++) { /* :211-215 */
    FT_RE(ix, 0) = FK(2 * ix, 0); FT_IM(ix, 0) = FK(2 * ix + 1, 0);
    FT_RE(ix, ny / 2) = 0.0; FT_IM(ix, ny / 2) = 0.0;
  }
  FT_RE(nx / 2, 0) = 0.0; FT_IM(nx / 2, 0) = 0.0; /* :218-219 */
  FT_RE(nx / 2, ny / 2) = 0.0; FT_IM(nx / 2, ny / 2) = 0.0;
  for (int iy = 1; iy <= ny / 2 - 1; iy++) { /* :222-231 */
    for (int ix = 0; ix <= nx / 2 - 1; ix++) {
      FT_RE(ix, iy) = FK(2 * ix, 2 * iy) - FK(2 * ix + 1, 2 * iy + 1);
      FT_IM(ix, iy) = FK(2 * ix, 2 * iy + 1) + FK(2 * ix + 1, 2 * iy);
      FT_RE(ix, ny - iy) = FK(2 * ix, 2 * iy) + FK(2 * ix + 1, 2 * iy + 1);
      FT_IM(ix, ny - iy) = -FK(2 * ix, 2 * iy + 1) + FK(2 * ix + 1, 2 * iy);
    }
    FT_RE(nx / 2, iy) = 0.0; FT_IM(nx / 2, iy) = 0.0;
    FT_RE(nx / 2, ny - iy) = 0.0; FT_IM(nx / 2, ny - iy) = 0.0;
  }
#undef FK
#undef FT_RE
#undef FT_IM
  o_irfft2(s, ft, f);
  free(ft);
}

/* ========================================================================= */
/* spectral operators                                                         */
/* ========================================================================= */
/* cofdx.f90:1-29 */
void o_cofdx(o_sim *s, const double *fk, double *out) {
  const int nx = s->nx, ny = s->ny;
  double scale1 = 2.0 * s->pi / s->p.xl;
  for (int ky = 0; ky < ny; ky++) {
    for (int kx = 0; kx <= nx - 2; kx += 2)
      out[(size_t)ky * nx + kx] = -fk[(size_t)ky * nx + kx + 1] * (double)(kx / 2) * scale1;
    for (int kx = 1; kx <= nx - 1; kx += 2)
      out[(size_t)ky * nx + kx] = fk[(size_t)ky * nx + kx - 1] * (double)((kx - 1) / 2) * scale1;
  }
}

/* cofdy.f90:1-30 */
void o_cofdy(o_sim *s, const double *fk, double *out) {
  const int nx = s->nx, ny = s->ny;
  double scale1 = 2.0 * s->pi / s->p.yl;
  for (int kx = 0; kx < nx; kx++) {
    for (int ky = 0; ky <= ny - 2; ky += 2)
      out[(size_t)ky * nx + kx] = -fk[(size_t)(ky + 1) * nx + kx] * (double)(ky / 2) * scale1;
    for (int ky = 1; ky <= ny - 1; ky += 2)
      out[(size_t)ky * nx + kx] = fk[(size_t)(ky - 1) * nx + kx] * (double)((ky - 1) / 2) * scale1;
  }
}

/* cofdxdx.f90:1-27 */
void o_cofdxdx(o_sim *s, const double *fk, double *out) {
  const int nx = s->nx, ny = s->ny;
  double a = 2.0 * s->pi / s->p.xl, scale = a * a;
  for (int ky = 0; ky < ny; ky++)
    for (int kx = 0; kx <= nx - 2; kx += 2) {
      double k = (double)(kx / 2), k2 = k * k;
      out[(size_t)ky * nx + kx] = -fk[(size_t)ky * nx + kx] * k2 * scale;
      out[(size_t)ky * nx + kx + 1] = -fk[(size_t)ky * nx + kx + 1] * k2 * scale;
    }
}

/* cofdydy.f90:1-27 */
void o_cofdydy(o_sim *s, const double *fk, double *out) {
  const int nx = s->nx, ny = s->ny;
  double a = 2.0 * s->pi / s->p.yl, scale = a * a;
  for (int kx = 0; kx < nx; kx++)
    for (int ky = 0; ky <= ny - 2; ky += 2) {
      double k = (double)(ky / 2), k2 = k * k;
      out[(size_t)ky * nx + kx] = -fk[(size_t)ky * nx + kx] * k2 * scale;
      out[(size_t)(ky + 1) * nx + kx] = -fk[(size_t)(ky + 1) * nx + kx] * k2 * scale;
    }
}

/* cofdxdy.f90:1-40 */
void o_cofdxdy(o_sim *s, const double *fk, double *out) {
  const int nx = s->nx, ny = s->ny;
  double scale = (2.0 * s->pi / s->p.xl) * (2.0 * s->pi / s->p.yl);
#define F(ix, iy) fk[(size_t)(iy) * nx + (ix)]
#define O(ix, iy) out[(size_t)(iy) * nx + (ix)]
  for (int ky = 0; ky <= ny / 2 - 1; ky++)
    for (int kx = 0; kx <= nx / 2 - 1; kx++) {
      double fac = -(double)kx * (double)ky * scale;
      O(2 * kx, 2 * ky) = fac * F(2 * kx + 1, 2 * ky + 1);
      O(2 * kx + 1, 2 * ky) = -fac * F(2 * kx, 2 * ky + 1);
      O(2 * kx, 2 * ky + 1) = -fac * F(2 * kx + 1, 2 * ky);
      O(2 * kx + 1, 2 * ky + 1) = fac * F(2 * kx, 2 * ky);
    }
#undef F
#undef O
}

/* poisson.f90:1-30 */
void o_poisson(o_sim *s, const double *fk, double *ans) {
  const int nx = s->nx, ny = s->ny;
#define F(ix, iy) fk[(size_t)(iy) * nx + (ix)]
#define A(ix, iy) ans[(size_t)(iy) * nx + (ix)]
  for (int ky = 0; ky <= ny / 2 - 1; ky++)
    for (int kx = 0; kx <= nx / 2 - 1; kx++) {
      if (kx == 0 && ky == 0) {
        A(0, 0) = 0.0; A(1, 0) = 0.0; A(0, 1) = 0.0; A(1, 1) = 0.0;
      } else {
        double quot = ((double)(kx * kx) * s->scalex + (double)(ky * ky) * s->scaley);
        A(2 * kx, 2 * ky) = F(2 * kx, 2 * ky) / quot;
        A(2 * kx + 1, 2 * ky) = F(2 * kx + 1, 2 * ky) / quot;
        A(2 * kx, 2 * ky + 1) = F(2 * kx, 2 * ky + 1) / quot;
        A(2 * kx + 1, 2 * ky + 1) = F(2 * kx + 1, 2 * ky + 1) / quot;
      }
    }
#undef F
#undef A
}

/* cal_vis.f90:1-36 */
void o_vis(o_sim *s, double dt, double *work) {
  const int nx = s->nx, ny = s->ny;
  double coefx = -(dt * s->p.nu * s->scalex);
  double coefy = -(dt * s->p.nu * s->scaley);
  for (int ky = 0; ky <= ny / 2 - 1; ky++)
    for (int kx = 0; kx <= nx / 2 - 1; kx++)
      work[(size_t)(2 * ky) * nx + 2 * kx] =
          exp(coefx * (double)(kx * kx) + coefy * (double)(ky * ky));
  for (int ky = 0; ky <= ny / 2 - 1; ky++)
    for (int kx = 0; kx <= nx / 2 - 1; kx++) {
      double v = work[(size_t)(2 * ky) * nx + 2 * kx];
      work[(size_t)(2 * ky) * nx + 2 * kx + 1] = v;
      work[(size_t)(2 * ky + 1) * nx + 2 * kx] = v;
      work[(size_t)(2 * ky + 1) * nx + 2 * kx + 1] = v;
    }
}

/* dealiase_mask.f90:1-33 */
void o_dealiase_mask(o_sim *s, double *d) {
  const int nx = s->nx, ny = s->ny;
  int kx_max = nx / 2 - 1, ky_max = ny / 2 - 1;
  double kx_trunc = 2.0 / 3.0 * (double)kx_max;
  double ky_trunc = 2.0 / 3.0 * (double)ky_max;
  for (size_t i = 0; i < (size_t)nx * ny; i++) d[i] = 1.0;
  for (int ky = 0; ky <= ky_max; ky++)
    for (int kx = 0; kx <= kx_max; kx++) {
      double a = (double)kx, b = (double)ky;
      if (a * a / (kx_trunc * kx_trunc) + b * b / (ky_trunc * ky_trunc) >= 1.0) {
        d[(size_t)(2 * ky) * nx + 2 * kx] = 0.0;
        d[(size_t)(2 * ky) * nx + 2 * kx + 1] = 0.0;
        d[(size_t)(2 * ky + 1) * nx + 2 * kx] = 0.0;
        d[(size_t)(2 * ky + 1) * nx + 2 * kx + 1] = 0.0;
      }
    }
}

/* ========================================================================= */
/* kinematics                                                                 */
/* ========================================================================= */
/* mean_velocity.f90:4-84 */
double o_mean_ux(o_sim *s, double time) {
  const o_params *p = &s->p;
  double tau1, tf;
  switch (p->iMeanVelocity) {
  case 1:
    tau1 = p->T_fullspeed;
    if (time < tau1) tf = (time * time * time) / (-0.5 * tau1 * tau1 * tau1) + 3 * (time * time) / (tau1 * tau1);
    else tf = 1.0;
    return cos(p->theta_inf) * tf;
  case 2:
    tau1 = p->T_fullspeed;
    if (time < tau1) tf = 0.5 * (1.0 - cos(0.5 * s->pi * time / tau1));
    else tf = 1.0;
    return cos(p->theta_inf) * tf;
  case 3: return cos(p->theta_inf);
  case 4: return 0.0;
  case 6: return time;
  case 11: return s->u_mean[0];
  case 22: return 0.1;
  default: return 0.0; /* undefined in the reference */
  }
}

double o_mean_uy(o_sim *s, double time) {
  const o_params *p = &s->p;
  double tau1, tf;
  switch (p->iMeanVelocity) {
  case 1:
    tau1 = p->T_fullspeed;
    if (time < tau1) tf = (time * time * time) / (-0.5 * tau1 * tau1 * tau1) + 3 * (time * time) / (tau1 * tau1);
    else tf = 1.0;
    return sin(p->theta_inf) * tf;
  case 2:
    tau1 = p->T_fullspeed;
    if (time < tau1) tf = 0.5 * (1.0 - cos(0.5 * s->pi * time / tau1));
    else tf = 1.0;
    return sin(p->theta_inf) * tf;
  case 3: return sin(p->theta_inf);
  case 4: return 0.0;
  case 6: return 0.0;
  case 11: return s->u_mean[1];
  default: return 0.0;
  }
}

/* mouvement.f90:5-241, plus two prescribed-motion cases that the reference
 * does not have (600: sinusoidal heave, 601: Fourier heave + pitch). */
/* periods of the heaving cases 401..413, mouvement.f90:81-94 */
static const double o_ts400[13] = {0.51, 0.34, 0.25, 0.20, 0.17, 0.14, 0.13, 0.11, 0.10, 0.09, 0.08, 0.08, 0.07};

/* soft-startup time: the cases 401..413 overwrite the global tau with their
   period on every mouvement call (mouvement.f90:102-104) */
static double o_tau(const o_sim *s, int im) {
  return (im >= 401 && im <= 413) ? o_ts400[im - 401] : s->p.tau;
}

void o_mouvement(o_sim *s, double time, const o_beam *b, double *alpha,
                 double *alpha_t, double *alpha_tt, double LE[6]) {
  const double pi = s->pi;
  int im = b->iMouvement;
  for (int i = 0; i < 6; i++) LE[i] = 0.0;
  *alpha = 0.0; *alpha_t = 0.0; *alpha_tt = 0.0;
  if (im == 0) { /* :19-28 */
    LE[0] = b->x0; LE[1] = b->y0;
    *alpha = b->AngleBeam * pi / 180.0;
  } else if ((im >= 201 && im <= 213) || (im >= 300 && im <= 350)) { /* :29-75 */
    static const double ts[13] = {1.0, 1.09, 1.20, 1.29, 1.33, 1.4, 1.5, 1.71, 2.0, 2.4, 3.0, 4.0, 1.62};
    double y_max = 0.5, t_max = (im < 300) ? ts[im - 200 - 1] : 1.62;
    double k, kt, ktt;
    if (time <= 1.0) {
      double a = -20.0, bb = 70.0, c = -84.0, d = 35.0;
      double t2 = time * time, t3 = t2 * time, t4 = t3 * time, t5 = t4 * time, t6 = t5 * time, t7 = t6 * time;
      k = a * t7 + bb * t6 + c * t5 + d * t4;
      kt = 7.0 * a * t6 + 6.0 * bb * t5 + 5.0 * c * t4 + 4.0 * d * t3;
      ktt = 42. * a * t5 + 30. * bb * t4 + 20. * c * t3 + 12. * d * t2;
    } else { k = 1.0; kt = 0.0; ktt = 0.0; }
    double arg = 2.0 * pi * time / t_max + pi / 2. + b->phase;
    double y = y_max * sin(arg);
    double yt = y_max * (2.0 * pi / t_max) * cos(arg);
    double w = 2.0 * pi / t_max;
    double ytt = -(y_max * (w * w) * sin(arg));
    LE[0] = b->x0; LE[1] = b->y0 + k * y; LE[2] = 0.0;
    LE[3] = kt * y + yt * k; LE[4] = 0.0; LE[5] = ktt * y + ytt * k + 2. * kt * yt;
    *alpha = b->AngleBeam * pi / 180.0;
  } else if (im >= 401 && im <= 413) { /* :77-117 (also sets tau = t_max, see o_tau) */
    double y_max = 0.0333, t_max = o_ts400[im - 401], w = 2.0 * pi / t_max;
    LE[0] = b->x0;
    LE[1] = b->y0 + y_max * sin(2.0 * pi * time / t_max + pi / 2);
    LE[3] = y_max * (2.0 * pi / t_max) * cos(2.0 * pi * time / t_max + pi / 2);
    LE[5] = -(y_max * (w * w) * sin(2.0 * pi * time / t_max + pi / 2));
    *alpha = b->AngleBeam * pi / 180.0;
  } else if (im == 2) { /* :118-133 */
    double y_max = 0.5, t_max = 2.5, w = 2.0 * pi / t_max;
    LE[0] = b->x0;
    LE[1] = b->y0 + y_max * sin(2.0 * pi * time / t_max + pi / 2);
    LE[3] = y_max * (2.0 * pi / t_max) * cos(2.0 * pi * time / t_max + pi / 2);
    LE[5] = -(y_max * (w * w) * sin(2.0 * pi * time / t_max + pi / 2));
    *alpha = b->AngleBeam * pi / 180.0;
  } else if (im == 3) { /* :134-147 */
    double angle_max = 80.0 * pi / 180.0, f = 0.5;
    LE[0] = b->x0; LE[1] = b->y0;
    *alpha = angle_max * sin(2.0 * pi * f * time);
    *alpha_t = angle_max * cos(2.0 * pi * f * time) * (2.0 * pi * f);
    double w = 2.0 * pi * f;
    *alpha_tt = -1.0 * angle_max * sin(2.0 * pi * f * time) * (w * w);
  } else if (im == 4) { /* :148-160 */
    double u0 = -1.0;
    LE[0] = b->x0 + u0 * time; LE[2] = u0; LE[1] = b->y0; LE[3] = 0.0;
    *alpha = b->AngleBeam * pi / 180.0;
  } else if (im == 600) {
    /* no reference counterpart: heave y = A sin(2 pi f t) */
    double A = s->p.heave_amp, w = 2.0 * pi * s->p.heave_freq;
    LE[0] = b->x0;
    LE[1] = b->y0 + A * sin(w * time);
    LE[3] = A * w * cos(w * time);
    LE[5] = -(A * (w * w) * sin(w * time));
    *alpha = b->AngleBeam * pi / 180.0;
  } else if (im == 601) {
    /* no reference counterpart: h(t) = sum_m a_m sin(2 pi m t/T) + b_m cos(...) - b_m,
       alpha(t) likewise with c_m, d_m; m = 1, 2; T = 1/heave_freq */
    const double *c = s->p.motion_coef;
    double w0 = 2.0 * pi * s->p.heave_freq;
    double h = 0, ht = 0, htt = 0, a = 0, at = 0, att = 0;
    for (int m = 1; m <= 2; m++) {
      double wm = w0 * m, sn = sin(wm * time), cs = cos(wm * time);
      double ah = c[2 * (m - 1)], bh = c[2 * (m - 1) + 1];
      double ap = c[4 + 2 * (m - 1)], bp = c[4 + 2 * (m - 1) + 1];
      h += ah * sn + bh * (cs - 1.0);
      ht += wm * (ah * cs - bh * sn);
      htt += -(wm * wm) * (ah * sn + bh * cs);
      a += ap * sn + bp * (cs - 1.0);
      at += wm * (ap * cs - bp * sn);
      att += -(wm * wm) * (ap * sn + bp * cs);
    }
    LE[0] = b->x0; LE[1] = b->y0 + h; LE[3] = ht; LE[5] = htt;
    *alpha = b->AngleBeam * pi / 180.0 + a; *alpha_t = at; *alpha_tt = att;
  }
}

/* integrate_position.f90:1-47 */
void o_integrate_position(o_sim *s, double time, o_beam *b) {
  double alpha, alpha_t, alpha_tt, LE[6];
  o_mouvement(s, time, b, &alpha, &alpha_t, &alpha_tt, LE);
  const int ns = s->ns;
  const double ds = s->ds;
  for (int i = 0; i < ns; i++) { b->x[i] = 0.0; b->y[i] = 0.0; b->vx[i] = 0.0; b->vy[i] = 0.0; }
  for (int sidx = 0; sidx < ns; sidx++) {
    double xrel = 0.0, yrel = 0.0, vxrel = 0.0, vyrel = 0.0;
    for (int i = 0; i <= sidx - 1; i++) {
      xrel = xrel + cos(b->theta[i]) * ds;
      yrel = yrel + sin(b->theta[i]) * ds;
      vxrel = vxrel - b->theta_dot[i] * sin(b->theta[i]) * ds;
      vyrel = vyrel + b->theta_dot[i] * cos(b->theta[i]) * ds;
    }
    b->x[sidx] = LE[0] + xrel * cos(alpha) - yrel * sin(alpha);
    b->y[sidx] = LE[1] + yrel * cos(alpha) + xrel * sin(alpha);
    b->vx[sidx] = LE[2] + cos(alpha) * (vxrel - alpha_t * yrel) - sin(alpha) * (vyrel + alpha_t * xrel);
    b->vy[sidx] = LE[3] + cos(alpha) * (vyrel + alpha_t * xrel) + sin(alpha) * (vxrel - alpha_t * yrel);
  }
}

/* init_beam.f90:48-169 for beam ib (0-based): the first at (x0, y0) with
 * phase 0; the second ("dragonfly", :50-76) at (x0 + 1.25, y0), the same
 * motion, which must be one of 300-307, with phase (iMotion - 300) pi / 4 */
void o_init_beam_k(o_sim *s, o_beam *b, int ib) {
  const int ns = s->ns;
  memset(b, 0, sizeof(*b));
  b->x = dalloc(ns); b->y = dalloc(ns); b->vx = dalloc(ns); b->vy = dalloc(ns);
  b->theta = dalloc(ns); b->theta_dot = dalloc(ns);
  b->pressure_new = dalloc(ns); b->tau_new = dalloc(ns);
  b->x0 = s->p.x0; b->y0 = s->p.y0; b->iMouvement = s->p.iMotion;
  b->AngleBeam = s->p.AngleBeam; b->phase = 0.0;
  if (ib == 1) {
    b->x0 = s->p.x0 + 1.25;
    switch (b->iMouvement) {
      case 300: b->phase = 0.00 * s->pi; break;
      case 301: b->phase = 0.25 * s->pi; break;
      case 302: b->phase = 0.50 * s->pi; break;
      case 303: b->phase = 0.75 * s->pi; break;
      case 304: b->phase = 1.00 * s->pi; break;
      case 305: b->phase = 1.25 * s->pi; break;
      case 306: b->phase = 1.50 * s->pi; break;
      case 307: b->phase = 1.75 * s->pi; break;
      default: fprintf(stderr, "oracle: error...iMotion bad bad (init_beam.f90:72-73)\n"); abort();
    }
  }
  double alpha, at, att, LE[6];
  o_mouvement(s, 0.0, b, &alpha, &at, &att, LE);
  for (int n = 0; n < ns; n++) {
    b->x[n] = LE[0] + (double)n * s->ds;
    b->y[n] = LE[1];
  }
  b->UnsteadyCorrectionsReady = 0;
  o_integrate_position(s, 0.0, b);
}

void o_init_beam(o_sim *s, o_beam *b) { o_init_beam_k(s, b, 0); }

void o_free_beam(o_beam *b) {
  free(b->x); free(b->y); free(b->vx); free(b->vy); free(b->theta);
  free(b->theta_dot); free(b->pressure_new); free(b->tau_new);
  memset(b, 0, sizeof(*b));
}

/* ========================================================================= */
/* mask (create_mask.f90)                                                     */
/* ========================================================================= */
/* SmoothStep: create_mask.f90:475-520 */
void o_smoothstep(o_sim *s, double *f, double x, double t, double h) {
  if (!s->sharp) {
    double GradientERF = fabs((exp(-(2.0 * 1.0) * (2.0 * 1.0)) - 1.0) / sqrt(s->pi));
    double delta = h * GradientERF;
    *f = 0.5 * (erf((t - x) / delta) + erf((x + t) / delta));
  } else {
    if (x <= t) *f = 1.0;
    if (x > t) *f = 0.0;
  }
}

/* ConvertAngle: create_mask.f90:548-559 */
static double convert_angle(double pi, double a) {
  return (a < 0.0) ? 2.0 * pi + a : a;
}

static inline int fnint(double x) { return (int)lround(x); }
static inline int imax(int a, int b) { return a > b ? a : b; }
static inline int imin(int a, int b) { return a < b ? a : b; }

/* DrawADot: create_mask.f90:663-704 (full-grid loop) */
static void draw_a_dot(o_sim *s, double px, double py, double vx, double vy,
                       double t, double N) {
  const int nx = s->nx, ny = s->ny;
  const double dx = s->dx, dy = s->dy;
  double hh = N * fmax(dx, dy);
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++) {
      double a = (double)i * dx - px, b = (double)j * dy - py;
      double R = sqrt(a * a + b * b);
      double temp;
      o_smoothstep(s, &temp, R, t, hh);
      size_t k = (size_t)j * nx + i;
      if (temp > s->mask[k]) { s->mask[k] = temp; s->maskvx[k] = vx; s->maskvy[k] = vy; }
    }
}

/* the reach of segments and hinges: create_mask.f90:579, :737
 * (t* = t + 2 N max(dx,dy)); with iSmoothCutoff = 1 (no reference
 * counterpart) t + 7 delta, delta = N max(dx,dy) GradientERF (:511-512) */
static double o_t_cut(const o_sim *s, double t, double N) {
  if (!s->p.iSmoothCutoff) return t + 2.0 * N * fmax(s->dx, s->dy);
  double GradientERF = fabs((exp(-(2.0 * 1.0) * (2.0 * 1.0)) - 1.0) / sqrt(s->pi));
  return t + 7.0 * (N * fmax(s->dx, s->dy) * GradientERF);
}

/* DrawHinge: create_mask.f90:563-656 */
static void draw_hinge(o_sim *s, double px, double py, double alpha1,
                       double alpha2, double vx, double vy, double t,
                       double N) {
  const int nx = s->nx, ny = s->ny;
  const double dx = s->dx, dy = s->dy, pi = s->pi;
  double t_star = o_t_cut(s, t, N);
  int ixmin = imax(fnint((px - t_star) / dx), 0);
  int ixmax = imin(fnint((px + t_star) / dx), nx - 1);
  int iymin = imax(fnint((py - t_star) / dy), 0);
  int iymax = imin(fnint((py + t_star) / dy), ny - 1);
  double beta1 = convert_angle(pi, alpha1) * 180.0 / pi;
  double beta2 = convert_angle(pi, alpha2) * 180.0 / pi;
  for (int i = ixmin; i <= ixmax; i++)
    for (int j = iymin; j <= iymax; j++) {
      double a = (double)i * dx - px, b = (double)j * dy - py;
      double R = sqrt(a * a + b * b);
      if (R <= t_star) {
        double gamma = atan2((double)j * dy - py, (double)i * dx - px);
        gamma = convert_angle(pi, gamma) * 180.0 / pi;
        int in1 = 0, in2 = 0;
        if (beta1 < 90.0) {
          if ((gamma >= 0.0 && gamma <= 90.0 + beta1) || (gamma >= 270.0 + beta1 && gamma <= 360.0)) in1 = 1;
        } else if (beta1 >= 90.0 && beta1 < 270.0) {
          if (gamma >= 270.0 + beta1 - 360.0 && gamma <= 90.0 + beta1) in1 = 1;
        } else if (beta1 >= 270.0) {
          if ((gamma >= 0.0 && gamma <= 90.0 + beta1 - 360.0) || (gamma >= 270.0 + beta1 - 360.0 && gamma <= 360.0)) in1 = 1;
        }
        if (beta2 <= 90.0) {
          if (gamma >= 90.0 + beta2 && gamma < 270.0 + beta2) in2 = 1;
        } else if (beta2 >= 90.0 && beta2 < 270.0) {
          if ((gamma >= 0.0 && gamma <= 270.0 + beta2 - 360.0) || (gamma >= 90.0 + beta2 && gamma <= 360.0)) in2 = 1;
        } else if (beta2 >= 270.0) {
          if (gamma >= 90.0 + beta2 - 360.0 && gamma < 270.0 + beta2 - 360.0) in2 = 1;
        }
        if (in1 && in2) {
          double temp;
          o_smoothstep(s, &temp, R, t, N * fmax(dx, dy));
          size_t k = (size_t)j * nx + i;
          if (temp > s->mask[k]) { s->mask[k] = temp; s->maskvx[k] = vx; s->maskvy[k] = vy; }
        }
      }
    }
}

/* DrawBeamSegment: create_mask.f90:708-780 */
static void draw_beam_segment(o_sim *s, double p1x, double p1y, double p2x,
                              double p2y, double v1x, double v1y, double v2x,
                              double v2y, double t, double N) {
  const int nx = s->nx, ny = s->ny;
  const double dx = s->dx, dy = s->dy, ds = s->ds;
  double t_star = o_t_cut(s, t, N);
  double alpha = -atan2((p2y - p1y), (p2x - p1x));
  int ixmin = imax(fnint((fmin(p1x, p2x) - t_star) / dx), 0);
  int ixmax = imin(fnint((fmax(p1x, p2x) + t_star) / dx), nx - 1);
  int iymin = imax(fnint((fmin(p1y, p2y) - t_star) / dy), 0);
  int iymax = imin(fnint((fmax(p1y, p2y) + t_star) / dy), ny - 1);
  double ca = cos(alpha), sa = sin(alpha);
  double p1sx = ca * p1x - sa * p1y, p1sy = sa * p1x + ca * p1y;
  double p2sx = ca * p2x - sa * p2y, p2sy = sa * p2x + ca * p2y;
  for (int i = ixmin; i <= ixmax; i++)
    for (int j = iymin; j <= iymax; j++) {
      double xs = (double)i * dx * ca - (double)j * dy * sa;
      double ys = (double)i * dx * sa + (double)j * dy * ca;
      if (xs >= p1sx && xs <= p2sx && ys <= (p1sy + t_star) && ys >= (p2sy - t_star)) {
        double temp;
        o_smoothstep(s, &temp, fabs(ys - p1sy), t, N * fmax(dx, dy));
        size_t k = (size_t)j * nx + i;
        if (temp > s->mask[k]) {
          s->mask[k] = temp;
          s->maskvx[k] = (v1x + (xs - p1sx) * (v2x - v1x) / ds);
          s->maskvy[k] = (v1y + (xs - p1sx) * (v2y - v1y) / ds);
        }
      }
    }
}

/* DrawSharpEnd: create_mask.f90:784-852 */
static void draw_sharp_end(o_sim *s, double pex, double pey, double alpha,
                           double vx, double vy, double t, double N) {
  const int nx = s->nx, ny = s->ny;
  const double dx = s->dx, dy = s->dy;
  double mdx = fmax(dx, dy);
  double t_star = t + 2.0 * N * mdx;
  double p1x = pex - 2.0 * mdx * cos(alpha), p1y = pey - 2.0 * mdx * sin(alpha);
  double p2x = pex + 4.0 * mdx * cos(alpha), p2y = pey + 4.0 * mdx * sin(alpha);
  int ixmin = imax(fnint((fmin(p1x, p2x) - t_star) / dx), 0);
  int ixmax = imin(fnint((fmax(p1x, p2x) + t_star) / dx), nx - 1);
  int iymin = imax(fnint((fmin(p1y, p2y) - t_star) / dy), 0);
  int iymax = imin(fnint((fmax(p1y, p2y) + t_star) / dy), ny - 1);
  double ca = cos(alpha), sa = sin(alpha);
  double p1sx = ca * p1x + sa * p1y, p1sy = -sa * p1x + ca * p1y;
  double p2sx = ca * p2x + sa * p2y, p2sy = -sa * p2x + ca * p2y;
  for (int i = ixmin; i <= ixmax; i++)
    for (int j = iymin; j <= iymax; j++) {
      double xs = (double)i * dx * ca + (double)j * dy * sa;
      double ys = -((double)i * dx * sa) + (double)j * dy * ca;
      if (xs >= p1sx && xs <= p2sx && ys <= (p1sy + t_star) && ys >= (p2sy - t_star)) {
        double tmp2, temp;
        o_smoothstep(s, &tmp2, xs - p1sx, 2.0 * mdx, N * mdx);
        o_smoothstep(s, &temp, fabs(ys - p1sy), t, N * mdx);
        size_t k = (size_t)j * nx + i;
        s->mask[k] = temp * tmp2;
        if (xs - p1sx >= 2.0 * mdx) { s->maskvx[k] = vx; s->maskvy[k] = vy; }
      }
    }
}

/* create_mask: create_mask.f90:2-218 (walls, cylinder, beam; sponges 4 and 22
 * are not restated) */
/* SmoothStep2: create_mask.f90:523-544 (cosine step of half width h) */
static double o_smoothstep2(const o_sim *s, double x, double t, double h) {
  if (x <= t - h) return 1.0;
  if ((t - h) < x && x < (t + h)) return 0.5 * (1. + cos((x - t + h) * s->pi / (2.0 * h)));
  return 0.0;
}

/* SmoothFrameAroundDomain: create_mask.f90:370-463.  A level set phi of the
 * distance to the domain's edges (rounded corners of radius 2 thickness),
 * then mask1 = SmoothStep2(phi; thickness/2, thickness/2), thickness =
 * 4 t_beam.  Unset phi is 1000. */
void o_frame_mask(o_sim *s, double *mask1) {
  const int nx = s->nx, ny = s->ny;
  const double dx = s->dx, dy = s->dy, xl = s->p.xl, yl = s->p.yl;
  const size_t NN = (size_t)nx * ny;
  double *phi = dalloc(NN);
  for (size_t k = 0; k < NN; k++) phi[k] = 1000.0;
  const double thickness = 4.0 * s->p.t_beam, R1 = 0.0, R2 = 2.0 * thickness, RR = R1 + R2;
  const int ixmin = fnint(RR / dx), ixmax = nx - fnint(RR / dx);
  const int iymin = fnint(RR / dy), iymax = ny - fnint(RR / dy);
#define PHI(ix, iy) phi[(size_t)(iy) * nx + (ix)]
  for (int ix = ixmin; ix <= ixmax; ix++)
    for (int iy = 0; iy <= ny - 1; iy++) {
      if (ix < 0 || ix >= nx) continue;
      const double y = (double)iy * dy;
      const double m = fmin(PHI(ix, iy), fmin(fabs(y - R1), fabs(y - (yl - R1))));
      PHI(ix, iy) = (iy <= fnint(R1 / dy) || iy >= ny - 1 - fnint(R1 / dy)) ? -m : +m;
    }
  for (int ix = 0; ix <= nx - 1; ix++)
    for (int iy = iymin; iy <= iymax; iy++) {
      if (iy < 0 || iy >= ny) continue;
      const double x = (double)ix * dx;
      const double m = fmin(PHI(ix, iy), fmin(fabs(x - R1), fabs(x - (xl - R1))));
      PHI(ix, iy) = (ix <= fnint(R1 / dx) || ix >= nx - 1 - fnint(R1 / dx)) ? -m : +m;
    }
  /* the four corners */
  const double cx[4] = {RR, RR, xl - RR, xl - RR}, cy[4] = {RR, yl - RR, yl - RR, RR};
  const int x0s[4] = {0, 0, ixmax, ixmax}, x1s[4] = {ixmin, ixmin, nx - 1, nx - 1};
  const int y0s[4] = {0, iymax, iymax, 0}, y1s[4] = {iymin, ny - 1, ny - 1, iymin};
  for (int c = 0; c < 4; c++)
    for (int ix = x0s[c]; ix <= x1s[c]; ix++)
      for (int iy = y0s[c]; iy <= y1s[c]; iy++) {
        if (ix < 0 || ix >= nx || iy < 0 || iy >= ny) continue;
        const double x = (double)ix * dx, y = (double)iy * dy;
        const double a = x - cx[c], b = y - cy[c];
        PHI(ix, iy) = fmin(PHI(ix, iy), -(sqrt(a * a + b * b) - R2));
      }
  for (size_t k = 0; k < NN; k++) mask1[k] = o_smoothstep2(s, phi[k], thickness / 2., thickness / 2.);
#undef PHI
  free(phi);
}

void o_create_mask(o_sim *s, double time, o_beam *b) {
  const int nx = s->nx, ny = s->ny, ns = s->ns;
  const double dy = s->dy, N = s->p.N_smooth;
  size_t NN = (size_t)nx * ny;
  double h_star = s->p.h_channel + N * dy;
  memset(s->mask, 0, NN * sizeof(double));
  memset(s->maskvx, 0, NN * sizeof(double));
  memset(s->maskvy, 0, NN * sizeof(double));
  if (s->p.iWalls > 0) { /* :34-62 */
    int iymax = fnint(h_star / dy);
    for (int j = 0; j <= iymax && j < ny; j++) {
      double temp;
      o_smoothstep(s, &temp, (double)j * dy, s->p.h_channel, N * dy);
      for (int i = 0; i < nx; i++) s->mask[(size_t)j * nx + i] = temp;
    }
    int iymin = fnint((s->p.yl - h_star) / dy);
    for (int j = ny - 1; j >= iymin && j >= 0; j--) {
      double temp;
      o_smoothstep(s, &temp, fabs((double)j * dy - s->p.yl), s->p.h_channel, N * dy);
      for (int i = 0; i < nx; i++) s->mask[(size_t)j * nx + i] = temp;
    }
  }
  if (s->p.iCylinder == 1) /* :67-69 */
    draw_a_dot(s, s->p.x0 - s->p.R_cylinder, s->p.y0, 0.0, 0.0, s->p.R_cylinder, N);
  for (int ib = 0; ib < s->nb; ib++, b++) { /* :74-119, do ib = 1, iBeam */
    double alpha, at, att, LE[6];
    o_mouvement(s, time, b, &alpha, &at, &att, LE);
    const double t = s->p.t_beam;
    for (int i = 0; i <= ns - 2; i++)
      draw_beam_segment(s, b->x[i], b->y[i], b->x[i + 1], b->y[i + 1], b->vx[i], b->vy[i],
                        b->vx[i + 1], b->vy[i + 1], t, N);
    if (s->p.iSharpTrailing == 1)
      draw_sharp_end(s, b->x[ns - 1], b->y[ns - 1], b->theta[ns - 1] + alpha, b->vx[ns - 1],
                     b->vy[ns - 1], t, N);
    for (int i = 1; i <= ns - 2; i++) {
      double alpha1 = atan2((b->y[i] - b->y[i - 1]), (b->x[i] - b->x[i - 1]));
      double alpha2 = atan2((b->y[i + 1] - b->y[i]), (b->x[i + 1] - b->x[i]));
      draw_hinge(s, b->x[i], b->y[i], alpha1, alpha2, b->vx[i], b->vy[i], t, N);
    }
    draw_a_dot(s, b->x[0], b->y[0], b->vx[0], b->vy[0], t, N);
    if (s->p.iSharpTrailing == 0)
      draw_a_dot(s, b->x[ns - 1], b->y[ns - 1], b->vx[ns - 1], b->vy[ns - 1], t, N);
  }
  if (s->p.iSponge == 4) { /* velocity sponge of a channel inflow, :130-184 */
    const double dx = s->dx, S = s->p.SpongeSize, ssz = S - 2.0 * N * dx;
    const double H_eff = s->p.yl - 2.0 * s->p.h_channel, mux = o_mean_ux(s, time);
    for (int half = 0; half < 2; half++) {
      const int ixmin = half ? fnint((0.5 * S) / dx) : 0;
      const int ixmax = half ? fnint((1.0 * S) / dx) : fnint((0.5 * S) / dx);
      for (int ix = ixmin; ix <= ixmax; ix++)
        for (int iy = 0; iy < ny; iy++) {
          double tmp;
          const double y_chan = (double)iy * dy - s->p.h_channel;
          if (half == 0) o_smoothstep(s, &tmp, 0.75 * S - (double)ix * dx, ssz * 0.5, 0.25 * ssz);
          else o_smoothstep(s, &tmp, (double)ix * dx - 0.5 * S, ssz * 0.5, N * dx);
          const size_t k = (size_t)iy * nx + ix;
          if (s->mask[k] < tmp) s->mask[k] = tmp;
          if (y_chan > 0.0 && y_chan < H_eff)
            s->maskvx[k] = 1.5 * mux * y_chan * (H_eff - y_chan) / ((0.5 * H_eff) * (0.5 * H_eff));
          else
            s->maskvx[k] = 0.0;
        }
    }
  }
  if (s->p.iSponge == 22) { /* frame around the domain, :187-202 */
    if (!s->frame_done) { o_frame_mask(s, s->mask_frame); s->frame_done = 1; }
    const double w = s->p.eps / fmax(1.0e-3, s->p.eps);
    const double mux = o_mean_ux(s, time), muy = o_mean_uy(s, time);
    for (size_t k = 0; k < NN; k++) s->mask[k] = s->mask[k] + s->mask_frame[k] * w;
    for (size_t k = 0; k < NN; k++)
      if (s->mask_frame[k] > 0.0) { s->maskvx[k] = mux; s->maskvy[k] = muy; }
  }
  for (size_t k = 0; k < NN; k++) s->mask[k] = s->mask[k] / s->p.eps; /* :210-214 */
}

/* create_sponge_mask: create_mask.f90:223-364 (types 0,1,2,3,10) */
void o_create_sponge_mask(o_sim *s) {
  const int nx = s->nx, ny = s->ny;
  const double dx = s->dx, dy = s->dy, N = s->p.N_smooth, xl = s->p.xl, yl = s->p.yl;
  double *ms = s->mask_sponge;
  size_t NN = (size_t)nx * ny;
  memset(ms, 0, NN * sizeof(double));
  int isp = s->p.iSponge;
  if (isp == 1) {
    double sss = s->p.SpongeSize, ssz = sss - 2.0 * N * dx, sb = xl - sss;
    int ixmax = fnint((xl - 0.5 * sss) / dx), ixmin = fnint(sb / dx);
    for (int ix = ixmin; ix <= ixmax; ix++) {
      double tmp, x = xl - 0.5 * sss - (double)ix * dx;
      o_smoothstep(s, &tmp, x, ssz * 0.5, N * dx);
      for (int iy = 0; iy < ny; iy++) ms[(size_t)iy * nx + ix] = tmp;
    }
    ixmin = fnint((xl - 0.5 * sss) / dx); ixmax = nx - 1;
    for (int ix = ixmin; ix <= ixmax; ix++) {
      double tmp, x = (double)ix * dx - xl + 0.5 * sss - 1.0 * dx;
      o_smoothstep(s, &tmp, x, ssz * 0.4, N * dx);
      for (int iy = 0; iy < ny; iy++) ms[(size_t)iy * nx + ix] = tmp;
    }
  } else if (isp == 22) { /* :333-349, after the startup mask */
    const int ixmin = fnint(s->p.SpongeSize / dx), iymin = fnint(s->p.SpongeSize / dy);
    const int ixmax = fnint((xl - s->p.SpongeSize) / dx), iymax = fnint((yl - s->p.SpongeSize) / dy);
    for (size_t k = 0; k < NN; k++) ms[k] = 1.0;
    for (int iy = iymin; iy <= iymax; iy++)
      for (int ix = ixmin; ix <= ixmax; ix++)
        if (ix >= 0 && ix < nx && iy >= 0 && iy < ny) ms[(size_t)iy * nx + ix] = 0.0;
    for (size_t k = 0; k < NN; k++)
      if (s->mask[k] > 0.0) ms[k] = 0.0;
  } else if (isp == 10) {
    int ixmin = fnint((xl - s->p.SpongeSize) / dx);
    for (int ix = ixmin; ix < nx; ix++)
      for (int iy = 0; iy < ny; iy++) ms[(size_t)iy * nx + ix] = 1.0;
  } else if (isp == 2) {
    double sss = s->p.SpongeSize, ssz = sss - 2.0 * N * dx;
    int ixmax = fnint((0.5 * sss) / dx);
    for (int ix = 0; ix <= ixmax; ix++) {
      double tmp, x = 0.5 * sss - (double)ix * dx;
      o_smoothstep(s, &tmp, x, ssz * 0.5, N * dx);
      for (int iy = 0; iy < ny; iy++) ms[(size_t)iy * nx + ix] = tmp;
    }
    int ixmin = fnint((0.5 * sss) / dx);
    ixmax = fnint((1.0 * sss) / dx);
    for (int ix = ixmin; ix <= ixmax; ix++) {
      double tmp, x = (double)ix * dx - 0.5 * sss;
      o_smoothstep(s, &tmp, x, ssz * 0.5, N * dx);
      for (int iy = 0; iy < ny; iy++) ms[(size_t)iy * nx + ix] = tmp;
    }
    for (int ix = 0; ix <= ixmax; ix++)
      for (int iy = 0; iy < ny; iy++) {
        size_t k = (size_t)iy * nx + ix;
        if (s->mask[k] * s->p.eps > 0.0) ms[k] = ms[k] * (1.0 - s->mask[k] * s->p.eps);
      }
  } else if (isp == 3) {
    double sss = s->p.SpongeSize, ssz = sss - N * dx;
    for (int ix = 0; ix < nx; ix++)
      for (int iy = 0; iy < ny; iy++) {
        double t1, t2, t3, t4;
        o_smoothstep(s, &t1, (double)ix * dx, ssz, N * dx);
        o_smoothstep(s, &t2, xl - (double)ix * dx, ssz, N * dx);
        o_smoothstep(s, &t3, (double)iy * dy, ssz, N * dy);
        o_smoothstep(s, &t4, yl - (double)iy * dy, ssz, N * dy);
        double v = t1 + t2 + t3 + t4;
        if (v > 1.0) v = 1.0;
        ms[(size_t)iy * nx + ix] = v;
      }
  }
  for (size_t k = 0; k < NN; k++) ms[k] = ms[k] / s->p.eps_sponge; /* :355 */
}

/* ========================================================================= */
/* fluid solver (NavierStokes.f90)                                            */
/* ========================================================================= */
/* cal_velocity: NavierStokes.f90:181-212 */
void o_cal_velocity(o_sim *s, double time, const double *vortk, double *u,
                    double *stream) {
  size_t N = (size_t)s->nx * s->ny;
  double *work2 = dalloc(N);
  double *u1 = u, *u2 = u + N;
  o_poisson(s, vortk, stream);
  o_cofdx(s, stream, work2);
  o_cofitxy(s, work2, u2);
  o_cofdy(s, stream, work2);
  o_cofitxy(s, work2, u1);
  double mx = s->p.U_mean_true * o_mean_ux(s, time);
  double my = s->p.U_mean_true * o_mean_uy(s, time);
  for (size_t k = 0; k < N; k++) {
    u1[k] = u1[k] + mx;
    u2[k] = -u2[k] + my;
  }
  free(work2);
}

/* cal_nlk: NavierStokes.f90:9-177 */
/* the iSponge = 10 part of the x-penalisation (NavierStokes.f90:103-121):
 * w = -mask_sponge * du1/dx (spectral derivative), and for ixmin <= ix <=
 * ixmax (ixmin = nint((xl - SpongeSize)/dx), ixmax = nx - 16) the trapezoid
 * integral of w from ixmin to ix along each row is added to pen1 */
void o_sponge10_term(o_sim *s, const double *u1, double *pen1) {
  const o_params *p = &s->p;
  const int nx = s->nx, ny = s->ny;
  const size_t N = (size_t)nx * ny;
  const double dx = s->dx;
  double *work2 = dalloc(N), *dk = dalloc(N);
  o_coftxy(s, u1, work2); o_cofdx(s, work2, dk); o_cofitxy(s, dk, work2);
  for (size_t k = 0; k < N; k++) work2[k] = -(s->mask_sponge[k] * work2[k]);
  const int ixmin = fnint((p->xl - p->SpongeSize) / dx), ixmax = nx - 1 - 15;
  for (int iy = 0; iy < ny; iy++) {
    const double *w = work2 + (size_t)iy * nx;
    for (int ix = ixmin; ix <= ixmax; ix++) {
      double summe = 0.0;
      for (int is = ixmin; is <= ix - 1; is++) summe = summe + 0.5 * (w[is] + w[is + 1]) * dx;
      pen1[(size_t)iy * nx + ix] = pen1[(size_t)iy * nx + ix] + summe;
    }
  }
  free(work2); free(dk);
}

void o_cal_nlk(o_sim *s, double time, double *dt1, double *nlk,
               const double *vortk, const double *u) {
  const o_params *p = &s->p;
  size_t N = (size_t)s->nx * s->ny;
  const double *u1 = u, *u2 = u + N;
  double dx = s->dx, dy = s->dy;
  double *work2 = dalloc(N), *vdx = dalloc(N), *vdy = dalloc(N);
  double *pen1 = dalloc(N), *pen2 = dalloc(N), *penvortk = dalloc(N);
  o_cofdx(s, vortk, work2); o_cofitxy(s, work2, vdx);     /* :40-41 */
  o_cofdy(s, vortk, work2); o_cofitxy(s, work2, vdy);     /* :43-44 */
  for (size_t k = 0; k < N; k++) work2[k] = -(u1[k] * vdx[k]) - u2[k] * vdy[k]; /* :48 */
  o_coftxy(s, work2, nlk);                                 /* :52 */
  double mx = 0.0;                                         /* :59-65 */
  for (size_t k = 0; k < N; k++) {
    double w = u1[k] * u1[k] + u2[k] * u2[k];
    if (k == 0 || w > mx) mx = w;
  }
  double u_max = sqrt(mx);
  double d = p->cfl * fmin(dx, dy) / u_max;                /* :66 */
  if (u_max < 1.e-10) d = 1.e-2;                           /* :69 */
  if (p->iBeam > 0 || p->iWalls > 0 || p->iCylinder > 0) d = fmin(0.9 * p->eps, d);
  if (p->dt_fixed > 0.0) d = fmin(p->dt_fixed, d);
  if ((p->Time_end - time) < d) d = p->Time_end - time;
  *dt1 = d;
  s->u_max = u_max;
  if (p->iBeam > 0 || p->iWalls > 0 || p->iCylinder > 0 || p->iSponge == 22) { /* :88-132 */
    for (size_t k = 0; k < N; k++) {
      pen1[k] = -(s->mask[k] * (u1[k] - s->maskvx[k]));
      pen2[k] = -(s->mask[k] * (u2[k] - s->maskvy[k]));
    }
    s->fm_set = 0;
    if (p->iMeanVelocity == 11) { /* :97-100 */
      double a = 0.0, c = 0.0;
      for (size_t k = 0; k < N; k++) { a = a + pen1[k]; c = c + pen2[k]; }
      s->fm_new[0] = a / (double)N;
      s->fm_new[1] = c / (double)N;
      s->fm_set = 1;
    }
    if (p->iSponge == 10) o_sponge10_term(s, u1, pen1); /* :103-121 */
    o_coftxy(s, pen2, work2); o_cofdx(s, work2, pen2);
    o_coftxy(s, pen1, work2); o_cofdy(s, work2, pen1);
  } else {
    memset(pen1, 0, N * sizeof(double)); memset(pen2, 0, N * sizeof(double));
  }
  if ((p->iSponge > 0 && p->iSponge < 4) || p->iSponge == 22) { /* :142-164 */
    memcpy(penvortk, vortk, N * sizeof(double));
    o_cofitxy(s, penvortk, work2);
    for (size_t k = 0; k < N; k++) work2[k] = s->mask_sponge[k] * work2[k];
    o_coftxy(s, work2, penvortk);
    for (size_t k = 0; k < N; k++) nlk[k] = nlk[k] + (pen2[k] - pen1[k]) - penvortk[k];
  } else {
    for (size_t k = 0; k < N; k++) nlk[k] = nlk[k] + (pen2[k] - pen1[k]);
  }
  free(work2); free(vdx); free(vdy); free(pen1); free(pen2); free(penvortk);
}

/* cal_drag of one beam: NavierStokes.f90:523-589 (the loop body) */
static void cal_drag_one(o_sim *s, double time, double dt1, const double *penal,
                         o_beam *b) {
  const int nx = s->nx, ny = s->ny, ns = s->ns;
  const double dx = s->dx, dy = s->dy, norm = dx * dy;
  const o_params *p = &s->p;
  size_t N = (size_t)nx * ny;
  double alpha, at, att, LE[6];
  o_mouvement(s, time, b, &alpha, &at, &att, LE);
  int ixmin, ixmax, iymin, iymax;
  if (p->iBeam > 0) {
    double ymax = b->y[0], ymin = b->y[0], xmin = b->x[0], xmax = b->x[0];
    for (int i = 1; i < ns; i++) {
      if (b->y[i] > ymax) ymax = b->y[i];
      if (b->y[i] < ymin) ymin = b->y[i];
      if (b->x[i] < xmin) xmin = b->x[i];
      if (b->x[i] > xmax) xmax = b->x[i];
    }
    iymax = (int)((fmax(ymax + 1.5 * p->t_beam, LE[1] + 2.0 * p->R_cylinder)) / dy);
    iymin = (int)((fmin(ymin - 1.5 * p->t_beam, LE[1] - 2.0 * p->R_cylinder)) / dy);
    ixmin = (int)(fmin((LE[0] - fmax(2.5 * p->R_cylinder, 1.5 * p->t_beam)), (xmin - 1.5 * p->t_beam)) / dx);
    ixmax = (int)((xmax + 1.5 * p->t_beam) / dx);
    if (p->iSmoothCutoff) { /* every pixel within t + 7 delta + max(dx,dy) of a node */
      double r = o_t_cut(s, p->t_beam, p->N_smooth) + fmax(dx, dy);
      ixmin = imin(ixmin, fnint((xmin - r) / dx));
      ixmax = imax(ixmax, fnint((xmax + r) / dx));
      iymin = imin(iymin, fnint((ymin - r) / dy));
      iymax = imax(iymax, fnint((ymax + r) / dy));
    }
  } else {
    iymax = ny - 1; iymin = 0; ixmax = nx - 1; ixmin = 0;
  }
  double h_star = p->h_channel + p->N_smooth * dy;
  if (p->iWalls > 0) {
    if (ixmin < 0) ixmin = 0;
    if (ixmax > nx - 1) ixmax = nx - 1;
    if (iymin < fnint(h_star / dy) + 2) iymin = fnint(h_star / dy) + 2;
    if (iymax > fnint((p->yl - h_star) / dy) - 2) iymax = fnint((p->yl - h_star) / dy) - 2;
  }
  if (ixmin < 0) ixmin = 0;
  if (ixmax > nx - 1) ixmax = nx - 1;
  if (iymin < 0) iymin = 0;
  if (iymax > ny - 1) iymax = ny - 1;
  double s1 = 0.0, s2 = 0.0, s3 = 0.0, s4 = 0.0;
  for (int iy = iymin; iy <= iymax; iy++)
    for (int ix = ixmin; ix <= ixmax; ix++) {
      size_t k = (size_t)iy * nx + ix;
      s1 += penal[k];
      s2 += penal[N + k];
      s3 += s->maskvx[k] * s->mask[k];
      s4 += s->maskvy[k] * s->mask[k];
    }
  b->Force[0] = s1 * norm;
  b->Force[1] = s2 * norm;
  b->drag_unst_new = s3 * norm * p->eps;
  b->lift_unst_new = s4 * norm * p->eps;
  if (b->UnsteadyCorrectionsReady) {
    b->Force_unst[0] = (b->drag_unst_new - b->drag_unst_old) / dt1;
    b->Force_unst[1] = (b->lift_unst_new - b->lift_unst_old) / dt1;
  } else {
    b->Force_unst[0] = 0.0; b->Force_unst[1] = 0.0;
  }
  b->drag_unst_old = b->drag_unst_new;
  b->lift_unst_old = b->lift_unst_new;
  b->UnsteadyCorrectionsReady = 1;
}

/* cal_drag: NavierStokes.f90:512-593; b points to beams(1:iBeam) */
void o_cal_drag(o_sim *s, double time, double dt1, const double *penal,
                o_beam *b) {
  for (int ib = 0; ib < s->nb; ib++) cal_drag_one(s, time, dt1, penal, &b[ib]); /* :527-591 */
}

/* cal_energy: NavierStokes.f90:222-263 (values kept instead of appended) */
static void cal_energy(o_sim *s, double time, double dt1, const double *vor,
                       const double *u, const double *penal) {
  size_t N = (size_t)s->nx * s->ny;
  const double dx = s->dx, dy = s->dy;
  const double *u1 = u, *u2 = u + N;
  double mx = o_mean_ux(s, time + dt1), my = o_mean_uy(s, time + dt1);
  double a = 0, b = 0, c = 0, d = 0, e = 0;
  for (size_t k = 0; k < N; k++) a += vor[k] * vor[k];
  for (size_t k = 0; k < N; k++) {
    double p1 = u1[k] - mx, p2 = u2[k] - my;
    b += (p1 * p1 + p2 * p2);
  }
  for (size_t k = 0; k < N; k++) {
    double r1 = u1[k] - s->maskvx[k], r2 = u2[k] - s->maskvy[k];
    c += s->mask[k] * (r1 * r1 + r2 * r2);
  }
  for (size_t k = 0; k < N; k++)
    d += s->mask[k] * (s->maskvx[k] * (u1[k] - s->maskvx[k]) + s->maskvy[k] * (u2[k] - s->maskvy[k]));
  for (size_t k = 0; k < N; k++) e += mx * penal[k] + my * penal[N + k];
  s->energies[0] = dx * dy * sqrt(a);
  s->energies[1] = 0.5 * dx * dy * b;
  s->energies[2] = s->p.nu * dx * dy * a;
  s->energies[3] = dx * dy * c;
  s->energies[4] = dx * dy * d;
  s->energies[5] = dx * dy * s->p.U_mean_true * e;
}

/* pressure: NavierStokes.f90:273-403 */
void o_pressure(o_sim *s, double time, double dt1, const double *vortk,
                double *press, double *u, o_beam *b) {
  size_t N = (size_t)s->nx * s->ny;
  double *u1 = u, *u2 = u + N;
  double *penal = dalloc(2 * N), *work1 = dalloc(N), *work2 = dalloc(N);
  double *p1 = penal, *p2 = penal + N;
  o_cal_velocity(s, time + dt1, vortk, u, work1);          /* :305 */
  for (size_t k = 0; k < N; k++) {                         /* :311-317 */
    p1[k] = s->mask[k] * (u1[k] - s->maskvx[k]);
    p2[k] = s->mask[k] * (u2[k] - s->maskvy[k]);
  }
  if (s->p.iBeam > 0) o_cal_drag(s, time, dt1, penal, b);  /* :324 */
  o_cofitxy(s, vortk, work1);                              /* :331 */
  for (size_t k = 0; k < N; k++) s->vort[k] = work1[k];
  for (size_t k = 0; k < N; k++)                           /* :670-683 */
    s->vor_mean[k] = (s->vor_mean[k] * time + dt1 * work1[k]) / (time + dt1);
  cal_energy(s, time, dt1, work1, u, penal);               /* :338 */
  if (s->p.iFLUSI == 1 || s->p.iSaveBeam > 0) {            /* :344-392 */
    for (size_t k = 0; k < N; k++) {
      p1[k] = p1[k] - work1[k] * u2[k];
      p2[k] = p2[k] + work1[k] * u1[k];
    }
    o_coftxy(s, p1, work2); o_cofdx(s, work2, p1);
    o_coftxy(s, p2, work2); o_cofdy(s, work2, p2);
    for (size_t k = 0; k < N; k++) work1[k] = p1[k] + p2[k];
    o_poisson(s, work1, press);
    o_cofitxy(s, press, work1);
    for (size_t k = 0; k < N; k++) press[k] = work1[k] - 0.5 * (u1[k] * u1[k] + u2[k] * u2[k]);
  } else {
    memset(press, 0, N * sizeof(double));
  }
  free(penal); free(work1); free(work2);
}

/* EvolveFluidExplicit: NavierStokes.f90:422-459 */
void o_evolve_explicit(o_sim *s, double time, double *dt1, int n0, int n1,
                       o_beam *b) {
  size_t N = (size_t)s->nx * s->ny;
  double *stream = dalloc(N);
  o_cal_velocity(s, time, s->vortk[n0], s->u, stream);
  o_cal_nlk(s, time, dt1, s->nlk[n0], s->vortk[n0], s->u);
  o_vis(s, *dt1, s->workvis[n1]);
  for (size_t k = 0; k < N; k++)
    s->vortk[n1][k] = (s->vortk[n0][k] + *dt1 * s->nlk[n0][k]) * s->workvis[n1][k] * s->dealiase[k];
  if (s->p.iMeanVelocity == 11) { /* :453-456 */
    if (s->fm_set) { s->fmean[n0][0] = s->fm_new[0]; s->fmean[n0][1] = s->fm_new[1]; }
    s->u_mean[0] = s->u_mean[0] + s->fmean[n0][0] * *dt1;
    s->u_mean[1] = s->u_mean[1] + s->fmean[n0][1] * *dt1;
  }
  o_pressure(s, time, *dt1, s->vortk[n1], s->press, s->u, b);
  free(stream);
}

/* ComputePressureSnapshot: NavierStokes.f90:596-665, for vortk of Fortran
 * level `level` at `time`, with the current mask */
void o_pressure_snapshot(o_sim *s, double time, int level, double *press) {
  const size_t N = (size_t)s->nx * s->ny;
  double *u = dalloc(2 * N), *work1 = dalloc(N), *work2 = dalloc(N), *wdx = dalloc(N), *wdy = dalloc(N);
  double *pen1 = dalloc(N), *pen2 = dalloc(N);
  const double *vortk = s->vortk[level];
  o_cal_velocity(s, time, vortk, u, work1);                                  /* :622 */
  const double *u1 = u, *u2 = u + N;
  for (size_t k = 0; k < N; k++) {                                           /* :625-626 */
    pen1[k] = s->mask[k] * (u1[k] - s->maskvx[k]);
    pen2[k] = s->mask[k] * (u2[k] - s->maskvy[k]);
  }
  o_cofitxy(s, vortk, work1);                                                /* :629 */
  o_coftxy(s, pen1, work2); o_cofdx(s, work2, pen1);                         /* :635-636 */
  o_coftxy(s, pen2, work2); o_cofdy(s, work2, pen2);                         /* :638-639 */
  for (size_t k = 0; k < N; k++) { wdx[k] = -work1[k] * u2[k]; wdy[k] = +work1[k] * u1[k]; } /* :644-645 */
  o_coftxy(s, wdx, work1); o_cofdx(s, work1, wdx);                           /* :648-649 */
  o_coftxy(s, wdy, work1); o_cofdy(s, work1, wdy);                           /* :650-651 */
  for (size_t k = 0; k < N; k++) work1[k] = ((pen1[k] + pen2[k]) + wdy[k]) + wdx[k]; /* :654-655 */
  o_poisson(s, work1, work2);                                                /* :659 */
  o_cofitxy(s, work2, work1);                                                /* :660 */
  for (size_t k = 0; k < N; k++) press[k] = work1[k] - 0.5 * (u1[k] * u1[k] + u2[k] * u2[k]); /* :661 */
  free(u); free(work1); free(work2); free(wdx); free(wdy); free(pen1); free(pen2);
}

/* EvolveFluidAB2: NavierStokes.f90:463-508 */
void o_evolve_ab2(o_sim *s, double time, double dt0, double *dt1, int n0,
                  int n1, o_beam *b) {
  size_t N = (size_t)s->nx * s->ny;
  for (size_t k = 0; k < N; k++) s->nlk[n1][k] = s->nlk[n1][k] * s->workvis[n0][k];
  o_cal_nlk(s, time, dt1, s->nlk[n0], s->vortk[n0], s->u);
  double d1 = *dt1;
  double b10 = d1 / dt0 * (0.5 * d1 + dt0);
  double b11 = -(0.5 * (d1 * d1)) / dt0;
  o_vis(s, d1, s->workvis[n1]);
  for (size_t k = 0; k < N; k++)
    s->vortk[n1][k] = (s->vortk[n0][k] + b10 * s->nlk[n0][k] + b11 * s->nlk[n1][k]) *
                      s->workvis[n1][k] * s->dealiase[k];
  if (s->p.iMeanVelocity == 11) { /* :501-503 */
    if (s->fm_set) { s->fmean[n0][0] = s->fm_new[0]; s->fmean[n0][1] = s->fm_new[1]; }
    for (int c = 0; c < 2; c++) s->u_mean[c] = s->u_mean[c] + b10 * s->fmean[n0][c] + b11 * s->fmean[n1][c];
  }
  o_pressure(s, time, d1, s->vortk[n1], s->press, s->u, b);
}

/* ========================================================================= */
/* beam forces (Interpolation.f90:6-101, BeamForces.f90:6-509)               */
/* ========================================================================= */
/* delta(typ="phi_4star"): Interpolation.f90:51-101 */
double o_delta_phi4star(double x, double dx, double pi) {
  double r = fabs(x / dx);
  if (r < 0.5) return (3. / 8.) + (pi / 32.) - 0.25 * (r * r);
  if (r >= 0.5 && r <= 1.5)
    return 0.25 + (1 - r) / 8. * sqrt(-2. + 8. * r - 4. * (r * r)) - asin(sqrt(2.) * (r - 1.)) / 8.;
  if (r >= 1.5 && r <= 2.5)
    return (17. / 16.) - (pi / 64.) - (3. * r / 4.) + ((r * r) / 8.) +
           (r - 2.) * sqrt(-14. + 16. * r - 4. * (r * r)) / 16. + asin(sqrt(2.) * (r - 2.)) / 16.;
  return 0.0;
}

/* DeltaInterpolation (box 0,0,xl-dx,yl-dy): Interpolation.f90:6-48 */
double o_delta_interp(o_sim *s, double xt, double yt, const double *field,
                      int *err) {
  const int nx = s->nx, ny = s->ny;
  double x2 = s->p.xl - s->dx, y2 = s->p.yl - s->dy;
  double dx = (x2 - 0.0) / (double)(nx - 1);
  double dy = (y2 - 0.0) / (double)(ny - 1);
  if (xt > x2 || xt < 0.0 || yt > y2 || yt < 0.0) { if (err) *err = 1; return 0.0; }
  int i = (int)((xt - 0.0) / dx), j = (int)((yt - 0.0) / dy);
  if (i - 4 < 0 || i + 4 > nx - 1 || j - 4 < 0 || j + 4 > ny - 1) { if (err) *err = 2; return 0.0; }
  double R1 = 0.0;
  for (int ix = i - 4; ix <= i + 4; ix++)
    for (int iy = j - 4; iy <= j + 4; iy++) {
      double xl_ = (double)ix * dx + 0.0, yl_ = (double)iy * dy + 0.0;
      R1 = R1 + o_delta_phi4star(fabs(xl_ - xt), dx, s->pi) * o_delta_phi4star(fabs(yl_ - yt), dy, s->pi) *
                    field[(size_t)iy * nx + ix];
    }
  return R1;
}

/* BeamSurfaceIntegral: BeamForces.f90:306-509 */
static void beam_surface_integral(o_sim *s, double time, o_beam *b,
                                  const double *u, const double *sa,
                                  const double *sb, const double *press, double *bsi) {
  const int nc = 128, ns = s->ns;
  const double dx = s->dx, dy = s->dy, pi = s->pi, ds = s->ds;
  const double mdx = fmax(dx, dy), rad = s->p.t_beam + s->p.N_smooth * mdx;
  double Ft[2] = {0, 0}, Fn[2] = {0, 0}, Fp[2] = {0, 0};
  double alpha, at, att, LE[6];
  int err = 0;
  if (s->p.iCylinder != 0) return;
  o_mouvement(s, time, b, &alpha, &at, &att, LE);
  double dphi = pi / (double)(nc - 1);
  double beta = b->theta[ns - 1] + alpha;
  double dL = dphi * rad;
  for (int i = 0; i < nc; i++) {
    double phi = -(0.5 * pi - beta) + (double)i * dphi;
    double xc = b->x[ns - 1] + rad * cos(phi), yc = b->y[ns - 1] + rad * sin(phi);
    double A = o_delta_interp(s, xc, yc, sa, &err), B = o_delta_interp(s, xc, yc, sb, &err);
    double P = o_delta_interp(s, xc, yc, press, &err);
    double sn = +A * cos(2.0 * phi) + B * sin(2.0 * phi);
    double st = -(A * sin(2.0 * phi)) + B * cos(2.0 * phi);
    double sp = -P;
    Ft[0] = Ft[0] + dL * st * (-sin(phi)); Ft[1] = Ft[1] + dL * st * (+cos(phi));
    Fn[0] = Fn[0] + dL * sn * (+cos(phi)); Fn[1] = Fn[1] + dL * sn * (+sin(phi));
    Fp[0] = Fp[0] + dL * sp * (+cos(phi)); Fp[1] = Fp[1] + dL * sp * (+sin(phi));
  }
  beta = b->theta[1] + alpha; /* BeamForces.f90:390 uses node 1 */
  for (int i = 0; i < nc; i++) {
    double phi = (0.5 * pi + beta) + (double)i * dphi;
    double xc = b->x[0] + rad * cos(phi), yc = b->y[0] + rad * sin(phi);
    double A = o_delta_interp(s, xc, yc, sa, &err), B = o_delta_interp(s, xc, yc, sb, &err);
    double P = o_delta_interp(s, xc, yc, press, &err);
    double sn = +A * cos(2.0 * phi) + B * sin(2.0 * phi);
    double st = -(A * sin(2.0 * phi)) + B * cos(2.0 * phi);
    double sp = -P;
    Ft[0] = Ft[0] + dL * st * (-sin(phi)); Ft[1] = Ft[1] + dL * st * (+cos(phi));
    Fn[0] = Fn[0] + dL * sn * (+cos(phi)); Fn[1] = Fn[1] + dL * sn * (+sin(phi));
    Fp[0] = Fp[0] + dL * sp * (+cos(phi)); Fp[1] = Fp[1] + dL * sp * (+sin(phi));
  }
  for (int i = 0; i < ns; i++) {
    double th = b->theta[i] + alpha;
    double xu = b->x[i] - rad * sin(th), yu = b->y[i] + rad * cos(th);
    double xu1 = b->x[i] - s->p.t_beam * sin(th), yu1 = b->y[i] + s->p.t_beam * cos(th);
    double A = o_delta_interp(s, xu, yu, sa, &err), B = o_delta_interp(s, xu, yu, sb, &err);
    double P = o_delta_interp(s, xu1, yu1, press, &err);
    double sn = -(A * cos(2.0 * th)) - B * sin(2.0 * th);
    double st = -(A * sin(2.0 * th)) + B * cos(2.0 * th);
    double sp = -P;
    Ft[0] = Ft[0] + ds * st * (+cos(th)); Ft[1] = Ft[1] + ds * st * (+sin(th));
    Fn[0] = Fn[0] + ds * sn * (-sin(th)); Fn[1] = Fn[1] + ds * sn * (+cos(th));
    Fp[0] = Fp[0] + ds * sp * (-sin(th)); Fp[1] = Fp[1] + ds * sp * (+cos(th));
  }
  for (int i = 0; i < ns; i++) {
    double th = b->theta[i] + alpha;
    double xb = b->x[i] + rad * sin(th), yb = b->y[i] - rad * cos(th);
    double xb1 = b->x[i] + s->p.t_beam * sin(th), yb1 = b->y[i] - s->p.t_beam * cos(th);
    double A = o_delta_interp(s, xb, yb, sa, &err), B = o_delta_interp(s, xb, yb, sb, &err);
    double P = o_delta_interp(s, xb1, yb1, press, &err);
    double sn = -(A * cos(2.0 * th)) - B * sin(2.0 * th);
    double st = -(A * sin(2.0 * th)) + B * cos(2.0 * th);
    double sp = -P;
    Ft[0] = Ft[0] + ds * st * (-cos(th)); Ft[1] = Ft[1] + ds * st * (-sin(th));
    Fn[0] = Fn[0] + ds * sn * (+sin(th)); Fn[1] = Fn[1] + ds * sn * (-cos(th));
    Fp[0] = Fp[0] + ds * sp * (+sin(th)); Fp[1] = Fp[1] + ds * sp * (-cos(th));
  }
  size_t N = (size_t)s->nx * s->ny;
  double sd = 0.0, sl = 0.0;
  for (size_t k = 0; k < N; k++) sd += s->mask[k] * (u[k] - s->maskvx[k]);
  for (size_t k = 0; k < N; k++) sl += s->mask[k] * (u[N + k] - s->maskvy[k]);
  bsi[0] = sd * dx * dy;
  bsi[1] = sl * dx * dy;
  bsi[2] = Ft[0] + Fn[0] + Fp[0];
  bsi[3] = Ft[1] + Fn[1] + Fp[1];
  bsi[4] = Ft[0]; bsi[5] = Fn[0]; bsi[6] = Fp[0];
  bsi[7] = Ft[1]; bsi[8] = Fn[1]; bsi[9] = Fp[1];
  if (err) fprintf(stderr, "oracle: DeltaInterpolation target outside field\n");
}

/* BeamSurfaceIntegral of the held beam on the given stress and pressure
 * fields (instead of the stresses GetForces derives from u); the 10 results
 * land in the diagnostics (o_get_diag bsi) */
void o_surface_integral(o_sim *s, double time, const double *u, const double *sa, const double *sb,
                        const double *press) {
  beam_surface_integral(s, time, &s->beams[0], u, sa, sb, press, s->bsi);
}

/* GetForces -> GetForcesInterp with SPL=3: BeamForces.f90:6-51, 63-288 */
static void get_forces_one(o_sim *s, double time, o_beam *b, const double *press,
                           const double *u, int ib);

/* GetForces: BeamForces.f90:6-51, do i = 1, iBeam (b points to
 * beams(1:iBeam)); the stresses (:105-124) are the same for every beam and
 * computed once */
void o_get_forces(o_sim *s, double time, o_beam *b, const double *press,
                  const double *u) {
  const size_t N = (size_t)s->nx * s->ny;
  const double nu = s->p.nu;
  if (s->p.iViscous == 1) { /* :105-124 */
    double *w1 = dalloc(N), *w2 = dalloc(N), *w3 = dalloc(N);
    o_coftxy(s, u, w1); o_cofdx(s, w1, w2); o_cofitxy(s, w2, s->stress_a);
    for (size_t k = 0; k < N; k++) s->stress_a[k] = 2.0 * nu * s->stress_a[k];
    o_coftxy(s, u, w1); o_cofdy(s, w1, w2); o_cofitxy(s, w2, w3);
    o_coftxy(s, u + N, w1); o_cofdx(s, w1, w2); o_cofitxy(s, w2, s->stress_b);
    for (size_t k = 0; k < N; k++) s->stress_b[k] = nu * (w3[k] + s->stress_b[k]);
    free(w1); free(w2); free(w3);
  }
  for (int ib = 0; ib < s->nb; ib++) get_forces_one(s, time, &b[ib], press, u, ib);
  memcpy(s->bsi, s->bsi_b[0], sizeof(s->bsi));
  s->fx_int = s->fx_int_b[0];
  s->fy_int = s->fy_int_b[0];
}

/* GetForcesInterp with SPL=3 for beam ib: BeamForces.f90:63-288 */
static void get_forces_one(o_sim *s, double time, o_beam *b, const double *press,
                           const double *u, int ib) {
  const int ns = s->ns;
  const double dx = s->dx, dy = s->dy, ds = s->ds;
  int visc = (s->p.iViscous == 1);
  double alpha, at, att, LE[6];
  int err = 0;
  o_mouvement(s, time, b, &alpha, &at, &att, LE);
  double *tau_beam = b->tau_new, *pb = b->pressure_new;
  double *sigma_beam = dalloc(ns), *theta = dalloc(ns);
  for (int n = 0; n < ns; n++) { tau_beam[n] = 0.0; pb[n] = 0.0; }
  if (visc) beam_surface_integral(s, time, b, u, s->stress_a, s->stress_b, press, s->bsi_b[ib]); /* :129 */
  for (int n = 0; n < ns; n++) theta[n] = b->theta[n] + alpha; /* :150 */
  const double mdx = fmax(dy, dx);
  for (int n = 0; n < ns; n++) { /* :153-216, SPL=3 */
    double rr = s->p.t_beam + s->p.N_smooth * mdx;
    double r1 = s->p.t_beam - 0.0 * s->p.N_smooth * mdx;
    double st = sin(theta[n]), ct = cos(theta[n]);
    double xu = b->x[n] - rr * st, yu = b->y[n] + rr * ct;
    double xb = b->x[n] + rr * st, yb = b->y[n] - rr * ct;
    double xu1 = b->x[n] - r1 * st, yu1 = b->y[n] + r1 * ct;
    double xb1 = b->x[n] + r1 * st, yb1 = b->y[n] - r1 * ct;
    pb[n] = o_delta_interp(s, xu1, yu1, press, &err) - o_delta_interp(s, xb1, yb1, press, &err);
    if (visc) {
      double A1 = o_delta_interp(s, xu, yu, s->stress_a, &err);
      double B1 = o_delta_interp(s, xu, yu, s->stress_b, &err);
      double A2 = o_delta_interp(s, xb, yb, s->stress_a, &err);
      double B2 = o_delta_interp(s, xb, yb, s->stress_b, &err);
      double c2 = cos(2.0 * theta[n]), s2 = sin(2.0 * theta[n]);
      double sn1 = -(A1 * c2) - B1 * s2, st1 = -(A1 * s2) + B1 * c2;
      double sn2 = -(A2 * c2) - B2 * s2, st2 = -(A2 * s2) + B2 * c2;
      tau_beam[n] = st1 - st2;
      sigma_beam[n] = -sn1 + sn2;
    }
  }
  if (visc) {
    for (int n = 0; n < ns; n++) pb[n] = pb[n] + sigma_beam[n];
  } else {
    for (int n = 0; n < ns; n++) tau_beam[n] = 0.0;
  }
  double soft; /* :246-252 */
  const double Tr = s->p.T_release, tau = o_tau(s, b->iMouvement);
  if (time <= Tr) soft = 0.0;
  else if (time > Tr && time < (Tr + tau)) {
    double a = time - Tr;
    soft = (a * a * a) / (-0.5 * (tau * tau * tau)) + 3. * (a * a) / (tau * tau);
  } else soft = 1.0;
  for (int n = 0; n < ns; n++) { pb[n] = pb[n] * soft; tau_beam[n] = tau_beam[n] * soft; }
  double fx = 0.0, fy = 0.0; /* :261-267 */
  for (int i = 0; i < ns; i++) {
    fx = fx - (sin(theta[i]) * pb[i] + cos(theta[i]) * tau_beam[i]) * ds;
    fy = fy + (cos(theta[i]) * pb[i] + sin(theta[i]) * tau_beam[i]) * ds;
  }
  b->Force_press[0] = fx; b->Force_press[1] = fy;
  s->fx_int_b[ib] = fx; s->fy_int_b[ib] = fy;
  if (err) fprintf(stderr, "oracle: DeltaInterpolation target outside field\n");
  free(sigma_beam); free(theta);
}

/* ========================================================================= */
/* time loop: time_step.f90:84-162 (startup) and :178-341 (AB2 loop), for   */
/* prescribed kinematics (iFLUSI=0, iMotion>0) or a fixed beam.             */
/* ========================================================================= */
void o_set_vortk(o_sim *s, const double *vortk_P) {
  size_t N = (size_t)s->nx * s->ny;
  memcpy(s->vortk[s->n0], vortk_P, N * sizeof(double));
}

static void require_no_fsi(o_sim *s) {
  if (s->p.iFLUSI == 1) {
    fprintf(stderr, "oracle: the solid solver (iFLUSI=1) is not restated\n");
    abort();
  }
}

void o_startup_step(o_sim *s) {
  o_beam *b = s->beams;
  require_no_fsi(s);
  if (!s->started) { /* time_step.f90:84,90 */
    o_create_mask(s, s->time, b);
    o_create_sponge_mask(s);
    s->started = 1;
  }
  if ((s->p.iSponge >= 4 && s->time < s->p.T_fullspeed) || s->p.iFLUSI == 1 || s->p.iMotion > 0)
    o_create_mask(s, s->time, b);
  o_evolve_explicit(s, s->time, &s->dt1, s->n0, s->n1, b);
  if ((s->p.iFLUSI == 1 || s->p.iSaveBeam > 0) && s->p.iBeam > 0)
    o_get_forces(s, s->time, b, s->press, s->u);
  if (s->p.iFLUSI == 0 && s->p.iMotion > 0) /* time_step.f90:215-218 */
    for (int ib = 0; ib < s->nb; ib++) o_integrate_position(s, s->time, &b[ib]);
  int t = s->n1; s->n1 = s->n0; s->n0 = t;
  s->time = s->time + s->dt1;
}

void o_ab2_step(o_sim *s) {
  o_beam *b = s->beams;
  require_no_fsi(s);
  s->dt0 = s->dt1;
  if ((s->p.iSponge >= 4 && s->time < s->p.T_fullspeed) || s->p.iFLUSI == 1 || s->p.iMotion > 0)
    o_create_mask(s, s->time, b);
  o_evolve_ab2(s, s->time, s->dt0, &s->dt1, s->n0, s->n1, b);
  if ((s->p.iFLUSI == 1 || s->p.iSaveBeam > 0) && s->p.iBeam > 0)
    o_get_forces(s, s->time, b, s->press, s->u);
  if (s->p.iFLUSI == 0 && s->p.iMotion > 0) /* time_step.f90:215-218 */
    for (int ib = 0; ib < s->nb; ib++) o_integrate_position(s, s->time, &b[ib]);
  int t = s->n1; s->n1 = s->n0; s->n0 = t;
  s->time = s->time + s->dt1;
  s->it++;
}

/* The per-step calls of time_step.f90:187-199 for a host that moves the beam
 * itself (iFLUSI=1: its solid solver writes the beam arrays between steps). */
void o_host_create_mask(o_sim *s, double time) {
  o_create_mask(s, time, s->beams);
  if (!s->started) { /* after the startup mask (time_step.f90:84,90): sponge 22 reads it */
    o_create_sponge_mask(s);
    s->started = 1;
  }
}
double o_host_evolve(o_sim *s, double time, double dt0, int explicit_step) {
  double dt1 = 0.0;
  if (explicit_step)
    o_evolve_explicit(s, time, &dt1, s->n0, s->n1, s->beams);
  else
    o_evolve_ab2(s, time, dt0, &dt1, s->n0, s->n1, s->beams);
  return dt1;
}
void o_host_get_forces(o_sim *s, double time) { o_get_forces(s, time, s->beams, s->press, s->u); }
void o_host_swap(o_sim *s) {
  int t = s->n1; s->n1 = s->n0; s->n0 = t;
}

void o_run(o_sim *s, int nsteps) {
  for (int k = 0; k < nsteps; k++) {
    if (s->time >= s->p.Time_end) break;
    if (!s->startup_done) {
      o_startup_step(s);
      s->startup_done = 1;
    } else {
      o_ab2_step(s);
    }
  }
}

/* ---- state access -------------------------------------------------------- */
static double *field_ptr(o_sim *s, int which) {
  switch (which) {
  case O_F_VORTK: return s->vortk[s->n0];
  case O_F_NLK: return s->nlk[s->n0];
  case O_F_WORKVIS: return s->workvis[s->n0];
  case O_F_U1: return s->u;
  case O_F_U2: return s->u + (size_t)s->nx * s->ny;
  case O_F_PRESS: return s->press;
  case O_F_MASK: return s->mask;
  case O_F_MASKVX: return s->maskvx;
  case O_F_MASKVY: return s->maskvy;
  case O_F_MASK_SPONGE: return s->mask_sponge;
  case O_F_VOR_MEAN: return s->vor_mean;
  case O_F_DEALIASE: return s->dealiase;
  case O_F_STRESS_A: return s->stress_a;
  case O_F_STRESS_B: return s->stress_b;
  case O_F_VORT: return s->vort;
  }
  return NULL;
}

void o_get_field(o_sim *s, int which, double *out) {
  double *p = field_ptr(s, which);
  if (p) memcpy(out, p, sizeof(double) * (size_t)s->nx * s->ny);
}

void o_set_field(o_sim *s, int which, const double *in) {
  double *p = field_ptr(s, which);
  if (p) memcpy(p, in, sizeof(double) * (size_t)s->nx * s->ny);
}

o_beam *o_get_beam(o_sim *s) { return &s->beams[0]; }
o_beam *o_get_beam_k(o_sim *s, int ib) { return (ib >= 0 && ib < s->nb) ? &s->beams[ib] : NULL; }
int o_nbeams(const o_sim *s) { return s->nb; }
/* BeamSurfaceIntegral and GetForcesInterp's force sums of beam ib from the
 * last o_get_forces: bsi[0..9], then fx_int, fy_int */
void o_get_beam_bsi(o_sim *s, int ib, double *out) {
  for (int i = 0; i < 10; i++) out[i] = s->bsi_b[ib][i];
  out[10] = s->fx_int_b[ib];
  out[11] = s->fy_int_b[ib];
}

void o_get_diag(o_sim *s, o_diag *d) {
  memset(d, 0, sizeof(*d));
  d->time = s->time; d->dt0 = s->dt0; d->dt1 = s->dt1; d->u_max = s->u_max;
  memcpy(d->energies, s->energies, sizeof(d->energies));
  memcpy(d->bsi, s->bsi, sizeof(d->bsi));
  d->fx_int = s->fx_int; d->fy_int = s->fy_int;
  d->n0 = s->n0; d->it = s->it; d->started = s->started;
  d->u_mean[0] = s->u_mean[0]; d->u_mean[1] = s->u_mean[1];
}

/* ---- synthetic inputs ---------------------------------------------------- */
uint64_t o_splitmix64(uint64_t *st) {
  uint64_t z = (*st += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

void o_normals(uint64_t seed, int n, double *out) {
  uint64_t st = seed;
  const double two53 = 1.0 / 9007199254740992.0;
  const double twopi = 6.283185307179586476925286766559005768;
  for (int i = 0; i < n; i += 2) {
    double u1 = (double)((o_splitmix64(&st) >> 11) + 1) * two53; /* (0,1] */
    double u2 = (double)(o_splitmix64(&st) >> 11) * two53;       /* [0,1) */
    double r = sqrt(-2.0 * log(u1));
    out[i] = r * cos(twopi * u2);
    if (i + 1 < n) out[i + 1] = r * sin(twopi * u2);
  }
}

void o_init_noise(o_sim *s, uint64_t seed, double amp, int kmax2,
                  double *vk) {
  const int nx = s->nx, ny = s->ny;
  size_t N = (size_t)nx * ny;
  double *f = dalloc(N);
  o_normals(seed, (int)N, f);
  for (size_t k = 0; k < N; k++) f[k] = amp * f[k];
  o_coftxy(s, f, vk);
  for (int ky = 0; ky <= ny / 2 - 1; ky++)
    for (int kx = 0; kx <= nx / 2 - 1; kx++)
      if (kx * kx + ky * ky > kmax2) {
        vk[(size_t)(2 * ky) * nx + 2 * kx] = 0.0;
        vk[(size_t)(2 * ky) * nx + 2 * kx + 1] = 0.0;
        vk[(size_t)(2 * ky + 1) * nx + 2 * kx] = 0.0;
        vk[(size_t)(2 * ky + 1) * nx + 2 * kx + 1] = 0.0;
      }
  for (size_t k = 0; k < N; k++) vk[k] = vk[k] * s->dealiase[k];
  free(f);
}
