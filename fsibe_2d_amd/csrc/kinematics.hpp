// kinematics.hpp -- host kinematics of the beam's leading edge and the
// beam's node positions: mouvement (mouvement.f90:5-241, plus the heave case
// 600 and the two-harmonic heave + pitch case 601) and integrate_position
// (integrate_position.f90:1-47).  Host code only; used by the fluid loop for
// prescribed motion and by the solid solver.
#pragma once
#include <cmath>

#include "../../include/fsibe_hip.h"

// periods of the heaving cases 401..413 (mouvement.f90:81-94)
static const double kTs400[13] = {0.51, 0.34, 0.25, 0.20, 0.17, 0.14, 0.13, 0.11, 0.10, 0.09, 0.08, 0.08, 0.07};

// soft-startup time: cases 401..413 overwrite the global tau with their
// period on every mouvement call (mouvement.f90:102-104)
static double tau_eff(const fsibe_params &p) {
  return (p.iMotion >= 401 && p.iMotion <= 413) ? kTs400[p.iMotion - 401] : p.tau;
}

static void mouvement_host(const fsibe_params &p, int im, double x0, double y0, double angle, double phase,
                           double time, double &alpha, double &alpha_t, double &alpha_tt, double LE[6]) {
  const double pi = 4.0 * atan(1.0);
  for (int i = 0; i < 6; ++i) LE[i] = 0.0;
  alpha = alpha_t = alpha_tt = 0.0;
  if (im == 0) {
    LE[0] = x0; LE[1] = y0;
    alpha = angle * pi / 180.0;
  } else if ((im >= 201 && im <= 213) || (im >= 300 && im <= 350)) {
    static const double ts[13] = {1.0, 1.09, 1.20, 1.29, 1.33, 1.4, 1.5, 1.71, 2.0, 2.4, 3.0, 4.0, 1.62};
    const double y_max = 0.5, t_max = (im < 300) ? ts[im - 201] : 1.62;
    double k, kt, ktt;
    if (time <= 1.0) {
      const double a = -20.0, b = 70.0, c = -84.0, d = 35.0;
      const double t2 = time * time, t3 = t2 * time, t4 = t3 * time, t5 = t4 * time, t6 = t5 * time,
                   t7 = t6 * time;
      k = a * t7 + b * t6 + c * t5 + d * t4;
      kt = 7.0 * a * t6 + 6.0 * b * t5 + 5.0 * c * t4 + 4.0 * d * t3;
      ktt = 42. * a * t5 + 30. * b * t4 + 20. * c * t3 + 12. * d * t2;
    } else {
      k = 1.0; kt = 0.0; ktt = 0.0;
    }
    const double arg = 2.0 * pi * time / t_max + pi / 2. + phase;
    const double w = 2.0 * pi / t_max;
    const double y = y_max * sin(arg), yt = y_max * (2.0 * pi / t_max) * cos(arg);
    const double ytt = -(y_max * (w * w) * sin(arg));
    LE[0] = x0; LE[1] = y0 + k * y; LE[3] = kt * y + yt * k; LE[5] = ktt * y + ytt * k + 2. * kt * yt;
    alpha = angle * pi / 180.0;
  } else if (im >= 401 && im <= 413) {
    const double y_max = 0.0333, t_max = kTs400[im - 401], w = 2.0 * pi / t_max;
    LE[0] = x0;
    LE[1] = y0 + y_max * sin(2.0 * pi * time / t_max + pi / 2);
    LE[3] = y_max * (2.0 * pi / t_max) * cos(2.0 * pi * time / t_max + pi / 2);
    LE[5] = -(y_max * (w * w) * sin(2.0 * pi * time / t_max + pi / 2));
    alpha = angle * pi / 180.0;
  } else if (im == 2) {
    const double y_max = 0.5, t_max = 2.5, w = 2.0 * pi / t_max;
    LE[0] = x0;
    LE[1] = y0 + y_max * sin(2.0 * pi * time / t_max + pi / 2);
    LE[3] = y_max * (2.0 * pi / t_max) * cos(2.0 * pi * time / t_max + pi / 2);
    LE[5] = -(y_max * (w * w) * sin(2.0 * pi * time / t_max + pi / 2));
    alpha = angle * pi / 180.0;
  } else if (im == 3) {
    const double angle_max = 80.0 * pi / 180.0, f = 0.5, w = 2.0 * pi * f;
    LE[0] = x0; LE[1] = y0;
    alpha = angle_max * sin(2.0 * pi * f * time);
    alpha_t = angle_max * cos(2.0 * pi * f * time) * (2.0 * pi * f);
    alpha_tt = -1.0 * angle_max * sin(2.0 * pi * f * time) * (w * w);
  } else if (im == 4) {
    const double u0 = -1.0;
    LE[0] = x0 + u0 * time; LE[2] = u0; LE[1] = y0; LE[3] = 0.0;
    alpha = angle * pi / 180.0;
  } else if (im == 600) {
    const double A = p.heave_amp, w = 2.0 * pi * p.heave_freq;
    LE[0] = x0;
    LE[1] = y0 + A * sin(w * time);
    LE[3] = A * w * cos(w * time);
    LE[5] = -(A * (w * w) * sin(w * time));
    alpha = angle * pi / 180.0;
  } else if (im == 601) {
    const double *c = p.motion_coef;
    const double w0 = 2.0 * pi * p.heave_freq;
    double h = 0, ht = 0, htt = 0, a = 0, at = 0, att = 0;
    for (int m = 1; m <= 2; ++m) {
      const double wm = w0 * m, sn = sin(wm * time), cs = cos(wm * time);
      const double ah = c[2 * (m - 1)], bh = c[2 * (m - 1) + 1];
      const double ap = c[4 + 2 * (m - 1)], bp = c[4 + 2 * (m - 1) + 1];
      h += ah * sn + bh * (cs - 1.0);
      ht += wm * (ah * cs - bh * sn);
      htt += -(wm * wm) * (ah * sn + bh * cs);
      a += ap * sn + bp * (cs - 1.0);
      at += wm * (ap * cs - bp * sn);
      att += -(wm * wm) * (ap * sn + bp * cs);
    }
    LE[0] = x0; LE[1] = y0 + h; LE[3] = ht; LE[5] = htt;
    alpha = angle * pi / 180.0 + a; alpha_t = at; alpha_tt = att;
  }
}

// integrate_position.f90:1-47.  The reference recomputes each node's prefix
// sums from i = 0; carrying them from node s to s+1 performs the same
// additions in the same order, so the result is identical at O(ns) cost.
// The beam's placement (type Solid x0, y0, phase) is an argument: the
// second beam of a tandem run sits at x0 + 1.25 with its own phase.
static void integrate_position_host(const fsibe_params &p, int im, double x0, double y0, double phase, double time,
                                    const double *theta, const double *theta_dot, double *x, double *y, double *vx,
                                    double *vy) {
  double alpha, at, att, LE[6];
  mouvement_host(p, im, x0, y0, p.AngleBeam, phase, time, alpha, at, att, LE);
  const int ns = p.ns;
  const double ds = p.length / (double)(ns - 1);
  const double ca = cos(alpha), sa = sin(alpha);
  double xrel = 0.0, yrel = 0.0, vxrel = 0.0, vyrel = 0.0;
  for (int s = 0; s < ns; ++s) {
    if (s > 0) {
      const int i = s - 1;
      const double c = cos(theta[i]), sn = sin(theta[i]);
      xrel = xrel + c * ds;
      yrel = yrel + sn * ds;
      vxrel = vxrel - theta_dot[i] * sn * ds;
      vyrel = vyrel + theta_dot[i] * c * ds;
    }
    x[s] = LE[0] + xrel * ca - yrel * sa;
    y[s] = LE[1] + yrel * ca + xrel * sa;
    vx[s] = LE[2] + ca * (vxrel - at * yrel) - sa * (vyrel + at * xrel);
    vy[s] = LE[3] + ca * (vyrel + at * xrel) + sa * (vxrel - at * yrel);
  }
}

// the first beam at the parameters' leading edge
static void integrate_position_host(const fsibe_params &p, int im, double time, const double *theta,
                                    const double *theta_dot, double *x, double *y, double *vx, double *vy) {
  integrate_position_host(p, im, p.x0, p.y0, 0.0, time, theta, theta_dot, x, y, vx, vy);
}

// placement of beam ib (init_beam.f90:44-76): the first at (x0, y0) with
// phase 0; the second, for the dragonfly motions 300..307 only, at
// (x0 + 1.25, y0) with phase (iMotion - 300) pi / 4.  false: no such beam.
static bool beam_placement(const fsibe_params &p, int ib, double &x0, double &y0, double &phase) {
  const double pi = 4.0 * atan(1.0);
  if (ib == 0) {
    x0 = p.x0; y0 = p.y0; phase = 0.0;
    return true;
  }
  if (ib != 1 || p.iMotion < 300 || p.iMotion > 307) return false;
  x0 = p.x0 + 1.25; y0 = p.y0;
  phase = 0.25 * (double)(p.iMotion - 300) * pi;
  return true;
}
