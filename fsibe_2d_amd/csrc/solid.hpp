// solid.hpp -- the host solid solver of the beam (SolidSolver.f90).
//
// One inextensible beam in the frame of its moving leading edge, unknowns the
// angle theta(s) and the tension T(s).  IBES_solver (SolidSolver.f90:99-414)
// takes one implicit step (implicit Euler, Crank-Nicolson or BDF2 with a
// Crank-Nicolson first step) by Newton iterations on the 2 ns + 4 unknowns
//     x = [theta(-1 .. ns+1), T(-1 .. ns-1)]
// (three ghost angles and one ghost tension carry the boundary conditions),
// the residual of F_nonlinear (:552-667) and the analytic Jacobian of Jacobi
// (:671-863).  The predictor is one explicit Euler step, EE1 (:1309-1331),
// through RHS_beameqn (:918-994) and Tension (:1000-1081).
//
// The reference factors the dense (2 ns + 4)^2 Jacobian (Solve_LGS "DIRECT",
// :502-520).  The system is banded once the unknowns are interleaved node by
// node (theta(j), T(j)) and each equation is placed at the node it belongs
// to: 7 sub- and 8 super-diagonals.  This solver factors that band with row
// interchanges (partial pivoting), O(ns) instead of O(ns^3), once per time
// step as a rule (see the Newton loop); the Newton iterates differ from the
// dense solves' by rounding only.
//
// Expressions keep the reference's operand order (Fortran evaluates a+b+c as
// (a+b)+c and x**2 as x*x); the library is compiled without floating-point
// contraction.
#pragma once
#include <algorithm>
#include <cmath>
#include <string>
#include <utility>
#include <vector>

#include "kinematics.hpp"

namespace ibes {

// time schemes (share_vars.f90:43)
constexpr int kEulerImplicit = 1, kCrankNicholson = 2, kRungeKutta4 = 3, kEulerExplicit = 4, kBDF2 = 5;
// band of the interleaved Newton matrix
constexpr int kKL = 7, kKU = 8, kBW = 2 * kKL + kKU + 1;
constexpr int kMaxIter = 49;          // IBES "emergency brake" (SolidSolver.f90:311-314)
constexpr double kErrorStop = 1.0e-7;  // SolidSolver.f90:126

// the solver's part of type(solid) (share_vars.f90:50-63)
struct Beam {
  int ns = 0;
  std::vector<double> x, y, vx, vy, theta, theta_dot, ax, ay;
  std::vector<double> pressure_old, pressure_new, tau_old, tau_new;
  std::vector<double> oldold;  // beam_oldold(0:ns-1, 1:6): column c at [c * ns]
  double dt_old = 0.0;
  bool StartupStep = true;
  double E_kinetic = 0.0, E_pot = 0.0, E_elastic = 0.0, Inertial_Force[2] = {0.0, 0.0};
  void resize(int n) {
    ns = n;
    for (auto *v : {&x, &y, &vx, &vy, &theta, &theta_dot, &ax, &ay, &pressure_old, &pressure_new, &tau_old, &tau_new})
      v->assign(n, 0.0);
    oldold.assign((size_t)6 * n, 0.0);
  }
};

// ---- Differentiate1D / create_diff_matrices (SolidSolver.f90:1085-1233) ----
// The difference matrices are banded; each row is applied as its stencil,
// terms added in column order as the dense matrix-vector product adds them.
// The coefficients are the matrices' entries, k / (2 dx) etc.
struct Diff {
  double dx = 0.0;
  // interior and boundary coefficients, already divided by the grid factor
  double c1i[2], c1b[3];
  double c2i[3], c2b[4];
  double c3i[4], c3b[6];
  double c4i[5], c4b[7];
  void init(double h) {
    dx = h;
    const double s1 = 2.0 * h, s2 = h * h, s3 = 2.0 * (h * h * h), s4 = (h * h) * (h * h);
    c1i[0] = -1. / s1; c1i[1] = 1. / s1;
    c1b[0] = -3. / s1; c1b[1] = 4. / s1; c1b[2] = -1. / s1;
    c2i[0] = 1. / s2; c2i[1] = -2. / s2; c2i[2] = 1. / s2;
    c2b[0] = 2. / s2; c2b[1] = -5. / s2; c2b[2] = 4. / s2; c2b[3] = -1. / s2;
    c3i[0] = -1. / s3; c3i[1] = 2. / s3; c3i[2] = -2. / s3; c3i[3] = 1. / s3;  // at i-2, i-1, i+1, i+2
    const double b3[6] = {-17.0 / 2., 71.0 / 2., -59., 49., -41.0 / 2., 7.0 / 2.};
    for (int k = 0; k < 6; ++k) c3b[k] = b3[k] / s3;
    c4i[0] = 1. / s4; c4i[1] = -4. / s4; c4i[2] = 6. / s4; c4i[3] = -4. / s4; c4i[4] = 1. / s4;
    const double b4[7] = {35.0 / 6., -31., 137.0 / 2., -242.0 / 3., 107.0 / 2., -19.0, 17.0 / 6.};
    for (int k = 0; k < 7; ++k) c4b[k] = b4[k] / s4;
  }
  // first derivative: rows 0 and N-1 one-sided of second order
  void d1(const double *f, double *o, int n) const {
    o[0] = c1b[0] * f[0] + c1b[1] * f[1] + c1b[2] * f[2];
    for (int i = 1; i < n - 1; ++i) o[i] = c1i[0] * f[i - 1] + c1i[1] * f[i + 1];
    o[n - 1] = (-c1b[2]) * f[n - 3] + (-c1b[1]) * f[n - 2] + (-c1b[0]) * f[n - 1];
  }
  void d2(const double *f, double *o, int n) const {
    o[0] = c2b[0] * f[0] + c2b[1] * f[1] + c2b[2] * f[2] + c2b[3] * f[3];
    for (int i = 1; i < n - 1; ++i) o[i] = c2i[0] * f[i - 1] + c2i[1] * f[i] + c2i[2] * f[i + 1];
    o[n - 1] = c2b[3] * f[n - 4] + c2b[2] * f[n - 3] + c2b[1] * f[n - 2] + c2b[0] * f[n - 1];
  }
  // third derivative: rows 0, 1 forward and N-2, N-1 backward one-sided
  void d3(const double *f, double *o, int n) const {
    for (int r = 0; r < 2; ++r) {
      double a = 0.0;
      for (int k = 0; k < 6; ++k) a = a + c3b[k] * f[r + k];
      o[r] = a;
    }
    for (int i = 2; i < n - 2; ++i)
      o[i] = c3i[0] * f[i - 2] + c3i[1] * f[i - 1] + c3i[2] * f[i + 1] + c3i[3] * f[i + 2];
    for (int r = n - 2; r < n; ++r) {  // D3(N,N-5..N) = -(row 0 reversed)
      const int e = (r == n - 1) ? n - 1 : n - 2;
      double a = 0.0;
      for (int k = 5; k >= 0; --k) a = a + (-c3b[k]) * f[e - k];
      o[r] = a;
    }
  }
  void d4(const double *f, double *o, int n) const {
    for (int r = 0; r < 2; ++r) {
      double a = 0.0;
      for (int k = 0; k < 7; ++k) a = a + c4b[k] * f[r + k];
      o[r] = a;
    }
    for (int i = 2; i < n - 2; ++i)
      o[i] = c4i[0] * f[i - 2] + c4i[1] * f[i - 1] + c4i[2] * f[i] + c4i[3] * f[i + 1] + c4i[4] * f[i + 2];
    for (int r = n - 2; r < n; ++r) {  // D4(N,N-6..N) = row 0 reversed
      const int e = (r == n - 1) ? n - 1 : n - 2;
      double a = 0.0;
      for (int k = 6; k >= 0; --k) a = a + c4b[k] * f[e - k];
      o[r] = a;
    }
  }
};

// ---- tridiagonal LU with row interchanges (the LAPACK dgttrf / dgttrs
// algorithm Tension calls, SolidSolver.f90:1064-1065) ----------------------
inline bool tridiag_solve(int n, double *dl, double *d, double *du, double *b) {
  std::vector<double> du2(n > 2 ? n - 2 : 1, 0.0);
  std::vector<int> ipiv(n);
  for (int i = 0; i < n; ++i) ipiv[i] = i;
  for (int i = 0; i < n - 1; ++i) {
    if (std::fabs(d[i]) >= std::fabs(dl[i])) {
      if (d[i] != 0.0) {
        const double fact = dl[i] / d[i];
        dl[i] = fact;
        d[i + 1] = d[i + 1] - fact * du[i];
      }
    } else {
      const double fact = d[i] / dl[i];
      d[i] = dl[i];
      dl[i] = fact;
      const double temp = du[i];
      du[i] = d[i + 1];
      d[i + 1] = temp - fact * d[i + 1];
      if (i < n - 2) {
        du2[i] = du[i + 1];
        du[i + 1] = -fact * du[i + 1];
      }
      ipiv[i] = i + 1;
    }
  }
  for (int i = 0; i < n; ++i)
    if (d[i] == 0.0) return false;
  for (int i = 0; i < n - 1; ++i) {  // L
    const int ip = ipiv[i];
    const double temp = b[i - ip + i + 1] - dl[i] * b[ip];
    b[i] = b[ip];
    b[i + 1] = temp;
  }
  b[n - 1] = b[n - 1] / d[n - 1];  // U
  if (n > 1) b[n - 2] = (b[n - 2] - du[n - 2] * b[n - 1]) / d[n - 2];
  for (int i = n - 3; i >= 0; --i) b[i] = (b[i] - du[i] * b[i + 1] - du2[i] * b[i + 2]) / d[i];
  return true;
}

// ---- banded Newton matrix with partial pivoting --------------------------
// Row r keeps columns r - kKL .. r + kKL + kKU (room for the fill of the row
// interchanges): row(r)[c] is entry (r, c).  last[r] is the last column of
// row r that may be nonzero; elimination and back substitution stop there.
// factor() keeps the multipliers of step j below the diagonal of column j
// and the interchange of step j in piv[j] (the storage of LAPACK's dgbtrf),
// so that a factorisation solves further right-hand sides.
struct Band {
  int n = 0;
  std::vector<double> a;
  std::vector<int> last, piv;
  std::vector<double> inv;  // 1 / U(j, j) after factor()
  void reset(int nn) {
    if (n != nn || a.size() != (size_t)nn * kBW) {
      n = nn;
      a.assign((size_t)n * kBW, 0.0);
      last.resize(n);
      piv.resize(n);
      inv.resize(n);
    } else {
      std::fill(a.begin(), a.end(), 0.0);
    }
    for (int r = 0; r < n; ++r) last[r] = r;
  }
  // flat index of entry (r, c) in a
  static size_t index(int r, int c) { return (size_t)r * kBW + kKL - r + c; }
  double *row(int r) { return a.data() + (size_t)r * kBW + kKL - r; }
  const double *row(int r) const { return a.data() + (size_t)r * kBW + kKL - r; }
  double &at(int r, int c) {
    if (c > last[r]) last[r] = c;
    return row(r)[c];
  }
  // y = A x (before factor())
  void matvec(const double *x, double *y) const {
    for (int r = 0; r < n; ++r) {
      const double *rr = row(r);
      double s = 0.0;
      for (int c = std::max(0, r - kKL); c <= last[r]; ++c) s = s + rr[c] * x[c];
      y[r] = s;
    }
  }
  // LU with row interchanges in place; false if singular
  bool factor() {
    for (int j = 0; j < n; ++j) {
      const int lst = std::min(n - 1, j + kKL);
      int p = j;
      double best = std::fabs(row(j)[j]);
      for (int i = j + 1; i <= lst; ++i) {
        const double v = std::fabs(row(i)[j]);
        if (v > best) { best = v; p = i; }
      }
      if (best == 0.0) return false;
      piv[j] = p;
      if (p != j) {
        double *rj = row(j), *rp = row(p);
        const int ce = std::max(last[j], last[p]);
        for (int c = j; c <= ce; ++c) std::swap(rj[c], rp[c]);
        std::swap(last[j], last[p]);
      }
      const double *rj = row(j);
      const double pv = rj[j];
      const int ce = last[j];
      for (int i = j + 1; i <= lst; ++i) {
        double *ri = row(i);
        if (ri[j] == 0.0) continue;
        const double l = ri[j] / pv;
        ri[j] = l;
        for (int c = j + 1; c <= ce; ++c) ri[c] = ri[c] - l * rj[c];
        if (ce > last[i]) last[i] = ce;
      }
    }
    for (int j = 0; j < n; ++j) inv[j] = 1.0 / row(j)[j];
    return true;
  }
  // b <- A^-1 b with the factors of factor()
  void solve_factored(double *b) const {
    for (int j = 0; j < n; ++j) {
      if (piv[j] != j) std::swap(b[j], b[piv[j]]);
      const int lst = std::min(n - 1, j + kKL);
      for (int i = j + 1; i <= lst; ++i) {
        const double l = row(i)[j];
        if (l != 0.0) b[i] = b[i] - l * b[j];
      }
    }
    // back substitution; the terms of b[j+2..] go into four partial sums
    // (the chain of one sum would wait on every subtraction), and only the
    // last term waits for b[j+1]
    for (int j = n - 1; j >= 0; --j) {
      const double *rj = row(j);
      double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
      int c = j + 2;
      for (; c + 3 <= last[j]; c += 4) {
        s0 = s0 + rj[c] * b[c];
        s1 = s1 + rj[c + 1] * b[c + 1];
        s2 = s2 + rj[c + 2] * b[c + 2];
        s3 = s3 + rj[c + 3] * b[c + 3];
      }
      for (; c <= last[j]; ++c) s0 = s0 + rj[c] * b[c];
      double s = b[j] - ((s0 + s1) + (s2 + s3));
      if (last[j] >= j + 1) s = s - rj[j + 1] * b[j + 1];
      b[j] = s * inv[j];
    }
  }
};

// The solver for one beam: parameters, scratch, and the step.
struct Solver {
  fsibe_params p;
  // the solved beam's placement (type Solid x0, y0, phase; init_beam.f90:44-76)
  double x0 = 0.0, y0 = 0.0, phase = 0.0;
  int ns = 0, nx = 0;     // nx = 2 ns + 4 unknowns
  double ds = 0.0, grav = 0.0;
  int method = kBDF2;     // TimeMethodSolid (restored after a startup step)
  Diff dn, de;            // stencils on ns points and on the ns + 3 extended points
  Band band, lu;          // the Jacobian being assembled; the factorisation in use
  // the Jacobian's sparsity is the same at every iteration: the band index
  // of each emitted entry, in emission order, and each row's last column
  std::vector<size_t> jofs;
  std::vector<int> jlast;
  int refactored = 0;     // Newton iterations after the first that factored afresh
  std::vector<int> rowpos, colpos;  // reference equation k / unknown l (1-based) -> band row / column
  // per-step context shared by F_nonlinear and Jacobi
  struct Ctx {
    int method = kBDF2;
    double time = 0, dt = 0, dt_old = 0, C1 = 0, C2 = 0, C3 = 0, C4 = 0;
    double alpha = 0, alpha_t = 0, alpha_tt = 0, LE[6] = {0, 0, 0, 0, 0, 0};
    const double *th_old = nullptr, *thd_old = nullptr, *th_oo = nullptr, *thd_oo = nullptr;
    const double *p = nullptr, *tau = nullptr;
    std::vector<double> tau_s, old_rhs, thd_new;
  } cx;
  std::vector<double> th_old_, thd_old_;  // theta, theta_dot at t_n
  // the predictor's RHS_beameqn at t_n (pre_th1_ = theta_dot, pre_T_ =
  // tension), computed by predict() before the new loads are known
  std::vector<double> pre_th1_, pre_T_;
  double pre_time_ = 0.0;
  bool pre_ok_ = false;
  std::string err;
  int iter = 0;            // Newton iterations of the last step (IBES_iter)
  bool brake = false;      // the step tripped the reference's emergency brake

  bool init(const fsibe_params &pp) {
    p = pp;
    ns = p.ns;
    x0 = p.x0; y0 = p.y0; phase = 0.0;
    jofs.clear();
    if (ns < 8) { err = "the solid solver needs ns >= 8 (one-sided fourth-derivative stencils)"; return false; }
    if (!(p.TimeMethodSolid == kEulerImplicit || p.TimeMethodSolid == kCrankNicholson || p.TimeMethodSolid == kBDF2)) {
      err = "TimeMethodSolid must be 1 (implicit Euler), 2 (Crank-Nicolson) or 5 (BDF2)";
      return false;
    }
    if (!(p.mue > 0.0)) { err = "mue must be > 0"; return false; }
    nx = 2 * ns + 4;
    ds = p.length / (double)(ns - 1);
    grav = p.grav;
    method = p.TimeMethodSolid;
    dn.init(ds);
    de.init(ds);
    band.reset(nx);
    lu.reset(nx);
    colpos.assign(nx + 1, 0);
    rowpos.assign(nx + 1, 0);
    // unknowns: theta(j) = x(j+2) at 2(j+1), T(j) = x(ns+5+j) at 2(j+1)+1,
    // theta(ns), theta(ns+1) at the end
    for (int j = -1; j <= ns + 1; ++j) colpos[j + 2] = (j <= ns - 1) ? 2 * (j + 1) : (j == ns ? 2 * ns + 2 : 2 * ns + 3);
    for (int j = -1; j <= ns - 1; ++j) colpos[ns + 5 + j] = 2 * (j + 1) + 1;
    // equations: each at the unknown it depends on most strongly, so that the
    // elimination rarely needs a row interchange: the leading-edge rows 3
    // (theta(-1)), 2 (T(-1)), 1 (theta(0)), 7 (T(0)); the theta and T
    // equations of node i at theta(i), T(i); the trailing-edge rows 8, 6, 4, 5
    rowpos[3] = 0; rowpos[2] = 1; rowpos[1] = 2; rowpos[7] = 3;
    for (int i = 1; i <= ns - 2; ++i) {
      rowpos[8 + i] = 2 * i + 2;
      rowpos[8 + (ns - 2) + i] = 2 * i + 3;
    }
    rowpos[8] = 2 * ns; rowpos[6] = 2 * ns + 1; rowpos[4] = 2 * ns + 2; rowpos[5] = 2 * ns + 3;
    cx.tau_s.assign(ns, 0.0);
    cx.old_rhs.assign(ns, 0.0);
    cx.thd_new.assign(ns, 0.0);
    return true;
  }

  void motion(double t, double &a, double &at, double &att, double LE[6]) const {
    mouvement_host(p, p.iMotion, x0, y0, p.AngleBeam, phase, t, a, at, att, LE);
  }

  // GravityImpulse (SolidSolver.f90:874-911): grav as a function of time
  void gravity_impulse(double time) { grav = gravity_at(p.iImpulse, time, grav); }
  static double gravity_at(int iImpulse, double time, double g) {
    const double pi = 4.0 * atan(1.0);
    if (iImpulse == 1 || iImpulse == 2) {
      const double T0 = 0.75, T1 = 0.85, amp = iImpulse == 1 ? 1.00 : 4.00;
      g = amp * sin(pi * (time - T0) / (T1 - T0));
      if (time < T0) g = 0.0;
      if (time > T1) g = 0.0;
    } else if (iImpulse == 3) {
      const double T0 = 2.0, T1 = 3.0;
      double k;
      if (time <= T0) {
        k = 0.0;
      } else if (time >= T0 && time <= T1) {
        const double a = -20.0, b = 70.0, c = -84.0, d = 35.0;
        const double t = (time - T0) / (T1 - T0);
        const double t2 = t * t, t4 = t2 * t2;
        k = a * (t4 * (t2 * t)) + b * (t4 * t2) + c * (t4 * t) + d * t4;
      } else {
        k = 1.0;
      }
      g = 0.7 * k;
    }
    return g;
  }

  // Tension (SolidSolver.f90:1000-1081): T and T_s of the angle field theta
  bool tension(double time, const double *th, const double *thd, const double *pr, const double *tau, double *T,
               double *T_s) {
    double a, at, att, LE[6];
    motion(time, a, at, att, LE);
    std::vector<double> th_s(ns), th_ss(ns), th_sss(ns), tau_s(ns), dm(ns), du(ns - 1), dl(ns - 1), rhs(ns);
    dn.d1(th, th_s.data(), ns);
    dn.d2(th, th_ss.data(), ns);
    dn.d3(th, th_sss.data(), ns);
    dn.d1(tau, tau_s.data(), ns);
    const double eta = p.eta, mue = p.mue;
    const double K1 = mue * (LE[4] * cos(a) + LE[5] * sin(a) + grav * sin(a)) - tau[0];
    const double T_s0 = -eta * th_ss[0] * th_s[0] + K1;
    const double ds2 = ds * ds;
    for (int i = 0; i < ns; ++i) dm[i] = -2.0 / ds2 - th_s[i] * th_s[i];
    for (int i = 0; i < ns - 1; ++i) { du[i] = 1.0 / ds2; dl[i] = 1.0 / ds2; }
    dm[ns - 1] = 1.0;
    dl[ns - 2] = 0.0;
    du[0] = 2.0 / ds2;
    for (int i = 0; i < ns; ++i) {
      const double w = thd[i] + at;
      rhs[i] = -(pr[i] * th_s[i]) - (2. * eta * th_s[i] * th_sss[i]) - (eta * (th_ss[i] * th_ss[i])) - mue * (w * w) -
               tau_s[i];
    }
    rhs[ns - 1] = 0.0;
    rhs[0] = rhs[0] + T_s0 * 2.0 / ds;
    for (int i = 0; i < ns; ++i) rhs[i] = rhs[i] / dm[i];
    for (int i = 0; i < ns - 1; ++i) {
      du[i] = du[i] / dm[i];
      dl[i] = dl[i] / dm[i + 1];
    }
    for (int i = 0; i < ns; ++i) dm[i] = 1.0;
    if (!tridiag_solve(ns, dl.data(), dm.data(), du.data(), rhs.data())) {
      err = "Tension: singular tridiagonal system";
      return false;
    }
    for (int i = 0; i < ns; ++i) T[i] = rhs[i];
    dn.d1(T, T_s, ns);
    T_s[0] = T_s0;
    return true;
  }

  // RHS_beameqn (SolidSolver.f90:918-994).  In: theta, theta_dot at `time`;
  // out: theta <- theta_dot, theta_dot <- theta_ddot, T the tension.
  bool rhs_beameqn(double time, double *th, double *thd, const double *pr, const double *tau, double *T) {
    double a, at, att, LE[6];
    motion(time, a, at, att, LE);
    th[0] = 0.0;
    thd[0] = 0.0;
    std::vector<double> T_s(ns);
    if (!tension(time, th, thd, pr, tau, T, T_s.data())) return false;
    const double eta = p.eta, mue = p.mue, ds2 = ds * ds;
    const int ne = ns + 3;
    std::vector<double> ext(ne), e1(ne), e2(ne), e3(ne), e4(ne), p_s(ns);
    const double K2 = pr[0] + mue * (LE[5] * cos(a) - LE[4] * sin(a) + grav * cos(a));
    const double C2 = 2.0 * ds * K2 - T[0] * th[1] + (eta / ds2) * (10.0 * th[0] - 12.0 * th[1] + 6.0 * th[2] - th[3]);
    ext[0] = C2 / ((3.0 * eta / ds2) - T[0]);
    for (int i = 0; i < ns; ++i) ext[i + 1] = th[i];
    const double A1 = th[ns - 3] / 12. - 2. * th[ns - 2] / 3.;
    const double A2 = -th[ns - 3] / 12. + 4. * th[ns - 2] / 3. - 5. * th[ns - 1] / 2.;
    ext[ns + 1] = 3. * (A1 - A2) / 2.;
    ext[ns + 2] = 12. * (2. * A1 - A2);
    de.d1(ext.data(), e1.data(), ne);
    de.d2(ext.data(), e2.data(), ne);
    de.d3(ext.data(), e3.data(), ne);
    de.d4(ext.data(), e4.data(), ne);
    dn.d1(pr, p_s.data(), ns);
    for (int i = 0; i < ns; ++i) {
      const double s1 = e1[i + 1], s2 = e2[i + 1], s4 = e4[i + 1];
      const double acc = (-p_s[i] - eta * s4 + s2 * (T[i] + eta * (s1 * s1)) + 2. * T_s[i] * s1 - mue * att -
                          p.sigma * thd[i] + tau[i] * s1) /
                         mue;
      th[i] = thd[i];
      thd[i] = acc;
    }
    thd[0] = 0.0;
    return true;
  }

  // time-scheme constants (SolidSolver.f90:577-594)
  static void scheme(int m, double dt, double dt_old, double &C1, double &C2, double &C3, double &C4) {
    if (m == kEulerImplicit) {
      C1 = 1.0; C2 = 0.0; C3 = 1.0; C4 = 0.0;
    } else if (m == kCrankNicholson) {
      C1 = 2.0; C2 = 1.0; C3 = 1.0; C4 = 0.0;
    } else {
      const double R = dt / dt_old;
      C1 = (1. + 2. * R) / (1. + R);
      C2 = 0.0;
      C3 = ((1. + R) * (1. + R)) / (1. + 2. * R);
      C4 = (-(R * R)) / (1. + 2. * R);
    }
  }

  // the step context of F_nonlinear / Jacobi: scheme, kinematics at time+dt,
  // tau_s, and theta_dot at the new level for the current iterate
  void set_context(double time, double dt, double dt_old, int m, const Beam &b, const double *th_old,
                   const double *thd_old) {
    cx.method = m;
    cx.time = time; cx.dt = dt; cx.dt_old = dt_old;
    scheme(m, dt, dt_old, cx.C1, cx.C2, cx.C3, cx.C4);
    motion(time + dt, cx.alpha, cx.alpha_t, cx.alpha_tt, cx.LE);
    cx.th_old = th_old; cx.thd_old = thd_old;
    cx.th_oo = b.oldold.data() + 4 * (size_t)ns;
    cx.thd_oo = b.oldold.data() + 5 * (size_t)ns;
    cx.p = b.pressure_new.data();
    cx.tau = b.tau_new.data();
    dn.d1(cx.tau, cx.tau_s.data(), ns);
  }
  void theta_dot_new(const double *th) {
    for (int i = 0; i < ns; ++i)
      cx.thd_new[i] = (cx.C1 / cx.dt) * (th[i] - cx.C3 * cx.th_old[i] - cx.C4 * cx.th_oo[i]) - cx.C2 * cx.thd_old[i];
  }

  // F_nonlinear (SolidSolver.f90:552-667).  x in the reference's order;
  // emit(k, value) for equations k = 1 .. 2 ns + 4.
  template <class Emit>
  void residual(const double *x, Emit &&emit) {
    const double *th = x + 1;            // th[-1] .. th[ns+1]
    const double *T = x + ns + 3 + 1;    // T[-1] .. T[ns-1]
    const double eta = p.eta, mue = p.mue, sigma = p.sigma;
    const double *pr = cx.p, *tau = cx.tau, *tau_s = cx.tau_s.data(), *orhs = cx.old_rhs.data();
    const double a = cx.alpha, at = cx.alpha_t, att = cx.alpha_tt;
    const double C1 = cx.C1, C2 = cx.C2, C3 = cx.C3, C4 = cx.C4, dt = cx.dt;
    theta_dot_new(th);
    const double *tdn = cx.thd_new.data();
    const double K1 = mue * (cx.LE[4] * cos(a) + cx.LE[5] * sin(a) + grav * sin(a)) - tau[0];
    const double K2 = pr[0] + mue * (cx.LE[5] * cos(a) - cx.LE[4] * sin(a) + grav * cos(a));
    const double ds2 = ds * ds, ds3 = ds * ds * ds, ds4 = ds2 * ds2;
    const int n = ns;
    emit(1, th[0]);
    emit(2, (T[1] - T[-1]) / (2.0 * ds) + (eta / (2.0 * ds3)) * (th[-1] - 2.0 * th[0] + th[1]) * (th[1] - th[-1]) - K1);
    emit(3, T[0] * (th[1] - th[-1]) / (2.0 * ds) -
                (eta / (2.0 * ds3)) * (-3.0 * th[-1] + 10.0 * th[0] - 12.0 * th[1] + 6.0 * th[2] - th[3]) - K2);
    emit(4, ((1.0 / 12.0) * th[n - 3] - (2.0 / 3.0) * th[n - 2] + (2.0 / 3.0) * th[n] - (1.0 / 12.0) * th[n + 1]) / ds);
    emit(5, ((-1.0 / 12.0) * th[n - 3] + (4.0 / 3.0) * th[n - 2] - 2.5 * th[n - 1] + (4.0 / 3.0) * th[n] +
             (-1.0 / 12.0) * th[n + 1]) /
                ds2);
    emit(6, T[n - 1]);
    {
      const double q = (th[1] - th[-1]) / (2.0 * ds);
      const double c = (th[1] - 2.0 * th[0] + th[-1]) / ds2;
      const double w = tdn[0] + at;
      emit(7, (T[-1] - 2.0 * T[0] + T[1]) / ds2 - T[0] * (q * q) + pr[0] * q +
                  (eta / ds4) * (th[1] - th[-1]) *
                      (-2.5 * th[0] + 9.0 * th[1] - 12.0 * th[2] + 7.0 * th[3] - 1.5 * th[4]) +
                  eta * (c * c) + mue * (w * w) - tau_s[0]);
    }
    {
      const int e = n - 1;
      const double q = (th[n] - th[n - 2]) / (2. * ds);
      const double in = -(3. * pr[e] - 4. * pr[e - 1] + pr[e - 2]) / (2. * ds) -
                        eta * (th[n + 1] - 4. * th[n] + 6. * th[e] - 4. * th[n - 2] + th[n - 3]) / ds4 +
                        (T[e] + eta * (q * q)) * (th[n] - 2. * th[e] + th[n - 2]) / ds2 +
                        (3. * T[e] - 4. * T[e - 1] + T[e - 2]) * (th[n] - th[n - 2]) / (2. * ds2) - att * (mue * mue) -
                        sigma * tdn[e] + tau[e] * (th[n] - th[n - 2]) / (2.0 * ds) + C2 * mue * orhs[e];
      emit(8, C3 * cx.thd_old[e] + C4 * cx.thd_oo[e] - tdn[e] + (dt / (C1 * mue)) * in);
    }
    for (int i = 1; i <= n - 2; ++i) {
      const double q = (th[i + 1] - th[i - 1]) / (2. * ds);
      const double in = -1.0 * (pr[i + 1] - pr[i - 1]) / (2. * ds) -
                        eta * (th[i + 2] - 4. * th[i + 1] + 6. * th[i] - 4. * th[i - 1] + th[i - 2]) / ds4 +
                        (T[i] + eta * (q * q)) * (th[i + 1] - 2. * th[i] + th[i - 1]) / ds2 +
                        (T[i + 1] - T[i - 1]) * (th[i + 1] - th[i - 1]) / (2.0 * ds2) - att * (mue * mue) -
                        sigma * tdn[i] + tau[i] * (th[i + 1] - th[i - 1]) / (2.0 * ds) + C2 * mue * orhs[i];
      emit(8 + i, C3 * cx.thd_old[i] + C4 * cx.thd_oo[i] - tdn[i] + (dt / (C1 * mue)) * in);
      const double q2 = (th[i + 1] - th[i - 1]) / (2.0 * ds);
      const double c = (th[i + 1] - 2.0 * th[i] + th[i - 1]) / ds2;
      const double w = tdn[i] + at;
      emit(8 + (n - 2) + i,
           (T[i - 1] - 2.0 * T[i] + T[i + 1]) / ds2 - T[i] * (q2 * q2) + pr[i] * (th[i + 1] - th[i - 1]) / (2.0 * ds) +
               eta * (th[i + 1] - th[i - 1]) * (th[i + 2] - 2.0 * th[i + 1] + 2.0 * th[i - 1] - th[i - 2]) /
                   (2.0 * ds4) +
               eta * (c * c) + mue * (w * w) - tau_s[i]);
    }
  }

  // Jacobi (SolidSolver.f90:671-863): emit(k, l, dF_k/dx_l), the entry the
  // reference stores as J(l, k).  The Jacobian only steers Newton's
  // iterates (the root is F's), so its divisions by the grid constants may
  // become products with their reciprocals (a rounding-level change).
  template <class Emit>
  void jacobian(const double *x, Emit &&emit) {
#pragma clang fp reciprocal(on)
    const double *th = x + 1;
    const double *T = x + ns + 3 + 1;
    const double eta = p.eta, mue = p.mue, sigma = p.sigma;
    const double *pr = cx.p, *tau = cx.tau;
    const double at = cx.alpha_t;
    const double C1 = cx.C1, dt = cx.dt;
    theta_dot_new(th);
    const double *tdn = cx.thd_new.data();
    const double D = C1 / dt;
    const double ds2 = ds * ds, ds3 = ds * ds * ds, ds4 = ds2 * ds2;
    const int n = ns, T0 = n + 4;
    emit(1, 2, 1.0);
    emit(2, 1, (eta / (2.0 * ds3)) * (-2.0 * th[-1] + 2.0 * th[0]));
    emit(2, 2, -(eta / ds3) * (th[1] - th[-1]));
    emit(2, 3, (eta / (2.0 * ds3)) * (-2.0 * th[0] + 2.0 * th[1]));
    emit(2, T0, -0.5 / ds);
    emit(2, T0 + 2, 0.5 / ds);
    emit(3, 1, (-0.5 * eta / ds3) * (-3.0) - T[0] / (2.0 * ds));
    emit(3, 2, (-0.5 * eta / ds3) * (10.0));
    emit(3, 3, (-0.5 * eta / ds3) * (-12.0) + T[0] / (2.0 * ds));
    emit(3, 4, (-0.5 * eta / ds3) * (6.0));
    emit(3, 5, (-0.5 * eta / ds3) * (-1.0));
    emit(3, T0 + 1, (th[1] - th[-1]) / (2.0 * ds));
    emit(4, n + 3, (-1.0 / 12.0) / ds);
    emit(4, n + 2, (2.0 / 3.0) / ds);
    emit(4, n, (-2.0 / 3.0) / ds);
    emit(4, n - 1, (1.0 / 12.0) / ds);
    emit(5, n + 3, (-1.0 / 12.0) / ds2);
    emit(5, n + 2, (4.0 / 3.0) / ds2);
    emit(5, n + 1, (-5.0 / 2.0) / ds2);
    emit(5, n, (4.0 / 3.0) / ds2);
    emit(5, n - 1, (-1.0 / 12.0) / ds2);
    emit(6, 2 * n + 4, 1.0);
    {
      const double st = -2.5 * th[0] + 9.0 * th[1] - 12.0 * th[2] + 7.0 * th[3] - 1.5 * th[4];
      const double cc = th[1] - 2.0 * th[0] + th[-1];
      const double dd = th[1] - th[-1];
      emit(7, 1, T[0] * dd / (2.0 * ds2) - pr[0] / (2.0 * ds) - (eta / ds4) * st + 2.0 * eta * cc / ds4);
      emit(7, 2, -2.5 * eta * dd / ds4 - 4.0 * eta * cc / ds4 + 2.0 * mue * (tdn[0] + at) * D);
      emit(7, 3, -T[0] * dd / (2.0 * ds2) + pr[0] / (2.0 * ds) + (eta / ds4) * st + 9.0 * eta * dd / ds4 +
                     2.0 * eta * cc / ds4);
      emit(7, 4, -12.0 * eta * dd / ds4);
      emit(7, 5, 7.0 * eta * dd / ds4);
      emit(7, 6, -1.5 * eta * dd / ds4);
      emit(7, n + 4, 1.0 / ds2);
      emit(7, n + 5, (-2.0 - (dd * dd) / 4.0) / ds2);
      emit(7, n + 6, 1.0 / ds2);
    }
    {
      const int e = n - 1;
      const double g = dt / (C1 * mue);
      const double dd = th[n] - th[n - 2];
      const double cc = th[n] - 2.0 * th[e] + th[n - 2];
      const double q = dd / (2.0 * ds);
      const double Tq = T[e] + eta * (q * q);
      const double tg = 3.0 * T[e] - 4.0 * T[e - 1] + T[e - 2];
      emit(8, 2 * n + 4, (dt / (C1 * mue * ds2)) * (cc + (3.0 / 2.0) * dd));
      emit(8, 2 * n + 3, -4.0 * g * dd / (2.0 * ds2));
      emit(8, 2 * n + 2, 1.0 * g * dd / (2.0 * ds2));
      emit(8, n + 3, -dt * eta / ((C1 * mue) * ds4));
      emit(8, n + 2, (dt / (C1 * mue * ds2)) * (4.0 * eta / ds2 + Tq + tau[e] * ds / (2.0) +
                                                   eta * cc * dd / (2.0 * ds2) + tg / 2.0));
      emit(8, n, (dt / (C1 * mue * ds2)) * (4.0 * eta / ds2 + Tq - tau[e] * ds / (2.0) - eta * cc * dd / (2.0 * ds2) -
                                              tg / 2.0));
      emit(8, n + 1, -D + g * (-6.0 * eta / ds4 - 2.0 * Tq / ds2 - D * sigma));
      emit(8, n - 1, -dt * eta / (C1 * mue * ds4));
    }
    const double g = dt / (C1 * mue);
    for (int i = 1; i <= n - 2; ++i) {
      const int k = 8 + i, l = i + 2, m = i + n + 5;
      const double dd = th[i + 1] - th[i - 1];
      const double cc = th[i + 1] - 2.0 * th[i] + th[i - 1];
      const double q = dd / (2.0 * ds);
      const double Tq = T[i] + eta * (q * q);
      emit(k, l, -D + g * (-6.0 * eta / ds4 - (2.0 / ds2) * Tq - D * sigma));
      emit(k, l - 1, g * (4.0 * eta / ds4 + Tq / ds2 - tau[i] / (2.0 * ds) - eta * cc * dd / (2.0 * ds4) -
                          (T[i + 1] - T[i - 1]) / (2.0 * ds2)));
      emit(k, l + 1, g * (4.0 * eta / ds4 + Tq / ds2 + tau[i] / (2.0 * ds) + eta * cc * dd / (2.0 * ds4) +
                          (T[i + 1] - T[i - 1]) / (2.0 * ds2)));
      emit(k, l - 2, -eta * dt / (C1 * mue * ds4));
      emit(k, l + 2, -eta * dt / (C1 * mue * ds4));
      emit(k, m, g * cc / ds2);
      emit(k, m - 1, -g * dd / (2.0 * ds2));
      emit(k, m + 1, g * dd / (2.0 * ds2));
    }
    for (int i = 1; i <= n - 2; ++i) {
      const int k = 8 + n - 2 + i, l = i + 2, m = i + n + 5;
      const double dd = th[i + 1] - th[i - 1];
      const double cc = th[i + 1] - 2.0 * th[i] + th[i - 1];
      const double st = th[i + 2] - 2.0 * th[i + 1] + 2.0 * th[i - 1] - th[i - 2];
      emit(k, l, -4.0 * eta * cc / ds4 + 2.0 * mue * (tdn[i] + at) * D);
      emit(k, l - 1, T[i] * dd / (2.0 * ds2) - pr[i] / (2.0 * ds) + eta * dd / ds4 - eta * st / (2.0 * ds4) +
                         2.0 * eta * cc / ds4);
      emit(k, l + 1, -T[i] * dd / (2.0 * ds2) + pr[i] / (2.0 * ds) - eta * dd / ds4 + eta * st / (2.0 * ds4) +
                         2.0 * eta * cc / ds4);
      emit(k, l - 2, -eta * dd / (2.0 * ds4));
      emit(k, l + 2, eta * dd / (2.0 * ds4));
      const double q = dd / (2.0 * ds);
      emit(k, m, -2.0 / ds2 - (q * q));
      emit(k, m - 1, 1.0 / ds2);
      emit(k, m + 1, 1.0 / ds2);
    }
  }

  // SolidEnergies (SolidSolver.f90:76-92)
  void energies(Beam &b) const {
    std::vector<double> th_s(ns);
    dn.d1(b.theta.data(), th_s.data(), ns);
    double sk = 0.0, sp = 0.0, se = 0.0, fx = 0.0, fy = 0.0;
    for (int i = 0; i < ns; ++i) {
      sk = sk + (b.vx[i] * b.vx[i] + b.vy[i] * b.vy[i]);
      sp = sp + (b.y[i] - y0);
      se = se + th_s[i] * th_s[i];
      fx = fx + b.ax[i];
      fy = fy + b.ay[i];
    }
    b.E_kinetic = p.mue * 0.5 * ds * sk;
    b.E_pot = p.mue * grav * ds * sp;
    b.E_elastic = p.eta * 0.5 * ds * se;
    b.Inertial_Force[0] = p.mue * ds * fx;
    b.Inertial_Force[1] = p.mue * ds * fy;
  }

  void integrate(double time, Beam &b) const {
    integrate_position_host(p, p.iMotion, x0, y0, phase, time, b.theta.data(), b.theta_dot.data(), b.x.data(),
                            b.y.data(), b.vx.data(), b.vy.data());
  }

  // the scheme of the next IBES step (a BDF2 run starts with Crank-Nicolson)
  int step_method(const Beam &b) const {
    return (b.StartupStep && method == kBDF2) ? kCrankNicholson : method;
  }

  // old_rhs of Crank-Nicolson: RHS_beameqn at t_n (SolidSolver.f90:220-224)
  bool prepare_old_rhs(double time, int m, const Beam &b) {
    if (m == kCrankNicholson) {
      std::vector<double> th(b.theta), T(ns);
      cx.old_rhs = b.theta_dot;
      if (!rhs_beameqn(time, th.data(), cx.old_rhs.data(), b.pressure_old.data(), b.tau_old.data(), T.data()))
        return false;
    } else {
      std::fill(cx.old_rhs.begin(), cx.old_rhs.end(), 0.0);
    }
    return true;
  }

  // the context of the step t_n -> t_n + dt for the beam b: scheme, old
  // levels, old_rhs (Crank-Nicolson), kinematics at t_n + dt
  bool begin(double time, double dt, const Beam &b, int m) {
    th_old_ = b.theta;
    thd_old_ = b.theta_dot;
    if (!prepare_old_rhs(time, m, b)) return false;
    set_context(time, dt, b.dt_old, m, b, th_old_.data(), thd_old_.data());
    return true;
  }

  // initial guess: an explicit Euler step with the loads at t_n, then the
  // ghost values of the boundary conditions (SolidSolver.f90:167-209)
  // The part of the predictor that needs only the state at t_n and the old
  // loads (RHS_beameqn with pressure_old, tau_old, gravity at t_n), so that
  // a host can compute it while the fluid step that brings the new loads
  // still runs.  guess() uses it when its time matches, else computes it.
  bool predict(double time, const Beam &b) {
    pre_ok_ = false;
    gravity_impulse(time);
    pre_th1_ = b.theta;
    std::vector<double> thd1(b.theta_dot);
    pre_T_.assign(ns, 0.0);
    if (!rhs_beameqn(time, pre_th1_.data(), thd1.data(), b.pressure_old.data(), b.tau_old.data(), pre_T_.data()))
      return false;
    pre_time_ = time;
    pre_ok_ = true;
    return true;
  }

  bool guess(double time, double dt, const Beam &b, double *x) {
    double alpha, at, att, LE[6];
    motion(time + dt, alpha, at, att, LE);
    if (!(pre_ok_ && pre_time_ == time) && !predict(time, b)) return false;
    pre_ok_ = false;  // used once
    const std::vector<double> &th1 = pre_th1_, &T = pre_T_;
    std::vector<double> tg(ns);
    for (int i = 0; i < ns; ++i) tg[i] = b.theta[i] + dt * th1[i];
    const double eta = p.eta, mue = p.mue, ds2 = ds * ds;
    double *thg = x + 1;  // thg[-1] .. thg[ns+1]
    for (int i = 0; i < ns; ++i) thg[i] = tg[i];
    const double K2 = b.pressure_new[0] + mue * (LE[5] * cos(alpha) - LE[4] * sin(alpha) + grav * cos(alpha));
    const double C2 = 2.0 * ds * K2 - T[0] * thg[1] + (eta / ds2) * (10.0 * thg[0] - 12.0 * thg[1] + 6.0 * thg[2] - thg[3]);
    thg[-1] = C2 / ((3.0 * eta / ds2) - T[0]);
    const double A1 = tg[ns - 3] / 12. - 2. * tg[ns - 2] / 3.;
    const double A2 = -tg[ns - 3] / 12. + 4. * tg[ns - 2] / 3. - 5. * tg[ns - 1] / 2.;
    thg[ns] = 3. * (A1 - A2) / 2.;
    thg[ns + 1] = 12. * (2. * A1 - A2);
    const double K1 = mue * (LE[4] * cos(alpha) + LE[5] * sin(alpha) + grav * sin(alpha)) - b.tau_new[0];
    const double th_ss0 = (2.0 * thg[0] - 5.0 * thg[1] + 4.0 * thg[2] - 1.0 * thg[3]) / ds2;
    const double th_s0 = (-1.5 * thg[0] + 2.0 * thg[1] - 0.5 * thg[2]) / ds;
    const double T_s0 = -eta * th_ss0 * th_s0 + K1;
    double *Tg = x + ns + 3 + 1;  // Tg[-1] .. Tg[ns-1]
    Tg[-1] = T[1] - 2.0 * ds * T_s0;
    for (int i = 0; i < ns; ++i) Tg[i] = T[i];
    return true;
  }

  // IBES_solver (SolidSolver.f90:99-414)
  bool ibes(double time, double dt, Beam &b) {
    const int m = step_method(b);
    b.StartupStep = false;
    const std::vector<double> x_old(b.x), y_old(b.y), vx_old(b.vx), vy_old(b.vy);
    std::vector<double> x(nx), dx(nx), rhs(nx), tmp(nx);
    if (!guess(time, dt, b, x.data())) return false;
    if (!begin(time, dt, b, m)) return false;
    const std::vector<double> &th_old = th_old_, &thd_old = thd_old_;
    const double *thg = x.data() + 1;
    // Newton (SolidSolver.f90:229-316).  The first iteration factors its
    // Jacobian J_0.  A later iteration first takes the chord step
    // dx = -J_0^-1 F_k with that factorisation.  It differs from the Newton
    // step -J_k^-1 F_k by about rho |dx|, rho = ||J_0^-1 (J_k - J_0)||, and
    // since F_k is what the previous step left, |dx| is about rho times the
    // previous |dx|: a chord step at most 1e-4 of the previous one (or below
    // the rounding of x itself, 1e-14 max|x|, as in a converged last
    // iteration) bounds rho by about 1e-4 and is taken, within 1e-8 of the
    // previous step of the Newton iterate, which the next iteration's
    // correction removes.  Otherwise the iteration assembles J_k: it solves
    // with J_0's factorisation and one step of iterative refinement against
    // J_k (a correction c of about rho |dx|, leaving rho^2 |dx|), or, if
    // |c| > 1e-4 |dx|, factors J_k instead.
    iter = 0;
    refactored = 0;
    double dm_prev = 0.0;
    for (;;) {
      residual(x.data(), [&](int k, double v) { rhs[rowpos[k]] = -1.0 * v; });
      bool have = false;
      double xm = 0.0;
      if (iter > 0) {
        for (int l = 0; l < nx; ++l) xm = std::max(xm, std::fabs(x[l]));
        dx = rhs;
        lu.solve_factored(dx.data());
        double dm = 0.0;
        for (int r = 0; r < nx; ++r) dm = std::max(dm, std::fabs(dx[r]));
        have = dm <= 1e-4 * dm_prev || dm <= 1e-14 * xm;
      }
      if (!have) {
        band.reset(nx);
        if (jofs.empty()) {
          bool outside = false;
          jacobian(x.data(), [&](int k, int l, double v) {
            const int r = rowpos[k], c = colpos[l];
            if (c < r - kKL || c > r + kKU) outside = true;
            else { band.at(r, c) = v; jofs.push_back(Band::index(r, c)); }
          });
          if (outside) { jofs.clear(); err = "IBES: Jacobian entry outside the band"; return false; }
          jlast = band.last;
        } else {
          double *a = band.a.data();
          size_t e = 0;
          jacobian(x.data(), [&](int, int, double v) { a[jofs[e++]] = v; });
          band.last = jlast;
        }
        bool fresh = iter == 0;
        if (!fresh) {
          // dx holds the chord step J_0^-1 (-F_k)
          band.matvec(dx.data(), tmp.data());
          double cm = 0.0, dm = 0.0;
          for (int r = 0; r < nx; ++r) tmp[r] = rhs[r] - tmp[r];
          lu.solve_factored(tmp.data());
          for (int r = 0; r < nx; ++r) {
            cm = std::max(cm, std::fabs(tmp[r]));
            dm = std::max(dm, std::fabs(dx[r]));
          }
          if (cm <= 1e-4 * dm || cm <= 1e-14 * xm) {
            for (int r = 0; r < nx; ++r) dx[r] = dx[r] + tmp[r];
          } else {
            fresh = true;
            ++refactored;
          }
        }
        if (fresh) {
          std::swap(lu, band);
          if (!lu.factor()) { err = "IBES: singular Newton matrix"; return false; }
          dx = rhs;
          lu.solve_factored(dx.data());
        }
      }
      dm_prev = 0.0;
      for (int r = 0; r < nx; ++r) dm_prev = std::max(dm_prev, std::fabs(dx[r]));
      ++iter;
      // dx is in band order; the update and the norms run in the reference's order
      double s2 = 0.0, x2 = 0.0;
      for (int l = 1; l <= nx; ++l) x[l - 1] = x[l - 1] + dx[colpos[l]];
      for (int l = 1; l <= nx; ++l) s2 = s2 + dx[colpos[l]] * dx[colpos[l]];
      for (int l = 0; l < nx; ++l) x2 = x2 + x[l] * x[l];
      for (int l = 0; l < nx; ++l)
        if (std::isnan(x[l])) { err = "IBES: NaN in the Newton iterate"; return false; }
      const double e = sqrt(s2), erel = std::fabs(sqrt(s2) / sqrt(x2));
      const bool done = (e < kErrorStop || erel < kErrorStop) && iter > 2;
      // the reference stops the run past 49 iterations, converged or not
      if (iter > kMaxIter) { err = "IBES: Newton did not converge in 49 iterations"; return false; }
      if (done) break;
    }
    // cut the ghost points, theta_dot of the scheme (SolidSolver.f90:322-348)
    for (int i = 0; i < ns; ++i) b.theta[i] = thg[i];
    double C1, C2s, C3, C4;
    scheme(m, dt, b.dt_old, C1, C2s, C3, C4);
    const double *th_oo = b.oldold.data() + 4 * (size_t)ns;
    for (int i = 0; i < ns; ++i)
      b.theta_dot[i] = (C1 / dt) * (b.theta[i] - C3 * th_old[i] - C4 * th_oo[i]) - C2s * thd_old[i];
    integrate(time + dt, b);
    for (int i = 0; i < ns; ++i) {
      b.ax[i] = (b.vx[i] - vx_old[i]) / dt;
      b.ay[i] = (b.vy[i] - vy_old[i]) / dt;
    }
    double jump = 0.0;
    for (int i = 0; i < ns; ++i) jump = std::max(jump, std::fabs(b.theta_dot[i] - thd_old[i]));
    brake = jump > 100.0;
    // beam_oldold = beam_old, dt_old = dt (SolidSolver.f90:405-411)
    const std::vector<double> *cols[6] = {&x_old, &y_old, &vx_old, &vy_old, &th_old, &thd_old};
    for (int c = 0; c < 6; ++c)
      for (int i = 0; i < ns; ++i) b.oldold[(size_t)c * ns + i] = (*cols[c])[i];
    b.dt_old = dt;
    return true;
  }

  // GravityImpulse(time); SolidSolverWrapper(time, dt, beams)
  // (time_step.f90:209-210, SolidSolver.f90:13-44); then the loads advance,
  // pressure_old = pressure_new (time_step.f90:225-228)
  bool step(double time, double dt, Beam &b) {
    gravity_impulse(time);
    if (!wrapper(time, dt, b)) return false;
    b.pressure_old = b.pressure_new;
    b.tau_old = b.tau_new;
    return true;
  }

  // SolidSolverWrapper (SolidSolver.f90:13-44) alone: an IBES step once the
  // beam is released, else the leading edge's motion only; SolidEnergies
  bool wrapper(double time, double dt, Beam &b) {
    brake = false;
    iter = 0;
    if (time > p.T_release) {
      if (!ibes(time, dt, b)) return false;
    } else {
      integrate(time + dt, b);
    }
    energies(b);
    return true;
  }
};

// init_beams for one beam (init_beam.f90:80-120) at placement (x0, y0,
// phase) and InitializeSolidSolver
inline void init_beam(const fsibe_params &p, Beam &b, double x0, double y0, double phase) {
  b.resize(p.ns);
  double alpha, at, att, LE[6];
  mouvement_host(p, p.iMotion, x0, y0, p.AngleBeam, phase, 0.0, alpha, at, att, LE);
  const double ds = p.length / (double)(p.ns - 1);
  for (int n = 0; n < p.ns; ++n) {
    b.x[n] = LE[0] + (double)n * ds;
    b.y[n] = LE[1];
  }
  b.dt_old = 0.0;
  b.StartupStep = true;
  integrate_position_host(p, p.iMotion, x0, y0, phase, 0.0, b.theta.data(), b.theta_dot.data(), b.x.data(),
                          b.y.data(), b.vx.data(), b.vy.data());
}

}  // namespace ibes
