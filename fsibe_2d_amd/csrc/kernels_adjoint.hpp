// kernels_adjoint.hpp -- the discrete adjoint of the fused fluid step for
// prescribed kinematics (no reference counterpart: the reference has no
// adjoint).  Each kernel is the transpose of one stage of kernels_step.hpp,
// run in reverse order, with the same FFT engine:
//
//   stage (forward)                       adjoint kernel (reverse order)
//   K6 rows  u = V(vortk(n1)), cal_drag   ak_rows_vel:  J-term, V^T rows
//   K5 cols  psi, i ky psi                ak_cols_ab2:  V^T columns, then
//   K4 cols  nlk, AB2 update                            the AB2 and nlk transposes
//   K2 rows  nonlinear + penalisation     ak_rows_nl:   products, masks
//   K1 cols  A, B of vortk(n0)            ak_cols_w0
//   K0 mask  create_mask                  ak_mask:      d(mask)/d(kinematics)
//
// The state of the reverse sweep: lW = dJ/d vortk (half spectrum), lNL =
// dJ/d nlk of the previous step, lu1/lu2 = dJ/du (physical).  Transposes
// over the real inner product of the stored arrays (the real and imaginary
// parts of each stored half-spectrum entry are two coordinates):
//   column inverse DFT (unnormalised)    -> column forward DFT (unnormalised)
//   column forward DFT x norm            -> norm x column inverse DFT
//   row c2r: f = Re X0 + 2 sum_k Re(Xk e^{i th})  -> X0 = R[f]_0, Xk = 2 R[f]_k
//   row r2c (unnormalised) R             -> f = c2r(Y), Y0 = X0, Yk = Xk / 2
//   spectral multiplier m                -> conj(m)
#pragma once
#include "kernels_mask.hpp"
#include "kernels_step.hpp"

namespace fsibe {

// ---- ak_rows_vel (transpose of K6's velocity, plus the objective) ----------
// J = cJ sum_window m (u1 - vx) at this step; lu1 += cJ m on the window;
// lm = cJ (u1 - vx), lvx = -cJ m there (the mask adjoints are assigned here on
// the whole mask box and added to by ak_rows_nl).  Then Ta = R^T[lu1],
// Tb = i kx R^T[lu2] (u1 = R[Cpsiy] + U, u2 = -R[i kx Cpsi] + U).
template <class FX>
__global__ void __launch_bounds__(FX::LB) ak_rows_vel(const double *__restrict__ lu1, const double *__restrict__ lu2,
                                                      const double *__restrict__ u1n, MaskBox mb, Window win,
                                                      double cJ, double *__restrict__ lm, double *__restrict__ lvx,
                                                      double *__restrict__ lvy, double2 *__restrict__ Ta,
                                                      double2 *__restrict__ Tb, FX F,
                                                      const double2 *__restrict__ twx, StepConsts c) {
#pragma clang fp contract(off)
  extern __shared__ double2 lds[];
  const int H = F.H(), nx = 2 * H, tid = threadIdx.x, nt = blockDim.x, PL = padlen(H);
  const int y = blockIdx.x;
  double2 *f0 = lds, *f1 = lds + PL;
  const bool mrow = mb.x1 >= mb.x0 && y >= mb.y0 && y <= mb.y1;
  const bool wrow = y >= win.y0 && y <= win.y1;
  for (int n = tid; n < H; n += nt) {
    double2 a = reinterpret_cast<const double2 *>(lu1 + (size_t)y * nx)[n];
    const double2 b = reinterpret_cast<const double2 *>(lu2 + (size_t)y * nx)[n];
    if (mrow) {
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int x = 2 * n + h;
        if (!mb.inside(x, y)) continue;
        const size_t qq = (size_t)(y - mb.y0) * mb.w + (x - mb.x0);
        double gm = 0.0, gvx = 0.0;
        if (wrow && x >= win.x0 && x <= win.x1) {
          const double m = mb.m[qq];
          const double uu = h ? reinterpret_cast<const double2 *>(u1n + (size_t)y * nx)[n].y
                              : reinterpret_cast<const double2 *>(u1n + (size_t)y * nx)[n].x;
          gm = cJ * (uu - mb.vx[qq]);
          gvx = -(cJ * m);
          if (h) a.y += cJ * m; else a.x += cJ * m;
        }
        lm[qq] = gm;
        lvx[qq] = gvx;
        lvy[qq] = 0.0;
      }
    }
    f0[pidx(n)] = a;
    f1[pidx(n)] = b;
  }
  __syncthreads();
  F.template fft<false, 2>(lds, twx, tid, nt);
  rows_r2c_post(lds, 2, H, PL, twx, tid, nt);
  for (int k = tid; k < H; k += nt) {
    const double wk = k == 0 ? 1.0 : 2.0;
    const double2 xa = f0[pidx(k)], xb = f1[pidx(k)];
    const double kf = (double)k * c.sx;
    Ta[(size_t)y * H + k] = make_double2(wk * xa.x, wk * xa.y);
    // i kx (wk xb)
    Tb[(size_t)y * H + k] = make_double2(-(kf * (wk * xb.y)), kf * (wk * xb.x));
  }
}

// ---- ak_cols_ab2 (transposes of K5's psi columns, K4's AB2 and nlk) ---------
// lW holds dJ/d vortk(n1) from the later steps; on return it holds the part
// of dJ/d vortk(n0) that passes through the AB2 update, lNL the part of
// dJ/d nlk(n-1), and Tc, Td the row-pass adjoints of Rc, Rpx.
template <class FY>
__global__ void __launch_bounds__(FY::LB) ak_cols_ab2(const double2 *__restrict__ Ta, const double2 *__restrict__ Tb,
                                                      double2 *__restrict__ lW, double2 *__restrict__ lNL,
                                                      double2 *__restrict__ Tc, double2 *__restrict__ Td,
                                                      const DevScalars *__restrict__ st, int explicit_step, FY F,
                                                      const double2 *__restrict__ twy, StepConsts c) {
#pragma clang fp contract(off)
  extern __shared__ double2 lds[];
  const int ny = F.NY(), TX = F.tx(), H = c.H, tid = threadIdx.x, nt = blockDim.x;
  const int kx0 = xcd_remap(blockIdx.x, gridDim.x) * TX;
  const int TT = 2 * TX;
  for (int e = tid; e < ny * TX; e += nt) {
    const int y = e / TX, cc = e - y * TX;
    lds[pidx(y) * TT + cc] = Ta[(size_t)y * H + kx0 + cc];
    lds[pidx(y) * TT + TX + cc] = Tb[(size_t)y * H + kx0 + cc];
  }
  __syncthreads();
  F.template fft<false, 2, 2>(lds, twy, tid, nt);
  const DevScalars s = *st;
  const int hy = ny >> 1;
  for (int e = tid; e < ny * TX; e += nt) {
    const int idx = e / TX, cc = e - idx * TX, kx = kx0 + cc;
    const size_t g = (size_t)idx * H + kx;
    double2 lnl = make_double2(0.0, 0.0), lw0 = make_double2(0.0, 0.0), lnl1 = make_double2(0.0, 0.0);
    if (idx != hy) {
      const int ky = ky_signed(idx, ny);
      const int aky = ky < 0 ? -ky : ky;
      const double kyf = (double)ky * c.sy;
      const double2 ya = lds[pidx(idx) * TT + cc], yb = lds[pidx(idx) * TT + TX + cc];
      // d psi = -i ky Ya + Yb, then the Poisson factor (psi = W1 / |k|^2)
      double2 lpsi = make_double2(kyf * ya.y + yb.x, -(kyf * ya.x) + yb.y);
      double2 lw1 = lW[g];
      if (!(kx == 0 && ky == 0)) {
        const double quot = ((double)(kx * kx) * c.scalex + (double)(aky * aky) * c.scaley);
        lw1.x += lpsi.x / quot;
        lw1.y += lpsi.y / quot;
      }
      const double kx2 = (double)(kx * kx), ky2 = (double)(aky * aky);
      const double v1 = exp(s.cx1 * kx2 + s.cy1 * ky2);
      const double d = dealias(c, kx, aky);
      const double f = v1 * d;
      const double2 ln0 = lNL[g];
      lw0 = make_double2(f * lw1.x, f * lw1.y);
      lnl = make_double2(s.b10 * lw0.x + ln0.x, s.b10 * lw0.y + ln0.y);
      if (!explicit_step) {
        const double v0 = exp(s.cx0 * kx2 + s.cy0 * ky2);
        lnl1 = make_double2(s.b11 * v0 * lw0.x, s.b11 * v0 * lw0.y);
      }
    }
    lW[g] = lw0;
    lNL[g] = lnl1;
    // nl = norm Cfwd(Rc) - i ky norm Cfwd(Rpx):  dRc = norm Cinv(lnl), dRpx = norm Cinv(i ky lnl)
    const double kyf = (double)ky_signed(idx, ny) * c.sy;
    lds[pidx(idx) * TT + cc] = make_double2(c.norm * lnl.x, c.norm * lnl.y);
    lds[pidx(idx) * TT + TX + cc] = make_double2(c.norm * (-(kyf * lnl.y)), c.norm * (kyf * lnl.x));
  }
  __syncthreads();
  F.template fft<true, 2, 2>(lds, twy, tid, nt);
  for (int e = tid; e < ny * TX; e += nt) {
    const int y = e / TX, cc = e - y * TX;
    Tc[(size_t)y * H + kx0 + cc] = lds[pidx(y) * TT + cc];
    Td[(size_t)y * H + kx0 + cc] = lds[pidx(y) * TT + TX + cc];
  }
}

// ---- ak_rows_nl (transpose of K2) -------------------------------------------
// Forward values omega, omega_y, omega_x (c2r of A, B, i kx A) and the
// adjoints of the products lr, lpy, lpx (Rc = R[r] + i kx R[py], Rpx =
// R[px]) in one six-sequence transform; then the product rule, and the
// weighted r2c of l(omega), l(omega_x), l(omega_y):  Te = dA, Tf = dB.
// lu1, lu2 (dJ/du of this step's velocity) are assigned; the mask adjoints
// lm, lvx, lvy are added to.
template <class FX>
__global__ void __launch_bounds__(FX::LB) ak_rows_nl(const double2 *__restrict__ Tc, const double2 *__restrict__ Td,
                                                     const double2 *__restrict__ A, const double2 *__restrict__ B,
                                                     const double *__restrict__ u1, const double *__restrict__ u2,
                                                     const double *__restrict__ ms, MaskBox mb, int penal_on,
                                                     double *__restrict__ lu1, double *__restrict__ lu2,
                                                     double *__restrict__ lm, double *__restrict__ lvx,
                                                     double *__restrict__ lvy, double2 *__restrict__ Te,
                                                     double2 *__restrict__ Tf, FX F, const double2 *__restrict__ twx,
                                                     StepConsts c) {
#pragma clang fp contract(off)
  extern __shared__ double2 lds[];
  const int H = F.H(), nx = 2 * H, tid = threadIdx.x, nt = blockDim.x, PL = padlen(H);
  const int y = blockIdx.x;
  double2 *f0 = lds, *f1 = lds + PL, *f2 = lds + 2 * PL;
  double2 *g0 = lds + 3 * PL, *g1 = lds + 4 * PL, *g2 = lds + 5 * PL;
  for (int k = tid; k < H; k += nt) {
    const double2 a = A[(size_t)y * H + k];
    const double kf = (double)k * c.sx;
    const int q = pidx(k);
    f0[q] = a;
    f1[q] = B[(size_t)y * H + k];
    f2[q] = make_double2(-kf * a.y, kf * a.x);
    // transposes of the unnormalised r2c: Y0 = X0, Yk = Xk / 2 (k >= 1)
    const double hk = k == 0 ? 1.0 : 0.5;
    const double2 tc = Tc[(size_t)y * H + k], td = Td[(size_t)y * H + k];
    g0[q] = make_double2(hk * tc.x, hk * tc.y);                   // l(r)
    g1[q] = make_double2(hk * (kf * tc.y), hk * (-(kf * tc.x)));  // l(py) = -i kx Tc
    g2[q] = make_double2(hk * td.x, hk * td.y);                   // l(px)
  }
  __syncthreads();
  rows_c2r_pre(lds, 6, H, PL, twx, tid, nt);
  F.template fft<true, 6>(lds, twx, tid, nt);
  const bool prow = penal_on && mb.x1 >= mb.x0 && y >= mb.y0 && y <= mb.y1;
  for (int n = tid; n < H; n += nt) {
    const int q = pidx(n);
    const double2 w = f0[q], wy = f1[q], wx = f2[q];
    const double2 lr = g0[q], lpy = g1[q], lpx = g2[q];
    const double2 a1 = reinterpret_cast<const double2 *>(u1 + (size_t)y * nx)[n];
    const double2 a2 = reinterpret_cast<const double2 *>(u2 + (size_t)y * nx)[n];
    double2 s = make_double2(0.0, 0.0);
    if (ms) s = reinterpret_cast<const double2 *>(ms + (size_t)y * nx)[n];
    // r = -u1 wx - u2 wy - s w
    double2 lu1v = make_double2(-(wx.x * lr.x), -(wx.y * lr.y));
    double2 lu2v = make_double2(-(wy.x * lr.x), -(wy.y * lr.y));
    const double2 lw = make_double2(-(s.x * lr.x), -(s.y * lr.y));
    const double2 lwx = make_double2(-(a1.x * lr.x), -(a1.y * lr.y));
    const double2 lwy = make_double2(-(a2.x * lr.x), -(a2.y * lr.y));
    if (prow) {
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int x = 2 * n + h;
        if (!mb.inside(x, y)) continue;
        const size_t qq = (size_t)(y - mb.y0) * mb.w + (x - mb.x0);
        const double m = mb.m[qq], vx = mb.vx[qq], vy = mb.vy[qq];
        const double uu1 = h ? a1.y : a1.x, uu2 = h ? a2.y : a2.x;
        const double gpx = h ? lpx.y : lpx.x, gpy = h ? lpy.y : lpy.x;
        // px = -m (u1 - vx), py = -m (u2 - vy)
        if (h) { lu1v.y += -(m * gpx); lu2v.y += -(m * gpy); }
        else { lu1v.x += -(m * gpx); lu2v.x += -(m * gpy); }
        lvx[qq] += m * gpx;
        lvy[qq] += m * gpy;
        lm[qq] += -((uu1 - vx) * gpx) - (uu2 - vy) * gpy;
      }
    }
    reinterpret_cast<double2 *>(lu1 + (size_t)y * nx)[n] = lu1v;
    reinterpret_cast<double2 *>(lu2 + (size_t)y * nx)[n] = lu2v;
    g0[q] = lw;
    g1[q] = lwy;
    g2[q] = lwx;
  }
  __syncthreads();
  // weighted r2c (transpose of c2r) of l(omega), l(omega_y), l(omega_x)
  F.template fft<false, 3>(g0, twx, tid, nt);
  rows_r2c_post(g0, 3, H, PL, twx, tid, nt);
  for (int k = tid; k < H; k += nt) {
    const int q = pidx(k);
    const double wk = k == 0 ? 1.0 : 2.0;
    const double kf = (double)k * c.sx;
    const double2 xw = g0[q], xwy = g1[q], xwx = g2[q];
    // dA = R^T l(w) + conj(i kx) R^T l(wx), dB = R^T l(wy)
    Te[(size_t)y * H + k] = make_double2(wk * (xw.x + kf * xwx.y), wk * (xw.y - kf * xwx.x));
    Tf[(size_t)y * H + k] = make_double2(wk * xwy.x, wk * xwy.y);
  }
}

// ---- ak_cols_w0 (transpose of K1) -------------------------------------------
// lW += Cfwd(Te) - i ky Cfwd(Tf)  (A = Cinv(W0), B = Cinv(i ky W0))
template <class FY>
__global__ void __launch_bounds__(FY::LB) ak_cols_w0(const double2 *__restrict__ Te, const double2 *__restrict__ Tf,
                                                     double2 *__restrict__ lW, FY F,
                                                     const double2 *__restrict__ twy, StepConsts c) {
#pragma clang fp contract(off)
  extern __shared__ double2 lds[];
  const int ny = F.NY(), TX = F.tx(), H = c.H, tid = threadIdx.x, nt = blockDim.x;
  const int kx0 = xcd_remap(blockIdx.x, gridDim.x) * TX;
  const int TT = 2 * TX;
  for (int e = tid; e < ny * TX; e += nt) {
    const int y = e / TX, cc = e - y * TX;
    lds[pidx(y) * TT + cc] = Te[(size_t)y * H + kx0 + cc];
    lds[pidx(y) * TT + TX + cc] = Tf[(size_t)y * H + kx0 + cc];
  }
  __syncthreads();
  F.template fft<false, 2, 2>(lds, twy, tid, nt);
  const int hy = ny >> 1;
  for (int e = tid; e < ny * TX; e += nt) {
    const int idx = e / TX, cc = e - idx * TX, kx = kx0 + cc;
    const size_t g = (size_t)idx * H + kx;
    if (idx == hy) {
      lW[g] = make_double2(0.0, 0.0);
      continue;
    }
    const double kyf = (double)ky_signed(idx, ny) * c.sy;
    const double2 ye = lds[pidx(idx) * TT + cc], yf = lds[pidx(idx) * TT + TX + cc];
    const double2 l = lW[g];
    lW[g] = make_double2(l.x + (ye.x + kyf * yf.y), l.y + (ye.y - kyf * yf.x));
  }
}

// ---- ak_mask: d(mask)/d(kinematics) of a rigid beam -------------------------
// For prescribed motion with theta = 0 the beam is the straight line from
// the leading edge LE = (x0, LEy) at angle alpha; node i sits at
// LE + s_i t (t = (cos a, sin a)), with velocity (-at s_i sin a,
// hdot + at s_i cos a).  The pixel's value came from primitive `win`
// (create_mask's drawing order: segments 0..ns-2, hinges 1..ns-2, the
// leading-edge dot, the trailing-edge dot, after `off` fixed primitives).
// Sums over the tile of lm d(m)/dq + lvx d(vx)/dq + lvy d(vy)/dq for
// q = (LEy, alpha, hdot, alpha_t); one record of 4 partial sums per tile.
struct RigidBeam {
  double LEx, LEy, alpha, hdot, at;
  const double *s;  // s_i of the nodes (integrate_position's sums)
  int ns, off;
};

__device__ __forceinline__ double smooth_step_d(const MaskConsts &c, double x, double t) {
#pragma clang fp contract(off)
  if (c.sharp) return 0.0;
  const double delta = c.h * c.gerf;
  const double a = (x + t) / delta, b = (t - x) / delta;
  return (exp(-(a * a)) - exp(-(b * b))) / (delta * sqrt(c.pi));
}

__global__ void __launch_bounds__(256) ak_mask(const int *__restrict__ win, const double *__restrict__ lm,
                                               const double *__restrict__ lvx, const double *__restrict__ lvy,
                                               MaskConsts c, RigidBeam rb, int bx0, int by0, int bw, int bh,
                                               double *__restrict__ part) {
#pragma clang fp contract(off)
  __shared__ double red[4 * 4];
  const int npx = bw * bh;
  const double ca = cos(rb.alpha), sa = sin(rb.alpha);
  double g[4] = {0.0, 0.0, 0.0, 0.0};
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < npx; q += gridDim.x * blockDim.x) {
    const int p = win[q];
    if (p < rb.off) continue;
    const int ly = q / bw, lx = q - ly * bw;
    const double x = (double)(bx0 + lx) * c.dx, y = (double)(by0 + ly) * c.dy;
    const double rx = x - rb.LEx, ry = y - rb.LEy;
    const double xi = ca * rx + sa * ry;    // along the beam
    const double d = -(sa * rx) + ca * ry;  // across the beam
    const double gm = lm[q] / c.eps, gvx = lvx[q], gvy = lvy[q];
    const int k = p - rb.off, ns = rb.ns;
    if (k <= ns - 2) {
      // segment: m = SS(|d|), (vx, vy) = v(xi)
      const double ad = fabs(d), sg = d < 0.0 ? -1.0 : 1.0;
      const double dv = smooth_step_d(c, ad, c.t) * sg;
      g[0] += gm * dv * (-ca) + gvx * (rb.at * sa * sa) + gvy * (-(rb.at * ca * sa));
      g[1] += gm * dv * (-xi) + gvx * (-(rb.at * (ca * xi + sa * d))) + gvy * (rb.at * (ca * d - sa * xi));
      g[2] += gvy;
      g[3] += gvx * (-(sa * xi)) + gvy * (ca * xi);
    } else {
      // hinge (node k - ns + 2) or dot (leading / trailing edge): m = SS(R)
      int node;
      if (k <= 2 * ns - 4) node = k - (ns - 1) + 1;
      else if (k == 2 * ns - 3) node = 0;
      else node = ns - 1;
      const double si = rb.s[node];
      const double px = rb.LEx + si * ca, py = rb.LEy + si * sa;
      const double ex = x - px, ey = y - py;
      const double R = sqrt(ex * ex + ey * ey);
      if (R > 0.0) {
        const double dv = smooth_step_d(c, R, c.t);
        // dR/dLEy = -ey / R, dR/dalpha = -(e . s_i n) / R with n = (-sa, ca)
        g[0] += gm * dv * (-ey / R);
        g[1] += gm * dv * (-(si * (-(sa * ex) + ca * ey)) / R);
      }
      g[1] += gvx * (-(rb.at * si * ca)) + gvy * (-(rb.at * si * sa));
      g[2] += gvy;
      g[3] += gvx * (-(si * sa)) + gvy * (si * ca);
    }
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < 4; ++i) g[i] = wave_sum(g[i]);
  if (lane == 0)
#pragma unroll
    for (int i = 0; i < 4; ++i) red[i * 4 + wid] = g[i];
  __syncthreads();
  if (threadIdx.x < 4) {
    const int i = threadIdx.x;
    part[(size_t)blockIdx.x * 4 + i] = ((red[i * 4] + red[i * 4 + 1]) + red[i * 4 + 2]) + red[i * 4 + 3];
  }
}

// device-to-device copies of n doubles / ints (tape and checkpoints)
__global__ void ak_copy_d(const double *__restrict__ src, double *__restrict__ dst, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    dst[i] = src[i];
}
__global__ void ak_copy_i(const int *__restrict__ src, int *__restrict__ dst, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    dst[i] = src[i];
}
__global__ void ak_zero(double *__restrict__ dst, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    dst[i] = 0.0;
}

}  // namespace fsibe
