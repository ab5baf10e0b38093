// kernels_step.hpp -- fused kernels of one fluid step (EvolveFluidAB2 /
// EvolveFluidExplicit + pressure + cal_drag + cal_energy + time_avg_vort,
// NavierStokes.f90:9-508) and of GetForces (BeamForces.f90:63-288).
//
// The reference's 19 FFT-based transforms per step are regrouped by linearity
// into one pass over rows and one pass over column tiles per stage; every
// elementwise operation runs inside the FFT kernel that produces or consumes
// its operands:
//
//   K1 cols  W0 -> A = Fy^-1[W0], B = Fy^-1[i ky W0]
//   K2 rows  A,B -> omega, d_x omega, d_y omega (3 c2r), with u, mask, sponge:
//            a = -u.grad(omega) - chi_s omega, (px,py) = -chi (u - u_s);
//            Rc = Fx[a] + i kx Fx[py], Rpx = Fx[px]          (3 r2c)
//   K3 1 WG  CFL dt1, AB2 weights, mean flow                  (cal_nlk :59-78)
//   K4 cols  nlk = N(Rc) - i ky N(Rpx); vortk(n1) = AB2 update (:478-498)
//   K5 cols  vortk(n1) -> Fy^-1 of omega, psi, i ky psi, -ky^2 psi (cal_velocity)
//   K6 rows  omega, u1, u2; vor_mean; energy/drag/penal partial sums;
//            pressure rhs q = chi(u-u_s) + omega x u -> Fx  (pressure :305-370)
//   K7 cols  div rhs -> poisson -> Fy^-1 on the rows near the beam
//   K8 rows  p = q - |u|^2/2 and the stresses on the rows near the beam
//   K9/10    reductions, Delta interpolation at the beam (GetForcesInterp)
#pragma once
#include "kernels.hpp"

namespace fsibe {

struct StepConsts {
  int nx, ny, H;
  double sx, sy;           // 2 pi / xl, 2 pi / yl
  double scalex, scaley;   // params.f90:140-141
  double norm;             // 1/(nx ny), cof_fftw33.f90:162
  double kxt2, kyt2;       // (2/3 kx_max)^2, dealiase_mask.f90:16-17
  int idealis;
  double nu, eps, dx, dy;
};

struct MaskBox {
  int x0, x1, y0, y1, w;  // inclusive box, pitch w; x1 < x0 means empty
  const double *m, *vx, *vy;
  __device__ __forceinline__ bool inside(int x, int y) const {
    return x >= x0 && x <= x1 && y >= y0 && y <= y1;
  }
};

// scalars computed on the device each step (no host round trip for dt1)
struct DevScalars {
  double dt1, u_max, b10, b11;
  double cx1, cy1, cx0, cy0;   // vis exponents -dt nu scale for dt1 and dt0
  double UMx, UMy;             // U_mean_true * Mean_u(time+dt1)
  double Mx, My;               // Mean_u(time+dt1)
  double time, dt0;
  double drag_unst_old, lift_unst_old;
  int ready, pad;
};

struct FlowParams {
  int iMeanVelocity, penal_on;
  double theta_inf, T_fullspeed, U_mean_true, cfl, eps, dt_fixed, Time_end, pi, dx, dy, nu,
      scalex, scaley;
};

// mean_velocity.f90:4-84 (case 11 is rejected on the host)
__device__ __forceinline__ void mean_u(const FlowParams &f, double time, double &ux, double &uy) {
#pragma clang fp contract(off)
  double tf = 1.0;
  switch (f.iMeanVelocity) {
    case 1:
      if (time < f.T_fullspeed) {
        const double t1 = f.T_fullspeed;
        tf = (time * time * time) / (-0.5 * t1 * t1 * t1) + 3 * (time * time) / (t1 * t1);
      }
      ux = cos(f.theta_inf) * tf; uy = sin(f.theta_inf) * tf; return;
    case 2:
      if (time < f.T_fullspeed) tf = 0.5 * (1.0 - cos(0.5 * f.pi * time / f.T_fullspeed));
      ux = cos(f.theta_inf) * tf; uy = sin(f.theta_inf) * tf; return;
    case 3: ux = cos(f.theta_inf); uy = sin(f.theta_inf); return;
    case 6: ux = time; uy = 0.0; return;
    case 22: ux = 0.1; uy = 0.0; return;
    default: ux = 0.0; uy = 0.0; return;
  }
}

// dealiase_mask.f90:21 for mode (kx, |ky|)
__device__ __forceinline__ double dealias(const StepConsts &c, int kx, int ky) {
#pragma clang fp contract(off)
  if (!c.idealis) return 1.0;
  const double a = (double)kx, b = (double)ky;
  return (a * a / c.kxt2 + b * b / c.kyt2 >= 1.0) ? 0.0 : 1.0;
}

// ---- deterministic block reductions (fixed shape for fixed blockDim) --------
__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// reduces NV values per thread to thread 0; red must hold NV*nwaves doubles
template <int NV>
__device__ void block_sum(double (&v)[NV], double *red) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
  for (int i = 0; i < NV; ++i) v[i] = wave_sum(v[i]);
  __syncthreads();
  if (lane == 0)
#pragma unroll
    for (int i = 0; i < NV; ++i) red[i * nw + wid] = v[i];
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      double s = 0.0;
      for (int w = 0; w < nw; ++w) s += red[i * nw + w];
      v[i] = s;
    }
  }
  __syncthreads();
}

// ============================================================================
// K1 / K5: inverse column transforms of one spectral field into several
// derived fields.  Grid: ntiles * nfield blocks, one (tile, field) each.
// field 0: F, 1: psi = F/|k|^2, 2: i ky psi, 3: -ky^2 psi, 4: i ky F
// ============================================================================
template <int EMAX>
__global__ void __launch_bounds__(kNT) k_cols_inv_fields(const double2 *__restrict__ W, double2 *__restrict__ o0,
                                                         double2 *__restrict__ o1, double2 *__restrict__ o2,
                                                         double2 *__restrict__ o3, int fsel0, int fsel1,
                                                         int fsel2, int fsel3, int nfield, FftPlan plan,
                                                         const double2 *__restrict__ twy, StepConsts c, int TX) {
  extern __shared__ double2 lds[];
  const int ny = plan.L, H = c.H, tid = threadIdx.x;
  const int lb = xcd_remap(blockIdx.x, gridDim.x);
  const int tile = lb / nfield, slot = lb - tile * nfield;
  const int kx0 = tile * TX;
  const int fsel = slot == 0 ? fsel0 : slot == 1 ? fsel1 : slot == 2 ? fsel2 : fsel3;
  double2 *out = slot == 0 ? o0 : slot == 1 ? o1 : slot == 2 ? o2 : o3;
  const int hy = ny >> 1;
  for (int e = tid; e < ny * TX; e += kNT) {
    const int idx = e / TX, cc = e - idx * TX, kx = kx0 + cc;
    const double2 w = W[(size_t)idx * H + kx];
    double2 v;
    if (fsel == 0) {
      v = w;
    } else if (fsel == 4) {
      const double kyf = (double)ky_signed(idx, ny) * c.sy;
      v = make_double2(-kyf * w.y, kyf * w.x);
    } else {
      const int ky = ky_signed(idx, ny);
      const int aky = ky < 0 ? -ky : ky;
      double2 psi;
      if ((kx == 0 && ky == 0) || idx == hy) {
        psi = make_double2(0.0, 0.0);
      } else {
        const double quot = ((double)(kx * kx) * c.scalex + (double)(aky * aky) * c.scaley);
        psi = make_double2(w.x / quot, w.y / quot);
      }
      const double kyf = (double)ky * c.sy;
      if (fsel == 1) v = psi;
      else if (fsel == 2) v = make_double2(-kyf * psi.y, kyf * psi.x);
      else v = make_double2(-(kyf * kyf) * psi.x, -(kyf * kyf) * psi.y);
    }
    lds[e] = v;
  }
  __syncthreads();
  lds_fft<EMAX, true, true>(lds, plan, TX, 1, TX, twy, 1, tid, kNT);
  for (int e = tid; e < ny * TX; e += kNT) {
    const int y = e / TX, cc = e - y * TX;
    out[(size_t)y * H + kx0 + cc] = lds[e];
  }
}

// ============================================================================
// K2: nonlinear term, penalisation and sponge in physical space (one row/WG)
// cal_nlk, NavierStokes.f90:39-52 and :88-164
// ============================================================================
template <int EMAX>
__global__ void __launch_bounds__(kNT) k_rows_nl(const double2 *__restrict__ A, const double2 *__restrict__ B,
                                                 const double *__restrict__ u1, const double *__restrict__ u2,
                                                 const double *__restrict__ ms, MaskBox mb, int penal_on,
                                                 double2 *__restrict__ Rc, double2 *__restrict__ Rpx, FftPlan plan,
                                                 const double2 *__restrict__ twx, StepConsts c) {
  extern __shared__ double2 lds[];
  const int H = c.H, nx = c.nx, tid = threadIdx.x;
  const int y = blockIdx.x;
  double2 *f0 = lds, *f1 = lds + H, *f2 = lds + 2 * H;
  const bool prow = penal_on && mb.x1 >= mb.x0 && y >= mb.y0 && y <= mb.y1;
  for (int k = tid; k < H; k += kNT) {
    const double2 a = A[(size_t)y * H + k];
    const double kf = (double)k * c.sx;
    f0[k] = a;
    f1[k] = B[(size_t)y * H + k];
    f2[k] = make_double2(-kf * a.y, kf * a.x);
  }
  __syncthreads();
  rows_c2r_pre(lds, 3, H, twx, tid, kNT);
  lds_fft<EMAX, true, false>(lds, plan, 3, H, 1, twx, 2, tid, kNT);
  for (int n = tid; n < H; n += kNT) {
    const double2 w = f0[n], wy = f1[n], wx = f2[n];
    const double2 a1 = reinterpret_cast<const double2 *>(u1 + (size_t)y * nx)[n];
    const double2 a2 = reinterpret_cast<const double2 *>(u2 + (size_t)y * nx)[n];
    double2 s = make_double2(0.0, 0.0);
    if (ms) s = reinterpret_cast<const double2 *>(ms + (size_t)y * nx)[n];
    double2 r;
    r.x = (-(a1.x * wx.x) - a2.x * wy.x) - s.x * w.x;
    r.y = (-(a1.y * wx.y) - a2.y * wy.y) - s.y * w.y;
    double2 px = make_double2(0.0, 0.0), py = make_double2(0.0, 0.0);
    if (prow) {
      const int x = 2 * n;
      if (mb.inside(x, y)) {
        const size_t q = (size_t)(y - mb.y0) * mb.w + (x - mb.x0);
        px.x = -(mb.m[q] * (a1.x - mb.vx[q]));
        py.x = -(mb.m[q] * (a2.x - mb.vy[q]));
      }
      if (mb.inside(x + 1, y)) {
        const size_t q = (size_t)(y - mb.y0) * mb.w + (x + 1 - mb.x0);
        px.y = -(mb.m[q] * (a1.y - mb.vx[q]));
        py.y = -(mb.m[q] * (a2.y - mb.vy[q]));
      }
    }
    f0[n] = r;
    f1[n] = px;
    f2[n] = py;
  }
  __syncthreads();
  lds_fft<EMAX, false, false>(lds, plan, 3, H, 1, twx, 2, tid, kNT);
  rows_r2c_post(lds, 3, H, twx, tid, kNT);
  for (int k = tid; k < H; k += kNT) {
    const double2 xa = f0[k], xpx = f1[k], xpy = f2[k];
    const double kf = (double)k * c.sx;
    Rc[(size_t)y * H + k] = make_double2(xa.x - kf * xpy.y, xa.y + kf * xpy.x);
    Rpx[(size_t)y * H + k] = xpx;
  }
}

// ============================================================================
// K3: time step and AB2 weights (cal_nlk NavierStokes.f90:59-78,
// EvolveFluidAB2 :488-489, vis cal_vis.f90:20-21), one workgroup
// ============================================================================
__global__ void __launch_bounds__(kNT) k_step_scalars(const double *__restrict__ umax_part, int nparts,
                                                      DevScalars *__restrict__ s, FlowParams f, double time,
                                                      double dt0, int explicit_step) {
#pragma clang fp contract(off)
  __shared__ double red[kNT / 64];
  double m = 0.0;
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) m = fmax(m, umax_part[i]);
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) m = fmax(m, red[w]);
    const double u_max = sqrt(m);
    double dt1 = f.cfl * fmin(f.dx, f.dy) / u_max;
    if (u_max < 1.e-10) dt1 = 1.e-2;
    if (f.penal_on) dt1 = fmin(0.9 * f.eps, dt1);
    if (f.dt_fixed > 0.0) dt1 = fmin(f.dt_fixed, dt1);
    if ((f.Time_end - time) < dt1) dt1 = f.Time_end - time;
    s->u_max = u_max;
    s->dt1 = dt1;
    s->time = time;
    s->dt0 = dt0;
    if (explicit_step) {
      s->b10 = dt1;
      s->b11 = 0.0;
    } else {
      s->b10 = dt1 / dt0 * (0.5 * dt1 + dt0);
      s->b11 = -(0.5 * (dt1 * dt1)) / dt0;
    }
    s->cx1 = -(dt1 * f.nu * f.scalex);
    s->cy1 = -(dt1 * f.nu * f.scaley);
    s->cx0 = -(dt0 * f.nu * f.scalex);
    s->cy0 = -(dt0 * f.nu * f.scaley);
    double mx, my;
    mean_u(f, time + dt1, mx, my);
    s->Mx = mx;
    s->My = my;
    s->UMx = f.U_mean_true * mx;
    s->UMy = f.U_mean_true * my;
  }
}

// ============================================================================
// K4: nlk(n0) and the integrating-factor AB2 / Euler update of vortk
// (EvolveFluidAB2 NavierStokes.f90:478-498, EvolveFluidExplicit :444-449)
// ============================================================================
template <int EMAX>
__global__ void __launch_bounds__(kNT) k_cols_ab2(const double2 *__restrict__ Rc, const double2 *__restrict__ Rpx,
                                                  const double2 *__restrict__ W0, const double2 *__restrict__ NL1,
                                                  double2 *__restrict__ NL0, double2 *__restrict__ W1,
                                                  const DevScalars *__restrict__ s, int explicit_step,
                                                  FftPlan plan, const double2 *__restrict__ twy, StepConsts c,
                                                  int TX) {
  extern __shared__ double2 lds[];
  const int ny = plan.L, H = c.H, tid = threadIdx.x;
  const int kx0 = xcd_remap(blockIdx.x, gridDim.x) * TX;
  const int TT = 2 * TX;
  for (int e = tid; e < ny * TX; e += kNT) {
    const int y = e / TX, cc = e - y * TX;
    lds[y * TT + cc] = Rc[(size_t)y * H + kx0 + cc];
    lds[y * TT + TX + cc] = Rpx[(size_t)y * H + kx0 + cc];
  }
  __syncthreads();
  lds_fft<EMAX, false, true>(lds, plan, TT, 1, TT, twy, 1, tid, kNT);
  const double b10 = s->b10, b11 = s->b11, cx1 = s->cx1, cy1 = s->cy1, cx0 = s->cx0, cy0 = s->cy0;
  const int hy = ny >> 1;
  for (int e = tid; e < ny * TX; e += kNT) {
    const int idx = e / TX, cc = e - idx * TX, kx = kx0 + cc;
    const size_t g = (size_t)idx * H + kx;
    double2 nl = make_double2(0.0, 0.0), w1 = make_double2(0.0, 0.0);
    if (idx != hy) {
      const int ky = ky_signed(idx, ny);
      const int aky = ky < 0 ? -ky : ky;
      const double kyf = (double)ky * c.sy;
      const double2 fc = lds[idx * TT + cc], fp = lds[idx * TT + TX + cc];
      const double2 Fc = make_double2(fc.x * c.norm, fc.y * c.norm);
      const double2 Fp = make_double2(fp.x * c.norm, fp.y * c.norm);
      nl = make_double2(Fc.x + kyf * Fp.y, Fc.y - kyf * Fp.x);
      const double kx2 = (double)(kx * kx), ky2 = (double)(aky * aky);
      const double v1 = exp(cx1 * kx2 + cy1 * ky2);
      const double d = dealias(c, kx, aky);
      const double2 w0 = W0[g];
      if (explicit_step) {
        w1.x = (w0.x + b10 * nl.x) * v1 * d;
        w1.y = (w0.y + b10 * nl.y) * v1 * d;
      } else {
        const double v0 = exp(cx0 * kx2 + cy0 * ky2);
        const double2 n1 = NL1[g];
        const double2 n1v = make_double2(n1.x * v0, n1.y * v0);
        w1.x = ((w0.x + b10 * nl.x) + b11 * n1v.x) * v1 * d;
        w1.y = ((w0.y + b10 * nl.y) + b11 * n1v.y) * v1 * d;
      }
    }
    NL0[g] = nl;
    W1[g] = w1;
  }
}

// ============================================================================
// K6: velocity, vorticity, diagnostics and the pressure right-hand side
// (pressure NavierStokes.f90:305-370, cal_velocity :181-212, cal_drag
// :567-575, time_avg_vort :670-683, cal_energy :222-263).  One row per WG.
// mode 0: velocity only (EvolveFluidExplicit's cal_velocity at time, :440);
// mode 1: full (with diagnostics); pres: also build the pressure rhs.
// Partial sums per row (NSUM):
//  0 sum w^2   1 sum |u-M|^2  2 porous  3 inflow_solid  4 inflow_mean
//  5 sum p1    6 sum p2 (whole grid)    7 sum p1  8 sum p2  9 sum chi vx
//  10 sum chi vy (cal_drag window)
// ============================================================================
constexpr int kNSUM = 11;
struct Window { int x0, x1, y0, y1; };

template <int EMAX>
__global__ void __launch_bounds__(kNT) k_rows_press(const double2 *__restrict__ Cw, const double2 *__restrict__ Cpsi,
                                                    const double2 *__restrict__ Cpsiy, double *__restrict__ u1,
                                                    double *__restrict__ u2, double *__restrict__ vor_mean,
                                                    MaskBox mb, Window win, double *__restrict__ sums,
                                                    double *__restrict__ umax_part, double2 *__restrict__ Rq,
                                                    double2 *__restrict__ Rqy, const DevScalars *__restrict__ s,
                                                    double UMx0, double UMy0, int mode, int pres, FftPlan plan,
                                                    const double2 *__restrict__ twx, StepConsts c) {
  extern __shared__ double2 lds[];
  __shared__ double red[kNSUM * (kNT / 64)];
  const int H = c.H, nx = c.nx, tid = threadIdx.x;
  const int y = blockIdx.x;
  double2 *f0 = lds, *f1 = lds + H, *f2 = lds + 2 * H;
  const int nf = mode ? 3 : 2;
  double2 *g1 = mode ? f1 : f0, *g2 = mode ? f2 : f1;  // psi_x, psi_y slots
  for (int k = tid; k < H; k += kNT) {
    const double kf = (double)k * c.sx;
    const double2 p = Cpsi[(size_t)y * H + k];
    if (mode) f0[k] = Cw[(size_t)y * H + k];
    g1[k] = make_double2(-kf * p.y, kf * p.x);
    g2[k] = Cpsiy[(size_t)y * H + k];
  }
  __syncthreads();
  rows_c2r_pre(lds, nf, H, twx, tid, kNT);
  lds_fft<EMAX, true, false>(lds, plan, nf, H, 1, twx, 2, tid, kNT);
  double UMx = UMx0, UMy = UMy0, Mx = 0.0, My = 0.0, time = 0.0, dt1 = 0.0;
  if (mode) {
    UMx = s->UMx; UMy = s->UMy; Mx = s->Mx; My = s->My; time = s->time; dt1 = s->dt1;
  }
  const bool mrow = mb.x1 >= mb.x0 && y >= mb.y0 && y <= mb.y1;
  const bool wrow = y >= win.y0 && y <= win.y1;
  double acc[kNSUM];
#pragma unroll
  for (int i = 0; i < kNSUM; ++i) acc[i] = 0.0;
  double um = 0.0;
  for (int n = tid; n < H; n += kNT) {
    const double2 gx = g1[n], uy1 = g2[n];
    double2 a1, a2;
    {
#pragma clang fp contract(off)
      a1 = make_double2(uy1.x + UMx, uy1.y + UMx);
      a2 = make_double2(-gx.x + UMy, -gx.y + UMy);
    }
    reinterpret_cast<double2 *>(u1 + (size_t)y * nx)[n] = a1;
    reinterpret_cast<double2 *>(u2 + (size_t)y * nx)[n] = a2;
    um = fmax(um, fmax(a1.x * a1.x + a2.x * a2.x, a1.y * a1.y + a2.y * a2.y));
    if (!mode) continue;
    const double2 w = f0[n];
    {
      double2 *vm = reinterpret_cast<double2 *>(vor_mean + (size_t)y * nx) + n;
      const double2 v = *vm;
      *vm = make_double2((v.x * time + dt1 * w.x) / (time + dt1), (v.y * time + dt1 * w.y) / (time + dt1));
    }
    acc[0] += w.x * w.x + w.y * w.y;
    {
      const double p = a1.x - Mx, q = a2.x - My, r = a1.y - Mx, t = a2.y - My;
      acc[1] += (p * p + q * q) + (r * r + t * t);
    }
    double2 q1 = make_double2(0.0, 0.0), q2 = make_double2(0.0, 0.0);
    if (mrow) {
      const int x = 2 * n;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        if (!mb.inside(x + h, y)) continue;
        const size_t qq = (size_t)(y - mb.y0) * mb.w + (x + h - mb.x0);
        const double m = mb.m[qq], vx = mb.vx[qq], vy = mb.vy[qq];
        const double uu1 = h ? a1.y : a1.x, uu2 = h ? a2.y : a2.x;
        const double r1 = uu1 - vx, r2 = uu2 - vy;
        const double p1 = m * r1, p2 = m * r2;
        acc[2] += m * (r1 * r1 + r2 * r2);
        acc[3] += m * (vx * r1 + vy * r2);
        acc[4] += Mx * p1 + My * p2;
        acc[5] += p1;
        acc[6] += p2;
        if (wrow && x + h >= win.x0 && x + h <= win.x1) {
          acc[7] += p1;
          acc[8] += p2;
          acc[9] += vx * m;
          acc[10] += vy * m;
        }
        if (h) { q1.y = p1; q2.y = p2; } else { q1.x = p1; q2.x = p2; }
      }
    }
    if (pres) {
      // penal(:,:,1) - vor*uy, penal(:,:,2) + vor*ux   (:357-358)
      f0[n] = make_double2(q1.x - w.x * a2.x, q1.y - w.y * a2.y);
      f1[n] = make_double2(q2.x + w.x * a1.x, q2.y + w.y * a1.y);
    }
  }
  // u_max partial (next step's CFL)
  {
    um = wave_max(um);
    __syncthreads();
    if ((tid & 63) == 0) red[tid >> 6] = um;
    __syncthreads();
    if (tid == 0) {
      double m = 0.0;
      for (int w = 0; w < kNT / 64; ++w) m = fmax(m, red[w]);
      umax_part[y] = m;
    }
    __syncthreads();
  }
  if (!mode) return;
  block_sum<kNSUM>(acc, red);
  if (tid == 0)
    for (int i = 0; i < kNSUM; ++i) sums[(size_t)y * kNSUM + i] = acc[i];
  if (!pres) return;
  lds_fft<EMAX, false, false>(lds, plan, 2, H, 1, twx, 2, tid, kNT);
  rows_r2c_post(lds, 2, H, twx, tid, kNT);
  for (int k = tid; k < H; k += kNT) {
    const double2 xq = f0[k];
    const double kf = (double)k * c.sx;
    Rq[(size_t)y * H + k] = make_double2(-kf * xq.y, kf * xq.x);
    Rqy[(size_t)y * H + k] = f1[k];
  }
}

// ============================================================================
// K7: divergence + poisson (pressure NavierStokes.f90:366-382), then the
// inverse column transform; rows [r0, r1] are written.
// ============================================================================
template <int EMAX>
__global__ void __launch_bounds__(kNT) k_cols_press(const double2 *__restrict__ Rq, const double2 *__restrict__ Rqy,
                                                    double2 *__restrict__ Cq, int r0, int r1, FftPlan plan,
                                                    const double2 *__restrict__ twy, StepConsts c, int TX) {
  extern __shared__ double2 lds[];
  const int ny = plan.L, H = c.H, tid = threadIdx.x;
  const int kx0 = xcd_remap(blockIdx.x, gridDim.x) * TX;
  const int TT = 2 * TX;
  for (int e = tid; e < ny * TX; e += kNT) {
    const int y = e / TX, cc = e - y * TX;
    lds[y * TT + cc] = Rq[(size_t)y * H + kx0 + cc];
    lds[y * TT + TX + cc] = Rqy[(size_t)y * H + kx0 + cc];
  }
  __syncthreads();
  lds_fft<EMAX, false, true>(lds, plan, TT, 1, TT, twy, 1, tid, kNT);
  const int hy = ny >> 1;
  for (int e = tid; e < ny * TX; e += kNT) {
    const int idx = e / TX, cc = e - idx * TX, kx = kx0 + cc;
    const int ky = ky_signed(idx, ny);
    const int aky = ky < 0 ? -ky : ky;
    double2 q = make_double2(0.0, 0.0);
    if (idx != hy && !(kx == 0 && ky == 0)) {
      const double kyf = (double)ky * c.sy;
      const double2 a = lds[idx * TT + cc], b = lds[idx * TT + TX + cc];
      const double2 fa = make_double2(a.x * c.norm, a.y * c.norm);
      const double2 fb = make_double2(b.x * c.norm, b.y * c.norm);
      const double2 rhs = make_double2(fa.x - kyf * fb.y, fa.y + kyf * fb.x);
      const double quot = ((double)(kx * kx) * c.scalex + (double)(aky * aky) * c.scaley);
      q = make_double2(rhs.x / quot, rhs.y / quot);
    }
    // this thread is the only reader and writer of slot (idx, cc)
    lds[idx * TT + cc] = q;
  }
  __syncthreads();
  lds_fft<EMAX, true, true>(lds, plan, TX, 1, TT, twy, 1, tid, kNT);
  for (int e = tid; e < ny * TX; e += kNT) {
    const int y = e / TX, cc = e - y * TX;
    if (y >= r0 && y <= r1) Cq[(size_t)y * H + kx0 + cc] = lds[y * TT + cc];
  }
}

// ============================================================================
// K8: static pressure p = q - |u|^2/2 (:390) and the viscous stresses
// stress_a = 2 nu d_x u1, stress_b = nu (d_y u1 + d_x u2) (BeamForces.f90:
// 105-117) on rows r0+blockIdx.x, written for columns [bx0, bx1] with pitch bw.
// Spectrally d_x u1 = i kx (i ky psi), d_y u1 + d_x u2 = (kx^2 - ky^2) psi.
// ============================================================================
template <int EMAX>
__global__ void __launch_bounds__(kNT) k_rows_box(const double2 *__restrict__ Cq, const double2 *__restrict__ Cpsi,
                                                  const double2 *__restrict__ Cpsiy,
                                                  const double2 *__restrict__ Cpsiyy, const double *__restrict__ u1,
                                                  const double *__restrict__ u2, double *__restrict__ press,
                                                  double *__restrict__ sa, double *__restrict__ sb, int r0, int bx0,
                                                  int bx1, int bw, int visc, FftPlan plan,
                                                  const double2 *__restrict__ twx, StepConsts c) {
  extern __shared__ double2 lds[];
  const int H = c.H, nx = c.nx, tid = threadIdx.x;
  const int y = r0 + blockIdx.x;
  double2 *f0 = lds, *f1 = lds + H, *f2 = lds + 2 * H;
  const int nf = visc ? 3 : 1;
  for (int k = tid; k < H; k += kNT) {
    f0[k] = Cq[(size_t)y * H + k];
    if (visc) {
      const double kf = (double)k * c.sx;
      const double2 py = Cpsiy[(size_t)y * H + k];
      const double2 p = Cpsi[(size_t)y * H + k], pyy = Cpsiyy[(size_t)y * H + k];
      f1[k] = make_double2(-kf * py.y, kf * py.x);
      f2[k] = make_double2((kf * kf) * p.x + pyy.x, (kf * kf) * p.y + pyy.y);
    }
  }
  __syncthreads();
  rows_c2r_pre(lds, nf, H, twx, tid, kNT);
  lds_fft<EMAX, true, false>(lds, plan, nf, H, 1, twx, 2, tid, kNT);
  const double *q = reinterpret_cast<const double *>(f0);
  const double *A = reinterpret_cast<const double *>(f1);
  const double *Bv = reinterpret_cast<const double *>(f2);
  const size_t row = (size_t)blockIdx.x * bw;
  for (int x = bx0 + tid; x <= bx1; x += kNT) {
#pragma clang fp contract(off)
    const double v1 = u1[(size_t)y * nx + x], v2 = u2[(size_t)y * nx + x];
    press[row + (x - bx0)] = q[x] - 0.5 * (v1 * v1 + v2 * v2);
    if (visc) {
      sa[row + (x - bx0)] = 2.0 * c.nu * A[x];
      sb[row + (x - bx0)] = c.nu * Bv[x];
    }
  }
}

// ============================================================================
// K9: reductions of the per-row partial sums: cal_drag (Force, Force_unst,
// NavierStokes.f90:567-589), cal_energy (:236-256), BeamSurfaceIntegral
// drag/lift (BeamForces.f90:483-486).  One workgroup; thread i sums quantity
// i over rows in order.
// ============================================================================
struct StepOutDev {
  double dt1, u_max;
  double Force[2], Force_unst[2], Force_press[2];
  double energies[6];
  double bsi[10];
  double fx_int, fy_int;
  int status, pad;
};

__global__ void k_reduce_step(const double *__restrict__ sums, int ny, DevScalars *__restrict__ s,
                              StepOutDev *__restrict__ out, double dxdy, double nu, double eps,
                              double U_mean_true, int beam_on) {
#pragma clang fp contract(off)
  __shared__ double tot[kNSUM];
  const int i = threadIdx.x;
  if (i < kNSUM) {
    double a = 0.0;
    for (int y = 0; y < ny; ++y) a += sums[(size_t)y * kNSUM + i];
    tot[i] = a;
  }
  __syncthreads();
  if (i == 0) {
    out->dt1 = s->dt1;
    out->u_max = s->u_max;
    out->energies[0] = dxdy * sqrt(tot[0]);
    out->energies[1] = 0.5 * dxdy * tot[1];
    out->energies[2] = nu * dxdy * tot[0];
    out->energies[3] = dxdy * tot[2];
    out->energies[4] = dxdy * tot[3];
    out->energies[5] = dxdy * U_mean_true * tot[4];
    out->bsi[0] = tot[5] * dxdy;
    out->bsi[1] = tot[6] * dxdy;
    if (beam_on) {
      out->Force[0] = tot[7] * dxdy;
      out->Force[1] = tot[8] * dxdy;
      const double dn = tot[9] * dxdy * eps, ln = tot[10] * dxdy * eps;
      if (s->ready) {
        out->Force_unst[0] = (dn - s->drag_unst_old) / s->dt1;
        out->Force_unst[1] = (ln - s->lift_unst_old) / s->dt1;
      } else {
        out->Force_unst[0] = 0.0;
        out->Force_unst[1] = 0.0;
      }
      s->drag_unst_old = dn;
      s->lift_unst_old = ln;
      s->ready = 1;
    }
  }
}

// ============================================================================
// K10: GetForcesInterp with SPL=3 (BeamForces.f90:150-270) and the surface
// integral (:306-489).  Interpolation targets are prepared on the host.
// ============================================================================
struct ForceGeom {
  int ns, nc, nt;           // beam nodes, circle points, targets
  int bx0, by0, bw, bh;     // box of the press/stress arrays
  double dxi, dyi, pi;      // DeltaInterpolation spacing (Interpolation.f90:15-16)
  double ds, dL, soft;
  int visc, bsi_on;
};

// phi_4star, Interpolation.f90:82-91
__device__ __forceinline__ double phi4star(double x, double dx, double pi) {
#pragma clang fp contract(off)
  const double r = fabs(x / dx);
  if (r < 0.5) return (3. / 8.) + (pi / 32.) - 0.25 * (r * r);
  if (r >= 0.5 && r <= 1.5)
    return 0.25 + (1 - r) / 8. * sqrt(-2. + 8. * r - 4. * (r * r)) - asin(sqrt(2.) * (r - 1.)) / 8.;
  if (r >= 1.5 && r <= 2.5)
    return (17. / 16.) - (pi / 64.) - (3. * r / 4.) + ((r * r) / 8.) +
           (r - 2.) * sqrt(-14. + 16. * r - 4. * (r * r)) / 16. + asin(sqrt(2.) * (r - 2.)) / 16.;
  return 0.0;
}

// targets: tg[3*t] = x, tg[3*t+1] = y, tg[3*t+2] = field id (0 press, 1 sa, 2 sb)
__global__ void k_interp(const double *__restrict__ tg, int nt, const double *__restrict__ press,
                         const double *__restrict__ sa, const double *__restrict__ sb, ForceGeom g,
                         double *__restrict__ val) {
#pragma clang fp contract(off)
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  const double xt = tg[3 * t], yt = tg[3 * t + 1];
  const int fid = (int)tg[3 * t + 2];
  const double *f = fid == 0 ? press : fid == 1 ? sa : sb;
  const int i = (int)((xt - 0.0) / g.dxi), j = (int)((yt - 0.0) / g.dyi);
  double R1 = 0.0;
  for (int ix = i - 4; ix <= i + 4; ++ix) {
    const double xl = (double)ix * g.dxi + 0.0;
    const double px = phi4star(fabs(xl - xt), g.dxi, g.pi);
    for (int iy = j - 4; iy <= j + 4; ++iy) {
      const double yl = (double)iy * g.dyi + 0.0;
      R1 = R1 + px * phi4star(fabs(yl - yt), g.dyi, g.pi) * f[(size_t)(iy - g.by0) * g.bw + (ix - g.bx0)];
    }
  }
  val[t] = R1;
}

// node data nd[16*n + ...]: 0 sin th, 1 cos th, 2 sin 2th, 3 cos 2th
// circle data cd[8*i + ...] (trailing i < nc, leading nc <= i < 2nc):
//   0 cos phi, 1 sin phi, 2 cos 2phi, 3 sin 2phi
// interpolated values val: node n: 6n+0 P(u1) 6n+1 P(b1) 6n+2 A(u) 6n+3 B(u)
//   6n+4 A(b) 6n+5 B(b); circle point i: 6ns + 3i + {0 A, 1 B, 2 P}
__global__ void k_forces_final(const double *__restrict__ nd, const double *__restrict__ cd,
                               const double *__restrict__ val, ForceGeom g, double *__restrict__ pbeam,
                               double *__restrict__ tbeam, StepOutDev *__restrict__ out) {
#pragma clang fp contract(off)
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int ns = g.ns;
  double fx = 0.0, fy = 0.0;
  for (int n = 0; n < ns; ++n) {
    const double st = nd[16 * n], ct = nd[16 * n + 1], s2 = nd[16 * n + 2], c2 = nd[16 * n + 3];
    double pb = val[6 * n] - val[6 * n + 1];
    double tb = 0.0;
    if (g.visc) {
      const double A1 = val[6 * n + 2], B1 = val[6 * n + 3], A2 = val[6 * n + 4], B2 = val[6 * n + 5];
      const double sn1 = -(A1 * c2) - B1 * s2, st1 = -(A1 * s2) + B1 * c2;
      const double sn2 = -(A2 * c2) - B2 * s2, st2 = -(A2 * s2) + B2 * c2;
      tb = st1 - st2;
      pb = pb + (-sn1 + sn2);
    }
    pb = pb * g.soft;
    tb = tb * g.soft;
    pbeam[n] = pb;
    tbeam[n] = tb;
    fx = fx - (st * pb + ct * tb) * g.ds;
    fy = fy + (ct * pb + st * tb) * g.ds;
  }
  out->Force_press[0] = fx;
  out->Force_press[1] = fy;
  out->fx_int = fx;
  out->fy_int = fy;
  if (!g.bsi_on) return;
  double Ft0 = 0, Ft1 = 0, Fn0 = 0, Fn1 = 0, Fp0 = 0, Fp1 = 0;
  for (int i = 0; i < 2 * g.nc; ++i) {
    const double cp = cd[8 * i], sp = cd[8 * i + 1], c2 = cd[8 * i + 2], s2 = cd[8 * i + 3];
    const double A = val[6 * ns + 3 * i], B = val[6 * ns + 3 * i + 1], P = val[6 * ns + 3 * i + 2];
    const double sn = +A * c2 + B * s2, stt = -(A * s2) + B * c2, spp = -P;
    Ft0 = Ft0 + g.dL * stt * (-sp); Ft1 = Ft1 + g.dL * stt * (+cp);
    Fn0 = Fn0 + g.dL * sn * (+cp);  Fn1 = Fn1 + g.dL * sn * (+sp);
    Fp0 = Fp0 + g.dL * spp * (+cp); Fp1 = Fp1 + g.dL * spp * (+sp);
  }
  for (int n = 0; n < ns; ++n) {  // top side, BeamForces.f90:433-452
    const double st = nd[16 * n], ct = nd[16 * n + 1], s2 = nd[16 * n + 2], c2 = nd[16 * n + 3];
    const double A = val[6 * n + 2], B = val[6 * n + 3], P = val[6 * n];
    const double sn = -(A * c2) - B * s2, stt = -(A * s2) + B * c2, spp = -P;
    Ft0 = Ft0 + g.ds * stt * (+ct); Ft1 = Ft1 + g.ds * stt * (+st);
    Fn0 = Fn0 + g.ds * sn * (-st);  Fn1 = Fn1 + g.ds * sn * (+ct);
    Fp0 = Fp0 + g.ds * spp * (-st); Fp1 = Fp1 + g.ds * spp * (+ct);
  }
  for (int n = 0; n < ns; ++n) {  // bottom side, :462-480
    const double st = nd[16 * n], ct = nd[16 * n + 1], s2 = nd[16 * n + 2], c2 = nd[16 * n + 3];
    const double A = val[6 * n + 4], B = val[6 * n + 5], P = val[6 * n + 1];
    const double sn = -(A * c2) - B * s2, stt = -(A * s2) + B * c2, spp = -P;
    Ft0 = Ft0 + g.ds * stt * (-ct); Ft1 = Ft1 + g.ds * stt * (-st);
    Fn0 = Fn0 + g.ds * sn * (+st);  Fn1 = Fn1 + g.ds * sn * (-ct);
    Fp0 = Fp0 + g.ds * spp * (+st); Fp1 = Fp1 + g.ds * spp * (-ct);
  }
  out->bsi[2] = Ft0 + Fn0 + Fp0;
  out->bsi[3] = Ft1 + Fn1 + Fp1;
  out->bsi[4] = Ft0; out->bsi[5] = Fn0; out->bsi[6] = Fp0;
  out->bsi[7] = Ft1; out->bsi[8] = Fn1; out->bsi[9] = Fp1;
}

}  // namespace fsibe
