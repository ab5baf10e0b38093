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
//   K3       CFL dt1, AB2 weights, mean flow (block 0 of K2)  (cal_nlk :59-78)
//   K4 cols  nlk = N(Rc) - i ky N(Rpx); vortk(n1) = AB2 update (:478-498)
//   K5 cols  vortk(n1) -> Fy^-1 of omega, psi, i ky psi, -ky^2 psi (cal_velocity)
//   K6 rows  omega, u1, u2; vor_mean; energy/drag/penal partial sums;
//            pressure rhs q = chi(u-u_s) + omega x u -> Fx  (pressure :305-370)
//   K7 cols  div rhs -> poisson -> Fy^-1 on the rows near the beam
//   K8 rows  p = q - |u|^2/2 and the stresses on the rows near the beam
//   K10      Delta interpolation at the beam and the per-node force terms
//            (GetForcesInterp); the host adds the per-row sums of K6 and
//            the per-node terms of K10, which the kernels write into the
//            pinned result record, in a fixed order
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
  // iMeanVelocity = 11: the mean velocity u_mean and the mean penalisation
  // force fmean of the two time levels (share_vars.f90:23-24), carried
  // from step to step (the step kernels leave them alone)
  double u_mean[2];
  double fmean[2][2];
};

// per-step results; lives in pinned host memory that the kernels write
// directly (two slots, so the host reads one step while the next runs)
struct StepOutDev {
  double dt1, u_max;
  double Force[2], Force_unst[2], Force_press[2];
  double energies[6];
  double bsi[10];
  double fx_int, fy_int;
  int status, pad;
};

struct FlowParams {
  int iMeanVelocity, penal_on;
  double theta_inf, T_fullspeed, U_mean_true, cfl, eps, dt_fixed, Time_end, pi, dx, dy, nu,
      scalex, scaley;
};

// mean_velocity.f90:4-84; case 11 (u_mean, zero mean pressure) returns 0
// here and k_mean11 sets the step's mean velocity once u_mean is advanced
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
// Wave reductions on the VALU with data-parallel primitives (DPP moves of the
// two 32-bit halves; 64-bit operations take no DPP operand on this ISA):
// pairs and halves of each quad, rotations by 4 and 8 within each row of 16
// lanes, then the row broadcasts 15 and 31, which leave the total in lane
// 63.  Lanes that a row mask leaves out take `old`, a neutral term (0 for a
// sum, the lane's own value for a maximum).  The order is fixed, so results
// are deterministic; lane 63's value is broadcast to the wave.  Every lane of
// the wave must be active.
template <int CTRL, int ROWMASK = 0xf>
__device__ __forceinline__ double dpp_move(double v, double old) {
  const long long b = __double_as_longlong(v), o = __double_as_longlong(old);
  const int lo = __builtin_amdgcn_update_dpp((int)(o & 0xffffffffll), (int)(b & 0xffffffffll), CTRL, ROWMASK, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp((int)(o >> 32), (int)(b >> 32), CTRL, ROWMASK, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
__device__ __forceinline__ double lane63(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), 63);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
constexpr int kDppQuad1032 = 0xB1, kDppQuad2301 = 0x4E;  // quad_perm [1,0,3,2], [2,3,0,1]
constexpr int kDppRor4 = 0x124, kDppRor8 = 0x128, kDppBcast15 = 0x142, kDppBcast31 = 0x143;
__device__ __forceinline__ double wave_sum(double v) {
  v = v + dpp_move<kDppQuad1032>(v, 0.0);
  v = v + dpp_move<kDppQuad2301>(v, 0.0);
  v = v + dpp_move<kDppRor4>(v, 0.0);
  v = v + dpp_move<kDppRor8>(v, 0.0);
  v = v + dpp_move<kDppBcast15, 0xa>(v, 0.0);
  v = v + dpp_move<kDppBcast31, 0xc>(v, 0.0);
  return lane63(v);
}
__device__ __forceinline__ double wave_max(double v) {
  v = fmax(v, dpp_move<kDppQuad1032>(v, v));
  v = fmax(v, dpp_move<kDppQuad2301>(v, v));
  v = fmax(v, dpp_move<kDppRor4>(v, v));
  v = fmax(v, dpp_move<kDppRor8>(v, v));
  v = fmax(v, dpp_move<kDppBcast15, 0xa>(v, v));
  v = fmax(v, dpp_move<kDppBcast31, 0xc>(v, v));
  return lane63(v);
}

// the wave sums of v[0..NV) into dst[i*nw + wid] (lane 0 writes)
template <int NV>
__device__ __forceinline__ void wave_sums(const double *v, double *dst, int lane, int wid, int nw) {
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    const double q = wave_sum(v[i]);
    if (lane == 0) dst[i * nw + wid] = q;
  }
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
// the derived field fsel of spectral value w at (row idx, column kx)
__device__ __forceinline__ double2 inv_field_value(int fsel, double2 w, int idx, int kx, int ny, const StepConsts &c) {
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
    if ((kx == 0 && ky == 0) || idx == (ny >> 1)) {
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
  return v;
}

template <class FY>
__device__ __forceinline__ void cols_inv_fields_body(const double2 *__restrict__ W, double2 *__restrict__ o0,
                                                            double2 *__restrict__ o1, double2 *__restrict__ o2,
                                                            double2 *__restrict__ o3, int fsel0, int fsel1,
                                                            int fsel2, int fsel3, int nfield, FY F,
                                                            const double2 *__restrict__ twy, StepConsts c, int r0_3,
                                                            int r1_3, int bid, int nblk) {
  extern __shared__ double2 lds[];
  const int ny = F.NY(), TX = F.tx(), H = c.H, tid = threadIdx.x, nt = blockDim.x;
  const int lb = xcd_remap(bid, nblk);
  const int tile = lb / nfield, slot = lb - tile * nfield;
  const int kx0 = tile * TX;
  const int fsel = slot == 0 ? fsel0 : slot == 1 ? fsel1 : slot == 2 ? fsel2 : fsel3;
  double2 *out = slot == 0 ? o0 : slot == 1 ? o1 : slot == 2 ? o2 : o3;
  for (int e = tid; e < ny * TX; e += nt) {
    const int idx = e / TX, cc = e - idx * TX, kx = kx0 + cc;
    lds[pidx(idx) * TX + cc] = inv_field_value(fsel, W[(size_t)idx * H + kx], idx, kx, ny, c);
  }
  __syncthreads();
  F.template fft<true, 1, 1>(lds, twy, tid, nt);
  // -ky^2 psi (field 3) feeds only the stresses on the force box's rows
  // [r0_3, r1_3]; its other rows are not stored
  const int ya = fsel == 3 ? r0_3 : 0, yb = fsel == 3 ? r1_3 : ny - 1;
  for (int e = tid; e < ny * TX; e += nt) {
    const int y = e / TX, cc = e - y * TX;
    if (y >= ya && y <= yb) out[(size_t)y * H + kx0 + cc] = lds[pidx(y) * TX + cc];
  }
}

// ============================================================================
// K2b (iSponge = 10 only): the outflow sponge of cal_nlk NavierStokes.f90:
// 103-121.  Row y of Cpy = Fy^-1[i ky psi] gives du1/dx = Fx^-1[i kx Cpy];
// w = -chi_s du1/dx; e(ix) = sum_{is=ixmin}^{ix-1} (w(is)+w(is+1))/2 dx for
// ixmin <= ix <= ixmax (summed by one thread in the reference's order),
// e = 0 elsewhere; then Rpx += Fx[e] (Rpx holds Fx[-chi (u1 - u_s)]).
// ============================================================================
template <class FX>
__global__ void __launch_bounds__(FX::LB) k_rows_sponge10(const double2 *__restrict__ Cpy,
                                                          const double *__restrict__ ms,
                                                          double2 *__restrict__ Rpx, int ixmin, int ixmax,
                                                          FX F, const double2 *__restrict__ twx, StepConsts c) {
  extern __shared__ double2 lds[];
  const int H = F.H(), nx = 2 * H, tid = threadIdx.x, nt = blockDim.x, PL = padlen(H);
  const int y = blockIdx.x;
  double2 *f0 = lds, *f1 = lds + PL;
  for (int k = tid; k < H; k += nt) {
    const double2 p = Cpy[(size_t)y * H + k];
    const double kf = (double)k * c.sx;
    f0[pidx(k)] = make_double2(-kf * p.y, kf * p.x);
    f1[pidx(k)] = make_double2(0.0, 0.0);
  }
  __syncthreads();
  rows_c2r_pre(lds, 1, H, PL, twx, tid, nt);
  F.template fft<true, 1>(lds, twx, tid, nt);
  for (int n = tid; n < H; n += nt) {
#pragma clang fp contract(off)
    const double2 d = f0[pidx(n)];
    const double2 m = reinterpret_cast<const double2 *>(ms + (size_t)y * nx)[n];
    f0[pidx(n)] = make_double2(-(m.x * d.x), -(m.y * d.y));
  }
  __syncthreads();
  if (tid == 0) {
#pragma clang fp contract(off)
    const double *w = reinterpret_cast<const double *>(f0);
    double *e = reinterpret_cast<double *>(f1);
    double summe = 0.0;
    for (int ix = ixmin; ix <= ixmax; ++ix) {
      e[2 * pidx(ix >> 1) + (ix & 1)] = summe;
      const int i1 = ix + 1;
      summe = summe + 0.5 * (w[2 * pidx(ix >> 1) + (ix & 1)] + w[2 * pidx(i1 >> 1) + (i1 & 1)]) * c.dx;
    }
  }
  __syncthreads();
  for (int n = tid; n < H; n += nt) f0[pidx(n)] = f1[pidx(n)];
  __syncthreads();
  F.template fft<false, 1>(lds, twx, tid, nt);
  rows_r2c_post(lds, 1, H, PL, twx, tid, nt);
  for (int k = tid; k < H; k += nt) {
    const double2 a = Rpx[(size_t)y * H + k], b = f0[pidx(k)];
    Rpx[(size_t)y * H + k] = make_double2(a.x + b.x, a.y + b.y);
  }
}

// ============================================================================
// K3: time step and AB2 weights (cal_nlk NavierStokes.f90:59-78,
// EvolveFluidAB2 :488-489, vis cal_vis.f90:20-21); run by block 0 of K2,
// whose inputs (the previous step's CFL maxima) are complete at launch
// ============================================================================
// Block-wide (any multiple of 64 threads); red holds >= blockDim/64 doubles.
__device__ void step_scalars(const double *__restrict__ umax_part, int nparts, DevScalars *__restrict__ s,
                             const FlowParams &f, double time, double dt0, int explicit_step, double *red,
                             StepOutDev *__restrict__ res) {
#pragma clang fp contract(off)
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
    // the host reads dt1 as soon as this kernel has finished
    res->dt1 = dt1;
    res->u_max = u_max;
    __threadfence_system();
  }
  __syncthreads();
}

// ============================================================================
// K2: nonlinear term, penalisation and sponge in physical space (one row/WG)
// cal_nlk, NavierStokes.f90:39-52 and :88-164
// ============================================================================
template <class FX>
__device__ __forceinline__ void rows_nl_body(const double2 *__restrict__ A, const double2 *__restrict__ B,
                                                    const double *__restrict__ u1, const double *__restrict__ u2,
                                                    const double *__restrict__ ms, MaskBox mb, int penal_on,
                                                    double2 *__restrict__ Rc, double2 *__restrict__ Rpx, FX F,
                                                    const double2 *__restrict__ twx, StepConsts c,
                                                    const double *__restrict__ umax_part, int nparts,
                                                    DevScalars *__restrict__ sc, FlowParams fl, double time,
                                                    double dt0, int explicit_step, StepOutDev *__restrict__ res, int y,
                                                    bool scal, double *__restrict__ rowpen, int rpx_zero) {
  extern __shared__ double2 lds[];
  __shared__ double red[16];
  if (scal) step_scalars(umax_part, nparts, sc, fl, time, dt0, explicit_step, red, res);
  const int H = F.H(), nx = 2 * H, tid = threadIdx.x, nt = blockDim.x, PL = padlen(H);
  double2 *f0 = lds, *f1 = lds + PL, *f2 = lds + 2 * PL;
  const bool prow = penal_on && mb.x1 >= mb.x0 && y >= mb.y0 && y <= mb.y1;
  double ps1 = 0.0, ps2 = 0.0;  // the row's sums of the penalisation term (rowpen)
  // rows: w, dw/dy, dw/dx = i kx w
  stage_rows_c2r<3>(lds, 3, H, PL, twx, tid, nt, [&](int f, int k) -> double2 {
    if (f == 1) return B[(size_t)y * H + k];
    const double2 a = A[(size_t)y * H + k];
    if (f == 0) return a;
    const double kf = (double)k * c.sx;
    return make_double2(-kf * a.y, kf * a.x);
  });
  F.template fft<true, 3>(lds, twx, tid, nt);
  for (int n = tid; n < H; n += nt) {
    const int q = pidx(n);
    const double2 w = f0[q], wy = f1[q], wx = f2[q];
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
        const size_t qq = (size_t)(y - mb.y0) * mb.w + (x - mb.x0);
        px.x = -(mb.m[qq] * (a1.x - mb.vx[qq]));
        py.x = -(mb.m[qq] * (a2.x - mb.vy[qq]));
      }
      if (mb.inside(x + 1, y)) {
        const size_t qq = (size_t)(y - mb.y0) * mb.w + (x + 1 - mb.x0);
        px.y = -(mb.m[qq] * (a1.y - mb.vx[qq]));
        py.y = -(mb.m[qq] * (a2.y - mb.vy[qq]));
      }
      ps1 += px.x + px.y;
      ps2 += py.x + py.y;
    }
    f0[q] = r;
    f1[q] = px;
    f2[q] = py;
  }
  if (rowpen) {  // f_mean of cal_nlk, NavierStokes.f90:97-100: per-row sums
    __shared__ double rp[32];
    ps1 = wave_sum(ps1);
    ps2 = wave_sum(ps2);
    if ((tid & 63) == 0) { rp[2 * (tid >> 6)] = ps1; rp[2 * (tid >> 6) + 1] = ps2; }
    __syncthreads();
    if (tid == 0) {
      double a = 0.0, b = 0.0;
      for (int w = 0; w < nt / 64; ++w) { a += rp[2 * w]; b += rp[2 * w + 1]; }
      rowpen[2 * y] = a;
      rowpen[2 * y + 1] = b;
    }
  }
  __syncthreads();
  if (prow) {
    F.template fft<false, 3>(lds, twx, tid, nt);
    rows_r2c_store<3>(lds, H, PL, twx, tid, nt, [&](int k, const double2 *x) {
      const double2 xa = x[0], xpx = x[1], xpy = x[2];
      const double kf = (double)k * c.sx;
      Rc[(size_t)y * H + k] = make_double2(xa.x - kf * xpy.y, xa.y + kf * xpy.x);
      Rpx[(size_t)y * H + k] = xpx;
    });
  } else {
    // no mask on this row: px = py = 0, whose transforms are 0, so only the
    // first field is transformed (a sequence rounds alike in any batch) and
    // Rc takes the same arithmetic with the zero spectra
    F.template fft<false, 1>(lds, twx, tid, nt);
    rows_r2c_store<1>(lds, H, PL, twx, tid, nt, [&](int k, const double2 *x) {
      const double2 xa = x[0];
      const double kf = (double)k * c.sx;
      Rc[(size_t)y * H + k] = make_double2(xa.x - kf * 0.0, xa.y + kf * 0.0);
      // Rpx off the mask rows is not read (K4 reads the mask rows only), so
      // its zeros are written only for the iSponge = 10 term, which adds
      // to every row (rpx_zero); null: no penalisation anywhere
      if (Rpx && rpx_zero) Rpx[(size_t)y * H + k] = make_double2(0.0, 0.0);
    });
  }
}

// K2, penalisation part of a row whose convective part went ahead: the
// ahead pass (rows_nl_body with an empty mask box) left Rc = Fx[r] (with
// the zero spectra of a row without mask, whose arithmetic leaves it
// unchanged) and Rpx = 0 on every row; on the mask rows this adds px, py =
// -chi (u - u_s) (cal_nlk NavierStokes.f90:88-96), their transforms and the
// row sums of the penalisation term, with the arithmetic of rows_nl_body's
// mask rows: Rc = Fx[r] - i kx Fx[py] (as (re - kx im, im + kx re)),
// Rpx = Fx[px].  Launched with rows_nl_body's thread count, so the row
// sums add in the same order.
template <class FX>
__device__ __forceinline__ void rows_pen_body(const double *__restrict__ u1, const double *__restrict__ u2, MaskBox mb,
                                              double2 *__restrict__ Rc, double2 *__restrict__ Rpx, FX F,
                                              const double2 *__restrict__ twx, StepConsts c, int y,
                                              double *__restrict__ rowpen) {
  extern __shared__ double2 lds[];
  const int H = F.H(), nx = 2 * H, tid = threadIdx.x, nt = blockDim.x, PL = padlen(H);
  double2 *f0 = lds, *f1 = lds + PL;
  double ps1 = 0.0, ps2 = 0.0;
  for (int n = tid; n < H; n += nt) {
    const int q = pidx(n);
    const double2 a1 = reinterpret_cast<const double2 *>(u1 + (size_t)y * nx)[n];
    const double2 a2 = reinterpret_cast<const double2 *>(u2 + (size_t)y * nx)[n];
    double2 px = make_double2(0.0, 0.0), py = make_double2(0.0, 0.0);
    const int x = 2 * n;
    if (mb.inside(x, y)) {
      const size_t qq = (size_t)(y - mb.y0) * mb.w + (x - mb.x0);
      px.x = -(mb.m[qq] * (a1.x - mb.vx[qq]));
      py.x = -(mb.m[qq] * (a2.x - mb.vy[qq]));
    }
    if (mb.inside(x + 1, y)) {
      const size_t qq = (size_t)(y - mb.y0) * mb.w + (x + 1 - mb.x0);
      px.y = -(mb.m[qq] * (a1.y - mb.vx[qq]));
      py.y = -(mb.m[qq] * (a2.y - mb.vy[qq]));
    }
    ps1 += px.x + px.y;
    ps2 += py.x + py.y;
    f0[q] = px;
    f1[q] = py;
  }
  if (rowpen) {  // f_mean of cal_nlk, NavierStokes.f90:97-100: per-row sums
    __shared__ double rp[32];
    ps1 = wave_sum(ps1);
    ps2 = wave_sum(ps2);
    if ((tid & 63) == 0) { rp[2 * (tid >> 6)] = ps1; rp[2 * (tid >> 6) + 1] = ps2; }
    __syncthreads();
    if (tid == 0) {
      double a = 0.0, b = 0.0;
      for (int w = 0; w < nt / 64; ++w) { a += rp[2 * w]; b += rp[2 * w + 1]; }
      rowpen[2 * y] = a;
      rowpen[2 * y + 1] = b;
    }
  }
  __syncthreads();
  F.template fft<false, 2>(lds, twx, tid, nt);
  rows_r2c_store<2>(lds, H, PL, twx, tid, nt, [&](int k, const double2 *x) {
    const double2 xpx = x[0], xpy = x[1];
    const double kf = (double)k * c.sx;
    const double2 xa = Rc[(size_t)y * H + k];
    Rc[(size_t)y * H + k] = make_double2(xa.x - kf * xpy.y, xa.y + kf * xpy.x);
    Rpx[(size_t)y * H + k] = xpx;
  });
}

// ============================================================================
// K4: nlk(n0) and the integrating-factor AB2 / Euler update of vortk
// (EvolveFluidAB2 NavierStokes.f90:478-498, EvolveFluidExplicit :444-449)
// ============================================================================
template <class FY>
__device__ __forceinline__ void cols_ab2_body(const double2 *__restrict__ Rc, const double2 *__restrict__ Rpx,
                                                     const double2 *__restrict__ W0, const double2 *__restrict__ NL1,
                                                     double2 *__restrict__ NL0, double2 *__restrict__ W1,
                                                     const DevScalars *__restrict__ s, int explicit_step, FY F,
                                                     const double2 *__restrict__ twy, StepConsts c, int py0, int py1,
                                                     int bid, int nblk) {
  extern __shared__ double2 lds[];
  const int ny = F.NY(), TX = F.tx(), H = c.H, tid = threadIdx.x, nt = blockDim.x;
  const int kx0 = xcd_remap(bid, nblk) * TX;
  // Rpx null: no penalisation anywhere in the simulation, its spectrum is 0,
  // and the tile holds Rc alone (half the LDS, so that 4096-point columns
  // fit two columns per tile).  Rows outside [py0, py1] of Rpx are 0 (the
  // mask box's rows hold the penalisation; K2 writes 0 elsewhere): they are
  // not read
  const bool px_on = Rpx != nullptr;
  const int TT = px_on ? 2 * TX : TX;
  for (int e = tid; e < ny * TX; e += nt) {
    const int y = e / TX, cc = e - y * TX;
    lds[pidx(y) * TT + cc] = Rc[(size_t)y * H + kx0 + cc];
    if (px_on)
      lds[pidx(y) * TT + TX + cc] = (y >= py0 && y <= py1) ? Rpx[(size_t)y * H + kx0 + cc] : make_double2(0.0, 0.0);
  }
  __syncthreads();
  if (px_on) F.template fft<false, 2, 2>(lds, twy, tid, nt);
  else F.template fft<false, 1, 1>(lds, twy, tid, nt);
  const double b10 = s->b10, b11 = s->b11, cx1 = s->cx1, cy1 = s->cy1, cx0 = s->cx0, cy0 = s->cy0;
  const int hy = ny >> 1;
  for (int e = tid; e < ny * TX; e += nt) {
    const int idx = e / TX, cc = e - idx * TX, kx = kx0 + cc;
    const size_t g = (size_t)idx * H + kx;
    double2 nl = make_double2(0.0, 0.0), w1 = make_double2(0.0, 0.0);
    if (idx != hy) {
      const int ky = ky_signed(idx, ny);
      const int aky = ky < 0 ? -ky : ky;
      const double kyf = (double)ky * c.sy;
      const double2 fc = lds[pidx(idx) * TT + cc];
      const double2 fp = px_on ? lds[pidx(idx) * TT + TX + cc] : make_double2(0.0, 0.0);
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
//  10 sum chi vy (the first beam's cal_drag window)   11..14 the same four
//  sums over the second beam's window (win2, empty with one beam; its
//  reductions are then skipped and the sums stay 0)
// ============================================================================
constexpr int kNSUM = 15;
constexpr int kNSUM1 = 11;  // the sums of a run with at most one beam
struct Window { int x0, x1, y0, y1; };

// CMODE, CPRES: mode and pres fixed at compile time (-1: the run-time
// arguments decide), so that the step's instance carries no dead branches
template <class FX, int CMODE = -1, int CPRES = -1>
__device__ __forceinline__ void rows_press_body(const double2 *__restrict__ Cw, const double2 *__restrict__ Cpsi,
                                                    const double2 *__restrict__ Cpsiy, double *__restrict__ u1,
                                                    double *__restrict__ u2, double *__restrict__ vor_mean,
                                                    MaskBox mb, Window win, Window win2, double *__restrict__ sums,
                                                    double *__restrict__ umax_part, double2 *__restrict__ Rq,
                                                    double2 *__restrict__ Rqy, const DevScalars *__restrict__ s,
                                                    double UMx0, double UMy0, int mode, int pres, FX F,
                                                    const double2 *__restrict__ twx, StepConsts c, int bid, int nblk) {
  extern __shared__ double2 lds[];
  __shared__ double red[(kNSUM + 1) * 16];
  if constexpr (CMODE >= 0) mode = CMODE;
  if constexpr (CPRES >= 0) pres = CPRES;
  const int H = F.H(), nx = 2 * H, tid = threadIdx.x, nt = blockDim.x, PL = padlen(H);
  const int y = bid;
  double2 *f0 = lds, *f1 = lds + PL, *f2 = lds + 2 * PL;
  const int nf = mode ? 3 : 2;
  double2 *g1 = mode ? f1 : f0, *g2 = mode ? f2 : f1;  // psi_x, psi_y slots
  // rows: (w,) dpsi/dx = i kx psi, dpsi/dy
  stage_rows_c2r<3>(lds, nf, H, PL, twx, tid, nt, [&](int f, int k) -> double2 {
    const int g = mode ? f : f + 1;
    if (g == 0) return Cw[(size_t)y * H + k];
    if (g == 2) return Cpsiy[(size_t)y * H + k];
    const double kf = (double)k * c.sx;
    const double2 p = Cpsi[(size_t)y * H + k];
    return make_double2(-kf * p.y, kf * p.x);
  });
  if (mode) F.template fft<true, 3>(lds, twx, tid, nt);
  else F.template fft<true, 2>(lds, twx, tid, nt);
  double UMx = UMx0, UMy = UMy0, Mx = 0.0, My = 0.0, time = 0.0, dt1 = 0.0;
  if (mode) {
    UMx = s->UMx; UMy = s->UMy; Mx = s->Mx; My = s->My; time = s->time; dt1 = s->dt1;
  }
  const bool mrow = mb.x1 >= mb.x0 && y >= mb.y0 && y <= mb.y1;
  const bool wrow = y >= win.y0 && y <= win.y1;
  const bool two = win2.x1 >= win2.x0;
  const bool wrow2 = two && y >= win2.y0 && y <= win2.y1;
  double acc[kNSUM];
#pragma unroll
  for (int i = 0; i < kNSUM; ++i) acc[i] = 0.0;
  double um = 0.0;
  for (int n = tid; n < H; n += nt) {
    const int qn = pidx(n);
    const double2 gx = g1[qn], uy1 = g2[qn];
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
    const double2 w = f0[qn];
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
        if (wrow2 && x + h >= win2.x0 && x + h <= win2.x1) {
          acc[11] += p1;
          acc[12] += p2;
          acc[13] += vx * m;
          acc[14] += vy * m;
        }
        if (h) { q1.y = p1; q2.y = p2; } else { q1.x = p1; q2.x = p2; }
      }
    }
    if (pres) {
      // penal(:,:,1) - vor*uy, penal(:,:,2) + vor*ux   (:357-358)
      f0[qn] = make_double2(q1.x - w.x * a2.x, q1.y - w.y * a2.y);
      f1[qn] = make_double2(q2.x + w.x * a1.x, q2.y + w.y * a1.y);
    }
  }
  // the u_max partial (next step's CFL) and the row sums in one pass: the
  // waves' values go to LDS, one barrier (which also orders the physics
  // loop's LDS writes before the forward transforms), then thread 0 adds
  // them wave by wave, as block_sum does.  Rows outside the mask box reduce
  // only the two energy sums (the mask sums 2..10 are exactly 0), rows of
  // one window the first kNSUM1 (11..14 are exactly 0).
  {
    const int lane = tid & 63, wid = tid >> 6, nw = nt >> 6;
    um = wave_max(um);
    if (lane == 0) red[wid] = um;
    const int nv = !mode ? 0 : !mrow ? 2 : two ? kNSUM : kNSUM1;
    if (nv == kNSUM) wave_sums<kNSUM>(acc, red + nw, lane, wid, nw);
    else if (nv == kNSUM1) wave_sums<kNSUM1>(acc, red + nw, lane, wid, nw);
    else if (nv == 2) wave_sums<2>(acc, red + nw, lane, wid, nw);
    __syncthreads();
    if (tid == 0) {
      double m = 0.0;
      for (int w = 0; w < nw; ++w) m = fmax(m, red[w]);
      umax_part[y] = m;
      if (mode)
        for (int i = 0; i < kNSUM; ++i) {
          double q = 0.0;
          if (i < nv)
            for (int w = 0; w < nw; ++w) q += red[nw + i * nw + w];
          sums[(size_t)y * kNSUM + i] = q;
        }
    }
  }
  if (!mode) return;
  if (!pres) return;
  F.template fft<false, 2>(lds, twx, tid, nt);
  rows_r2c_store<2>(lds, H, PL, twx, tid, nt, [&](int k, const double2 *x) {
    const double kf = (double)k * c.sx;
    Rq[(size_t)y * H + k] = make_double2(-kf * x[0].y, kf * x[0].x);
    Rqy[(size_t)y * H + k] = x[1];
  });
}


// ============================================================================
// K7: divergence + poisson (pressure NavierStokes.f90:366-382), then the
// inverse column transform; rows [r0, r1] are written.
// ============================================================================
template <class FY>
__device__ __forceinline__ void cols_press_body(const double2 *__restrict__ Rq, const double2 *__restrict__ Rqy,
                                                       double2 *__restrict__ Cq, int r0, int r1, FY F,
                                                       const double2 *__restrict__ twy, StepConsts c, int bid, int nblk) {
  extern __shared__ double2 lds[];
  const int ny = F.NY(), TX = F.tx(), H = c.H, tid = threadIdx.x, nt = blockDim.x;
  const int kx0 = xcd_remap(bid, nblk) * TX;
  const int TT = 2 * TX;
  for (int e = tid; e < ny * TX; e += nt) {
    const int y = e / TX, cc = e - y * TX;
    lds[pidx(y) * TT + cc] = Rq[(size_t)y * H + kx0 + cc];
    lds[pidx(y) * TT + TX + cc] = Rqy[(size_t)y * H + kx0 + cc];
  }
  __syncthreads();
  F.template fft<false, 2, 2>(lds, twy, tid, nt);
  const int hy = ny >> 1;
  for (int e = tid; e < ny * TX; e += nt) {
    const int idx = e / TX, cc = e - idx * TX, kx = kx0 + cc;
    const int ky = ky_signed(idx, ny);
    const int aky = ky < 0 ? -ky : ky;
    double2 q = make_double2(0.0, 0.0);
    if (idx != hy && !(kx == 0 && ky == 0)) {
      const double kyf = (double)ky * c.sy;
      const double2 a = lds[pidx(idx) * TT + cc], b = lds[pidx(idx) * TT + TX + cc];
      const double2 fa = make_double2(a.x * c.norm, a.y * c.norm);
      const double2 fb = make_double2(b.x * c.norm, b.y * c.norm);
      const double2 rhs = make_double2(fa.x - kyf * fb.y, fa.y + kyf * fb.x);
      const double quot = ((double)(kx * kx) * c.scalex + (double)(aky * aky) * c.scaley);
      q = make_double2(rhs.x / quot, rhs.y / quot);
    }
    // this thread is the only reader and writer of slot (idx, cc)
    lds[pidx(idx) * TT + cc] = q;
  }
  __syncthreads();
  F.template fft<true, 1, 2>(lds, twy, tid, nt);
  for (int e = tid; e < ny * TX; e += nt) {
    const int y = e / TX, cc = e - y * TX;
    if (y >= r0 && y <= r1) Cq[(size_t)y * H + kx0 + cc] = lds[pidx(y) * TT + cc];
  }
}

// ============================================================================
// K8: static pressure p = q - |u|^2/2 (:390) and the viscous stresses
// stress_a = 2 nu d_x u1, stress_b = nu (d_y u1 + d_x u2) (BeamForces.f90:
// 105-117) on rows r0+blockIdx.x, written for columns [bx0, bx1] with pitch bw.
// Spectrally d_x u1 = i kx (i ky psi), d_y u1 + d_x u2 = (kx^2 - ky^2) psi.
// ============================================================================
template <class FX>
__device__ __forceinline__ void rows_box_body(const double2 *__restrict__ Cq, const double2 *__restrict__ Cpsi,
                                          const double2 *__restrict__ Cpsiy, const double2 *__restrict__ Cpsiyy,
                                          const double *__restrict__ u1, const double *__restrict__ u2,
                                          double *__restrict__ press, double *__restrict__ sa,
                                          double *__restrict__ sb, int r0, int bx0, int bx1, int bw, int visc,
                                          FX F, const double2 *__restrict__ twx, StepConsts c, int bid, int nblk) {
  extern __shared__ double2 lds[];
  const int H = F.H(), nx = 2 * H, tid = threadIdx.x, nt = blockDim.x, PL = padlen(H);
  const int y = r0 + bid;
  double2 *f0 = lds, *f1 = lds + PL, *f2 = lds + 2 * PL;
  const int nf = visc ? 3 : 1;
  // rows: q, and with viscosity d2psi/dxdy = i kx dpsi/dy and
  // d2psi/dy2 - d2psi/dx2 = kx^2 psi + d2psi/dy2
  stage_rows_c2r<3>(lds, nf, H, PL, twx, tid, nt, [&](int f, int k) -> double2 {
    if (f == 0) return Cq[(size_t)y * H + k];
    const double kf = (double)k * c.sx;
    if (f == 1) {
      const double2 py = Cpsiy[(size_t)y * H + k];
      return make_double2(-kf * py.y, kf * py.x);
    }
    const double2 p = Cpsi[(size_t)y * H + k], pyy = Cpsiyy[(size_t)y * H + k];
    return make_double2((kf * kf) * p.x + pyy.x, (kf * kf) * p.y + pyy.y);
  });
  if (visc) F.template fft<true, 3>(lds, twx, tid, nt);
  else F.template fft<true, 1>(lds, twx, tid, nt);
  const size_t row = (size_t)bid * bw;
  for (int x = bx0 + tid; x <= bx1; x += nt) {
#pragma clang fp contract(off)
    const int q = pidx(x >> 1), h = x & 1;
    const double qv = h ? f0[q].y : f0[q].x;
    const double v1 = u1[(size_t)y * nx + x], v2 = u2[(size_t)y * nx + x];
    press[row + (x - bx0)] = qv - 0.5 * (v1 * v1 + v2 * v2);
    if (visc) {
      sa[row + (x - bx0)] = 2.0 * c.nu * (h ? f1[q].y : f1[q].x);
      sb[row + (x - bx0)] = c.nu * (h ? f2[q].y : f2[q].x);
    }
  }
}

// ============================================================================
// K10: GetForcesInterp with SPL=3 (BeamForces.f90:150-270) and the surface
// integral (:306-489).  Interpolation targets are prepared on the host.
// ============================================================================
struct ForceGeom {
  int ns, nc, nt;           // beam nodes, circle points (per beam), targets (all beams)
  int bx0, by0, bw, bh;     // box of the press/stress arrays
  double dxi, dyi, pi;      // DeltaInterpolation spacing (Interpolation.f90:15-16)
  double ds, dL, soft;
  int visc, bsi_on;
  int nb, nx, ny, pad;      // beams; grid (the stencil is kept inside it)
  double rr, r1, dphi;      // target offsets t+N max(dx,dy) and t (BeamForces.f90:150-151), pi/(nc-1)
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

// targets: tg[3*t] = x, tg[3*t+1] = y, tg[3*t+2] = field id (0 press, 1 sa, 2 sb);
// beam b's targets start at b (6ns + 6nc): there node n owns targets
// 6n .. 6n+5 (P(u1) P(b1) A(u) B(u) A(b) B(b)), circle point i (trailing
// i < nc, leading nc <= i < 2nc) owns 6ns + 3i + {0 A, 1 B, 2 P}.
// node data nd[4 (b ns + n) + ...]: sin th, cos th, sin 2th, cos 2th; circle
// data cd[4 (2 b nc + i) + ...]: cos phi, sin phi, cos 2phi, sin 2phi;
// loads pbeam/tbeam[b ns + n].
//
// One block of 64 threads per unit: a beam node (6 targets) or a pair of
// circle points (6 targets); beam b owns units b (ns + nc) .. (b+1)(ns + nc) - 1.  Thread (t, a) evaluates the kernel weights
// wx[a], wy[a] and the 9 products of stencil column a (DeltaInterpolation,
// Interpolation.f90:26-44); one thread per target adds the 81 products in
// the reference's order (ix outer, iy inner).  Thread 0 then forms the
// unit's loads and its terms of the eight force sums (GetForcesInterp
// BeamForces.f90:204-267, BeamSurfaceIntegral :354-480) and writes them to
// part[8 unit + k] in the pinned result record; the host adds them in unit
// order.
constexpr int kForceThreads = 64;

__device__ __forceinline__ void interp6(const double *__restrict__ tg, int t0, int ntu,
                                        const double *__restrict__ press, const double *__restrict__ sa,
                                        const double *__restrict__ sb, const ForceGeom &g, double *vals,
                                        double *__restrict__ val) {
#pragma clang fp contract(off)
  __shared__ double wxs[6][9], wys[6][9], prod[6][81];
  const int lt = threadIdx.x / 9, a = threadIdx.x - 9 * lt;
  const bool act = lt < ntu;
  int i = 0, j = 0;
  const double *f = press;
  if (act) {
    const int t = t0 + lt;
    const double xt = tg[3 * t], yt = tg[3 * t + 1];
    const int fid = (int)tg[3 * t + 2];
    f = fid == 0 ? press : fid == 1 ? sa : sb;
    i = (int)((xt - 0.0) / g.dxi);
    j = (int)((yt - 0.0) / g.dyi);
    // the host's box holds every stencil with a cell to spare; the clamp
    // keeps a stencil that the device's rounding moved by a cell inside it
    // (phi4star vanishes beyond 2.5 cells, so a moved stencil adds the same
    // nonzero terms in the same order)
    i = min(max(i, g.bx0 + 4), g.bx0 + g.bw - 5);
    j = min(max(j, g.by0 + 4), g.by0 + g.bh - 5);
    wxs[lt][a] = phi4star(fabs(((double)(i - 4 + a) * g.dxi + 0.0) - xt), g.dxi, g.pi);
    wys[lt][a] = phi4star(fabs(((double)(j - 4 + a) * g.dyi + 0.0) - yt), g.dyi, g.pi);
  }
  __syncthreads();
  if (act) {
    const double *col = f + (size_t)(j - 4 - g.by0) * g.bw + (i - 4 - g.bx0 + a);
    const double wx = wxs[lt][a];
#pragma unroll
    for (int b = 0; b < 9; ++b) prod[lt][9 * a + b] = wx * wys[lt][b] * col[(size_t)b * g.bw];
  }
  __syncthreads();
  if (act && a == 0) {
    double R1 = 0.0;
    for (int q = 0; q < 81; ++q) R1 = R1 + prod[lt][q];
    vals[lt] = R1;
    val[t0 + lt] = R1;
  }
  __syncthreads();
}

__device__ __forceinline__ void forces_body(const double *__restrict__ tg, const double *__restrict__ press,
                                                          const double *__restrict__ sa, const double *__restrict__ sb,
                                                          ForceGeom g, double *__restrict__ val,
                                                          const double *__restrict__ nd, const double *__restrict__ cd,
                                                          double *__restrict__ pbeam, double *__restrict__ tbeam,
                                                          double *__restrict__ part, int bid, int nblk) {
#pragma clang fp contract(off)
  __shared__ double vals[6];
  const int ns = g.ns, per = ns + g.nc, b = bid / per, u = bid - b * per;
  const bool node = u < ns;
  const int t0 = b * 6 * per + (node ? 6 * u : 6 * ns + 6 * (u - ns));
  nd += (size_t)4 * b * ns;
  cd += (size_t)8 * b * g.nc;
  pbeam += (size_t)b * ns;
  tbeam += (size_t)b * ns;
  interp6(tg, t0, 6, press, sa, sb, g, vals, val);
  if (threadIdx.x == 0) {
    // acc: 0 fx, 1 fy, 2 Ft0, 3 Ft1, 4 Fn0, 5 Fn1, 6 Fp0, 7 Fp1
    double acc[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (node) {
      const double st = nd[4 * u], ct = nd[4 * u + 1], s2 = nd[4 * u + 2], c2 = nd[4 * u + 3];
      double pb = vals[0] - vals[1];
      double tb = 0.0;
      const double A1 = vals[2], B1 = vals[3], A2 = vals[4], B2 = vals[5];
      if (g.visc) {  // BeamForces.f90:206-215, :237
        const double sn1 = -(A1 * c2) - B1 * s2, st1 = -(A1 * s2) + B1 * c2;
        const double sn2 = -(A2 * c2) - B2 * s2, st2 = -(A2 * s2) + B2 * c2;
        tb = st1 - st2;
        pb = pb + (-sn1 + sn2);
      }
      pb = pb * g.soft;  // :254-255
      tb = tb * g.soft;
      pbeam[u] = pb;
      tbeam[u] = tb;
      acc[0] = -((st * pb + ct * tb) * g.ds);  // :265-266
      acc[1] = (ct * pb + st * tb) * g.ds;
      if (g.bsi_on) {
        {  // top side, :433-452
          const double A = A1, B = B1, P = vals[0];
          const double sn = -(A * c2) - B * s2, stt = -(A * s2) + B * c2, spp = -P;
          acc[2] += g.ds * stt * (+ct); acc[3] += g.ds * stt * (+st);
          acc[4] += g.ds * sn * (-st);  acc[5] += g.ds * sn * (+ct);
          acc[6] += g.ds * spp * (-st); acc[7] += g.ds * spp * (+ct);
        }
        {  // bottom side, :462-480
          const double A = A2, B = B2, P = vals[1];
          const double sn = -(A * c2) - B * s2, stt = -(A * s2) + B * c2, spp = -P;
          acc[2] += g.ds * stt * (-ct); acc[3] += g.ds * stt * (-st);
          acc[4] += g.ds * sn * (+st);  acc[5] += g.ds * sn * (-ct);
          acc[6] += g.ds * spp * (+st); acc[7] += g.ds * spp * (-ct);
        }
      }
    } else {
      for (int k = 0; k < 2; ++k) {  // half circles, :354-420
        const int i = 2 * (u - ns) + k;
        const double cp = cd[4 * i], sp = cd[4 * i + 1], c2 = cd[4 * i + 2], s2 = cd[4 * i + 3];
        const double A = vals[3 * k], B = vals[3 * k + 1], P = vals[3 * k + 2];
        const double sn = +A * c2 + B * s2, stt = -(A * s2) + B * c2, spp = -P;
        acc[2] += g.dL * stt * (-sp); acc[3] += g.dL * stt * (+cp);
        acc[4] += g.dL * sn * (+cp);  acc[5] += g.dL * sn * (+sp);
        acc[6] += g.dL * spp * (+cp); acc[7] += g.dL * spp * (+sp);
      }
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) part[8 * bid + k] = acc[k];
  }
}

// ============================================================================
// Entry points.  k_*: one simulation, arguments by value (grid.x = rows or
// tiles).  kb_*: a batch of simulations of one grid size, one argument
// record per simulation at blockIdx.y; grid.x is the largest per-simulation
// block count and blocks beyond a simulation's own count return at once.
// ============================================================================
template <class FY>
__global__ void __launch_bounds__(FY::LB) k_cols_inv_fields(const double2 *__restrict__ W, double2 *__restrict__ o0,
                                                            double2 *__restrict__ o1, double2 *__restrict__ o2,
                                                            double2 *__restrict__ o3, int fsel0, int fsel1,
                                                            int fsel2, int fsel3, int nfield, FY F,
                                                            const double2 *__restrict__ twy, StepConsts c, int r0_3,
                                                            int r1_3) {
  cols_inv_fields_body(W, o0, o1, o2, o3, fsel0, fsel1, fsel2, fsel3, nfield, F, twy, c, r0_3, r1_3, blockIdx.x,
                       gridDim.x);
}
template <class FX>
__global__ void __launch_bounds__(FX::LB) k_rows_nl(const double2 *__restrict__ A, const double2 *__restrict__ B,
                                                    const double *__restrict__ u1, const double *__restrict__ u2,
                                                    const double *__restrict__ ms, MaskBox mb, int penal_on,
                                                    double2 *__restrict__ Rc, double2 *__restrict__ Rpx, FX F,
                                                    const double2 *__restrict__ twx, StepConsts c,
                                                    const double *__restrict__ umax_part, int nparts,
                                                    DevScalars *__restrict__ sc, FlowParams fl, double time,
                                                    double dt0, int explicit_step, StepOutDev *__restrict__ res,
                                                    int ra, int na, int rb, int scal, double *__restrict__ rowpen,
                                                    int rpx_zero) {
  // rows ra .. ra+na-1, then rb onwards (two row ranges; one when rb is unused)
  const int b = blockIdx.x;
  rows_nl_body(A, B, u1, u2, ms, mb, penal_on, Rc, Rpx, F, twx, c, umax_part, nparts, sc, fl, time, dt0,
               explicit_step, res, b < na ? ra + b : rb + (b - na), scal && b == 0, rowpen, rpx_zero);
}

// K2's penalisation rows y0 .. y0 + nrow - 1 after the ahead pass; block
// nrow (the last) computes the K3 scalars
template <class FX>
__global__ void __launch_bounds__(FX::LB) k_rows_pen(const double *__restrict__ u1, const double *__restrict__ u2,
                                                     MaskBox mb, double2 *__restrict__ Rc, double2 *__restrict__ Rpx,
                                                     FX F, const double2 *__restrict__ twx, StepConsts c, int y0,
                                                     int nrow, double *__restrict__ rowpen,
                                                     const double *__restrict__ umax_part, int nparts,
                                                     DevScalars *__restrict__ sc, FlowParams fl, double time,
                                                     double dt0, int explicit_step, StepOutDev *__restrict__ res) {
  if ((int)blockIdx.x >= nrow) {
    __shared__ double red[16];
    step_scalars(umax_part, nparts, sc, fl, time, dt0, explicit_step, red, res);
    return;
  }
  rows_pen_body(u1, u2, mb, Rc, Rpx, F, twx, c, y0 + blockIdx.x, rowpen);
}

// iMeanVelocity = 11 (zero mean pressure): f_mean = sum(penal) / (nx ny)
// from the row sums (cal_nlk, NavierStokes.f90:97-100) into fmean(:,n0),
// then u_mean advanced as EvolveFluidExplicit (:453-456) or EvolveFluidAB2
// (:501-503) do, and the step's mean velocity Mean_u(time+dt1) = u_mean set
// for the velocity kernel.  One block of 256 threads.
__global__ void __launch_bounds__(256) k_mean11(const double *__restrict__ rowpen, int ny, double inv_n,
                                                DevScalars *__restrict__ s, int n0, int n1, int explicit_step,
                                                double U_mean_true) {
#pragma clang fp contract(off)
  __shared__ double rp[8];
  double a = 0.0, b = 0.0;
  if (rowpen)
    for (int y = threadIdx.x; y < ny; y += blockDim.x) { a += rowpen[2 * y]; b += rowpen[2 * y + 1]; }
  a = wave_sum(a);
  b = wave_sum(b);
  if ((threadIdx.x & 63) == 0) { rp[2 * (threadIdx.x >> 6)] = a; rp[2 * (threadIdx.x >> 6) + 1] = b; }
  __syncthreads();
  if (threadIdx.x != 0) return;
  if (rowpen) {  // penalisation off: cal_nlk leaves f_mean as it was
    s->fmean[n0][0] = ((rp[0] + rp[2]) + (rp[4] + rp[6])) * inv_n;
    s->fmean[n0][1] = ((rp[1] + rp[3]) + (rp[5] + rp[7])) * inv_n;
  }
  double u0 = s->u_mean[0], u1 = s->u_mean[1];
  if (explicit_step) {
    u0 = u0 + s->fmean[n0][0] * s->dt1;
    u1 = u1 + s->fmean[n0][1] * s->dt1;
  } else {
    u0 = u0 + s->b10 * s->fmean[n0][0] + s->b11 * s->fmean[n1][0];
    u1 = u1 + s->b10 * s->fmean[n0][1] + s->b11 * s->fmean[n1][1];
  }
  s->u_mean[0] = u0;
  s->u_mean[1] = u1;
  s->Mx = u0;
  s->My = u1;
  s->UMx = U_mean_true * u0;
  s->UMy = U_mean_true * u1;
}
template <class FY>
__global__ void __launch_bounds__(FY::LB) k_cols_ab2(const double2 *__restrict__ Rc, const double2 *__restrict__ Rpx,
                                                     const double2 *__restrict__ W0, const double2 *__restrict__ NL1,
                                                     double2 *__restrict__ NL0, double2 *__restrict__ W1,
                                                     const DevScalars *__restrict__ s, int explicit_step, FY F,
                                                     const double2 *__restrict__ twy, StepConsts c, int py0, int py1) {
  cols_ab2_body(Rc, Rpx, W0, NL1, NL0, W1, s, explicit_step, F, twy, c, py0, py1, blockIdx.x, gridDim.x);
}
template <class FX, int CMODE = -1, int CPRES = -1>
__global__ void __launch_bounds__(FX::LB) k_rows_press(const double2 *__restrict__ Cw, const double2 *__restrict__ Cpsi,
                                                       const double2 *__restrict__ Cpsiy, double *__restrict__ u1,
                                                       double *__restrict__ u2, double *__restrict__ vor_mean,
                                                       MaskBox mb, Window win, Window win2,
                                                       double *__restrict__ sums,
                                                       double *__restrict__ umax_part, double2 *__restrict__ Rq,
                                                       double2 *__restrict__ Rqy, const DevScalars *__restrict__ s,
                                                       double UMx0, double UMy0, int mode, int pres, FX F,
                                                       const double2 *__restrict__ twx, StepConsts c) {
  rows_press_body<FX, CMODE, CPRES>(Cw, Cpsi, Cpsiy, u1, u2, vor_mean, mb, win, win2, sums, umax_part, Rq, Rqy, s, UMx0,
                                    UMy0, mode, pres,
                  F, twx, c, blockIdx.x, gridDim.x);
}
template <class FY>
__global__ void __launch_bounds__(FY::LB) k_cols_press(const double2 *__restrict__ Rq, const double2 *__restrict__ Rqy,
                                                       double2 *__restrict__ Cq, int r0, int r1, FY F,
                                                       const double2 *__restrict__ twy, StepConsts c) {
  cols_press_body(Rq, Rqy, Cq, r0, r1, F, twy, c, blockIdx.x, gridDim.x);
}
template <class FX>
__global__ void __launch_bounds__(FX::LB) k_rows_box(const double2 *__restrict__ Cq, const double2 *__restrict__ Cpsi,
                                                     const double2 *__restrict__ Cpsiy,
                                                     const double2 *__restrict__ Cpsiyy, const double *__restrict__ u1,
                                                     const double *__restrict__ u2, double *__restrict__ press,
                                                     double *__restrict__ sa, double *__restrict__ sb, int r0, int bx0,
                                                     int bx1, int bw, int visc, FX F,
                                                     const double2 *__restrict__ twx, StepConsts c) {
  rows_box_body(Cq, Cpsi, Cpsiy, Cpsiyy, u1, u2, press, sa, sb, r0, bx0, bx1, bw, visc, F, twx, c, blockIdx.x,
                gridDim.x);
}
__global__ void __launch_bounds__(kForceThreads) k_forces(const double *__restrict__ tg, const double *__restrict__ press,
                                                          const double *__restrict__ sa, const double *__restrict__ sb,
                                                          ForceGeom g, double *__restrict__ val,
                                                          const double *__restrict__ nd, const double *__restrict__ cd,
                                                          double *__restrict__ pbeam, double *__restrict__ tbeam,
                                                          double *__restrict__ part) {
  forces_body(tg, press, sa, sb, g, val, nd, cd, pbeam, tbeam, part, blockIdx.x, gridDim.x);
}

// ---- per-simulation argument records of the batched entry points ----------
struct ColsInvArgs {
  const double2 *W;
  double2 *o[4];
  int fsel[4], nfield, nblk;
  int r0_3, r1_3;  // rows of field 3 (-ky^2 psi) that are stored
  StepConsts c;
};
struct RowsNlArgs {
  const double2 *A, *B;
  const double *u1, *u2, *ms;
  MaskBox mb;
  int penal_on, nparts, explicit_step, nblk;
  double2 *Rc, *Rpx;
  StepConsts c;
  const double *umax_part;
  DevScalars *sc;
  FlowParams fl;
  double time, dt0;
  StepOutDev *res;
  int rpx_zero;  // write Rpx's zero rows (iSponge = 10 adds to every row)
};
struct ColsAb2Args {
  const double2 *Rc, *Rpx, *W0, *NL1;
  double2 *NL0, *W1;
  const DevScalars *s;
  int explicit_step, nblk;
  int py0, py1;  // rows of Rpx that may be nonzero
  StepConsts c;
};
struct RowsPressArgs {
  const double2 *Cw, *Cpsi, *Cpsiy;
  double *u1, *u2, *vor_mean, *sums, *umax_part;
  MaskBox mb;
  Window win, win2;
  double2 *Rq, *Rqy;
  const DevScalars *s;
  double UMx0, UMy0;
  int mode, pres, nblk, pad;
  StepConsts c;
};
struct ColsPressArgs {
  const double2 *Rq, *Rqy;
  double2 *Cq;
  int r0, r1, nblk, pad;
  StepConsts c;
};
struct RowsBoxArgs {
  const double2 *Cq, *Cpsi, *Cpsiy, *Cpsiyy;
  const double *u1, *u2;
  double *press, *sa, *sb;
  int r0, bx0, bx1, bw, visc, nblk;
  StepConsts c;
};
struct ForcesArgs {
  const double *tg, *press, *sa, *sb, *nd, *cd;
  double *val, *pbeam, *tbeam, *part;
  ForceGeom g;
  int nblk, pad;
};

template <class FY>
__global__ void __launch_bounds__(FY::LB) kb_cols_inv_fields(const ColsInvArgs *__restrict__ arr, FY F,
                                                             const double2 *__restrict__ twy) {
  const ColsInvArgs &a = arr[blockIdx.y];
  if ((int)blockIdx.x >= a.nblk) return;
  cols_inv_fields_body(a.W, a.o[0], a.o[1], a.o[2], a.o[3], a.fsel[0], a.fsel[1], a.fsel[2], a.fsel[3], a.nfield, F,
                       twy, a.c, a.r0_3, a.r1_3, blockIdx.x, a.nblk);
}
template <class FX>
__global__ void __launch_bounds__(FX::LB) kb_rows_nl(const RowsNlArgs *__restrict__ arr, FX F,
                                                     const double2 *__restrict__ twx) {
  const RowsNlArgs &a = arr[blockIdx.y];
  if ((int)blockIdx.x >= a.nblk) return;
  rows_nl_body(a.A, a.B, a.u1, a.u2, a.ms, a.mb, a.penal_on, a.Rc, a.Rpx, F, twx, a.c, a.umax_part, a.nparts, a.sc,
               a.fl, a.time, a.dt0, a.explicit_step, a.res, (int)blockIdx.x, blockIdx.x == 0, nullptr, a.rpx_zero);
}
template <class FY>
__global__ void __launch_bounds__(FY::LB) kb_cols_ab2(const ColsAb2Args *__restrict__ arr, FY F,
                                                      const double2 *__restrict__ twy) {
  const ColsAb2Args &a = arr[blockIdx.y];
  if ((int)blockIdx.x >= a.nblk) return;
  cols_ab2_body(a.Rc, a.Rpx, a.W0, a.NL1, a.NL0, a.W1, a.s, a.explicit_step, F, twy, a.c, a.py0, a.py1, blockIdx.x,
                a.nblk);
}
template <class FX>
__global__ void __launch_bounds__(FX::LB) kb_rows_press(const RowsPressArgs *__restrict__ arr, FX F,
                                                        const double2 *__restrict__ twx) {
  const RowsPressArgs &a = arr[blockIdx.y];
  if ((int)blockIdx.x >= a.nblk) return;
  rows_press_body(a.Cw, a.Cpsi, a.Cpsiy, a.u1, a.u2, a.vor_mean, a.mb, a.win, a.win2, a.sums, a.umax_part, a.Rq, a.Rqy, a.s,
                  a.UMx0, a.UMy0, a.mode, a.pres, F, twx, a.c, blockIdx.x, a.nblk);
}
template <class FY>
__global__ void __launch_bounds__(FY::LB) kb_cols_press(const ColsPressArgs *__restrict__ arr, FY F,
                                                        const double2 *__restrict__ twy) {
  const ColsPressArgs &a = arr[blockIdx.y];
  if ((int)blockIdx.x >= a.nblk) return;
  cols_press_body(a.Rq, a.Rqy, a.Cq, a.r0, a.r1, F, twy, a.c, blockIdx.x, a.nblk);
}
template <class FX>
__global__ void __launch_bounds__(FX::LB) kb_rows_box(const RowsBoxArgs *__restrict__ arr, FX F,
                                                      const double2 *__restrict__ twx) {
  const RowsBoxArgs &a = arr[blockIdx.y];
  if ((int)blockIdx.x >= a.nblk) return;
  rows_box_body(a.Cq, a.Cpsi, a.Cpsiy, a.Cpsiyy, a.u1, a.u2, a.press, a.sa, a.sb, a.r0, a.bx0, a.bx1, a.bw, a.visc, F,
                twx, a.c, blockIdx.x, a.nblk);
}
__global__ void __launch_bounds__(kForceThreads) kb_forces(const ForcesArgs *__restrict__ arr) {
  const ForcesArgs &a = arr[blockIdx.y];
  if ((int)blockIdx.x >= a.nblk) return;
  forces_body(a.tg, a.press, a.sa, a.sb, a.g, a.val, a.nd, a.cd, a.pbeam, a.tbeam, a.part, blockIdx.x, a.nblk);
}

}  // namespace fsibe
