// kernels_mask.hpp -- create_mask (create_mask.f90:2-218) as one per-pixel
// kernel.  The host turns the beam into a list of primitives in the order the
// reference draws them (segments :91-93, sharp end :97-99, hinges :103-107,
// leading-edge dot :111, trailing-edge dot :115-117; the cylinder dot :67-69
// first) with the geometry that does not depend on the pixel precomputed
// (rotation, rotated end points, nint bounding boxes, hinge angles).  Each
// pixel of the mask box then applies every primitive whose box contains it,
// in order, with the reference's replacement rule (strict >, except the
// sharp end, which overwrites).  Pixels outside the box are 0 in all three
// masks: the box is the union of the primitives' boxes, and a dot changes no
// pixel at distance >= t + 7 delta (both erf terms are exactly -+1 there).
#pragma once
#include <hip/hip_runtime.h>

namespace fsibe {

enum PrimType { PRIM_SEG = 1, PRIM_HINGE = 2, PRIM_DOT = 3, PRIM_SHARP = 4 };

struct Prim {
  int type, ix0, ix1, iy0, iy1, pad0, pad1, pad2;
  double a[12];
};

struct MaskConsts {
  double dx, dy, t, h, gerf, inv_eps_unused, ds, pi;
  int sharp;
  double eps;
  // velocity sponge of a channel inflow (iSponge = 4, create_mask.f90:130-184)
  int vs4 = 0, vs_ixmid = 0, vs_ixmax = 0;
  double vs_S = 0.0, vs_ssz = 0.0, vs_Heff = 0.0, h_channel = 0.0, N = 0.0;
  // frame around the domain (iSponge = 22, :187-202): mask += frame w
  const double *frame = nullptr;
  int nx_frame = 0;
  double frame_M = 1.0;  // max(1e-3, eps)
  // Mean_u(time) of the mask's time; mean_dev (2 doubles on the device)
  // instead when iMeanVelocity = 11 (u_mean lives on the device)
  double mux = 0.0, muy = 0.0;
  const double *mean_dev = nullptr;
};

// SmoothStep, create_mask.f90:475-520 (gerf = |(exp(-4)-1)/sqrt(pi)|)
__device__ __forceinline__ double smooth_step(const MaskConsts &c, double x, double t, double h) {
#pragma clang fp contract(off)
  if (!c.sharp) {
    const double delta = h * c.gerf;
    return 0.5 * (erf((t - x) / delta) + erf((x + t) / delta));
  }
  return (x <= t) ? 1.0 : 0.0;
}

// ConvertAngle(gamma)*180/pi, create_mask.f90:548-559, :603
__device__ __forceinline__ double angle_deg(double g, double pi) {
#pragma clang fp contract(off)
  const double a = (g < 0.0) ? 2.0 * pi + g : g;
  return a * 180.0 / pi;
}

// One workgroup of four waves per 8x8 tile of the mask box (lane = pixel).
// The bounding boxes of all primitives are first copied to LDS.  Then, in
// passes over the primitive list, the ones whose boxes meet the tile are
// compacted in drawing order (ballot) and staged in LDS; the four waves
// evaluate the staged primitives' profile values at the tile's pixels in
// parallel (the erf / sqrt / atan2 work, which does not depend on the running
// maximum), and wave 0 applies them in drawing order with the reference's
// replacement rule.  A value of -1 marks "primitive does not touch this
// pixel" (every profile value is >= 0 and so is the running maximum); the
// sharp end marks "not touched" with -2 and otherwise overwrites.
constexpr int kMaskTile = 8;
constexpr int kMaskPix = kMaskTile * kMaskTile;  // pixels per tile = lanes of a wave
// waves per tile: 8 for one simulation's mask (its ~200 tiles leave most
// CUs free, so more waves per tile shorten the kernel), 4 for a batch's
// (thousands of tiles fill the GPU, and smaller workgroups pack better)
constexpr int kMaskWaves = 8, kMaskWavesBatch = 4;
constexpr int kMaskThreads = kMaskPix * kMaskWaves, kMaskThreadsBatch = kMaskPix * kMaskWavesBatch;
constexpr int kMaxTilePrims = 1024;  // >= number of primitives (host checks)
constexpr int kMaskStage = 64;       // primitives staged per pass (one pass for a tile of the beam)

// profile value of primitive P at (xi, yj), or -1 / -2 (see above)
__device__ __forceinline__ double prim_value(const Prim &P, const MaskConsts &c, int i, int j, double xi,
                                             double yj) {
#pragma clang fp contract(off)
  const double none = P.type == PRIM_SHARP ? -2.0 : -1.0;
  if (i < P.ix0 || i > P.ix1 || j < P.iy0 || j > P.iy1) return none;
  const double *a = P.a;
  if (P.type == PRIM_SEG) {
    // DrawBeamSegment :764-778; a = p1s(2) p2s(2) cos sin t_star v1(2) v2(2)
    const double xs = xi * a[4] - yj * a[5];
    const double ys = xi * a[5] + yj * a[4];
    if (xs >= a[0] && xs <= a[2] && ys <= (a[1] + a[6]) && ys >= (a[3] - a[6]))
      return smooth_step(c, fabs(ys - a[1]), c.t, c.h);
    return -1.0;
  }
  if (P.type == PRIM_HINGE) {
    // DrawHinge :595-653; a = px py beta1 beta2 t_star vx vy d1(2) d2(2)
    const double ddx = xi - a[0], ddy = yj - a[1];
    const double R = sqrt(ddx * ddx + ddy * ddy);
    // The sector test below is "angle to segment 1 <= 90 deg and angle to
    // segment 2 >= 90 deg", i.e. d1.v >= 0 and d2.v <= 0 with d1, d2 the
    // unit directions of the two segments.  Pixels that fail it by more
    // than 1e-9 R (far beyond the rounding of the degree comparisons) are
    // skipped; the rest take the reference's exact atan2 test.
    const double dot1 = a[8] * ddx + a[9] * ddy, dot2 = a[10] * ddx + a[11] * ddy;
    if (!(R <= a[4]) || dot1 < -1e-9 * R || dot2 > 1e-9 * R) return -1.0;
    const double gamma = angle_deg(atan2(yj - a[1], xi - a[0]), c.pi);
    const double b1 = a[2], b2 = a[3];
    bool in1 = false, in2 = false;
    if (b1 < 90.0) {
      in1 = (gamma >= 0.0 && gamma <= 90.0 + b1) || (gamma >= 270.0 + b1 && gamma <= 360.0);
    } else if (b1 >= 90.0 && b1 < 270.0) {
      in1 = (gamma >= 270.0 + b1 - 360.0 && gamma <= 90.0 + b1);
    } else if (b1 >= 270.0) {
      in1 = (gamma >= 0.0 && gamma <= 90.0 + b1 - 360.0) || (gamma >= 270.0 + b1 - 360.0 && gamma <= 360.0);
    }
    if (b2 <= 90.0) {
      in2 = (gamma >= 90.0 + b2 && gamma < 270.0 + b2);
    } else if (b2 >= 90.0 && b2 < 270.0) {
      in2 = (gamma >= 0.0 && gamma <= 270.0 + b2 - 360.0) || (gamma >= 90.0 + b2 && gamma <= 360.0);
    } else if (b2 >= 270.0) {
      in2 = (gamma >= 90.0 + b2 - 360.0 && gamma < 270.0 + b2 - 360.0);
    }
    return (in1 && in2) ? smooth_step(c, R, c.t, c.h) : -1.0;
  }
  if (P.type == PRIM_DOT) {
    // DrawADot :690-701; a = px py vx vy t
    const double ddx = xi - a[0], ddy = yj - a[1];
    const double R = sqrt(ddx * ddx + ddy * ddy);
    return smooth_step(c, R, a[4], c.h);
  }
  // PRIM_SHARP, DrawSharpEnd :825-847; a = p1s(2) p2s(2) cos sin t_star vx vy mdx
  const double xs = xi * a[4] + yj * a[5];
  const double ys = -(xi * a[5]) + yj * a[4];
  if (xs >= a[0] && xs <= a[2] && ys <= (a[1] + a[6]) && ys >= (a[3] - a[6])) {
    const double tmp2 = smooth_step(c, xs - a[0], 2.0 * a[9], c.h);
    const double temp = smooth_step(c, fabs(ys - a[1]), c.t, c.h);
    return temp * tmp2;
  }
  return -2.0;
}

// ---- the force kernel's tables (on the device) ------------------------------
// From the beam block X[nb ns], Y[nb ns], TH[nb ns], alpha[nb] (beam b's
// nodes from b ns) that the host writes with the primitives, workgroups past
// the mask's tiles build the force kernel's tables: per node
// (GetForcesInterp BeamForces.f90:150-165) th = theta + alpha, nd = sin th,
// cos th, sin 2th, cos 2th and the targets P(u1) P(b1) A(u) B(u) A(b) B(b)
// at X -+ r sin th, Y +- r cos th (r1 for the pressure pair, rr for the
// stresses); per circle point (BeamSurfaceIntegral :351-362, :389-400) phi,
// cd = cos phi, sin phi, cos 2phi, sin 2phi and the targets A, B, P at the
// end node + rr (cos phi, sin phi).  Table layout: beam b's targets from
// b (6 ns + 6 nc) (node n: 6n .. 6n+5, circle point i: 6 ns + 3i + {0 A,
// 1 B, 2 P}), nd[4 (b ns + n)], cd[4 (2 b nc + i)].
struct BeamGeo {
  const double *beam;  // the beam block in device memory
  int ns, nb, nc;      // nodes per beam, beams, circle points per beam (0: none)
  int fwg;             // workgroups past the tiles that build the force tables (0: no forces)
  double pi, rr, r1, dphi;
  double *tg, *nd, *cd;
};

// the force tables, one thread per node or circle point (item w)
__device__ __forceinline__ void force_tables(const BeamGeo &g, int w) {
#pragma clang fp contract(off)
  const int ns = g.ns, nc = g.nc, per = ns + 2 * nc;
  if (w >= g.nb * per) return;
  const int b = w / per, k = w - b * per, nsb = g.nb * ns;
  const double *X = g.beam + b * ns, *Y = g.beam + nsb + b * ns, *TH = g.beam + 2 * nsb + b * ns;
  const double alpha = g.beam[3 * nsb + b];
  double *tg = g.tg + (size_t)3 * b * (6 * ns + 6 * nc);
  if (k < ns) {
    const double th = TH[k] + alpha;
    const double st = sin(th), ct = cos(th);
    double *nd = g.nd + 4 * ((size_t)b * ns + k);
    nd[0] = st; nd[1] = ct; nd[2] = sin(2.0 * th); nd[3] = cos(2.0 * th);
    const double xu = X[k] - g.rr * st, yu = Y[k] + g.rr * ct, xb = X[k] + g.rr * st, yb = Y[k] - g.rr * ct;
    const double xu1 = X[k] - g.r1 * st, yu1 = Y[k] + g.r1 * ct, xb1 = X[k] + g.r1 * st, yb1 = Y[k] - g.r1 * ct;
    const double t6[18] = {xu1, yu1, 0.0, xb1, yb1, 0.0, xu, yu, 1.0, xu, yu, 2.0, xb, yb, 1.0, xb, yb, 2.0};
    double *o = tg + 18 * k;
#pragma unroll
    for (int q = 0; q < 18; ++q) o[q] = t6[q];
  } else {
    const int i = k - ns;  // circle point: i < nc trailing end, else leading end
    double phi, px, py;
    if (i < nc) {
      const double beta = TH[ns - 1] + alpha;
      phi = -(0.5 * g.pi - beta) + (double)i * g.dphi;
      px = X[ns - 1]; py = Y[ns - 1];
    } else {
      const double beta = TH[1] + alpha;  // BeamForces.f90:390 takes node 1
      phi = (0.5 * g.pi + beta) + (double)(i - nc) * g.dphi;
      px = X[0]; py = Y[0];
    }
    const double cp = cos(phi), sp = sin(phi);
    double *cd = g.cd + 4 * ((size_t)2 * b * nc + i);
    cd[0] = cp; cd[1] = sp; cd[2] = cos(2.0 * phi); cd[3] = sin(2.0 * phi);
    const double xc = px + g.rr * cp, yc = py + g.rr * sp;
    double *o = tg + 18 * ns + 9 * i;
    o[0] = xc; o[1] = yc; o[2] = 1.0;
    o[3] = xc; o[4] = yc; o[5] = 2.0;
    o[6] = xc; o[7] = yc; o[8] = 0.0;
  }
}

// owin (may be null): index of the primitive that set each pixel's value
// (-1: none, the wall or zero), read by the adjoint of the mask
template <int WAVES>
__device__ __forceinline__ void mask_body(const Prim *__restrict__ prims, int np, const double *__restrict__ wallrow,
                                          const MaskConsts &c, int bx0, int by0, int bw, int bh,
                                          double *__restrict__ om, double *__restrict__ ovx,
                                          double *__restrict__ ovy, int *__restrict__ owin, int tbx, int tby) {
#pragma clang fp contract(off)
  extern __shared__ int4 bnd[];  // ix0, ix1, iy0, iy1 of the np primitives (dynamic LDS, np entries)
  __shared__ Prim sp[kMaskStage];
  __shared__ int list[kMaskStage];
  __shared__ double val[kMaskStage][kMaskPix];
  __shared__ int nstage, pnext;
  const int tid = threadIdx.x, lane = tid & (kMaskPix - 1), wv = tid / kMaskPix;
  const int tx0 = bx0 + tbx * kMaskTile, ty0 = by0 + tby * kMaskTile;
  const int tx1 = tx0 + kMaskTile - 1, ty1 = ty0 + kMaskTile - 1;
  constexpr int NTHR = kMaskPix * WAVES;
  for (int p = tid; p < np; p += NTHR) {
    const int4 h = reinterpret_cast<const int4 *>(prims + p)[0];  // type ix0 ix1 iy0
    bnd[p] = make_int4(h.y, h.z, h.w, prims[p].iy1);
  }
  if (tid == 0) pnext = 0;
  __syncthreads();
  const int lx = tbx * kMaskTile + (lane & (kMaskTile - 1));
  const int ly = tby * kMaskTile + (lane / kMaskTile);
  const bool valid = lx < bw && ly < bh;
  const int i = bx0 + lx, j = by0 + ly;
  const double xi = (double)i * c.dx, yj = (double)j * c.dy;
  double m = (valid && wallrow) ? wallrow[j] : 0.0, mvx = 0.0, mvy = 0.0;
  int wi = -1;
  while (true) {
    // wave 0 compacts the next primitives that meet the tile, in order
    if (wv == 0) {
      int p0 = pnext, n = 0;
      while (p0 < np) {
        const int p = p0 + lane;
        bool f = false;
        if (p < np) {
          const int4 b = bnd[p];
          f = !(b.y < tx0 || b.x > tx1 || b.w < ty0 || b.z > ty1);
        }
        const unsigned long long bal = __ballot(f);
        const int cnt = __popcll(bal);
        if (n + cnt > kMaskStage) {  // take the first kMaskStage - n of this group
          const int take = kMaskStage - n;
          const int pre = __popcll(bal & ((1ull << lane) - 1ull));
          if (f && pre < take) list[n + pre] = p;
          // resume after the last primitive taken
          unsigned long long bb = bal;
          int lastlane = -1;
          for (int k = 0; k < take; ++k) {
            lastlane = __ffsll((long long)bb) - 1;
            bb &= bb - 1ull;
          }
          p0 = p0 + lastlane + 1;
          n = kMaskStage;
          break;
        }
        if (f) list[n + __popcll(bal & ((1ull << lane) - 1ull))] = p;
        n += cnt;
        p0 += kMaskPix;
      }
      if (lane == 0) {
        nstage = n;
        pnext = p0 < np ? p0 : np;
      }
    }
    __syncthreads();
    // read once per pass: wave 0 rewrites both at the start of the next pass
    const int n = nstage, pn = pnext;
    if (n == 0) break;
    for (int e = tid; e < n * (int)(sizeof(Prim) / 8); e += NTHR) {
      const int q = e / (int)(sizeof(Prim) / 8), w = e - q * (int)(sizeof(Prim) / 8);
      reinterpret_cast<double *>(sp + q)[w] = reinterpret_cast<const double *>(prims + list[q])[w];
    }
    __syncthreads();
    for (int q = wv; q < n; q += WAVES) val[q][lane] = prim_value(sp[q], c, i, j, xi, yj);
    __syncthreads();
    if (wv == 0 && valid) {
      for (int q = 0; q < n; ++q) {
        const Prim &P = sp[q];
        const double v = val[q][lane];
        const double *a = P.a;
        if (P.type == PRIM_SHARP) {
          if (v > -1.5) {
            m = v;
            wi = list[q];
            const double xs = xi * a[4] + yj * a[5];
            if (xs - a[0] >= 2.0 * a[9]) { mvx = a[7]; mvy = a[8]; }
          }
        } else if (v > m) {
          m = v;
          wi = list[q];
          if (P.type == PRIM_SEG) {
            const double xs = xi * a[4] - yj * a[5];
            mvx = (a[7] + (xs - a[0]) * (a[9] - a[7]) / c.ds);
            mvy = (a[8] + (xs - a[0]) * (a[10] - a[8]) / c.ds);
          } else if (P.type == PRIM_HINGE) {
            mvx = a[5]; mvy = a[6];
          } else {
            mvx = a[2]; mvy = a[3];
          }
        }
      }
    }
    __syncthreads();  // the next pass overwrites list, sp and val
    if (pn >= np) break;
  }
  if (wv != 0 || !valid) return;
  if (c.vs4 || c.frame) {
    const double mux = c.mean_dev ? c.mean_dev[0] : c.mux, muy = c.mean_dev ? c.mean_dev[1] : c.muy;
    if (c.vs4) {
      // left half (x = 0.75 S - i dx, step of half width S/4), then the right
      // half (x = i dx - 0.5 S, the smoothing layer); the middle column gets
      // both; maskvx is the channel's parabola wherever the sponge is drawn
      const double y_chan = (double)j * c.dy - c.h_channel;
      const double prof = (y_chan > 0.0 && y_chan < c.vs_Heff)
                              ? 1.5 * mux * y_chan * (c.vs_Heff - y_chan) / ((0.5 * c.vs_Heff) * (0.5 * c.vs_Heff))
                              : 0.0;
      if (i >= 0 && i <= c.vs_ixmid) {
        const double tmp = smooth_step(c, 0.75 * c.vs_S - (double)i * c.dx, c.vs_ssz * 0.5, 0.25 * c.vs_ssz);
        if (m < tmp) m = tmp;
        mvx = prof;
      }
      if (i >= c.vs_ixmid && i <= c.vs_ixmax) {
        const double tmp = smooth_step(c, (double)i * c.dx - 0.5 * c.vs_S, c.vs_ssz * 0.5, c.N * c.dx);
        if (m < tmp) m = tmp;
        mvx = prof;
      }
    }
    if (c.frame) {
      const double fr = c.frame[(size_t)j * (size_t)c.nx_frame + i];
      m = m + (fr * c.eps) / c.frame_M;
      if (fr > 0.0) { mvx = mux; mvy = muy; }
    }
  }
  const size_t q = (size_t)ly * bw + lx;
  om[q] = m / c.eps;  // :210-214
  ovx[q] = mvx;
  ovy[q] = mvy;
  if (owin) owin[q] = wi;
}

// grid: the tiles (tiles_y rows of gridDim.x), then one more row whose first
// bg.fwg workgroups build the force tables
__global__ void __launch_bounds__(kMaskThreads) k_mask(const Prim *__restrict__ prims, int np,
                                              const double *__restrict__ wallrow, MaskConsts c, BeamGeo bg,
                                              int tiles_y, int bx0, int by0, int bw, int bh,
                                              double *__restrict__ om, double *__restrict__ ovx,
                                              double *__restrict__ ovy, int *__restrict__ owin) {
  if ((int)blockIdx.y >= tiles_y) {
    if ((int)blockIdx.x < bg.fwg) force_tables(bg, blockIdx.x * kMaskThreads + threadIdx.x);
    return;
  }
  if ((int)blockIdx.x * kMaskTile >= bw) return;  // the grid is as wide as the force row needs
  mask_body<kMaskWaves>(prims, np, wallrow, c, bx0, by0, bw, bh, om, ovx, ovy, owin, blockIdx.x, blockIdx.y);
}

// batched: one record per simulation at blockIdx.y, tiles flattened on x
struct MaskArgs {
  const Prim *prims;
  const double *wallrow;
  double *om, *ovx, *ovy;
  MaskConsts c;
  BeamGeo bg;
  int np, bx0, by0, bw, bh, tiles_x, tiles, pad;
};
// blocks past a member's tiles: its first bg.fwg build its force tables
__global__ void __launch_bounds__(kMaskThreadsBatch) kb_mask(const MaskArgs *__restrict__ arr) {
  const MaskArgs &a = arr[blockIdx.y];
  if ((int)blockIdx.x >= a.tiles) {
    const int w = (int)blockIdx.x - a.tiles;
    if (w < a.bg.fwg) force_tables(a.bg, w * kMaskThreadsBatch + threadIdx.x);
    return;
  }
  const int tby = blockIdx.x / a.tiles_x, tbx = blockIdx.x - tby * a.tiles_x;
  mask_body<kMaskWavesBatch>(a.prims, a.np, a.wallrow, a.c, a.bx0, a.by0, a.bw, a.bh, a.om, a.ovx, a.ovy, nullptr,
                             tbx, tby);
}

}  // namespace fsibe
