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

// One workgroup per 16x16 tile of the mask box.  The primitives whose boxes
// meet the tile are first compacted, in drawing order, into an LDS list
// (ballot + prefix count per chunk of 256), so each pixel visits only those.
constexpr int kMaskTile = 8;
constexpr int kMaskThreads = kMaskTile * kMaskTile;
constexpr int kMaxTilePrims = 1024;  // >= number of primitives (host checks)

__global__ void __launch_bounds__(kMaskThreads) k_mask(const Prim *__restrict__ prims, int np,
                                              const double *__restrict__ wallrow, MaskConsts c, int bx0, int by0,
                                              int bw, int bh, double *__restrict__ om, double *__restrict__ ovx,
                                              double *__restrict__ ovy) {
#pragma clang fp contract(off)
  __shared__ int list[kMaxTilePrims];
  __shared__ int wcount[kMaskThreads / 64];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int tx0 = bx0 + blockIdx.x * kMaskTile, ty0 = by0 + blockIdx.y * kMaskTile;
  const int tx1 = tx0 + kMaskTile - 1, ty1 = ty0 + kMaskTile - 1;
  int nlist = 0;
  for (int c0 = 0; c0 < np; c0 += kMaskThreads) {
    const int p = c0 + tid;
    bool f = false;
    if (p < np) {
      const Prim &P = prims[p];
      f = !(P.ix1 < tx0 || P.ix0 > tx1 || P.iy1 < ty0 || P.iy0 > ty1);
    }
    const unsigned long long b = __ballot(f);
    const int pre = __popcll(b & ((1ull << lane) - 1ull));
    if (lane == 0) wcount[wid] = __popcll(b);
    __syncthreads();
    int off = nlist;
    for (int w = 0; w < wid; ++w) off += wcount[w];
    if (f && off + pre < kMaxTilePrims) list[off + pre] = p;
    int tot = 0;
    for (int w = 0; w < kMaskThreads / 64; ++w) tot += wcount[w];
    __syncthreads();
    nlist += tot;
  }
  nlist = nlist < kMaxTilePrims ? nlist : kMaxTilePrims;
  const int lx = blockIdx.x * kMaskTile + (tid & (kMaskTile - 1));
  const int ly = blockIdx.y * kMaskTile + (tid / kMaskTile);
  if (lx >= bw || ly >= bh) return;
  const int i = bx0 + lx, j = by0 + ly;
  double m = wallrow ? wallrow[j] : 0.0, mvx = 0.0, mvy = 0.0;
  const double xi = (double)i * c.dx, yj = (double)j * c.dy;
  for (int q = 0; q < nlist; ++q) {
    // every lane of the wave reads the same entry: keep it in a scalar register
    const Prim &P = prims[__builtin_amdgcn_readfirstlane(list[q])];
    if (i < P.ix0 || i > P.ix1 || j < P.iy0 || j > P.iy1) continue;
    const double *a = P.a;
    if (P.type == PRIM_SEG) {
      // DrawBeamSegment :764-778; a = p1s(2) p2s(2) cos sin t_star v1(2) v2(2)
      const double xs = xi * a[4] - yj * a[5];
      const double ys = xi * a[5] + yj * a[4];
      if (xs >= a[0] && xs <= a[2] && ys <= (a[1] + a[6]) && ys >= (a[3] - a[6])) {
        const double temp = smooth_step(c, fabs(ys - a[1]), c.t, c.h);
        if (temp > m) {
          m = temp;
          mvx = (a[7] + (xs - a[0]) * (a[9] - a[7]) / c.ds);
          mvy = (a[8] + (xs - a[0]) * (a[10] - a[8]) / c.ds);
        }
      }
    } else if (P.type == PRIM_HINGE) {
      // DrawHinge :595-653; a = px py beta1 beta2 t_star vx vy d1(2) d2(2)
      const double ddx = xi - a[0], ddy = yj - a[1];
      const double R = sqrt(ddx * ddx + ddy * ddy);
      // The sector test below is "angle to segment 1 <= 90 deg and angle to
      // segment 2 >= 90 deg", i.e. d1.v >= 0 and d2.v <= 0 with d1, d2 the
      // unit directions of the two segments.  Pixels that fail it by more
      // than 1e-9 R (far beyond the rounding of the degree comparisons) are
      // skipped; the rest take the reference's exact atan2 test.
      const double dot1 = a[8] * ddx + a[9] * ddy, dot2 = a[10] * ddx + a[11] * ddy;
      if (R <= a[4] && !(dot1 < -1e-9 * R || dot2 > 1e-9 * R)) {
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
        if (in1 && in2) {
          const double temp = smooth_step(c, R, c.t, c.h);
          if (temp > m) { m = temp; mvx = a[5]; mvy = a[6]; }
        }
      }
    } else if (P.type == PRIM_DOT) {
      // DrawADot :690-701; a = px py vx vy t
      const double ddx = xi - a[0], ddy = yj - a[1];
      const double R = sqrt(ddx * ddx + ddy * ddy);
      const double temp = smooth_step(c, R, a[4], c.h);
      if (temp > m) { m = temp; mvx = a[2]; mvy = a[3]; }
    } else if (P.type == PRIM_SHARP) {
      // DrawSharpEnd :825-847; a = p1s(2) p2s(2) cos sin t_star vx vy mdx
      const double xs = xi * a[4] + yj * a[5];
      const double ys = -(xi * a[5]) + yj * a[4];
      if (xs >= a[0] && xs <= a[2] && ys <= (a[1] + a[6]) && ys >= (a[3] - a[6])) {
        const double tmp2 = smooth_step(c, xs - a[0], 2.0 * a[9], c.h);
        const double temp = smooth_step(c, fabs(ys - a[1]), c.t, c.h);
        m = temp * tmp2;
        if (xs - a[0] >= 2.0 * a[9]) { mvx = a[7]; mvy = a[8]; }
      }
    }
  }
  const size_t q = (size_t)ly * bw + lx;
  om[q] = m / c.eps;  // :210-214
  ovx[q] = mvx;
  ovy[q] = mvy;
}

}  // namespace fsibe
