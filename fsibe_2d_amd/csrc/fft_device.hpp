// fft_device.hpp -- workgroup-level FP64 complex FFTs staged in LDS (gfx950).
//
// A workgroup transforms `batch` independent sequences of length L held in
// LDS.  Element e of sequence b lives at buf[b*bs + e*es].  Each pass is one
// in-place Stockham radix-R step: every thread loads its butterflies into
// registers, the workgroup synchronises, and every thread writes the results
// back to the autosorted positions, so a single LDS buffer suffices and the
// result is in natural order after the last pass.
//
// Radices 8, 4, 2 and 3 are supported; the host picks the pass sequence
// (FftPlan).  Twiddles come from a global table tw[m] = exp(-2 pi i m / Ltab)
// read with stride Ltab/L (L1/L2-resident: at most 64 KiB per length).
#pragma once
#include <hip/hip_runtime.h>

namespace fsibe {

struct FftPlan {
  int L;        // transform length
  int np;       // number of passes
  int r[16];    // radices, pass order
};

__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
// multiply by -i (forward) or +i (inverse)
template <bool INV>
__device__ __forceinline__ double2 mul_mi(double2 a) {
  return INV ? make_double2(-a.y, a.x) : make_double2(a.y, -a.x);
}

// ---- small DFTs: y_q = sum_r v_r exp(-+2 pi i r q / R) -----------------------
template <int R, bool INV> struct Dft;

template <bool INV> struct Dft<2, INV> {
  __device__ __forceinline__ static void run(double2 *v) {
    double2 a = cadd(v[0], v[1]), b = csub(v[0], v[1]);
    v[0] = a; v[1] = b;
  }
};

template <bool INV> struct Dft<4, INV> {
  __device__ __forceinline__ static void run(double2 *v) {
    double2 a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
    double2 a2 = cadd(v[1], v[3]), a3 = mul_mi<INV>(csub(v[1], v[3]));
    v[0] = cadd(a0, a2); v[2] = csub(a0, a2);
    v[1] = cadd(a1, a3); v[3] = csub(a1, a3);
  }
};

template <bool INV> struct Dft<8, INV> {
  __device__ __forceinline__ static void run(double2 *v) {
    const double h = 0.70710678118654752440;  // sqrt(1/2)
    double2 a[4] = {v[0], v[2], v[4], v[6]};
    double2 b[4] = {v[1], v[3], v[5], v[7]};
    Dft<4, INV>::run(a);
    Dft<4, INV>::run(b);
    // b1 *= w8, b2 *= w8^2, b3 *= w8^3 with w8 = exp(-+ i pi/4)
    double2 t1 = INV ? make_double2(h * (b[1].x - b[1].y), h * (b[1].x + b[1].y))
                     : make_double2(h * (b[1].x + b[1].y), h * (b[1].y - b[1].x));
    double2 t2 = mul_mi<INV>(b[2]);
    double2 t3 = INV ? make_double2(-h * (b[3].x + b[3].y), h * (b[3].x - b[3].y))
                     : make_double2(h * (b[3].y - b[3].x), -h * (b[3].x + b[3].y));
    v[0] = cadd(a[0], b[0]); v[4] = csub(a[0], b[0]);
    v[1] = cadd(a[1], t1);   v[5] = csub(a[1], t1);
    v[2] = cadd(a[2], t2);   v[6] = csub(a[2], t2);
    v[3] = cadd(a[3], t3);   v[7] = csub(a[3], t3);
  }
};

template <bool INV> struct Dft<3, INV> {
  __device__ __forceinline__ static void run(double2 *v) {
    const double s3 = 0.86602540378443864676;  // sin(2 pi / 3)
    double2 t = cadd(v[1], v[2]), d = csub(v[1], v[2]);
    double2 m = make_double2(v[0].x - 0.5 * t.x, v[0].y - 0.5 * t.y);
    // e = (sign i s3) d, sign = -1 forward
    double2 e = INV ? make_double2(-s3 * d.y, s3 * d.x) : make_double2(s3 * d.y, -s3 * d.x);
    v[0] = cadd(v[0], t);
    v[1] = cadd(m, e);
    v[2] = csub(m, e);
  }
};

// ---- one in-place Stockham pass ------------------------------------------------
// Work item w in [0, L/R * batch) is decoded as (b, j) with b fastest when
// BFAST (column tiles: b is the contiguous index) and j fastest otherwise.
template <int R, int MAXIT, bool INV, bool BFAST>
__device__ __forceinline__ void stockham_pass(double2 *__restrict__ buf, int L, int batch, int bs,
                                              int es, int Ns, const double2 *__restrict__ tw,
                                              int twmul, int tid, int nthr) {
  const int m = L / R;
  const int total = m * batch;
  const int tstep = (L / (Ns * R)) * twmul;
  double2 v[MAXIT][R];
  int bb[MAXIT], jj[MAXIT];
#pragma unroll
  for (int it = 0; it < MAXIT; ++it) {
    const int w = tid + it * nthr;
    if (w < total) {
      int b, j;
      if (BFAST) { b = w % batch; j = w / batch; } else { j = w % m; b = w / m; }
      bb[it] = b; jj[it] = j;
      const int k = j % Ns;
      double2 *p = buf + b * bs + j * es;
#pragma unroll
      for (int r = 0; r < R; ++r) v[it][r] = p[r * m * es];
      if (Ns > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          double2 t = tw[r * k * tstep];
          if (INV) t.y = -t.y;
          v[it][r] = cmul(v[it][r], t);
        }
      }
      Dft<R, INV>::run(v[it]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < MAXIT; ++it) {
    const int w = tid + it * nthr;
    if (w < total) {
      const int j = jj[it];
      const int k = j % Ns;
      const int base = (j / Ns) * Ns * R + k;
      double2 *p = buf + bb[it] * bs + base * es;
#pragma unroll
      for (int r = 0; r < R; ++r) p[r * Ns * es] = v[it][r];
    }
  }
  __syncthreads();
}

// Runs every pass of `plan`.  EMAX bounds L*batch/nthr (elements per thread);
// the host guarantees it.  Caller must __syncthreads() before (data in LDS);
// the function ends with a barrier.
template <int EMAX, bool INV, bool BFAST>
__device__ void lds_fft(double2 *__restrict__ buf, const FftPlan &plan, int batch, int bs, int es,
                        const double2 *__restrict__ tw, int twmul, int tid, int nthr) {
  int Ns = 1;
  for (int p = 0; p < plan.np; ++p) {
    const int R = plan.r[p];
    switch (R) {
      case 8: stockham_pass<8, (EMAX + 7) / 8, INV, BFAST>(buf, plan.L, batch, bs, es, Ns, tw, twmul, tid, nthr); break;
      case 4: stockham_pass<4, (EMAX + 3) / 4, INV, BFAST>(buf, plan.L, batch, bs, es, Ns, tw, twmul, tid, nthr); break;
      case 2: stockham_pass<2, (EMAX + 1) / 2, INV, BFAST>(buf, plan.L, batch, bs, es, Ns, tw, twmul, tid, nthr); break;
      case 3: stockham_pass<3, (EMAX + 2) / 3 + 1, INV, BFAST>(buf, plan.L, batch, bs, es, Ns, tw, twmul, tid, nthr); break;
      default: break;
    }
    Ns *= R;
  }
}

}  // namespace fsibe
