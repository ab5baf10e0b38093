// fft_device.hpp -- workgroup-level FP64 complex FFTs staged in LDS (gfx950).
//
// A workgroup transforms `batch` independent sequences of length L held in
// LDS.  Element e of sequence b lives at buf[b*bs + pidx(e)*es]; pidx inserts
// one empty slot after every 8 elements, which spreads the stride-8 writes of
// the first radix-8 pass over all 64 LDS banks.  Each pass is one in-place
// Stockham radix-R step: every thread loads its butterflies into registers,
// the workgroup synchronises, and every thread writes the results back to the
// autosorted positions; the result is in natural order after the last pass.
//
// Radices 16, 8, 4, 2 and 3 are supported; the host picks the pass sequence
// (FftPlan).  Twiddles: w = exp(-2 pi i k/(Ns R)) is read once per butterfly
// from the table tw[m] = exp(-2 pi i m / Ltab) (stride Ltab/L) and its powers
// w^2 .. w^(R-1) are formed by complex multiplication (twiddle_row).
#pragma once
#include <hip/hip_runtime.h>

namespace fsibe {

struct FftPlan {
  int L;        // transform length
  int np;       // number of passes
  int r[16];    // radices, pass order
  // the radices again, 4 bits per pass (radix_code), for the device: read
  // with shifts, where an index into r[] at run time would place the plan
  // in per-lane scratch memory
  unsigned long long code;
};

// 4-bit code of a radix: 2 -> 0, 3 -> 1, 4 -> 2, 8 -> 3, 16 -> 4
__host__ __device__ __forceinline__ constexpr int radix_code(int R) {
  return R == 2 ? 0 : R == 3 ? 1 : R == 4 ? 2 : R == 8 ? 3 : 4;
}

// padded position of element e and padded length of a sequence of L elements
__host__ __device__ __forceinline__ constexpr int pidx(int e) { return e + (e >> 3); }
__host__ __device__ __forceinline__ constexpr int padlen(int L) { return L > 0 ? pidx(L - 1) + 1 : 0; }

__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
// Complex product with its two fused multiply-adds written out.  The
// library builds with contraction off, so every fused operation of the FFT
// is explicit and every instantiation (compile-time or runtime plan, any
// tile width) rounds exactly alike.
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(fma(a.x, b.x, -(a.y * b.y)), fma(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
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
    // b1 *= w8 and b3 *= w8^3 (w8 = exp(-+ i pi/4)) as h * s1, h * s3, each
    // folded into the sums below by a fused multiply-add; b2 *= w8^2 = -+i
    const double2 s1 = INV ? make_double2(b[1].x - b[1].y, b[1].x + b[1].y)
                           : make_double2(b[1].x + b[1].y, b[1].y - b[1].x);
    const double2 s3 = INV ? make_double2(-(b[3].x + b[3].y), b[3].x - b[3].y)
                           : make_double2(b[3].y - b[3].x, -(b[3].x + b[3].y));
    const double2 t2 = mul_mi<INV>(b[2]);
    v[0] = cadd(a[0], b[0]); v[4] = csub(a[0], b[0]);
    v[1] = make_double2(fma(h, s1.x, a[1].x), fma(h, s1.y, a[1].y));
    v[5] = make_double2(fma(-h, s1.x, a[1].x), fma(-h, s1.y, a[1].y));
    v[2] = cadd(a[2], t2);   v[6] = csub(a[2], t2);
    v[3] = make_double2(fma(h, s3.x, a[3].x), fma(h, s3.y, a[3].y));
    v[7] = make_double2(fma(-h, s3.x, a[3].x), fma(-h, s3.y, a[3].y));
  }
};

// 16 = 4 x 4: radix-4 DFTs of the four stride-4 subsequences, the twiddles
// W16^(q k1), then radix-4 DFTs across them; output k1 + 4 k2
template <bool INV> struct Dft<16, INV> {
  __device__ __forceinline__ static void run(double2 *v) {
    const double c1 = 0.92387953251128675613, s1 = 0.38268343236508977173;  // cos, sin (pi/8)
    const double h = 0.70710678118654752440;                                // sqrt(1/2)
    double2 a[4][4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
#pragma unroll
      for (int r = 0; r < 4; ++r) a[q][r] = v[q + 4 * r];
      Dft<4, INV>::run(a[q]);
    }
    // W16^m = exp(-+ 2 pi i m / 16) for m = 1, 2, 3, 6, 9 (m = 4 is -+i)
    const double2 w1 = INV ? make_double2(c1, s1) : make_double2(c1, -s1);
    const double2 w2 = INV ? make_double2(h, h) : make_double2(h, -h);
    const double2 w3 = INV ? make_double2(s1, c1) : make_double2(s1, -c1);
    const double2 w6 = INV ? make_double2(-h, h) : make_double2(-h, -h);
    const double2 w9 = INV ? make_double2(-c1, -s1) : make_double2(-c1, s1);
    a[1][1] = cmul(a[1][1], w1); a[1][2] = cmul(a[1][2], w2); a[1][3] = cmul(a[1][3], w3);
    a[2][1] = cmul(a[2][1], w2); a[2][2] = mul_mi<INV>(a[2][2]); a[2][3] = cmul(a[2][3], w6);
    a[3][1] = cmul(a[3][1], w3); a[3][2] = cmul(a[3][2], w6); a[3][3] = cmul(a[3][3], w9);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
      double2 b[4] = {a[0][k1], a[1][k1], a[2][k1], a[3][k1]};
      Dft<4, INV>::run(b);
#pragma unroll
      for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = b[k2];
    }
  }
};

template <bool INV> struct Dft<3, INV> {
  __device__ __forceinline__ static void run(double2 *v) {
    const double s3 = 0.86602540378443864676;  // sin(2 pi / 3)
    const double2 t = cadd(v[1], v[2]), d = csub(v[1], v[2]);
    const double2 m = make_double2(fma(-0.5, t.x, v[0].x), fma(-0.5, t.y, v[0].y));
    // v1 = m + e, v2 = m - e with e = (sign i s3) d, sign = -1 forward
    const double ex = INV ? -s3 : s3;  // e.x = ex * d.y, e.y = -ex * d.x
    v[0] = cadd(v[0], t);
    v[1] = make_double2(fma(ex, d.y, m.x), fma(-ex, d.x, m.y));
    v[2] = make_double2(fma(-ex, d.y, m.x), fma(ex, d.x, m.y));
  }
};

// v[r] *= w1^r for r = 1 .. R-1.  The powers come by binary splitting,
// w^r = w^(r/2) w^(r - r/2), so each is at most ceil(log2 r) products from
// w1 (a chain w^r = w^(r-1) w1 would be r-1 products deep, with that many
// roundings and a serial dependency); every pass kind calls this, so every
// instance of a transform rounds alike.
template <int R>
__device__ __forceinline__ void twiddle_row(double2 *v, double2 w1) {
  double2 w[R];
  w[1] = w1;
#pragma unroll
  for (int r = 2; r < R; ++r) w[r] = cmul(w[r / 2], w[r - r / 2]);
#pragma unroll
  for (int r = 1; r < R; ++r) v[r] = cmul(v[r], w[r]);
}

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
      double2 *p = buf + b * bs;
#pragma unroll
      for (int r = 0; r < R; ++r) v[it][r] = p[pidx(j + r * m) * es];
      if (Ns > 1) {
        double2 w1 = tw[k * tstep];
        if (INV) w1.y = -w1.y;
        twiddle_row<R>(v[it], w1);
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
      double2 *p = buf + bb[it] * bs;
#pragma unroll
      for (int r = 0; r < R; ++r) p[pidx(base + r * Ns) * es] = v[it][r];
    }
  }
  __syncthreads();
}

// Runs every pass of `plan`.  EMAX bounds L*batch/nthr (elements per thread);
// the host guarantees it, and gives plans with a radix-16 pass only to
// instances with EMAX >= 16 (the others leave the 16-point butterfly out:
// it would set their register budget).  Caller must __syncthreads() before
// (data in LDS); the function ends with a barrier.
template <int EMAX, bool INV, bool BFAST>
__device__ void lds_fft(double2 *__restrict__ buf, const FftPlan &plan, int batch, int bs, int es,
                        const double2 *__restrict__ tw, int twmul, int tid, int nthr) {
  int Ns = 1;
  const unsigned long long code = plan.code;
  for (int p = 0; p < plan.np; ++p) {
    switch ((int)((code >> (4 * p)) & 15ull)) {
      case 4:
        if constexpr (EMAX >= 16)
          stockham_pass<16, (EMAX + 15) / 16, INV, BFAST>(buf, plan.L, batch, bs, es, Ns, tw, twmul, tid, nthr);
        Ns *= 16;
        break;
      case 3: stockham_pass<8, (EMAX + 7) / 8, INV, BFAST>(buf, plan.L, batch, bs, es, Ns, tw, twmul, tid, nthr); Ns *= 8; break;
      case 2: stockham_pass<4, (EMAX + 3) / 4, INV, BFAST>(buf, plan.L, batch, bs, es, Ns, tw, twmul, tid, nthr); Ns *= 4; break;
      case 0: stockham_pass<2, (EMAX + 1) / 2, INV, BFAST>(buf, plan.L, batch, bs, es, Ns, tw, twmul, tid, nthr); Ns *= 2; break;
      case 1: stockham_pass<3, (EMAX + 2) / 3, INV, BFAST>(buf, plan.L, batch, bs, es, Ns, tw, twmul, tid, nthr); Ns *= 3; break;
      default: break;
    }
  }
}

// ---- compile-time plans ------------------------------------------------------------
// The same pass sequence as the host's runtime planner: factors 3 first, then
// 16 or 2 or 4 for the power of two left after the eights, then 8s.
__host__ __device__ constexpr int cp_n3(int L) {
  int c = 0;
  while (L > 0 && L % 3 == 0) { L /= 3; ++c; }
  return c;
}
__host__ __device__ constexpr int cp_e2(int L) {
  while (L > 0 && L % 3 == 0) L /= 3;
  int c = 0;
  while (L > 0 && L % 2 == 0) { L /= 2; ++c; }
  return c;
}
__host__ __device__ constexpr int cp_np(int L) {
  const int n3 = cp_n3(L), e2 = cp_e2(L), a = e2 / 3, b = e2 % 3;
  return n3 + a + ((b == 1 && a >= 1) ? 0 : (b != 0 ? 1 : 0));
}
__host__ __device__ constexpr int cp_radix(int L, int p) {
  const int n3 = cp_n3(L), e2 = cp_e2(L), a = e2 / 3, b = e2 % 3;
  if (p < n3) return 3;
  p -= n3;
  if (b == 1 && a >= 1) return p < 1 ? 16 : 8;
  if (b == 1) return p < 1 ? 2 : 8;
  if (b == 2) return p < 1 ? 4 : 8;
  return 8;
}

// In-place Stockham FFT with every size, stride and trip count known at
// compile time (index arithmetic folds to shifts and constant multiplies).
// Element e of sequence b at buf[b*BS + pidx(e)*ES]; NT threads take part.
template <int L, bool INV, bool BFAST, int BATCH, int BS, int ES, int NT, int TWM, int P = 0, int NS = 1>
__device__ __forceinline__ void fft_ct(double2 *__restrict__ buf, const double2 *__restrict__ tw, int tid) {
  if constexpr (P < cp_np(L)) {
    constexpr int R = cp_radix(L, P);
    constexpr int M = L / R;
    constexpr int TOTAL = M * BATCH;
    constexpr int ITEMS = (TOTAL + NT - 1) / NT;
    constexpr int TSTEP = (L / (NS * R)) * TWM;
    double2 v[ITEMS][R];
    int bq[ITEMS], jq[ITEMS];
#pragma unroll
    for (int it = 0; it < ITEMS; ++it) {
      const int w = tid + it * NT;
      if (TOTAL % NT == 0 || w < TOTAL) {
        int b, j;
        if constexpr (BFAST) { b = w % BATCH; j = w / BATCH; } else { j = w % M; b = w / M; }
        bq[it] = b; jq[it] = j;
        double2 *p = buf + b * BS;
#pragma unroll
        for (int r = 0; r < R; ++r) v[it][r] = p[pidx(j + r * M) * ES];
        if constexpr (NS > 1) {
          double2 w1 = tw[(j % NS) * TSTEP];
          if (INV) w1.y = -w1.y;
          twiddle_row<R>(v[it], w1);
        }
        Dft<R, INV>::run(v[it]);
      }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < ITEMS; ++it) {
      const int w = tid + it * NT;
      if (TOTAL % NT == 0 || w < TOTAL) {
        const int j = jq[it];
        const int base = (j / NS) * NS * R + (j % NS);
        double2 *p = buf + bq[it] * BS;
#pragma unroll
        for (int r = 0; r < R; ++r) p[pidx(base + r * NS) * ES] = v[it][r];
      }
    }
    __syncthreads();
    fft_ct<L, INV, BFAST, BATCH, BS, ES, NT, TWM, P + 1, NS * R>(buf, tw, tid);
  }
}

// ---- transform policies: the kernels are written once against these -------------
// Row transforms (length H, sequences of padlen(H) back to back, twiddle table
// of length 2H).  Runtime plan:
template <int EMAX> struct RowRT {
  static constexpr int LB = EMAX <= 8 ? 1024 : 256;
  FftPlan plan;
  __device__ __forceinline__ int H() const { return plan.L; }
  template <bool INV, int NSEQ>
  __device__ __forceinline__ void fft(double2 *lds, const double2 *tw, int tid, int nt) const {
    lds_fft<EMAX, INV, false>(lds, plan, NSEQ, padlen(plan.L), 1, tw, 2, tid, nt);
  }
};
// compile-time plan, launched with NTC threads
template <int HC, int NTC> struct RowCT {
  static constexpr int LB = NTC;
  FftPlan plan;
  __device__ __forceinline__ int H() const { return HC; }
  template <bool INV, int NSEQ>
  __device__ __forceinline__ void fft(double2 *lds, const double2 *tw, int tid, int) const {
    fft_ct<HC, INV, false, NSEQ, padlen(HC), 1, NTC, 2>(lds, tw, tid);
  }
};
// Column tiles of TX columns (length NY); a call transforms BM*TX sequences
// with element stride EM*TX (column c of field f at offset f*TX + c).
template <int EMAX> struct ColRT {
  static constexpr int LB = EMAX <= 8 ? 1024 : 256;
  FftPlan plan;
  int TX;
  __device__ __forceinline__ int NY() const { return plan.L; }
  __device__ __forceinline__ int tx() const { return TX; }
  template <bool INV, int BM, int EM>
  __device__ __forceinline__ void fft(double2 *lds, const double2 *tw, int tid, int nt) const {
    lds_fft<EMAX, INV, true>(lds, plan, BM * TX, 1, EM * TX, tw, 1, tid, nt);
  }
};
template <int NYC, int TXC, int NTC> struct ColCT {
  static constexpr int LB = NTC;
  FftPlan plan;
  int TX;
  __device__ __forceinline__ int NY() const { return NYC; }
  __device__ __forceinline__ int tx() const { return TXC; }
  template <bool INV, int BM, int EM>
  __device__ __forceinline__ void fft(double2 *lds, const double2 *tw, int tid, int) const {
    fft_ct<NYC, INV, true, BM * TXC, 1, EM * TXC, NTC, 1>(lds, tw, tid);
  }
};

}  // namespace fsibe
