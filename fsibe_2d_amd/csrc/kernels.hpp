// kernels.hpp -- HIP kernels of the FSIBE_2D fluid step for gfx950.
//
// Layouts in HBM
//   real field     double  f[ny][nx]        (reference f(0:nx-1,0:ny-1))
//   packed field   double  fk[ny][nx]       (reference coftxy layout)
//   half spectrum  double2 c[ny][H], H=nx/2 (complex coefficient of mode
//                  (kx, ky) at c[ky mod ny][kx]; kx=nx/2 and ky=ny/2 dropped)
// The packed layout is only used at the Tier-A boundary; the step keeps its
// spectral state as half spectra, where every spectral operator of the
// reference (cofdx, cofdy, poisson, vis, dealiase) is a per-mode complex
// scale: cofdx multiplies c(kx, +-ky) by i kx 2pi/xl, cofdy by i (+-ky) 2pi/yl.
#pragma once
#include <hip/hip_runtime.h>

#include "fft_device.hpp"

namespace fsibe {

constexpr int kNT = 256;  // threads per workgroup (4 wavefronts)

// signed y wavenumber of half-spectrum row idx; the Nyquist row maps to 0
__device__ __forceinline__ int ky_signed(int idx, int ny) {
  const int h = ny >> 1;
  return idx < h ? idx : (idx > h ? idx - ny : 0);
}

// XCD-aware block index: the 8 XCDs receive workgroups round-robin, so
// consecutive logical tiles (which share 128-byte lines of the column-major
// access) are given to blocks that run on the same XCD and its L2.
__device__ __forceinline__ int xcd_remap(int bid, int nblk) {
  if ((nblk & 7) != 0) return bid;
  const int per = nblk >> 3;
  return (bid & 7) * per + (bid >> 3);
}

// ---- real <-> half-length complex conversions (in place, per row) -------------
// Forward: Z = FFT_H(x[2n] + i x[2n+1]) -> X[k] = 1/2 (Z_k + conj Z_{H-k})
//          - i/2 e^{-2 pi i k/nx} (Z_k - conj Z_{H-k}), k < H (X[H] dropped).
// (fused multiply-adds explicit, as in cmul)
__device__ __forceinline__ double2 r2c_one(double2 a, double2 b, double2 w) {
  const double2 bc = cconj(b);
  const double2 e = make_double2(0.5 * (a.x + bc.x), 0.5 * (a.y + bc.y));
  const double2 d = make_double2(0.5 * (a.x - bc.x), 0.5 * (a.y - bc.y));
  const double pr = fma(w.x, d.x, -(w.y * d.y)), pi = fma(w.x, d.y, w.y * d.x);
  return make_double2(e.x + pi, e.y - pr);
}

// Inverse (unnormalised c2r, X[H] = 0, Im X[0] ignored):
// Z_k = (X_k + conj X_{H-k}) + i e^{+2 pi i k/nx} (X_k - conj X_{H-k}).
__device__ __forceinline__ double2 c2r_one(double2 a, double2 b, double2 wconj) {
  const double2 bc = cconj(b);
  const double2 e = make_double2(a.x + bc.x, a.y + bc.y);
  const double2 d = make_double2(a.x - bc.x, a.y - bc.y);
  const double pr = fma(wconj.x, d.x, -(wconj.y * d.y)), pi = fma(wconj.x, d.y, wconj.y * d.x);
  return make_double2(e.x - pi, e.y + pr);
}

// Launch bounds: up to 1024 threads when at most 8 elements per thread are
// held in registers during a pass, 256 threads otherwise.
#define FSIBE_LB(EMAX) __launch_bounds__((EMAX) <= 8 ? 1024 : 256)

// rows: `nrows` sequences of length H, sequence r at buf + r*PL (PL = padlen(H))
__device__ void rows_r2c_post(double2 *buf, int nrows, int H, int PL, const double2 *__restrict__ twx,
                              int tid, int nthr) {
  const int hp = (H >> 1) + 1;
  for (int e = tid; e < nrows * hp; e += nthr) {
    const int r = e / hp, k = e - r * hp;
    double2 *Z = buf + r * PL;
    const int q = (k == 0) ? 0 : H - k;
    const double2 a = Z[pidx(k)], b = Z[pidx(q)];
    const double2 xk = r2c_one(a, b, twx[k]);
    if (q != k) Z[pidx(q)] = r2c_one(b, a, twx[q]);
    Z[pidx(k)] = xk;
  }
  __syncthreads();
}

__device__ void rows_c2r_pre(double2 *buf, int nrows, int H, int PL, const double2 *__restrict__ twx,
                             int tid, int nthr) {
  const int hp = (H >> 1) + 1;
  for (int e = tid; e < nrows * hp; e += nthr) {
    const int r = e / hp, k = e - r * hp;
    double2 *X = buf + r * PL;
    if (k == 0) {
      const double a = X[0].x;  // x-DC: real part only, x-Nyquist is zero
      X[0] = make_double2(a, a);
    } else {
      const int q = H - k;
      const double2 a = X[pidx(k)], b = X[pidx(q)];
      const double2 zk = c2r_one(a, b, cconj(twx[k]));
      if (q != k) X[pidx(q)] = c2r_one(b, a, cconj(twx[q]));
      X[pidx(k)] = zk;
    }
  }
  __syncthreads();
}

// rows_c2r_pre fused into the staging of the rows: each thread takes the
// pair (k, H-k) of every row, reads both half-spectrum entries through
// val(f, k) (global loads and the field's own spectral factor) and writes
// the pair's pre-processed values, as rows_c2r_pre computes them, to row f
// at buf + f*PL.  One LDS write per element instead of a write, a read and
// a write, and one barrier fewer.
template <int NFMAX, class Val>
__device__ __forceinline__ void stage_rows_c2r(double2 *buf, int nf, int H, int PL, const double2 *__restrict__ twx,
                                               int tid, int nthr, Val val) {
  const int hp = (H >> 1) + 1;
  for (int k = tid; k < hp; k += nthr) {
    if (k == 0) {
#pragma unroll
      for (int f = 0; f < NFMAX; ++f)
        if (f < nf) {
          const double a = val(f, 0).x;  // x-DC: real part only, x-Nyquist is zero
          buf[f * PL] = make_double2(a, a);
        }
      continue;
    }
    const int q = H - k;
    const double2 wk = cconj(twx[k]), wq = cconj(twx[q]);
#pragma unroll
    for (int f = 0; f < NFMAX; ++f)
      if (f < nf) {
        const double2 a = val(f, k), b = q != k ? val(f, q) : a;
        double2 *X = buf + f * PL;
        if (q != k) X[pidx(q)] = c2r_one(b, a, wq);
        X[pidx(k)] = c2r_one(a, b, wk);
      }
  }
  __syncthreads();
}

// rows_r2c_post fused into the final stores: each thread takes the pair
// (k, H-k) of the NF rows, forms their spectra as rows_r2c_post does and
// hands them to out(k, x[NF]) and out(H-k, x[NF]) (once for k = 0 and
// k = H/2).  No LDS writes and no barrier after the last pass.
template <int NF, class Out>
__device__ __forceinline__ void rows_r2c_store(const double2 *buf, int H, int PL, const double2 *__restrict__ twx,
                                               int tid, int nthr, Out out) {
  const int hp = (H >> 1) + 1;
  for (int k = tid; k < hp; k += nthr) {
    const int q = (k == 0) ? 0 : H - k;
    const double2 wk = twx[k], wq = twx[q];
    double2 xk[NF], xq[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      const double2 a = buf[f * PL + pidx(k)], b = buf[f * PL + pidx(q)];
      xk[f] = r2c_one(a, b, wk);
      xq[f] = r2c_one(b, a, wq);
    }
    out(k, xk);
    if (q != k) out(q, xq);
  }
}

// ============================================================================
// Tier A kernels (packed layout at the boundary)
// ============================================================================
// real rows -> half spectra (unnormalised), one row per block
template <class FX>
__global__ void __launch_bounds__(FX::LB) k_rows_r2c(const double *__restrict__ f, double2 *__restrict__ out, FX F,
                                                     const double2 *__restrict__ twx) {
  extern __shared__ double2 lds[];
  const int H = F.H(), nx = 2 * H, tid = threadIdx.x, nt = blockDim.x;
  const int row = blockIdx.x;
  for (int n = tid; n < H; n += nt) lds[pidx(n)] = reinterpret_cast<const double2 *>(f + (size_t)row * nx)[n];
  __syncthreads();
  F.template fft<false, 1>(lds, twx, tid, nt);
  rows_r2c_post(lds, 1, H, padlen(H), twx, tid, nt);
  for (int n = tid; n < H; n += nt) out[(size_t)row * H + n] = lds[pidx(n)];
}

template <class FX>
__global__ void __launch_bounds__(FX::LB) k_rows_c2r(const double2 *__restrict__ in, double *__restrict__ f, FX F,
                                                     const double2 *__restrict__ twx) {
  extern __shared__ double2 lds[];
  const int H = F.H(), nx = 2 * H, tid = threadIdx.x, nt = blockDim.x;
  const int row = blockIdx.x;
  for (int n = tid; n < H; n += nt) lds[pidx(n)] = in[(size_t)row * H + n];
  __syncthreads();
  rows_c2r_pre(lds, 1, H, padlen(H), twx, tid, nt);
  F.template fft<true, 1>(lds, twx, tid, nt);
  for (int n = tid; n < H; n += nt) reinterpret_cast<double2 *>(f + (size_t)row * nx)[n] = lds[pidx(n)];
}

// half spectra -> column FFT -> x norm -> packed layout (cof_fftw33.f90:165-189)
// column tile layout: element y of column c at lds[pidx(y)*TX + c]
template <class FY>
__global__ void __launch_bounds__(FY::LB) k_cols_fwd_pack(const double2 *__restrict__ in, double *__restrict__ fk,
                                                          FY F, const double2 *__restrict__ twy, int H,
                                                          double norm) {
  extern __shared__ double2 lds[];
  const int ny = F.NY(), TX = F.tx(), nx = 2 * H, tid = threadIdx.x, nt = blockDim.x;
  const int kx0 = xcd_remap(blockIdx.x, gridDim.x) * TX;
  for (int e = tid; e < ny * TX; e += nt) {
    const int y = e / TX, c = e - y * TX;
    lds[pidx(y) * TX + c] = in[(size_t)y * H + kx0 + c];
  }
  __syncthreads();
  F.template fft<false, 1, 1>(lds, twy, tid, nt);
  const int hy = ny >> 1;
  for (int e = tid; e < hy * TX; e += nt) {
    const int ky = e / TX, c = e - ky * TX, kx = kx0 + c;
    double2 r0, r1;
    if (ky == 0) {
      const double2 v = lds[c];
      r0 = make_double2(v.x * norm, v.y * norm);
      r1 = make_double2(0.0, 0.0);
    } else {
      const double2 p = lds[pidx(ky) * TX + c], m = lds[pidx(ny - ky) * TX + c];
      r0 = make_double2(0.5 * (p.x + m.x) * norm, 0.5 * (p.y + m.y) * norm);
      r1 = make_double2(0.5 * (p.y - m.y) * norm, 0.5 * (-p.x + m.x) * norm);
    }
    reinterpret_cast<double2 *>(fk + (size_t)(2 * ky) * nx)[kx] = r0;
    reinterpret_cast<double2 *>(fk + (size_t)(2 * ky + 1) * nx)[kx] = r1;
  }
}

// packed layout -> half spectra (cof_fftw33.f90:211-231) -> inverse column FFT
template <class FY>
__global__ void __launch_bounds__(FY::LB) k_cols_unpack_inv(const double *__restrict__ fk, double2 *__restrict__ out,
                                                            FY F, const double2 *__restrict__ twy, int H) {
  extern __shared__ double2 lds[];
  const int ny = F.NY(), TX = F.tx(), nx = 2 * H, tid = threadIdx.x, nt = blockDim.x;
  const int kx0 = xcd_remap(blockIdx.x, gridDim.x) * TX;
  const int hy = ny >> 1;
  for (int e = tid; e < hy * TX; e += nt) {
    const int ky = e / TX, c = e - ky * TX, kx = kx0 + c;
    const double2 ab = reinterpret_cast<const double2 *>(fk + (size_t)(2 * ky) * nx)[kx];
    const double2 cd = reinterpret_cast<const double2 *>(fk + (size_t)(2 * ky + 1) * nx)[kx];
    if (ky == 0) {
      lds[c] = ab;
      lds[pidx(hy) * TX + c] = make_double2(0.0, 0.0);
    } else {
      lds[pidx(ky) * TX + c] = make_double2(ab.x - cd.y, cd.x + ab.y);
      lds[pidx(ny - ky) * TX + c] = make_double2(ab.x + cd.y, -cd.x + ab.y);
    }
  }
  __syncthreads();
  F.template fft<true, 1, 1>(lds, twy, tid, nt);
  for (int e = tid; e < ny * TX; e += nt) {
    const int y = e / TX, c = e - y * TX;
    out[(size_t)y * H + kx0 + c] = lds[pidx(y) * TX + c];
  }
}

// packed-layout operators, one thread per (even, odd) pair of x -------------
struct PackedDims {
  int nx, ny;
  double sx, sy;          // 2 pi / xl, 2 pi / yl
  double scalex, scaley;  // (2 pi / L)^2 as in params.f90:140-141
};

// cofdx.f90:19-26
__global__ void k_cofdx(const double *__restrict__ fk, double *__restrict__ out, PackedDims d) {
#pragma clang fp contract(off)
  const int hx = d.nx >> 1;
  const size_t n = (size_t)hx * d.ny;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
    const int kx = (int)(e % hx);
    const double2 v = reinterpret_cast<const double2 *>(fk)[e];
    const double k = (double)kx;
    reinterpret_cast<double2 *>(out)[e] = make_double2(-v.y * k * d.sx, v.x * k * d.sx);
  }
}

// cofdy.f90:19-27: pairs along y
__global__ void k_cofdy(const double *__restrict__ fk, double *__restrict__ out, PackedDims d) {
#pragma clang fp contract(off)
  const int hy = d.ny >> 1;
  const size_t n = (size_t)d.nx * hy;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
    const int x = (int)(e % d.nx), l = (int)(e / d.nx);
    const double a = fk[(size_t)(2 * l) * d.nx + x], b = fk[(size_t)(2 * l + 1) * d.nx + x];
    const double k = (double)l;
    out[(size_t)(2 * l) * d.nx + x] = -b * k * d.sy;
    out[(size_t)(2 * l + 1) * d.nx + x] = a * k * d.sy;
  }
}

// cofdxdx.f90, cofdydy.f90, cofdxdy.f90, poisson.f90: per 2x2 block
// op 2: -kx^2 sx^2, 3: -ky^2 sy^2, 4: d/dxdy, 5: poisson
__global__ void k_block_op(const double *__restrict__ fk, double *__restrict__ out, PackedDims d, int op) {
#pragma clang fp contract(off)
  const int hx = d.nx >> 1, hy = d.ny >> 1;
  const size_t n = (size_t)hx * hy;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
    const int kx = (int)(e % hx), ky = (int)(e / hx);
    const double2 r0 = reinterpret_cast<const double2 *>(fk + (size_t)(2 * ky) * d.nx)[kx];
    const double2 r1 = reinterpret_cast<const double2 *>(fk + (size_t)(2 * ky + 1) * d.nx)[kx];
    double2 o0, o1;
    if (op == 2) {
      const double k = (double)kx, k2 = k * k, sc = d.sx * d.sx;
      o0 = make_double2(-r0.x * k2 * sc, -r0.y * k2 * sc);
      o1 = make_double2(-r1.x * k2 * sc, -r1.y * k2 * sc);
    } else if (op == 3) {
      const double k = (double)ky, k2 = k * k, sc = d.sy * d.sy;
      o0 = make_double2(-r0.x * k2 * sc, -r0.y * k2 * sc);
      o1 = make_double2(-r1.x * k2 * sc, -r1.y * k2 * sc);
    } else if (op == 4) {
      const double fac = -(double)kx * (double)ky * (d.sx * d.sy);
      // out(2kx,2ky)=fac*f(2kx+1,2ky+1); out(2kx+1,2ky)=-fac*f(2kx,2ky+1)
      // out(2kx,2ky+1)=-fac*f(2kx+1,2ky); out(2kx+1,2ky+1)=fac*f(2kx,2ky)
      o0 = make_double2(fac * r1.y, -fac * r1.x);
      o1 = make_double2(-fac * r0.y, fac * r0.x);
    } else {
      if (kx == 0 && ky == 0) {
        o0 = make_double2(0.0, 0.0);
        o1 = o0;
      } else {
        const double quot = ((double)(kx * kx) * d.scalex + (double)(ky * ky) * d.scaley);
        o0 = make_double2(r0.x / quot, r0.y / quot);
        o1 = make_double2(r1.x / quot, r1.y / quot);
      }
    }
    reinterpret_cast<double2 *>(out + (size_t)(2 * ky) * d.nx)[kx] = o0;
    reinterpret_cast<double2 *>(out + (size_t)(2 * ky + 1) * d.nx)[kx] = o1;
  }
}

// cal_vis.f90:16-33
__global__ void k_vis(double *__restrict__ work, PackedDims d, double coefx, double coefy) {
#pragma clang fp contract(off)
  const int hx = d.nx >> 1, hy = d.ny >> 1;
  const size_t n = (size_t)hx * hy;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
    const int kx = (int)(e % hx), ky = (int)(e / hx);
    const double v = exp(coefx * (double)(kx * kx) + coefy * (double)(ky * ky));
    reinterpret_cast<double2 *>(work + (size_t)(2 * ky) * d.nx)[kx] = make_double2(v, v);
    reinterpret_cast<double2 *>(work + (size_t)(2 * ky + 1) * d.nx)[kx] = make_double2(v, v);
  }
}

// packed <-> half spectrum conversions (no normalisation) ---------------------
__global__ void k_packed_to_half(const double *__restrict__ fk, double2 *__restrict__ c, int nx, int ny) {
#pragma clang fp contract(off)
  const int H = nx >> 1, hy = ny >> 1;
  const size_t n = (size_t)H * hy;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
    const int kx = (int)(e % H), ky = (int)(e / H);
    const double2 ab = reinterpret_cast<const double2 *>(fk + (size_t)(2 * ky) * nx)[kx];
    const double2 cd = reinterpret_cast<const double2 *>(fk + (size_t)(2 * ky + 1) * nx)[kx];
    if (ky == 0) {
      c[kx] = ab;
      c[(size_t)hy * H + kx] = make_double2(0.0, 0.0);
    } else {
      c[(size_t)ky * H + kx] = make_double2(ab.x - cd.y, cd.x + ab.y);
      c[(size_t)(ny - ky) * H + kx] = make_double2(ab.x + cd.y, -cd.x + ab.y);
    }
  }
}

__global__ void k_half_to_packed(const double2 *__restrict__ c, double *__restrict__ fk, int nx, int ny) {
#pragma clang fp contract(off)
  const int H = nx >> 1, hy = ny >> 1;
  const size_t n = (size_t)H * hy;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
    const int kx = (int)(e % H), ky = (int)(e / H);
    double2 r0, r1;
    if (ky == 0) {
      r0 = c[kx];
      r1 = make_double2(0.0, 0.0);
    } else {
      const double2 p = c[(size_t)ky * H + kx], m = c[(size_t)(ny - ky) * H + kx];
      r0 = make_double2(0.5 * (p.x + m.x), 0.5 * (p.y + m.y));
      r1 = make_double2(0.5 * (p.y - m.y), 0.5 * (-p.x + m.x));
    }
    reinterpret_cast<double2 *>(fk + (size_t)(2 * ky) * nx)[kx] = r0;
    reinterpret_cast<double2 *>(fk + (size_t)(2 * ky + 1) * nx)[kx] = r1;
  }
}

// Copy n 16-byte words of the per-step staging record from pinned, mapped
// host memory to device memory.  A kernel on the step's own stream instead of
// a DMA copy: the step's next kernel then follows without a hand-off between
// the copy engine and the compute queue.
__global__ void k_stage_copy(const double2 *__restrict__ src, double2 *__restrict__ dst, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) dst[i] = src[i];
}

// The same copy, queued before the host has written src: each workgroup
// waits until the host stores a value >= seq into gate[0] (host memory),
// then copies.  The wait is bounded: after max_ticks of the constant-rate
// wall clock it stores seq into gate[1], which the host checks after the
// step, and copies anyway.
__global__ void k_stage_copy_gated(const double2 *__restrict__ src, double2 *__restrict__ dst, int n,
                                   long long *gate, long long seq, long long max_ticks) {
  if (threadIdx.x == 0) {
    const long long t0 = wall_clock64();
    while (__hip_atomic_load(gate, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM) < seq) {
      if (wall_clock64() - t0 > max_ticks) {
        __hip_atomic_store(gate + 1, seq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
  }
  __syncthreads();
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) dst[i] = src[i];
}

// Several staging copies in one launch: segment blockIdx.y copies n 16-byte
// words from src (mapped host memory) to dst (device memory).
struct StageSeg {
  const double2 *src;
  double2 *dst;
  long long n;
};
constexpr int kMaxStageSegs = 128;
struct StageSegs {
  StageSeg s[kMaxStageSegs];
};
__global__ void k_stage_gather(StageSegs segs) {
  const StageSeg g = segs.s[blockIdx.y];
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < g.n; i += (long long)gridDim.x * blockDim.x)
    g.dst[i] = g.src[i];
}

}  // namespace fsibe
