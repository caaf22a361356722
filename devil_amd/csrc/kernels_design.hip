// Device construction of the MFMA operand tables of a design (the layouts
// documented on dv::TileDesign), one thread per (16- or 32-cell block, lane).
// K columns are described by ColSpec: a < 0 a zero column, b < 0 the design
// column x_a, otherwise the product x_a x_b.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace dv {

__device__ __forceinline__ double kcol(const double* __restrict__ X, int64_t n, ColSpec c,
                                       int64_t cell) {
  if (c.a < 0 || cell >= n) return 0.0;
  const double xa = X[cell + c.a * n];
  return c.b < 0 ? xa : xa * X[cell + c.b * n];
}

__device__ __forceinline__ uint16_t bf16_rne(float x) {
  uint32_t u = __float_as_uint(x);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

// FP64 design (16-cell blocks): Xp rows permuted (row m = cell 4 (m%4) + m/4),
// Kp, offp and eop; with f32 != 0 also the FP32 copies (natural row order).
__global__ void k_build_tables(const double* __restrict__ X, const double* __restrict__ off,
                               const double* __restrict__ eo, int64_t n, int p,
                               const ColSpec* __restrict__ spec, int ncol, int nt,
                               int64_t nblk, double* __restrict__ Xp, double* __restrict__ offp,
                               double* __restrict__ eop, double* __restrict__ Kp, int f32,
                               float* __restrict__ Xf, float* __restrict__ offf,
                               float* __restrict__ Kf) {
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= nblk * 64) return;
  const int64_t b = id / 64;
  const int l = (int)(id % 64);
  const int NJ = (p + 3) / 4;
  for (int j = 0; j < NJ; ++j) {
    const int f = 4 * j + l / 16;
    const int m = l % 16;
    const int64_t cp = 16 * b + 4 * (m % 4) + m / 4;
    const int64_t cn = 16 * b + m;
    Xp[(b * NJ + j) * 64 + l] = (f < p && cp < n) ? X[cp + f * n] : 0.0;
    if (f32) Xf[(b * NJ + j) * 64 + l] = (f < p && cn < n) ? (float)X[cn + f * n] : 0.0f;
  }
  if (l < 16) {
    const int64_t c = 16 * b + l;
    offp[c] = c < n ? off[c] : 0.0;
    eop[c] = c < n ? eo[c] : 0.0;
    if (f32) offf[c] = c < n ? (float)off[c] : 0.0f;
  }
  for (int t = 0; t < nt; ++t) {
    const int col = 16 * t + l % 16;
    const ColSpec cs = col < ncol ? spec[col] : ColSpec{-1, -1};
    for (int r = 0; r < 4; ++r) {
      const int64_t cell = 16 * b + 4 * (l / 16) + r;
      const double v = kcol(X, n, cs, cell);
      Kp[((b * nt + t) * 64 + l) * 4 + r] = v;
      if (f32) Kf[((b * nt + t) * 64 + l) * 4 + r] = (float)v;
    }
  }
}

// bf16 mixed design (32-cell steps): xab, offl, kg, kh (spec of the lin16
// column layout: tile 0 = the p design columns, tiles 1.. = pair columns).
__global__ void k_build_tables_bf16(const double* __restrict__ X, const double* __restrict__ off,
                                    int64_t n, int p, const ColSpec* __restrict__ spec,
                                    int ncol, int nth, int64_t nst, uint16_t* __restrict__ xab,
                                    float* __restrict__ offl, uint16_t* __restrict__ kg,
                                    uint16_t* __restrict__ kh, uint16_t* __restrict__ kh2) {
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= nst * 64) return;
  const int64_t st = id / 64;
  const int l = (int)(id % 64);
  const int NE = eta_blocks(p);
  const double L2E = 1.44269504088896340736;
  for (int u = 0; u < 2; ++u)
    for (int j = 0; j < NE; ++j)
      for (int i = 0; i < 8; ++i) {
        const int64_t cell = st * 32 + 16 * u + l % 16;
        const int e = 32 * j + 8 * (l / 16) + i;
        const int f = e / kEtaTerms, term = e % kEtaTerms;
        uint16_t v = 0;
        if (cell < n && f < p) {
          double rem = X[cell + f * n];
          for (int q = 0; q <= eta_x_piece(term); ++q) {  // x = x1 + x2 + x3
            v = bf16_rne((float)rem);
            rem -= (double)__uint_as_float((uint32_t)v << 16);
          }
        }
        xab[(((st * 2 + u) * NE + j) * 64 + l) * 8 + i] = v;
      }
  if (l < 32) {
    const int64_t c = st * 32 + l;
    offl[c] = c < n ? (float)(off[c] * L2E) : 0.0f;
  }
  for (int i = 0; i < 8; ++i) {
    const int64_t cell = st * 32 + (i < 4 ? 4 * (l / 16) + i : 16 + 4 * (l / 16) + (i - 4));
    const int col = l % 16;
    double x = (cell < n && col < p) ? X[cell + col * n] : 0.0;
    for (int q = 0; q < 3; ++q) {  // x = x1 + x2 + x3 (each rounded from the rest)
      const uint16_t xq = bf16_rne((float)x);
      kg[((st * 3 + q) * 64 + l) * 8 + i] = xq;
      x -= (double)__uint_as_float((uint32_t)xq << 16);
    }
    for (int t = 0; t < nth; ++t) {
      const int kc = 16 * (t + 1) + l % 16;
      const ColSpec cs = kc < ncol ? spec[kc] : ColSpec{-1, -1};
      const double v = kcol(X, n, cs, cell);
      const uint16_t v1 = bf16_rne((float)v);
      kh[((st * nth + t) * 64 + l) * 8 + i] = v1;
      kh2[((st * nth + t) * 64 + l) * 8 + i] =
          bf16_rne((float)(v - (double)__uint_as_float((uint32_t)v1 << 16)));
    }
  }
}

// Init-pass operands: Zp[blk][lane][r] = M[16 blk + 4 (lane/16) + r][lane%16]
// - shift[lane%16] (M n x p column-major, shift may be null), and the
// per-cell 1 / exp(offset) and log(sf) padded to 16 nblk.
__global__ void k_build_init(const double* __restrict__ M, const double* __restrict__ shift,
                             const double* __restrict__ eo, const double* __restrict__ logsf,
                             int64_t n, int p, int64_t nblk, double* __restrict__ Zp,
                             double* __restrict__ einv, double* __restrict__ logsfp) {
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= nblk * 64) return;
  const int64_t b = id / 64;
  const int l = (int)(id % 64);
  if (Zp) {
    const int col = l % 16;
    for (int r = 0; r < 4; ++r) {
      const int64_t cell = 16 * b + 4 * (l / 16) + r;
      double v = 0.0;
      if (cell < n && col < p) v = M[cell + col * n] - (shift ? shift[col] : 0.0);
      Zp[(b * 64 + l) * 4 + r] = v;
    }
  }
  if (l < 16) {
    const int64_t c = 16 * b + l;
    einv[c] = c < n ? 1.0 / eo[c] : 0.0;
    if (logsfp) logsfp[c] = c < n ? logsf[c] : 0.0;
  }
}

// ---------------------------------------------------------------------------
// Column statistics and QR of the design on the device (deterministic:
// per-block partials in fixed order, combined on the host).
// ---------------------------------------------------------------------------
constexpr int kStatBlocks = 64;

template <typename T>
__device__ __forceinline__ T block_sum256(T v, T* sh) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (t < o) sh[t] += sh[t + o];
    __syncthreads();
  }
  const T r = sh[0];
  __syncthreads();
  return r;
}

// grid (kStatBlocks, C): per candidate column c and cell range x: the sum of
// mixed (bits, cell) words (an order-free hash), whether any value is
// non-zero, and the sum of the values.
__global__ __launch_bounds__(256) void k_col_stats(const double* __restrict__ X, int64_t n,
                                                   const ColSpec* __restrict__ cand,
                                                   uint64_t* __restrict__ ph,
                                                   uint64_t* __restrict__ pnz,
                                                   double* __restrict__ ps) {
  __shared__ uint64_t shu[256];
  __shared__ double shd[256];
  const int c = blockIdx.y;
  const int x = blockIdx.x;
  const ColSpec cs = cand[c];
  const int64_t lo = n * x / kStatBlocks, hi = n * (x + 1) / kStatBlocks;
  uint64_t h = 0, nz = 0;
  double su = 0.0;
  for (int64_t i = lo + threadIdx.x; i < hi; i += 256) {
    const double v = kcol(X, n, cs, i);
    const uint64_t u = (uint64_t)__double_as_longlong(v);
    const uint64_t k = (u ^ ((uint64_t)i * 0x9E3779B97F4A7C15ULL)) * 0xBF58476D1CE4E5B9ULL;
    h += k ^ (k >> 31);
    nz |= (v != 0.0);
    su += v;
  }
  h = block_sum256<uint64_t>(h, shu);
  nz = block_sum256<uint64_t>(nz, shu);
  su = block_sum256<double>(su, shd);
  if (threadIdx.x == 0) {
    ph[c * kStatBlocks + x] = h;
    pnz[c * kStatBlocks + x] = nz;
    ps[c * kStatBlocks + x] = su;
  }
}

// per block: number of cells where columns c1 and c2 differ
__global__ __launch_bounds__(256) void k_col_mismatch(const double* __restrict__ X, int64_t n,
                                                      ColSpec c1, ColSpec c2,
                                                      uint64_t* __restrict__ out) {
  __shared__ uint64_t shu[256];
  const int x = blockIdx.x;
  const int64_t lo = n * x / kStatBlocks, hi = n * (x + 1) / kStatBlocks;
  uint64_t m = 0;
  for (int64_t i = lo + threadIdx.x; i < hi; i += 256) m += kcol(X, n, c1, i) != kcol(X, n, c2, i);
  m = block_sum256<uint64_t>(m, shu);
  if (threadIdx.x == 0) out[x] = m;
}

// per block: the packed upper triangle of M'M over its cell range
__global__ __launch_bounds__(256) void k_gram(const double* __restrict__ M, int64_t n, int p,
                                              double* __restrict__ out) {
  __shared__ double shd[256];
  const int x = blockIdx.x;
  const int64_t lo = n * x / kStatBlocks, hi = n * (x + 1) / kStatBlocks;
  int q = 0;
  for (int a = 0; a < p; ++a)
    for (int b = a; b < p; ++b, ++q) {
      double su = 0.0;
      for (int64_t i = lo + threadIdx.x; i < hi; i += 256) su += M[i + a * n] * M[i + b * n];
      su = block_sum256<double>(su, shd);
      if (threadIdx.x == 0) out[x * 55 + q] = su;
    }
}

// M <- M R^-1 row by row (R p x p upper, column-major)
__global__ void k_right_solve(double* __restrict__ M, int64_t n, int p,
                              const double* __restrict__ R) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double row[10];
  for (int j = 0; j < p; ++j) {
    double s = M[i + j * n];
    for (int k = 0; k < j; ++k) s -= row[k] * R[k + j * p];
    row[j] = s / R[j + j * p];
  }
  for (int j = 0; j < p; ++j) M[i + j * n] = row[j];
}

hipError_t launch_col_stats(hipStream_t s, const double* X, int64_t n, const ColSpec* cand,
                            int C, uint64_t* ph, uint64_t* pnz, double* ps) {
  hipLaunchKernelGGL(k_col_stats, dim3(kStatBlocks, C), dim3(256), 0, s, X, n, cand, ph, pnz, ps);
  return hipGetLastError();
}

hipError_t launch_col_mismatch(hipStream_t s, const double* X, int64_t n, ColSpec c1, ColSpec c2,
                               uint64_t* out) {
  hipLaunchKernelGGL(k_col_mismatch, dim3(kStatBlocks), dim3(256), 0, s, X, n, c1, c2, out);
  return hipGetLastError();
}

hipError_t launch_gram(hipStream_t s, const double* M, int64_t n, int p, double* out) {
  hipLaunchKernelGGL(k_gram, dim3(kStatBlocks), dim3(256), 0, s, M, n, p, out);
  return hipGetLastError();
}

hipError_t launch_right_solve(hipStream_t s, double* M, int64_t n, int p, const double* R) {
  hipLaunchKernelGGL(k_right_solve, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, M, n, p,
                     R);
  return hipGetLastError();
}

hipError_t launch_build_init(hipStream_t s, const double* M, const double* shift,
                             const double* eo, const double* logsf, int64_t n, int p,
                             int64_t nblk, double* Zp, double* einv, double* logsfp) {
  const int64_t threads = nblk * 64;
  hipLaunchKernelGGL(k_build_init, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, s, M,
                     shift, eo, logsf, n, p, nblk, Zp, einv, logsfp);
  return hipGetLastError();
}

hipError_t launch_build_tables(hipStream_t s, const double* X, const double* off,
                               const double* eo, int64_t n, int p, const ColSpec* spec, int ncol,
                               int nt, int64_t nblk, double* Xp, double* offp, double* eop,
                               double* Kp, int f32, float* Xf, float* offf, float* Kf) {
  const int64_t threads = nblk * 64;
  hipLaunchKernelGGL(k_build_tables, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, s, X,
                     off, eo, n, p, spec, ncol, nt, nblk, Xp, offp, eop, Kp, f32, Xf, offf, Kf);
  return hipGetLastError();
}

hipError_t launch_build_tables_bf16(hipStream_t s, const double* X, const double* off, int64_t n,
                                    int p, const ColSpec* spec, int ncol, int nth, int64_t nst,
                                    uint16_t* xab, float* offl, uint16_t* kg, uint16_t* kh,
                                    uint16_t* kh2) {
  const int64_t threads = nst * 64;
  hipLaunchKernelGGL(k_build_tables_bf16, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0,
                     s, X, off, n, p, spec, ncol, nth, nst, xab, offl, kg, kh, kh2);
  return hipGetLastError();
}

}  // namespace dv
