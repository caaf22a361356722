// gfx950 FP64 IRLS with the contractions on the matrix cores.
//
// One IRLS round (src/beta.cpp:38-46) over the still-active genes is two
// launches:
//
//   k_irls_accum  one wavefront per (16-gene tile, cell chunk).  For every
//                 16-cell block:
//                   eta' = -X beta      v_mfma_f64_16x16x4f64, D[cell][gene]
//                   w = exp(eta' - off), q = (k + y) w / (1 + k w)      (VALU)
//                   acc += q' K         v_mfma_f64_16x16x4f64 per 16 columns
//                 where K holds the distinct columns x_a x_b (a <= b) and x_a
//                 of the design (the Khatri-Rao product, deduplicated on the
//                 host).  The tile's 16 x NCP partial sums are written per
//                 chunk.
//   k_irls_update one wavefront per active gene: sums its chunk partials in
//                 chunk order, forms A = k X'QX and g = k (X'q - X'1), solves
//                 with LU (partial pivoting), updates beta and applies the
//                 reference stopping rules; unconverged genes are appended to
//                 the next active list.
//
// Chunking is fixed per fit and a gene's row of an MFMA product depends only
// on that gene's operand row, so results do not depend on the active-list
// order or on how many genes share a tile.
//
// v_mfma_f64_16x16x4f64 lane layout (checked on gfx950): lane l supplies
// A[l % 16][l / 16] and B[l / 16][l % 16]; register r of the result holds
// D[l / 16 + 4 r][l % 16].  Within a 16-cell block the rows of the eta
// product are permuted (row m = cell 4 (m % 4) + m / 4) so that lane l ends
// up with the four consecutive cells 4 (l / 16) + r of gene slot l % 16:
// its counts are one 16-byte load and its offsets four consecutive doubles.
#include <hip/hip_runtime.h>

#include "../../include/devil_hip.h"
#include "device_common.hpp"
#include "kernels.hpp"

namespace dv {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef int i4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ d4 mfma16x16x4(double a, double b, d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// exp(x) = 2^n e^r with n = rint(x / ln 2) and r = x - n ln 2 (Cody-Waite,
// two constants), e^r by its degree-12 Taylor polynomial (|r| <= 0.347,
// truncation < 2.5e-16 relative), scaled with v_ldexp_f64 which saturates to
// inf / 0 outside the double range.
__device__ __forceinline__ double exp_f64(double x) {
  const double n = rint(x * 1.4426950408889634);
  double r = fma(-n, 6.93147180369123816490e-01, x);
  r = fma(-n, 1.90821492927058770002e-10, r);
  double p = 1.0 / 479001600.0;
  p = fma(p, r, 1.0 / 39916800.0);
  p = fma(p, r, 1.0 / 3628800.0);
  p = fma(p, r, 1.0 / 362880.0);
  p = fma(p, r, 1.0 / 40320.0);
  p = fma(p, r, 1.0 / 5040.0);
  p = fma(p, r, 1.0 / 720.0);
  p = fma(p, r, 1.0 / 120.0);
  p = fma(p, r, 1.0 / 24.0);
  p = fma(p, r, 1.0 / 6.0);
  p = fma(p, r, 0.5);
  p = fma(p, r, 1.0);
  p = fma(p, r, 1.0);
  return ldexp(p, (int)n);
}

// 1/d from v_rcp_f64 refined by two Newton steps (about 1 ulp).
__device__ __forceinline__ double rcp_f64(double d) {
  double r = __builtin_amdgcn_rcp(d);
  double e = fma(-d, r, 1.0);
  r = fma(r, e, r);
  e = fma(-d, r, 1.0);
  return fma(r, e, r);
}

// Four consecutive counts of a row as doubles.
template <typename CT>
__device__ __forceinline__ void load4(const CT* p, bool vec, double (&y)[4]);

template <>
__device__ __forceinline__ void load4<int32_t>(const int32_t* p, bool vec, double (&y)[4]) {
  if (vec) {
    const i4v v = *reinterpret_cast<const i4v*>(p);
#pragma unroll
    for (int r = 0; r < 4; ++r) y[r] = (double)v[r];
  } else {
#pragma unroll
    for (int r = 0; r < 4; ++r) y[r] = (double)p[r];
  }
}

template <>
__device__ __forceinline__ void load4<float>(const float* p, bool vec, double (&y)[4]) {
  if (vec) {
    const f4v v = *reinterpret_cast<const f4v*>(p);
#pragma unroll
    for (int r = 0; r < 4; ++r) y[r] = (double)v[r];
  } else {
#pragma unroll
    for (int r = 0; r < 4; ++r) y[r] = (double)p[r];
  }
}

template <>
__device__ __forceinline__ void load4<double>(const double* p, bool vec, double (&y)[4]) {
  if (vec) {
    const d4 v = *reinterpret_cast<const d4*>(p);
#pragma unroll
    for (int r = 0; r < 4; ++r) y[r] = v[r];
  } else {
#pragma unroll
    for (int r = 0; r < 4; ++r) y[r] = p[r];
  }
}

// Operands of one 16-cell block for one lane.
template <int NJ, int NTMAX>
struct BlockOps {
  double y[4];
  double xa[NJ];
  d4 off;
  d4 kb[NTMAX];
};

template <int NJ, int NTMAX, typename CT>
__device__ __forceinline__ void load_block(const TileDesign& td, const CT* yrow, bool vec,
                                           int64_t b, int lane, int h, int NT,
                                           BlockOps<NJ, NTMAX>& o) {
  const int64_t c = b * 16 + 4 * h;
  if (c + 4 <= td.n) {
    load4<CT>(yrow + c, vec, o.y);
  } else {
#pragma unroll
    for (int r = 0; r < 4; ++r) o.y[r] = (c + r < td.n) ? (double)yrow[c + r] : 0.0;
  }
#pragma unroll
  for (int j = 0; j < NJ; ++j) o.xa[j] = td.Xp[(b * NJ + j) * 64 + lane];
  o.off = *reinterpret_cast<const d4*>(td.offp + c);
#pragma unroll
  for (int t = 0; t < NTMAX; ++t)
    if (t < NT) o.kb[t] = *reinterpret_cast<const d4*>(td.Kp + ((b * NT + t) * 64 + lane) * 4);
}

// lane l of a wave: gene slot s = l % 16, cell group h = l / 16; its cells in
// block b are 16 b + 4 h + r.
template <int P, typename CT>
__device__ __forceinline__ void irls_tile_chunk(const TileDesign& td, const CT* __restrict__ Y,
                                                int64_t ld, const int32_t* __restrict__ active,
                                                int nA, const double* __restrict__ beta,
                                                const double* __restrict__ kvec, int tile,
                                                int chunk, double* __restrict__ part) {
  constexpr int NJ = (P + 3) / 4;
  constexpr int NTMAX = ((P * (P + 1)) / 2 + P + 15) / 16;
  const int lane = threadIdx.x & 63;
  const int s = lane & 15, h = lane >> 4;
  const int slot = tile * 16 + s;
  const bool live = slot < nA;
  const int g = live ? active[slot] : active[tile * 16];
  const int NT = td.nt;
  // B operand of the eta product: -beta[g][4 j + h]
  double bneg[NJ];
#pragma unroll
  for (int j = 0; j < NJ; ++j) {
    const int f = 4 * j + h;
    bneg[j] = (live && f < P) ? -beta[(int64_t)g * P + f] : 0.0;
  }
  const double k = live ? kvec[g] : 0.0;
  const CT* yrow = Y + (int64_t)g * ld;
  const bool vec = (ld & 3) == 0 && ((reinterpret_cast<uintptr_t>(Y) & 15) == 0);
  d4 acc[NTMAX];
#pragma unroll
  for (int t = 0; t < NTMAX; ++t) acc[t] = d4{0.0, 0.0, 0.0, 0.0};
  const int64_t b0 = (int64_t)chunk * td.chunk_blocks;
  const int64_t b1 = min((int64_t)(chunk + 1) * td.chunk_blocks, td.nblk);
  BlockOps<NJ, NTMAX> cur, nxt;
  load_block<NJ, NTMAX, CT>(td, yrow, vec, b0, lane, h, NT, cur);
  for (int64_t b = b0; b < b1; ++b) {
    if (b + 1 < b1) load_block<NJ, NTMAX, CT>(td, yrow, vec, b + 1, lane, h, NT, nxt);
    d4 eta = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int j = 0; j < NJ; ++j) eta = mfma16x16x4(cur.xa[j], bneg[j], eta);
    double q[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const double w = exp_f64(eta[r] - cur.off[r]);
      const double d = fma(k, w, 1.0);
      q[r] = ((k + cur.y[r]) * w) * rcp_f64(d);
    }
#pragma unroll
    for (int t = 0; t < NTMAX; ++t) {
      if (t < NT) {
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[t] = mfma16x16x4(q[r], cur.kb[t][r], acc[t]);
      }
    }
    cur = nxt;
  }
  // partial sums: register r holds gene slot h + 4 r, column 16 t + s
  double* out = part + ((int64_t)tile * td.nchunk + chunk) * 16 * td.ncp;
#pragma unroll
  for (int t = 0; t < NTMAX; ++t) {
    if (t < NT) {
#pragma unroll
      for (int r = 0; r < 4; ++r) out[(h + 4 * r) * td.ncp + 16 * t + s] = acc[t][r];
    }
  }
}

template <int P>
__global__ __launch_bounds__(256) void k_irls_accum(TileDesign td, Counts c,
                                                    const int32_t* __restrict__ active,
                                                    const int32_t* __restrict__ nactive,
                                                    const double* __restrict__ beta,
                                                    const double* __restrict__ kvec,
                                                    double* __restrict__ part) {
  const int nA = *nactive;
  const int ntile = (nA + 15) / 16;
  const int wave = threadIdx.x >> 6;
  const int chunk = blockIdx.x % td.nchunk;
  const int tile = (blockIdx.x / td.nchunk) * 4 + wave;
  if (tile >= ntile) return;
  if (c.dtype == 0)
    irls_tile_chunk<P>(td, static_cast<const int32_t*>(c.ptr), c.ld, active, nA, beta, kvec,
                       tile, chunk, part);
  else if (c.dtype == 1)
    irls_tile_chunk<P>(td, static_cast<const float*>(c.ptr), c.ld, active, nA, beta, kvec,
                       tile, chunk, part);
  else
    irls_tile_chunk<P>(td, static_cast<const double*>(c.ptr), c.ld, active, nA, beta, kvec,
                       tile, chunk, part);
}

// One wavefront per active gene: lanes sum the chunk partials of columns
// lane and lane + 64 (coalesced, in chunk order) into LDS; every lane then
// runs the same p x p solve.
template <int P>
__global__ __launch_bounds__(64) void k_irls_update(TileDesign td,
                                                    const int32_t* __restrict__ active,
                                                    const int32_t* __restrict__ nactive,
                                                    const double* __restrict__ part,
                                                    double* __restrict__ beta,
                                                    const double* __restrict__ kvec,
                                                    int max_iter, double eps,
                                                    int32_t* __restrict__ iters,
                                                    int32_t* __restrict__ status,
                                                    int32_t* __restrict__ next,
                                                    int32_t* __restrict__ nnext) {
  __shared__ double acc[DEVIL_NCP_MAX];
  const int nA = *nactive;
  const int i = blockIdx.x;
  if (i >= nA) return;
  const int lane = threadIdx.x;
  const int g = active[i];
  const int tile = i >> 4, s = i & 15;
  const double* src = part + (int64_t)tile * td.nchunk * 16 * td.ncp + (int64_t)s * td.ncp;
  for (int col = lane; col < td.ncp; col += 64) {
    double a = 0.0;
    for (int ch = 0; ch < td.nchunk; ++ch) a += src[(int64_t)ch * 16 * td.ncp + col];
    acc[col] = a;
  }
  __syncthreads();
  const double k = kvec[g];
  double M[P][P], v[P];
#pragma unroll
  for (int a = 0; a < P; ++a) {
    const int gi = td.kidx_lin[a];
    v[a] = k * ((gi >= 0 ? acc[gi] : 0.0) - td.sx[a]);
#pragma unroll
    for (int bb = a; bb < P; ++bb) {
      const int pi = td.kidx_pair[tri<P>(a, bb)];
      const double e = k * (pi >= 0 ? acc[pi] : 0.0);
      M[a][bb] = e;
      M[bb][a] = e;
    }
  }
  if (lane != 0) return;
  int st = status[g];
  int perm[P];
  if (lu_factor<P>(M, perm)) st |= DEVIL_GENE_SINGULAR;
  lu_solve<P>(M, perm, v);
  double mx = 0.0;
  bool has_nan = false;
#pragma unroll
  for (int j = 0; j < P; ++j) {
    beta[(int64_t)g * P + j] += v[j];
    const double d = fabs(v[j]);
    if (isnan(d)) has_nan = true;
    else mx = fmax(mx, d);
  }
  bool conv = !has_nan && mx < eps;
  const int it = iters[g] + 1;
  iters[g] = it;
  if (isnan(v[0])) {
    conv = true;
    st |= DEVIL_GENE_NAN;
  }
  if (!conv && it >= max_iter) st |= DEVIL_GENE_MAXITER;
  status[g] = st;
  if (!conv && it < max_iter) next[atomicAdd(nnext, 1)] = g;
}

__global__ void k_iota(int32_t* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = i;
}

#define DV_TILE_P(p, KERNEL, GRID, BLOCK, STREAM, ...)                               \
  switch (p) {                                                                      \
    case 1: hipLaunchKernelGGL(KERNEL<1>, GRID, BLOCK, 0, STREAM, __VA_ARGS__); break; \
    case 2: hipLaunchKernelGGL(KERNEL<2>, GRID, BLOCK, 0, STREAM, __VA_ARGS__); break; \
    case 3: hipLaunchKernelGGL(KERNEL<3>, GRID, BLOCK, 0, STREAM, __VA_ARGS__); break; \
    case 4: hipLaunchKernelGGL(KERNEL<4>, GRID, BLOCK, 0, STREAM, __VA_ARGS__); break; \
    case 5: hipLaunchKernelGGL(KERNEL<5>, GRID, BLOCK, 0, STREAM, __VA_ARGS__); break; \
    case 6: hipLaunchKernelGGL(KERNEL<6>, GRID, BLOCK, 0, STREAM, __VA_ARGS__); break; \
    case 7: hipLaunchKernelGGL(KERNEL<7>, GRID, BLOCK, 0, STREAM, __VA_ARGS__); break; \
    case 8: hipLaunchKernelGGL(KERNEL<8>, GRID, BLOCK, 0, STREAM, __VA_ARGS__); break; \
    case 9: hipLaunchKernelGGL(KERNEL<9>, GRID, BLOCK, 0, STREAM, __VA_ARGS__); break; \
    case 10: hipLaunchKernelGGL(KERNEL<10>, GRID, BLOCK, 0, STREAM, __VA_ARGS__); break; \
    default: return hipErrorInvalidValue;                                           \
  }

hipError_t launch_irls_accum(hipStream_t s, const TileDesign& td, Counts c, int p,
                             const int32_t* active, const int32_t* nactive, int nA_bound,
                             const double* beta, const double* kvec, double* part) {
  if (nA_bound <= 0) return hipSuccess;
  const int ntile = (nA_bound + 15) / 16;
  const int groups = (ntile + 3) / 4;
  const dim3 grid((unsigned)((int64_t)groups * td.nchunk));
  DV_TILE_P(p, k_irls_accum, grid, dim3(256), s, td, c, active, nactive, beta, kvec, part);
  return hipGetLastError();
}

hipError_t launch_irls_update(hipStream_t s, const TileDesign& td, int p,
                              const int32_t* active, const int32_t* nactive, int nA_bound,
                              const double* part, double* beta, const double* kvec,
                              int max_iter, double eps, int32_t* iters, int32_t* status,
                              int32_t* next, int32_t* nnext) {
  if (nA_bound <= 0) return hipSuccess;
  const dim3 grid((unsigned)nA_bound);
  DV_TILE_P(p, k_irls_update, grid, dim3(64), s, td, active, nactive, part, beta, kvec,
            max_iter, eps, iters, status, next, nnext);
  return hipGetLastError();
}

hipError_t launch_iota(hipStream_t s, int32_t* p, int n) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_iota, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n);
  return hipGetLastError();
}

}  // namespace dv
