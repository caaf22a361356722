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
#include "tile_common.hpp"

namespace dv {

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
      q[r] = ((k + cur.y[r]) * w) * rcp1_f64(d);
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

// ---------------------------------------------------------------------------
// Mixed precision accumulate.  Same tile / chunk / partial-sum structure; per
// element in FP32: mu = 2^((x'beta + off) log2 e) (v_exp_f32),
// q = (k + y) / (k + mu) (v_rcp_f32) -- the reference's m w written with mu,
// which keeps the right limits where FP32 mu under- or overflows -- and
// q' K on v_mfma_f32_16x16x4f32 with FP32 accumulators flushed into FP64
// every 16 blocks (256 cells).
// v_mfma_f32_16x16x4f32 (checked on gfx950): lane l supplies A[l%16][l/16]
// and B[l/16][l%16]; register r of the result holds D[4 (l/16) + r][l%16].
// ---------------------------------------------------------------------------
typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f4 mfma32(float a, float b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

template <int NJ, int NTMAX>
struct BlockOpsF {
  float y[4];
  float xa[NJ];
  f4 off;
  f4 kb[NTMAX];
};

template <typename CT>
__device__ __forceinline__ void load4f(const CT* p, bool vec, float (&y)[4]) {
  double t[4];
  load4<CT>(p, vec, t);
#pragma unroll
  for (int r = 0; r < 4; ++r) y[r] = (float)t[r];
}

template <int NJ, int NTMAX, typename CT>
__device__ __forceinline__ void load_block_f(const TileDesign& td, const CT* yrow, bool vec,
                                             int64_t b, int lane, int h, int NT,
                                             BlockOpsF<NJ, NTMAX>& o) {
  const int64_t c = b * 16 + 4 * h;
  if (c + 4 <= td.n) {
    load4f<CT>(yrow + c, vec, o.y);
  } else {
#pragma unroll
    for (int r = 0; r < 4; ++r) o.y[r] = (c + r < td.n) ? (float)yrow[c + r] : 0.0f;
  }
#pragma unroll
  for (int j = 0; j < NJ; ++j) o.xa[j] = td.Xf[(b * NJ + j) * 64 + lane];
  o.off = *reinterpret_cast<const f4*>(td.offf + c);
#pragma unroll
  for (int t = 0; t < NTMAX; ++t)
    if (t < NT) o.kb[t] = *reinterpret_cast<const f4*>(td.Kf + ((b * NT + t) * 64 + lane) * 4);
}

template <int P, typename CT>
__device__ __forceinline__ void irls_tile_chunk_mixed(const TileDesign& td,
                                                      const CT* __restrict__ Y, int64_t ld,
                                                      const int32_t* __restrict__ active, int nA,
                                                      const double* __restrict__ beta,
                                                      const double* __restrict__ kvec, int tile,
                                                      int chunk, double* __restrict__ part) {
  constexpr int NJ = (P + 3) / 4;
  constexpr int NTMAX = ((P * (P + 1)) / 2 + P + 15) / 16;
  constexpr int kSpan = 16;  // blocks per FP32 accumulation span
  const int lane = threadIdx.x & 63;
  const int s = lane & 15, h = lane >> 4;
  const int slot = tile * 16 + s;
  const bool live = slot < nA;
  const int g = live ? active[slot] : active[tile * 16];
  const int NT = td.nt;
  float bpos[NJ];
#pragma unroll
  for (int j = 0; j < NJ; ++j) {
    const int f = 4 * j + h;
    bpos[j] = (live && f < P) ? (float)beta[(int64_t)g * P + f] : 0.0f;
  }
  const float k = live ? (float)kvec[g] : 0.0f;  // k is a float in the reference
  const CT* yrow = Y + (int64_t)g * ld;
  const bool vec = (ld & 3) == 0 && ((reinterpret_cast<uintptr_t>(Y) & 15) == 0);
  f4 acc[NTMAX];
  double acc64[NTMAX][4];
#pragma unroll
  for (int t = 0; t < NTMAX; ++t) {
    acc[t] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int r = 0; r < 4; ++r) acc64[t][r] = 0.0;
  }
  const int64_t b0 = (int64_t)chunk * td.chunk_blocks;
  const int64_t b1 = min((int64_t)(chunk + 1) * td.chunk_blocks, td.nblk);
  BlockOpsF<NJ, NTMAX> cur, nxt;
  load_block_f<NJ, NTMAX, CT>(td, yrow, vec, b0, lane, h, NT, cur);
  int span = 0;
  for (int64_t b = b0; b < b1; ++b) {
    if (b + 1 < b1) load_block_f<NJ, NTMAX, CT>(td, yrow, vec, b + 1, lane, h, NT, nxt);
    f4 eta = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int j = 0; j < NJ; ++j) eta = mfma32(cur.xa[j], bpos[j], eta);
    float q[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const float mu = __builtin_amdgcn_exp2f((eta[r] + cur.off[r]) * 1.44269504088896341f);
      q[r] = (k + cur.y[r]) * __builtin_amdgcn_rcpf(k + mu);
    }
#pragma unroll
    for (int t = 0; t < NTMAX; ++t) {
      if (t < NT) {
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[t] = mfma32(q[r], cur.kb[t][r], acc[t]);
      }
    }
    if (++span == kSpan || b + 1 == b1) {
      span = 0;
#pragma unroll
      for (int t = 0; t < NTMAX; ++t) {
#pragma unroll
        for (int r = 0; r < 4; ++r) acc64[t][r] += (double)acc[t][r];
        acc[t] = f4{0.f, 0.f, 0.f, 0.f};
      }
    }
    cur = nxt;
  }
  // register r holds gene slot 4 h + r, column 16 t + s
  double* out = part + ((int64_t)tile * td.nchunk + chunk) * 16 * td.ncp;
#pragma unroll
  for (int t = 0; t < NTMAX; ++t) {
    if (t < NT) {
#pragma unroll
      for (int r = 0; r < 4; ++r) out[(4 * h + r) * td.ncp + 16 * t + s] = acc64[t][r];
    }
  }
}

template <int P>
__global__ __launch_bounds__(256) void k_irls_accum_mixed(TileDesign td, Counts c,
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
    irls_tile_chunk_mixed<P>(td, static_cast<const int32_t*>(c.ptr), c.ld, active, nA, beta,
                             kvec, tile, chunk, part);
  else if (c.dtype == 1)
    irls_tile_chunk_mixed<P>(td, static_cast<const float*>(c.ptr), c.ld, active, nA, beta,
                             kvec, tile, chunk, part);
  else
    irls_tile_chunk_mixed<P>(td, static_cast<const double*>(c.ptr), c.ld, active, nA, beta,
                             kvec, tile, chunk, part);
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
                                                    int32_t* __restrict__ nnext, double band,
                                                    int32_t* __restrict__ recheck,
                                                    int32_t* __restrict__ nrecheck) {
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
    const double d = fabs(v[j]);
    if (isnan(d)) has_nan = true;
    else mx = fmax(mx, d);
  }
  // partials from the mixed-precision pass: a step whose stopping decision
  // could differ from FP64 arithmetic (max |delta| within band * eps of eps,
  // or NaN) is left undone and the gene is redone in FP64 this iteration
  if (band > 0.0 && (has_nan || fabs(mx - eps) <= band * eps)) {
    recheck[atomicAdd(nrecheck, 1)] = g;
    return;
  }
#pragma unroll
  for (int j = 0; j < P; ++j) beta[(int64_t)g * P + j] += v[j];
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

hipError_t launch_irls_accum_mixed(hipStream_t s, const TileDesign& td, Counts c, int p,
                                   const int32_t* active, const int32_t* nactive, int nA_bound,
                                   const double* beta, const double* kvec, double* part) {
  if (nA_bound <= 0) return hipSuccess;
  const int ntile = (nA_bound + 15) / 16;
  const int groups = (ntile + 3) / 4;
  const dim3 grid((unsigned)((int64_t)groups * td.nchunk));
  DV_TILE_P(p, k_irls_accum_mixed, grid, dim3(256), s, td, c, active, nactive, beta, kvec, part);
  return hipGetLastError();
}

hipError_t launch_irls_update(hipStream_t s, const TileDesign& td, int p,
                              const int32_t* active, const int32_t* nactive, int nA_bound,
                              const double* part, double* beta, const double* kvec,
                              int max_iter, double eps, int32_t* iters, int32_t* status,
                              int32_t* next, int32_t* nnext, double band, int32_t* recheck,
                              int32_t* nrecheck) {
  if (nA_bound <= 0) return hipSuccess;
  const dim3 grid((unsigned)nA_bound);
  DV_TILE_P(p, k_irls_update, grid, dim3(64), s, td, active, nactive, part, beta, kvec,
            max_iter, eps, iters, status, next, nnext, band, recheck, nrecheck);
  return hipGetLastError();
}

hipError_t launch_iota(hipStream_t s, int32_t* p, int n) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_iota, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n);
  return hipGetLastError();
}

}  // namespace dv
