// gfx950 passes of the "new" / "I" overdispersion (fit_overdispersion_cppp,
// src/new_overdispersion.cpp:195-481), in lock-step rounds like the MLE:
//
//   k_newod_init    per gene: the moment start k = clamp(1 / kappa_mom) from
//                   the setup pass's sums, and the driver's first request
//   k_newod_accum   one wavefront per (16-gene tile, cell chunk): at each
//                   active gene's requested k the per-cell sums of F_k,
//                   sum log1p(k w), sum (k + y) w / (1 + k w) with
//                   w = exp(-(X beta + off)), sum X beta, and sum
//                   digamma(k + y) for genes without a count table
//   k_newod_update  one wavefront per active gene: the table's digamma sum,
//                   C1, F_k (newod_driver.hpp) and one step of the driver
#include <hip/hip_runtime.h>

#include "../../include/devil_hip.h"
#include "device_common.hpp"
#include "kernels.hpp"
#include "newod_driver.hpp"
#include "tile_common.hpp"

namespace dv {

constexpr int kNodScal = 4;  // sum log1p(k w), sum (k + y) w / (1 + k w), sum eta, sum digamma

__global__ void k_newod_init(int64_t G, const double* __restrict__ raw_start,
                             NewodState* __restrict__ st, double* __restrict__ kreq,
                             int32_t* __restrict__ act, int32_t* __restrict__ nact,
                             int max_iter, double eps) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= G) return;
  // kappa_mom = (var y - ybar) / ybar^2, 0.01 when not finite or <= 0;
  // k = 1 / kappa_mom clamped to [1e-4, 1e4] (:233-249)
  double km = raw_start[g];
  if (!isfinite(km) || km <= 0.0) km = 0.01;
  double k = 1.0 / km;
  if (!isfinite(k)) k = 1e-4;
  NewodState s;
  s.pc = 0;
  s.err = 0;
  s.k = fmax(1e-4, fmin(1e4, k));
  newod_advance(s, 0.0, max_iter, eps);  // the first request
  st[g] = s;
  kreq[g] = s.pending;
  act[atomicAdd(nact, 1)] = (int32_t)g;
}

template <int P, typename CT, bool VEC>
__device__ __forceinline__ void newod_tile_chunk(const TileDesign& td, const CT* __restrict__ Y,
                                                 int64_t ld, const int32_t* __restrict__ active,
                                                 int nA, const double* __restrict__ beta,
                                                 const double* __restrict__ kreq,
                                                 const int32_t* __restrict__ tab_n, int tile,
                                                 int chunk, double* __restrict__ part) {
  constexpr int NJ = (P + 3) / 4;
  const int lane = threadIdx.x & 63;
  const int s = lane & 15, h = lane >> 4;
  const int slot = tile * 16 + s;
  const bool live = slot < nA;
  const int g = live ? active[slot] : active[tile * 16];
  double bpos[NJ];
#pragma unroll
  for (int j = 0; j < NJ; ++j) {
    const int f = 4 * j + h;
    bpos[j] = (live && f < P) ? beta[(int64_t)g * P + f] : 0.0;
  }
  const double k = kreq[g];
  const bool cell_psi = tab_n[g] < 0;
  const CT* yrow = Y + (int64_t)g * ld;
  const CT* yh = yrow + 4 * h;
  double sa = 0.0, sb = 0.0, se = 0.0, sd = 0.0;
  const int64_t b0 = (int64_t)chunk * td.chunk_blocks;
  const int64_t b1 = min((int64_t)(chunk + 1) * td.chunk_blocks, td.nblk);
  typedef TileBlock<NJ, 0, CT> Blk;  // no K operands: eta and offsets only
  auto body = [&](const Blk& cur, int64_t b) {
    d4 eta = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int j = 0; j < NJ; ++j) eta = mfma16x16x4(cur.xa[j], bpos[j], eta);
    const int64_t c0 = b * 16 + 4 * h;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      if (c0 + r < td.n) {
        const double w = exp_f64(-(eta[r] + cur.off[r]));
        const double y = (double)cur.y[r];
        const double kw = k * w;
        sa += log1p_f64(kw);
        sb += (k + y) * w * rcp_f64(1.0 + kw);
        se += eta[r];
        if (cell_psi) sd += digamma(k + y);
      }
    }
  };
  auto load = [&](int64_t b, Blk& o) {
    load_tile_block<NJ, 0, CT, VEC>(td, yh, b, lane, h, 0, td.offp, o);
  };
  auto load_tail = [&](int64_t b, Blk& o) {
    load_tile_block_tail<NJ, 0, CT>(td, yrow, b, lane, h, 0, td.offp, o);
  };
  pipeline<Blk>(b0, min(b1, td.n / 16), b1, load, load_tail, body);
#pragma unroll
  for (int o = 16; o <= 32; o <<= 1) {
    sa += __shfl_xor(sa, o, 64);
    sb += __shfl_xor(sb, o, 64);
    se += __shfl_xor(se, o, 64);
    sd += __shfl_xor(sd, o, 64);
  }
  if (h == 0) {
    double* out = part + (((int64_t)tile * td.nchunk + chunk) * 16 + s) * kNodScal;
    out[0] = sa;
    out[1] = sb;
    out[2] = se;
    out[3] = sd;
  }
}

template <int P>
__global__ __launch_bounds__(256) void k_newod_accum(TileDesign td, Counts c,
                                                     const int32_t* __restrict__ active,
                                                     const int32_t* __restrict__ nactive,
                                                     const double* __restrict__ beta,
                                                     const double* __restrict__ kreq,
                                                     const int32_t* __restrict__ tab_n,
                                                     double* __restrict__ part) {
  const int nA = *nactive;
  const int ntile = (nA + 15) / 16;
  const int wave = wave_index();
  int chunk, group;
  tile_coords(td.nchunk, chunk, group);
  const int tile = group * 4 + wave;
  if (tile >= ntile) return;
  with_counts(c, [&](auto Y, auto vec) {
    typedef std::remove_cv_t<std::remove_pointer_t<decltype(Y)>> CT;
    newod_tile_chunk<P, CT, decltype(vec)::value>(td, Y, c.ld, active, nA, beta, kreq, tab_n,
                                                  tile, chunk, part);
  });
}

// threads of the update: the count-table digamma sum of one gene spreads over
// four waves (a lone gene's long tail of evaluations is latency-bound)
constexpr int kNodUpdThreads = 256;

__global__ __launch_bounds__(kNodUpdThreads) void k_newod_update(int nchunk, const int32_t* __restrict__ active,
                                                     const int32_t* __restrict__ nactive,
                                                     const double* __restrict__ part,
                                                     const int32_t* __restrict__ tab_n,
                                                     const int32_t* __restrict__ tab_u,
                                                     const int32_t* __restrict__ tab_f,
                                                     NewodState* __restrict__ st,
                                                     double* __restrict__ kreq, int64_t n,
                                                     int max_iter, double eps,
                                                     double* __restrict__ theta,
                                                     int32_t* __restrict__ iters,
                                                     int32_t* __restrict__ status,
                                                     int32_t* __restrict__ next,
                                                     int32_t* __restrict__ nnext) {
  const int nA = *nactive;
  const int i = blockIdx.x;
  if (i >= nA) return;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int g = active[i];
  const int tile = i >> 4, s = i & 15;
  __shared__ double red[2][kNodUpdThreads / 64];
  // the chunks' partials (wave 0): 16 lanes per scalar, then a reduction
  // within each group of 16
  static_assert(kNodScal * 16 == 64, "one 16-lane group per scalar");
  double a = 0.0;
  if (wave == 0) {
    const int q = lane >> 4, l = lane & 15;
    const double* src = part + ((int64_t)tile * nchunk * 16 + s) * kNodScal + q;
    for (int ch = l; ch < nchunk; ch += 16) a += src[(int64_t)ch * 16 * kNodScal];
#pragma unroll
    for (int o = 8; o >= 1; o >>= 1) a += __shfl_xor(a, o, 64);
  }
  const double sa = __shfl(a, 0, 64), sb = __shfl(a, 16, 64), se = __shfl(a, 32, 64);
  double sd = __shfl(a, 48, 64);
  const double k = kreq[g];
  // digamma(k + y) over the count table; cells with y = 0 give digamma(k)
  const int tn = tab_n[g];
  if (tn >= 0) {  // uniform over the block
    double tp = 0.0, nz = 0.0;
    for (int j = t; j < tn; j += kNodUpdThreads) {
      const double f = (double)tab_f[(int64_t)g * kMleTabCap + j];
      tp += f * digamma(k + (double)tab_u[(int64_t)g * kMleTabCap + j]);
      nz += f;
    }
    tp = wave_sum(tp);
    nz = wave_sum(nz);
    if (lane == 0) {
      red[0][wave] = tp;
      red[1][wave] = nz;
    }
    __syncthreads();
    if (wave == 0) {
      tp = 0.0;
      nz = 0.0;
#pragma unroll
      for (int w = 0; w < kNodUpdThreads / 64; ++w) {
        tp += red[0][w];
        nz += red[1][w];
      }
      sd = tp + ((double)n - nz) * digamma(k);
    }
  }
  if (wave != 0) return;
  // wave 0: C1 and F_k (the same values in every lane), then lane 0 steps the
  // driver
  const double C1 = se - (sd - sa) + sb;
  int ferr = 0;
  const double F = nod_F_from_C1_wave((double)n, C1, eps, ferr);
  if (lane != 0) return;
  NewodState sg = st[g];
  if (ferr) sg.err = ferr;
  const int done = newod_advance(sg, F, max_iter, eps);
  st[g] = sg;
  if (done) {
    theta[g] = 1.0 / sg.k;
    iters[g] = sg.it + 1;
    if (sg.err) status[g] |= DEVIL_GENE_OD_FAILED;
  } else {
    kreq[g] = sg.pending;
    next[atomicAdd(nnext, 1)] = g;
  }
}

#define DV_NOD_P(p, KERNEL, GRID, BLOCK, STREAM, ...)                               \
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

hipError_t launch_newod_init(hipStream_t s, int64_t G, const double* raw_start,
                             NewodState* st, double* kreq, int32_t* act, int32_t* nact,
                             int max_iter, double eps) {
  if (G <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_newod_init, dim3((unsigned)((G + 63) / 64)), dim3(64), 0, s, G,
                     raw_start, st, kreq, act, nact, max_iter, eps);
  return hipGetLastError();
}

hipError_t launch_newod_accum(hipStream_t s, const TileDesign& td, Counts c, int p,
                              const int32_t* active, const int32_t* nactive, int nA_bound,
                              const double* beta, const double* kreq, const int32_t* tab_n,
                              double* part) {
  if (nA_bound <= 0) return hipSuccess;
  const int ntile = (nA_bound + 15) / 16;
  const int groups = (ntile + 3) / 4;
  const dim3 grid((unsigned)((int64_t)groups * td.nchunk));
  DV_NOD_P(p, k_newod_accum, grid, dim3(256), s, td, c, active, nactive, beta, kreq, tab_n,
           part);
  return hipGetLastError();
}

hipError_t launch_newod_update(hipStream_t s, int nchunk, const int32_t* active,
                               const int32_t* nactive, int nA_bound, const double* part,
                               const MleState& ms, NewodState* st, double* kreq, int64_t n,
                               int max_iter, double eps, int32_t* next, int32_t* nnext) {
  if (nA_bound <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_newod_update, dim3((unsigned)nA_bound), dim3(kNodUpdThreads), 0, s,
                     nchunk, active, nactive, part, ms.tab_n, ms.tab_u, ms.tab_f, st, kreq, n, max_iter, eps,
                     ms.theta, ms.iters, ms.status, next, nnext);
  return hipGetLastError();
}

}  // namespace dv
