// gfx950 MLE overdispersion with Cox-Reid adjustment, tiled like the IRLS.
//
// R/overdispersion.R:25-115 runs nlminb on the log-likelihood of
// src/overdispersion.cpp:53-88, its score (:95-178) and the score's
// derivative (:186-242).  Every gene runs one of two drivers, all genes in
// lock-step rounds of one evaluation each:
//   DEVIL_MLE_NLMINB (default)  nlminb's PORT iteration and stopping rules
//                               restated for one parameter (port_driver.hpp)
//   DEVIL_MLE_ROOT              the safeguarded Newton iteration on
//                               score(log theta) = 0 to |dx| < 1e-9
//
//   k_mle_setup   one wavefront per gene: sum y, sum (y-1) y, max y, the
//                 all-zero rule and the start value (var(y) - ybar) / ybar^2.
//   k_mle_accum   one wavefront per (16-gene tile, cell chunk) evaluates, at
//                 each gene's current log theta, the per-cell sums of the
//                 score and its derivative and the three Cox-Reid Gram
//                 matrices X' diag(w) X, X' diag(-w^2) X, X' diag(2 w^3) X
//                 with w = mu / (1 + mu theta) (v_mfma_f64_16x16x4f64 over the
//                 Khatri-Rao columns, as in the IRLS).
//   k_mle_update  one wavefront per active gene: forms the log-likelihood,
//                 score and derivative, applies the far-left rule
//                 (score(log 1e-8) < 0 -> 0) on the first round and one step
//                 of the gene's driver after that.
//
// Per cell everything is written with r = 1 / (1 + mu theta):
//   y / (mu + 1/theta) = y theta r,  (y - mu) / (mu + 1/theta) = (y - mu) theta r,
//   (mu^2 theta + y) / (1 + mu theta)^2 = (mu mu theta + y) r^2,  w = mu r,
// and the digamma / trigamma sums and log(1 + x) - x / (1 + x) are taken in
// their cancellation-free forms (dpsi_diff, dpsi1_diff, nb_f).  The
// log-likelihood is taken as
//   sum_i ldiff(1/theta, y_i) - sum_i (y_i + 1/theta) log1p(mu_i theta) + CR,
// ldiff(a, y) = lgamma(a + y) - lgamma(a) - y log(a): the same value as the
// reference's lgamma sums, whose y log(1/theta) and n/theta log(theta) terms
// cancel exactly.
#include <hip/hip_runtime.h>

#include "../../include/devil_hip.h"
#include "device_common.hpp"
#include "kernels.hpp"
#include "port_driver.hpp"
#include "tile_common.hpp"

namespace dv {

// lgamma(a + y) - lgamma(a) - y log(a) for a > 0, y >= 0 without the
// cancellation of the two lgamma values: a sum of log1p(j / a) for small
// integer y, else Stirling's series differenced term by term after shifting
// a to >= 10 (ldiff(a, y) = ldiff(a + 1, y) + y log1p(1/a) - log1p(y/a)).
__device__ inline double ldiff(double a, double y) {
  if (y == 0.0) return 0.0;
  if (y <= 16.0 && y == floor(y)) {
    double s = 0.0;
    for (int j = 1; j < (int)y; ++j) s += log1p(j / a);
    return s;
  }
  double s = 0.0;
  while (a < 10.0) {
    s += y * log1p(1.0 / a) - log1p(y / a);
    a += 1.0;
  }
  // (a + y - 1/2) log1p(t) - y = a (log1p(t) - t) + (y - 1/2) log1p(t), t = y / a,
  // with log1p(t) - t by its series below t = 0.1
  const double t = y / a;
  double lm;
  if (t < 0.1) {
    double q = 0.0;
    for (int k = 24; k >= 2; --k) q = ((k & 1) ? 1.0 : -1.0) / k + t * q;
    lm = t * t * q;
  } else {
    lm = log1p(t) - t;
  }
  return s + a * lm + (y - 0.5) * log1p(t) + lgamma_corr(a + y) - lgamma_corr(a);
}

enum { kMleFarLeft = 0, kMleNewton = 1, kMleDone = 2 };

// columns of a gene's partial row: 8 scalars then the three Gram tiles
constexpr int kMleScal = 8;  // sc, l1, l2, dg, tg, ld (cell ldiff), sum y L, sum L

// Per-gene table of the distinct positive counts (value, frequency), the
// make_table_if_small idea of src/overdispersion.cpp:32-46: the digamma,
// trigamma and lgamma sums then cost O(#distinct) per pass instead of O(n).
// Every integer count below kMleHist has its bin, so only genes with a count
// of 4096 or more (or fractional counts) stay on the per-cell path
// (tab_n = -1).
constexpr int kMleHist = 4096;
static_assert(kMleTabCap == kMleHist, "one table entry per possible bin");
constexpr int kSetupThreads = 256;

// One workgroup of 256 threads per gene, one read of its counts: the
// histogram in LDS, sum y, sum y^2 (integer counts: exact, 64 and 128 bits),
// max y and the zero count; float counts take a second read for the
// variance about the mean.
template <int P>
__global__ __launch_bounds__(kSetupThreads) void k_mle_setup(Counts c, int64_t G, int64_t n,
                                                             MleState ms) {
  __shared__ int32_t hist[kMleHist];
  __shared__ int32_t big;
  __shared__ int64_t w1[kSetupThreads / 64], wz[kSetupThreads / 64];
  __shared__ uint64_t wlo[kSetupThreads / 64], whi[kSetupThreads / 64];
  __shared__ double wd1[kSetupThreads / 64], wmx[kSetupThreads / 64];
  __shared__ int32_t wcnt[kSetupThreads / 64];
  const int64_t g = blockIdx.x;
  if (g >= G) return;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  constexpr int NW = kSetupThreads / 64;
  for (int k = t; k < kMleHist; k += kSetupThreads) hist[k] = 0;
  if (t == 0) big = 0;
  __syncthreads();
  const void* row = count_row(c.ptr, c.dtype, g, c.ld);
  const bool integer = c.dtype == DEVIL_COUNTS_I32;
  int64_t s1i = 0, nz = 0;
  uint64_t s2lo = 0, s2hi = 0;
  double s1d = 0.0, my = 0.0;
  for (int64_t i = t; i < n; i += kSetupThreads) {
    double y;
    if (integer) {
      const int64_t yi = static_cast<const int32_t*>(row)[i];
      s1i += yi;
      const uint64_t q = (uint64_t)(yi * yi);
      s2lo += q;
      s2hi += s2lo < q ? 1 : 0;
      y = (double)yi;
    } else {
      y = load_count(row, c.dtype, i);
      s1d += y;
    }
    my = fmax(my, y);
    nz += y == 0.0 ? 1 : 0;
    if (y > 0.0) {
      if (y < (double)kMleHist && y == floor(y)) atomicAdd(&hist[(int)y], 1);
      else big = 1;
    }
  }
  // wave reductions, then across the four waves
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    s1i += __shfl_xor(s1i, o, 64);
    nz += __shfl_xor(nz, o, 64);
    const uint64_t lo = __shfl_xor(s2lo, o, 64), hi = __shfl_xor(s2hi, o, 64);
    s2lo += lo;
    s2hi += hi + (s2lo < lo ? 1 : 0);
    s1d += __shfl_xor(s1d, o, 64);
    my = fmax(my, __shfl_xor(my, o, 64));
  }
  if (lane == 0) {
    w1[wave] = s1i;
    wz[wave] = nz;
    wlo[wave] = s2lo;
    whi[wave] = s2hi;
    wd1[wave] = s1d;
    wmx[wave] = my;
  }
  __syncthreads();
  s1i = 0;
  nz = 0;
  s1d = 0.0;
  my = 0.0;
  unsigned __int128 s2 = 0;
#pragma unroll
  for (int w = 0; w < NW; ++w) {
    s1i += w1[w];
    nz += wz[w];
    s2 += ((unsigned __int128)whi[w] << 64) | wlo[w];
    s1d += wd1[w];
    my = fmax(my, wmx[w]);
  }
  // compact the non-zero bins (ascending value) into the gene's table
  {
    constexpr int per = kMleHist / kSetupThreads;
    int cnt = 0;
    for (int k = t * per; k < (t + 1) * per; ++k) cnt += hist[k] > 0 ? 1 : 0;
    int pre = cnt;  // inclusive prefix sum within the wave
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int v = __shfl_up(pre, o, 64);
      if (lane >= o) pre += v;
    }
    if (lane == 63) wcnt[wave] = pre;
    __syncthreads();
    int base = 0, total = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      if (w < wave) base += wcnt[w];
      total += wcnt[w];
    }
    const bool ok = big == 0 && total <= kMleTabCap;
    if (ok) {
      int pos = base + pre - cnt;
      int32_t* tu = ms.tab_u + g * kMleTabCap;
      int32_t* tf = ms.tab_f + g * kMleTabCap;
      for (int k = t * per; k < (t + 1) * per; ++k)
        if (hist[k] > 0) {
          tu[pos] = k;
          tf[pos] = hist[k];
          ++pos;
        }
    }
    if (t == 0) ms.tab_n[g] = ok ? total : -1;
  }
  if (nz == n) {  // all(y == 0) -> 0 (R/overdispersion.R:28-30)
    if (t == 0) {
      ms.mode[g] = kMleDone;
      ms.theta[g] = 0.0;
      ms.iters[g] = 0;
      ms.status[g] |= DEVIL_GENE_MLE_ZERO;
    }
    return;
  }
  double s1, v, spy;
  if (integer) {
    // var with n - 1 from the exact sums: (n s2 - s1^2) / (n (n - 1))
    s1 = (double)s1i;
    const __int128 num = (__int128)n * (__int128)s2 - (__int128)s1i * (__int128)s1i;
    v = ((double)num / (double)n) / (double)(n - 1);
    spy = (double)(s2 - (unsigned __int128)s1i);  // sum (y - 1) y
  } else {
    s1 = s1d;
    const double ybar = s1 / (double)n;
    double vv = 0.0, sp = 0.0;
    for (int64_t i = t; i < n; i += kSetupThreads) {
      const double y = load_count(row, c.dtype, i);
      vv += (y - ybar) * (y - ybar);
      sp += (y - 1) * y;
    }
    vv = wave_sum(vv);
    sp = wave_sum(sp);
    __syncthreads();
    if (lane == 0) {
      wd1[wave] = vv;
      wmx[wave] = sp;
    }
    __syncthreads();
    vv = 0.0;
    sp = 0.0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      vv += wd1[w];
      sp += wmx[w];
    }
    v = vv / (double)(n - 1);
    spy = sp;
  }
  const double ybar = s1 / (double)n;
  double start = (v - ybar) / (ybar * ybar);
  if (t == 0 && ms.raw_start) ms.raw_start[g] = start;
  if (isnan(start) || start <= 0) start = 0.5;
  if (t == 0) {
    ms.mode[g] = kMleFarLeft;
    ms.x[g] = log(1e-8);
    ms.start[g] = start;
    ms.sum_y[g] = s1;
    ms.sum_py[g] = spy;
    ms.max_y[g] = my;
    ms.iters[g] = 0;
  }
}

__global__ void k_mle_collect(const int32_t* __restrict__ mode, int G, int32_t* __restrict__ act,
                              int32_t* __restrict__ nact) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g < G && mode[g] != kMleDone) act[atomicAdd(nact, 1)] = g;
}

template <int P, typename CT, bool VEC>
__device__ __forceinline__ void mle_tile_chunk(const TileDesign& td, const CT* __restrict__ Y,
                                               int64_t ld, const int32_t* __restrict__ active,
                                               int nA, const double* __restrict__ beta,
                                               const double* __restrict__ xs,
                                               const int32_t* __restrict__ tab_n, int tile,
                                               int chunk, bool want_f,
                                               double* __restrict__ part) {
  constexpr int NJ = (P + 3) / 4;
  constexpr int NTMAX = ((P * (P + 1)) / 2 + P + 15) / 16;
  const int lane = threadIdx.x & 63;
  const int s = lane & 15, h = lane >> 4;
  const int slot = tile * 16 + s;
  const bool live = slot < nA;
  const int g = live ? active[slot] : active[tile * 16];
  const int NT = td.nt;
  double bpos[NJ];
#pragma unroll
  for (int j = 0; j < NJ; ++j) {
    const int f = 4 * j + h;
    bpos[j] = (live && f < P) ? beta[(int64_t)g * P + f] : 0.0;
  }
  const double theta = exp(xs[g]);
  const double tn1 = 1.0 / theta;
  const bool cell_psi = tab_n[g] < 0;
  const CT* yrow = Y + (int64_t)g * ld;
  const CT* yh = yrow + 4 * h;
  d4 a0[NTMAX], a1[NTMAX], a2[NTMAX];
#pragma unroll
  for (int t = 0; t < NTMAX; ++t) {
    a0[t] = d4{0.0, 0.0, 0.0, 0.0};
    a1[t] = a0[t];
    a2[t] = a0[t];
  }
  double sc = 0.0, l1 = 0.0, l2 = 0.0, dg = 0.0, tg = 0.0, lg = 0.0, syl = 0.0, sl = 0.0;
  const int64_t b0 = (int64_t)chunk * td.chunk_blocks;
  const int64_t b1 = min((int64_t)(chunk + 1) * td.chunk_blocks, td.nblk);
  typedef TileBlock<NJ, NTMAX, CT> Blk;
  auto body = [&](const Blk& cur, int64_t b) {
    d4 eta = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int j = 0; j < NJ; ++j) eta = mfma16x16x4(cur.xa[j], bpos[j], eta);
    const int64_t c0 = b * 16 + 4 * h;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      double mu = exp_f64(eta[r] + cur.off[r]);
      if (mu == 0.0) mu = 1e-6;
      const double y = (double)cur.y[r];
      const double mt = mu * theta;
      const double rr = rcp_f64(1.0 + mt);
      const double L = log1p_f64(mt);
      const double f = (mt < 1e-2) ? nb_f(mt) : L - mt * rr;
      const double w = mu * rr;
      const double w1 = -1.0 * w * w;
      const double w2 = -2.0 * w1 * w;
      // the three Gram sums take this cell row at once (each accumulator still
      // adds r = 0..3 in order), so no per-row weights stay live
#pragma unroll
      for (int t = 0; t < NTMAX; ++t) {
        if (t < NT) {
          a0[t] = mfma16x16x4(w, cur.kb[t][r], a0[t]);
          a1[t] = mfma16x16x4(w1, cur.kb[t][r], a1[t]);
          a2[t] = mfma16x16x4(w2, cur.kb[t][r], a2[t]);
        }
      }
      if (c0 + r < td.n) {
        sc += f + y * theta * rr;
        l1 += L + (y - mu) * theta * rr;
        l2 += (mu * mt + y) * rr * rr;
        syl += y * L;
        sl += L;
        if (cell_psi && y != 0.0) {
          dg += dpsi_diff(tn1, y);
          tg += dpsi1_diff(tn1, y);
          if (want_f) lg += ldiff(tn1, y);
        }
      }
    }
  };
  auto load = [&](int64_t b, Blk& o) {
    load_tile_block<NJ, NTMAX, CT, VEC>(td, yh, b, lane, h, NT, td.offp, o);
  };
  auto load_tail = [&](int64_t b, Blk& o) {
    load_tile_block_tail<NJ, NTMAX, CT>(td, yrow, b, lane, h, NT, td.offp, o);
  };
  pipeline<Blk>(b0, min(b1, td.n / 16), b1, load, load_tail, body);  // trailing chunks may be empty
  // the four lanes of a gene slot (h = 0..3) hold disjoint cells
#pragma unroll
  for (int o = 16; o <= 32; o <<= 1) {
    sc += __shfl_xor(sc, o, 64);
    l1 += __shfl_xor(l1, o, 64);
    l2 += __shfl_xor(l2, o, 64);
    dg += __shfl_xor(dg, o, 64);
    tg += __shfl_xor(tg, o, 64);
    lg += __shfl_xor(lg, o, 64);
    syl += __shfl_xor(syl, o, 64);
    sl += __shfl_xor(sl, o, 64);
  }
  const int W = kMleScal + 3 * td.ncp;
  double* out = part + ((int64_t)tile * td.nchunk + chunk) * 16 * W;
  if (h == 0) {
    double* o = out + s * W;
    o[0] = sc;
    o[1] = l1;
    o[2] = l2;
    o[3] = dg;
    o[4] = tg;
    o[5] = lg;
    o[6] = syl;
    o[7] = sl;
  }
#pragma unroll
  for (int t = 0; t < NTMAX; ++t) {
    if (t < NT) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        double* o = out + (h + 4 * r) * W + kMleScal + 16 * t + s;
        o[0] = a0[t][r];
        o[td.ncp] = a1[t][r];
        o[2 * td.ncp] = a2[t][r];
      }
    }
  }
}

template <int P>
__global__ __launch_bounds__(256) void k_mle_accum(TileDesign td, Counts c,
                                                   const int32_t* __restrict__ active,
                                                   const int32_t* __restrict__ nactive,
                                                   const double* __restrict__ beta,
                                                   const double* __restrict__ xs,
                                                   const int32_t* __restrict__ tab_n,
                                                   int want_f, double* __restrict__ part) {
  const int nA = *nactive;
  const int ntile = (nA + 15) / 16;
  const int wave = wave_index();
  int chunk, group;
  tile_coords(td.nchunk, chunk, group);
  const int tile = group * 4 + wave;
  if (tile >= ntile) return;
  with_counts(c, [&](auto Y, auto vec) {
    typedef std::remove_cv_t<std::remove_pointer_t<decltype(Y)>> CT;
    mle_tile_chunk<P, CT, decltype(vec)::value>(td, Y, c.ld, active, nA, beta, xs, tab_n, tile,
                                                chunk, want_f != 0, part);
  });
}

// The update runs on kMleUpdThreads threads per gene: the chunk partials in
// kMleUpdSlices interleaved slices per column and the count-table sums spread
// over every thread (a lone gene's last rounds are latency-bound), then one
// lane steps the driver.
constexpr int kMleUpdThreads = 256;
constexpr int kMleUpdSlices = 4;

template <int P>
__global__ __launch_bounds__(kMleUpdThreads) void k_mle_update(
    TileDesign td, const int32_t* __restrict__ active, const int32_t* __restrict__ nactive,
    const double* __restrict__ part, MleState ms, int64_t n, int max_iter, PortParams pp,
    int32_t* __restrict__ next, int32_t* __restrict__ nnext) {
  __shared__ double acc[kMleScal + 3 * DEVIL_NCP_MAX];
  __shared__ double slice[kMleUpdSlices][kMleScal + 3 * DEVIL_NCP_MAX];
  __shared__ double red[3][kMleUpdThreads / 64];
  const int nA = *nactive;
  const int i = blockIdx.x;
  if (i >= nA) return;  // uniform over the block
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int g = active[i];
  const int tile = i >> 4, s = i & 15;
  const int W = kMleScal + 3 * td.ncp;
  const double* src = part + (int64_t)tile * td.nchunk * 16 * W + (int64_t)s * W;
  for (int idx = t; idx < kMleUpdSlices * W; idx += kMleUpdThreads) {
    const int col = idx % W, q = idx / W;
    double a = 0.0;
    for (int ch = q; ch < td.nchunk; ch += kMleUpdSlices) a += src[(int64_t)ch * 16 * W + col];
    slice[q][col] = a;
  }
  __syncthreads();
  for (int col = t; col < W; col += kMleUpdThreads) {
    double a = 0.0;
#pragma unroll
    for (int q = 0; q < kMleUpdSlices; ++q) a += slice[q][col];
    acc[col] = a;
  }
  const double x = ms.x[g];
  const double theta = exp(x);
  const double tn1 = 1.0 / theta;
  const int mode0 = ms.mode[g];
  const bool port = ms.driver == kMleDriverNlminb && mode0 == kMleNewton;
  // digamma / trigamma (and, for the nlminb driver, lgamma) difference sums
  // over the distinct-count table
  double tdg = 0.0, ttg = 0.0, tld = 0.0;
  const int tn = ms.tab_n[g];
  for (int k = t; k < tn; k += kMleUpdThreads) {
    const double u = (double)ms.tab_u[(int64_t)g * kMleTabCap + k];
    const double f = (double)ms.tab_f[(int64_t)g * kMleTabCap + k];
    tdg += f * dpsi_diff(tn1, u);
    ttg += f * dpsi1_diff(tn1, u);
    if (port) tld += f * ldiff(tn1, u);
  }
  tdg = wave_sum(tdg);
  ttg = wave_sum(ttg);
  tld = wave_sum(tld);
  if (lane == 0) {
    red[0][wave] = tdg;
    red[1][wave] = ttg;
    red[2][wave] = tld;
  }
  __syncthreads();
  if (t != 0) return;
  tdg = ttg = tld = 0.0;
#pragma unroll
  for (int w = 0; w < kMleUpdThreads / 64; ++w) {
    tdg += red[0][w];
    ttg += red[1][w];
    tld += red[2][w];
  }
  const double sc = acc[0], l1 = acc[1], l2 = acc[2];
  const double dg = acc[3] + tdg, tg = acc[4] + ttg;
  const double* B0 = acc + kMleScal;
  const double* B1 = B0 + td.ncp;
  const double* B2 = B1 + td.ncp;
  double Bm[P][P], Dm[P][P], D2m[P][P], Bi[P][P];
#pragma unroll
  for (int a = 0; a < P; ++a)
#pragma unroll
    for (int bb = a; bb < P; ++bb) {
      const int pi = td.kidx_pair[tri<P>(a, bb)];
      const double v0 = (pi >= 0 ? B0[pi] : 0.0) + (a == bb ? 1e-6 : 0.0);
      const double v1 = pi >= 0 ? B1[pi] : 0.0;
      const double v2 = pi >= 0 ? B2[pi] : 0.0;
      Bm[a][bb] = v0;
      Bm[bb][a] = v0;
      Dm[a][bb] = v1;
      Dm[bb][a] = v1;
      D2m[a][bb] = v2;
      D2m[bb][a] = v2;
    }
  chol_inverse<P>(Bm, Bi);
  double tr_bdb = 0.0;
#pragma unroll
  for (int a = 0; a < P; ++a)
#pragma unroll
    for (int bb = 0; bb < P; ++bb) tr_bdb += Bi[a][bb] * Dm[bb][a];
  const double cr_s = -0.5 * tr_bdb * 0.99;
  const double sum_y = ms.sum_y[g], spy = ms.sum_py[g], max_y = ms.max_y[g];
  const double corr = tn1 > 1e5 ? spy / (2 * tn1) : 0.0;
  const double dterm = (max_y * 1e6 < tn1) ? sum_y - corr : fmin(dg * tn1, sum_y - corr);
  const double score = sc * tn1 - dterm + cr_s * theta;
  const int mode = mode0;
  int it = ms.iters[g] + 1;
  ms.iters[g] = it;
  if (mode == kMleFarLeft) {
    if (score < 0) {  // R/overdispersion.R:40-46
      ms.mode[g] = kMleDone;
      ms.theta[g] = 0.0;
      ms.status[g] |= DEVIL_GENE_MLE_ZERO;
      return;
    }
    // the root driver brackets in [log 1e-8, log 1e16]; nlminb's box is
    // [log 1e-16, log 1e16] (R/overdispersion.R:72)
    const double lo = ms.driver == kMleDriverNlminb ? pp.lo : log(1e-8), hi = log(1e16);
    ms.lo[g] = lo;
    ms.hi[g] = hi;
    ms.x[g] = fmin(fmax(log(ms.start[g]), lo), hi);
    ms.mode[g] = kMleNewton;
    if (ms.driver == kMleDriverNlminb) {
      PortState& st = ms.port[g];
      st.phase = kPortStart;
      st.xe = ms.x[g];
      st.code = 0;
    }
    if (max_iter > 0) {
      next[atomicAdd(nnext, 1)] = g;
    } else {
      ms.mode[g] = kMleDone;
      ms.theta[g] = exp(ms.x[g]);
    }
    return;
  }
  // derivative of the score (src/overdispersion.cpp:186-242)
  double DiDb[P][P];
#pragma unroll
  for (int a = 0; a < P; ++a)
#pragma unroll
    for (int bb = 0; bb < P; ++bb) {
      double t = 0.0;
#pragma unroll
      for (int k = 0; k < P; ++k) t += Bi[a][k] * Dm[k][bb];
      DiDb[a][bb] = t;
    }
  double ddetb = 0.0, tr_sq = 0.0, tr_d2 = 0.0;
#pragma unroll
  for (int a = 0; a < P; ++a) {
    ddetb += DiDb[a][a];
#pragma unroll
    for (int k = 0; k < P; ++k) {
      tr_sq += DiDb[a][k] * DiDb[k][a];
      tr_d2 += Bi[a][k] * D2m[k][a];
    }
  }
  const double d2detb = ddetb * ddetb - tr_sq + tr_d2;
  const double cr_t = (0.5 * ddetb * ddetb - 0.5 * d2detb) * 0.99;
  const double cr_t2 = -0.5 * ddetb * 0.99;
  const double tn2 = 1.0 / (theta * theta);
  const double ll = -2 * tn1 * (l1 - dg) + (l2 + tn2 * tg);
  const double d = ll + cr_t * (theta * theta) + (l1 - dg) * tn1 + cr_t2 * theta;
  if (port) {
    // log-likelihood (src/overdispersion.cpp:53-88): the Cox-Reid term from
    // the LU factors (partial pivoting) of X' diag(w) X without the ridge,
    // each pivot's log clamped below at log(1e-50)
    double Lm[P][P];
    int perm[P];
#pragma unroll
    for (int a = 0; a < P; ++a)
#pragma unroll
      for (int bb = a; bb < P; ++bb) {
        const int pi = td.kidx_pair[tri<P>(a, bb)];
        const double v0 = pi >= 0 ? B0[pi] : 0.0;
        Lm[a][bb] = v0;
        Lm[bb][a] = v0;
      }
    lu_factor<P>(Lm, perm);
    double ldet = 0.0;
#pragma unroll
    for (int a = 0; a < P; ++a) {
      const double e = Lm[a][a];
      ldet += e < 1e-50 ? log(1e-50) : log(e);
    }
    const double loglik = acc[5] + tld - (acc[6] + tn1 * acc[7]) - 0.5 * ldet * 0.99;
    PortState st = ms.port[g];
    const int done = port_advance(st, -loglik, -score, -d, pp);
    ms.port[g] = st;
    if (done || it - 1 >= kPtEvalMax + 1) {
      ms.mode[g] = kMleDone;
      ms.theta[g] = exp(st.x);
      ms.x[g] = st.x;
      if (!(st.code >= 3 && st.code <= 6)) ms.status[g] |= DEVIL_GENE_MLE_NOCONV;
    } else {
      ms.x[g] = st.xe;
      next[atomicAdd(nnext, 1)] = g;
    }
    return;
  }
  // one safeguarded Newton step (same driver as the oracle)
  bool done = false;
  double xn = x;
  if (!isfinite(score) || !isfinite(d) || score == 0.0) {
    done = true;
  } else {
    double lo = ms.lo[g], hi = ms.hi[g];
    if (score > 0) lo = fmax(lo, x);
    else hi = fmin(hi, x);
    xn = (d < 0) ? x - score / d : (score > 0 ? x + 1.0 : x - 1.0);
    if (!(xn > lo && xn < hi)) xn = 0.5 * (lo + hi);
    ms.lo[g] = lo;
    ms.hi[g] = hi;
    ms.x[g] = xn;
    if (fabs(xn - x) < 1e-9) done = true;
  }
  if (it - 1 >= max_iter) done = true;  // Newton passes = iters - 1 (far-left pass)
  if (done) {
    ms.mode[g] = kMleDone;
    ms.theta[g] = exp(xn);
  } else {
    next[atomicAdd(nnext, 1)] = g;
  }
}

#define DV_MLE_P(p, KERNEL, GRID, BLOCK, STREAM, ...)                                \
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

hipError_t launch_mle_setup(hipStream_t s, Counts c, int64_t G, int64_t n, int p,
                            const MleState& ms) {
  if (G <= 0) return hipSuccess;
  DV_MLE_P(p, k_mle_setup, dim3((unsigned)G), dim3(kSetupThreads), s, c, G, n, ms);
  return hipGetLastError();
}

hipError_t launch_mle_collect(hipStream_t s, const int32_t* mode, int G, int32_t* act,
                              int32_t* nact) {
  if (G <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_mle_collect, dim3((unsigned)((G + 255) / 256)), dim3(256), 0, s, mode,
                     G, act, nact);
  return hipGetLastError();
}

hipError_t launch_mle_accum(hipStream_t s, const TileDesign& td, Counts c, int p,
                            const int32_t* active, const int32_t* nactive, int nA_bound,
                            const double* beta, const double* xs, const int32_t* tab_n,
                            int want_f, double* part) {
  if (nA_bound <= 0) return hipSuccess;
  const int ntile = (nA_bound + 15) / 16;
  const int groups = (ntile + 3) / 4;
  const dim3 grid((unsigned)((int64_t)groups * td.nchunk));
  DV_MLE_P(p, k_mle_accum, grid, dim3(256), s, td, c, active, nactive, beta, xs, tab_n, want_f,
           part);
  return hipGetLastError();
}

hipError_t launch_mle_update(hipStream_t s, const TileDesign& td, int p, const int32_t* active,
                             const int32_t* nactive, int nA_bound, const double* part,
                             const MleState& ms, int64_t n, int max_iter, double rel_tol,
                             int32_t* next, int32_t* nnext) {
  if (nA_bound <= 0) return hipSuccess;
  PortParams pp;
  pp.lo = log(1e-16);
  pp.up = log(1e16);
  pp.rfctol = rel_tol;
  pp.sctol = rel_tol;
  pp.iter_max = max_iter;
  DV_MLE_P(p, k_mle_update, dim3((unsigned)nA_bound), dim3(kMleUpdThreads), s, td, active,
           nactive, part, ms, n, max_iter, pp, next, nnext);
  return hipGetLastError();
}

}  // namespace dv
