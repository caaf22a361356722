// gfx950 FP64 tile passes after the IRLS: method-of-moments overdispersion
// (src/mom_overdispersion.cpp:24-53) and bread + clustered meat
// (src/variance.cpp:22-47, :186-228; R/main.R:481-492).
//
// Same wavefront tiling as the IRLS accumulate (kernels_tile.hip): a wave
// owns 16 genes and one chunk of cells; eta = X beta on
// v_mfma_f64_16x16x4f64 with the permuted design rows, so lane l holds the
// four consecutive cells 4 (l / 16) + r of gene slot l % 16.  Chunks come
// from a chunk table (cell ranges [c0, c1) and their cluster), aligned to
// the patient blocks when the fit is clustered; cells of a 16-cell block
// outside the chunk's range are masked.
//
//   k_mom_accum    per cell mu = s_i exp(x'beta); sums (y - mu)^2 - mu and
//                  mu^2 per (gene, chunk)
//   k_mom_finish   theta = max(0, n/(n-p) num/den), 0 when den <= 0
//   k_bread_accum  w = (y theta + 1) mu / (1 + theta mu)^2 against every K
//                  column and r = (y - mu) / (1 + mu theta) against the design
//                  columns (K tile 0), per (gene, chunk)
//   k_bread_finish H = sum of the w partials, bread = H^-1 (LU), S_c = sum of
//                  the r partials of cluster c's chunks, meat = adj/n sum_c
//                  S_c S_c', sandwich = bread meat bread n
#include <hip/hip_runtime.h>

#include "../../include/devil_hip.h"
#include "device_common.hpp"
#include "kernels.hpp"
#include "tile_common.hpp"

namespace dv {

template <int P, typename CT, bool VEC>
__device__ __forceinline__ void mom_tile_chunk(const TileDesign& td, ChunkTable ct,
                                               const CT* __restrict__ Y, int64_t ld, int64_t G,
                                               const double* __restrict__ beta, int tile,
                                               int chunk, double* __restrict__ part,
                                               const double* __restrict__ e2) {
  constexpr int NJ = (P + 3) / 4;
  const int lane = threadIdx.x & 63;
  const int s = lane & 15, h = lane >> 4;
  const int64_t g0 = (int64_t)tile * 16 + s;
  const bool live = g0 < G;
  const int64_t g = live ? g0 : (int64_t)tile * 16;
  double bpos[NJ];
#pragma unroll
  for (int j = 0; j < NJ; ++j) {
    const int f = 4 * j + h;
    bpos[j] = (live && f < P) ? beta[g * P + f] : 0.0;
  }
  const CT* yrow = Y + g * ld;
  const CT* yh = yrow + 4 * h;
  const int64_t c0 = ct.c0[chunk], c1 = ct.c1[chunk];
  const int64_t b0 = c0 / 16, b1 = (c1 + 15) / 16;
  double num = 0.0, den = 0.0;
  typedef TileBlock<NJ, 0, CT> Blk;  // off holds exp(offset)
  // blocks inside [c0, c1) skip the per-cell range test
  auto body_as = [&](const Blk& cur, int64_t b, auto inner) {
    constexpr bool kInner = decltype(inner)::value;
    d4 eta = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int j = 0; j < NJ; ++j) eta = mfma16x16x4(cur.xa[j], bpos[j], eta);
    const int64_t cb = b * 16 + 4 * h;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      if (kInner || (cb + r >= c0 && cb + r < c1)) {
        const double mu = cur.off[r] * exp_tab(eta[r], e2);
        const double d = (double)cur.y[r] - mu;
        num += d * d - mu;
        den += mu * mu;
      }
    }
  };
  auto body = [&](const Blk& cur, int64_t b) {
    if (b * 16 >= c0 && b * 16 + 16 <= c1) body_as(cur, b, std::true_type{});
    else body_as(cur, b, std::false_type{});
  };
  auto load = [&](int64_t b, Blk& o) {
    load_tile_block<NJ, 0, CT, VEC>(td, yh, b, lane, h, 0, td.eop, o);
  };
  auto load_tail = [&](int64_t b, Blk& o) {
    load_tile_block_tail<NJ, 0, CT>(td, yrow, b, lane, h, 0, td.eop, o);
  };
  if (c0 < c1) pipeline<Blk>(b0, min(b1, td.n / 16), b1, load, load_tail, body);
#pragma unroll
  for (int o = 16; o <= 32; o <<= 1) {
    num += __shfl_xor(num, o, 64);
    den += __shfl_xor(den, o, 64);
  }
  if (h == 0) {
    double* out = part + (((int64_t)tile * ct.nchunk + chunk) * 16 + s) * 2;
    out[0] = num;
    out[1] = den;
  }
}

// HC: also the HC0 meat sum, sum_i r_i^2 x_i x_i' with r = (y - mu) / (1 + theta mu)
// (src/variance.cpp:64-85, 157-178), in ncp more partial columns.
template <int P, typename CT, bool VEC, bool HC>
__device__ __forceinline__ void bread_tile_chunk(const TileDesign& td, ChunkTable ct,
                                                 const CT* __restrict__ Y, int64_t ld,
                                                 int64_t G, const double* __restrict__ beta,
                                                 const double* __restrict__ theta, int tile,
                                                 int chunk, double* __restrict__ part,
                                                 const double* __restrict__ e2) {
  constexpr int NJ = (P + 3) / 4;
  constexpr int NTMAX = ((P * (P + 1)) / 2 + P + 15) / 16;
  const int lane = threadIdx.x & 63;
  const int s = lane & 15, h = lane >> 4;
  const int64_t g0 = (int64_t)tile * 16 + s;
  const bool live = g0 < G;
  const int64_t g = live ? g0 : (int64_t)tile * 16;
  const int NT = td.nt;
  double bpos[NJ];
#pragma unroll
  for (int j = 0; j < NJ; ++j) {
    const int f = 4 * j + h;
    bpos[j] = (live && f < P) ? beta[g * P + f] : 0.0;
  }
  const double th = live ? theta[g] : 0.0;
  const CT* yrow = Y + g * ld;
  const CT* yh = yrow + 4 * h;
  d4 aw[NTMAX], ah[HC ? NTMAX : 1], ar;
#pragma unroll
  for (int t = 0; t < NTMAX; ++t) aw[t] = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int t = 0; t < (HC ? NTMAX : 1); ++t) ah[t] = d4{0.0, 0.0, 0.0, 0.0};
  ar = d4{0.0, 0.0, 0.0, 0.0};
  const int64_t c0 = ct.c0[chunk], c1 = ct.c1[chunk];
  const int64_t b0 = c0 / 16, b1 = c0 < c1 ? (c1 + 15) / 16 : b0;
  typedef TileBlock<NJ, NTMAX, CT> Blk;  // off holds exp(offset)
  auto body = [&](const Blk& cur, int64_t b) {
    d4 eta = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int j = 0; j < NJ; ++j) eta = mfma16x16x4(cur.xa[j], bpos[j], eta);
    const int64_t cb = b * 16 + 4 * h;
    double w[4], rr[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const double mu = cur.off[r] * exp_tab(eta[r], e2);
      const double y = (double)cur.y[r];
      const double d = fma(th, mu, 1.0);
      const double id = rcp_f64(d);
      const bool in = cb + r >= c0 && cb + r < c1;
      w[r] = in ? fma(y, th, 1.0) * mu * id * id : 0.0;
      rr[r] = in ? (y - mu) * id : 0.0;
    }
#pragma unroll
    for (int t = 0; t < NTMAX; ++t) {
      if (t < NT) {
#pragma unroll
        for (int r = 0; r < 4; ++r) aw[t] = mfma16x16x4(w[r], cur.kb[t][r], aw[t]);
      }
    }
    if constexpr (HC) {
#pragma unroll
      for (int t = 0; t < NTMAX; ++t) {
        if (t < NT) {
#pragma unroll
          for (int r = 0; r < 4; ++r) ah[t] = mfma16x16x4(rr[r] * rr[r], cur.kb[t][r], ah[t]);
        }
      }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) ar = mfma16x16x4(rr[r], cur.kb[0][r], ar);
  };
  auto load = [&](int64_t b, Blk& o) {
    load_tile_block<NJ, NTMAX, CT, VEC>(td, yh, b, lane, h, NT, td.eop, o);
  };
  auto load_tail = [&](int64_t b, Blk& o) {
    load_tile_block_tail<NJ, NTMAX, CT>(td, yrow, b, lane, h, NT, td.eop, o);
  };
  pipeline<Blk>(b0, min(b1, td.n / 16), b1, load, load_tail, body);
  // register r holds gene slot h + 4 r, column 16 t + s
  const int W = td.ncp + 16 + (HC ? td.ncp : 0);
  double* out = part + ((int64_t)tile * ct.nchunk + chunk) * 16 * W;
#pragma unroll
  for (int t = 0; t < NTMAX; ++t) {
    if (t < NT) {
#pragma unroll
      for (int r = 0; r < 4; ++r) out[(h + 4 * r) * W + 16 * t + s] = aw[t][r];
    }
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) out[(h + 4 * r) * W + td.ncp + s] = ar[r];
  if constexpr (HC) {
#pragma unroll
    for (int t = 0; t < NTMAX; ++t) {
      if (t < NT) {
#pragma unroll
        for (int r = 0; r < 4; ++r) out[(h + 4 * r) * W + td.ncp + 16 + 16 * t + s] = ah[t][r];
      }
    }
  }
}

template <int P>
__global__ __launch_bounds__(256) void k_mom_accum(TileDesign td, ChunkTable ct, Counts c,
                                                   int64_t G, const double* __restrict__ beta,
                                                   double* __restrict__ part) {
  __shared__ double e2[64];
  exp_tab_fill(e2, threadIdx.x, blockDim.x);
  __syncthreads();
  const int wave = wave_index();
  int chunk, group;
  tile_coords(ct.nchunk, chunk, group);
  const int tile = group * 4 + wave;
  if ((int64_t)tile * 16 >= G) return;
  with_counts(c, [&](auto Y, auto vec) {
    typedef std::remove_cv_t<std::remove_pointer_t<decltype(Y)>> CT;
    mom_tile_chunk<P, CT, decltype(vec)::value>(td, ct, Y, c.ld, G, beta, tile, chunk, part,
                                                e2);
  });
}

template <int P>
__global__ __launch_bounds__(256) void k_mom_finish(ChunkTable ct, int64_t G, int64_t n,
                                                    const double* __restrict__ part,
                                                    double* __restrict__ theta,
                                                    int32_t* __restrict__ status) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= G) return;
  const int64_t tile = g / 16, s = g % 16;
  double num = 0.0, den = 0.0;
  for (int ch = 0; ch < ct.nchunk; ++ch) {
    const double* src = part + ((tile * ct.nchunk + ch) * 16 + s) * 2;
    num += src[0];
    den += src[1];
  }
  const double corr = (double)n / (double)(n - P);
  double th = 0.0;
  if (den > 0.0) {
    th = corr * num / den;
    if (th < 0.0) th = 0.0;
  }
  theta[g] = th;
  if (status && !(den > 0.0)) status[g] |= DEVIL_GENE_MOM_DEN0;
}

template <int P, bool HC>
__global__ __launch_bounds__(256) void k_bread_accum(TileDesign td, ChunkTable ct, Counts c,
                                                     int64_t G, const double* __restrict__ beta,
                                                     const double* __restrict__ theta,
                                                     double* __restrict__ part) {
  __shared__ double e2[64];
  exp_tab_fill(e2, threadIdx.x, blockDim.x);
  __syncthreads();
  const int wave = wave_index();
  int chunk, group;
  tile_coords(ct.nchunk, chunk, group);
  const int tile = group * 4 + wave;
  if ((int64_t)tile * 16 >= G) return;
  with_counts(c, [&](auto Y, auto vec) {
    typedef std::remove_cv_t<std::remove_pointer_t<decltype(Y)>> CT;
    bread_tile_chunk<P, CT, decltype(vec)::value, HC>(td, ct, Y, c.ld, G, beta, theta, tile,
                                                      chunk, part, e2);
  });
}

// One wavefront per gene.  C = number of clusters (0: no sandwich; -1: the
// HC0 meat, sum_i r_i^2 x_i x_i' / n, from the partials' last ncp columns).
// bread_given: the bread is already in `bread` (the mixed path's): only the
// meat and the sandwich are formed.
template <int P>
__global__ __launch_bounds__(64) void k_bread_finish(TileDesign td, ChunkTable ct, int64_t G,
                                                     int64_t n, int C,
                                                     const double* __restrict__ part,
                                                     double* __restrict__ bread,
                                                     double* __restrict__ sandwich,
                                                     int32_t* __restrict__ status,
                                                     int bread_given) {
  constexpr int NPAIR = P * (P + 1) / 2;
  __shared__ double accw[DEVIL_NCP_MAX];
  __shared__ double acch[DEVIL_NCP_MAX];
  __shared__ double sc[16];
  __shared__ double meat[NPAIR];
  const int64_t g = blockIdx.x;
  if (g >= G) return;
  const int lane = threadIdx.x;
  const bool hc0 = C < 0;
  const int W = td.ncp + 16 + (hc0 ? td.ncp : 0);
  const double* src = part + (g / 16) * ct.nchunk * 16 * W + (g % 16) * W;
  for (int col = lane; col < td.ncp && !bread_given; col += 64) {
    double a = 0.0, m = 0.0;
    for (int ch = 0; ch < ct.nchunk; ++ch) {
      a += src[(int64_t)ch * 16 * W + col];
      if (hc0) m += src[(int64_t)ch * 16 * W + td.ncp + 16 + col];
    }
    accw[col] = a;
    if (hc0) acch[col] = m;
  }
  // lane t < NPAIR owns meat entry t = tri(a, b)
  int ma = 0, mb = 0;
  if (lane < NPAIR) {
    int t = lane, a = 0;
    while (t >= P - a) {
      t -= P - a;
      ++a;
    }
    ma = a;
    mb = a + t;
    meat[lane] = 0.0;
  }
  if (C > 0 && sandwich) {
    int cl = ct.cl[0];
    double S = 0.0;
    for (int ch = 0; ch <= ct.nchunk; ++ch) {
      const bool flush = ch == ct.nchunk || ct.cl[ch] != cl;
      if (flush) {
        if (lane < 16) sc[lane] = S;
        __syncthreads();
        if (lane < NPAIR) {
          const int ia = td.kidx_lin[ma], ib = td.kidx_lin[mb];
          meat[lane] += (ia >= 0 ? sc[ia] : 0.0) * (ib >= 0 ? sc[ib] : 0.0);
        }
        __syncthreads();
        S = 0.0;
        if (ch == ct.nchunk) break;
        cl = ct.cl[ch];
      }
      if (lane < 16) S += src[(int64_t)ch * 16 * W + td.ncp + lane];
    }
  }
  __syncthreads();
  if (lane != 0) return;
  double Hm[P][P], Bi[P][P];
  double* out = bread + g * P * P;
  if (bread_given) {
#pragma unroll
    for (int a = 0; a < P; ++a)
#pragma unroll
      for (int bb = 0; bb < P; ++bb) Bi[a][bb] = out[a + bb * P];
  } else {
#pragma unroll
    for (int a = 0; a < P; ++a)
#pragma unroll
      for (int bb = a; bb < P; ++bb) {
        const int pi = td.kidx_pair[tri<P>(a, bb)];
        const double v = pi >= 0 ? accw[pi] : 0.0;
        Hm[a][bb] = v;
        Hm[bb][a] = v;
      }
    const int sing = lu_inverse<P>(Hm, Bi);
#pragma unroll
    for (int a = 0; a < P; ++a)
#pragma unroll
      for (int bb = 0; bb < P; ++bb) out[a + bb * P] = Bi[a][bb];
    if (status && sing) status[g] |= DEVIL_GENE_SINGULAR;
  }
  if (C != 0 && sandwich) {
    const double adj = (C > 1) ? (double)C / (C - 1.0) : 1.0;
    const double scl = (hc0 ? 1.0 : adj) / (double)n;
    double Me[P][P], T1[P][P];
#pragma unroll
    for (int a = 0; a < P; ++a)
#pragma unroll
      for (int bb = a; bb < P; ++bb) {
        double v;
        if (hc0) {
          const int pi = td.kidx_pair[tri<P>(a, bb)];
          v = scl * (pi >= 0 ? acch[pi] : 0.0);
        } else {
          v = scl * meat[tri<P>(a, bb)];
        }
        Me[a][bb] = v;
        Me[bb][a] = v;
      }
#pragma unroll
    for (int a = 0; a < P; ++a)
#pragma unroll
      for (int bb = 0; bb < P; ++bb) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < P; ++k) s += Bi[a][k] * Me[k][bb];
        T1[a][bb] = s;
      }
    double* so = sandwich + g * P * P;
#pragma unroll
    for (int a = 0; a < P; ++a)
#pragma unroll
      for (int bb = 0; bb < P; ++bb) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < P; ++k) s += T1[a][k] * Bi[k][bb];
        so[a + bb * P] = s * (double)n;
      }
  }
}

// ---------------------------------------------------------------------------
// Init pass (R/overdispersion.R:129-135, R/beta.R:13-20, R/main.R:404-406,
// src/beta_init_univariate.cpp:13-95) on the same tiling: per (gene, chunk)
// the count sums sum y, sum y^2 (exact 64-bit integers for int32 counts) and
// z' Z on v_mfma_f64_16x16x4f64, where Z holds the p columns of the
// projection (thin Q of X for the least-squares start, X - mean(X) for the
// univariate start) and z = log1p(y / s) or log(y + 0.5) - log s.
// ---------------------------------------------------------------------------
// ZM: 0 no projection (alpha only, rough start), 1 least squares (QR),
// 2 univariate.
template <typename CT, int ZM, bool VEC>
__device__ __forceinline__ void init_tile_chunk(const TileDesign& td, InitTile it, ChunkTable ct,
                                                const CT* __restrict__ Y, int64_t ld, int64_t G,
                                                int tile, int chunk, const LogTabEntry* ltab,
                                                double* __restrict__ part) {
  const int lane = threadIdx.x & 63;
  const int s = lane & 15, h = lane >> 4;
  const int64_t g0 = (int64_t)tile * 16 + s;
  const int64_t g = g0 < G ? g0 : (int64_t)tile * 16;
  const CT* yrow = Y + g * ld;
  const CT* yh = yrow + 4 * h;
  const int64_t c0 = ct.c0[chunk], c1 = ct.c1[chunk];
  const int64_t b0 = c0 / 16, b1 = c0 < c1 ? (c1 + 15) / 16 : b0;
  constexpr bool kInt = sizeof(CT) == 4 && CT(0.5) == CT(0);  // integer counts
  d4 az = d4{0.0, 0.0, 0.0, 0.0};
  // integer counts: sum y in int64, sum y^2 in 128 bits (lo, hi words; a
  // block's four squares, each below 2^62, add up without overflow in 64)
  int64_t s1i = 0;
  uint64_t s2lo = 0, s2hi = 0;
  double s1d = 0.0, s2d = 0.0;
  struct Blk {
    CT y[4];
    d4 zb;   // projection columns of the block (B operand)
    d4 aux;  // exp(-offset) (QR) or offset (univariate) of the lane's cells
  };
  auto load_tables = [&](int64_t b, Blk& o) {
    if constexpr (ZM != 0) {
      o.zb = reinterpret_cast<const d4*>(it.Zp + b * 256)[lane];
      o.aux = reinterpret_cast<const d4*>((ZM == 1 ? it.einv : it.logsf) + b * 16)[h];
    }
  };
  auto load = [&](int64_t b, Blk& o) {
    load_tables(b, o);
    load4_fast<CT, VEC>(yh + b * 16, o.y);
  };
  auto load_tail = [&](int64_t b, Blk& o) {
    load_tables(b, o);
    load4_tail<CT>(yrow, b * 16 + 4 * h, td.n, o.y);
  };
  // blocks inside [c0, c1) (all but a chunk's first and last) skip the
  // per-cell range test; log1p(0) = 0 exactly, so zero counts need no branch
  auto body_as = [&](const Blk& o, int64_t b, auto inner) {
    constexpr bool kInner = decltype(inner)::value;
    const int64_t cb = b * 16 + 4 * h;
    double z[4];
    uint64_t sq = 0;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const bool in = kInner || (cb + r >= c0 && cb + r < c1);
      const double y = in ? (double)o.y[r] : 0.0;
      if constexpr (ZM == 1) {
        z[r] = log1p_tab(y * o.aux[r], ltab);
      } else if constexpr (ZM == 2) {
        z[r] = in ? log_tab(y + 0.5, 0.0, ltab) - o.aux[r] : 0.0;
      } else {
        z[r] = 0.0;
      }
      if constexpr (kInt) {
        const int64_t yi = in ? (int64_t)o.y[r] : 0;
        s1i += yi;
        sq += (uint64_t)(yi * yi);
      } else {
        s1d += y;
        s2d += y * y;
      }
    }
    if constexpr (kInt) {
      s2lo += sq;
      s2hi += s2lo < sq ? 1 : 0;
    }
    if constexpr (ZM != 0) {
#pragma unroll
      for (int r = 0; r < 4; ++r) az = mfma16x16x4(z[r], o.zb[r], az);
    }
  };
  auto body = [&](const Blk& o, int64_t b) {
    if (b * 16 >= c0 && b * 16 + 16 <= c1) body_as(o, b, std::true_type{});
    else body_as(o, b, std::false_type{});
  };
  pipeline<Blk>(b0, min(b1, td.n / 16), b1, load, load_tail, body);
#pragma unroll
  for (int o = 16; o <= 32; o <<= 1) {
    s1i += __shfl_xor(s1i, o, 64);
    const uint64_t lo = __shfl_xor(s2lo, o, 64), hi = __shfl_xor(s2hi, o, 64);
    s2lo += lo;
    s2hi += hi + (s2lo < lo ? 1 : 0);
    s1d += __shfl_xor(s1d, o, 64);
    s2d += __shfl_xor(s2d, o, 64);
  }
  // register r holds gene slot h + 4 r, column s
  double* out = part + ((int64_t)tile * ct.nchunk + chunk) * 16 * 20;
#pragma unroll
  for (int r = 0; r < 4; ++r) out[(h + 4 * r) * 20 + s] = az[r];
  if (h == 0) {
    double* o = out + s * 20 + 16;
    o[0] = s1d;
    if constexpr (kInt) reinterpret_cast<uint64_t*>(o)[1] = s2hi;
    else o[1] = s2d;
    reinterpret_cast<int64_t*>(o)[2] = s1i;
    reinterpret_cast<uint64_t*>(o)[3] = s2lo;
  }
}

template <int ZM>
__global__ __launch_bounds__(256) void k_init_accum(TileDesign td, InitTile it, ChunkTable ct,
                                                    Counts c, int64_t G,
                                                    double* __restrict__ part) {
  __shared__ LogTabEntry ltab[kLogTab];
  if constexpr (ZM != 0) {
    log_tab_fill(ltab, threadIdx.x, blockDim.x);
    __syncthreads();
  }
  const int wave = wave_index();
  int chunk, group;
  tile_coords(ct.nchunk, chunk, group);
  const int tile = group * 4 + wave;
  if ((int64_t)tile * 16 >= G) return;
  with_counts(c, [&](auto Y, auto vec) {
    typedef std::remove_cv_t<std::remove_pointer_t<decltype(Y)>> CT;
    init_tile_chunk<CT, ZM, decltype(vec)::value>(td, it, ct, Y, c.ld, G, tile, chunk, ltab,
                                                  part);
  });
}

// n * s2 - s1^2 exactly (128-bit; s2 < n 2^62, so n * s2 fits for n < 2^32),
// then one rounding
__device__ __forceinline__ double exact_ssd(int64_t s1, unsigned __int128 s2, int64_t n) {
  const __int128 v = (__int128)n * (__int128)s2 - (__int128)s1 * (__int128)s1;
  return (double)v;
}

template <int P>
__global__ __launch_bounds__(256) void k_init_finish(InitTile it, ChunkTable ct, int64_t G,
                                                     int64_t n, int integer_counts,
                                                     const double* __restrict__ part,
                                                     double* __restrict__ alpha0,
                                                     double* __restrict__ beta0) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= G) return;
  const int64_t tile = g / 16, s = g % 16;
  double w[P], s1d = 0.0, s2d = 0.0;
  int64_t s1i = 0;
  unsigned __int128 s2i = 0;
#pragma unroll
  for (int j = 0; j < P; ++j) w[j] = 0.0;
  for (int ch = 0; ch < ct.nchunk; ++ch) {
    const double* src = part + ((tile * ct.nchunk + ch) * 16 + s) * 20;
#pragma unroll
    for (int j = 0; j < P; ++j) w[j] += src[j];
    s1d += src[16];
    s1i += reinterpret_cast<const int64_t*>(src)[18];
    if (integer_counts) {
      const uint64_t hi = reinterpret_cast<const uint64_t*>(src)[17];
      const uint64_t lo = reinterpret_cast<const uint64_t*>(src)[19];
      s2i += ((unsigned __int128)hi << 64) | lo;
    } else {
      s2d += src[17];
    }
  }
  double s1, ssd;  // sum y and sum (y - mean)^2
  if (integer_counts) {
    s1 = (double)s1i;
    ssd = exact_ssd(s1i, s2i, n) / (double)n;
  } else {
    s1 = s1d;
    ssd = s2d - s1d * s1d / (double)n;
  }
  const double mean = s1 / (double)n;
  if (it.want_alpha) {
    const double v = ssd / (double)(n - 1);
    const double d = (v - it.moi * mean) / (mean * mean);
    alpha0[g] = (isnan(d) || d < 0.0) ? 0.01 : d;
  }
  if (it.mode == kInitNone) return;
  double b[P];
  if (it.mode == kInitQR) {
#pragma unroll
    for (int a = P - 1; a >= 0; --a) {
      double t = w[a];
#pragma unroll
      for (int k = a + 1; k < P; ++k) t -= it.R[a + k * P] * b[k];
      b[a] = t / it.R[a + a * P];
    }
  } else if (it.mode == kInitRough) {
    b[0] = log1p(mean);
#pragma unroll
    for (int j = 1; j < P; ++j) b[j] = 0.0;
  } else {
#pragma unroll
    for (int j = 0; j < P; ++j) b[j] = w[j] / it.denom[j];
    if (it.intercept_first) b[0] = log((s1 + 1e-8) / (it.sum_sf + 1e-8));
  }
#pragma unroll
  for (int j = 0; j < P; ++j) beta0[g * P + j] = b[j];
}

#define DV_POST_P(p, KERNEL, GRID, BLOCK, STREAM, ...)                               \
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

// the same dispatch for kernels with a second (bool) template argument
#define DV_POST_P2_CASE(P_, KERNEL, B, GRID, BLOCK, STREAM, ...)    \
  case P_: {                                                        \
    auto k_ = KERNEL<P_, B>;                                        \
    hipLaunchKernelGGL(k_, GRID, BLOCK, 0, STREAM, __VA_ARGS__);    \
  } break;
#define DV_POST_P2(p, KERNEL, B, GRID, BLOCK, STREAM, ...)                     \
  switch (p) {                                                                \
    DV_POST_P2_CASE(1, KERNEL, B, GRID, BLOCK, STREAM, __VA_ARGS__)           \
    DV_POST_P2_CASE(2, KERNEL, B, GRID, BLOCK, STREAM, __VA_ARGS__)           \
    DV_POST_P2_CASE(3, KERNEL, B, GRID, BLOCK, STREAM, __VA_ARGS__)           \
    DV_POST_P2_CASE(4, KERNEL, B, GRID, BLOCK, STREAM, __VA_ARGS__)           \
    DV_POST_P2_CASE(5, KERNEL, B, GRID, BLOCK, STREAM, __VA_ARGS__)           \
    DV_POST_P2_CASE(6, KERNEL, B, GRID, BLOCK, STREAM, __VA_ARGS__)           \
    DV_POST_P2_CASE(7, KERNEL, B, GRID, BLOCK, STREAM, __VA_ARGS__)           \
    DV_POST_P2_CASE(8, KERNEL, B, GRID, BLOCK, STREAM, __VA_ARGS__)           \
    DV_POST_P2_CASE(9, KERNEL, B, GRID, BLOCK, STREAM, __VA_ARGS__)           \
    DV_POST_P2_CASE(10, KERNEL, B, GRID, BLOCK, STREAM, __VA_ARGS__)          \
    default: return hipErrorInvalidValue;                                     \
  }

hipError_t launch_mom_tiled(hipStream_t s, const TileDesign& td, const ChunkTable& ct, Counts c,
                            int64_t G, int64_t n, int p, const double* beta, double* part,
                            double* theta, int32_t* status) {
  if (G <= 0) return hipSuccess;
  const int64_t groups = ((G + 15) / 16 + 3) / 4;
  DV_POST_P(p, k_mom_accum, dim3((unsigned)(groups * ct.nchunk)), dim3(256), s, td, ct, c, G,
            beta, part);
  DV_POST_P(p, k_mom_finish, dim3((unsigned)((G + 255) / 256)), dim3(256), s, ct, G, n, part,
            theta, status);
  return hipGetLastError();
}

hipError_t launch_bread_tiled(hipStream_t s, const TileDesign& td, const ChunkTable& ct,
                              Counts c, int64_t G, int64_t n, int p, const double* beta,
                              const double* theta, int C, double* part, double* bread,
                              double* sandwich, int32_t* status) {
  if (G <= 0) return hipSuccess;
  const int64_t groups = ((G + 15) / 16 + 3) / 4;
  if (C < 0) {
    DV_POST_P2(p, k_bread_accum, true, dim3((unsigned)(groups * ct.nchunk)), dim3(256), s, td,
               ct, c, G, beta, theta, part);
  } else {
    DV_POST_P2(p, k_bread_accum, false, dim3((unsigned)(groups * ct.nchunk)), dim3(256), s, td,
               ct, c, G, beta, theta, part);
  }
  DV_POST_P(p, k_bread_finish, dim3((unsigned)G), dim3(64), s, td, ct, G, n, C, part, bread,
            sandwich, status, 0);
  return hipGetLastError();
}

hipError_t launch_meat_finish(hipStream_t s, const TileDesign& td, const ChunkTable& ct,
                              int64_t G, int64_t n, int p, int C, const double* spart,
                              const double* bread, double* sandwich) {
  if (G <= 0) return hipSuccess;
  if (C <= 0 || !sandwich) return hipErrorInvalidValue;
  // the score partials alone: 16 columns per row, design column a at a
  TileDesign tc = td;
  tc.ncp = 0;
  DV_POST_P(p, k_bread_finish, dim3((unsigned)G), dim3(64), s, tc, ct, G, n, C, spart,
            const_cast<double*>(bread), sandwich, nullptr, 1);
  return hipGetLastError();
}

hipError_t launch_init_tiled(hipStream_t s, const TileDesign& td, const InitTile& it,
                             const ChunkTable& ct, Counts c, int64_t G, int64_t n, int p,
                             double* part, double* alpha0, double* beta0) {
  if (G <= 0) return hipSuccess;
  const int64_t groups = ((G + 15) / 16 + 3) / 4;
  const dim3 grid((unsigned)(groups * ct.nchunk));
  if (it.mode == kInitQR)
    hipLaunchKernelGGL(k_init_accum<1>, grid, dim3(256), 0, s, td, it, ct, c, G, part);
  else if (it.mode == kInitUnivariate)
    hipLaunchKernelGGL(k_init_accum<2>, grid, dim3(256), 0, s, td, it, ct, c, G, part);
  else
    hipLaunchKernelGGL(k_init_accum<0>, grid, dim3(256), 0, s, td, it, ct, c, G, part);
  DV_POST_P(p, k_init_finish, dim3((unsigned)((G + 255) / 256)), dim3(256), s, it, ct, G, n,
            c.dtype == 0 ? 1 : 0, part, alpha0, beta0);
  return hipGetLastError();
}

}  // namespace dv
