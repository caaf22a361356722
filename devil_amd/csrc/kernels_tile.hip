// gfx950 FP64 IRLS with the contractions on the matrix cores.
//
// One IRLS round (src/beta.cpp:38-46) over the still-active genes is two
// launches:
//
//   k_irls_accum  one wavefront per (16-gene tile, cell chunk).  For every
//                 16-cell block:
//                   eta' = -X beta      v_mfma_f64_16x16x4f64, D[cell][gene]
//                   w = exp(eta' - off), r = m w - 1 = (y w - 1) / (1 + k w)
//                   acc += r' K         v_mfma_f64_16x16x4f64 per 16 columns
//                 where K holds the distinct columns x_a x_b (a <= b) and x_a
//                 of the design (the Khatri-Rao product, deduplicated on the
//                 host).  The tile's 16 x NCP partial sums are written per
//                 chunk.
//   k_irls_update one wavefront per active gene: sums its chunk partials in
//                 chunk order, forms A = k (X'diag(r)X + X'X) and g = k X'r
//                 (the column sums of K are precomputed in FP64), solves
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
//
// The file also holds the mixed-precision accumulates on the same tiling
// (k_irls_accum_mixed: FP32 MFMA; k_irls_accum_bf16: bf16 MFMA from split
// operands; k_irls_accum_ws / _ws2: the bf16 arithmetic with operands staged
// in LDS by direct-to-LDS loads), the mixed bread (k_bread_accum_ws) and,
// with patients, the bread and per-patient score sums in one pass
// (k_bread_meat_ws), and the joint launches that serve many problems.
#include <hip/hip_runtime.h>

#include "../../include/devil_hip.h"
#include "device_common.hpp"
#include "kernels.hpp"
#include "tile_common.hpp"

namespace dv {

// lane l of a wave: gene slot s = l % 16, cell group h = l / 16; its cells in
// block b are 16 b + 4 h + r.
template <int P, typename CT, bool VEC>
__device__ __forceinline__ void irls_tile_chunk(const TileDesign& td, const CT* __restrict__ Y,
                                                int64_t ld, const int32_t* __restrict__ active,
                                                int nA, const double* __restrict__ beta,
                                                const double* __restrict__ kvec, int tile,
                                                int chunk, double* __restrict__ part,
                                                const double* __restrict__ e2) {
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
  const CT* yh = yrow + 4 * h;
  d4 acc[NTMAX];
#pragma unroll
  for (int t = 0; t < NTMAX; ++t) acc[t] = d4{0.0, 0.0, 0.0, 0.0};
  const int64_t b0 = (int64_t)chunk * td.chunk_blocks;
  const int64_t b1 = min((int64_t)(chunk + 1) * td.chunk_blocks, td.nblk);
  typedef TileBlock<NJ, NTMAX, CT> Blk;
  auto body = [&](const Blk& cur, int64_t) {
    d4 eta = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int j = 0; j < NJ; ++j) eta = mfma16x16x4(cur.xa[j], bneg[j], eta);
    // m w - 1 = (y w - 1) / (1 + k w), summed without the n-term
    // cancellation of sum (m w) - sum 1
    double q[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const double w = exp_tab(eta[r] - cur.off[r], e2);
      const double d = fma(k, w, 1.0);
      q[r] = fma((double)cur.y[r], w, -1.0) * rcp1_f64(d);
    }
#pragma unroll
    for (int t = 0; t < NTMAX; ++t) {
      if (t < NT) {
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[t] = mfma16x16x4(q[r], cur.kb[t][r], acc[t]);
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
// r = (y - mu) / (k + mu) (v_rcp_f32) -- the reference's m w - 1 written with
// mu; mu = 0 gives y / k as in FP64, an overflowed mu gives NaN, which sends
// the gene to the FP64 recheck -- and r' K on v_mfma_f32_16x16x4f32 with FP32
// accumulators flushed into FP64 every 4 blocks (64 cells).
// v_mfma_f32_16x16x4f32 (checked on gfx950): lane l supplies A[l%16][l/16]
// and B[l/16][l%16]; register r of the result holds D[4 (l/16) + r][l%16].
// ---------------------------------------------------------------------------
typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f4 mfma32(float a, float b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

template <int NJ, int NTMAX, typename CT>
struct BlockOpsF {
  CT y[4];
  float xa[NJ];
  f4 off;
  f4 kb[NTMAX];
};

template <int NJ, int NTMAX, typename CT>
__device__ __forceinline__ void load_block_f(const TileDesign& td, int64_t b, int lane, int h,
                                             int NT, BlockOpsF<NJ, NTMAX, CT>& o) {
  const float* xs = td.Xf + b * (NJ * 64);
#pragma unroll
  for (int j = 0; j < NJ; ++j) o.xa[j] = xs[j * 64 + lane];
  o.off = reinterpret_cast<const f4*>(td.offf + b * 16)[h];
  const f4* ks = reinterpret_cast<const f4*>(td.Kf) + b * NT * 64;
#pragma unroll
  for (int t = 0; t < NTMAX; ++t) {
    const int tt = t < NT ? t : NT - 1;
    o.kb[t] = ks[tt * 64 + lane];
  }
}

template <int P, typename CT, bool VEC>
__device__ __forceinline__ void irls_tile_chunk_mixed(const TileDesign& td,
                                                      const CT* __restrict__ Y, int64_t ld,
                                                      const int32_t* __restrict__ active, int nA,
                                                      const double* __restrict__ beta,
                                                      const double* __restrict__ kvec, int tile,
                                                      int chunk, double* __restrict__ part) {
  constexpr int NJ = (P + 3) / 4;
  constexpr int NTMAX = ((P * (P + 1)) / 2 + P + 15) / 16;
  constexpr int kSpan = 4;  // blocks per FP32 accumulation span (64 cells)
  const int lane = threadIdx.x & 63;
  const int s = lane & 15, h = lane >> 4;
  const int slot = tile * 16 + s;
  const bool live = slot < nA;
  const int g = live ? active[slot] : active[tile * 16];
  const int NT = td.nt;
  // beta = bhi + blo in two floats: rounding beta to one float would shift
  // every gene's evaluation point by 2^-24 |beta| per component, and through
  // the iteration's coupling that reaches the small coefficients
  float bhi[NJ], blo[NJ];
#pragma unroll
  for (int j = 0; j < NJ; ++j) {
    const int f = 4 * j + h;
    const double bd = (live && f < P) ? beta[(int64_t)g * P + f] : 0.0;
    bhi[j] = (float)bd;
    blo[j] = (float)(bd - (double)bhi[j]);
  }
  const float k = live ? (float)kvec[g] : 0.0f;  // k is a float in the reference
  const CT* yrow = Y + (int64_t)g * ld;
  const CT* yh = yrow + 4 * h;
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
  typedef BlockOpsF<NJ, NTMAX, CT> Blk;
  int span = 0;
  float tmin = 0.0f;  // smallest log2(mu) seen by this lane
  auto body = [&](const Blk& cur, int64_t b) {
    f4 eta = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
      eta = mfma32(cur.xa[j], blo[j], eta);
      eta = mfma32(cur.xa[j], bhi[j], eta);
    }
    // m w - 1 = (y - mu) / (k + mu): no cancellation against the design sums,
    // and relative FP32 error even where mu << k
    float q[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const float t2 = (eta[r] + cur.off[r]) * 1.44269504088896341f;
      tmin = fminf(tmin, t2);
      const float mu = __builtin_amdgcn_exp2f(t2);
      q[r] = ((float)cur.y[r] - mu) * __builtin_amdgcn_rcpf(k + mu);
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
  };
  auto load = [&](int64_t b, Blk& o) {
    load_block_f<NJ, NTMAX, CT>(td, b, lane, h, NT, o);  // tables first
    load4_fast<CT, VEC>(yh + b * 16, o.y);
  };
  auto load_tail = [&](int64_t b, Blk& o) {
    load_block_f<NJ, NTMAX, CT>(td, b, lane, h, NT, o);
    load4_tail<CT>(yrow, b * 16 + 4 * h, td.n, o.y);
  };
  pipeline<Blk>(b0, min(b1, td.n / 16), b1, load, load_tail, body);  // trailing chunks may be empty
  // register r holds gene slot 4 h + r, column 16 t + s
  // exp(-eta - off) overflows in FP64 where log2(mu) < -1024: the reference
  // then takes a NaN step, so such a gene's partials are made NaN and the
  // gene is redone in FP64.  Register r of this lane holds gene slot 4 h + r.
  tmin = fminf(tmin, __shfl_xor(tmin, 16, 64));
  tmin = fminf(tmin, __shfl_xor(tmin, 32, 64));
  bool poison[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) poison[r] = __shfl(tmin, 4 * h + r, 64) < -1024.0f;
  double* out = part + ((int64_t)tile * td.nchunk + chunk) * 16 * td.ncp;
#pragma unroll
  for (int t = 0; t < NTMAX; ++t) {
    if (t < NT) {
#pragma unroll
      for (int r = 0; r < 4; ++r)
        out[(4 * h + r) * td.ncp + 16 * t + s] = poison[r] ? __builtin_nan("") : acc64[t][r];
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
  const int wave = wave_index();
  int chunk, group;
  tile_coords(td.nchunk, chunk, group);
  const int tile = group * 4 + wave;
  if (tile >= ntile) return;
  with_counts(c, [&](auto Y, auto vec) {
    typedef std::remove_cv_t<std::remove_pointer_t<decltype(Y)>> CT;
    irls_tile_chunk_mixed<P, CT, decltype(vec)::value>(td, Y, c.ld, active, nA, beta, kvec, tile,
                                                       chunk, part);
  });
}

// ---------------------------------------------------------------------------
// Mixed precision accumulate, bf16 variant (gfx950 v_mfma_f32_16x16x32_bf16).
// A step is 32 cells (two 16-cell halves); lane l = (gene slot s = l % 16,
// group h = l / 16) owns cells 4h + r and 16 + 4h + r, r < 4.
//   eta: per half, sum over split entries of A (cells x 32 entries) B (entries
//        x genes): x = x1 + x2 + x3 and log2(e) beta = b1 + b2 + b3 in bf16,
//        entries (x1 b1, x1 b2, x1 b3, x2 b1, x2 b2, x3 b1) per feature (below
//        2^-25 |x beta| from the product), FP32 accumulation; register r holds
//        cell 4h + r of the half.
//   r = (y - mu) / (k + mu), mu = 2^(eta + off log2 e)             (FP32 VALU)
//   gradient tile (the p design columns): r' X from r = r1 + r2 + r3 and
//        X = X1 + X2 + X3 in bf16, the six products of total order <= 2
//        (below 2^-26 relative error per product: FP32-product accuracy)
//   pair tiles (x_a x_b): three bf16 products r1 K1 + r1 K2 + r2 K1 with
//        r = r1 + r2 and K = K1 + K2 split into bf16 (about 16 significant
//        bits per product), k index (h, i) -> cell
//        c(h, i) = i < 4 ? 4h + i : 16 + 4h + i - 4
// FP32 accumulators are flushed to FP64 every 4 steps (128 cells).
// ---------------------------------------------------------------------------
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ f4 mfma_bf(bf8 a, bf8 b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

typedef float f2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf2 __attribute__((ext_vector_type(2)));
typedef unsigned int u4 __attribute__((ext_vector_type(4)));

// x = x1 + x2 + x3 in bf16, each piece the rounding of what the earlier ones
// leave (the differences are exact in FP32), so |x - x1 - x2 - x3| <= 2^-27
// |x|.  Pairs are converted together (one v_cvt_pk_bf16_f32 each) and read
// back as FP32 from the packed word (shift / mask).
__device__ __forceinline__ float bf_lo_f(unsigned int u) { return __builtin_bit_cast(float, u << 16); }
__device__ __forceinline__ float bf_hi_f(unsigned int u) {
  return __builtin_bit_cast(float, u & 0xffff0000u);
}
__device__ __forceinline__ unsigned int bf_pack(float a, float b) {
  return __builtin_bit_cast(unsigned int, __builtin_convertvector(f2{a, b}, bf2));
}

__device__ __forceinline__ void bf_split8_2(const float (&x)[8], bf8& x1, bf8& x2) {
  u4 a, b;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const unsigned int u = bf_pack(x[2 * i], x[2 * i + 1]);
    a[i] = u;
    b[i] = bf_pack(x[2 * i] - bf_lo_f(u), x[2 * i + 1] - bf_hi_f(u));
  }
  x1 = __builtin_bit_cast(bf8, a);
  x2 = __builtin_bit_cast(bf8, b);
}

__device__ __forceinline__ void bf_split8_3(const float (&x)[8], bf8& x1, bf8& x2, bf8& x3) {
  u4 a, b, c;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const unsigned int u = bf_pack(x[2 * i], x[2 * i + 1]);
    const float d0 = x[2 * i] - bf_lo_f(u), d1 = x[2 * i + 1] - bf_hi_f(u);
    const unsigned int v = bf_pack(d0, d1);
    a[i] = u;
    b[i] = v;
    c[i] = bf_pack(d0 - bf_lo_f(v), d1 - bf_hi_f(v));
  }
  x1 = __builtin_bit_cast(bf8, a);
  x2 = __builtin_bit_cast(bf8, b);
  x3 = __builtin_bit_cast(bf8, c);
}

// Operands of one 32-cell step for lane (s, h): counts (as stored) of cells
// 4h + r and 16 + 4h + r, the eta A operands of both halves, the offsets
// (times log2 e), the gradient-tile and pair-tile B operands.
template <int NE, int NTH>
struct BfTables {
  bf8 xa[2 * NE];
  f4 off0, off1;
  bf8 kg[3];
  bf8 k1[NTH], k2[NTH];
};

template <typename CT>
struct BfCounts {
  CT y[8];
};

// One step's contractions of the cell scalars rr (residuals or bread
// weights) of cells c(h, i).  Gradient tile (the design columns), GRAD3:
// r and x each in three bf16 pieces, the six products of total order <= 2
// (smallest first); what is dropped and the split residues are below 2^-26
// of |r x|, under the rounding of an FP32 product.  Without GRAD3 (the
// bread, whose design columns hold pair products too) and in the pair
// tiles: r1 K1 + r2 K1 + r1 K2 (about 16 significant bits per product).
template <typename S>
__device__ __forceinline__ void bf_grad3(const float (&rr)[8], const S& o, f4& ag) {
  bf8 r1, r2, r3;
  bf_split8_3(rr, r1, r2, r3);
  ag = mfma_bf(r3, o.kg[0], ag);
  ag = mfma_bf(r2, o.kg[1], ag);
  ag = mfma_bf(r1, o.kg[2], ag);
  ag = mfma_bf(r2, o.kg[0], ag);
  ag = mfma_bf(r1, o.kg[1], ag);
  ag = mfma_bf(r1, o.kg[0], ag);
}

template <int NTH, bool GRAD3, typename S>
__device__ __forceinline__ void bf_accumulate(const float (&rr)[8], const S& o, int nth, f4& ag,
                                              f4 (&ah)[NTH]) {
  bf8 r1, r2, r3;
  if constexpr (GRAD3) {
    bf_split8_3(rr, r1, r2, r3);
    ag = mfma_bf(r3, o.kg[0], ag);
    ag = mfma_bf(r2, o.kg[1], ag);
    ag = mfma_bf(r1, o.kg[2], ag);
    ag = mfma_bf(r2, o.kg[0], ag);
    ag = mfma_bf(r1, o.kg[1], ag);
    ag = mfma_bf(r1, o.kg[0], ag);
  } else {
    bf_split8_2(rr, r1, r2);
    ag = mfma_bf(r2, o.kg[0], ag);
    ag = mfma_bf(r1, o.kg[1], ag);
    ag = mfma_bf(r1, o.kg[0], ag);
  }
#pragma unroll
  for (int t = 0; t < NTH; ++t)
    if (t < nth) {
      ah[t] = mfma_bf(r2, o.k1[t], ah[t]);
      ah[t] = mfma_bf(r1, o.k2[t], ah[t]);
      ah[t] = mfma_bf(r1, o.k1[t], ah[t]);
    }
}

template <int NE, int NTH, typename CT>
struct BfStep {
  bf8 xa[2 * NE];
  f4 off0, off1;
  bf8 kg[3];
  bf8 k1[NTH], k2[NTH];
  CT y[8];
};

// Counts of one step for lane (s, h): cells 4 h + r and 16 + 4 h + r.
// yh = the gene's row + 4 h.  Table addresses are a uniform per-step base
// plus the lane's offset, so they need no 64-bit vector arithmetic.
template <typename CT, bool VEC>
__device__ __forceinline__ void load_bf_y(const CT* yh, int64_t st, CT (&y)[8]) {
  CT t0[4], t1[4];
  load4_fast<CT, VEC>(yh + st * 32, t0);
  load4_fast<CT, VEC>(yh + st * 32 + 16, t1);
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    y[r] = t0[r];
    y[4 + r] = t1[r];
  }
}

// The row's partial last step.
template <typename CT>
__device__ __forceinline__ void load_bf_y_tail(const TileDesign& td, const CT* yrow, int64_t st,
                                               int h, CT (&y)[8]) {
  CT t0[4], t1[4];
  load4_tail<CT>(yrow, st * 32 + 4 * h, td.n, t0);
  load4_tail<CT>(yrow, st * 32 + 16 + 4 * h, td.n, t1);
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    y[r] = t0[r];
    y[4 + r] = t1[r];
  }
}

template <int NE, typename CT, typename S>
__device__ __forceinline__ void load_bf_eta(const TileDesign& td, int64_t st, int lane, int h,
                                            S& o) {
  const bf8* xs = reinterpret_cast<const bf8*>(td.xab) + st * (2 * NE * 64);
#pragma unroll
  for (int j = 0; j < 2 * NE; ++j) o.xa[j] = xs[j * 64 + lane];
  const f4* os = reinterpret_cast<const f4*>(td.offl + st * 32);
  o.off0 = os[h];
  o.off1 = os[4 + h];
}

template <int NTH, typename S>
__device__ __forceinline__ void load_bf_k(const TileDesign& td, int64_t st, int lane, int nth,
                                          S& o) {
  const bf8* gs = reinterpret_cast<const bf8*>(td.kg) + st * (3 * 64);
#pragma unroll
  for (int q = 0; q < 3; ++q) o.kg[q] = gs[q * 64 + lane];
  const bf8* k1s = reinterpret_cast<const bf8*>(td.kh) + st * nth * 64;
  const bf8* k2s = reinterpret_cast<const bf8*>(td.kh2) + st * nth * 64;
#pragma unroll
  for (int t = 0; t < NTH; ++t) {
    // unused tiles load a used one (or tile 0 of a zero table): no branch
    const int tt = t < nth ? t : (nth > 0 ? nth - 1 : 0);
    o.k1[t] = k1s[tt * 64 + lane];
    o.k2[t] = k2s[tt * 64 + lane];
  }
}

// B operand of eta for gene slot s: entries 32 j + 8 h + i, entry e = 6 f + term
// holding the bf16 piece of log2(e) beta_f that term selects (see above).
template <int P, int NE>
__device__ __forceinline__ void bf_beta_operand(const double* __restrict__ beta, int g, bool live,
                                                int h, bf8 (&bb)[NE]) {
#pragma unroll
  for (int j = 0; j < NE; ++j) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int e = 32 * j + 8 * h + i;
      const int f = e / kEtaTerms, term = e % kEtaTerms;
      float v = 0.0f;
      if (live && f < P) {
        double rem = beta[(int64_t)g * P + f] * 1.44269504088896340736;
        const int want = eta_b_piece(term);
        float piece = 0.0f;
        for (int q = 0; q <= want; ++q) {
          piece = (float)(__bf16)(float)rem;
          rem -= (double)piece;
        }
        v = piece;
      }
      bb[j][i] = (__bf16)v;
    }
  }
}

// eta of the step's 8 cells and their residuals r = (y - mu) / (k + mu).
template <int NE, typename CT, typename S>
__device__ __forceinline__ void bf_residuals(const S& o, const CT (&y)[8], const bf8 (&bb)[NE],
                                             float k, float& tmin, float (&rr)[8]) {
  f4 eta0 = f4{0.f, 0.f, 0.f, 0.f}, eta1 = eta0;
#pragma unroll
  for (int j = 0; j < NE; ++j) eta0 = mfma_bf(o.xa[j], bb[j], eta0);
#pragma unroll
  for (int j = 0; j < NE; ++j) eta1 = mfma_bf(o.xa[NE + j], bb[j], eta1);
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    const float t2 = (r < 4 ? eta0[r] : eta1[r - 4]) + (r < 4 ? o.off0[r] : o.off1[r - 4]);
    tmin = fminf(tmin, t2);
    const float mu = __builtin_amdgcn_exp2f(t2);
    rr[r] = ((float)y[r] - mu) * __builtin_amdgcn_rcpf(k + mu);
  }
}

// Bread weights of the step's 8 cells, w = (y th + 1) mu / (1 + th mu)^2
// (src/variance.cpp:22-47); tmax tracks the largest log2(mu) (FP32 exp2
// overflows past 128 where FP64 does not).
template <int NE, typename CT, typename S>
__device__ __forceinline__ void bf_bread_weights(const S& o, const CT (&y)[8],
                                                 const bf8 (&bb)[NE], float th, float& tmax,
                                                 float (&ww)[8]) {
  f4 eta0 = f4{0.f, 0.f, 0.f, 0.f}, eta1 = eta0;
#pragma unroll
  for (int j = 0; j < NE; ++j) eta0 = mfma_bf(o.xa[j], bb[j], eta0);
#pragma unroll
  for (int j = 0; j < NE; ++j) eta1 = mfma_bf(o.xa[NE + j], bb[j], eta1);
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    const float t2 = (r < 4 ? eta0[r] : eta1[r - 4]) + (r < 4 ? o.off0[r] : o.off1[r - 4]);
    tmax = fmaxf(tmax, t2);
    const float mu = __builtin_amdgcn_exp2f(t2);
    const float rd = __builtin_amdgcn_rcpf(fmaf(th, mu, 1.0f));
    ww[r] = fmaf((float)y[r], th, 1.0f) * (mu * rd) * rd;
  }
}

// The bread weights and the score residuals r = (y - mu) / (1 + th mu)
// (src/variance.cpp:186-228) of the step's 8 cells from one mu.
template <int NE, typename CT, typename S>
__device__ __forceinline__ void bf_bread_score(const S& o, const CT (&y)[8], const bf8 (&bb)[NE],
                                               float th, float& tmax, float (&ww)[8],
                                               float (&rs)[8]) {
  f4 eta0 = f4{0.f, 0.f, 0.f, 0.f}, eta1 = eta0;
#pragma unroll
  for (int j = 0; j < NE; ++j) eta0 = mfma_bf(o.xa[j], bb[j], eta0);
#pragma unroll
  for (int j = 0; j < NE; ++j) eta1 = mfma_bf(o.xa[NE + j], bb[j], eta1);
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    const float t2 = (r < 4 ? eta0[r] : eta1[r - 4]) + (r < 4 ? o.off0[r] : o.off1[r - 4]);
    tmax = fmaxf(tmax, t2);
    const float mu = __builtin_amdgcn_exp2f(t2);
    const float rd = __builtin_amdgcn_rcpf(fmaf(th, mu, 1.0f));
    const float yf = (float)y[r];
    ww[r] = fmaf(yf, th, 1.0f) * (mu * rd) * rd;
    rs[r] = (yf - mu) * rd;
  }
}

template <int P, typename CT, bool VEC>
__device__ __forceinline__ void irls_tile_chunk_bf(const TileDesign& td, const CT* __restrict__ Y,
                                                   int64_t ld, const int32_t* __restrict__ active,
                                                   int nA, const double* __restrict__ beta,
                                                   const double* __restrict__ kvec, int tile,
                                                   int chunk, double* __restrict__ part) {
  constexpr int NE = eta_blocks(P);
  constexpr int NTH = ((P * (P + 1)) / 2 + 15) / 16;  // pair tiles at most
  constexpr int kSpan = 4;                            // steps per FP32 span (gradient tile)
  const int lane = threadIdx.x & 63;
  const int s = lane & 15, h = lane >> 4;
  const int slot = tile * 16 + s;
  const bool live = slot < nA;
  const int g = live ? active[slot] : active[tile * 16];
  const int nth = td.nt - 1;
  bf8 bb[NE];
  bf_beta_operand<P, NE>(beta, g, live, h, bb);
  const float k = live ? (float)kvec[g] : 0.0f;
  const CT* yrow = Y + (int64_t)g * ld;
  // gradient tile: FP32 span sums flushed to FP64; pair tiles: FP32 over the
  // chunk (relative error ~1e-6 of a Hessian entry).  Each accumulator's
  // MFMAs are issued back to back (dependent accumulation runs without
  // stalls; interleaving accumulators measured slower).
  f4 ag0 = f4{0.f, 0.f, 0.f, 0.f};
  double ag64[4] = {0.0, 0.0, 0.0, 0.0};
  f4 ah[NTH];
#pragma unroll
  for (int t = 0; t < NTH; ++t) ah[t] = f4{0.f, 0.f, 0.f, 0.f};
  float tmin = 0.0f;  // smallest log2(mu) seen by this lane
  typedef BfStep<NE, NTH, CT> Step;
  const CT* yh = yrow + 4 * h;
  auto load = [&](int64_t st, Step& o) {
    load_bf_eta<NE, CT>(td, st, lane, h, o);
    load_bf_k<NTH>(td, st, lane, nth, o);
    load_bf_y<CT, VEC>(yh, st, o.y);
  };
  auto load_tail = [&](int64_t st, Step& o) {
    load_bf_eta<NE, CT>(td, st, lane, h, o);
    load_bf_k<NTH>(td, st, lane, nth, o);
    load_bf_y_tail<CT>(td, yrow, st, h, o.y);
  };
  int span = 0;
  auto step = [&](const Step& o, int64_t) {
    float rr[8];
    bf_residuals<NE, CT>(o, o.y, bb, k, tmin, rr);
    bf_accumulate<NTH, true>(rr, o, nth, ag0, ah);
    if (++span == kSpan) {
      span = 0;
#pragma unroll
      for (int r = 0; r < 4; ++r) ag64[r] += (double)ag0[r];
      ag0 = f4{0.f, 0.f, 0.f, 0.f};
    }
  };
  const int64_t s0 = (int64_t)chunk * td.chunk_blocks;
  const int64_t s1 = min((int64_t)(chunk + 1) * td.chunk_blocks, td.nblk);
  pipeline<Step>(s0, min(s1, td.n / 32), s1, load, load_tail, step);  // trailing chunks may be empty
#pragma unroll
  for (int r = 0; r < 4; ++r) ag64[r] += (double)ag0[r];
  // exp(-eta - off) overflows in FP64 where log2(mu) < -1024: the reference
  // then takes a NaN step, so such a gene's partials are made NaN and the
  // gene is redone in FP64.  Register r of this lane holds gene slot 4 h + r.
  tmin = fminf(tmin, __shfl_xor(tmin, 16, 64));
  tmin = fminf(tmin, __shfl_xor(tmin, 32, 64));
  bool poison[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) poison[r] = __shfl(tmin, 4 * h + r, 64) < -1024.0f;
  double* out = part + ((int64_t)tile * td.nchunk + chunk) * 16 * td.ncp;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const double nanv = __builtin_nan("");
    out[(4 * h + r) * td.ncp + s] = poison[r] ? nanv : ag64[r];
#pragma unroll
    for (int t = 0; t < NTH; ++t)
      if (t < nth) out[(4 * h + r) * td.ncp + 16 * (t + 1) + s] = poison[r] ? nanv : (double)ah[t][r];
  }
}

template <int P>
__global__ __launch_bounds__(256) void k_irls_accum_bf16(TileDesign td, Counts c,
                                                         const int32_t* __restrict__ active,
                                                         const int32_t* __restrict__ nactive,
                                                         const double* __restrict__ beta,
                                                         const double* __restrict__ kvec,
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
    irls_tile_chunk_bf<P, CT, decltype(vec)::value>(td, Y, c.ld, active, nA, beta, kvec, tile,
                                                    chunk, part);
  });
}

// ---------------------------------------------------------------------------
// Wave-specialised bf16 variant: the bf16 kernel's arithmetic with every
// operand staged in LDS by global_load_lds_dwordx4 (no registers held by
// loads in flight).  Loads of one wave complete in issue order, so a wave
// that loads both the tables (mostly L2 hits) and the counts (HBM) cannot
// keep counts in flight for longer than its table prefetch distance; here
// the four compute waves (16-gene tiles) load only their counts, three
// steps ahead into a private 4-slot ring, and read the tables of a step from
// a shared 3-slot ring that wave 4 fills two steps ahead.  One workgroup
// barrier per step: after barrier k the compute waves read table slot k % 3
// while the loader overwrites the slot read in step k - 1.
// ---------------------------------------------------------------------------
typedef __attribute__((address_space(3))) void* lds_ptr_t;

// Compute waves (16-gene tiles) per workgroup of the wave-specialised kernel
// (one more wave loads the tables): as many as the LDS holds, at most 15
// (a workgroup of 1024 threads).  Table ring: 3 slots of (2 NE + 3 + 2 NTH)
// KB + 128 B; each compute wave: 4 count slots of 2 KB.
__host__ __device__ constexpr int ws_waves(int p, int tpw = 1) {
  const int ne = eta_blocks(p), nth = (p * (p + 1) / 2 + 15) / 16;
  const int tables = 3 * ((2 * ne + 3 + 2 * nth) * 1024 + 128);
  const int fit = (160 * 1024 - tables) / (tpw * 4 * 2048);
  return fit < 15 / tpw ? fit : 15 / tpw;
}

template <int NE, int NTH>
struct WsLayout {
  static constexpr int kX = 0;                          // 2 NE x 1 KB: eta A operands
  static constexpr int kG = 2 * NE * 1024;              // 3 x 1 KB: gradient tile (3 pieces)
  static constexpr int kH1 = kG + 3 * 1024;             // NTH x 1 KB: pair tiles (hi)
  static constexpr int kH2 = kH1 + NTH * 1024;          // NTH x 1 KB: pair tiles (lo)
  static constexpr int kOff = kH2 + NTH * 1024;         // 128 B: 32 offsets (x log2 e)
  static constexpr int kSlot = kOff + 128;
  static constexpr int kSlots = 3;
  static constexpr int kLoads = 2 * NE + 3 + 2 * NTH + 1;  // loads per step
  // per compute wave: 4 count slots of 2 x 1 KB (cells 4 h + r, 16 + 4 h + r)
  static constexpr int kY = kSlots * kSlot;
  static constexpr int kYSlot = 2048;
  static constexpr int kYWave = 4 * kYSlot;
};

// s_waitcnt vmcnt(N) leaving expcnt and lgkmcnt unconstrained (gfx9 encoding).
template <int N>
__device__ __forceinline__ void wait_vmcnt() {
  static_assert(N >= 0 && N < 64, "vmcnt range");
  __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4) | (15 << 8));
}

// Workgroup barrier for waves with LDS-DMA loads in flight.  __syncthreads'
// release fence would wait for all of them (they write LDS); instead: this
// wave's LDS reads are complete (lgkmcnt(0)), the barrier, and an acquire
// fence that keeps later LDS reads after it.  Completion of the DMA loads a
// step needs is made explicit with wait_vmcnt by the wave that issued them.
__device__ __forceinline__ void ws_barrier() {
  __builtin_amdgcn_s_waitcnt(0xC07F);  // vmcnt 63, expcnt 7, lgkmcnt 0
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

__device__ __forceinline__ void lds_dma16(const void* g, unsigned char* l) {
  __builtin_amdgcn_global_load_lds(g, (lds_ptr_t)l, 16, 0, 0);
}

// 4 bytes per lane, written to l + 4 lane
__device__ __forceinline__ void lds_dma4(const void* g, unsigned char* l) {
  __builtin_amdgcn_global_load_lds(g, (lds_ptr_t)l, 4, 0, 0);
}

template <int NE, int NTH>
__device__ __forceinline__ void ws_loader(const TileDesign& td, int64_t s0, int64_t s1,
                                          unsigned char* ring) {
  typedef WsLayout<NE, NTH> Lay;
  const int lane = threadIdx.x & 63;
  const int nth = td.nt - 1;
  const int64_t nsteps = s1 - s0;
  auto issue = [&](int64_t st, int slot) {
    unsigned char* b = ring + slot * Lay::kSlot;
    const bf8* xs = reinterpret_cast<const bf8*>(td.xab) + st * (2 * NE * 64);
#pragma unroll
    for (int j = 0; j < 2 * NE; ++j) lds_dma16(xs + j * 64 + lane, b + Lay::kX + j * 1024);
    const bf8* gs = reinterpret_cast<const bf8*>(td.kg) + st * (3 * 64);
#pragma unroll
    for (int q = 0; q < 3; ++q) lds_dma16(gs + q * 64 + lane, b + Lay::kG + q * 1024);
    const bf8* k1s = reinterpret_cast<const bf8*>(td.kh) + st * nth * 64;
    const bf8* k2s = reinterpret_cast<const bf8*>(td.kh2) + st * nth * 64;
#pragma unroll
    for (int t = 0; t < NTH; ++t) {
      // a fixed number of loads per step (the waits below count them):
      // unused tiles load a used one again, to the same place
      const int tt = t < nth ? t : (nth > 0 ? nth - 1 : 0);
      lds_dma16(k1s + tt * 64 + lane, b + Lay::kH1 + tt * 1024);
      lds_dma16(k2s + tt * 64 + lane, b + Lay::kH2 + tt * 1024);
    }
    if (lane < 8) lds_dma16(reinterpret_cast<const f4*>(td.offl + st * 32) + lane, b + Lay::kOff);
  };
  // steps s0 and s0 + 1 in flight; step s0 complete before barrier 0
  if (nsteps > 0) issue(s0, 0);
  if (nsteps > 1) {
    issue(s0 + 1, 1);
    wait_vmcnt<Lay::kLoads>();
  } else {
    wait_vmcnt<0>();
  }
  ws_barrier();
  for (int64_t k = 0; k < nsteps; ++k) {
    // step s0 + k + 2 into the slot read in step k - 1; step s0 + k + 1
    // complete before the barrier that ends step k
    if (k + 2 < nsteps) {
      issue(s0 + k + 2, (int)((k + 2) % Lay::kSlots));
      wait_vmcnt<Lay::kLoads>();
    } else {
      wait_vmcnt<0>();
    }
    ws_barrier();
  }
}

// MODE 0: IRLS accumulate (gpar = k, residuals r); MODE 1: bread (gpar =
// theta, weights w): the same contractions of a per-cell scalar with the
// design columns and the pair columns; MODE 2: the bread and, against the
// design columns, the score residuals of the cells [c0, c1) of one patient
// (cells of the chunk's first and last step outside the range are masked)
// into spart.  nchunk: chunks of the partial layout.  A16: every count row starts 16-byte
// aligned, so a lane's four cells arrive by one 16-byte direct-to-LDS load
// (lane l's gene l % 16, cells 4 (l / 16) .. + 3, at byte 16 l of each KB);
// otherwise four 4-byte direct-to-LDS loads per KB, load e covering genes
// 4 e .. 4 e + 3 with 16 lanes (64 contiguous bytes) each, so gene s's 16
// cells sit at dwords 16 s .. 16 s + 15 and lane (s, h) reads its four at
// byte 64 s + 16 h.
template <int P, int MODE, int TPW, bool A16, typename CT>
__device__ __forceinline__ void ws_compute(const TileDesign& td, const CT* __restrict__ Y,
                                           int64_t ld, const int32_t* __restrict__ active,
                                           int nA, const double* __restrict__ beta,
                                           const double* __restrict__ kvec, int tile0,
                                           int ntile, int chunk, int nchunk, int64_t s0,
                                           int64_t s1, int64_t c0, int64_t c1,
                                           unsigned char* ring, double* __restrict__ part,
                                           double* __restrict__ spart) {
  constexpr int NE = eta_blocks(P);
  constexpr int NTH = ((P * (P + 1)) / 2 + 15) / 16;
  constexpr int kSpan = 4;
  typedef WsLayout<NE, NTH> Lay;
  const int lane = threadIdx.x & 63;
  const int s = lane & 15, h = lane >> 4;
  const int nth = td.nt - 1;
  // the wave's tiles tile0 .. tile0 + TPW - 1; a wave whose first tile is
  // live runs all of them (tiles past the last one repeat the first one's
  // rows, so every tile issues the same loads, and write nothing)
  const bool wave_live = tile0 < ntile;
  bool tile_live[TPW];
  int g[TPW];
  bf8 bb[TPW][NE];
  float k[TPW];
  const CT* yrow[TPW];
  const CT* yh[TPW];
#pragma unroll
  for (int u = 0; u < TPW; ++u) {
    const int tile = tile0 + u;
    tile_live[u] = tile < ntile;
    const int slot = (tile_live[u] ? tile : tile0) * 16 + s;
    const bool live = wave_live && slot < nA;
    g[u] = wave_live ? (live ? active[slot] : active[(tile_live[u] ? tile : tile0) * 16]) : 0;
    bf_beta_operand<P, NE>(beta, g[u], live, h, bb[u]);
    k[u] = live ? (float)kvec[g[u]] : 0.0f;
    yrow[u] = Y + (int64_t)g[u] * ld;
    yh[u] = yrow[u] + 4 * h;
  }
  // 4-byte loads: in load e, lane l fetches cell l % 16 of the gene of lane
  // 4 e + l / 16 (lanes 0..15 hold the 16 genes in lane order)
  const CT* ydw[TPW][4];
#pragma unroll
  for (int u = 0; u < TPW; ++u) {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      if constexpr (A16) {
        ydw[u][e] = yh[u];
      } else {
        const int gs = __shfl(g[u], 4 * e + (lane >> 4), 64);
        ydw[u][e] = Y + (int64_t)gs * ld + (lane & 15);
      }
    }
  }
  constexpr int kLoadsPerStep = A16 ? 2 : 8;  // count loads per tile and step
  const int64_t nsteps = s1 - s0;
  f4 ag0[TPW];
  double ag64[TPW][4];
  f4 ah[TPW][NTH];
  float tmin[TPW];
  f4 as0[TPW];         // MODE 2: score sums of the FP32 span
  double as64[TPW][4];
#pragma unroll
  for (int u = 0; u < TPW; ++u) {
    ag0[u] = f4{0.f, 0.f, 0.f, 0.f};
    as0[u] = f4{0.f, 0.f, 0.f, 0.f};
    tmin[u] = 0.0f;
#pragma unroll
    for (int r = 0; r < 4; ++r) ag64[u][r] = as64[u][r] = 0.0;
#pragma unroll
    for (int t = 0; t < NTH; ++t) ah[u][t] = f4{0.f, 0.f, 0.f, 0.f};
  }
  int span = 0;
  typedef BfCounts<CT> Cnt;
  unsigned char* yring = ring + Lay::kY + wave_index() * (TPW * Lay::kYWave);
  // counts of step st into count slot q of every tile (lane l's 16 bytes at
  // l * 16)
  auto issuey = [&](int64_t st, int q) {
#pragma unroll
    for (int u = 0; u < TPW; ++u) {
      unsigned char* yr = yring + u * Lay::kYWave + q * Lay::kYSlot;
      if constexpr (A16) {
        lds_dma16(yh[u] + st * 32, yr);
        lds_dma16(yh[u] + st * 32 + 16, yr + 1024);
      } else {
#pragma unroll
        for (int half = 0; half < 2; ++half)
#pragma unroll
          for (int e = 0; e < 4; ++e)
            lds_dma4(ydw[u][e] + st * 32 + 16 * half, yr + 1024 * half + 256 * e);
      }
    }
  };
  auto ready = [&](int q, Cnt (&o)[TPW]) {
    typedef typename Vec4Of<CT>::T V4;
#pragma unroll
    for (int u = 0; u < TPW; ++u) {
      const unsigned char* yr = yring + u * Lay::kYWave + q * Lay::kYSlot;
      const int at = A16 ? lane : 4 * s + h;
      const V4 a = reinterpret_cast<const V4*>(yr)[at];
      const V4 b = reinterpret_cast<const V4*>(yr + 1024)[at];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        o[u].y[r] = a[r];
        o[u].y[4 + r] = b[r];
      }
    }
  };
  // step s0 + k: tables from ring slot k % 3 (read once for the wave's
  // tiles), counts from the count rings
  auto step = [&](int64_t kk, const Cnt (&yc)[TPW]) {
    const unsigned char* b = ring + (int)(kk % Lay::kSlots) * Lay::kSlot;
    BfTables<NE, NTH> o;
    const bf8* lx = reinterpret_cast<const bf8*>(b + Lay::kX);
#pragma unroll
    for (int j = 0; j < 2 * NE; ++j) o.xa[j] = lx[j * 64 + lane];
    const f4* lo = reinterpret_cast<const f4*>(b + Lay::kOff);
    o.off0 = lo[h];
    o.off1 = lo[4 + h];
    const bf8* lg = reinterpret_cast<const bf8*>(b + Lay::kG);
#pragma unroll
    for (int q = 0; q < 3; ++q) o.kg[q] = lg[q * 64 + lane];
    const bf8* l1 = reinterpret_cast<const bf8*>(b + Lay::kH1);
    const bf8* l2 = reinterpret_cast<const bf8*>(b + Lay::kH2);
#pragma unroll
    for (int t = 0; t < NTH; ++t)
      if (t < nth) {
        o.k1[t] = l1[t * 64 + lane];
        o.k2[t] = l2[t * 64 + lane];
      }
#pragma unroll
    for (int u = 0; u < TPW; ++u) {
      float rr[8];
      if constexpr (MODE == 0) {
        bf_residuals<NE, CT>(o, yc[u].y, bb[u], k[u], tmin[u], rr);
      } else if constexpr (MODE == 1) {
        bf_bread_weights<NE, CT>(o, yc[u].y, bb[u], k[u], tmin[u], rr);
      } else {
        float rs[8];
        bf_bread_score<NE, CT>(o, yc[u].y, bb[u], k[u], tmin[u], rr, rs);
        const int64_t cell0 = (s0 + kk) * 32;
        if (cell0 < c0 || cell0 + 32 > c1) {
#pragma unroll
          for (int r = 0; r < 8; ++r) {
            const int64_t cell = cell0 + (r < 4 ? 4 * h + r : 16 + 4 * h + r - 4);
            if (cell < c0 || cell >= c1) rr[r] = rs[r] = 0.0f;
          }
        }
        bf_grad3(rs, o, as0[u]);
      }
      bf_accumulate<NTH, MODE == 0>(rr, o, nth, ag0[u], ah[u]);
    }
    if (++span == kSpan) {
      span = 0;
#pragma unroll
      for (int u = 0; u < TPW; ++u) {
#pragma unroll
        for (int r = 0; r < 4; ++r) ag64[u][r] += (double)ag0[u][r];
        ag0[u] = f4{0.f, 0.f, 0.f, 0.f};
        if constexpr (MODE == 2) {
#pragma unroll
          for (int r = 0; r < 4; ++r) as64[u][r] += (double)as0[u][r];
          as0[u] = f4{0.f, 0.f, 0.f, 0.f};
        }
      }
    }
  };
  // Steps [s0, s1f) lie inside the rows.  Their counts are in flight three
  // steps ahead: s0 .. s0 + 2 before barrier 0, then in step k the load of
  // step s0 + k + 3 (clamped to s1f - 1) into count slot (k + 3) % 4, whose
  // step k - 1 was read.  This wave's vector-memory instructions are these
  // loads alone (kLoadsPerStep per tile and step), so vmcnt(3 kLoadsPerStep
  // TPW) after issuing step k + 3 leaves steps k + 1 .. k + 3 in flight and
  // step k in LDS.  A row's
  // partial last step comes after them and loads its counts when it runs.
  const int64_t s1f = max(s0, min(s1, td.n / 32));
  const int64_t nf = s1f - s0;
  if (wave_live && nf > 0) {
#pragma unroll
    for (int d = 0; d < 3; ++d) issuey(min(s0 + d, s1f - 1), d);
  }
  ws_barrier();
  int64_t kk = 0;
  for (; kk + 4 <= nf; kk += 4) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (wave_live) {
        issuey(min(s0 + kk + u + 3, s1f - 1), (u + 3) & 3);
        wait_vmcnt<3 * kLoadsPerStep * TPW>();
        Cnt yc[TPW];
        ready(u, yc);
        step(kk + u, yc);
      }
      ws_barrier();
    }
  }
  // at most three full steps left, in count slots 0 .. 2; every load of
  // this wave completes here (none may land in LDS after the wave ends)
  wait_vmcnt<0>();
#pragma unroll
  for (int u = 0; u < 3; ++u) {
    if (kk + u < nf) {
      if (wave_live) {
        Cnt yc[TPW];
        ready(u, yc);
        step(kk + u, yc);
      }
      ws_barrier();
    }
  }
  for (kk = nf; kk < nsteps; ++kk) {
    if (wave_live) {
      Cnt t[TPW];
#pragma unroll
      for (int u = 0; u < TPW; ++u) load_bf_y_tail<CT>(td, yrow[u], s0 + kk, h, t[u].y);
      step(kk, t);
    }
    ws_barrier();
  }
  if (!wave_live) return;
#pragma unroll
  for (int u = 0; u < TPW; ++u) {
    if (!tile_live[u]) continue;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      ag64[u][r] += (double)ag0[u][r];
      as64[u][r] += (double)as0[u][r];
    }
    // IRLS: tmin is the smallest log2(mu), poison below -1024 (FP64's
    // exp(-eta - off) overflows there); bread: tmin holds the largest,
    // poison from 126 on (FP32 mu overflows)
    float tm = tmin[u];
    bool poison[4];
    if constexpr (MODE == 0) {
      tm = fminf(tm, __shfl_xor(tm, 16, 64));
      tm = fminf(tm, __shfl_xor(tm, 32, 64));
#pragma unroll
      for (int r = 0; r < 4; ++r) poison[r] = __shfl(tm, 4 * h + r, 64) < -1024.0f;
    } else {
      tm = fmaxf(tm, __shfl_xor(tm, 16, 64));
      tm = fmaxf(tm, __shfl_xor(tm, 32, 64));
#pragma unroll
      for (int r = 0; r < 4; ++r) poison[r] = !(__shfl(tm, 4 * h + r, 64) < 126.0f);
    }
    double* out = part + ((int64_t)(tile0 + u) * nchunk + chunk) * 16 * td.ncp;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const double nanv = __builtin_nan("");
      out[(4 * h + r) * td.ncp + s] = poison[r] ? nanv : ag64[u][r];
#pragma unroll
      for (int t = 0; t < NTH; ++t)
        if (t < nth)
          out[(4 * h + r) * td.ncp + 16 * (t + 1) + s] = poison[r] ? nanv : (double)ah[u][t][r];
      if constexpr (MODE == 2)
        spart[(((int64_t)(tile0 + u) * nchunk + chunk) * 16 + 4 * h + r) * 16 + s] =
            poison[r] ? nanv : as64[u][r];
    }
  }
}

template <int P, int MODE, int TPW>
__device__ __forceinline__ void ws_body(const TileDesign& td, const Counts& c,
                                        const int32_t* __restrict__ active,
                                        const int32_t* __restrict__ nactive,
                                        const double* __restrict__ beta,
                                        const double* __restrict__ gpar,
                                        double* __restrict__ part, int wb, int nwb,
                                        const ChunkTable& ct = ChunkTable{},
                                        double* __restrict__ spart = nullptr) {
  constexpr int NE = eta_blocks(P);
  constexpr int NTH = ((P * (P + 1)) / 2 + 15) / 16;
  typedef WsLayout<NE, NTH> Lay;
  constexpr int kWsWaves = ws_waves(P, TPW);
  __shared__ __attribute__((aligned(16)))
  unsigned char ring[Lay::kY + kWsWaves * TPW * Lay::kYWave];
  const int nA = *nactive;
  const int ntile = (nA + 15) / 16;
  const int wave = wave_index();
  int chunk, group;
  const int nchunk = MODE == 2 ? ct.nchunk : td.nchunk;
  tile_coords_at(wb, nwb, nchunk, chunk, group);
  if (group * kWsWaves * TPW >= ntile) return;  // the whole workgroup
  int64_t s0, s1, c0 = 0, c1 = 0;
  if constexpr (MODE == 2) {
    // the steps holding the chunk's cells [c0, c1)
    c0 = ct.c0[chunk];
    c1 = ct.c1[chunk];
    s0 = c0 / 32;
    s1 = c0 < c1 ? min((c1 + 31) / 32, td.nblk) : s0;
  } else {
    s0 = (int64_t)chunk * td.chunk_blocks;
    s1 = max(s0, min((int64_t)(chunk + 1) * td.chunk_blocks, td.nblk));
  }
  if (wave == kWsWaves) {
    ws_loader<NE, NTH>(td, s0, s1, ring);
    return;
  }
  const int tile0 = (group * kWsWaves + wave) * TPW;
  const bool a16 = (c.ld & 3) == 0 && (reinterpret_cast<uintptr_t>(c.ptr) & 15) == 0;
  if (a16) {
    if (c.dtype == 0)
      ws_compute<P, MODE, TPW, true>(td, static_cast<const int32_t*>(c.ptr), c.ld, active, nA,
                                     beta, gpar, tile0, ntile, chunk, nchunk, s0, s1, c0, c1,
                                     ring, part, spart);
    else
      ws_compute<P, MODE, TPW, true>(td, static_cast<const float*>(c.ptr), c.ld, active, nA,
                                     beta, gpar, tile0, ntile, chunk, nchunk, s0, s1, c0, c1,
                                     ring, part, spart);
  } else if constexpr (TPW == 1) {
    // (two tiles per wave would keep 64 loads in flight, past what vmcnt
    // counts: the launchers run misaligned rows with one tile per wave)
    if (c.dtype == 0)
      ws_compute<P, MODE, TPW, false>(td, static_cast<const int32_t*>(c.ptr), c.ld, active, nA,
                                      beta, gpar, tile0, ntile, chunk, nchunk, s0, s1, c0, c1,
                                      ring, part, spart);
    else
      ws_compute<P, MODE, TPW, false>(td, static_cast<const float*>(c.ptr), c.ld, active, nA,
                                      beta, gpar, tile0, ntile, chunk, nchunk, s0, s1, c0, c1,
                                      ring, part, spart);
  }
}

// Counts must be 4-byte elements in 16-byte aligned rows (the launchers fall
// back to the register kernels otherwise).
template <int P>
__global__ __launch_bounds__(64 * (ws_waves(P) + 1)) void k_irls_accum_ws(
    TileDesign td, Counts c, const int32_t* __restrict__ active,
    const int32_t* __restrict__ nactive, const double* __restrict__ beta,
    const double* __restrict__ kvec, double* __restrict__ part) {
  ws_body<P, 0, 1>(td, c, active, nactive, beta, kvec, part, blockIdx.x, gridDim.x);
}

// The same with two 16-gene tiles per compute wave: half the waves, each
// reading a step's tables once for 32 genes and holding two independent
// chains of MFMA and VALU work.
template <int P>
__global__ __launch_bounds__(64 * (ws_waves(P, 2) + 1)) void k_irls_accum_ws2(
    TileDesign td, Counts c, const int32_t* __restrict__ active,
    const int32_t* __restrict__ nactive, const double* __restrict__ beta,
    const double* __restrict__ kvec, double* __restrict__ part) {
  ws_body<P, 0, 2>(td, c, active, nactive, beta, kvec, part, blockIdx.x, gridDim.x);
}

// Mixed-precision bread partials: per (16-gene tile, chunk) the sums of
// w x_a (design columns) and w K (pair columns) with FP32 element math,
// bf16x3 pair products and FP64 partial sums (genes in active order).
template <int P>
__global__ __launch_bounds__(64 * (ws_waves(P) + 1)) void k_bread_accum_ws(
    TileDesign td, Counts c, const int32_t* __restrict__ genes,
    const int32_t* __restrict__ ngenes, const double* __restrict__ beta,
    const double* __restrict__ theta, double* __restrict__ part) {
  ws_body<P, 1, 1>(td, c, genes, ngenes, beta, theta, part, blockIdx.x, gridDim.x);
}

// The same with the per-patient score sums of the clustered meat: chunks are
// the patient-aligned cell ranges of ct; spart[tile][chunk][16][16] holds the
// sums of r x_a (design column a at column a) with FP32 element math, bf16x3
// products and FP64 sums over FP32 spans of four steps.
template <int P>
__global__ __launch_bounds__(64 * (ws_waves(P) + 1)) void k_bread_meat_ws(
    TileDesign td, ChunkTable ct, Counts c, const int32_t* __restrict__ genes,
    const int32_t* __restrict__ ngenes, const double* __restrict__ beta,
    const double* __restrict__ theta, double* __restrict__ part, double* __restrict__ spart) {
  ws_body<P, 2, 1>(td, c, genes, ngenes, beta, theta, part, blockIdx.x, gridDim.x, ct, spart);
}

// bread = H^-1 with H[a][b] = the sum over chunks of the pair column of
// (a, b) (src/variance.cpp:22-47); one 64-thread block per gene.  A NaN
// partial (FP32 range exceeded) counts in *fallback: the caller then runs
// the FP64 bread for the problem (singular[] is merged into the status only
// when it does not).
template <int P>
__global__ __launch_bounds__(64) void k_bread_finish_tile(TileDesign td, int64_t G,
                                                          const double* __restrict__ part,
                                                          double* __restrict__ bread,
                                                          int32_t* __restrict__ singular,
                                                          int32_t* __restrict__ fallback) {
  __shared__ double acc[DEVIL_NCP_MAX];
  const int64_t g = blockIdx.x;
  if (g >= G) return;
  const int lane = threadIdx.x;
  const int64_t tile = g >> 4, s = g & 15;
  const double* src = part + tile * td.nchunk * 16 * td.ncp + s * td.ncp;
  for (int col = lane; col < td.ncp; col += 64) {
    double a = 0.0;
    for (int ch = 0; ch < td.nchunk; ++ch) a += src[(int64_t)ch * 16 * td.ncp + col];
    acc[col] = a;
  }
  __syncthreads();
  if (lane != 0) return;
  double Hm[P][P], Bi[P][P];
  bool bad = false;
#pragma unroll
  for (int a = 0; a < P; ++a)
#pragma unroll
    for (int bb = a; bb < P; ++bb) {
      const int pi = td.kidx_pair[tri<P>(a, bb)];
      const double v = pi >= 0 ? acc[pi] : 0.0;
      bad |= isnan(v);
      Hm[a][bb] = v;
      Hm[bb][a] = v;
    }
  if (bad) {
    atomicAdd(fallback, 1);
    return;
  }
  const int sing = lu_inverse<P>(Hm, Bi);
  double* out = bread + g * P * P;
#pragma unroll
  for (int a = 0; a < P; ++a)
#pragma unroll
    for (int bb = 0; bb < P; ++bb) out[a + bb * P] = Bi[a][bb];
  singular[g] = sing ? 1 : 0;
}

// status[g] |= bit where flag[g] is set
__global__ void k_merge_flags(int32_t* __restrict__ status, const int32_t* __restrict__ flag,
                              int64_t G, int32_t bit) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g < G && flag[g]) status[g] |= bit;
}

// FP64 accumulate of workgroup wb of nwb (4 tiles, one per wave); e2: the
// exponential table, filled by the caller
template <int P>
__device__ __forceinline__ void irls_accum_body(const TileDesign& td, const Counts& c,
                                                const int32_t* __restrict__ active,
                                                const int32_t* __restrict__ nactive,
                                                const double* __restrict__ beta,
                                                const double* __restrict__ kvec,
                                                double* __restrict__ part, const double* e2,
                                                int wb, int nwb) {
  const int nA = *nactive;
  const int ntile = (nA + 15) / 16;
  const int wave = wave_index();
  int chunk, group;
  tile_coords_at(wb, nwb, td.nchunk, chunk, group);
  const int tile = group * 4 + wave;
  if (tile >= ntile) return;
  with_counts(c, [&](auto Y, auto vec) {
    typedef std::remove_cv_t<std::remove_pointer_t<decltype(Y)>> CT;
    irls_tile_chunk<P, CT, decltype(vec)::value>(td, Y, c.ld, active, nA, beta, kvec, tile,
                                                 chunk, part, e2);
  });
}

template <int P>
__global__ __launch_bounds__(256) void k_irls_accum(TileDesign td, Counts c,
                                                    const int32_t* __restrict__ active,
                                                    const int32_t* __restrict__ nactive,
                                                    const double* __restrict__ beta,
                                                    const double* __restrict__ kvec,
                                                    double* __restrict__ part) {
  __shared__ double e2[64];
  exp_tab_fill(e2, threadIdx.x, blockDim.x);
  __syncthreads();
  irls_accum_body<P>(td, c, active, nactive, beta, kvec, part, e2, blockIdx.x, gridDim.x);
}

// One workgroup per active gene.  Its threads sum the chunk partials in
// kUpdSlices interleaved slices per column (two accumulators each, so a
// round of up to 512 chunks keeps many loads in flight), the slices are
// added in LDS, and one lane runs the p x p solve.
constexpr int kUpdThreads = 256;
constexpr int kUpdSlices = 16;

// The update of active-list entry i (one workgroup of kUpdThreads).
template <int P>
__device__ __forceinline__ void irls_update_body(const TileDesign& td, int i,
                                                 const int32_t* __restrict__ active,
                                                 const int32_t* __restrict__ nactive,
                                                 const double* __restrict__ part,
                                                 double* __restrict__ beta,
                                                 const double* __restrict__ kvec, int max_iter,
                                                 double eps, int32_t* __restrict__ iters,
                                                 int32_t* __restrict__ status,
                                                 int32_t* __restrict__ next,
                                                 int32_t* __restrict__ nnext, double band,
                                                 int32_t* __restrict__ recheck,
                                                 int32_t* __restrict__ nrecheck,
                                                 double* __restrict__ bprev, int polish) {
  __shared__ double acc[DEVIL_NCP_MAX];
  __shared__ double slice[kUpdSlices][DEVIL_NCP_MAX];
  const int nA = *nactive;
  if (i >= nA) return;
  const int lane = threadIdx.x;
  const int g = active[i];
  const int tile = i >> 4, s = i & 15;
  const int ncp = td.ncp;
  const int64_t stride = (int64_t)16 * ncp;
  const double* src = part + (int64_t)tile * td.nchunk * stride + (int64_t)s * ncp;
  for (int idx = lane; idx < kUpdSlices * ncp; idx += kUpdThreads) {
    const int col = idx % ncp, q = idx / ncp;
    const double* sc = src + col;
    double a0 = 0.0, a1 = 0.0;
    int ch = q;
    for (; ch + kUpdSlices < td.nchunk; ch += 2 * kUpdSlices) {
      a0 += sc[(int64_t)ch * stride];
      a1 += sc[(int64_t)(ch + kUpdSlices) * stride];
    }
    if (ch < td.nchunk) a0 += sc[(int64_t)ch * stride];
    slice[q][col] = a0 + a1;
  }
  __syncthreads();
  for (int col = lane; col < ncp; col += kUpdThreads) {
    double a = 0.0;
#pragma unroll
    for (int q = 0; q < kUpdSlices; ++q) a += slice[q][col];
    acc[col] = a;
  }
  __syncthreads();
  if (lane >= 64) return;
  const double k = kvec[g];
  double M[P][P], v[P];
#pragma unroll
  for (int a = 0; a < P; ++a) {
    const int gi = td.kidx_lin[a];
    v[a] = k * (gi >= 0 ? acc[gi] : 0.0);
#pragma unroll
    for (int bb = a; bb < P; ++bb) {
      const int pi = td.kidx_pair[tri<P>(a, bb)];
      const double e = k * (pi >= 0 ? acc[pi] + td.ksum[pi] : 0.0);
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
  // polish: the partials were taken at the gene's beta before its last
  // (mixed) step; that step is redone in FP64, decisions unchanged
  if (polish) {
    if (!has_nan) {
#pragma unroll
      for (int j = 0; j < P; ++j) beta[(int64_t)g * P + j] = bprev[(int64_t)g * P + j] + v[j];
    }
    return;
  }
  // partials from the mixed-precision pass: a step whose stopping decision
  // could differ from FP64 arithmetic (max |delta| within band * eps of eps,
  // or NaN) is left undone; the gene leaves the mixed rounds and is refit in
  // FP64 from its starting beta afterwards
  if (band > 0.0 && (has_nan || fabs(mx - eps) <= band * eps)) {
    recheck[atomicAdd(nrecheck, 1)] = g;
    return;
  }
  if (bprev) {
#pragma unroll
    for (int j = 0; j < P; ++j) bprev[(int64_t)g * P + j] = beta[(int64_t)g * P + j];
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

template <int P>
__global__ __launch_bounds__(kUpdThreads) void k_irls_update(TileDesign td,
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
                                                    int32_t* __restrict__ nrecheck,
                                                    double* __restrict__ bprev, int polish) {
  irls_update_body<P>(td, blockIdx.x, active, nactive, part, beta, kvec, max_iter, eps, iters,
                      status, next, nnext, band, recheck, nrecheck, bprev, polish);
}

// ---------------------------------------------------------------------------
// Joint rounds: one accumulate launch (per design width and kernel) and one
// update launch (per design width) per round for many problems.  Item q owns
// workgroups [wg0, wg0 + nwg) of the accumulate grid (wg0 a multiple of 8, so
// the XCD walk of tile_coords holds within each problem) and blocks
// [ub0, ub0 + nub) of the update grid; items are in increasing wg0 and ub0.
__device__ __forceinline__ int joint_find(const JointItem* __restrict__ it, int n, int b,
                                          bool upd) {
  int q = 0;
  while (q + 1 < n && (upd ? it[q + 1].ub0 : it[q + 1].wg0) <= b) ++q;
  return q;
}

template <int P>
__global__ __launch_bounds__(256) void k_irls_accum_joint(const JointItem* __restrict__ items,
                                                          int nitems) {
  __shared__ double e2[64];
  exp_tab_fill(e2, threadIdx.x, blockDim.x);
  __syncthreads();
  const int b = blockIdx.x;
  const JointItem& J = items[joint_find(items, nitems, b, false)];
  const int wb = b - J.wg0;
  if (wb >= J.nwg) return;  // padding to the next multiple of 8
  irls_accum_body<P>(J.td, J.c, J.active, J.nactive, J.beta, J.kvec, J.part, e2, wb, J.nwg);
}

template <int P>
__global__ __launch_bounds__(64 * (ws_waves(P) + 1)) void k_irls_accum_ws_joint(
    const JointItem* __restrict__ items, int nitems) {
  const int b = blockIdx.x;
  const JointItem& J = items[joint_find(items, nitems, b, false)];
  const int wb = b - J.wg0;
  if (wb >= J.nwg) return;
  ws_body<P, 0, 1>(J.td, J.c, J.active, J.nactive, J.beta, J.kvec, J.part, wb, J.nwg);
}

template <int P>
__global__ __launch_bounds__(kUpdThreads) void k_irls_update_joint(
    const JointItem* __restrict__ items, int nitems) {
  const int b = blockIdx.x;
  const JointItem& J = items[joint_find(items, nitems, b, true)];
  const int i = b - J.ub0;
  if (i >= J.nub) return;
  irls_update_body<P>(J.td, i, J.active, J.nactive, J.part, J.ubeta, J.kvec, J.max_iter, J.eps,
                      J.iters, J.status, J.next, J.nnext, J.band, J.recheck, J.nrecheck,
                      J.bprev, J.polish);
}

// genes list[0..n): beta <- beta0, iteration count and status cleared
__global__ void k_restart_genes(const int32_t* __restrict__ list, int n, int p,
                                const double* __restrict__ beta0, double* __restrict__ beta,
                                int32_t* __restrict__ iters, int32_t* __restrict__ status,
                                double* __restrict__ bprev) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t g = list[i];
  for (int j = 0; j < p; ++j) beta[g * p + j] = beta0[g * p + j];
  if (bprev) bprev[g * p] = __builtin_nan("");  // refit in FP64: nothing to polish
  iters[g] = 0;
  status[g] = 0;
}

hipError_t launch_restart_genes(hipStream_t s, const int32_t* list, int n, int p,
                                const double* beta0, double* beta, int32_t* iters,
                                int32_t* status, double* bprev) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_restart_genes, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, list, n,
                     p, beta0, beta, iters, status, bprev);
  return hipGetLastError();
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

hipError_t launch_irls_accum_bf16(hipStream_t s, const TileDesign& td, Counts c, int p,
                                  const int32_t* active, const int32_t* nactive, int nA_bound,
                                  const double* beta, const double* kvec, double* part) {
  if (nA_bound <= 0) return hipSuccess;
  const int ntile = (nA_bound + 15) / 16;
  const int groups = (ntile + 3) / 4;
  const dim3 grid((unsigned)((int64_t)groups * td.nchunk));
  DV_TILE_P(p, k_irls_accum_bf16, grid, dim3(256), s, td, c, active, nactive, beta, kvec, part);
  return hipGetLastError();
}

hipError_t launch_irls_accum_ws(hipStream_t s, const TileDesign& td, Counts c, int p,
                                const int32_t* active, const int32_t* nactive, int nA_bound,
                                const double* beta, const double* kvec, double* part, int tpw) {
  if (nA_bound <= 0) return hipSuccess;
  const int ntile = (nA_bound + 15) / 16;
  // 8-byte counts have no direct-to-LDS form here: the register kernel
  if (c.dtype == 2)
    return launch_irls_accum_bf16(s, td, c, p, active, nactive, nA_bound, beta, kvec, part);
  // 4-byte counts must start 4-byte aligned (every row then does); rows not
  // 16-byte aligned run with one tile per wave
  if ((reinterpret_cast<uintptr_t>(c.ptr) & 3) != 0) return hipErrorInvalidValue;
  if ((c.ld & 3) != 0 || (reinterpret_cast<uintptr_t>(c.ptr) & 15) != 0) tpw = 1;
  const int nw = ws_waves(p, tpw);
  const int per_group = nw * tpw;
  const int wgroups = (ntile + per_group - 1) / per_group;
  const dim3 grid((unsigned)((int64_t)wgroups * td.nchunk));
  if (tpw == 2) {
    DV_TILE_P(p, k_irls_accum_ws2, grid, dim3(64 * (nw + 1)), s, td, c, active, nactive, beta,
              kvec, part);
  } else {
    DV_TILE_P(p, k_irls_accum_ws, grid, dim3(64 * (nw + 1)), s, td, c, active, nactive, beta,
              kvec, part);
  }
  return hipGetLastError();
}

int ws_tiles_per_workgroup(int p, int tpw) { return ws_waves(p, tpw) * tpw; }

hipError_t launch_irls_accum_joint(hipStream_t s, int kind, int p, const JointItem* items,
                                   int nitems, int total_wg) {
  if (nitems <= 0 || total_wg <= 0) return hipSuccess;
  if (kind == kJointWs) {
    const int nw = ws_waves(p, 1);
    DV_TILE_P(p, k_irls_accum_ws_joint, dim3((unsigned)total_wg), dim3(64 * (nw + 1)), s, items,
              nitems);
  } else {
    DV_TILE_P(p, k_irls_accum_joint, dim3((unsigned)total_wg), dim3(256), s, items, nitems);
  }
  return hipGetLastError();
}

hipError_t launch_irls_update_joint(hipStream_t s, int p, const JointItem* items, int nitems,
                                    int total_ub) {
  if (nitems <= 0 || total_ub <= 0) return hipSuccess;
  DV_TILE_P(p, k_irls_update_joint, dim3((unsigned)total_ub), dim3(kUpdThreads), s, items,
            nitems);
  return hipGetLastError();
}

hipError_t launch_merge_flags(hipStream_t s, int32_t* status, const int32_t* flag, int64_t G,
                              int32_t bit) {
  if (G <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_merge_flags, dim3((unsigned)((G + 255) / 256)), dim3(256), 0, s, status,
                     flag, G, bit);
  return hipGetLastError();
}

hipError_t launch_bread_mixed(hipStream_t s, const TileDesign& td, Counts c, int p,
                              const int32_t* genes, const int32_t* ngenes, int64_t G,
                              const double* beta, const double* theta, double* part,
                              double* bread, int32_t* singular, int32_t* fallback) {
  if (G <= 0) return hipSuccess;
  if (c.dtype == 2 || (reinterpret_cast<uintptr_t>(c.ptr) & 3) != 0) return hipErrorInvalidValue;
  const int ntile = (int)((G + 15) / 16);
  const int nw = ws_waves(p);
  const int wgroups = (ntile + nw - 1) / nw;
  const dim3 grid((unsigned)((int64_t)wgroups * td.nchunk));
  DV_TILE_P(p, k_bread_accum_ws, grid, dim3(64 * (nw + 1)), s, td, c, genes, ngenes, beta, theta,
            part);
  DV_TILE_P(p, k_bread_finish_tile, dim3((unsigned)G), dim3(64), s, td, G, part, bread, singular,
            fallback);
  return hipGetLastError();
}

hipError_t launch_bread_meat_mixed(hipStream_t s, const TileDesign& td, const ChunkTable& ct,
                                   Counts c, int p, const int32_t* genes, const int32_t* ngenes,
                                   int64_t G, const double* beta, const double* theta,
                                   double* part, double* spart, double* bread, int32_t* singular,
                                   int32_t* fallback) {
  if (G <= 0) return hipSuccess;
  if (c.dtype == 2 || (reinterpret_cast<uintptr_t>(c.ptr) & 3) != 0 || ct.nchunk <= 0)
    return hipErrorInvalidValue;
  const int ntile = (int)((G + 15) / 16);
  const int nw = ws_waves(p);
  const int wgroups = (ntile + nw - 1) / nw;
  const dim3 grid((unsigned)((int64_t)wgroups * ct.nchunk));
  DV_TILE_P(p, k_bread_meat_ws, grid, dim3(64 * (nw + 1)), s, td, ct, c, genes, ngenes, beta,
            theta, part, spart);
  TileDesign tc = td;  // the bread partials are laid out over ct's chunks
  tc.nchunk = ct.nchunk;
  DV_TILE_P(p, k_bread_finish_tile, dim3((unsigned)G), dim3(64), s, tc, G, part, bread, singular,
            fallback);
  return hipGetLastError();
}

// Split reduction of a round's chunk partials for rounds of few tiles (many
// chunks each): workgroup (tile, r) sums chunks [r L, (r + 1) L) of the
// tile's 16 x ncp partials into out[tile][r][16][ncp], L = ceil(nchunk / R),
// every thread's loads independent of each other.  The update then sums the
// R slices (a design with nchunk = R over out).
constexpr int kRedThreads = 256;

__global__ __launch_bounds__(kRedThreads) void k_irls_reduce(int nchunk, int ncp, int R,
                                                             const int32_t* __restrict__ nactive,
                                                             const double* __restrict__ part,
                                                             double* __restrict__ out) {
  const int ntile = (*nactive + 15) / 16;
  const int tile = blockIdx.x / R, r = blockIdx.x % R;
  if (tile >= ntile) return;
  const int L = (nchunk + R - 1) / R;
  const int c0 = r * L, c1 = min(nchunk, c0 + L);
  const int64_t stride = (int64_t)16 * ncp;
  const double* src = part + (int64_t)tile * nchunk * stride;
  double* dst = out + ((int64_t)tile * R + r) * stride;
  for (int idx = threadIdx.x; idx < 16 * ncp; idx += kRedThreads) {
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    int ch = c0;
    for (; ch + 4 <= c1; ch += 4) {
#pragma unroll
      for (int u = 0; u < 4; ++u) a[u] += src[(int64_t)(ch + u) * stride + idx];
    }
    for (; ch < c1; ++ch) a[0] += src[(int64_t)ch * stride + idx];
    dst[idx] = (a[0] + a[1]) + (a[2] + a[3]);
  }
}

hipError_t launch_irls_reduce(hipStream_t s, int nchunk, int ncp, int R, int ntile_bound,
                              const int32_t* nactive, const double* part, double* out) {
  if (ntile_bound <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_irls_reduce, dim3((unsigned)(ntile_bound * R)), dim3(kRedThreads), 0, s,
                     nchunk, ncp, R, nactive, part, out);
  return hipGetLastError();
}

hipError_t launch_irls_update(hipStream_t s, const TileDesign& td, int p,
                              const int32_t* active, const int32_t* nactive, int nA_bound,
                              const double* part, double* beta, const double* kvec,
                              int max_iter, double eps, int32_t* iters, int32_t* status,
                              int32_t* next, int32_t* nnext, double band, int32_t* recheck,
                              int32_t* nrecheck, double* bprev, int polish) {
  if (nA_bound <= 0) return hipSuccess;
  const dim3 grid((unsigned)nA_bound);
  DV_TILE_P(p, k_irls_update, grid, dim3(kUpdThreads), s, td, active, nactive, part, beta, kvec,
            max_iter, eps, iters, status, next, nnext, band, recheck, nrecheck, bprev, polish);
  return hipGetLastError();
}

__global__ void k_fill_f64(double* __restrict__ x, int64_t n, double v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}

hipError_t launch_fill_f64(hipStream_t s, double* x, int64_t n, double v) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_f64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x, n, v);
  return hipGetLastError();
}

// Genes whose last mixed step is redone in FP64: a saved pre-step beta (not
// NaN: refit genes and genes never stepped keep NaN), no NaN / max-iteration
// status, and some coefficient below `below` in magnitude (where the mixed
// trajectory's ~5e-7 absolute error is a large relative one).
__global__ void k_polish_select(int64_t G, int p, const double* __restrict__ beta,
                                const double* __restrict__ bprev,
                                const int32_t* __restrict__ status, double below,
                                int32_t* __restrict__ list, int32_t* __restrict__ nlist) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= G) return;
  if (isnan(bprev[g * p])) return;
  if (status[g] & (DEVIL_GENE_NAN | DEVIL_GENE_MAXITER)) return;
  bool small = false;
  for (int j = 0; j < p; ++j) small = small || fabs(beta[g * p + j]) < below;
  if (small) list[atomicAdd(nlist, 1)] = (int32_t)g;
}

hipError_t launch_polish_select(hipStream_t s, int64_t G, int p, const double* beta,
                                const double* bprev, const int32_t* status, double below,
                                int32_t* list, int32_t* nlist) {
  if (G <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_polish_select, dim3((unsigned)((G + 255) / 256)), dim3(256), 0, s, G, p,
                     beta, bprev, status, below, list, nlist);
  return hipGetLastError();
}

hipError_t launch_iota(hipStream_t s, int32_t* p, int n) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_iota, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n);
  return hipGetLastError();
}

}  // namespace dv
