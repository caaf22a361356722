// Operand loading shared by the MFMA tile kernels (IRLS and MLE passes):
// a lane of a 16-gene tile reads, for 16-cell block b, its four consecutive
// counts, the eta-product A operands, the four offsets and the K operands.
#pragma once
#include <hip/hip_runtime.h>

#include "device_common.hpp"
#include "kernels.hpp"

namespace dv {

// (chunk, gene group) of this workgroup in a grid of groups x nchunk.  Blocks
// are dealt to the 8 XCDs round-robin by blockIdx; with nchunk a multiple of 8
// the XCD b % 8 takes the chunks congruent to it and walks them chunk-major,
// so the workgroups resident on one XCD share one chunk's operand tables in
// that XCD's L2.
__device__ __forceinline__ void tile_coords(int nchunk, int& chunk, int& group) {
  const int b = blockIdx.x;
  const int groups = gridDim.x / nchunk;
  if ((nchunk & 7) == 0) {
    const int x = b & 7, i = b >> 3;
    chunk = x + 8 * (i / groups);
    group = i % groups;
  } else {
    chunk = b % nchunk;
    group = b / nchunk;
  }
}

// Four consecutive counts of a row in their storage type, for kernels that
// load a block ahead of its use: converting right after the load would make
// the wave wait for the load there, so the conversion is left to the use.
// Cells at or past n read the row's cell n - 1 (a finite count of the same
// gene); the operand tables of padded cells are zero, so such a cell adds
// exactly zero to every product.
template <typename CT>
struct Vec4Of;
template <>
struct Vec4Of<int32_t> { typedef i4v T; };
template <>
struct Vec4Of<float> { typedef f4v T; };
template <>
struct Vec4Of<double> { typedef d4 T; };

template <typename CT>
__device__ __forceinline__ void load4_raw(const CT* row, int64_t c, int64_t n, bool vec,
                                          CT (&v)[4]) {
  if (c + 4 <= n) {
    if (vec) {
      const typename Vec4Of<CT>::T t = *reinterpret_cast<const typename Vec4Of<CT>::T*>(row + c);
#pragma unroll
      for (int r = 0; r < 4; ++r) v[r] = t[r];
    } else {
#pragma unroll
      for (int r = 0; r < 4; ++r) v[r] = row[c + r];
    }
  } else {
#pragma unroll
    for (int r = 0; r < 4; ++r) v[r] = row[min(c + r, n - 1)];
  }
}

// body(ops, i) for the steps i of [i0, i1), the operands of step i + 1
// loading while step i computes, into two buffers in turn.  Loads return in
// issue order, so the prefetch is kept free of what would make the wait for
// the current operands also wait for it: it is unconditional (the last step
// loads itself again), the buffers are never copied, and a loader issues its
// count loads (from HBM) after its operand-table loads (mostly from L2).
template <typename Ops, typename Load, typename Body>
__device__ __forceinline__ void pipeline(int64_t i0, int64_t i1, Load&& load, Body&& body) {
  if (i0 >= i1) return;
  Ops A, B;
  load(i0, A);
  for (int64_t i = i0; i < i1; i += 2) {
    load(min(i + 1, i1 - 1), B);
    body(A, i);
    if (i + 1 >= i1) break;
    load(min(i + 2, i1 - 1), A);
    body(B, i + 1);
  }
}

// As pipeline, with the counts of step i + 2 in flight while step i
// computes: the operand tables (read from L2) are loaded a step ahead and the
// counts (read from HBM) two steps ahead, issued after the tables, so that
// the wait for a step's tables never waits for counts issued later.  Steps
// run in groups of 4 (four count buffers, two table buffers, fixed indices);
// the last group's leftover steps load and run one at a time.
template <typename T, typename Y, typename LoadT, typename LoadY, typename Body>
__device__ __forceinline__ void pipeline_ty(int64_t i0, int64_t i1, LoadT&& loadt, LoadY&& loady,
                                            Body&& body) {
  if (i0 >= i1) return;
  T ta[2];
  Y ya[4];
  int64_t i = i0;
  if (i1 - i0 >= 4) {
    loadt(i0, ta[0]);
    loady(i0, ya[0]);
    loady(i0 + 1, ya[1]);
    const int64_t last = i1 - 1;
    for (; i + 4 <= i1; i += 4) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        loadt(min(i + u + 1, last), ta[(u + 1) & 1]);
        loady(min(i + u + 2, last), ya[(u + 2) & 3]);
        body(ta[u & 1], ya[u], i + u);
      }
    }
  }
  for (; i < i1; ++i) {
    loadt(i, ta[0]);
    loady(i, ya[0]);
    body(ta[0], ya[0], i);
  }
}

// Operands of one 16-cell block for one lane: counts (as stored), the eta
// A operands, four offsets (log or exp scale, by the caller's table) and the
// first NTMAX K tiles (none when NTMAX is 0).
template <int NJ, int NTMAX, typename CT>
struct TileBlock {
  CT y[4];
  double xa[NJ];
  d4 off;
  d4 kb[NTMAX > 0 ? NTMAX : 1];
};

// Counts of 16-cell block b for lane (s, h); yh = the gene's row + 4 h.
// Blocks wholly inside the row (all but the last) take the unchecked loads;
// the branch is uniform.
template <typename CT>
__device__ __forceinline__ void load_block_y(const TileDesign& td, const CT* yrow, const CT* yh,
                                             bool vec, int64_t b, int h, CT (&y)[4]) {
  if (b * 16 + 16 <= td.n) {
    const CT* p = yh + b * 16;
    if (vec) {
      const typename Vec4Of<CT>::T t = *reinterpret_cast<const typename Vec4Of<CT>::T*>(p);
#pragma unroll
      for (int r = 0; r < 4; ++r) y[r] = t[r];
    } else {
#pragma unroll
      for (int r = 0; r < 4; ++r) y[r] = p[r];
    }
  } else {
    load4_raw<CT>(yrow, b * 16 + 4 * h, td.n, vec, y);
  }
}

// Table addresses are a uniform per-block base plus the lane's offset (no
// 64-bit vector arithmetic).
template <int NJ, int NTMAX, typename CT>
__device__ __forceinline__ void load_tile_block(const TileDesign& td, const CT* yrow,
                                                const CT* yh, bool vec, int64_t b, int lane,
                                                int h, int NT, const double* __restrict__ offs,
                                                TileBlock<NJ, NTMAX, CT>& o) {
  const double* xs = td.Xp + b * (NJ * 64);
#pragma unroll
  for (int j = 0; j < NJ; ++j) o.xa[j] = xs[j * 64 + lane];
  o.off = reinterpret_cast<const d4*>(offs + b * 16)[h];
  if constexpr (NTMAX > 0) {
    const d4* ks = reinterpret_cast<const d4*>(td.Kp) + b * NT * 64;
#pragma unroll
    for (int t = 0; t < NTMAX; ++t) {
      const int tt = t < NT ? t : NT - 1;  // unused tiles load tile NT - 1 (no branch)
      o.kb[t] = ks[tt * 64 + lane];
    }
  }
  load_block_y<CT>(td, yrow, yh, vec, b, h, o.y);
}

}  // namespace dv
