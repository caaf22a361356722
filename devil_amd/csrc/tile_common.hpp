// Operand loading shared by the MFMA tile kernels (IRLS and MLE passes):
// a lane of a 16-gene tile reads, for 16-cell block b, its four consecutive
// counts, the eta-product A operands, the four offsets and the K operands.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "device_common.hpp"
#include "kernels.hpp"

namespace dv {

// (chunk, gene group) of this workgroup in a grid of groups x nchunk.  Blocks
// are dealt to the 8 XCDs round-robin by blockIdx; with nchunk a multiple of 8
// the XCD b % 8 takes the chunks congruent to it and walks them chunk-major,
// so the workgroups resident on one XCD share one chunk's operand tables in
// that XCD's L2.
// tile_coords_at: the same for workgroup b of nb (a range of a larger grid
// that starts at a multiple of 8, as the joint launches lay them out).
__device__ __forceinline__ void tile_coords_at(int b, int nb, int nchunk, int& chunk,
                                               int& group) {
  const int groups = nb / nchunk;
  if ((nchunk & 7) == 0) {
    const int x = b & 7, i = b >> 3;
    chunk = x + 8 * (i / groups);
    group = i % groups;
  } else {
    chunk = b % nchunk;
    group = b / nchunk;
  }
}

__device__ __forceinline__ void tile_coords(int nchunk, int& chunk, int& group) {
  tile_coords_at(blockIdx.x, gridDim.x, nchunk, chunk, group);
}

// Counts are loaded in their storage type and converted where they are used:
// converting right after the load would make the wave wait for the load
// there, and a prefetch would hide nothing.
template <typename CT>
struct Vec4Of;
template <>
struct Vec4Of<int32_t> { typedef i4v T; };
template <>
struct Vec4Of<float> { typedef f4v T; };
template <>
struct Vec4Of<double> { typedef d4 T; };

// Four consecutive counts at p, inside the row.  VEC: rows are 16-byte
// aligned with a multiple of 4 elements, one vector load.
template <typename CT, bool VEC>
__device__ __forceinline__ void load4_fast(const CT* p, CT (&v)[4]) {
  if constexpr (VEC) {
    const typename Vec4Of<CT>::T t = *reinterpret_cast<const typename Vec4Of<CT>::T*>(p);
#pragma unroll
    for (int r = 0; r < 4; ++r) v[r] = t[r];
  } else {
#pragma unroll
    for (int r = 0; r < 4; ++r) v[r] = p[r];
  }
}

// Counts of cells c .. c + 3 of a row of n cells, for the row's last block:
// cells at or past n read cell n - 1 (a finite count of the same gene); the
// operand tables of padded cells are zero, so such a cell adds exactly zero
// to every product, and the passes that use counts directly mask them.
template <typename CT>
__device__ __forceinline__ void load4_tail(const CT* row, int64_t c, int64_t n, CT (&v)[4]) {
#pragma unroll
  for (int r = 0; r < 4; ++r) v[r] = row[min(c + r, n - 1)];
}

// (count type, VEC) dispatch of a tile pass on its counts.
template <typename F>
__device__ __forceinline__ void with_counts(const Counts& c, F&& f) {
  const bool vec = (c.ld & 3) == 0 && (reinterpret_cast<uintptr_t>(c.ptr) & 15) == 0;
  if (c.dtype == 0) {
    const int32_t* y = static_cast<const int32_t*>(c.ptr);
    if (vec) f(y, std::true_type{});
    else f(y, std::false_type{});
  } else if (c.dtype == 1) {
    const float* y = static_cast<const float*>(c.ptr);
    if (vec) f(y, std::true_type{});
    else f(y, std::false_type{});
  } else {
    const double* y = static_cast<const double*>(c.ptr);
    if (vec) f(y, std::true_type{});
    else f(y, std::false_type{});
  }
}

// body(ops, i) for the steps i of [i0, i1).  The steps of [i0, i1f) lie
// inside every row: their operands are loaded a step ahead, into two buffers
// in turn.  Loads return in issue order, so nothing may make the wait for
// the current operands also wait for the prefetch: the prefetch is
// unconditional (the last step loads itself again) and branch-free, the
// buffers are never copied, and a loader issues its count loads (from HBM)
// after its operand-table loads (mostly from L2).  Steps from i1f on (a
// row's partial last step) load with load_tail when they run.
template <typename Ops, typename Load, typename LoadTail, typename Body>
__device__ __forceinline__ void pipeline(int64_t i0, int64_t i1f, int64_t i1, Load&& load,
                                         LoadTail&& load_tail, Body&& body) {
  if (i0 < i1f) {
    Ops A, B;
    load(i0, A);
    for (int64_t i = i0; i < i1f; i += 2) {
      load(min(i + 1, i1f - 1), B);
      body(A, i);
      if (i + 1 >= i1f) break;
      load(min(i + 2, i1f - 1), A);
      body(B, i + 1);
    }
  }
  for (int64_t i = max(i0, i1f); i < i1; ++i) {
    Ops T;
    load_tail(i, T);
    body(T, i);
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

// Table addresses are a uniform per-block base plus the lane's offset (no
// 64-bit vector arithmetic).
template <int NJ, int NTMAX, typename CT>
__device__ __forceinline__ void load_tile_tables(const TileDesign& td, int64_t b, int lane, int h,
                                                 int NT, const double* __restrict__ offs,
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
}

// Block b inside the row; yh = the gene's row + 4 h.
template <int NJ, int NTMAX, typename CT, bool VEC>
__device__ __forceinline__ void load_tile_block(const TileDesign& td, const CT* yh, int64_t b,
                                                int lane, int h, int NT,
                                                const double* __restrict__ offs,
                                                TileBlock<NJ, NTMAX, CT>& o) {
  load_tile_tables<NJ, NTMAX, CT>(td, b, lane, h, NT, offs, o);
  load4_fast<CT, VEC>(yh + b * 16, o.y);
}

// The row's partial last block.
template <int NJ, int NTMAX, typename CT>
__device__ __forceinline__ void load_tile_block_tail(const TileDesign& td, const CT* yrow,
                                                     int64_t b, int lane, int h, int NT,
                                                     const double* __restrict__ offs,
                                                     TileBlock<NJ, NTMAX, CT>& o) {
  load_tile_tables<NJ, NTMAX, CT>(td, b, lane, h, NT, offs, o);
  load4_tail<CT>(yrow, b * 16 + 4 * h, td.n, o.y);
}

}  // namespace dv
