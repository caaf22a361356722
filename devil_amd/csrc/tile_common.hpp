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

}  // namespace dv
