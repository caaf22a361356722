// gfx950 passes that prepare the count matrix for the fit:
//
//   size factors   calculate_sf (R/size_factors.R:37-127): per cell the sum
//                  over genes (normed_sum) or the mean and minimum of
//                  log1p(y) (psinorm), then the division by the geometric
//                  mean.  One streaming pass: a workgroup owns 1024
//                  consecutive cells and one slice of the genes; every load
//                  instruction of a wavefront reads 256 contiguous bytes of
//                  one gene row.  Slice partials are summed in a fixed order
//                  (integer counts: exact 64-bit sums), so results do not
//                  depend on scheduling.
//   group cells    group_data (R/group_data.R:38-69): out[g, i] = in[g, ord[i]]
//                  (writes coalesced; reads follow the stable cluster order,
//                  contiguous within each cluster).
//   CSC densify    dgCMatrix (column = cell) counts into the gene-major dense
//                  matrix the fit streams: zero fill, then one wavefront per
//                  cell scatters its non-zeros.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "../../include/devil_hip.h"
#include "device_common.hpp"
#include "kernels.hpp"

namespace dv {

constexpr int kSfCellsPerBlock = 1024;  // 256 threads x 4 cells

template <typename CT, int METHOD>
__device__ __forceinline__ void sf_slice(const CT* __restrict__ Y, int64_t ld, int64_t g0,
                                         int64_t g1, int64_t n, int64_t i0,
                                         int64_t* __restrict__ isum, double* __restrict__ dsum,
                                         double* __restrict__ dmin) {
  constexpr bool kInt = sizeof(CT) == 4 && CT(0.5) == CT(0);
  int64_t si[4] = {0, 0, 0, 0};
  double sd[4] = {0.0, 0.0, 0.0, 0.0};
  double mn[4] = {INFINITY, INFINITY, INFINITY, INFINITY};
  bool in[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) in[k] = i0 + 256 * k < n;
  for (int64_t g = g0; g < g1; ++g) {
    const CT* row = Y + g * ld + i0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (!in[k]) continue;
      const CT y = row[256 * k];
      if constexpr (METHOD == 0) {
        if constexpr (kInt) si[k] += (int64_t)y;
        else sd[k] += (double)y;
      } else {
        sd[k] += log1p_f64((double)y);
        mn[k] = fmin(mn[k], (double)y);
      }
    }
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (!in[k]) continue;
    const int64_t i = i0 + 256 * k;
    if constexpr (METHOD == 0 && kInt) isum[i] = si[k];
    else dsum[i] = sd[k];
    if constexpr (METHOD == 1) dmin[i] = mn[k];
  }
}

// grid (cell blocks, slices): partials [slice][cell]
template <int METHOD>
__global__ __launch_bounds__(256) void k_sf_partial(Counts c, int64_t G, int64_t n,
                                                    int slices, int64_t* __restrict__ isum,
                                                    double* __restrict__ dsum,
                                                    double* __restrict__ dmin) {
  const int64_t i0 = (int64_t)blockIdx.x * kSfCellsPerBlock + threadIdx.x;
  const int sl = blockIdx.y;
  const int64_t per = (G + slices - 1) / slices;
  const int64_t g0 = (int64_t)sl * per, g1 = min(G, g0 + per);
  const int64_t off = (int64_t)sl * n;
  if (c.dtype == 0)
    sf_slice<int32_t, METHOD>(static_cast<const int32_t*>(c.ptr), c.ld, g0, g1, n, i0,
                              isum + off, dsum + off, dmin + off);
  else if (c.dtype == 1)
    sf_slice<float, METHOD>(static_cast<const float*>(c.ptr), c.ld, g0, g1, n, i0,
                            isum + off, dsum + off, dmin + off);
  else
    sf_slice<double, METHOD>(static_cast<const double*>(c.ptr), c.ld, g0, g1, n, i0,
                             isum + off, dsum + off, dmin + off);
}

// per cell: the statistic (column sum, or mean - min of log1p) and its log;
// zero_flag set when a normed_sum column sums to 0
template <int METHOD>
__global__ void k_sf_cell(int64_t G, int64_t n, int slices, int integer_counts,
                          const int64_t* __restrict__ isum, const double* __restrict__ dsum,
                          const double* __restrict__ dmin, double* __restrict__ sf,
                          double* __restrict__ logsf, int32_t* __restrict__ zero_flag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double v;
  if (METHOD == 0) {
    if (integer_counts) {
      int64_t s = 0;
      for (int sl = 0; sl < slices; ++sl) s += isum[(int64_t)sl * n + i];
      v = (double)s;
    } else {
      double s = 0.0;
      for (int sl = 0; sl < slices; ++sl) s += dsum[(int64_t)sl * n + i];
      v = s;
    }
    if (v == 0.0) atomicOr(zero_flag, 1);
  } else {
    double s = 0.0, m = INFINITY;
    for (int sl = 0; sl < slices; ++sl) {
      s += dsum[(int64_t)sl * n + i];
      m = fmin(m, dmin[(int64_t)sl * n + i]);
    }
    v = s / (double)G - log1p(m);
  }
  sf[i] = v;
  logsf[i] = log(v);
}

// one workgroup: mean of logsf (fixed-order sum), then sf /= exp(mean)
__global__ __launch_bounds__(1024) void k_sf_scale(int64_t n, const double* __restrict__ logsf,
                                                   double* __restrict__ sf) {
  __shared__ double red[1024];
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) s += logsf[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  const double gm = exp(red[0] / (double)n);
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) sf[i] = sf[i] / gm;
}

hipError_t launch_size_factors(hipStream_t s, Counts c, int64_t G, int64_t n, int method,
                               int slices, int64_t* isum, double* dsum, double* dmin,
                               double* sf, double* logsf, int32_t* zero_flag) {
  if (n <= 0) return hipSuccess;
  const dim3 grid((unsigned)((n + kSfCellsPerBlock - 1) / kSfCellsPerBlock), (unsigned)slices);
  const dim3 g1((unsigned)((n + 255) / 256));
  if (method == 0) {
    hipLaunchKernelGGL(k_sf_partial<0>, grid, dim3(256), 0, s, c, G, n, slices, isum, dsum,
                       dmin);
    hipLaunchKernelGGL(k_sf_cell<0>, g1, dim3(256), 0, s, G, n, slices, c.dtype == 0 ? 1 : 0,
                       isum, dsum, dmin, sf, logsf, zero_flag);
  } else {
    hipLaunchKernelGGL(k_sf_partial<1>, grid, dim3(256), 0, s, c, G, n, slices, isum, dsum,
                       dmin);
    hipLaunchKernelGGL(k_sf_cell<1>, g1, dim3(256), 0, s, G, n, slices, c.dtype == 0 ? 1 : 0,
                       isum, dsum, dmin, sf, logsf, zero_flag);
  }
  hipLaunchKernelGGL(k_sf_scale, dim3(1), dim3(1024), 0, s, n, logsf, sf);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// group_data: out[g, i] = in[g, ord[i]]; grid (cell blocks of 1024, genes)
// ---------------------------------------------------------------------------
template <typename CT>
__global__ __launch_bounds__(256) void k_group_cells(const CT* __restrict__ in, int64_t ld_in,
                                                     int64_t G, int64_t n,
                                                     const int64_t* __restrict__ ord,
                                                     CT* __restrict__ out, int64_t ld_out) {
  const int64_t g = blockIdx.y + (int64_t)blockIdx.z * 65535;
  if (g >= G) return;
  const CT* src = in + g * ld_in;
  CT* dst = out + g * ld_out;
  const int64_t i0 = (int64_t)blockIdx.x * 1024 + threadIdx.x;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int64_t i = i0 + 256 * k;
    if (i < n) dst[i] = src[ord[i]];
  }
}

hipError_t launch_group_cells(hipStream_t s, Counts c, int64_t G, int64_t n,
                              const int64_t* ord, void* out, int64_t ld_out) {
  if (G <= 0 || n <= 0) return hipSuccess;
  const int64_t gz = (G + 65534) / 65535;
  const dim3 grid((unsigned)((n + 1023) / 1024), (unsigned)std::min<int64_t>(G, 65535),
                  (unsigned)gz);
  if (c.dtype == 0)
    hipLaunchKernelGGL(k_group_cells<int32_t>, grid, dim3(256), 0, s,
                       static_cast<const int32_t*>(c.ptr), c.ld, G, n, ord,
                       static_cast<int32_t*>(out), ld_out);
  else if (c.dtype == 1)
    hipLaunchKernelGGL(k_group_cells<float>, grid, dim3(256), 0, s,
                       static_cast<const float*>(c.ptr), c.ld, G, n, ord,
                       static_cast<float*>(out), ld_out);
  else
    hipLaunchKernelGGL(k_group_cells<double>, grid, dim3(256), 0, s,
                       static_cast<const double*>(c.ptr), c.ld, G, n, ord,
                       static_cast<double*>(out), ld_out);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// CSC (column = cell) -> gene-major dense; out is zero-filled first
// ---------------------------------------------------------------------------
template <typename OT>
__global__ __launch_bounds__(256) void k_csc_scatter(const int64_t* __restrict__ colptr,
                                                     const int32_t* __restrict__ rowind,
                                                     const double* __restrict__ values,
                                                     int64_t n, OT* __restrict__ out,
                                                     int64_t ld) {
  const int64_t i = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= n) return;
  const int lane = threadIdx.x & 63;
  for (int64_t k = colptr[i] + lane; k < colptr[i + 1]; k += 64)
    out[(int64_t)rowind[k] * ld + i] = (OT)values[k];
}

template <typename OT>
__global__ void k_fill_zero(OT* __restrict__ out, int64_t G, int64_t n, int64_t ld) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t g = t / n, i = t - g * n;
  if (g < G) out[g * ld + i] = (OT)0;
}

hipError_t launch_csc_to_dense(hipStream_t s, const int64_t* colptr, const int32_t* rowind,
                               const double* values, int64_t G, int64_t n, int out_dtype,
                               void* out, int64_t ld) {
  if (G <= 0 || n <= 0) return hipSuccess;
  const dim3 gz((unsigned)((G * n + 255) / 256)), gs((unsigned)((n + 3) / 4));
  if (out_dtype == 0) {
    hipLaunchKernelGGL(k_fill_zero<int32_t>, gz, dim3(256), 0, s, static_cast<int32_t*>(out),
                       G, n, ld);
    hipLaunchKernelGGL(k_csc_scatter<int32_t>, gs, dim3(256), 0, s, colptr, rowind, values, n,
                       static_cast<int32_t*>(out), ld);
  } else if (out_dtype == 1) {
    hipLaunchKernelGGL(k_fill_zero<float>, gz, dim3(256), 0, s, static_cast<float*>(out), G, n,
                       ld);
    hipLaunchKernelGGL(k_csc_scatter<float>, gs, dim3(256), 0, s, colptr, rowind, values, n,
                       static_cast<float*>(out), ld);
  } else {
    hipLaunchKernelGGL(k_fill_zero<double>, gz, dim3(256), 0, s, static_cast<double*>(out), G,
                       n, ld);
    hipLaunchKernelGGL(k_csc_scatter<double>, gs, dim3(256), 0, s, colptr, rowind, values, n,
                       static_cast<double*>(out), ld);
  }
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// rows of 4-byte counts copied into rows of ld_dst elements (a multiple of 4,
// so every destination row starts 16-byte aligned); grid (cell blocks of
// 1024, genes)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_pad_rows(const uint32_t* __restrict__ src,
                                                  int64_t ld_src, int64_t G, int64_t n,
                                                  uint32_t* __restrict__ dst, int64_t ld_dst) {
  const int64_t g = blockIdx.y + (int64_t)blockIdx.z * 65535;
  if (g >= G) return;
  const uint32_t* s = src + g * ld_src;
  uint32_t* d = dst + g * ld_dst;
  const int64_t i0 = (int64_t)blockIdx.x * 1024 + threadIdx.x;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int64_t i = i0 + 256 * k;
    if (i < n) d[i] = s[i];
  }
}

hipError_t launch_pad_rows(hipStream_t s, const void* src, int64_t ld_src, int64_t G, int64_t n,
                           void* dst, int64_t ld_dst) {
  if (G <= 0 || n <= 0) return hipSuccess;
  const dim3 grid((unsigned)((n + 1023) / 1024), (unsigned)std::min<int64_t>(G, 65535),
                  (unsigned)((G + 65534) / 65535));
  hipLaunchKernelGGL(k_pad_rows, grid, dim3(256), 0, s, static_cast<const uint32_t*>(src),
                     ld_src, G, n, static_cast<uint32_t*>(dst), ld_dst);
  return hipGetLastError();
}

// argument checks on the device: bad_flag |= 1 when an index is out of range
__global__ void k_check_order(const int64_t* __restrict__ ord, int64_t n,
                              int32_t* __restrict__ bad_flag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && (ord[i] < 0 || ord[i] >= n)) atomicOr(bad_flag, 1);
}

__global__ void k_check_csc(const int64_t* __restrict__ colptr, const int32_t* __restrict__ rowind,
                            int64_t G, int64_t n, int64_t nnz, int32_t* __restrict__ bad_flag) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n && (colptr[t] > colptr[t + 1] || colptr[t] < 0 || colptr[t + 1] > nnz))
    atomicOr(bad_flag, 1);
  if (t == 0 && colptr[0] != 0) atomicOr(bad_flag, 1);
  for (int64_t k = t; k < nnz; k += (int64_t)gridDim.x * blockDim.x)
    if (rowind[k] < 0 || rowind[k] >= G) atomicOr(bad_flag, 2);
}

hipError_t launch_check_order(hipStream_t s, const int64_t* ord, int64_t n, int32_t* flag) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_check_order, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, ord, n,
                     flag);
  return hipGetLastError();
}

hipError_t launch_check_csc(hipStream_t s, const int64_t* colptr, const int32_t* rowind,
                            int64_t G, int64_t n, int64_t nnz, int32_t* flag) {
  const int64_t need_n = (n + 255) / 256;
  const int64_t need_z = std::min<int64_t>((nnz + 255) / 256, 65536);
  const int64_t blocks = std::max<int64_t>(1, std::max(need_n, need_z));
  hipLaunchKernelGGL(k_check_csc, dim3((unsigned)blocks), dim3(256), 0, s, colptr, rowind, G,
                     n, nnz, flag);
  return hipGetLastError();
}

}  // namespace dv
