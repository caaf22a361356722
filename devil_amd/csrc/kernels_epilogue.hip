// gfx950 per-gene epilogue kernels: k as the reference's float holds it, the
// Wald test with its t tail in log space, Benjamini-Hochberg by a device radix
// sort and suffix-minimum scan, and small layout helpers (transposes, fills).
// None of them reads the count matrix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <hipcub/hipcub.hpp>

#include "../../include/devil_hip.h"
#include "device_common.hpp"
#include "kernels.hpp"

namespace dv {

// k as src/beta.cpp:25-28 holds it: the float argument k = (float)alpha0,
// then k = 1.0 / k stored back into the float.
__global__ void k_k_from_alpha(const double* __restrict__ alpha0, int64_t G,
                               double* __restrict__ k) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= G) return;
  const float kf = (float)alpha0[g];
  const float kk = (float)(1.0 / (double)kf);
  k[g] = (double)kk;
}

// ---------------------------------------------------------------------------
// test_de (R/test.R:94-133), one thread per gene, then Benjamini-Hochberg.
// ---------------------------------------------------------------------------
__global__ void k_test_de(const double* __restrict__ beta, const double* __restrict__ bread,
                          const double* __restrict__ sandwich, int64_t G, int p, double df,
                          const double* __restrict__ contrast, double max_lfc,
                          double* __restrict__ pval, double* __restrict__ logp,
                          double* __restrict__ lfc) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= G) return;
  double l = 0.0;
  for (int j = 0; j < p; ++j) l += beta[g * p + j] * contrast[j];
  const double* H = bread + g * p * p;
  double v0 = 0.0;
  for (int a = 0; a < p; ++a)
    for (int bb = 0; bb < p; ++bb) v0 += contrast[a] * H[a + bb * p] * contrast[bb];
  double lp = log_pt2_upper(fabs(l) / sqrt(v0), df);
  if (sandwich) {
    const double* S = sandwich + g * p * p;
    double v1 = 0.0;
    for (int a = 0; a < p; ++a)
      for (int bb = 0; bb < p; ++bb) v1 += contrast[a] * S[a + bb * p] * contrast[bb];
    const double lp1 = log_pt2_upper(fabs(l) / sqrt(v1), df);
    if (isnan(lp) || isnan(lp1)) lp = __builtin_nan("");
    else if (lp1 > lp) lp = lp1;
  }
  pval[g] = exp(lp);
  if (logp) logp[g] = lp;
  double lf = l / log(2.0);
  if (lf >= max_lfc) lf = max_lfc;
  if (lf <= -max_lfc) lf = -max_lfc;
  lfc[g] = lf;
}

// BH by its step-up definition: rank_j = #{k : p_k <= p_j} over non-NaN;
// adj_i = min(1, min_{j : p_j >= p_i} (m / rank_j) * p_j), which equals
// stats::p.adjust's pmin(1, cummin(n / i * p[o])) including ties.
// Benjamini-Hochberg over the non-NaN p-values (m of them, stats::p.adjust's
// lazy n): sorted ascending, adj(k) = min over k' >= k of m / (k' + 1) p(k'),
// capped at 1 (R/test.R:133 -> stats::p.adjust "BH").  Ties need no care:
// the minimum over a tied group's positions is taken at its last one, which
// is the m / rank p of the step-up definition.  NaN p-values sort last (key
// 2) and keep NaN.
__global__ void k_bh_keys(const double* __restrict__ p, int64_t G, double* __restrict__ key,
                          int32_t* __restrict__ idx) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= G) return;
  const double v = p[i];
  key[i] = isnan(v) ? 2.0 : v;
  idx[i] = (int32_t)i;
}

__global__ void k_fill_i32_bh(int32_t* m) { *m = 0; }

// m = the number of sorted keys below 2, written by the thread at the last one
__global__ void k_bh_count(const double* __restrict__ ks, int64_t G, int32_t* __restrict__ m) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= G) return;
  if (ks[k] < 1.5 && (k + 1 == G || !(ks[k + 1] < 1.5))) *m = (int32_t)(k + 1);
}

// vrev[G - 1 - k] = m / (k + 1) p(k) for k < m, +inf beyond (so an inclusive
// min-scan of vrev is the suffix minimum)
__global__ void k_bh_terms(const double* __restrict__ ks, int64_t G, const int32_t* __restrict__ m,
                           double* __restrict__ vrev) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= G) return;
  const int32_t mm = *m;
  vrev[G - 1 - k] = k < mm ? (double)mm / (double)(k + 1) * ks[k] : __builtin_huge_val();
}

__global__ void k_bh_scatter(const double* __restrict__ p, const int32_t* __restrict__ is,
                             const double* __restrict__ scan, int64_t G,
                             const int32_t* __restrict__ m, double* __restrict__ adj) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= G) return;
  const int32_t i = is[k];
  adj[i] = k < *m ? fmin(scan[G - 1 - k], 1.0) : p[i];
}

namespace {
struct MinOp {
  __device__ __forceinline__ double operator()(double a, double b) const { return fmin(a, b); }
};
}  // namespace

size_t bh_scratch_bytes(int64_t G) {
  if (G <= 0) return 0;
  size_t sort_b = 0, scan_b = 0;
  (void)hipcub::DeviceRadixSort::SortPairs(nullptr, sort_b, (const double*)nullptr,
                                           (double*)nullptr, (const int32_t*)nullptr,
                                           (int32_t*)nullptr, (int)G);
  (void)hipcub::DeviceScan::InclusiveScan(nullptr, scan_b, (const double*)nullptr,
                                          (double*)nullptr, MinOp(), (int)G);
  const size_t tmp = std::max(sort_b, scan_b);
  // the blocks launch_bh takes, each rounded to 256 bytes: keys, sorted keys,
  // vrev, scan (doubles); idx, sorted idx (int32); m; the library's scratch
  auto r256 = [](size_t v) { return ((v + 255) / 256) * 256; };
  return 4 * r256(8 * (size_t)G) + 2 * r256(4 * (size_t)G) + r256(4) + r256(tmp);
}

hipError_t launch_bh(hipStream_t s, const double* pval, int64_t G, double* adj, void* scratch,
                     size_t scratch_bytes) {
  if (G <= 0) return hipSuccess;
  char* b = static_cast<char*>(scratch);
  auto take = [&](size_t bytes) {
    char* q = b;
    b += ((bytes + 255) / 256) * 256;
    return q;
  };
  double* key = reinterpret_cast<double*>(take(8 * G));
  double* ks = reinterpret_cast<double*>(take(8 * G));
  double* vrev = reinterpret_cast<double*>(take(8 * G));
  double* scan = reinterpret_cast<double*>(take(8 * G));
  int32_t* idx = reinterpret_cast<int32_t*>(take(4 * G));
  int32_t* is = reinterpret_cast<int32_t*>(take(4 * G));
  int32_t* m = reinterpret_cast<int32_t*>(take(4));
  void* tmp = take(0);
  const size_t used = (size_t)(b - static_cast<char*>(scratch));
  if (used > scratch_bytes) return hipErrorInvalidValue;
  size_t tmp_bytes = scratch_bytes - used;
  const dim3 grid((unsigned)((G + 255) / 256)), blk(256);
  hipLaunchKernelGGL(k_bh_keys, grid, blk, 0, s, pval, G, key, idx);
  hipError_t e = hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, key, ks, idx, is, (int)G, 0,
                                                    64, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_fill_i32_bh, dim3(1), dim3(1), 0, s, m);
  hipLaunchKernelGGL(k_bh_count, grid, blk, 0, s, ks, G, m);
  hipLaunchKernelGGL(k_bh_terms, grid, blk, 0, s, ks, G, m, vrev);
  tmp_bytes = scratch_bytes - used;
  e = hipcub::DeviceScan::InclusiveScan(tmp, tmp_bytes, vrev, scan, MinOp(), (int)G, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_bh_scatter, grid, blk, 0, s, pval, is, scan, G, m, adj);
  return hipGetLastError();
}

// stats::p.adjust for the other methods test_de can pass (R/test.R:127).
// The p-values are sorted ascending once (NaN last, as in launch_bh); with m
// the number of non-NaN ones and k the 0-based ascending rank, in R's order
// of operations:
//   holm        (m - k) p(k), running maximum from k = 0, capped at 1
//   hochberg    (m - k) p(k), minimum over k' >= k, capped at 1
//   BY          q m / (k + 1) p(k), minimum over k' >= k, capped at 1
//   bonferroni  m p, capped at 1
//   hommel      pa(k) = max(min_k' m p(k') / (k' + 1),
//                           max_{j=2..m-1} min(j p(min(k, m - j)), q1(j))),
//               q1(j) = min_{i=2..j} j p(m - j + i - 1) / i, then max(pa, p)
// (the loop of the R function unrolled: its q is rewritten in full for each
// j, so pa is the maximum of the per-j vectors).
__global__ void k_padj_terms(const double* __restrict__ ks, int m, int method, double q,
                             double* __restrict__ t) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m) return;
  const double pk = ks[k];
  double v;
  if (method == kPadjHolm) {
    t[k] = (double)(m - k) * pk;  // forward order, max-scanned
    return;
  }
  if (method == kPadjHochberg) v = (double)(m - k) * pk;
  else v = q * (double)m / (double)(k + 1) * pk;  // BY (q = 1 for BH)
  t[m - 1 - k] = v;  // reversed, so a min-scan is the suffix minimum
}

__global__ void k_padj_scatter(const double* __restrict__ p, const double* __restrict__ ks,
                               const int32_t* __restrict__ is, const double* __restrict__ scan,
                               int64_t G, int m, int method, double* __restrict__ adj) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= G) return;
  const int32_t i = is[k];
  if (k >= m || m <= 1 || method == kPadjNone) {
    adj[i] = p[i];
    return;
  }
  double v;
  if (method == kPadjBonferroni) v = (double)m * ks[k];
  else if (method == kPadjHolm) v = scan[k];
  else v = scan[m - 1 - k];
  adj[i] = fmin(v, 1.0);
}

// hommel: min_k m p(k) / (k + 1) by one workgroup, into *pa0
__global__ __launch_bounds__(1024) void k_hommel_pa0(const double* __restrict__ ks, int m,
                                                     double* __restrict__ pa0) {
  __shared__ double red[16];
  double v = __builtin_huge_val();
  for (int k = threadIdx.x; k < m; k += 1024) v = fmin(v, (double)m * ks[k] / (double)(k + 1));
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < 16; ++w) v = fmin(v, red[w]);
    *pa0 = v;
  }
}

// hommel: q1[j] for j = 2 .. m - 1
__global__ void k_hommel_q1(const double* __restrict__ ks, int m, double* __restrict__ q1) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x + 2;
  if (j >= m) return;
  const double dj = (double)j;
  double v = __builtin_huge_val();
  for (int i = 2; i <= j; ++i) v = fmin(v, dj * ks[m - j + i - 1] / (double)i);
  q1[j] = v;
}

// hommel: adjusted p of ascending rank k, scattered to the gene
__global__ void k_hommel_pa(const double* __restrict__ ks, const int32_t* __restrict__ is, int m,
                            const double* __restrict__ q1, const double* __restrict__ pa0,
                            double* __restrict__ adj) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m) return;
  double pa = *pa0;
  for (int j = 2; j < m; ++j) pa = fmax(pa, fmin((double)j * ks[min(k, m - j)], q1[j]));
  adj[is[k]] = fmax(pa, ks[k]);
}

namespace {
struct MaxOp {
  __device__ __forceinline__ double operator()(double a, double b) const { return fmax(a, b); }
};
}  // namespace

hipError_t launch_padj_sort(hipStream_t s, const double* pval, int64_t G, void* scratch,
                            size_t scratch_bytes, PadjBufs& b) {
  char* c = static_cast<char*>(scratch);
  auto take = [&](size_t bytes) {
    char* q = c;
    c += ((bytes + 255) / 256) * 256;
    return q;
  };
  double* key = reinterpret_cast<double*>(take(8 * G));
  b.ks = reinterpret_cast<double*>(take(8 * G));
  b.t = reinterpret_cast<double*>(take(8 * G));
  b.scan = reinterpret_cast<double*>(take(8 * G));
  int32_t* idx = reinterpret_cast<int32_t*>(take(4 * G));
  b.is = reinterpret_cast<int32_t*>(take(4 * G));
  b.m = reinterpret_cast<int32_t*>(take(4));
  b.tmp = take(0);
  const size_t used = (size_t)(c - static_cast<char*>(scratch));
  if (used > scratch_bytes) return hipErrorInvalidValue;
  b.tmp_bytes = scratch_bytes - used;
  const dim3 grid((unsigned)((G + 255) / 256)), blk(256);
  hipLaunchKernelGGL(k_bh_keys, grid, blk, 0, s, pval, G, key, idx);
  size_t tb = b.tmp_bytes;
  hipError_t e = hipcub::DeviceRadixSort::SortPairs(b.tmp, tb, key, b.ks, idx, b.is, (int)G, 0,
                                                    64, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_fill_i32_bh, dim3(1), dim3(1), 0, s, b.m);
  hipLaunchKernelGGL(k_bh_count, grid, blk, 0, s, b.ks, G, b.m);
  return hipGetLastError();
}

hipError_t launch_padj_finish(hipStream_t s, const double* pval, int64_t G, int m, int method,
                              double q, const PadjBufs& b, double* adj) {
  const dim3 grid((unsigned)((G + 255) / 256)), blk(256);
  if (method == kPadjHommel && m > 2) {
    const dim3 gm((unsigned)((m + 255) / 256));
    hipLaunchKernelGGL(k_hommel_pa0, dim3(1), dim3(1024), 0, s, b.ks, m, b.scan);
    hipLaunchKernelGGL(k_hommel_q1, gm, blk, 0, s, b.ks, m, b.t);
    hipLaunchKernelGGL(k_padj_scatter, grid, blk, 0, s, pval, b.ks, b.is, b.scan, G, 0,
                       kPadjNone, adj);  // NaN entries (and all, overwritten below)
    hipLaunchKernelGGL(k_hommel_pa, gm, blk, 0, s, b.ks, b.is, m, b.t, b.scan, adj);
    return hipGetLastError();
  }
  if (method == kPadjHommel) method = kPadjHochberg;  // m == 2, as p.adjust
  if (m > 1 && method != kPadjNone && method != kPadjBonferroni) {
    const dim3 gm((unsigned)((m + 255) / 256));
    hipLaunchKernelGGL(k_padj_terms, gm, blk, 0, s, b.ks, m, method, q, b.t);
    size_t tb = b.tmp_bytes;
    hipError_t e = method == kPadjHolm
                       ? hipcub::DeviceScan::InclusiveScan(b.tmp, tb, b.t, b.scan, MaxOp(), m, s)
                       : hipcub::DeviceScan::InclusiveScan(b.tmp, tb, b.t, b.scan, MinOp(), m, s);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_padj_scatter, grid, blk, 0, s, pval, b.ks, b.is, b.scan, G, m, method,
                     adj);
  return hipGetLastError();
}

// dst[i] = src[i], i < n: device counts into pinned host slots (a kernel in
// the stream rather than a copy-engine transfer, so the next round's kernels
// need not wait for a copy engine)
__global__ void k_copy_counts(const int32_t* __restrict__ src, int n, int32_t* __restrict__ dst) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
}

hipError_t launch_copy_counts(hipStream_t s, const int32_t* src, int n, int32_t* dst) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_copy_counts, dim3(1), dim3(n < 256 ? 64 : 256), 0, s, src, n, dst);
  return hipGetLastError();
}

// G x p column-major (R layout) <-> gene-major G x p (row g holds its p values)
__global__ void k_transpose_gp(const double* __restrict__ src, int64_t G, int p, int to_gene,
                               double* __restrict__ dst) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= G * p) return;
  const int64_t g = t / p;
  const int j = (int)(t - g * p);
  if (to_gene) dst[g * p + j] = src[g + j * G];
  else dst[g + j * G] = src[g * p + j];
}

__global__ void k_fill_i32(int32_t* p, int64_t count, int32_t v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = v;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------


hipError_t launch_k_from_alpha(hipStream_t s, const double* alpha0, int64_t G, double* k) {
  if (G <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_k_from_alpha, dim3((unsigned)((G + 255) / 256)), dim3(256), 0, s,
                     alpha0, G, k);
  return hipGetLastError();
}





hipError_t launch_test_de(hipStream_t s, const double* beta, const double* bread,
                          const double* sandwich, int64_t G, int p, double df,
                          const double* contrast, double max_lfc, double* pval,
                          double* logp, double* lfc, void* bh_scratch, size_t bh_bytes,
                          double* adj) {
  if (G <= 0) return hipSuccess;
  const dim3 g1((unsigned)((G + 255) / 256));
  hipLaunchKernelGGL(k_test_de, g1, dim3(256), 0, s, beta, bread, sandwich, G, p, df,
                     contrast, max_lfc, pval, logp, lfc);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return launch_bh(s, pval, G, adj, bh_scratch, bh_bytes);
}

hipError_t launch_transpose_gp(hipStream_t s, const double* src, int64_t G, int p,
                               int to_gene_major, double* dst) {
  if (G <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_transpose_gp, dim3((unsigned)((G * p + 255) / 256)), dim3(256), 0, s,
                     src, G, p, to_gene_major, dst);
  return hipGetLastError();
}

hipError_t launch_fill_i32(hipStream_t s, int32_t* p, int64_t count, int32_t v) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, p,
                     count, v);
  return hipGetLastError();
}

}  // namespace dv
