// Internal launcher interface between the C ABI (devil_hip.cpp) and the
// gfx950 kernels.  All pointers are device pointers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dv {

struct Counts {
  const void* ptr;
  int dtype;   // 0 int32, 1 float32, 2 float64
  int64_t ld;  // elements between genes
};

// init modes for launch_init (beta part); -1 = no beta0
enum { kInitNone = -1, kInitQR = 0, kInitRough = 1, kInitUnivariate = 2 };

struct InitParams {
  int want_alpha;           // write alpha0
  double moi;               // 1 / mean(exp(offset))
  int init_mode;            // kInit*
  const double* eo;         // n: exp(offset)
  const double* Q;          // n x p column-major thin Q of X (QR mode)
  const double* R;          // p x p column-major upper R (QR mode)
  const double* X;          // n x p (univariate mode)
  const double* logsf;      // n: log(sf) (univariate mode)
  const double* xmean;      // p (univariate mode)
  const double* denom;      // p (univariate mode)
  int intercept_first;      // univariate mode
  double sum_sf;            // univariate mode
};

hipError_t launch_init(hipStream_t s, Counts c, int64_t G, int64_t n, int p,
                       const InitParams& ip, double* alpha0, double* beta0);

// k: per-gene k as src/beta.cpp holds it (float(1/float(alpha0)), widened)
hipError_t launch_k_from_alpha(hipStream_t s, const double* alpha0, int64_t G, double* k);

hipError_t launch_irls_fp64(hipStream_t s, Counts c, int64_t G, int64_t n, int p,
                            const double* X, const double* off, const double* k,
                            double* beta, int max_iter, double eps, int32_t* iters,
                            int32_t* status);

hipError_t launch_mom(hipStream_t s, Counts c, int64_t G, int64_t n, int p,
                      const double* X, const double* beta, const double* eo,
                      double* theta, int32_t* status);

hipError_t launch_bread_meat(hipStream_t s, Counts c, int64_t G, int64_t n, int p,
                             const double* X, const double* beta, const double* theta,
                             const double* eo, const int32_t* ends, int C, double* bread,
                             double* sandwich, int32_t* status);

hipError_t launch_mle(hipStream_t s, Counts c, int64_t G, int64_t n, int p,
                      const double* X, const double* beta, const double* off,
                      int max_iter, double* theta, int32_t* iters, int32_t* status);

hipError_t launch_test_de(hipStream_t s, const double* beta, const double* bread,
                          const double* sandwich, int64_t G, int p, double df,
                          const double* contrast, double max_lfc, double* pval,
                          double* logp, double* lfc, int32_t* rank_scratch,
                          double* adj);

// G x p column-major <-> gene-major (to_gene_major = 1: col-major -> gene-major)
hipError_t launch_transpose_gp(hipStream_t s, const double* src, int64_t G, int p,
                               int to_gene_major, double* dst);

hipError_t launch_fill_i32(hipStream_t s, int32_t* p, int64_t count, int32_t v);

}  // namespace dv
