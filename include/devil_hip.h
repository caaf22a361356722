/*
 * devil_hip.h -- C ABI of the MI355X (gfx950) per-gene NB-GLM fit.
 *
 * This is the drop-in boundary for devil's fitting hot path.  Each entry
 * point replaces one reference interface (paths relative to the devil
 * 0.99.0 tree):
 *
 *   devil_fit                 gpu_fit / cpu_fit          R/main.R:277-371, 375-523
 *                             via Rcpp beta_fit_gpu      src/beta.cpp:82-87
 *                             -> beta_fit_gpu_external   cuda/src/batch.hpp:18-27
 *   devil_estimate_dispersion estimate_dispersion        R/overdispersion.R:129-135
 *   devil_init_beta           init_beta (QR)             R/beta.R:13-20
 *                             rough init                 R/main.R:404-406
 *                             univariate init            src/beta_init_univariate.cpp:13-95
 *   devil_beta_fit            beta_fit, over all genes   src/beta.cpp:25-51
 *   devil_mom_dispersion      estimate_mom_dispersion_cpp src/mom_overdispersion.cpp:6-56
 *   devil_fit_dispersion_mle  fit_dispersion             R/overdispersion.R:25-115
 *                             (score/dscore of           src/overdispersion.cpp:95-242)
 *   devil_bread_meat          compute_hessian            src/variance.cpp:22-47
 *                             compute_clustered_meat_fast src/variance.cpp:186-228
 *                             sandwich assembly          R/main.R:481-492
 *   devil_test_de             test_de                    R/test.R:85-134
 *   devil_sandwich_hc0        compute_sandwich (HC0)     R/variance.R:16-26,
 *                             compute_meat               src/variance.cpp:157-178
 *   devil_size_factors        calculate_sf               R/size_factors.R:37-127
 *   devil_group_cells         group_data (counts)        R/group_data.R:38-69
 *   devil_csc_to_dense        sparse assays of fit_devil R/main.R:1-9
 *
 * Conventions
 *   - Every function returns DEVIL_OK (0) or a DEVIL_E_* code; nothing calls
 *     exit() or abort().  devil_last_error(ctx) gives the message of the last
 *     failure on that context.
 *   - counts are gene-major: gene g's n counts start at element g*counts_ld.
 *   - X is n x p column-major (R's matrix layout, zero-copy from a REALSXP).
 *   - beta is G x p column-major (R layout).  Per-gene p x p matrices are
 *     column-major and contiguous: M[g*p*p + a + b*p].
 *   - devil_fit takes host or device arrays (devil_problem.mem); the stage
 *     entry points take device pointers only.
 *   - A context is not thread-safe; distinct contexts may be used from
 *     distinct host threads.  Calls are synchronous: outputs are complete on
 *     return.
 */
#ifndef DEVIL_HIP_H
#define DEVIL_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* return codes */
#define DEVIL_OK 0
#define DEVIL_E_ARG 1          /* invalid argument (pointer, size, mode, p out of range) */
#define DEVIL_E_HIP 2          /* HIP runtime error */
#define DEVIL_E_NOMEM 3        /* device or host allocation failed */
#define DEVIL_E_RANK 4         /* design matrix rank deficient (QR init) */
#define DEVIL_E_CLUSTERS 5     /* cluster ends not increasing or last end != n */
#define DEVIL_E_NODEVICE 6     /* no HIP device / device index out of range */
#define DEVIL_E_UNSUPPORTED 7  /* valid request this library does not implement */

/* per-gene status bits (devil_outputs.status) */
#define DEVIL_GENE_MAXITER 1   /* beta loop stopped at max_iter (src/beta.cpp:37) */
#define DEVIL_GENE_NAN 2       /* beta loop stopped on a NaN step (src/beta.cpp:45) */
#define DEVIL_GENE_SINGULAR 4  /* a p x p factorisation met an exactly zero pivot */
#define DEVIL_GENE_MOM_DEN0 8  /* MOM denominator <= 0, theta = 0 */
#define DEVIL_GENE_MLE_ZERO 16 /* MLE special case returned 0 (all-zero gene or score(1e-8) < 0) */
#define DEVIL_GENE_MLE_NOCONV 32 /* the nlminb driver ended with convergence != 0 (singular or
                                    false convergence, evaluation or iteration limit): the
                                    reference reruns nlminb without the Hessian
                                    (R/overdispersion.R:75-110), which this library does not
                                    restate; theta is the first run's par */

#define DEVIL_GENE_OD_FAILED 64 /* "new" overdispersion: the reference stops with an error here
                                     (C1 <= n, a Laplace mode or quadrature failure, f not finite
                                     at the start); theta = NaN */

#define DEVIL_PMAX 10          /* largest design width this library supports */

enum { DEVIL_MEM_HOST = 0, DEVIL_MEM_DEVICE = 1 };
enum { DEVIL_COUNTS_I32 = 0, DEVIL_COUNTS_F32 = 1, DEVIL_COUNTS_F64 = 2 };
enum { DEVIL_INIT_QR = 0, DEVIL_INIT_ROUGH = 1, DEVIL_INIT_UNIVARIATE = 2 };
enum { DEVIL_OD_NONE = 0, DEVIL_OD_MOM = 1, DEVIL_OD_MLE = 2,
       DEVIL_OD_NEW = 3 /* "new" / "I": fit_overdispersion_cppp, src/new_overdispersion.cpp */ };
enum { DEVIL_PREC_FP64 = 0, DEVIL_PREC_MIXED = 1 };

typedef struct devil_ctx devil_ctx;

/* One context per host thread: owns a HIP stream and device workspaces. */
int devil_create(devil_ctx **ctx, int hip_device);
void devil_destroy(devil_ctx *ctx);
const char *devil_strerror(int code);
const char *devil_last_error(const devil_ctx *ctx);
/* Library version string, e.g. "devil_hip 0.1.0 gfx950". */
const char *devil_version(void);

/* ------------------------------------------------------------------------ */
/* Whole fit: R/main.R:375-523 cpu_fit semantics (the numerical contract),   */
/* called where gpu_fit (R/main.R:277-371) calls beta_fit_gpu.              */
/* ------------------------------------------------------------------------ */
typedef struct {
  int64_t G, n;
  int32_t p;
  int32_t mem;                  /* DEVIL_MEM_* : where every array below lives */
  const void *counts;           /* G x n, gene-major */
  int32_t counts_dtype;         /* DEVIL_COUNTS_* */
  int64_t counts_ld;            /* elements between consecutive genes (>= n) */
  const double *X;              /* n x p column-major */
  const double *offset;         /* n: off + log(size_factors) (R/size_factors.R:172-184) */
  const int32_t *cluster_ends;  /* cumulative 1-based ends (R/main.R:193), or NULL */
  int32_t n_clusters;           /* 0 = no clustered sandwich */
  const double *disp_init;      /* G (init_overdispersion), or NULL -> estimate_dispersion */
  const double *beta_init;      /* G x p column-major, or NULL -> init_mode */
  int32_t init_mode;            /* DEVIL_INIT_* (R/main.R:404-409) */
  int32_t od_mode;              /* DEVIL_OD_* (R/main.R:431-468) */
  int32_t max_iter;             /* R/main.R:175 */
  double tolerance;             /* R/main.R:176; used as a float, src/beta.cpp:25 */
  int32_t precision;            /* DEVIL_PREC_* */
} devil_problem;

typedef struct {                /* caller-allocated, same memory space as the problem */
  double *beta;                 /* G x p column-major */
  double *theta;                /* G: overdispersion (MOM / MLE / 0) */
  int32_t *beta_iters;          /* G */
  int32_t *theta_iters;         /* G: MLE passes; kappa_iters for "new"; 0 for MOM / none */
  double *bread;                /* G x p x p: beta_sandwiches_null */
  double *sandwich;             /* G x p x p, or NULL: beta_sandwiches (clusters only) */
  int32_t *status;              /* G, or NULL: DEVIL_GENE_* bits */
} devil_outputs;

/* Fits n_problems independent problems (probs[j] -> outs[j]); problems may
 * differ in G, n, p, design, clusters and mode.  With more than one problem
 * they run concurrently on DEVIL_TUNE_CONCURRENCY streams.  Returns the first
 * failing problem's code (devil_last_error names the problem). */
int devil_fit(devil_ctx *ctx, const devil_problem *probs, devil_outputs *outs,
              int32_t n_problems);

/* ------------------------------------------------------------------------ */
/* Stage entry points (device pointers; counts as in devil_problem).         */
/* ------------------------------------------------------------------------ */

/* Preparation of the count matrix (device pointers). */

/* calculate_sf (R/size_factors.R:37-127): sf[i] for each cell i.
 * NORMED_SUM: column sums over the genes divided by their geometric mean
 * (DEVIL_E_ARG with the reference's message when a column sums to 0);
 * PSINORM: mean_g log1p(y) - min_g log1p(y), divided by its geometric mean;
 * EDGER: DEVIL_E_UNSUPPORTED (the reference calls edgeR's calcNormFactors).
 * G <= 1 gives unit factors, as the reference. */
enum { DEVIL_SF_NORMED_SUM = 0, DEVIL_SF_PSINORM = 1, DEVIL_SF_EDGER = 2 };
int devil_size_factors(devil_ctx *ctx, const void *counts, int32_t counts_dtype, int64_t G,
                       int64_t n, int64_t counts_ld, int32_t method, double *sf);

/* group_data (R/group_data.R:38-69), the count part: out[g, i] = counts[g,
 * order[i]] for the 0-based cell order (R's order() of the cluster codes);
 * out has the counts' dtype and out_ld elements per gene.  An order entry
 * outside [0, n) returns DEVIL_E_ARG before anything is read. */
int devil_group_cells(devil_ctx *ctx, const void *counts, int32_t counts_dtype, int64_t G,
                      int64_t n, int64_t counts_ld, const int64_t *order, void *out,
                      int64_t out_ld);

/* A dgCMatrix / CSC G x n count matrix (R/main.R:1-9 accepts sparse assays:
 * column = cell; colptr n + 1 entries from 0, rowind 0-based gene indices,
 * values as R stores them) into the dense gene-major matrix the fit streams
 * (out_dtype DEVIL_COUNTS_*; values converted by a C cast).  The structure is
 * checked on the device first (DEVIL_E_ARG, nothing written). */
int devil_csc_to_dense(devil_ctx *ctx, const int64_t *colptr, const int32_t *rowind,
                       const double *values, int64_t G, int64_t n, int32_t out_dtype,
                       void *out, int64_t out_ld);

/* R/overdispersion.R:129-135: alpha0[g] = (var_g - ybar_g/mean(exp(off))) / ybar_g^2,
 * var with n-1; NaN or negative -> 0.01. */
int devil_estimate_dispersion(devil_ctx *ctx, const void *counts, int32_t counts_dtype,
                              int64_t G, int64_t n, int64_t counts_ld,
                              const double *offset, double *alpha0);

/* R/beta.R:13-20 (mode QR), R/main.R:404-406 (ROUGH),
 * src/beta_init_univariate.cpp:13-95 (UNIVARIATE, with sf = exp(offset)). */
int devil_init_beta(devil_ctx *ctx, const void *counts, int32_t counts_dtype, int64_t G,
                    int64_t n, int64_t counts_ld, const double *X, int32_t p,
                    const double *offset, int32_t init_mode, double *beta0);

/* src/beta.cpp:25-51 for every gene: beta (G x p col-major) is updated in place;
 * alpha0 holds the per-gene dispersion the reference passes as `float k`. */
int devil_beta_fit(devil_ctx *ctx, const void *counts, int32_t counts_dtype, int64_t G,
                   int64_t n, int64_t counts_ld, const double *X, int32_t p,
                   double *beta, const double *offset, const double *alpha0,
                   int32_t max_iter, double tolerance, int32_t precision,
                   int32_t *iters, int32_t *status);

/* src/mom_overdispersion.cpp:6-56 with sf = exp(offset). */
int devil_mom_dispersion(devil_ctx *ctx, const void *counts, int32_t counts_dtype,
                         int64_t G, int64_t n, int64_t counts_ld, const double *X,
                         int32_t p, const double *beta, const double *offset,
                         double *theta);

/* R/overdispersion.R:25-115 with Cox-Reid adjustment: special cases and start
 * value as the reference, then the context's MLE driver (DEVIL_TUNE_MLE_DRIVER):
 * by default stats::nlminb(rel.tol = tolerance, iter.max = max_iter)'s PORT
 * iteration restated (its stopping rules included), or the root of the score.
 * iters: passes over the counts per gene. */
int devil_fit_dispersion_mle(devil_ctx *ctx, const void *counts, int32_t counts_dtype,
                             int64_t G, int64_t n, int64_t counts_ld, const double *X,
                             int32_t p, const double *beta, const double *offset,
                             int32_t max_iter, double tolerance, double *theta,
                             int32_t *iters);

/* src/variance.cpp:22-47 (bread) and :186-228 (meat) in one pass; sandwich =
 * bread * meat * bread * n (R/main.R:491) when cluster_ends != NULL. */
int devil_bread_meat(devil_ctx *ctx, const void *counts, int32_t counts_dtype, int64_t G,
                     int64_t n, int64_t counts_ld, const double *X, int32_t p,
                     const double *beta, const double *theta, const double *offset,
                     const int32_t *cluster_ends, int32_t n_clusters, double *bread,
                     double *sandwich, int32_t *status);

/* The unclustered sandwich of compute_sandwich (R/variance.R:16-26) with the
 * HC0 meat of compute_meat / compute_scores (src/variance.cpp:64-85, 157-178):
 * meat = (1/n) sum_i s_i s_i', s_i = x_i (y_i - mu_i) / (1 + theta mu_i);
 * sandwich = bread meat bread n.  The bread is compute_hessian with theta as
 * devil_bread_meat (and cpu_fit) use it; the reference's compute_sandwich
 * passes 1/overdispersion there, which its callers do not rely on. */
int devil_sandwich_hc0(devil_ctx *ctx, const void *counts, int32_t counts_dtype, int64_t G,
                       int64_t n, int64_t counts_ld, const double *X, int32_t p,
                       const double *beta, const double *theta, const double *offset,
                       double *bread, double *sandwich, int32_t *status);

/* R/test.R:85-134: p = 2*pt(|c'b| / sqrt(c'Vc), n-2, upper); clustered:
 * max(p_sandwich, p_bread); BH over non-NaN p; lfc = c'b/log(2) clipped to
 * +-max_lfc.  sandwich may be NULL.  log_pval (natural log, exact where p
 * underflows) may be NULL. */
int devil_test_de(devil_ctx *ctx, const double *beta, const double *bread,
                  const double *sandwich, int64_t G, int32_t p, int64_t n,
                  const double *contrast, double max_lfc, double *pval,
                  double *log_pval, double *adj_pval, double *lfc);

/* stats::p.adjust(pval, method) on the device, for the pval_adjust_method
 * values test_de passes through (R/test.R:85,127; stats' own p.adjust.R).
 * NaN p-values stay NaN and the adjustment runs over the others (n = their
 * count); adj may alias pval.  devil_test_de itself applies "BH" (its
 * default). */
enum { DEVIL_PADJ_BH = 0, DEVIL_PADJ_BY = 1, DEVIL_PADJ_HOLM = 2, DEVIL_PADJ_HOCHBERG = 3,
       DEVIL_PADJ_HOMMEL = 4, DEVIL_PADJ_BONFERRONI = 5, DEVIL_PADJ_NONE = 6 };
int devil_p_adjust(devil_ctx *ctx, const double *pval, int64_t G, int32_t method,
                   double *adj);

/* devil_test_de without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Arguments and results are devil_test_de's.  Its scratch (the gene-major
 * beta, 8 G p bytes, and the Benjamini-Hochberg sort buffers, about 40 G
 * bytes plus the sort's temporary storage) lives in a fixed 16 MiB device
 * array per GPU, held until the process ends: G up to about 100000 genes at
 * p = 10; larger G returns DEVIL_E_ARG.  Every call on a GPU shares that
 * GPU's array, so calls that overlap on one GPU (several contexts or host
 * threads) race. */
int devil_test_de_noalloc(devil_ctx *ctx, const double *beta, const double *bread,
                          const double *sandwich, int64_t G, int32_t p, int64_t n,
                          const double *contrast, double max_lfc, double *pval,
                          double *log_pval, double *adj_pval, double *lfc);

/* ------------------------------------------------------------------------ */
/* Instrumentation (no counterpart in the reference).                        */
/* ------------------------------------------------------------------------ */
enum { DEVIL_STAGE_INIT = 0, DEVIL_STAGE_IRLS = 1, DEVIL_STAGE_DISPERSION = 2,
       DEVIL_STAGE_BREAD_MEAT = 3, DEVIL_STAGE_TEST_DE = 4, DEVIL_N_STAGES = 5 };

/* Enable (1) or disable (0) per-stage timing with HIP events recorded on the
 * context's stream around each stage's kernels. */
int devil_set_timing(devil_ctx *ctx, int enable);
/* Device milliseconds per stage of the last devil_fit (stages 0-3) and the
 * last devil_test_de (stage 4); ms must hold DEVIL_N_STAGES doubles. */
int devil_stage_times(const devil_ctx *ctx, double *ms);

/* Kernel choices of the mixed-precision path (set per context; the library
 * defaults are what the autotuner found best on gfx950). */
enum { DEVIL_TUNE_MIXED_KERNEL = 0,  /* DEVIL_MIXED_* (default AUTO) */
       DEVIL_TUNE_CHUNK_CELLS = 1,   /* cells per chunk of the tile passes, 0 = autotune */
       DEVIL_TUNE_CONCURRENCY = 2,   /* streams for multi-problem fits (default 8) */
       DEVIL_TUNE_MLE_DRIVER = 3,    /* DEVIL_MLE_* (default NLMINB) */
       DEVIL_TUNE_PAD_COPY = 4,      /* 1 (default): a mixed fit of device counts whose rows
                                        are not 16-byte aligned reads a copy with padded rows,
                                        kept by the context for the next fit (freed by
                                        devil_destroy or by setting 0); 0: never */
       DEVIL_TUNE_JOINT = 5,         /* 0 (default): devil_fit with several problems runs
                                        each problem's IRLS rounds on its own (the problems
                                        overlap on the streams); 1: jointly, in units of
                                        problems of one design width (about n_problems / 2
                                        streams problems each), one accumulate launch per
                                        kernel and one update launch per round and unit;
                                        N > 1: units of at most N problems */
       DEVIL_TUNE_GRAPH = 6,         /* 1: the IRLS rounds whose active genes fit one
                                        column of workgroups (16 x tiles per workgroup) run
                                        as replays of a captured two-round hipGraph
                                        (accumulate, clear, update, count copy, twice);
                                        0 (default): launched one by one, queued ahead of the
                                        count reads (the same speed; per-kernel timing) */
       DEVIL_TUNE_BATCH_GENES = 7,   /* host-resident counts are uploaded and fitted in gene
                                        batches sized by free device memory (up to 8 GiB);
                                        N > 0 caps a batch at N genes, 0 (default): no cap */
       DEVIL_N_TUNE = 8 };
/* MLE overdispersion drivers.  NLMINB: the reference's stats::nlminb call
 * (R/overdispersion.R:53-73) restated for one parameter -- its trust-region
 * Newton steps and its stopping rules (relative function convergence at
 * rel.tol = tolerance), so theta is where nlminb stops, not the root of the
 * score.  ROOT: safeguarded Newton on score(log theta) = 0 to |dx| < 1e-9. */
enum { DEVIL_MLE_NLMINB = 0, DEVIL_MLE_ROOT = 1 };
enum { DEVIL_MIXED_F32 = 0,   /* eta and all contractions on v_mfma_f32_16x16x4f32 */
       DEVIL_MIXED_BF16 = 1,  /* eta, design columns (six products of three-piece
                                 splits) and pair columns (three products) on
                                 v_mfma_f32_16x16x32_bf16 */
       DEVIL_MIXED_AUTO = 2,  /* autotuner: times the candidates on the problem's
                                 first genes once per (p, n, count type) and keeps
                                 the fastest for the context */
       DEVIL_MIXED_BF16_WS = 3, /* DEVIL_MIXED_BF16 arithmetic on wave-specialised
                                  workgroups: a loader wavefront stages the design
                                  tables in LDS, the compute wavefronts stage their
                                  counts in LDS three steps ahead (direct-to-LDS
                                  loads); 4-byte counts in 16-byte aligned rows,
                                  else DEVIL_MIXED_BF16 runs */
       DEVIL_MIXED_BF16_WS2 = 4 /* DEVIL_MIXED_BF16_WS with two 16-gene tiles per
                                   compute wavefront (half the wavefronts; same
                                   results bit for bit); not among the autotuner's
                                   candidates */ };
int devil_set_tuning(devil_ctx *ctx, int32_t key, int64_t value);

/* Per-launch timing of the count-streaming kernels (the IRLS accumulate
 * passes and the MLE score pass), kept while timing is enabled and reset by
 * each devil_fit. */
enum { DEVIL_KERNEL_IRLS_FP64 = 0, DEVIL_KERNEL_IRLS_MIXED = 1,
       DEVIL_KERNEL_MLE = 2,  /* the MLE score pass (k_mle_accum) */
       DEVIL_N_KERNELS = 3 };
/* launches: number of launches; ms: summed device milliseconds (HIP events
 * on the context's stream around each launch); bytes: summed algorithmic
 * bytes (count element size x n x active genes of each launch). */
int devil_kernel_stats(const devil_ctx *ctx, int32_t kernel, int64_t *launches, double *ms,
                       double *bytes);

/* Facts about the last devil_fit: genes the mixed path refit in FP64 (their
 * stopping step came within 1e-3 eps of the threshold, or was NaN), and the
 * mixed kernel (DEVIL_MIXED_*, -1 for FP64) and chunk size it ran with. */
enum { DEVIL_INFO_REPLAYED = 0, DEVIL_INFO_MIXED_KERNEL = 1, DEVIL_INFO_CHUNK_CELLS = 2,
       DEVIL_INFO_POLISHED = 3, /* genes whose last mixed IRLS step was redone in FP64 */
       DEVIL_INFO_ALIGNED_COPY = 4, /* 1 if the mixed fit read a copy of device counts whose
                                       rows it padded to 16-byte alignment */
       DEVIL_INFO_JOINED = 5 /* problems of the last multi-problem fit whose IRLS ran in the
                                joint rounds */ };
int devil_fit_info(const devil_ctx *ctx, int32_t key, int64_t *value);

#ifdef __cplusplus
}
#endif
#endif
