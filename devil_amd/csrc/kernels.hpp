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

// init modes of the init pass (beta part); -1 = no beta0
enum { kInitNone = -1, kInitQR = 0, kInitRough = 1, kInitUnivariate = 2 };

// Design laid out for the MFMA IRLS kernels (kernels_tile.hip).
#define DEVIL_NCP_MAX 80  // 16 * ceil((10*11/2 + 10) / 16)
// Entries per design column of the bf16 eta operands: x and log2(e) beta
// each in three bf16 pieces (each rounded from what the earlier leave), and
// the six products of total order <= 2: (x1 b1, x1 b2, x1 b3, x2 b1, x2 b2,
// x3 b1).  What is dropped and the split residues are below 2^-25 |x beta|.
constexpr int kEtaTerms = 6;
__host__ __device__ constexpr int eta_blocks(int p) { return (kEtaTerms * p + 31) / 32; }
__host__ __device__ constexpr int eta_x_piece(int t) { return t < 3 ? 0 : (t < 5 ? 1 : 2); }
__host__ __device__ constexpr int eta_b_piece(int t) { return t < 3 ? t : (t < 5 ? t - 3 : 0); }

struct TileDesign {
  int64_t n;             // cells
  int64_t nblk;          // 16-cell blocks (n padded to a multiple of 16)
  int64_t chunk_blocks;  // blocks per chunk
  int nchunk;            // chunks; partial sums are kept per chunk
  int nt;                // 16-column tiles of K
  int ncp;               // 16 * nt
  const double* Xp;      // [blk][j][lane] = X[16 blk + pi(lane%16)][4 j + lane/16] (0-padded),
                         // pi(m) = 4 (m % 4) + m / 4
  const double* offp;    // off padded with zeros to 16 nblk
  const double* Kp;      // [blk][t][lane][r] = K[16 blk + 4 (lane/16) + r][16 t + lane%16]
  // the same operands in FP32 for the mixed-precision IRLS (null unless built);
  // rows of the eta product follow the f32 MFMA output layout (identity order)
  const float* Xf;       // [blk][j][lane] = X[16 blk + lane%16][4 j + lane/16]
  const float* offf;     // off padded with zeros to 16 nblk
  const float* Kf;       // [blk][t][lane][r] = K[16 blk + 4 (lane/16) + r][16 t + lane%16]
  // bf16 mixed design: blocks are 32-cell steps, K column tile 0 holds the p
  // design columns (FP32 products), tiles 1.. the pair columns (bf16 products);
  // c(h, i) = i < 4 ? 4 h + i : 16 + 4 h + i - 4 is the cell of k index (h, i)
  const uint16_t* xab;   // [st][u][j][lane][8] bf16 eta entries 32 j + 8 (lane/16) + i of
                         // cell 32 st + 16 u + lane%16: per feature f, entries 6 f + t
                         // hold (x1, x1, x1, x2, x2, x3), x = x1 + x2 + x3 (eta_pieces)
  const float* offl;     // off * log2(e), padded with zeros to 32 nblk
  const uint16_t* kg;    // [st][q][lane][8] bf16 piece q of X[32 st + c(lane/16, i)][lane%16]:
                         // x = x1 + x2 + x3, each piece rounded from what the earlier leave
  const uint16_t* kh;    // [st][t][lane][8] bf16 K1 of K[32 st + c(lane/16, i)][16 (t+1) + lane%16]
  const uint16_t* kh2;   // the same layout, bf16 K2 = K - K1
  const double* eop;     // exp(off) padded with zeros to 16 nblk (FP64 design)
  int kidx_pair[55];     // K column of x_a x_b (packed upper triangle), -1 = zero column
  int kidx_lin[10];      // K column of x_a, -1 = zero column
  double ksum[DEVIL_NCP_MAX];  // sum over cells of each K column
};

// Column of K: a < 0 zero, b < 0 the design column x_a, else x_a x_b.
struct ColSpec {
  int a, b;
};
hipError_t launch_build_tables(hipStream_t s, const double* X, const double* off,
                               const double* eo, int64_t n, int p, const ColSpec* spec, int ncol,
                               int nt, int64_t nblk, double* Xp, double* offp, double* eop,
                               double* Kp, int f32, float* Xf, float* offf, float* Kf);
// Design statistics on the device: partials per kStatBlocks cell ranges.
constexpr int kStatBlocksHost = 64;
hipError_t launch_col_stats(hipStream_t s, const double* X, int64_t n, const ColSpec* cand,
                            int C, uint64_t* ph, uint64_t* pnz, double* ps);
hipError_t launch_col_mismatch(hipStream_t s, const double* X, int64_t n, ColSpec c1, ColSpec c2,
                               uint64_t* out);
hipError_t launch_gram(hipStream_t s, const double* M, int64_t n, int p, double* out);  // [64][55]
hipError_t launch_right_solve(hipStream_t s, double* M, int64_t n, int p, const double* R);
hipError_t launch_build_init(hipStream_t s, const double* M, const double* shift,
                             const double* eo, const double* logsf, int64_t n, int p,
                             int64_t nblk, double* Zp, double* einv, double* logsfp);
hipError_t launch_build_tables_bf16(hipStream_t s, const double* X, const double* off, int64_t n,
                                    int p, const ColSpec* spec, int ncol, int nth, int64_t nst,
                                    uint16_t* xab, float* offl, uint16_t* kg, uint16_t* kh,
                                    uint16_t* kh2);

// Cell ranges of the post-IRLS tile passes, aligned to the patient blocks of a
// clustered fit (chunks of one cluster are consecutive; trailing chunks may be
// empty, c0 == c1).
struct ChunkTable {
  const int64_t* c0;
  const int64_t* c1;
  const int32_t* cl;
  int nchunk;
};

// Init pass operands (kernels_post.hip): Zp in the K-operand lane layout of
// one 16-column tile, [blk][lane][r] = Z[16 blk + 4 (lane/16) + r][lane%16].
struct InitTile {
  int want_alpha;
  double moi;            // 1 / mean(exp(offset))
  int mode;              // kInit*
  const double* Zp;      // projection columns (QR: thin Q; univariate: X - mean)
  const double* einv;    // exp(-offset), padded to 16 nblk
  const double* logsf;   // offset, padded (univariate)
  const double* R;       // p x p column-major upper R (QR)
  const double* denom;   // p (univariate)
  int intercept_first;   // univariate
  double sum_sf;         // univariate
};
hipError_t launch_init_tiled(hipStream_t s, const TileDesign& td, const InitTile& it,
                             const ChunkTable& ct, Counts c, int64_t G, int64_t n, int p,
                             double* part, double* alpha0, double* beta0);

// MOM overdispersion and bread (+ clustered meat, sandwich) as FP64 MFMA tile
// passes (kernels_post.hip); part holds ceil(G/16) * nchunk * 16 * (ncp + 16)
// doubles.
hipError_t launch_mom_tiled(hipStream_t s, const TileDesign& td, const ChunkTable& ct, Counts c,
                            int64_t G, int64_t n, int p, const double* beta, double* part,
                            double* theta, int32_t* status);
hipError_t launch_bread_tiled(hipStream_t s, const TileDesign& td, const ChunkTable& ct,
                              Counts c, int64_t G, int64_t n, int p, const double* beta,
                              const double* theta, int C, double* part, double* bread,
                              double* sandwich, int32_t* status);
// meat and sandwich from per-patient score partials spart[tile][chunk][16][16]
// (design column a at column a) and a bread already in `bread`
hipError_t launch_meat_finish(hipStream_t s, const TileDesign& td, const ChunkTable& ct,
                              int64_t G, int64_t n, int p, int C, const double* spart,
                              const double* bread, double* sandwich);

hipError_t launch_irls_accum(hipStream_t s, const TileDesign& td, Counts c, int p,
                             const int32_t* active, const int32_t* nactive, int nA_bound,
                             const double* beta, const double* kvec, double* part);
// Mixed precision: FP32 transcendentals and v_mfma_f32_16x16x4f32 products,
// FP32 accumulation flushed to FP64 every 256 cells, FP64 chunk partials (the
// layout k_irls_update reads).
hipError_t launch_irls_accum_mixed(hipStream_t s, const TileDesign& td, Counts c, int p,
                                   const int32_t* active, const int32_t* nactive, int nA_bound,
                                   const double* beta, const double* kvec, double* part);

// bf16 variant on the 32-cell-step design.
hipError_t launch_irls_accum_bf16(hipStream_t s, const TileDesign& td, Counts c, int p,
                                  const int32_t* active, const int32_t* nactive, int nA_bound,
                                  const double* beta, const double* kvec, double* part);

// Wave-specialised bf16 variant (producer / consumer wavefronts, same design).
// 16-gene tiles per workgroup of the LDS-staged kernels at p columns with
// tpw tiles per compute wave.
int ws_tiles_per_workgroup(int p, int tpw);

hipError_t launch_irls_accum_ws(hipStream_t s, const TileDesign& td, Counts c, int p,
                                const int32_t* active, const int32_t* nactive, int nA_bound,
                                const double* beta, const double* kvec, double* part,
                                int tpw = 1);

// Mixed-precision bread (no clusters) on the bf16 operand tables td: the
// wave-specialised kernel with bread weights, then H^-1 per gene.  Genes are
// genes[0 .. *ngenes) (G of them, in order); 4-byte counts in 16-byte aligned
// rows only (hipErrorInvalidValue otherwise).  *fallback (zeroed by the
// caller) counts genes whose FP32 element range was exceeded: the caller then
// runs the FP64 bread.  singular[g] = 1 where H was singular.
hipError_t launch_bread_mixed(hipStream_t s, const TileDesign& td, Counts c, int p,
                              const int32_t* genes, const int32_t* ngenes, int64_t G,
                              const double* beta, const double* theta, double* part,
                              double* bread, int32_t* singular, int32_t* fallback);
// The same over the patient-aligned chunks of ct, also writing the per-patient
// score partials (spart, for launch_meat_finish).
hipError_t launch_bread_meat_mixed(hipStream_t s, const TileDesign& td, const ChunkTable& ct,
                                   Counts c, int p, const int32_t* genes, const int32_t* ngenes,
                                   int64_t G, const double* beta, const double* theta,
                                   double* part, double* spart, double* bread, int32_t* singular,
                                   int32_t* fallback);
// status[g] |= bit where flag[g] != 0.
hipError_t launch_merge_flags(hipStream_t s, int32_t* status, const int32_t* flag, int64_t G,
                              int32_t bit);

hipError_t launch_restart_genes(hipStream_t s, const int32_t* list, int n, int p,
                                const double* beta0, double* beta, int32_t* iters,
                                int32_t* status, double* bprev = nullptr);

// Sums the chunk partials of each active tile into R slices per tile
// (out[tile][R][16][ncp]); for rounds of few tiles and many chunks.
hipError_t launch_irls_reduce(hipStream_t s, int nchunk, int ncp, int R, int ntile_bound,
                              const int32_t* nactive, const double* part, double* out);

hipError_t launch_irls_update(hipStream_t s, const TileDesign& td, int p,
                              const int32_t* active, const int32_t* nactive, int nA_bound,
                              const double* part, double* beta, const double* kvec,
                              int max_iter, double eps, int32_t* iters, int32_t* status,
                              int32_t* next, int32_t* nnext, double band = 0.0,
                              int32_t* recheck = nullptr, int32_t* nrecheck = nullptr,
                              double* bprev = nullptr, int polish = 0);
hipError_t launch_fill_f64(hipStream_t s, double* x, int64_t n, double v);
// genes to polish (see k_polish_select) appended to list / *nlist
hipError_t launch_polish_select(hipStream_t s, int64_t G, int p, const double* beta,
                                const double* bprev, const int32_t* status, double below,
                                int32_t* list, int32_t* nlist);
hipError_t launch_iota(hipStream_t s, int32_t* p, int n);

// One problem's share of a joint round (kernels_tile.hip): its round shape,
// counts, active list and outputs, its workgroups of the accumulate grid and
// its blocks of the update grid.
struct JointItem {
  TileDesign td;
  Counts c;
  const int32_t* active;
  const int32_t* nactive;
  double* beta;         // beta the accumulate reads (polish: beta before the last step)
  const double* kvec;
  double* part;
  double* ubeta;        // beta the update writes
  int32_t* iters;
  int32_t* status;
  int32_t* next;
  int32_t* nnext;
  int32_t* recheck;
  int32_t* nrecheck;
  double* bprev;
  double eps, band;
  int max_iter, polish;
  int wg0, nwg;         // accumulate workgroups [wg0, wg0 + nwg)
  int ub0, nub;         // update blocks [ub0, ub0 + nub)
};
enum { kJointFp64 = 0, kJointWs = 1 };
hipError_t launch_irls_accum_joint(hipStream_t s, int kind, int p, const JointItem* items,
                                   int nitems, int total_wg);
hipError_t launch_irls_update_joint(hipStream_t s, int p, const JointItem* items, int nitems,
                                    int total_ub);

struct PortState;   // port_driver.hpp
struct NewodState;  // newod_driver.hpp

// Per-gene state of the tiled MLE (kernels_mle.hip); gene-major device arrays.
struct MleState {
  double* x = nullptr;       // current log(theta)
  double* lo = nullptr;      // Newton bracket
  double* hi = nullptr;
  double* start = nullptr;   // (var(y) - ybar) / ybar^2, or 0.5
  double* sum_y = nullptr;
  double* sum_py = nullptr;  // sum (y - 1) y
  double* max_y = nullptr;
  int32_t* mode = nullptr;   // 0 far-left pass pending, 1 Newton, 2 done
  int32_t* iters = nullptr;  // score passes
  double* theta = nullptr;   // result
  int32_t* status = nullptr; // DEVIL_GENE_* bits (or-ed)
  int32_t* tab_n = nullptr;  // distinct positive counts in the gene's table, -1 = none
  int32_t* tab_u = nullptr;  // [g * kMleTabCap + k] value
  int32_t* tab_f = nullptr;  // [g * kMleTabCap + k] frequency
  PortState* port = nullptr; // per-gene state of the nlminb driver
  int driver = 0;            // kMleDriver*
  double* raw_start = nullptr;  // (var(y) - ybar) / ybar^2 before any fallback, or NULL
};
constexpr int kMleTabCap = 4096;  // one entry per integer count below 4096
enum { kMleDriverNlminb = 0, kMleDriverRoot = 1 };

hipError_t launch_mle_setup(hipStream_t s, Counts c, int64_t G, int64_t n, int p,
                            const MleState& ms);
hipError_t launch_mle_collect(hipStream_t s, const int32_t* mode, int G, int32_t* act,
                              int32_t* nact);
hipError_t launch_mle_accum(hipStream_t s, const TileDesign& td, Counts c, int p,
                            const int32_t* active, const int32_t* nactive, int nA_bound,
                            const double* beta, const double* xs, const int32_t* tab_n,
                            int want_f, double* part);
hipError_t launch_mle_update(hipStream_t s, const TileDesign& td, int p, const int32_t* active,
                             const int32_t* nactive, int nA_bound, const double* part,
                             const MleState& ms, int64_t n, int max_iter, double rel_tol,
                             int32_t* next, int32_t* nnext);

// k: per-gene k as src/beta.cpp holds it (float(1/float(alpha0)), widened)
hipError_t launch_k_from_alpha(hipStream_t s, const double* alpha0, int64_t G, double* k);

// "new" / "I" overdispersion (kernels_newod.hip)
hipError_t launch_newod_init(hipStream_t s, int64_t G, const double* raw_start,
                             NewodState* st, double* kreq, int32_t* act, int32_t* nact,
                             int max_iter, double eps);
hipError_t launch_newod_accum(hipStream_t s, const TileDesign& td, Counts c, int p,
                              const int32_t* active, const int32_t* nactive, int nA_bound,
                              const double* beta, const double* kreq, const int32_t* tab_n,
                              double* part);
hipError_t launch_newod_update(hipStream_t s, int nchunk, const int32_t* active,
                               const int32_t* nactive, int nA_bound, const double* part,
                               const MleState& ms, NewodState* st, double* kreq, int64_t n,
                               int max_iter, double eps, int32_t* next, int32_t* nnext);

hipError_t launch_test_de(hipStream_t s, const double* beta, const double* bread,
                          const double* sandwich, int64_t G, int p, double df,
                          const double* contrast, double max_lfc, double* pval,
                          double* logp, double* lfc, void* bh_scratch, size_t bh_bytes,
                          double* adj);

// Benjamini-Hochberg adjustment of pval (device) into adj through a device
// sort; scratch of bh_scratch_bytes(G) bytes.
size_t bh_scratch_bytes(int64_t G);
hipError_t launch_bh(hipStream_t s, const double* pval, int64_t G, double* adj, void* scratch,
                     size_t scratch_bytes);

// G x p column-major <-> gene-major (to_gene_major = 1: col-major -> gene-major)
hipError_t launch_transpose_gp(hipStream_t s, const double* src, int64_t G, int p,
                               int to_gene_major, double* dst);

hipError_t launch_fill_i32(hipStream_t s, int32_t* p, int64_t count, int32_t v);
hipError_t launch_copy_counts(hipStream_t s, const int32_t* src, int n, int32_t* dst);
// stats::p.adjust methods (DEVIL_PADJ_* values) on the device
enum { kPadjBH = 0, kPadjBY = 1, kPadjHolm = 2, kPadjHochberg = 3, kPadjHommel = 4,
       kPadjBonferroni = 5, kPadjNone = 6 };
struct PadjBufs {
  double* ks;       // ascending p (NaN last)
  int32_t* is;      // their genes
  int32_t* m;       // non-NaN count (device)
  double* t;        // terms / hommel q1
  double* scan;     // scanned terms / hommel pa0
  void* tmp;
  size_t tmp_bytes;
};
// sort into b (scratch of bh_scratch_bytes(G)), then, with m read back,
// finish one method into adj
hipError_t launch_padj_sort(hipStream_t s, const double* pval, int64_t G, void* scratch,
                            size_t scratch_bytes, PadjBufs& b);
hipError_t launch_padj_finish(hipStream_t s, const double* pval, int64_t G, int m, int method,
                              double q, const PadjBufs& b, double* adj);


// count-matrix preparation (kernels_prep.hip)
hipError_t launch_size_factors(hipStream_t s, Counts c, int64_t G, int64_t n, int method,
                               int slices, int64_t* isum, double* dsum, double* dmin,
                               double* sf, double* logsf, int32_t* zero_flag);
hipError_t launch_group_cells(hipStream_t s, Counts c, int64_t G, int64_t n,
                              const int64_t* ord, void* out, int64_t ld_out);
hipError_t launch_csc_to_dense(hipStream_t s, const int64_t* colptr, const int32_t* rowind,
                               const double* values, int64_t G, int64_t n, int out_dtype,
                               void* out, int64_t ld);
hipError_t launch_pad_rows(hipStream_t s, const void* src, int64_t ld_src, int64_t G, int64_t n,
                           void* dst, int64_t ld_dst);
hipError_t launch_check_order(hipStream_t s, const int64_t* ord, int64_t n, int32_t* flag);
hipError_t launch_check_csc(hipStream_t s, const int64_t* colptr, const int32_t* rowind,
                            int64_t G, int64_t n, int64_t nnz, int32_t* flag);

}  // namespace dv
