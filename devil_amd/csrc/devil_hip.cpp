// C ABI of the gfx950 NB-GLM fit (include/devil_hip.h).
//
// devil_fit runs R/main.R:375-523 (cpu_fit) on the GPU:
//   init pass (alpha0, beta0) -> k = float(1/float(alpha0)) -> IRLS ->
//   dispersion (MOM | MLE | none) -> bread + clustered meat pass.
// Per-gene intermediates live on the device in gene-major layout; host-memory
// problems stream their counts in gene batches sized to the free HBM.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/devil_hip.h"
#include "kernels.hpp"

struct devil_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  int timing = 0;
  int tune[DEVIL_N_TUNE] = {DEVIL_MIXED_BF16};
  double stage_ms[DEVIL_N_STAGES] = {0, 0, 0, 0, 0};
  // per-launch records of the IRLS accumulate kernels (timing on)
  struct Launch {
    hipEvent_t a, b;
    int kind;
    double bytes;
  };
  std::vector<hipEvent_t> ev_pool;
  std::vector<Launch> launches;
  int64_t k_n[DEVIL_N_KERNELS] = {0, 0};
  double k_ms[DEVIL_N_KERNELS] = {0, 0};
  double k_bytes[DEVIL_N_KERNELS] = {0, 0};
};

namespace {

int fail(devil_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->err = msg;
  return code;
}

#define DV_HIP(call)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess)                                                              \
      return fail(ctx, DEVIL_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Device allocation owned by the caller's scope.
struct DevBuf {
  void* p = nullptr;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  hipError_t alloc(size_t bytes) {
    if (p) {
      (void)hipFree(p);
      p = nullptr;
    }
    if (bytes == 0) bytes = 8;
    return hipMalloc(&p, bytes);
  }
  template <typename T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

#define DV_ALLOC(buf, bytes)                                                             \
  do {                                                                                   \
    hipError_t e_ = (buf).alloc(bytes);                                                  \
    if (e_ != hipSuccess)                                                                \
      return fail(ctx, DEVIL_E_NOMEM, std::string("hipMalloc ") + std::to_string(bytes) + \
                                          " bytes: " + hipGetErrorString(e_));           \
  } while (0)

size_t count_elem_size(int dtype) { return dtype == DEVIL_COUNTS_F64 ? 8 : 4; }

// Householder QR of the n x p column-major X: thin Q (n x p) and upper R.
// Returns false when a diagonal of R is below 1e-7 of the largest, the rank
// tolerance R's qr() uses (LINPACK dqrdc2 tol = 1e-7).
bool householder_qr(const double* X, int64_t n, int p, std::vector<double>& Q,
                    std::vector<double>& R) {
  std::vector<double> A(X, X + (size_t)n * p), V((size_t)n * p, 0.0);
  std::vector<double> tau(p, 0.0);
  bool ok = true;
  for (int k = 0; k < p; ++k) {
    double nrm = 0.0;
    for (int64_t i = k; i < n; ++i) nrm += A[i + k * n] * A[i + k * n];
    nrm = std::sqrt(nrm);
    if (nrm == 0.0) {
      ok = false;
      continue;
    }
    const double alpha = A[k + k * n] > 0 ? -nrm : nrm;
    for (int64_t i = k; i < n; ++i) V[i + k * n] = A[i + k * n];
    V[k + k * n] -= alpha;
    double vn = 0.0;
    for (int64_t i = k; i < n; ++i) vn += V[i + k * n] * V[i + k * n];
    tau[k] = vn > 0 ? 2.0 / vn : 0.0;
    for (int j = k; j < p; ++j) {
      double s = 0.0;
      for (int64_t i = k; i < n; ++i) s += V[i + k * n] * A[i + j * n];
      s *= tau[k];
      for (int64_t i = k; i < n; ++i) A[i + j * n] -= s * V[i + k * n];
    }
  }
  R.assign((size_t)p * p, 0.0);
  for (int a = 0; a < p; ++a)
    for (int b = a; b < p; ++b) R[a + b * p] = A[a + b * n];
  Q.assign((size_t)n * p, 0.0);
  for (int j = 0; j < p; ++j) Q[j + j * n] = 1.0;
  for (int k = p - 1; k >= 0; --k) {
    if (tau[k] == 0.0) continue;
    for (int j = 0; j < p; ++j) {
      double s = 0.0;
      for (int64_t i = k; i < n; ++i) s += V[i + k * n] * Q[i + j * n];
      s *= tau[k];
      for (int64_t i = k; i < n; ++i) Q[i + j * n] -= s * V[i + k * n];
    }
  }
  double dmax = 0.0;
  for (int a = 0; a < p; ++a) dmax = std::max(dmax, std::fabs(R[a + a * p]));
  for (int a = 0; a < p; ++a)
    if (!(std::fabs(R[a + a * p]) > 1e-7 * dmax)) ok = false;
  return ok;
}

// Host copies of the small per-cell inputs plus what the init pass needs.
struct Design {
  int64_t n = 0;
  int p = 0;
  std::vector<double> X, off, eo, logsf, Q, R, xmean, denom;
  double moi = 0.0, sum_sf = 0.0;
  int intercept_first = 0;
};

// Device copies of the design, offsets and cluster ends.
struct DeviceDesign {
  DevBuf X, off, eo, Q, R, logsf, xmean, denom, ends;
};

int prepare_design(devil_ctx* ctx, const double* X, const double* off, int64_t n, int p,
                   int mem, int init_mode, Design& d) {
  d.n = n;
  d.p = p;
  d.X.resize((size_t)n * p);
  d.off.resize(n);
  if (mem == DEVIL_MEM_DEVICE) {
    DV_HIP(hipMemcpy(d.X.data(), X, sizeof(double) * n * p, hipMemcpyDeviceToHost));
    DV_HIP(hipMemcpy(d.off.data(), off, sizeof(double) * n, hipMemcpyDeviceToHost));
  } else {
    std::memcpy(d.X.data(), X, sizeof(double) * n * p);
    std::memcpy(d.off.data(), off, sizeof(double) * n);
  }
  d.eo.resize(n);
  double se = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    d.eo[i] = std::exp(d.off[i]);  // exp_offset, R/main.R:390
    se += d.eo[i];
  }
  d.moi = 1.0 / (se / (double)n);  // R/overdispersion.R:130
  if (init_mode == DEVIL_INIT_QR) {
    if (!householder_qr(d.X.data(), n, p, d.Q, d.R))
      return fail(ctx, DEVIL_E_RANK, "design matrix is rank deficient (QR init, R/beta.R:14)");
  } else if (init_mode == DEVIL_INIT_UNIVARIATE) {
    // src/beta_init_univariate.cpp:48-84
    double mx = 0.0;
    for (int64_t i = 0; i < n; ++i) mx = std::max(mx, std::fabs(d.X[i] - 1.0));
    d.intercept_first = mx < 1e-10;
    d.logsf.resize(n);
    d.sum_sf = 0.0;
    for (int64_t i = 0; i < n; ++i) {
      d.logsf[i] = std::log(d.eo[i]);
      d.sum_sf += d.eo[i];
    }
    d.xmean.assign(p, 0.0);
    d.denom.assign(p, 0.0);
    for (int j = 0; j < p; ++j) {
      double s = 0.0;
      for (int64_t i = 0; i < n; ++i) s += d.X[i + j * n];
      d.xmean[j] = s / (double)n;
      double q = 0.0;
      for (int64_t i = 0; i < n; ++i) {
        const double c = d.X[i + j * n] - d.xmean[j];
        q += c * c;
      }
      d.denom[j] = q <= 0.0 ? 1e-12 : q;
    }
  }
  return DEVIL_OK;
}

template <typename T>
int upload(devil_ctx* ctx, DevBuf& b, const std::vector<T>& v) {
  DV_ALLOC(b, sizeof(T) * v.size());
  if (!v.empty())
    DV_HIP(hipMemcpyAsync(b.p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice,
                          ctx->stream));
  return DEVIL_OK;
}

int upload_design(devil_ctx* ctx, const Design& d, DeviceDesign& dd) {
  int rc;
  if ((rc = upload(ctx, dd.X, d.X))) return rc;
  if ((rc = upload(ctx, dd.off, d.off))) return rc;
  if ((rc = upload(ctx, dd.eo, d.eo))) return rc;
  if (!d.Q.empty()) {
    if ((rc = upload(ctx, dd.Q, d.Q))) return rc;
    if ((rc = upload(ctx, dd.R, d.R))) return rc;
  }
  if (!d.logsf.empty()) {
    if ((rc = upload(ctx, dd.logsf, d.logsf))) return rc;
    if ((rc = upload(ctx, dd.xmean, d.xmean))) return rc;
    if ((rc = upload(ctx, dd.denom, d.denom))) return rc;
  }
  return DEVIL_OK;
}

dv::InitParams init_params(const Design& d, const DeviceDesign& dd, int want_alpha,
                           int init_mode) {
  dv::InitParams ip{};
  ip.want_alpha = want_alpha;
  ip.moi = d.moi;
  ip.init_mode = init_mode;
  ip.eo = dd.eo.as<double>();
  ip.Q = dd.Q.as<double>();
  ip.R = dd.R.as<double>();
  ip.X = dd.X.as<double>();
  ip.logsf = dd.logsf.as<double>();
  ip.xmean = dd.xmean.as<double>();
  ip.denom = dd.denom.as<double>();
  ip.intercept_first = d.intercept_first;
  ip.sum_sf = d.sum_sf;
  return ip;
}

int check_clusters(devil_ctx* ctx, const int32_t* ends, int C, int64_t n) {
  int64_t prev = 0;
  for (int c = 0; c < C; ++c) {
    if (ends[c] < prev)
      return fail(ctx, DEVIL_E_CLUSTERS,
                  "cluster ends must be non-decreasing (R/main.R:192-193)");
    prev = ends[c];
  }
  if (C > 0 && prev != n)
    return fail(ctx, DEVIL_E_CLUSTERS, "last cluster end must equal n");
  return DEVIL_OK;
}

int check_counts_args(devil_ctx* ctx, const void* counts, int dtype, int64_t G, int64_t n,
                      int64_t ld) {
  if (G < 0 || n < 2) return fail(ctx, DEVIL_E_ARG, "need G >= 0 and n >= 2");
  if (G > 0 && !counts) return fail(ctx, DEVIL_E_ARG, "counts is NULL");
  if (dtype < DEVIL_COUNTS_I32 || dtype > DEVIL_COUNTS_F64)
    return fail(ctx, DEVIL_E_ARG, "unknown counts dtype");
  if (ld < n) return fail(ctx, DEVIL_E_ARG, "counts_ld < n");
  if (G > 0x7fffffffLL) return fail(ctx, DEVIL_E_ARG, "G above 2^31-1");
  return DEVIL_OK;
}

int check_p(devil_ctx* ctx, int p) {
  if (p < 1 || p > DEVIL_PMAX)
    return fail(ctx, DEVIL_E_ARG, "p must be in [1, " + std::to_string(DEVIL_PMAX) + "]");
  return DEVIL_OK;
}

// Records HIP events around stages on the context stream; elapsed times are
// summed into ctx->stage_ms once the stream has been synchronised.
struct StageTimer {
  devil_ctx* ctx;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> marks;
  explicit StageTimer(devil_ctx* c) : ctx(c) {}
  ~StageTimer() {
    for (auto& m : marks) {
      (void)hipEventDestroy(m.second.first);
      (void)hipEventDestroy(m.second.second);
    }
  }
  hipEvent_t begin(int stage) {
    if (!ctx->timing) return nullptr;
    hipEvent_t a = nullptr, b = nullptr;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return nullptr;
    (void)hipEventRecord(a, ctx->stream);
    marks.push_back({stage, {a, b}});
    return b;
  }
  void end(hipEvent_t b) {
    if (b) (void)hipEventRecord(b, ctx->stream);
  }
  // call after the stream is synchronised
  void collect() {
    for (auto& m : marks) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, m.second.first, m.second.second) == hipSuccess)
        ctx->stage_ms[m.first] += ms;
    }
  }
};

// Host construction of the MFMA operand layouts (dv::TileDesign).
enum MixedKernel { kMixedNone = 0, kMixedF32 = 1, kMixedBf16 = 2 };

int mixed_kernel(const devil_ctx* ctx) {
  return ctx->tune[DEVIL_TUNE_MIXED_KERNEL] == DEVIL_MIXED_F32 ? kMixedF32 : kMixedBf16;
}

struct ColSet {
  std::vector<std::vector<double>> cols;
  int kidx_pair[55];
  int kidx_lin[10];
};

// Columns of K: x_a x_b (a <= b) and x_a of the design, exact duplicates
// merged (an intercept makes x_0 x_a == x_a, a 0/1 column squares to itself)
// and all-zero columns dropped.  lin16: the p design columns sit at 0..p-1
// as they are and the pair columns start at column 16.
void build_columns(const double* X, int64_t n, int p, bool lin16, ColSet& cs) {
  std::vector<uint64_t> hashes;
  auto hash_of = [&](const std::vector<double>& v) {
    uint64_t hsh = 1469598103934665603ULL;
    for (double x : v) {
      uint64_t u;
      std::memcpy(&u, &x, 8);
      hsh = (hsh ^ u) * 1099511628211ULL;
    }
    return hsh;
  };
  auto add_col = [&](std::vector<double>&& v) -> int {
    bool zero = true;
    for (double x : v)
      if (x != 0.0) {
        zero = false;
        break;
      }
    if (zero) return -1;
    const uint64_t hsh = hash_of(v);
    for (size_t c = 0; c < cs.cols.size(); ++c)
      if (hashes[c] == hsh && std::memcmp(cs.cols[c].data(), v.data(), sizeof(double) * n) == 0)
        return (int)c;
    cs.cols.push_back(std::move(v));
    hashes.push_back(hsh);
    return (int)cs.cols.size() - 1;
  };
  if (lin16) {
    for (int a = 0; a < p; ++a) {
      std::vector<double> v(X + a * n, X + (a + 1) * n);
      hashes.push_back(hash_of(v));
      cs.cols.push_back(std::move(v));
      cs.kidx_lin[a] = a;
    }
    while (cs.cols.size() < 16) {
      cs.cols.emplace_back(n, 0.0);
      hashes.push_back(~0ULL);
    }
  } else {
    for (int a = 0; a < p; ++a)
      cs.kidx_lin[a] = add_col(std::vector<double>(X + a * n, X + (a + 1) * n));
  }
  int t = 0;
  for (int a = 0; a < p; ++a)
    for (int b = a; b < p; ++b) {
      std::vector<double> v(n);
      for (int64_t i = 0; i < n; ++i) v[i] = X[i + a * n] * X[i + b * n];
      cs.kidx_pair[t++] = add_col(std::move(v));
    }
}

// Shape of a design with cells_per_block-cell blocks over the given columns.
void shape_design(const ColSet& cs, int64_t n, int p, int cells_per_block, dv::TileDesign& td) {
  std::memset(&td, 0, sizeof(td));
  for (int a = 0; a < p; ++a) td.kidx_lin[a] = cs.kidx_lin[a];
  for (int q = 0; q < p * (p + 1) / 2; ++q) td.kidx_pair[q] = cs.kidx_pair[q];
  for (size_t c = 0; c < (size_t)DEVIL_NCP_MAX; ++c) {
    long double sum = 0.0L;
    if (c < cs.cols.size())
      for (int64_t i = 0; i < n; ++i) sum += cs.cols[c][i];
    td.ksum[c] = (double)sum;
  }
  const int ncol = std::max<int>(1, (int)cs.cols.size());
  td.nt = (ncol + 15) / 16;
  td.ncp = 16 * td.nt;
  td.n = n;
  td.nblk = (n + cells_per_block - 1) / cells_per_block;
  const int64_t min_blocks = 2048 / cells_per_block;
  td.chunk_blocks = std::max<int64_t>(min_blocks, (td.nblk + 63) / 64);
  td.nchunk = (int)((td.nblk + td.chunk_blocks - 1) / td.chunk_blocks);
  if (td.nchunk >= 8) {
    // a multiple of 8 chunks: with blocks dealt round-robin to the 8 XCDs and
    // chunk = blockIdx % nchunk, each chunk's operands then live in one
    // XCD's L2 (trailing chunks may come out empty)
    td.nchunk = 8 * ((td.nchunk + 7) / 8);
    td.chunk_blocks = (td.nblk + td.nchunk - 1) / td.nchunk;
  }
}

uint16_t to_bf16(float x) {  // round to nearest even
  uint32_t u;
  std::memcpy(&u, &x, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

float from_bf16(uint16_t b) {
  const uint32_t u = (uint32_t)b << 16;
  float x;
  std::memcpy(&x, &u, 4);
  return x;
}

// cell of MFMA k index (h, i) within a 32-cell step of the bf16 design
int step_cell(int h, int i) { return i < 4 ? 4 * h + i : 16 + 4 * h + (i - 4); }

struct TileHost {
  dv::TileDesign td{};   // FP64 design (16-cell blocks, packed columns)
  dv::TileDesign tdm{};  // mixed design (copy of td for FP32, 32-cell steps for bf16)
  int mixed = kMixedNone;
  std::vector<double> Xp, offp, Kp;
  std::vector<float> Xf, offf, Kf;  // FP32 kernel operands
  std::vector<uint16_t> xab, kh;    // bf16 kernel operands
  std::vector<float> offl, kg;
};

void build_tile_design(const double* X, const double* off, int64_t n, int p, TileHost& th,
                       int mixed = kMixedNone) {
  th.mixed = mixed;
  ColSet cs;
  build_columns(X, n, p, false, cs);
  dv::TileDesign& td = th.td;
  shape_design(cs, n, p, 16, td);
  const auto& cols = cs.cols;
  const int NJ = (p + 3) / 4;
  th.Xp.assign((size_t)td.nblk * NJ * 64, 0.0);
  th.offp.assign((size_t)td.nblk * 16, 0.0);
  th.Kp.assign((size_t)td.nblk * td.nt * 256, 0.0);
  for (int64_t i = 0; i < n; ++i) th.offp[i] = off[i];
  for (int64_t b = 0; b < td.nblk; ++b) {
    // eta-product A operand: row m of the block is cell 4 (m % 4) + m / 4
    for (int j = 0; j < NJ; ++j)
      for (int l = 0; l < 64; ++l) {
        const int m = l % 16;
        const int64_t cell = 16 * b + 4 * (m % 4) + m / 4;
        const int f = 4 * j + l / 16;
        if (cell < n && f < p) th.Xp[(b * NJ + j) * 64 + l] = X[cell + f * n];
      }
    // contraction B operand: lane l, step r -> cell 4 (l / 16) + r, column l % 16
    for (int tt = 0; tt < td.nt; ++tt)
      for (int l = 0; l < 64; ++l)
        for (int r = 0; r < 4; ++r) {
          const int64_t cell = 16 * b + 4 * (l / 16) + r;
          const int col = 16 * tt + l % 16;
          if (cell < n && col < (int)cols.size())
            th.Kp[((b * td.nt + tt) * 64 + l) * 4 + r] = cols[col][cell];
        }
  }
  if (mixed == kMixedF32) {
    // FP32 copies; the eta product rows keep the natural cell order (the f32
    // MFMA returns rows 4 (l / 16) + r in lane l)
    th.tdm = td;
    th.Xf.assign((size_t)td.nblk * NJ * 64, 0.0f);
    th.offf.assign(th.offp.begin(), th.offp.end());
    th.Kf.assign(th.Kp.begin(), th.Kp.end());
    for (int64_t b = 0; b < td.nblk; ++b)
      for (int j = 0; j < NJ; ++j)
        for (int l = 0; l < 64; ++l) {
          const int64_t cell = 16 * b + l % 16;
          const int f = 4 * j + l / 16;
          if (cell < n && f < p) th.Xf[(b * NJ + j) * 64 + l] = (float)X[cell + f * n];
        }
  } else if (mixed == kMixedBf16) {
    ColSet cm;
    build_columns(X, n, p, true, cm);
    dv::TileDesign& tm = th.tdm;
    shape_design(cm, n, p, 32, tm);
    const int64_t nst = tm.nblk;
    const int NE = (7 * p + 31) / 32;
    const int nth = tm.nt - 1;
    const double L2E = 1.44269504088896340736;
    th.xab.assign((size_t)nst * 2 * NE * 64 * 8, 0);
    th.offl.assign((size_t)nst * 32, 0.0f);
    th.kg.assign((size_t)nst * 64 * 8, 0.0f);
    th.kh.assign((size_t)nst * std::max(nth, 0) * 64 * 8, 0);
    for (int64_t i = 0; i < n; ++i) th.offl[i] = (float)(off[i] * L2E);
    for (int64_t st = 0; st < nst; ++st) {
      for (int u = 0; u < 2; ++u)
        for (int j = 0; j < NE; ++j)
          for (int l = 0; l < 64; ++l)
            for (int i = 0; i < 8; ++i) {
              const int64_t cell = st * 32 + 16 * u + l % 16;
              const int e = 32 * j + 8 * (l / 16) + i;
              const int f = e / 7, term = e % 7;
              if (cell >= n || f >= p) continue;
              const double x = X[cell + f * n];
              const uint16_t x1 = to_bf16((float)x);
              const uint16_t x2 = to_bf16((float)(x - (double)from_bf16(x1)));
              th.xab[((((size_t)st * 2 + u) * NE + j) * 64 + l) * 8 + i] = term < 4 ? x1 : x2;
            }
      for (int l = 0; l < 64; ++l)
        for (int i = 0; i < 8; ++i) {
          const int64_t cell = st * 32 + step_cell(l / 16, i);
          const int col = l % 16;
          if (cell < n && col < p) th.kg[((size_t)st * 64 + l) * 8 + i] = (float)X[cell + col * n];
        }
      for (int t = 0; t < nth; ++t)
        for (int l = 0; l < 64; ++l)
          for (int i = 0; i < 8; ++i) {
            const int64_t cell = st * 32 + step_cell(l / 16, i);
            const int col = 16 * (t + 1) + l % 16;
            if (cell < n && col < (int)cm.cols.size())
              th.kh[(((size_t)st * nth + t) * 64 + l) * 8 + i] = to_bf16((float)cm.cols[col][cell]);
          }
    }
  }
}

struct DeviceTile {
  DevBuf Xp, offp, Kp, Xf, offf, Kf, xab, offl, kg, kh;
};

int upload_tile(devil_ctx* ctx, TileHost& th, DeviceTile& dt) {
  int rc;
  if ((rc = upload(ctx, dt.Xp, th.Xp))) return rc;
  if ((rc = upload(ctx, dt.offp, th.offp))) return rc;
  if ((rc = upload(ctx, dt.Kp, th.Kp))) return rc;
  th.td.Xp = dt.Xp.as<double>();
  th.td.offp = dt.offp.as<double>();
  th.td.Kp = dt.Kp.as<double>();
  if (th.mixed == kMixedF32) {
    if ((rc = upload(ctx, dt.Xf, th.Xf))) return rc;
    if ((rc = upload(ctx, dt.offf, th.offf))) return rc;
    if ((rc = upload(ctx, dt.Kf, th.Kf))) return rc;
    th.tdm.Xf = dt.Xf.as<float>();
    th.tdm.offf = dt.offf.as<float>();
    th.tdm.Kf = dt.Kf.as<float>();
  } else if (th.mixed == kMixedBf16) {
    if ((rc = upload(ctx, dt.xab, th.xab))) return rc;
    if ((rc = upload(ctx, dt.offl, th.offl))) return rc;
    if ((rc = upload(ctx, dt.kg, th.kg))) return rc;
    if (!th.kh.empty() && (rc = upload(ctx, dt.kh, th.kh))) return rc;
    th.tdm.xab = dt.xab.as<uint16_t>();
    th.tdm.offl = dt.offl.as<float>();
    th.tdm.kg = dt.kg.as<float>();
    th.tdm.kh = th.kh.empty() ? nullptr : dt.kh.as<uint16_t>();
  }
  return DEVIL_OK;
}

// IRLS over genes [0, G) of the counts view: rounds of accumulate + update
// over the still-active genes until none is left (src/beta.cpp:37).
// beta, k, iters, status are gene-major device arrays for these G genes.
//
// With timing on, each accumulate launch is bracketed by a pair of events;
// the pairs are read once the caller has synchronised (collect_launches).
hipEvent_t pool_event(devil_ctx* ctx) {
  if (!ctx->ev_pool.empty()) {
    hipEvent_t e = ctx->ev_pool.back();
    ctx->ev_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

struct LaunchTimer {
  devil_ctx* ctx;
  hipEvent_t a = nullptr;
  int kind;
  double bytes;
  LaunchTimer(devil_ctx* c, int k, double b) : ctx(c), kind(k), bytes(b) {
    if (!ctx->timing) return;
    a = pool_event(ctx);
    if (a && hipEventRecord(a, ctx->stream) != hipSuccess) {
      ctx->ev_pool.push_back(a);
      a = nullptr;
    }
  }
  ~LaunchTimer() {
    if (!a) return;
    hipEvent_t b = pool_event(ctx);
    if (!b || hipEventRecord(b, ctx->stream) != hipSuccess) {
      ctx->ev_pool.push_back(a);
      if (b) ctx->ev_pool.push_back(b);
      return;
    }
    ctx->launches.push_back({a, b, kind, bytes});
  }
};

// after a stream synchronisation: fold the recorded launches into the stats
void collect_launches(devil_ctx* ctx) {
  for (auto& l : ctx->launches) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, l.a, l.b) == hipSuccess) {
      ctx->k_n[l.kind] += 1;
      ctx->k_ms[l.kind] += ms;
      ctx->k_bytes[l.kind] += l.bytes;
    }
    ctx->ev_pool.push_back(l.a);
    ctx->ev_pool.push_back(l.b);
  }
  ctx->launches.clear();
}

//
// Mixed precision: each round accumulates in FP32 (FP64 partials); a gene
// whose FP32-based step lands within kMixedBand * eps of the stopping
// threshold (or is NaN) is not updated but recomputed in FP64 in the same
// round, so stopping decisions and iteration counts follow FP64 arithmetic.
constexpr double kMixedBand = 1e-3;

int run_irls_tiled(devil_ctx* ctx, const TileHost& tl, dv::Counts c, int64_t G, int p,
                   double* beta, const double* kvec, int max_iter, double eps,
                   int32_t* iters, int32_t* status) {
  const dv::TileDesign& td = tl.td;
  const dv::TileDesign& tdm = tl.tdm;
  const bool mixed = tl.mixed != kMixedNone;
  hipStream_t s = ctx->stream;
  const size_t es = count_elem_size(c.dtype);
  DV_HIP(dv::launch_fill_i32(s, iters, G, 0));
  DV_HIP(dv::launch_fill_i32(s, status, G, 0));
  if (max_iter <= 0 || G == 0) {
    if (max_iter <= 0) DV_HIP(dv::launch_fill_i32(s, status, G, DEVIL_GENE_MAXITER));
    return DEVIL_OK;
  }
  DevBuf act0, act1, cnt, part, rchk;
  DV_ALLOC(act0, sizeof(int32_t) * G);
  DV_ALLOC(act1, sizeof(int32_t) * G);
  DV_ALLOC(cnt, sizeof(int32_t) * 3);
  if (mixed) DV_ALLOC(rchk, sizeof(int32_t) * G);
  int32_t* nrchk = cnt.as<int32_t>() + 2;
  const int64_t ntile = (G + 15) / 16;
  size_t part_per_tile = (size_t)td.nchunk * 16 * td.ncp;
  if (mixed) part_per_tile = std::max(part_per_tile, (size_t)tdm.nchunk * 16 * tdm.ncp);
  DV_ALLOC(part, sizeof(double) * (size_t)ntile * part_per_tile);
  DV_HIP(dv::launch_iota(s, act0.as<int32_t>(), (int)G));
  int32_t* cur = act0.as<int32_t>();
  int32_t* nxt = act1.as<int32_t>();
  int32_t* ncur = cnt.as<int32_t>();
  int32_t* nnxt = cnt.as<int32_t>() + 1;
  DV_HIP(dv::launch_fill_i32(s, ncur, 1, (int32_t)G));
  int nA = (int)G;
  for (int it = 0; it < max_iter && nA > 0; ++it) {
    int32_t h[2] = {0, 0};
    const double row_bytes = (double)es * (double)td.n;
    if (!mixed) {
      {
        LaunchTimer lt(ctx, DEVIL_KERNEL_IRLS_FP64, row_bytes * nA);
        DV_HIP(dv::launch_irls_accum(s, td, c, p, cur, ncur, nA, beta, kvec, part.as<double>()));
      }
      DV_HIP(dv::launch_fill_i32(s, nnxt, 1, 0));
      DV_HIP(dv::launch_irls_update(s, td, p, cur, ncur, nA, part.as<double>(), beta, kvec,
                                    max_iter, eps, iters, status, nxt, nnxt));
    } else {
      {
        LaunchTimer lt(ctx, DEVIL_KERNEL_IRLS_MIXED, row_bytes * nA);
        if (tl.mixed == kMixedBf16)
          DV_HIP(dv::launch_irls_accum_bf16(s, tdm, c, p, cur, ncur, nA, beta, kvec,
                                            part.as<double>()));
        else
          DV_HIP(dv::launch_irls_accum_mixed(s, tdm, c, p, cur, ncur, nA, beta, kvec,
                                             part.as<double>()));
      }
      DV_HIP(dv::launch_fill_i32(s, nnxt, 1, 0));
      DV_HIP(dv::launch_fill_i32(s, nrchk, 1, 0));
      DV_HIP(dv::launch_irls_update(s, tdm, p, cur, ncur, nA, part.as<double>(), beta, kvec,
                                    max_iter, eps, iters, status, nxt, nnxt, kMixedBand,
                                    rchk.as<int32_t>(), nrchk));
      DV_HIP(hipMemcpyAsync(&h[1], nrchk, sizeof(int32_t), hipMemcpyDeviceToHost, s));
      DV_HIP(hipStreamSynchronize(s));
      if (h[1] > 0) {
        LaunchTimer lt(ctx, DEVIL_KERNEL_IRLS_FP64, row_bytes * h[1]);
        DV_HIP(dv::launch_irls_accum(s, td, c, p, rchk.as<int32_t>(), nrchk, h[1], beta, kvec,
                                     part.as<double>()));
        DV_HIP(dv::launch_irls_update(s, td, p, rchk.as<int32_t>(), nrchk, h[1],
                                      part.as<double>(), beta, kvec, max_iter, eps, iters,
                                      status, nxt, nnxt));
      }
    }
    DV_HIP(hipMemcpyAsync(&h[0], nnxt, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    DV_HIP(hipStreamSynchronize(s));
    collect_launches(ctx);
    nA = h[0];
    std::swap(cur, nxt);
    std::swap(ncur, nnxt);
  }
  return DEVIL_OK;
}

// MLE overdispersion over genes [0, G): setup pass, then rounds of
// accumulate + update over the genes still iterating (the first round is the
// far-left check of R/overdispersion.R:40-46).  theta, iters are outputs;
// status is or-ed.
int run_mle_tiled(devil_ctx* ctx, const dv::TileDesign& td, dv::Counts c, int64_t G, int64_t n,
                  int p, const double* beta, int max_iter, double* theta, int32_t* iters,
                  int32_t* status) {
  hipStream_t s = ctx->stream;
  if (G == 0) return DEVIL_OK;
  DevBuf dbl, ints, act0, act1, cnt, part, tab;
  DV_ALLOC(dbl, sizeof(double) * G * 7);
  DV_ALLOC(ints, sizeof(int32_t) * G * 2);
  DV_ALLOC(tab, sizeof(int32_t) * G * dv::kMleTabCap * 2);
  DV_ALLOC(act0, sizeof(int32_t) * G);
  DV_ALLOC(act1, sizeof(int32_t) * G);
  DV_ALLOC(cnt, sizeof(int32_t) * 2);
  const int W = 5 + 3 * td.ncp;
  const int64_t ntile = (G + 15) / 16;
  DV_ALLOC(part, sizeof(double) * (size_t)ntile * td.nchunk * 16 * W);
  dv::MleState ms;
  double* d = dbl.as<double>();
  ms.x = d;
  ms.lo = d + G;
  ms.hi = d + 2 * G;
  ms.start = d + 3 * G;
  ms.sum_y = d + 4 * G;
  ms.sum_py = d + 5 * G;
  ms.max_y = d + 6 * G;
  ms.mode = ints.as<int32_t>();
  ms.tab_n = ints.as<int32_t>() + G;
  ms.tab_u = tab.as<int32_t>();
  ms.tab_f = tab.as<int32_t>() + G * dv::kMleTabCap;
  ms.iters = iters;
  ms.theta = theta;
  ms.status = status;
  DV_HIP(dv::launch_mle_setup(s, c, G, n, p, ms));
  int32_t* cur = act0.as<int32_t>();
  int32_t* nxt = act1.as<int32_t>();
  int32_t* ncur = cnt.as<int32_t>();
  int32_t* nnxt = cnt.as<int32_t>() + 1;
  DV_HIP(dv::launch_fill_i32(s, ncur, 1, 0));
  DV_HIP(dv::launch_mle_collect(s, ms.mode, (int)G, cur, ncur));
  int32_t nA = 0;
  DV_HIP(hipMemcpyAsync(&nA, ncur, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  DV_HIP(hipStreamSynchronize(s));
  for (int round = 0; round <= max_iter && nA > 0; ++round) {
    DV_HIP(dv::launch_mle_accum(s, td, c, p, cur, ncur, nA, beta, ms.x, ms.tab_n,
                                part.as<double>()));
    DV_HIP(dv::launch_fill_i32(s, nnxt, 1, 0));
    DV_HIP(dv::launch_mle_update(s, td, p, cur, ncur, nA, part.as<double>(), ms, n, max_iter,
                                 nxt, nnxt));
    int32_t h = 0;
    DV_HIP(hipMemcpyAsync(&h, nnxt, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    DV_HIP(hipStreamSynchronize(s));
    nA = h;
    std::swap(cur, nxt);
    std::swap(ncur, nnxt);
  }
  return DEVIL_OK;
}

int set_device(devil_ctx* ctx) {
  if (!ctx) return DEVIL_E_ARG;
  DV_HIP(hipSetDevice(ctx->device));
  return DEVIL_OK;
}

}  // namespace

extern "C" {

const char* devil_version(void) { return "devil_hip 0.1.0 gfx950"; }

const char* devil_strerror(int code) {
  switch (code) {
    case DEVIL_OK: return "success";
    case DEVIL_E_ARG: return "invalid argument";
    case DEVIL_E_HIP: return "HIP runtime error";
    case DEVIL_E_NOMEM: return "out of memory";
    case DEVIL_E_RANK: return "design matrix is rank deficient";
    case DEVIL_E_CLUSTERS: return "invalid cluster ends";
    case DEVIL_E_NODEVICE: return "no HIP device";
    case DEVIL_E_UNSUPPORTED: return "unsupported request";
    default: return "unknown error";
  }
}

const char* devil_last_error(const devil_ctx* ctx) { return ctx ? ctx->err.c_str() : ""; }

int devil_create(devil_ctx** out, int hip_device) {
  if (!out) return DEVIL_E_ARG;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return DEVIL_E_NODEVICE;
  if (hip_device < 0 || hip_device >= count) return DEVIL_E_NODEVICE;
  if (hipSetDevice(hip_device) != hipSuccess) return DEVIL_E_HIP;
  devil_ctx* ctx = new (std::nothrow) devil_ctx();
  if (!ctx) return DEVIL_E_NOMEM;
  ctx->device = hip_device;
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
    delete ctx;
    return DEVIL_E_HIP;
  }
  *out = ctx;
  return DEVIL_OK;
}

void devil_destroy(devil_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) {
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipStreamDestroy(ctx->stream);
  }
  for (auto& l : ctx->launches) {
    (void)hipEventDestroy(l.a);
    (void)hipEventDestroy(l.b);
  }
  for (hipEvent_t e : ctx->ev_pool) (void)hipEventDestroy(e);
  delete ctx;
}

int devil_fit(devil_ctx* ctx, const devil_problem* pr, devil_outputs* out) {
  int rc;
  if ((rc = set_device(ctx))) return rc;
  if (!pr || !out) return fail(ctx, DEVIL_E_ARG, "NULL problem or outputs");
  const int64_t G = pr->G, n = pr->n;
  const int p = pr->p;
  if ((rc = check_p(ctx, p))) return rc;
  if ((rc = check_counts_args(ctx, pr->counts, pr->counts_dtype, G, n, pr->counts_ld)))
    return rc;
  if (pr->mem != DEVIL_MEM_HOST && pr->mem != DEVIL_MEM_DEVICE)
    return fail(ctx, DEVIL_E_ARG, "mem must be DEVIL_MEM_HOST or DEVIL_MEM_DEVICE");
  if (!pr->X || !pr->offset) return fail(ctx, DEVIL_E_ARG, "X or offset is NULL");
  if (pr->init_mode < DEVIL_INIT_QR || pr->init_mode > DEVIL_INIT_UNIVARIATE)
    return fail(ctx, DEVIL_E_ARG, "unknown init_mode");
  if (pr->od_mode < DEVIL_OD_NONE || pr->od_mode > DEVIL_OD_MLE)
    return fail(ctx, DEVIL_E_ARG, "unknown od_mode (R/main.R:467)");
  if (pr->max_iter < 0) return fail(ctx, DEVIL_E_ARG, "max_iter < 0");
  if (pr->precision != DEVIL_PREC_FP64 && pr->precision != DEVIL_PREC_MIXED)
    return fail(ctx, DEVIL_E_ARG, "unknown precision");
  if (pr->n_clusters < 0 || (pr->n_clusters > 0 && !pr->cluster_ends))
    return fail(ctx, DEVIL_E_ARG, "n_clusters > 0 needs cluster_ends");
  if (G > 0 && (!out->beta || !out->theta || !out->beta_iters || !out->theta_iters ||
                !out->bread))
    return fail(ctx, DEVIL_E_ARG, "an output pointer is NULL");
  if (G == 0) return DEVIL_OK;
  const int mem = pr->mem;
  const int C = pr->n_clusters;

  std::vector<int32_t> ends_h(C);
  if (C > 0) {
    if (mem == DEVIL_MEM_DEVICE)
      DV_HIP(hipMemcpy(ends_h.data(), pr->cluster_ends, sizeof(int32_t) * C,
                       hipMemcpyDeviceToHost));
    else
      std::memcpy(ends_h.data(), pr->cluster_ends, sizeof(int32_t) * C);
    if ((rc = check_clusters(ctx, ends_h.data(), C, n))) return rc;
  }
  const int init_mode = pr->beta_init ? dv::kInitNone : pr->init_mode;
  Design d;
  if ((rc = prepare_design(ctx, pr->X, pr->offset, n, p, mem, init_mode, d))) return rc;
  DeviceDesign dd;
  if ((rc = upload_design(ctx, d, dd))) return rc;
  if (C > 0 && (rc = upload(ctx, dd.ends, ends_h))) return rc;
  TileHost tl;
  build_tile_design(d.X.data(), d.off.data(), n, p, tl,
                    pr->precision == DEVIL_PREC_MIXED ? mixed_kernel(ctx) : kMixedNone);
  DeviceTile dt;
  if ((rc = upload_tile(ctx, tl, dt))) return rc;

  // per-gene device state (gene-major)
  DevBuf alpha0, kvec, beta, theta, iters, thit, status, bread, sandwich;
  DV_ALLOC(alpha0, sizeof(double) * G);
  DV_ALLOC(kvec, sizeof(double) * G);
  DV_ALLOC(beta, sizeof(double) * G * p);
  DV_ALLOC(theta, sizeof(double) * G);
  DV_ALLOC(iters, sizeof(int32_t) * G);
  DV_ALLOC(thit, sizeof(int32_t) * G);
  DV_ALLOC(status, sizeof(int32_t) * G);
  DV_ALLOC(bread, sizeof(double) * G * p * p);
  const bool want_sw = C > 0 && out->sandwich;
  if (want_sw) DV_ALLOC(sandwich, sizeof(double) * G * p * p);

  hipStream_t s = ctx->stream;
  if (pr->disp_init) {
    DV_HIP(hipMemcpyAsync(alpha0.p, pr->disp_init, sizeof(double) * G,
                          mem == DEVIL_MEM_DEVICE ? hipMemcpyDeviceToDevice
                                                  : hipMemcpyHostToDevice,
                          s));
  }
  if (pr->beta_init) {
    DevBuf tmp;
    DV_ALLOC(tmp, sizeof(double) * G * p);
    DV_HIP(hipMemcpyAsync(tmp.p, pr->beta_init, sizeof(double) * G * p,
                          mem == DEVIL_MEM_DEVICE ? hipMemcpyDeviceToDevice
                                                  : hipMemcpyHostToDevice,
                          s));
    DV_HIP(dv::launch_transpose_gp(s, tmp.as<double>(), G, p, 1, beta.as<double>()));
    DV_HIP(hipStreamSynchronize(s));
  }

  // counts: device-resident as given, or streamed in gene batches
  const size_t es = count_elem_size(pr->counts_dtype);
  int64_t batch = G;
  DevBuf cbuf;
  if (mem == DEVIL_MEM_HOST) {
    size_t freeb = 0, totalb = 0;
    DV_HIP(hipMemGetInfo(&freeb, &totalb));
    const size_t row_bytes = (size_t)pr->counts_ld * es;
    const size_t budget = std::min<size_t>(freeb / 2, (size_t)8 << 30);
    batch = std::max<int64_t>(1, std::min<int64_t>(G, (int64_t)(budget / row_bytes)));
    DV_ALLOC(cbuf, row_bytes * batch);
  }
  const double eps = (double)(float)pr->tolerance;  // src/beta.cpp:25 float eps
  const dv::InitParams ip = init_params(d, dd, pr->disp_init ? 0 : 1, init_mode);
  for (int k = 0; k < DEVIL_STAGE_TEST_DE; ++k) ctx->stage_ms[k] = 0.0;
  for (int k = 0; k < DEVIL_N_KERNELS; ++k) {
    ctx->k_n[k] = 0;
    ctx->k_ms[k] = ctx->k_bytes[k] = 0.0;
  }
  StageTimer tm(ctx);
  hipEvent_t ev;
  for (int64_t g0 = 0; g0 < G; g0 += batch) {
    const int64_t Gb = std::min(batch, G - g0);
    dv::Counts c;
    c.dtype = pr->counts_dtype;
    c.ld = pr->counts_ld;
    if (mem == DEVIL_MEM_HOST) {
      DV_HIP(hipMemcpyAsync(cbuf.p,
                            static_cast<const char*>(pr->counts) + (size_t)g0 * pr->counts_ld * es,
                            (size_t)Gb * pr->counts_ld * es, hipMemcpyHostToDevice, s));
      c.ptr = cbuf.p;
    } else {
      c.ptr = static_cast<const char*>(pr->counts) + (size_t)g0 * pr->counts_ld * es;
    }
    double* a0 = alpha0.as<double>() + g0;
    double* b0 = beta.as<double>() + g0 * p;
    ev = tm.begin(DEVIL_STAGE_INIT);
    DV_HIP(dv::launch_init(s, c, Gb, n, p, ip, a0, b0));
    DV_HIP(dv::launch_k_from_alpha(s, a0, Gb, kvec.as<double>() + g0));
    tm.end(ev);
    ev = tm.begin(DEVIL_STAGE_IRLS);
    if ((rc = run_irls_tiled(ctx, tl, c, Gb, p, b0, kvec.as<double>() + g0, pr->max_iter,
                             eps, iters.as<int32_t>() + g0, status.as<int32_t>() + g0)))
      return rc;
    tm.end(ev);
    double* th = theta.as<double>() + g0;
    ev = tm.begin(DEVIL_STAGE_DISPERSION);
    if (pr->od_mode == DEVIL_OD_MOM) {
      DV_HIP(dv::launch_mom(s, c, Gb, n, p, dd.X.as<double>(), b0, dd.eo.as<double>(), th,
                            status.as<int32_t>() + g0));
      DV_HIP(dv::launch_fill_i32(s, thit.as<int32_t>() + g0, Gb, 0));
    } else if (pr->od_mode == DEVIL_OD_MLE) {
      if ((rc = run_mle_tiled(ctx, tl.td, c, Gb, n, p, b0, pr->max_iter, th,
                              thit.as<int32_t>() + g0, status.as<int32_t>() + g0)))
        return rc;
    } else {
      DV_HIP(hipMemsetAsync(th, 0, sizeof(double) * Gb, s));
      DV_HIP(dv::launch_fill_i32(s, thit.as<int32_t>() + g0, Gb, 0));
    }
    tm.end(ev);
    ev = tm.begin(DEVIL_STAGE_BREAD_MEAT);
    DV_HIP(dv::launch_bread_meat(s, c, Gb, n, p, dd.X.as<double>(), b0, th,
                                 dd.eo.as<double>(), C > 0 ? dd.ends.as<int32_t>() : nullptr,
                                 C, bread.as<double>() + g0 * p * p,
                                 want_sw ? sandwich.as<double>() + g0 * p * p : nullptr,
                                 status.as<int32_t>() + g0));
    tm.end(ev);
  }

  // outputs: beta back to R's column-major layout
  DevBuf bcol;
  DV_ALLOC(bcol, sizeof(double) * G * p);
  DV_HIP(dv::launch_transpose_gp(s, beta.as<double>(), G, p, 0, bcol.as<double>()));
  const hipMemcpyKind kind =
      mem == DEVIL_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
  DV_HIP(hipMemcpyAsync(out->beta, bcol.p, sizeof(double) * G * p, kind, s));
  DV_HIP(hipMemcpyAsync(out->theta, theta.p, sizeof(double) * G, kind, s));
  DV_HIP(hipMemcpyAsync(out->beta_iters, iters.p, sizeof(int32_t) * G, kind, s));
  DV_HIP(hipMemcpyAsync(out->theta_iters, thit.p, sizeof(int32_t) * G, kind, s));
  DV_HIP(hipMemcpyAsync(out->bread, bread.p, sizeof(double) * G * p * p, kind, s));
  if (want_sw)
    DV_HIP(hipMemcpyAsync(out->sandwich, sandwich.p, sizeof(double) * G * p * p, kind, s));
  if (out->status)
    DV_HIP(hipMemcpyAsync(out->status, status.p, sizeof(int32_t) * G, kind, s));
  DV_HIP(hipStreamSynchronize(s));
  tm.collect();
  return DEVIL_OK;
}

int devil_set_timing(devil_ctx* ctx, int enable) {
  if (!ctx) return DEVIL_E_ARG;
  ctx->timing = enable ? 1 : 0;
  return DEVIL_OK;
}

int devil_stage_times(const devil_ctx* ctx, double* ms) {
  if (!ctx || !ms) return DEVIL_E_ARG;
  for (int k = 0; k < DEVIL_N_STAGES; ++k) ms[k] = ctx->stage_ms[k];
  return DEVIL_OK;
}

int devil_set_tuning(devil_ctx* ctx, int32_t key, int64_t value) {
  if (!ctx || key < 0 || key >= DEVIL_N_TUNE) return DEVIL_E_ARG;
  if (key == DEVIL_TUNE_MIXED_KERNEL && value != DEVIL_MIXED_F32 && value != DEVIL_MIXED_BF16)
    return fail(ctx, DEVIL_E_ARG, "unknown mixed kernel");
  ctx->tune[key] = (int)value;
  return DEVIL_OK;
}

int devil_kernel_stats(const devil_ctx* ctx, int32_t kernel, int64_t* launches, double* ms,
                       double* bytes) {
  if (!ctx || kernel < 0 || kernel >= DEVIL_N_KERNELS) return DEVIL_E_ARG;
  if (launches) *launches = ctx->k_n[kernel];
  if (ms) *ms = ctx->k_ms[kernel];
  if (bytes) *bytes = ctx->k_bytes[kernel];
  return DEVIL_OK;
}

// ---------------------------------------------------------------------------
// stage entry points (device pointers)
// ---------------------------------------------------------------------------

int devil_estimate_dispersion(devil_ctx* ctx, const void* counts, int32_t dtype, int64_t G,
                              int64_t n, int64_t ld, const double* offset, double* alpha0) {
  int rc;
  if ((rc = set_device(ctx))) return rc;
  if ((rc = check_counts_args(ctx, counts, dtype, G, n, ld))) return rc;
  if (!offset || (G > 0 && !alpha0)) return fail(ctx, DEVIL_E_ARG, "NULL pointer");
  std::vector<double> off(n);
  DV_HIP(hipMemcpy(off.data(), offset, sizeof(double) * n, hipMemcpyDeviceToHost));
  double se = 0.0;
  for (int64_t i = 0; i < n; ++i) se += std::exp(off[i]);
  dv::InitParams ip{};
  ip.want_alpha = 1;
  ip.moi = 1.0 / (se / (double)n);
  ip.init_mode = dv::kInitNone;
  dv::Counts c{counts, dtype, ld};
  DV_HIP(dv::launch_init(ctx->stream, c, G, n, 1, ip, alpha0, nullptr));
  DV_HIP(hipStreamSynchronize(ctx->stream));
  return DEVIL_OK;
}

int devil_init_beta(devil_ctx* ctx, const void* counts, int32_t dtype, int64_t G, int64_t n,
                    int64_t ld, const double* X, int32_t p, const double* offset,
                    int32_t init_mode, double* beta0) {
  int rc;
  if ((rc = set_device(ctx))) return rc;
  if ((rc = check_p(ctx, p))) return rc;
  if ((rc = check_counts_args(ctx, counts, dtype, G, n, ld))) return rc;
  if (!X || !offset || (G > 0 && !beta0)) return fail(ctx, DEVIL_E_ARG, "NULL pointer");
  if (init_mode < DEVIL_INIT_QR || init_mode > DEVIL_INIT_UNIVARIATE)
    return fail(ctx, DEVIL_E_ARG, "unknown init_mode");
  Design d;
  if ((rc = prepare_design(ctx, X, offset, n, p, DEVIL_MEM_DEVICE, init_mode, d))) return rc;
  DeviceDesign dd;
  if ((rc = upload_design(ctx, d, dd))) return rc;
  DevBuf tmp;
  DV_ALLOC(tmp, sizeof(double) * G * p);
  dv::Counts c{counts, dtype, ld};
  const dv::InitParams ip = init_params(d, dd, 0, init_mode);
  DV_HIP(dv::launch_init(ctx->stream, c, G, n, p, ip, nullptr, tmp.as<double>()));
  DV_HIP(dv::launch_transpose_gp(ctx->stream, tmp.as<double>(), G, p, 0, beta0));
  DV_HIP(hipStreamSynchronize(ctx->stream));
  return DEVIL_OK;
}

int devil_beta_fit(devil_ctx* ctx, const void* counts, int32_t dtype, int64_t G, int64_t n,
                   int64_t ld, const double* X, int32_t p, double* beta, const double* offset,
                   const double* alpha0, int32_t max_iter, double tolerance, int32_t precision,
                   int32_t* iters, int32_t* status) {
  int rc;
  if ((rc = set_device(ctx))) return rc;
  if ((rc = check_p(ctx, p))) return rc;
  if ((rc = check_counts_args(ctx, counts, dtype, G, n, ld))) return rc;
  if (!X || !offset || (G > 0 && (!beta || !alpha0 || !iters)))
    return fail(ctx, DEVIL_E_ARG, "NULL pointer");
  if (max_iter < 0) return fail(ctx, DEVIL_E_ARG, "max_iter < 0");
  if (precision != DEVIL_PREC_FP64 && precision != DEVIL_PREC_MIXED)
    return fail(ctx, DEVIL_E_ARG, "unknown precision");
  if (G == 0) return DEVIL_OK;
  hipStream_t s = ctx->stream;
  DevBuf bg, kv;
  DV_ALLOC(bg, sizeof(double) * G * p);
  DV_ALLOC(kv, sizeof(double) * G);
  DV_HIP(dv::launch_transpose_gp(s, beta, G, p, 1, bg.as<double>()));
  DV_HIP(dv::launch_k_from_alpha(s, alpha0, G, kv.as<double>()));
  dv::Counts c{counts, dtype, ld};
  std::vector<double> Xh((size_t)n * p), offh(n);
  DV_HIP(hipMemcpy(Xh.data(), X, sizeof(double) * n * p, hipMemcpyDeviceToHost));
  DV_HIP(hipMemcpy(offh.data(), offset, sizeof(double) * n, hipMemcpyDeviceToHost));
  TileHost th;
  build_tile_design(Xh.data(), offh.data(), n, p, th,
                    precision == DEVIL_PREC_MIXED ? mixed_kernel(ctx) : kMixedNone);
  DeviceTile dt;
  if ((rc = upload_tile(ctx, th, dt))) return rc;
  DevBuf st_local;
  int32_t* st = status;
  if (!st) {
    DV_ALLOC(st_local, sizeof(int32_t) * G);
    st = st_local.as<int32_t>();
  }
  if ((rc = run_irls_tiled(ctx, th, c, G, p, bg.as<double>(), kv.as<double>(), max_iter,
                           (double)(float)tolerance, iters, st)))
    return rc;
  DV_HIP(dv::launch_transpose_gp(s, bg.as<double>(), G, p, 0, beta));
  DV_HIP(hipStreamSynchronize(s));
  return DEVIL_OK;
}

int devil_mom_dispersion(devil_ctx* ctx, const void* counts, int32_t dtype, int64_t G,
                         int64_t n, int64_t ld, const double* X, int32_t p, const double* beta,
                         const double* offset, double* theta) {
  int rc;
  if ((rc = set_device(ctx))) return rc;
  if ((rc = check_p(ctx, p))) return rc;
  if ((rc = check_counts_args(ctx, counts, dtype, G, n, ld))) return rc;
  if (!X || !offset || (G > 0 && (!beta || !theta)))
    return fail(ctx, DEVIL_E_ARG, "NULL pointer");
  if (G == 0) return DEVIL_OK;
  hipStream_t s = ctx->stream;
  std::vector<double> off(n);
  DV_HIP(hipMemcpy(off.data(), offset, sizeof(double) * n, hipMemcpyDeviceToHost));
  for (auto& v : off) v = std::exp(v);
  DevBuf eo, bg;
  if ((rc = upload(ctx, eo, off))) return rc;
  DV_ALLOC(bg, sizeof(double) * G * p);
  DV_HIP(dv::launch_transpose_gp(s, beta, G, p, 1, bg.as<double>()));
  dv::Counts c{counts, dtype, ld};
  DV_HIP(dv::launch_mom(s, c, G, n, p, X, bg.as<double>(), eo.as<double>(), theta, nullptr));
  DV_HIP(hipStreamSynchronize(s));
  return DEVIL_OK;
}

int devil_fit_dispersion_mle(devil_ctx* ctx, const void* counts, int32_t dtype, int64_t G,
                             int64_t n, int64_t ld, const double* X, int32_t p,
                             const double* beta, const double* offset, int32_t max_iter,
                             double* theta, int32_t* iters) {
  int rc;
  if ((rc = set_device(ctx))) return rc;
  if ((rc = check_p(ctx, p))) return rc;
  if ((rc = check_counts_args(ctx, counts, dtype, G, n, ld))) return rc;
  if (!X || !offset || (G > 0 && (!beta || !theta || !iters)))
    return fail(ctx, DEVIL_E_ARG, "NULL pointer");
  if (G == 0) return DEVIL_OK;
  hipStream_t s = ctx->stream;
  DevBuf bg, st;
  DV_ALLOC(bg, sizeof(double) * G * p);
  DV_ALLOC(st, sizeof(int32_t) * G);
  DV_HIP(dv::launch_transpose_gp(s, beta, G, p, 1, bg.as<double>()));
  DV_HIP(dv::launch_fill_i32(s, st.as<int32_t>(), G, 0));
  dv::Counts c{counts, dtype, ld};
  std::vector<double> Xh((size_t)n * p), offh(n);
  DV_HIP(hipMemcpy(Xh.data(), X, sizeof(double) * n * p, hipMemcpyDeviceToHost));
  DV_HIP(hipMemcpy(offh.data(), offset, sizeof(double) * n, hipMemcpyDeviceToHost));
  TileHost th;
  build_tile_design(Xh.data(), offh.data(), n, p, th);
  DeviceTile dt;
  if ((rc = upload_tile(ctx, th, dt))) return rc;
  if ((rc = run_mle_tiled(ctx, th.td, c, G, n, p, bg.as<double>(), max_iter, theta, iters,
                          st.as<int32_t>())))
    return rc;
  DV_HIP(hipStreamSynchronize(s));
  return DEVIL_OK;
}

int devil_bread_meat(devil_ctx* ctx, const void* counts, int32_t dtype, int64_t G, int64_t n,
                     int64_t ld, const double* X, int32_t p, const double* beta,
                     const double* theta, const double* offset, const int32_t* cluster_ends,
                     int32_t n_clusters, double* bread, double* sandwich, int32_t* status) {
  int rc;
  if ((rc = set_device(ctx))) return rc;
  if ((rc = check_p(ctx, p))) return rc;
  if ((rc = check_counts_args(ctx, counts, dtype, G, n, ld))) return rc;
  if (!X || !offset || (G > 0 && (!beta || !theta || !bread)))
    return fail(ctx, DEVIL_E_ARG, "NULL pointer");
  if (n_clusters < 0 || (n_clusters > 0 && !cluster_ends))
    return fail(ctx, DEVIL_E_ARG, "n_clusters > 0 needs cluster_ends");
  if (G == 0) return DEVIL_OK;
  hipStream_t s = ctx->stream;
  if (n_clusters > 0) {
    std::vector<int32_t> e(n_clusters);
    DV_HIP(hipMemcpy(e.data(), cluster_ends, sizeof(int32_t) * n_clusters,
                     hipMemcpyDeviceToHost));
    if ((rc = check_clusters(ctx, e.data(), n_clusters, n))) return rc;
  }
  std::vector<double> off(n);
  DV_HIP(hipMemcpy(off.data(), offset, sizeof(double) * n, hipMemcpyDeviceToHost));
  for (auto& v : off) v = std::exp(v);
  DevBuf eo, bg;
  if ((rc = upload(ctx, eo, off))) return rc;
  DV_ALLOC(bg, sizeof(double) * G * p);
  DV_HIP(dv::launch_transpose_gp(s, beta, G, p, 1, bg.as<double>()));
  dv::Counts c{counts, dtype, ld};
  DV_HIP(dv::launch_bread_meat(s, c, G, n, p, X, bg.as<double>(), theta, eo.as<double>(),
                               cluster_ends, n_clusters, bread,
                               n_clusters > 0 ? sandwich : nullptr, status));
  DV_HIP(hipStreamSynchronize(s));
  return DEVIL_OK;
}

int devil_test_de(devil_ctx* ctx, const double* beta, const double* bread,
                  const double* sandwich, int64_t G, int32_t p, int64_t n,
                  const double* contrast, double max_lfc, double* pval, double* log_pval,
                  double* adj_pval, double* lfc) {
  int rc;
  if ((rc = set_device(ctx))) return rc;
  if ((rc = check_p(ctx, p))) return rc;
  if (G < 0 || (G > 0 && (!beta || !bread || !contrast || !pval || !adj_pval || !lfc)))
    return fail(ctx, DEVIL_E_ARG, "NULL pointer or G < 0");
  if (G == 0) return DEVIL_OK;
  hipStream_t s = ctx->stream;
  DevBuf bg, rank;
  DV_ALLOC(bg, sizeof(double) * G * p);
  DV_ALLOC(rank, sizeof(int32_t) * G);
  ctx->stage_ms[DEVIL_STAGE_TEST_DE] = 0.0;
  StageTimer tm(ctx);
  hipEvent_t ev = tm.begin(DEVIL_STAGE_TEST_DE);
  DV_HIP(dv::launch_transpose_gp(s, beta, G, p, 1, bg.as<double>()));
  DV_HIP(dv::launch_test_de(s, bg.as<double>(), bread, sandwich, G, p, (double)n - 2.0,
                            contrast, max_lfc, pval, log_pval, lfc, rank.as<int32_t>(),
                            adj_pval));
  tm.end(ev);
  DV_HIP(hipStreamSynchronize(s));
  tm.collect();
  return DEVIL_OK;
}

}  // extern "C"
