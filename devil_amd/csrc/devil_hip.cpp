// C ABI of the gfx950 NB-GLM fit (include/devil_hip.h).
//
// devil_fit runs R/main.R:375-523 (cpu_fit) on the GPU:
//   init pass (alpha0, beta0) -> k = float(1/float(alpha0)) -> IRLS ->
//   dispersion (MOM | MLE | none) -> bread + clustered meat pass.
// Per-gene intermediates live on the device in gene-major layout; host-memory
// problems stream their counts in gene batches sized to the free HBM.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <memory>
#include <atomic>
#include <map>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/devil_hip.h"
#include "kernels.hpp"
#include "newod_driver.hpp"
#include "port_driver.hpp"

// Device blocks of a context kept for reuse after the scope that used them
// ends, so repeated fits do not pay hipMalloc / hipFree (which synchronises
// the device).  Every access to a context's blocks is ordered on that
// context's stream, so work queued by the next call cannot touch a reused
// block before the earlier call's work on it has finished.
struct BlockCache {
  static constexpr size_t kMaxCached = size_t(6) << 30;
  std::multimap<size_t, void*> blocks;  // size -> block
  size_t cached = 0;
  // the smallest cached block of at least bytes (and not far larger)
  void* take(size_t bytes, size_t& got) {
    auto it = blocks.lower_bound(bytes);
    if (it == blocks.end() || it->first > 2 * bytes + (size_t(1) << 20)) return nullptr;
    void* p = it->second;
    got = it->first;
    cached -= got;
    blocks.erase(it);
    return p;
  }
  void give(void* p, size_t bytes) {
    if (cached + bytes > kMaxCached) {
      (void)hipFree(p);
      return;
    }
    blocks.emplace(bytes, p);
    cached += bytes;
  }
  void clear() {
    for (auto& b : blocks) (void)hipFree(b.second);
    blocks.clear();
    cached = 0;
  }
};

struct devil_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  void* scratch = nullptr;  // the device's no-allocation scratch array (its address)
  BlockCache cache;
  std::string err;
  int timing = 0;
  // kernel, chunk cells (0 auto), streams, MLE driver, aligned copies allowed
  int tune[DEVIL_N_TUNE] = {DEVIL_MIXED_AUTO, 0, 8, DEVIL_MLE_NLMINB, 1, 0, 0, 0};
  std::map<uint64_t, int> tuned;      // autotuner choices by problem shape
  std::map<uint64_t, int> tuned_joint;  // the same for the joint rounds (kernel fixed)
  std::vector<devil_ctx*> subs;       // per-stream contexts of multi-problem fits
  int64_t replayed = 0;                // genes refit in FP64 by the mixed path (last fit)
  int64_t polished = 0;                // genes whose last mixed step was redone in FP64
  int64_t replay_sum = 0;
  int64_t polish_sum = 0;
  int aligned_copy = 0;                // the last fit read an aligned copy of its counts
  // the aligned copy's buffer, kept for the next fit (a fresh hipMalloc of
  // tens of GB costs far more than the copy itself); freed by devil_destroy
  void* pad_p = nullptr;
  size_t pad_size = 0;
  int last_mk = 0;                     // mixed kernel and chunk cells of the last fit
  int64_t last_cc = 0;
  int64_t k_sum_n[DEVIL_N_KERNELS] = {0, 0};
  double k_sum_ms[DEVIL_N_KERNELS] = {0, 0};
  double k_sum_bytes[DEVIL_N_KERNELS] = {0, 0};
  double stage_ms[DEVIL_N_STAGES] = {0, 0, 0, 0, 0};
  // per-launch records of the IRLS accumulate kernels (timing on)
  struct Launch {
    hipEvent_t a, b;
    int kind;
    double bytes;
  };
  std::vector<hipEvent_t> ev_pool;
  std::vector<Launch> launches;
  int64_t k_n[DEVIL_N_KERNELS] = {};
  double k_ms[DEVIL_N_KERNELS] = {};
  double k_bytes[DEVIL_N_KERNELS] = {};
  // IRLS round counts read back without a stream synchronisation
  int32_t* round_counts = nullptr;  // pinned host slots
  int32_t* round_counts_dev = nullptr;  // the same slots as the device addresses them
  hipEvent_t round_ev[2] = {nullptr, nullptr};
  int64_t joined = 0;                 // problems of the last fit in the joint rounds
  // joint rounds of many problems: pinned counts and round descriptors
  void* joint_host = nullptr;
  void* joint_host_dev = nullptr;
  size_t joint_host_bytes = 0;
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

// Device allocation owned by the caller's scope; its block returns to the
// context's cache when the scope ends.
struct DevBuf {
  void* p = nullptr;
  size_t size = 0;
  BlockCache* cache = nullptr;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), size(o.size), cache(o.cache) { o.p = nullptr; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      size = o.size;
      cache = o.cache;
      o.p = nullptr;
    }
    return *this;
  }
  ~DevBuf() { release(); }
  void release() {
    if (!p) return;
    if (cache) cache->give(p, size);
    else (void)hipFree(p);
    p = nullptr;
  }
  hipError_t alloc(BlockCache* c, size_t bytes) {
    release();
    bytes = 256 * ((std::max<size_t>(bytes, 8) + 255) / 256);
    cache = c;
    size = bytes;
    if (c && (p = c->take(bytes, size))) return hipSuccess;
    hipError_t e = hipMalloc(&p, bytes);
    if (e == hipErrorOutOfMemory && c && c->cached) {  // cached blocks back to the device
      c->clear();
      e = hipMalloc(&p, bytes);
    }
    if (e != hipSuccess) p = nullptr;
    return e;
  }
  template <typename T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

#define DV_ALLOC(buf, bytes)                                                             \
  do {                                                                                   \
    hipError_t e_ = (buf).alloc(&ctx->cache, bytes);                                     \
    if (e_ != hipSuccess)                                                                \
      return fail(ctx, DEVIL_E_NOMEM, std::string("hipMalloc ") + std::to_string(bytes) + \
                                          " bytes: " + hipGetErrorString(e_));           \
  } while (0)

size_t count_elem_size(int dtype) { return dtype == DEVIL_COUNTS_F64 ? 8 : 4; }

// Host loops over cells split across threads ([lo, hi) ranges).
template <typename F>
void parallel_for(int64_t n, F&& fn) {
  const int64_t hw = std::max<int64_t>(1, std::thread::hardware_concurrency());
  const int64_t T = std::min<int64_t>({hw, 16, std::max<int64_t>(1, n / 65536)});
  if (T <= 1) {
    fn((int64_t)0, n);
    return;
  }
  std::vector<std::thread> th;
  th.reserve(T);
  for (int64_t t = 0; t < T; ++t) th.emplace_back([&, t] { fn(n * t / T, n * (t + 1) / T); });
  for (auto& x : th) x.join();
}

// Upper Cholesky factor of the p x p SPD matrix A (column-major): A = R'R.
bool cholesky_upper(const double* A, int p, double* R) {
  for (int j = 0; j < p; ++j)
    for (int i = 0; i < p; ++i) R[i + j * p] = 0.0;
  for (int j = 0; j < p; ++j) {
    double d = A[j + j * p];
    for (int k = 0; k < j; ++k) d -= R[k + j * p] * R[k + j * p];
    if (!(d > 0.0)) return false;
    const double r = std::sqrt(d);
    R[j + j * p] = r;
    for (int i = j + 1; i < p; ++i) {
      double s = A[j + i * p];
      for (int k = 0; k < j; ++k) s -= R[k + j * p] * R[k + i * p];
      R[j + i * p] = s / r;
    }
  }
  return true;
}

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


template <typename T>
int upload(devil_ctx* ctx, DevBuf& b, const std::vector<T>& v) {
  DV_ALLOC(b, sizeof(T) * v.size());
  if (!v.empty())
    DV_HIP(hipMemcpyAsync(b.p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice,
                          ctx->stream));
  return DEVIL_OK;
}

int upload_raw(devil_ctx* ctx, DevBuf& b, const void* src, size_t bytes, hipMemcpyKind kind) {
  DV_ALLOC(b, bytes);
  if (bytes) DV_HIP(hipMemcpyAsync(b.p, src, bytes, kind, ctx->stream));
  return DEVIL_OK;
}

// Host copies of the small per-cell inputs plus what the init pass needs.
struct Design {
  int64_t n = 0;
  int p = 0;
  std::vector<double> X;  // host copy of X, made only when a host step needs it
  std::vector<double> off, eo, logsf, R, xmean, denom;
  double moi = 0.0, sum_sf = 0.0;
  int intercept_first = 0;
};

// Device copies of the design, offsets and init operands.
struct DeviceDesign {
  DevBuf X, off, eo, Q, R, logsf, xmean, denom;
  const double* x = nullptr;  // device X (the caller's own when it is device memory)
};

void combine_gram(const std::vector<double>& part, int p, std::vector<double>& G) {
  G.assign((size_t)p * p, 0.0);
  int q = 0;
  for (int a = 0; a < p; ++a)
    for (int b = a; b < p; ++b, ++q) {
      long double su = 0.0L;
      for (int x = 0; x < dv::kStatBlocksHost; ++x) su += part[x * 55 + q];
      G[a + b * p] = G[b + a * p] = (double)su;
    }
}

// Thin QR of the device X (n x p) by Cholesky QR applied twice on the GPU:
// X'X = R1'R1, Q1 = X R1^-1, Q1'Q1 = R2'R2, Q = Q1 R2^-1, R = R2 R1.  ok is
// false when a Gram matrix is not positive definite or a diagonal of R is
// below 1e-7 of the largest (the rank tolerance of R's qr(), LINPACK dqrdc2).
int qr_device(devil_ctx* ctx, const double* Xd, int64_t n, int p, DevBuf& Q,
              std::vector<double>& R, bool& ok) {
  hipStream_t s = ctx->stream;
  ok = false;
  DV_ALLOC(Q, sizeof(double) * n * p);
  DV_HIP(hipMemcpyAsync(Q.p, Xd, sizeof(double) * n * p, hipMemcpyDeviceToDevice, s));
  DevBuf part, Rd;
  DV_ALLOC(part, sizeof(double) * dv::kStatBlocksHost * 55);
  DV_ALLOC(Rd, sizeof(double) * p * p);
  std::vector<double> hp((size_t)dv::kStatBlocksHost * 55), G, R1((size_t)p * p),
      R2((size_t)p * p);
  auto gram = [&]() -> int {
    DV_HIP(dv::launch_gram(s, Q.as<double>(), n, p, part.as<double>()));
    DV_HIP(hipMemcpyAsync(hp.data(), part.p, sizeof(double) * hp.size(), hipMemcpyDeviceToHost,
                          s));
    DV_HIP(hipStreamSynchronize(s));
    combine_gram(hp, p, G);
    return DEVIL_OK;
  };
  int rc;
  if ((rc = gram())) return rc;
  if (!cholesky_upper(G.data(), p, R1.data())) return DEVIL_OK;
  DV_HIP(hipMemcpyAsync(Rd.p, R1.data(), sizeof(double) * p * p, hipMemcpyHostToDevice, s));
  DV_HIP(dv::launch_right_solve(s, Q.as<double>(), n, p, Rd.as<double>()));
  if ((rc = gram())) return rc;
  if (!cholesky_upper(G.data(), p, R2.data())) return DEVIL_OK;
  DV_HIP(hipMemcpyAsync(Rd.p, R2.data(), sizeof(double) * p * p, hipMemcpyHostToDevice, s));
  DV_HIP(dv::launch_right_solve(s, Q.as<double>(), n, p, Rd.as<double>()));
  DV_HIP(hipStreamSynchronize(s));
  R.assign((size_t)p * p, 0.0);
  for (int j = 0; j < p; ++j)
    for (int i = 0; i <= j; ++i) {
      double su = 0.0;
      for (int k = i; k <= j; ++k) su += R2[i + k * p] * R1[k + j * p];
      R[i + j * p] = su;
    }
  double dmax = 0.0;
  for (int a = 0; a < p; ++a) dmax = std::max(dmax, std::fabs(R[a + a * p]));
  ok = true;
  for (int a = 0; a < p; ++a)
    if (!(std::fabs(R[a + a * p]) > 1e-7 * dmax)) ok = false;
  return DEVIL_OK;
}

int host_copy_of_x(devil_ctx* ctx, const double* X, int mem, Design& d) {
  if (!d.X.empty()) return DEVIL_OK;
  d.X.resize((size_t)d.n * d.p);
  if (mem == DEVIL_MEM_DEVICE)
    DV_HIP(hipMemcpy(d.X.data(), X, sizeof(double) * d.n * d.p, hipMemcpyDeviceToHost));
  else
    std::memcpy(d.X.data(), X, sizeof(double) * d.n * d.p);
  return DEVIL_OK;
}

// Offsets, exp(offset) and the init operands on host and device; X stays on
// the device (the caller's pointer when it is device memory).
int prepare_design(devil_ctx* ctx, const double* X, const double* off, int64_t n, int p,
                   int mem, int init_mode, Design& d, DeviceDesign& dd) {
  d.n = n;
  d.p = p;
  d.off.resize(n);
  if (mem == DEVIL_MEM_DEVICE)
    DV_HIP(hipMemcpy(d.off.data(), off, sizeof(double) * n, hipMemcpyDeviceToHost));
  else
    std::memcpy(d.off.data(), off, sizeof(double) * n);
  d.eo.resize(n);
  parallel_for(n, [&](int64_t lo, int64_t hi) {
    for (int64_t i = lo; i < hi; ++i) d.eo[i] = std::exp(d.off[i]);  // exp_offset, R/main.R:390
  });
  double se = 0.0;
  for (int64_t i = 0; i < n; ++i) se += d.eo[i];
  d.moi = 1.0 / (se / (double)n);  // R/overdispersion.R:130
  int rc;
  if (mem == DEVIL_MEM_DEVICE) {
    dd.x = X;
  } else {
    if ((rc = upload_raw(ctx, dd.X, X, sizeof(double) * n * p, hipMemcpyHostToDevice))) return rc;
    dd.x = dd.X.as<double>();
  }
  if ((rc = upload(ctx, dd.off, d.off))) return rc;
  if ((rc = upload(ctx, dd.eo, d.eo))) return rc;
  if (init_mode == DEVIL_INIT_QR) {
    bool ok = false;
    if ((rc = qr_device(ctx, dd.x, n, p, dd.Q, d.R, ok))) return rc;
    if (!ok) {  // ill-conditioned for Cholesky QR: Householder on the host
      if ((rc = host_copy_of_x(ctx, X, mem, d))) return rc;
      std::vector<double> Qh;
      if (!householder_qr(d.X.data(), n, p, Qh, d.R))
        return fail(ctx, DEVIL_E_RANK, "design matrix is rank deficient (QR init, R/beta.R:14)");
      if ((rc = upload(ctx, dd.Q, Qh))) return rc;
      DV_HIP(hipStreamSynchronize(ctx->stream));
    }
    if ((rc = upload(ctx, dd.R, d.R))) return rc;
  } else if (init_mode == DEVIL_INIT_UNIVARIATE) {
    // src/beta_init_univariate.cpp:48-84
    if ((rc = host_copy_of_x(ctx, X, mem, d))) return rc;
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
      double su = 0.0;
      for (int64_t i = 0; i < n; ++i) su += d.X[i + j * n];
      d.xmean[j] = su / (double)n;
      double q = 0.0;
      for (int64_t i = 0; i < n; ++i) {
        const double c = d.X[i + j * n] - d.xmean[j];
        q += c * c;
      }
      d.denom[j] = q <= 0.0 ? 1e-12 : q;
    }
    if ((rc = upload(ctx, dd.logsf, d.logsf))) return rc;
    if ((rc = upload(ctx, dd.xmean, d.xmean))) return rc;
    if ((rc = upload(ctx, dd.denom, d.denom))) return rc;
  }
  return DEVIL_OK;
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
    if (hipEventCreateWithFlags(&a, hipEventDisableSystemFence) != hipSuccess ||
        hipEventCreateWithFlags(&b, hipEventDisableSystemFence) != hipSuccess)
      return nullptr;
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
enum MixedKernel { kMixedNone = 0, kMixedF32 = 1, kMixedBf16 = 2, kMixedBf16Ws = 3,
                   kMixedBf16Ws2 = 4 };

// 4-byte counts in 16-byte aligned rows (what the LDS-staged kernels load).
bool counts_vec4(const dv::Counts& c) {
  return c.dtype != DEVIL_COUNTS_F64 && (c.ld & 3) == 0 &&
         (reinterpret_cast<uintptr_t>(c.ptr) & 15) == 0;
}

// 4-byte counts (what the LDS-staged kernels read, 16 or 4 bytes per load)
bool counts_ws(const dv::Counts& c) { return c.dtype != DEVIL_COUNTS_F64; }

// 16-gene tiles per workgroup of a round's accumulate kernel.
int accum_tiles_per_wg(int kind, int p, const dv::Counts& c) {
  if (!counts_ws(c)) return 4;
  if (kind == kMixedBf16Ws) return dv::ws_tiles_per_workgroup(p, 1);
  if (kind == kMixedBf16Ws2) return dv::ws_tiles_per_workgroup(p, counts_vec4(c) ? 2 : 1);
  return 4;
}

int mixed_kernel(const devil_ctx* ctx) {
  switch (ctx->tune[DEVIL_TUNE_MIXED_KERNEL]) {
    case DEVIL_MIXED_BF16: return kMixedBf16;
    case DEVIL_MIXED_BF16_WS: return kMixedBf16Ws;
    case DEVIL_MIXED_BF16_WS2: return kMixedBf16Ws2;
    default: return kMixedF32;
  }
}

// the bf16 kernels share one operand layout (32-cell steps)
bool bf16_layout(int mixed) {
  return mixed == kMixedBf16 || mixed == kMixedBf16Ws || mixed == kMixedBf16Ws2;
}

hipError_t launch_mixed_accum(int kind, hipStream_t s, const dv::TileDesign& td, dv::Counts c,
                              int p, const int32_t* cur, const int32_t* ncur, int nA,
                              const double* beta, const double* kvec, double* part) {
  if (kind == kMixedBf16)
    return dv::launch_irls_accum_bf16(s, td, c, p, cur, ncur, nA, beta, kvec, part);
  if (kind == kMixedBf16Ws)
    return dv::launch_irls_accum_ws(s, td, c, p, cur, ncur, nA, beta, kvec, part, 1);
  if (kind == kMixedBf16Ws2)
    return dv::launch_irls_accum_ws(s, td, c, p, cur, ncur, nA, beta, kvec, part, 2);
  return dv::launch_irls_accum_mixed(s, td, c, p, cur, ncur, nA, beta, kvec, part);
}

// Columns of K: x_a x_b (a <= b) and x_a of the design, exact duplicates
// merged (an intercept makes x_0 x_a == x_a, a 0/1 column squares to itself)
// and all-zero columns dropped.  lin16: the p design columns sit at 0..p-1
// as they are and the pair columns start at column 16.
struct ColPlan {
  std::vector<dv::ColSpec> spec;
  std::vector<double> ksum;  // sum over cells of each column
  int kidx_pair[55];
  int kidx_lin[10];
};

// Statistics of every candidate column computed on the device (Xd: device
// n x p X); equal hashes are confirmed by a cell-by-cell comparison.
int plan_columns(devil_ctx* ctx, const double* Xd, int64_t n, int p, bool lin16, ColPlan& cp) {
  hipStream_t s = ctx->stream;
  std::vector<dv::ColSpec> cand;
  for (int a = 0; a < p; ++a) cand.push_back({a, -1});
  for (int a = 0; a < p; ++a)
    for (int b = a; b < p; ++b) cand.push_back({a, b});
  const int C = (int)cand.size();
  const int B = dv::kStatBlocksHost;
  DevBuf dc, dh, dnz, dsum, dmm;
  int rc;
  if ((rc = upload(ctx, dc, cand))) return rc;
  DV_ALLOC(dh, sizeof(uint64_t) * C * B);
  DV_ALLOC(dnz, sizeof(uint64_t) * C * B);
  DV_ALLOC(dsum, sizeof(double) * C * B);
  DV_ALLOC(dmm, sizeof(uint64_t) * B);
  DV_HIP(dv::launch_col_stats(s, Xd, n, dc.as<dv::ColSpec>(), C, dh.as<uint64_t>(),
                              dnz.as<uint64_t>(), dsum.as<double>()));
  std::vector<uint64_t> ph((size_t)C * B), pnz((size_t)C * B);
  std::vector<double> psum((size_t)C * B);
  DV_HIP(hipMemcpyAsync(ph.data(), dh.p, sizeof(uint64_t) * C * B, hipMemcpyDeviceToHost, s));
  DV_HIP(hipMemcpyAsync(pnz.data(), dnz.p, sizeof(uint64_t) * C * B, hipMemcpyDeviceToHost, s));
  DV_HIP(hipMemcpyAsync(psum.data(), dsum.p, sizeof(double) * C * B, hipMemcpyDeviceToHost, s));
  DV_HIP(hipStreamSynchronize(s));
  std::vector<uint64_t> hash(C, 0), nonzero(C, 0);
  std::vector<double> sum(C, 0.0);
  for (int c = 0; c < C; ++c) {
    long double su = 0.0L;
    for (int x = 0; x < B; ++x) {
      hash[c] += ph[(size_t)c * B + x];
      nonzero[c] += pnz[(size_t)c * B + x];
      su += psum[(size_t)c * B + x];
    }
    sum[c] = (double)su;
  }
  std::vector<int> members;  // candidate behind each column (-1: padding)
  std::vector<uint64_t> mm(B);
  int err = DEVIL_OK;
  auto equal = [&](int c1, int c2) -> bool {
    if (dv::launch_col_mismatch(s, Xd, n, cand[c1], cand[c2], dmm.as<uint64_t>()) != hipSuccess ||
        hipMemcpyAsync(mm.data(), dmm.p, sizeof(uint64_t) * B, hipMemcpyDeviceToHost, s) !=
            hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
      err = fail(ctx, DEVIL_E_HIP, "column comparison failed");
      return false;
    }
    uint64_t tot = 0;
    for (uint64_t v : mm) tot += v;
    return tot == 0;
  };
  auto add = [&](int c) -> int {
    if (!nonzero[c]) return -1;
    for (size_t k = 0; k < cp.spec.size(); ++k)
      if (members[k] >= 0 && hash[members[k]] == hash[c] && equal(members[k], c)) return (int)k;
    cp.spec.push_back(cand[c]);
    cp.ksum.push_back(sum[c]);
    members.push_back(c);
    return (int)cp.spec.size() - 1;
  };
  if (lin16) {
    for (int a = 0; a < p; ++a) {
      cp.spec.push_back(cand[a]);
      cp.ksum.push_back(sum[a]);
      members.push_back(a);
      cp.kidx_lin[a] = a;
    }
    while (cp.spec.size() < 16) {
      cp.spec.push_back({-1, -1});
      cp.ksum.push_back(0.0);
      members.push_back(-1);
    }
  } else {
    for (int a = 0; a < p; ++a) cp.kidx_lin[a] = add(a);
  }
  for (int q = 0; q < p * (p + 1) / 2; ++q) cp.kidx_pair[q] = add(p + q);
  return err;
}

// Chunks of about chunk_blocks blocks.  Chunk boundaries fall on multiples of
// kChunkAlign blocks, where the mixed kernels' FP32 spans restart, so every
// chunking sums the same FP32 spans and differs only in the order of FP64
// sums.  A multiple of 8 chunks: with blocks dealt round-robin to the 8 XCDs
// and chunk = blockIdx % nchunk, each chunk's operands then live in one XCD's
// L2 (trailing chunks may come out empty).
constexpr int64_t kChunkAlign = 4;

void set_chunks(dv::TileDesign& td, int64_t chunk_blocks) {
  auto align = [](int64_t b) { return kChunkAlign * ((b + kChunkAlign - 1) / kChunkAlign); };
  td.chunk_blocks = align(std::max<int64_t>(1, chunk_blocks));
  td.nchunk = (int)((td.nblk + td.chunk_blocks - 1) / td.chunk_blocks);
  if (td.nchunk >= 8) {
    td.nchunk = 8 * ((td.nchunk + 7) / 8);
    td.chunk_blocks = align((td.nblk + td.nchunk - 1) / td.nchunk);
  }
}

// Shape of a design with cells_per_block-cell blocks over the given columns.
void shape_design(const ColPlan& cp, int64_t n, int p, int cells_per_block, dv::TileDesign& td,
                  int64_t chunk_cells = 2048) {
  std::memset(&td, 0, sizeof(td));
  for (int a = 0; a < p; ++a) td.kidx_lin[a] = cp.kidx_lin[a];
  for (int q = 0; q < p * (p + 1) / 2; ++q) td.kidx_pair[q] = cp.kidx_pair[q];
  for (size_t c = 0; c < cp.ksum.size() && c < (size_t)DEVIL_NCP_MAX; ++c)
    td.ksum[c] = cp.ksum[c];
  const int ncol = std::max<int>(1, (int)cp.spec.size());
  td.nt = (ncol + 15) / 16;
  td.ncp = 16 * td.nt;
  td.n = n;
  td.nblk = (n + cells_per_block - 1) / cells_per_block;
  const int64_t min_blocks = std::max<int64_t>(1, chunk_cells / cells_per_block);
  set_chunks(td, std::max<int64_t>(min_blocks, (td.nblk + 63) / 64));
}


// Operand tables of a fit's design on the device (built by kernels from the
// device copies of X, offset and exp(offset)).
struct TileHost {
  dv::TileDesign td{};   // FP64 design (16-cell blocks, packed columns)
  dv::TileDesign tdm{};  // mixed design (td for FP32, 32-cell steps for bf16)
  int mixed = kMixedNone;
  DevBuf spec, specm, Xp, offp, eop, Kp, Xf, offf, Kf, xab, offl, kg, kh, kh2;
};

// Xd, offd, eod: device n x p X, offset and exp(offset).
int make_tile(devil_ctx* ctx, const double* Xd, const double* offd, const double* eod,
              int64_t n, int p, int mixed, TileHost& th, int64_t chunk_cells = 2048) {
  hipStream_t s = ctx->stream;
  th.mixed = mixed;
  ColPlan cp;
  int rc;
  if ((rc = plan_columns(ctx, Xd, n, p, false, cp))) return rc;
  dv::TileDesign& td = th.td;
  shape_design(cp, n, p, 16, td, chunk_cells);
  if ((rc = upload(ctx, th.spec, cp.spec))) return rc;
  const int NJ = (p + 3) / 4;
  DV_ALLOC(th.Xp, sizeof(double) * td.nblk * NJ * 64);
  DV_ALLOC(th.offp, sizeof(double) * td.nblk * 16);
  DV_ALLOC(th.eop, sizeof(double) * td.nblk * 16);
  DV_ALLOC(th.Kp, sizeof(double) * td.nblk * td.nt * 256);
  const bool f32 = mixed == kMixedF32;
  if (f32) {
    DV_ALLOC(th.Xf, sizeof(float) * td.nblk * NJ * 64);
    DV_ALLOC(th.offf, sizeof(float) * td.nblk * 16);
    DV_ALLOC(th.Kf, sizeof(float) * td.nblk * td.nt * 256);
  }
  DV_HIP(dv::launch_build_tables(s, Xd, offd, eod, n, p, th.spec.as<dv::ColSpec>(),
                                 (int)cp.spec.size(), td.nt, td.nblk, th.Xp.as<double>(),
                                 th.offp.as<double>(), th.eop.as<double>(), th.Kp.as<double>(),
                                 f32 ? 1 : 0, th.Xf.as<float>(), th.offf.as<float>(),
                                 th.Kf.as<float>()));
  td.Xp = th.Xp.as<double>();
  td.offp = th.offp.as<double>();
  td.eop = th.eop.as<double>();
  td.Kp = th.Kp.as<double>();
  if (f32) {
    th.tdm = td;
    th.tdm.Xf = th.Xf.as<float>();
    th.tdm.offf = th.offf.as<float>();
    th.tdm.Kf = th.Kf.as<float>();
  } else if (bf16_layout(mixed)) {
    ColPlan cm;
    if ((rc = plan_columns(ctx, Xd, n, p, true, cm))) return rc;
    dv::TileDesign& tm = th.tdm;
    shape_design(cm, n, p, 32, tm, chunk_cells);
    if ((rc = upload(ctx, th.specm, cm.spec))) return rc;
    const int64_t nst = tm.nblk;
    const int NE = dv::eta_blocks(p);
    const int nth = tm.nt - 1;
    DV_ALLOC(th.xab, sizeof(uint16_t) * nst * 2 * NE * 64 * 8);
    DV_ALLOC(th.offl, sizeof(float) * nst * 32);
    DV_ALLOC(th.kg, sizeof(uint16_t) * nst * 3 * 64 * 8);
    // at least one pair tile: the kernels' operand loads are unconditional
    // (a design without pair columns reads zeros it never uses)
    const int64_t nth_alloc = std::max(nth, 1);
    DV_ALLOC(th.kh, sizeof(uint16_t) * nst * nth_alloc * 64 * 8);
    DV_ALLOC(th.kh2, sizeof(uint16_t) * nst * nth_alloc * 64 * 8);
    if (nth == 0) {
      const int64_t words = nst * 64 * 8 / 2;  // uint16 pairs
      DV_HIP(dv::launch_fill_i32(s, reinterpret_cast<int32_t*>(th.kh.p), words, 0));
      DV_HIP(dv::launch_fill_i32(s, reinterpret_cast<int32_t*>(th.kh2.p), words, 0));
    }
    DV_HIP(dv::launch_build_tables_bf16(s, Xd, offd, n, p, th.specm.as<dv::ColSpec>(),
                                        (int)cm.spec.size(), nth, nst, th.xab.as<uint16_t>(),
                                        th.offl.as<float>(), th.kg.as<uint16_t>(),
                                        th.kh.as<uint16_t>(), th.kh2.as<uint16_t>()));
    tm.xab = th.xab.as<uint16_t>();
    tm.offl = th.offl.as<float>();
    tm.kg = th.kg.as<uint16_t>();
    tm.kh = th.kh.as<uint16_t>();
    tm.kh2 = th.kh2.as<uint16_t>();
    tm.eop = td.eop;
  }
  return DEVIL_OK;
}

// Operands of the init tile pass (mode kInitNone: alpha0 only), built on the
// device from the uploaded design (dd: X, eo, Q for QR, logsf and xmean for
// the univariate start).
struct InitHost {
  DevBuf dZ, dE, dL, dR, dDen, dShift;
  dv::InitTile it{};
};

int build_init_tile(devil_ctx* ctx, const Design& d, const DeviceDesign& dd,
                    const dv::TileDesign& td, int want_alpha, int mode, InitHost& ih) {
  const int64_t n = d.n;
  const int p = d.p;
  dv::InitTile& it = ih.it;
  it.want_alpha = want_alpha;
  it.moi = d.moi;
  it.mode = mode;
  it.intercept_first = d.intercept_first;
  it.sum_sf = d.sum_sf;
  const size_t padded = (size_t)td.nblk * 16;
  const bool qr = mode == dv::kInitQR, uni = mode == dv::kInitUnivariate;
  DV_ALLOC(ih.dE, sizeof(double) * padded);
  if (qr || uni) DV_ALLOC(ih.dZ, sizeof(double) * padded * 16);
  if (uni) DV_ALLOC(ih.dL, sizeof(double) * padded);
  DV_HIP(dv::launch_build_init(ctx->stream, qr ? dd.Q.as<double>() : dd.x,
                               uni ? dd.xmean.as<double>() : nullptr, dd.eo.as<double>(),
                               uni ? dd.logsf.as<double>() : nullptr, n, p, td.nblk,
                               (qr || uni) ? ih.dZ.as<double>() : nullptr, ih.dE.as<double>(),
                               uni ? ih.dL.as<double>() : nullptr));
  it.einv = ih.dE.as<double>();
  it.Zp = (qr || uni) ? ih.dZ.as<double>() : nullptr;
  it.logsf = uni ? ih.dL.as<double>() : nullptr;
  it.R = qr ? dd.R.as<double>() : nullptr;
  it.denom = uni ? dd.denom.as<double>() : nullptr;
  return DEVIL_OK;
}

// FP64 tile design from a device-resident design matrix and offset.
int tile_from_device(devil_ctx* ctx, const double* X, const double* offset, int64_t n, int p,
                     TileHost& th, int mixed = kMixedNone) {
  std::vector<double> eo(n);
  DV_HIP(hipMemcpy(eo.data(), offset, sizeof(double) * n, hipMemcpyDeviceToHost));
  parallel_for(n, [&](int64_t lo, int64_t hi) {
    for (int64_t i = lo; i < hi; ++i) eo[i] = std::exp(eo[i]);
  });
  DevBuf eod;
  int rc;
  if ((rc = upload(ctx, eod, eo))) return rc;
  if ((rc = make_tile(ctx, X, offset, eod.as<double>(), n, p, mixed, th))) return rc;
  DV_HIP(hipStreamSynchronize(ctx->stream));  // eod is freed on return
  return DEVIL_OK;
}

// Chunk table of the post-IRLS passes: cell ranges of at most about
// max(2048, n / 64) cells, never straddling a patient block; the count is
// rounded up to a multiple of 8 with empty chunks (one XCD per chunk).
struct ChunkHost {
  std::vector<int64_t> c0, c1;
  std::vector<int32_t> cl;
  DevBuf d0, d1, dcl;
  dv::ChunkTable ct{};
};

int build_chunks(devil_ctx* ctx, int64_t n, const int32_t* ends, int C, ChunkHost& ch) {
  const int64_t target = std::max<int64_t>(2048, (n + 63) / 64);
  const int nc = C > 0 ? C : 1;
  int64_t start = 0;
  for (int k = 0; k < nc; ++k) {
    const int64_t end = C > 0 ? (int64_t)ends[k] : n;
    const int64_t len = end - start;
    const int64_t pieces = (len + target - 1) / target;
    for (int64_t q = 0; q < pieces; ++q) {
      ch.c0.push_back(start + len * q / pieces);
      ch.c1.push_back(start + len * (q + 1) / pieces);
      ch.cl.push_back(k);
    }
    start = end;
  }
  if (ch.c0.empty() || ch.c0.size() >= 8)
    while (ch.c0.empty() || ch.c0.size() % 8) {
      ch.c0.push_back(n);
      ch.c1.push_back(n);
      ch.cl.push_back(nc - 1);
    }
  int rc;
  if ((rc = upload(ctx, ch.d0, ch.c0))) return rc;
  if ((rc = upload(ctx, ch.d1, ch.c1))) return rc;
  if ((rc = upload(ctx, ch.dcl, ch.cl))) return rc;
  ch.ct.c0 = ch.d0.as<int64_t>();
  ch.ct.c1 = ch.d1.as<int64_t>();
  ch.ct.cl = ch.dcl.as<int32_t>();
  ch.ct.nchunk = (int)ch.c0.size();
  return DEVIL_OK;
}

// partial sums of the post-IRLS passes; hc0: the bread pass also keeps the
// HC0 meat columns
size_t post_part_doubles(const dv::TileDesign& td, const ChunkHost& ch, int64_t G,
                         bool hc0 = false) {
  return (size_t)((G + 15) / 16) * ch.ct.nchunk * 16 * (td.ncp + 16 + (hc0 ? td.ncp : 0));
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
  // timing only: no system-scope fence (no L2 writeback at every record,
  // which would also lengthen the measured kernels)
  hipEvent_t e = nullptr;
  if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) return nullptr;
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
    if (l.kind >= 0 && hipEventElapsedTime(&ms, l.a, l.b) == hipSuccess) {
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
// Mixed precision: each round accumulates in FP32/bf16 (FP64 partials).  A
// gene whose step lands within kMixedBand * eps of the stopping threshold (or
// is NaN) leaves the mixed rounds; once they end, those genes are refit from
// their starting beta with the FP64 rounds.  Stopping decisions closer to the
// threshold than the mixed trajectory's accuracy (~1e-8 in the step) are
// therefore always FP64 decisions on an FP64 trajectory.
constexpr double kMixedBand = 1e-3;
// mixed genes with a coefficient below this magnitude get their last step
// redone in FP64
constexpr double kPolishBelow = 5e-3;
constexpr int64_t kMixedMinCells = 4096;

// Chunking of one round over nA active genes.  A round whose workgroups
// (tiles_per_wg 16-gene tiles by nchunk chunks) cannot fill the GPU at the
// design's chunking gets shorter chunks, so the few genes left in the last
// rounds (and the FP64 refits) stream their rows with many waves rather than
// one wave per 1/64 of the row.  Bounded by the partial buffer (part_cap
// doubles) and by kMaxRoundChunks partials per gene for the update to sum.
constexpr int64_t kRoundWorkgroups = 1024;
constexpr int64_t kMaxRoundChunks = 512;
constexpr int64_t kMinRoundBlocks = 8;
// rounds of at most kSplitTiles tiles with more than kSplitSlices chunks sum
// their partials in kSplitSlices slices per tile first (k_irls_reduce)
constexpr int kSplitTiles = 16;
constexpr int kSplitSlices = 64;

dv::TileDesign round_shape(const dv::TileDesign& td, int nA, int tiles_per_wg,
                           size_t part_cap, int width = 0) {
  const int64_t ntile = (nA + 15) / 16;
  const int64_t groups = (ntile + tiles_per_wg - 1) / tiles_per_wg;
  if (ntile == 0 || groups * td.nchunk >= kRoundWorkgroups) return td;
  // doubles per gene and chunk in the partial buffer (the Gram columns)
  const int64_t w = width > 0 ? width : td.ncp;
  const int64_t cap = (int64_t)(part_cap / ((size_t)ntile * 16 * w));
  const int64_t want = std::min({(kRoundWorkgroups + groups - 1) / groups, kMaxRoundChunks, cap});
  if (want <= td.nchunk) return td;
  dv::TileDesign tr = td;
  set_chunks(tr, std::max(kMinRoundBlocks, (td.nblk + want - 1) / want));
  return (tr.nchunk > td.nchunk && tr.nchunk <= cap) ? tr : td;
}

// Pinned host slots and events through which irls_rounds reads the active
// counts (made on a context's first fit).
int round_sync_setup(devil_ctx* ctx) {
  if (!ctx->round_counts) {
    void* h = nullptr;
    DV_HIP(hipHostMalloc(&h, 64, hipHostMallocDefault));
    ctx->round_counts = static_cast<int32_t*>(h);
    void* d = nullptr;
    DV_HIP(hipHostGetDevicePointer(&d, h, 0));
    ctx->round_counts_dev = static_cast<int32_t*>(d);
  }
  for (hipEvent_t& e : ctx->round_ev)
    if (!e) DV_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  return DEVIL_OK;
}

// Rounds of accumulate + update over the active list act (count ncnt on the
// device, nA on the host) until none is left.  With band > 0 the mixed kernel
// runs and genes near the threshold are appended to replay / nreplay.  part
// holds part_cap doubles.
//
// The host never waits for the round it has just queued: round k's grid is
// sized for the count of round k - 1's list (lists only shrink), which round
// k - 2's update wrote and the host reads once that update has finished, so
// the GPU always has the next round queued.  A round queued for a list that
// has meanwhile emptied exits at once (the kernels read the count on the
// device).  The launch records get the exact count of their round.

int irls_rounds(devil_ctx* ctx, const TileHost& tl, bool mixed, dv::Counts c, int p,
                double* beta, const double* kvec, int max_iter, double eps, int32_t* iters,
                int32_t* status, int32_t* act, int32_t* act_spare, int32_t* ncnt,
                int32_t* ncnt_spare, int nA, double* part, size_t part_cap, int32_t* replay,
                int32_t* nreplay, double* bprev = nullptr) {
  const dv::TileDesign& tfull = mixed ? tl.tdm : tl.td;
  const int tiles_per_wg = accum_tiles_per_wg(mixed ? tl.mixed : kMixedNone, p, c);
  const int kind = mixed ? DEVIL_KERNEL_IRLS_MIXED : DEVIL_KERNEL_IRLS_FP64;
  hipStream_t s = ctx->stream;
  int rc;
  if ((rc = round_sync_setup(ctx))) return rc;
  const double row_bytes = (double)count_elem_size(c.dtype) * (double)tfull.n;
  int32_t* cur = act;
  int32_t* nxt = act_spare;
  int32_t* ncur = ncnt;
  int32_t* nnxt = ncnt_spare;
  volatile int32_t* slot = ctx->round_counts;
  DevBuf red;
  DV_ALLOC(red, sizeof(double) * kSplitTiles * kSplitSlices * 16 * tfull.ncp);
  // one round over list (a, na) into (b, nb), its next count copied to *out
  auto queue_round = [&](const dv::TileDesign& td, int bound, int32_t* a, int32_t* na,
                         int32_t* b, int32_t* nb, int32_t* out) -> int {
    if (!mixed)
      DV_HIP(dv::launch_irls_accum(s, td, c, p, a, na, bound, beta, kvec, part));
    else
      DV_HIP(launch_mixed_accum(tl.mixed, s, td, c, p, a, na, bound, beta, kvec, part));
    DV_HIP(dv::launch_fill_i32(s, nb, 1, 0));
    const int ntile_b = (bound + 15) / 16;
    if (ntile_b <= kSplitTiles && td.nchunk > kSplitSlices) {
      DV_HIP(dv::launch_irls_reduce(s, td.nchunk, td.ncp, kSplitSlices, ntile_b, na, part,
                                    red.as<double>()));
      dv::TileDesign tu = td;
      tu.nchunk = kSplitSlices;
      DV_HIP(dv::launch_irls_update(s, tu, p, a, na, bound, red.as<double>(), beta, kvec,
                                    max_iter, eps, iters, status, b, nb,
                                    mixed ? kMixedBand : 0.0, replay, nreplay, bprev));
    } else {
      DV_HIP(dv::launch_irls_update(s, td, p, a, na, bound, part, beta, kvec, max_iter, eps,
                                    iters, status, b, nb, mixed ? kMixedBand : 0.0, replay,
                                    nreplay, bprev));
    }
    DV_HIP(dv::launch_copy_counts(s, nb, 1, ctx->round_counts_dev + (out - ctx->round_counts)));
    return DEVIL_OK;
  };
  std::vector<int> count{nA};       // count[k]: genes of round k's list
  std::vector<long> rec;            // round k's launch record (-1: none)
  struct GraphRec {
    long idx;
    int r0;                         // its first round
  };
  std::vector<GraphRec> grec;       // graph replays: two rounds each
  int k = 0;
  bool graphed = false;
  for (; k < max_iter; ++k) {
    if (k >= 2) {
      DV_HIP(hipEventSynchronize(ctx->round_ev[k & 1]));
      count.push_back((int)slot[k & 1]);  // round k - 2's update: round k - 1's list
    }
    const int bound = k == 0 ? nA : count[k - 1];
    if (bound <= 0) break;
    // Once the list fits one workgroup column of tiles, round_shape no longer
    // depends on its length: the rest of the rounds can replay one graph.
    const int graph_genes = 16 * tiles_per_wg;
    if (k >= 2 && ctx->tune[DEVIL_TUNE_GRAPH] && bound <= graph_genes && max_iter - k >= 4) {
      // The tail: rounds k, k + 1, ... as replays of two captured rounds
      // (the list and counter buffers swap twice per replay, so one graph
      // serves every replay), with the grid and chunking of round k.  Two
      // graphs that differ only in their count slots alternate, so a
      // replay's counts are read while the next one runs.
      DV_HIP(hipEventSynchronize(ctx->round_ev[(k - 1) & 1]));
      count.push_back((int)slot[(k - 1) & 1]);  // round k's list
      const int gb = count[k];
      if (gb <= 0) break;
      const dv::TileDesign td = round_shape(tfull, gb, tiles_per_wg, part_cap);
      hipGraph_t g[2] = {nullptr, nullptr};
      hipGraphExec_t ge[2] = {nullptr, nullptr};
      auto cleanup = [&]() {
        for (int q = 0; q < 2; ++q) {
          if (ge[q]) (void)hipGraphExecDestroy(ge[q]);
          if (g[q]) (void)hipGraphDestroy(g[q]);
        }
      };
      for (int q = 0; q < 2 && rc == DEVIL_OK; ++q) {
        int32_t* sl = ctx->round_counts + 2 + 2 * q;
        if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) {
          rc = fail(ctx, DEVIL_E_HIP, "hipStreamBeginCapture");
          break;
        }
        int qrc = queue_round(td, gb, cur, ncur, nxt, nnxt, sl);
        if (qrc == DEVIL_OK) qrc = queue_round(td, gb, nxt, nnxt, cur, ncur, sl + 1);
        const hipError_t e = hipStreamEndCapture(s, &g[q]);
        if (qrc != DEVIL_OK) rc = qrc;
        else if (e != hipSuccess) rc = fail(ctx, DEVIL_E_HIP, "hipStreamEndCapture");
        else if (hipGraphInstantiate(&ge[q], g[q], nullptr, nullptr, 0) != hipSuccess)
          rc = fail(ctx, DEVIL_E_HIP, "hipGraphInstantiate");
      }
      if (rc != DEVIL_OK) {
        cleanup();
        return rc;
      }
      auto launch = [&](int j) -> int {
        const size_t nrec = ctx->launches.size();
        {
          LaunchTimer lt(ctx, kind, 0.0);
          DV_HIP(hipGraphLaunch(ge[j & 1], s));
        }
        grec.push_back({ctx->launches.size() > nrec ? (long)nrec : -1, k + 2 * j});
        DV_HIP(hipEventRecord(ctx->round_ev[j & 1], s));
        return DEVIL_OK;
      };
      // replay j runs rounds k + 2j, k + 2j + 1; its slots hold the counts of
      // rounds k + 2j + 1 and k + 2j + 2
      const int nrep_max = (max_iter - k + 1) / 2;
      int launched = 0;
      for (; launched < std::min(2, nrep_max) && rc == DEVIL_OK; ++launched) rc = launch(launched);
      for (int j = 0; rc == DEVIL_OK && j < launched; ++j) {
        if (hipEventSynchronize(ctx->round_ev[j & 1]) != hipSuccess) {
          rc = fail(ctx, DEVIL_E_HIP, "hipEventSynchronize");
          break;
        }
        const int32_t* sl = ctx->round_counts + 2 + 2 * (j & 1);
        count.push_back((int)sl[0]);
        count.push_back((int)sl[1]);
        if (sl[1] > 0 && launched < nrep_max) rc = launch(launched++);
      }
      if (rc == DEVIL_OK && hipStreamSynchronize(s) != hipSuccess)
        rc = fail(ctx, DEVIL_E_HIP, "hipStreamSynchronize");
      cleanup();
      if (rc != DEVIL_OK) return rc;
      graphed = true;
      break;
    }
    const dv::TileDesign td = round_shape(tfull, bound, tiles_per_wg, part_cap);
    const size_t nrec = ctx->launches.size();
    {
      LaunchTimer lt(ctx, kind, 0.0);
      if (!mixed)
        DV_HIP(dv::launch_irls_accum(s, td, c, p, cur, ncur, bound, beta, kvec, part));
      else
        DV_HIP(launch_mixed_accum(tl.mixed, s, td, c, p, cur, ncur, bound, beta, kvec, part));
    }
    rec.push_back(ctx->launches.size() > nrec ? (long)nrec : -1);
    DV_HIP(dv::launch_fill_i32(s, nnxt, 1, 0));
    const int ntile_b = (bound + 15) / 16;
    if (ntile_b <= kSplitTiles && td.nchunk > kSplitSlices) {
      DV_HIP(dv::launch_irls_reduce(s, td.nchunk, td.ncp, kSplitSlices, ntile_b, ncur, part,
                                    red.as<double>()));
      dv::TileDesign tu = td;
      tu.nchunk = kSplitSlices;
      DV_HIP(dv::launch_irls_update(s, tu, p, cur, ncur, bound, red.as<double>(), beta, kvec,
                                    max_iter, eps, iters, status, nxt, nnxt,
                                    mixed ? kMixedBand : 0.0, replay, nreplay, bprev));
    } else {
      DV_HIP(dv::launch_irls_update(s, td, p, cur, ncur, bound, part, beta, kvec, max_iter,
                                    eps, iters, status, nxt, nnxt, mixed ? kMixedBand : 0.0,
                                    replay, nreplay, bprev));
    }
    DV_HIP(dv::launch_copy_counts(s, nnxt, 1, ctx->round_counts_dev + (k & 1)));
    DV_HIP(hipEventRecord(ctx->round_ev[k & 1], s));
    std::swap(cur, nxt);
    std::swap(ncur, nnxt);
  }
  DV_HIP(hipStreamSynchronize(s));
  if (!graphed) {
    // counts of the lists the loop did not read (rounds k - 2 and k - 1)
    for (int j = std::max(0, k - 2); j < k; ++j)
      if ((int)count.size() == j + 1) count.push_back((int)slot[j & 1]);
  }
  // exact bytes per launch; launches over an emptied list are dropped
  for (size_t j = 0; j < rec.size(); ++j) {
    if (rec[j] < 0) continue;
    ctx->launches[rec[j]].bytes = row_bytes * (double)count[j];
    if (count[j] == 0) ctx->launches[rec[j]].kind = -1;
  }
  // a graph replay: its two rounds' accumulate bytes over the whole replay
  // (updates included in its time)
  for (const GraphRec& g : grec) {
    if (g.idx < 0) continue;
    const int a = g.r0 < (int)count.size() ? count[g.r0] : 0;
    const int b = g.r0 + 1 < (int)count.size() ? count[g.r0 + 1] : 0;
    ctx->launches[g.idx].bytes = row_bytes * (double)(a + b);
    if (a + b == 0) ctx->launches[g.idx].kind = -1;
  }
  collect_launches(ctx);
  return DEVIL_OK;
}

// Lock-step rounds of an overdispersion driver (MLE, "new") over the active
// list, with the count reads of irls_rounds: the host never waits for the
// round it has just queued, so the short rounds of the last genes run back to
// back.  accum(td, bound, k, cur, ncur) queues round k's accumulate pass
// (timed as `kind`), update(td, bound, k, cur, ncur, nxt, nnxt) its update,
// which appends the genes still running to nxt.  Lists only shrink.  *left =
// genes still active after max_rounds rounds.
template <class Accum, class Update>
int driver_rounds(devil_ctx* ctx, const dv::TileDesign& tfull, size_t part_cap, int width,
                  int kind, double row_bytes, int32_t* cur, int32_t* nxt, int32_t* ncur,
                  int32_t* nnxt, int nA, int64_t max_rounds, Accum&& accum, Update&& update,
                  int* left) {
  hipStream_t s = ctx->stream;
  int rc;
  if ((rc = round_sync_setup(ctx))) return rc;
  volatile int32_t* slot = ctx->round_counts;
  std::vector<int> count{nA};  // count[k]: genes of round k's list
  std::vector<long> rec;       // round k's launch record (-1: none)
  int64_t k = 0;
  for (; k < max_rounds; ++k) {
    if (k >= 2) {
      DV_HIP(hipEventSynchronize(ctx->round_ev[k & 1]));
      count.push_back((int)slot[k & 1]);  // round k - 2's update: round k - 1's list
    }
    const int bound = k == 0 ? nA : count[k - 1];
    if (bound <= 0) break;
    const dv::TileDesign td = round_shape(tfull, bound, 4, part_cap, width);
    const size_t nrec = ctx->launches.size();
    {
      LaunchTimer lt(ctx, kind, 0.0);
      if ((rc = accum(td, bound, k, cur, ncur))) return rc;
    }
    rec.push_back(ctx->launches.size() > nrec ? (long)nrec : -1);
    DV_HIP(dv::launch_fill_i32(s, nnxt, 1, 0));
    if ((rc = update(td, bound, k, cur, ncur, nxt, nnxt))) return rc;
    DV_HIP(dv::launch_copy_counts(s, nnxt, 1, ctx->round_counts_dev + (k & 1)));
    DV_HIP(hipEventRecord(ctx->round_ev[k & 1], s));
    std::swap(cur, nxt);
    std::swap(ncur, nnxt);
  }
  DV_HIP(hipStreamSynchronize(s));
  for (int64_t j = std::max<int64_t>(0, k - 2); j < k; ++j)
    if ((int64_t)count.size() == j + 1) count.push_back((int)slot[j & 1]);
  for (size_t j = 0; j < rec.size(); ++j) {
    if (rec[j] < 0) continue;
    ctx->launches[rec[j]].bytes = row_bytes * (double)count[j];
    if (count[j] == 0) ctx->launches[rec[j]].kind = -1;
  }
  collect_launches(ctx);
  *left = (int64_t)count.size() > k ? count[k] : 0;
  return DEVIL_OK;
}

int run_irls_tiled(devil_ctx* ctx, const TileHost& tl, dv::Counts c, int64_t G, int p,
                   double* beta, const double* kvec, int max_iter, double eps,
                   int32_t* iters, int32_t* status) {
  const dv::TileDesign& td = tl.td;
  const dv::TileDesign& tdm = tl.tdm;
  const bool mixed = tl.mixed != kMixedNone;
  hipStream_t s = ctx->stream;
  DV_HIP(dv::launch_fill_i32(s, iters, G, 0));
  DV_HIP(dv::launch_fill_i32(s, status, G, 0));
  if (max_iter <= 0 || G == 0) {
    if (max_iter <= 0) DV_HIP(dv::launch_fill_i32(s, status, G, DEVIL_GENE_MAXITER));
    return DEVIL_OK;
  }
  DevBuf act0, act1, cnt, part, rep, beta0;
  DV_ALLOC(act0, sizeof(int32_t) * G);
  DV_ALLOC(act1, sizeof(int32_t) * G);
  DV_ALLOC(cnt, sizeof(int32_t) * 4);
  const int64_t ntile = (G + 15) / 16;
  size_t part_per_tile = (size_t)td.nchunk * 16 * td.ncp;
  if (mixed) part_per_tile = std::max(part_per_tile, (size_t)tdm.nchunk * 16 * tdm.ncp);
  const size_t part_cap = (size_t)ntile * part_per_tile;
  DV_ALLOC(part, sizeof(double) * part_cap);
  int32_t* c0 = cnt.as<int32_t>();
  DV_HIP(dv::launch_iota(s, act0.as<int32_t>(), (int)G));
  DV_HIP(dv::launch_fill_i32(s, c0, 1, (int32_t)G));
  int rc;
  if (!mixed)
    return irls_rounds(ctx, tl, false, c, p, beta, kvec, max_iter, eps, iters, status,
                       act0.as<int32_t>(), act1.as<int32_t>(), c0, c0 + 1, (int)G,
                       part.as<double>(), part_cap, nullptr, nullptr);
  DevBuf bprev;
  DV_ALLOC(rep, sizeof(int32_t) * G);
  DV_ALLOC(beta0, sizeof(double) * G * p);
  DV_ALLOC(bprev, sizeof(double) * G * p);
  DV_HIP(hipMemcpyAsync(beta0.p, beta, sizeof(double) * G * p, hipMemcpyDeviceToDevice, s));
  DV_HIP(dv::launch_fill_f64(s, bprev.as<double>(), G * p, std::nan("")));
  int32_t* nrep = c0 + 2;
  DV_HIP(dv::launch_fill_i32(s, nrep, 1, 0));
  if ((rc = irls_rounds(ctx, tl, true, c, p, beta, kvec, max_iter, eps, iters, status,
                        act0.as<int32_t>(), act1.as<int32_t>(), c0, c0 + 1, (int)G,
                        part.as<double>(), part_cap, rep.as<int32_t>(), nrep,
                        bprev.as<double>())))
    return rc;
  int32_t nr = 0;
  DV_HIP(hipMemcpyAsync(&nr, nrep, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  DV_HIP(hipStreamSynchronize(s));
  ctx->replayed += nr;
  if (nr > 0) {
    // FP64 refit of the replay genes from their starting beta
    DV_HIP(dv::launch_restart_genes(s, rep.as<int32_t>(), nr, p, beta0.as<double>(), beta,
                                    iters, status, bprev.as<double>()));
    if ((rc = irls_rounds(ctx, tl, false, c, p, beta, kvec, max_iter, eps, iters, status,
                          rep.as<int32_t>(), act1.as<int32_t>(), nrep, c0 + 3, nr,
                          part.as<double>(), part_cap, nullptr, nullptr)))
      return rc;
  }
  // The mixed trajectory leaves ~5e-7 absolute on beta; where a coefficient
  // is small that is a large relative error.  Such genes get their last step
  // redone in FP64 from the beta it started at (the stopping decisions and
  // iteration counts are not touched).
  int32_t* list = act0.as<int32_t>();
  DV_HIP(dv::launch_fill_i32(s, c0, 1, 0));
  DV_HIP(dv::launch_polish_select(s, G, p, beta, bprev.as<double>(), status, kPolishBelow, list,
                                  c0));
  int32_t np = 0;
  DV_HIP(hipMemcpyAsync(&np, c0, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  DV_HIP(hipStreamSynchronize(s));
  ctx->polished += np;
  if (np == 0) return DEVIL_OK;
  const dv::TileDesign td1 = round_shape(td, np, 4, part_cap);
  DV_HIP(dv::launch_irls_accum(s, td1, c, p, list, c0, np, bprev.as<double>(), kvec,
                               part.as<double>()));
  DV_HIP(dv::launch_irls_update(s, td1, p, list, c0, np, part.as<double>(), beta, kvec, max_iter,
                                eps, iters, status, act1.as<int32_t>(), c0 + 1, 0.0, nullptr,
                                nullptr, bprev.as<double>(), 1));
  return DEVIL_OK;
}

// MLE overdispersion over genes [0, G): setup pass, then rounds of
// accumulate + update over the genes still iterating (the first round is the
// far-left check of R/overdispersion.R:40-46).  theta, iters are outputs;
// status is or-ed.
int run_mle_tiled(devil_ctx* ctx, const dv::TileDesign& td, dv::Counts c, int64_t G, int64_t n,
                  int p, const double* beta, int max_iter, double tolerance, double* theta,
                  int32_t* iters, int32_t* status) {
  hipStream_t s = ctx->stream;
  if (G == 0) return DEVIL_OK;
  const int driver = ctx->tune[DEVIL_TUNE_MLE_DRIVER] == DEVIL_MLE_ROOT ? dv::kMleDriverRoot
                                                                       : dv::kMleDriverNlminb;
  DevBuf dbl, ints, act0, act1, cnt, part, tab, port;
  DV_ALLOC(dbl, sizeof(double) * G * 7);
  DV_ALLOC(ints, sizeof(int32_t) * G * 2);
  DV_ALLOC(tab, sizeof(int32_t) * G * dv::kMleTabCap * 2);
  DV_ALLOC(act0, sizeof(int32_t) * G);
  DV_ALLOC(act1, sizeof(int32_t) * G);
  DV_ALLOC(cnt, sizeof(int32_t) * 2);
  if (driver == dv::kMleDriverNlminb) DV_ALLOC(port, sizeof(dv::PortState) * G);
  const int W = 8 + 3 * td.ncp;
  const int64_t ntile = (G + 15) / 16;
  DV_ALLOC(part, sizeof(double) * (size_t)ntile * td.nchunk * 16 * W);
  dv::MleState ms{};  // raw_start stays NULL: the MLE uses the fallback start
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
  ms.port = port.as<dv::PortState>();
  ms.driver = driver;
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
  const double row_bytes = (double)count_elem_size(c.dtype) * (double)n;
  const size_t part_cap = (size_t)ntile * td.nchunk * 16 * W;
  // passes: the far-left check, then at most max_iter + 1 Newton passes (root
  // driver) or eval.max = 200 objective evaluations (nlminb driver)
  const int max_rounds = std::max(max_iter, 200) + 2;
  int left = 0;
  return driver_rounds(
      ctx, td, part_cap, W, DEVIL_KERNEL_MLE, row_bytes, cur, nxt, ncur, nnxt, nA, max_rounds,
      [&](const dv::TileDesign& tr, int bound, int64_t k, int32_t* a, int32_t* na) -> int {
        DV_HIP(dv::launch_mle_accum(s, tr, c, p, a, na, bound, beta, ms.x, ms.tab_n,
                                    driver == dv::kMleDriverNlminb && k > 0,
                                    part.as<double>()));
        return DEVIL_OK;
      },
      [&](const dv::TileDesign& tr, int bound, int64_t, int32_t* a, int32_t* na, int32_t* b,
          int32_t* nb) -> int {
        DV_HIP(dv::launch_mle_update(s, tr, p, a, na, bound, part.as<double>(), ms, n, max_iter,
                                     tolerance, b, nb));
        return DEVIL_OK;
      },
      &left);
}

// "new" / "I" overdispersion over genes [0, G): the setup pass (count table,
// moment start), then rounds of one F_k evaluation per active gene until
// every gene's driver has finished.  theta, iters (kappa_iters) are outputs;
// status is or-ed.
int run_newod_tiled(devil_ctx* ctx, const dv::TileDesign& td, dv::Counts c, int64_t G,
                    int64_t n, int p, const double* beta, int max_iter, double tolerance,
                    double* theta, int32_t* iters, int32_t* status) {
  hipStream_t s = ctx->stream;
  if (G == 0) return DEVIL_OK;
  DevBuf dbl, ints, tab, act0, act1, cnt, part, st, kreq, scr;
  DV_ALLOC(dbl, sizeof(double) * G * 8);
  DV_ALLOC(ints, sizeof(int32_t) * G * 2);
  DV_ALLOC(tab, sizeof(int32_t) * G * dv::kMleTabCap * 2);
  DV_ALLOC(act0, sizeof(int32_t) * G);
  DV_ALLOC(act1, sizeof(int32_t) * G);
  DV_ALLOC(cnt, sizeof(int32_t) * 2);
  DV_ALLOC(st, sizeof(dv::NewodState) * G);
  DV_ALLOC(kreq, sizeof(double) * G);
  DV_ALLOC(scr, sizeof(int32_t) * G * 2 + sizeof(double) * G);
  const int64_t ntile = (G + 15) / 16;
  const size_t part_cap = (size_t)ntile * td.nchunk * 16 * 4;
  DV_ALLOC(part, sizeof(double) * part_cap);
  // the setup pass of the MLE: count table and (var y - ybar) / ybar^2; its
  // MLE special cases land in scratch (the "new" driver runs every gene)
  dv::MleState ms{};
  double* d = dbl.as<double>();
  ms.x = d;
  ms.lo = d + G;
  ms.hi = d + 2 * G;
  ms.start = d + 3 * G;
  ms.sum_y = d + 4 * G;
  ms.sum_py = d + 5 * G;
  ms.max_y = d + 6 * G;
  ms.raw_start = d + 7 * G;
  ms.mode = ints.as<int32_t>();
  ms.tab_n = ints.as<int32_t>() + G;
  ms.tab_u = tab.as<int32_t>();
  ms.tab_f = tab.as<int32_t>() + G * dv::kMleTabCap;
  ms.iters = scr.as<int32_t>();
  ms.status = scr.as<int32_t>() + G;
  ms.theta = reinterpret_cast<double*>(scr.as<int32_t>() + 2 * G);
  ms.port = nullptr;
  ms.driver = dv::kMleDriverRoot;
  DV_HIP(dv::launch_fill_f64(s, ms.raw_start, G, std::nan("")));
  DV_HIP(dv::launch_fill_i32(s, ms.status, G, 0));
  DV_HIP(dv::launch_mle_setup(s, c, G, n, p, ms));
  // the newod kernels write the real outputs
  ms.iters = iters;
  ms.status = status;
  ms.theta = theta;
  int32_t* cur = act0.as<int32_t>();
  int32_t* nxt = act1.as<int32_t>();
  int32_t* ncur = cnt.as<int32_t>();
  int32_t* nnxt = cnt.as<int32_t>() + 1;
  DV_HIP(dv::launch_fill_i32(s, ncur, 1, 0));
  DV_HIP(dv::launch_newod_init(s, G, ms.raw_start, st.as<dv::NewodState>(), kreq.as<double>(),
                               cur, ncur, max_iter, tolerance));
  int32_t nA = (int32_t)G;
  const double row_bytes = (double)count_elem_size(c.dtype) * (double)n;
  // every evaluation is one round; the driver's evaluations are bounded by
  // its loops (Newton start, Aitken steps with backtracking, the bracket
  // search and at most 60 Brent steps)
  const int64_t max_rounds = 16 + (int64_t)std::max(max_iter, 1) * 40 + 100;
  int left = 0;
  int rc = driver_rounds(
      ctx, td, part_cap, 4, DEVIL_KERNEL_MLE, row_bytes, cur, nxt, ncur, nnxt, nA, max_rounds,
      [&](const dv::TileDesign& tr, int bound, int64_t, int32_t* a, int32_t* na) -> int {
        DV_HIP(dv::launch_newod_accum(s, tr, c, p, a, na, bound, beta, kreq.as<double>(),
                                      ms.tab_n, part.as<double>()));
        return DEVIL_OK;
      },
      [&](const dv::TileDesign& tr, int bound, int64_t, int32_t* a, int32_t* na, int32_t* b,
          int32_t* nb) -> int {
        DV_HIP(dv::launch_newod_update(s, tr.nchunk, a, na, bound, part.as<double>(), ms,
                                       st.as<dv::NewodState>(), kreq.as<double>(), n, max_iter,
                                       tolerance, b, nb));
        return DEVIL_OK;
      },
      &left);
  if (rc) return rc;
  if (left > 0) return fail(ctx, DEVIL_E_HIP, "new overdispersion: evaluation rounds exhausted");
  return DEVIL_OK;
}

// Autotuner of the mixed-precision IRLS: times one accumulate pass over the
// first min(G, 1024) genes for each candidate kernel and chunk size (those the
// context leaves open) and returns the fastest.
uint64_t tune_key(int64_t n, int p, int dtype) {
  int lg = 0;
  while ((int64_t(1) << lg) < n) ++lg;
  return ((uint64_t)lg << 16) | ((uint64_t)p << 8) | (uint64_t)dtype;
}


// Mixed-precision bread of genes [0, G) on the bf16 operand tables; with C >
// 0 patients (chunk table ch, aligned to the patient blocks) the same pass
// also sums each patient's scores and the clustered sandwich follows.  ok =
// false when some gene's FP32 element range was exceeded: the caller then runs
// the FP64 bread (and meat) for the problem.
int bread_mixed(devil_ctx* ctx, const TileHost& tl, const ChunkHost& ch, dv::Counts c,
                int64_t G, int64_t n, int p, int C, const double* beta, const double* theta,
                double* bread, double* sandwich, int32_t* status, bool& ok) {
  hipStream_t s = ctx->stream;
  const dv::TileDesign& tm = tl.tdm;
  const bool meat = C > 0;
  const int nchunk = meat ? ch.ct.nchunk : tm.nchunk;
  const size_t tiles = (size_t)((G + 15) / 16);
  DevBuf genes, cnt, sing, part, spart;
  DV_ALLOC(genes, sizeof(int32_t) * G);
  DV_ALLOC(sing, sizeof(int32_t) * G);
  DV_ALLOC(cnt, sizeof(int32_t) * 2);
  DV_ALLOC(part, sizeof(double) * tiles * nchunk * 16 * tm.ncp);
  if (meat) DV_ALLOC(spart, sizeof(double) * tiles * nchunk * 16 * 16);
  DV_HIP(dv::launch_iota(s, genes.as<int32_t>(), (int)G));
  DV_HIP(dv::launch_fill_i32(s, cnt.as<int32_t>(), 1, (int32_t)G));
  DV_HIP(dv::launch_fill_i32(s, cnt.as<int32_t>() + 1, 1, 0));
  if (meat)
    DV_HIP(dv::launch_bread_meat_mixed(s, tm, ch.ct, c, p, genes.as<int32_t>(),
                                       cnt.as<int32_t>(), G, beta, theta, part.as<double>(),
                                       spart.as<double>(), bread, sing.as<int32_t>(),
                                       cnt.as<int32_t>() + 1));
  else
    DV_HIP(dv::launch_bread_mixed(s, tm, c, p, genes.as<int32_t>(), cnt.as<int32_t>(), G, beta,
                                  theta, part.as<double>(), bread, sing.as<int32_t>(),
                                  cnt.as<int32_t>() + 1));
  int32_t nbad = 0;
  DV_HIP(hipMemcpyAsync(&nbad, cnt.as<int32_t>() + 1, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  DV_HIP(hipStreamSynchronize(s));
  ok = nbad == 0;
  if (ok) {
    DV_HIP(dv::launch_merge_flags(s, status, sing.as<int32_t>(), G, DEVIL_GENE_SINGULAR));
    if (meat)
      DV_HIP(dv::launch_meat_finish(s, tm, ch.ct, G, n, p, C, spart.as<double>(), bread,
                                    sandwich));
  }
  // part and spart are freed on return
  DV_HIP(hipStreamSynchronize(s));
  return DEVIL_OK;
}

int autotune_mixed(devil_ctx* ctx, const DeviceDesign& dd, int64_t n, int p, dv::Counts c,
                   int64_t G, const double* beta, const double* kvec, int& mk, int64_t& cc) {
  hipStream_t s = ctx->stream;
  // enough genes that every candidate fills the GPU much as the full fit's
  // first rounds do
  const int64_t Gs = std::min<int64_t>(G, 16384);
  // the usual winners first: a later candidate must be 2% faster to replace
  // the choice, so timing noise does not flip it
  std::vector<int> kinds;
  if (ctx->tune[DEVIL_TUNE_MIXED_KERNEL] == DEVIL_MIXED_AUTO)
    kinds = {kMixedBf16Ws, kMixedBf16, kMixedF32};  // two-tile waves measured slower on C3
  else kinds = {mixed_kernel(ctx)};
  std::vector<int64_t> cells;
  if (ctx->tune[DEVIL_TUNE_CHUNK_CELLS] == 0) cells = {2048, 1024, 4096};
  else cells = {(int64_t)ctx->tune[DEVIL_TUNE_CHUNK_CELLS]};
  DevBuf act, cnt;
  DV_ALLOC(act, sizeof(int32_t) * Gs);
  DV_ALLOC(cnt, sizeof(int32_t));
  DV_HIP(dv::launch_iota(s, act.as<int32_t>(), (int)Gs));
  DV_HIP(dv::launch_fill_i32(s, cnt.as<int32_t>(), 1, (int32_t)Gs));
  hipEvent_t e0 = pool_event(ctx), e1 = pool_event(ctx);
  if (!e0 || !e1) return fail(ctx, DEVIL_E_HIP, "hipEventCreate");
  float best = 1e30f;
  int rc = DEVIL_OK;
  for (int kind : kinds)
    for (int64_t cl : cells) {
      TileHost th;
      if ((rc = make_tile(ctx, dd.x, dd.off.as<double>(), dd.eo.as<double>(), n, p, kind, th, cl)))
        break;
      const dv::TileDesign& tm = th.tdm;
      DevBuf part;
      DV_ALLOC(part, sizeof(double) * (size_t)((Gs + 15) / 16) * tm.nchunk * 16 * tm.ncp);
      auto launch = [&]() {
        return launch_mixed_accum(kind, s, tm, c, p, act.as<int32_t>(), cnt.as<int32_t>(),
                                  (int)Gs, beta, kvec, part.as<double>());
      };
      DV_HIP(launch());
      DV_HIP(hipEventRecord(e0, s));
      DV_HIP(launch());
      DV_HIP(launch());
      DV_HIP(hipEventRecord(e1, s));
      DV_HIP(hipEventSynchronize(e1));
      float ms = 0.f;
      DV_HIP(hipEventElapsedTime(&ms, e0, e1));
      if (ms < 0.98f * best) {
        best = ms;
        mk = kind;
        cc = cl;
      }
    }
  ctx->ev_pool.push_back(e0);
  ctx->ev_pool.push_back(e1);
  return rc;
}

int set_device(devil_ctx* ctx) {
  if (!ctx) return DEVIL_E_ARG;
  DV_HIP(hipSetDevice(ctx->device));
  return DEVIL_OK;
}

}  // namespace

// Scratch of devil_test_de_noalloc: a fixed device array, one copy per GPU,
// held until the process ends.  In host code the name is only a stand-in for
// the device symbol: its address on each device is looked up once, with that
// device current (devil_create), and only that address is handed to kernels.
static constexpr size_t kNoAllocScratchBytes = (size_t)16 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// the address of this device's scratch array (the device must be current)
int scratch_address(int device, void** ptr) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (device < 0 || device >= kMaxDevices) return DEVIL_E_NODEVICE;
  if (!g_scratch_ptr[device]) {
    void* q = nullptr;
    if (hipGetSymbolAddress(&q, HIP_SYMBOL(scratch)) != hipSuccess || !q) return DEVIL_E_HIP;
    g_scratch_ptr[device] = q;
  }
  *ptr = g_scratch_ptr[device];
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
  int rc = scratch_address(hip_device, &ctx->scratch);
  if (rc != DEVIL_OK) {
    delete ctx;
    return rc;
  }
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
  for (devil_ctx* sub : ctx->subs) devil_destroy(sub);
  for (hipEvent_t e : ctx->round_ev)
    if (e) (void)hipEventDestroy(e);
  if (ctx->round_counts) (void)hipHostFree(ctx->round_counts);
  if (ctx->joint_host) (void)hipHostFree(ctx->joint_host);
  if (ctx->pad_p) (void)hipFree(ctx->pad_p);
  ctx->cache.clear();
  delete ctx;
}

}  // extern "C"

namespace {

// One problem on ctx's stream (the body of devil_fit for n_problems == 1).
// One problem of a fit across its phases: set-up (validation, design tables,
// buffers), per gene batch the init pass, the IRLS and the stages after it,
// then the outputs.  fit_one runs them in order; devil_fit with several
// problems runs the IRLS of the eligible ones jointly (run_irls_joint).
struct ProbRun {
  const devil_problem* pr = nullptr;
  devil_outputs* out = nullptr;
  int64_t G = 0, n = 0;
  int p = 0, mem = 0, C = 0, init_mode = 0, mk = kMixedNone;
  bool mixed = false, autotune = false, want_sw = false, pad_host = false, run_init = false;
  int64_t cc = 2048, batch = 0, ld_pad = 0, dev_ld = 0;
  uint64_t key = 0;
  size_t es = 4;
  double eps = 0.0;
  std::vector<int32_t> ends_h;
  Design d;
  DeviceDesign dd;
  TileHost tl;
  ChunkHost chunks;
  InitHost ih;
  DevBuf alpha0, kvec, beta, theta, iters, thit, status, bread, sandwich, cbuf, stage, post;
  const void* dev_counts = nullptr;
  int64_t g0 = 0, Gb = 0;  // the current gene batch
  dv::Counts c{};
};

// Validation, design tables, per-gene buffers, counts layout.  G == 0 leaves
// r.G = 0 (nothing else to do).
int run_setup(devil_ctx* ctx, const devil_problem* pr, devil_outputs* out, ProbRun& r) {
  int rc;
  if ((rc = set_device(ctx))) return rc;
  if (!pr || !out) return fail(ctx, DEVIL_E_ARG, "NULL problem or outputs");
  r.pr = pr;
  r.out = out;
  const int64_t G = pr->G, n = pr->n;
  const int p = pr->p;
  r.G = G;
  r.n = n;
  r.p = p;
  if ((rc = check_p(ctx, p))) return rc;
  if ((rc = check_counts_args(ctx, pr->counts, pr->counts_dtype, G, n, pr->counts_ld)))
    return rc;
  if (pr->mem != DEVIL_MEM_HOST && pr->mem != DEVIL_MEM_DEVICE)
    return fail(ctx, DEVIL_E_ARG, "mem must be DEVIL_MEM_HOST or DEVIL_MEM_DEVICE");
  if (!pr->X || !pr->offset) return fail(ctx, DEVIL_E_ARG, "X or offset is NULL");
  if (pr->init_mode < DEVIL_INIT_QR || pr->init_mode > DEVIL_INIT_UNIVARIATE)
    return fail(ctx, DEVIL_E_ARG, "unknown init_mode");
  if (pr->od_mode < DEVIL_OD_NONE || pr->od_mode > DEVIL_OD_NEW)
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
  r.mem = mem;
  r.C = C;

  r.ends_h.assign(C, 0);
  if (C > 0) {
    if (mem == DEVIL_MEM_DEVICE)
      DV_HIP(hipMemcpy(r.ends_h.data(), pr->cluster_ends, sizeof(int32_t) * C,
                       hipMemcpyDeviceToHost));
    else
      std::memcpy(r.ends_h.data(), pr->cluster_ends, sizeof(int32_t) * C);
    if ((rc = check_clusters(ctx, r.ends_h.data(), C, n))) return rc;
  }
  r.init_mode = pr->beta_init ? dv::kInitNone : pr->init_mode;
  if ((rc = prepare_design(ctx, pr->X, pr->offset, n, p, mem, r.init_mode, r.d, r.dd)))
    return rc;
  // mixed kernel and chunk size: fixed by the context, remembered from an
  // earlier fit of this shape, or tuned after the first batch's init pass
  // mixed precision pays where the counts stream is long: problems with fewer
  // than kMixedMinCells cells run the FP64 kernels (exact; their cost is small)
  r.mixed = pr->precision == DEVIL_PREC_MIXED && n >= kMixedMinCells;
  r.mk = r.mixed ? mixed_kernel(ctx) : kMixedNone;
  r.cc = ctx->tune[DEVIL_TUNE_CHUNK_CELLS] > 0 ? ctx->tune[DEVIL_TUNE_CHUNK_CELLS] : 2048;
  r.key = tune_key(n, p, pr->counts_dtype);
  r.autotune = r.mixed && (ctx->tune[DEVIL_TUNE_MIXED_KERNEL] == DEVIL_MIXED_AUTO ||
                           ctx->tune[DEVIL_TUNE_CHUNK_CELLS] == 0);
  if (r.autotune && ctx->tuned.count(r.key)) {
    const int v = ctx->tuned[r.key];
    if (ctx->tune[DEVIL_TUNE_MIXED_KERNEL] == DEVIL_MIXED_AUTO) r.mk = v & 0xff;
    if (ctx->tune[DEVIL_TUNE_CHUNK_CELLS] == 0) r.cc = (int64_t)(v >> 8);
    r.autotune = false;
  }
  if ((rc = make_tile(ctx, r.dd.x, r.dd.off.as<double>(), r.dd.eo.as<double>(), n, p,
                      r.autotune ? kMixedNone : r.mk, r.tl, r.cc)))
    return rc;

  // per-gene device state (gene-major)
  DV_ALLOC(r.alpha0, sizeof(double) * G);
  DV_ALLOC(r.kvec, sizeof(double) * G);
  DV_ALLOC(r.beta, sizeof(double) * G * p);
  DV_ALLOC(r.theta, sizeof(double) * G);
  DV_ALLOC(r.iters, sizeof(int32_t) * G);
  DV_ALLOC(r.thit, sizeof(int32_t) * G);
  DV_ALLOC(r.status, sizeof(int32_t) * G);
  DV_ALLOC(r.bread, sizeof(double) * G * p * p);
  r.want_sw = C > 0 && out->sandwich;
  if (r.want_sw) DV_ALLOC(r.sandwich, sizeof(double) * G * p * p);

  hipStream_t s = ctx->stream;
  if (pr->disp_init) {
    DV_HIP(hipMemcpyAsync(r.alpha0.p, pr->disp_init, sizeof(double) * G,
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
    DV_HIP(dv::launch_transpose_gp(s, tmp.as<double>(), G, p, 1, r.beta.as<double>()));
    DV_HIP(hipStreamSynchronize(s));
  }

  // counts: device-resident as given, or streamed in gene batches.  4-byte
  // counts are read with 16-byte loads (and staged in LDS by the mixed
  // kernels) only when every row starts 16-byte aligned: host batches land in
  // rows padded to a multiple of 4 elements, and a mixed fit of misaligned
  // device rows first copies them so (when HBM holds the copy)
  r.es = count_elem_size(pr->counts_dtype);
  const size_t es = r.es;
  r.ld_pad = es == 4 ? ((n + 3) & ~int64_t(3)) : pr->counts_ld;
  const int64_t ld_pad = r.ld_pad;
  r.batch = G;
  r.dev_counts = pr->counts;
  r.dev_ld = pr->counts_ld;
  r.pad_host = mem == DEVIL_MEM_HOST && pr->counts_ld != ld_pad;
  if (mem == DEVIL_MEM_HOST) {
    size_t freeb = 0, totalb = 0;
    DV_HIP(hipMemGetInfo(&freeb, &totalb));
    const size_t row_bytes =
        (size_t)ld_pad * es + (r.pad_host ? (size_t)pr->counts_ld * es : 0);
    const size_t budget = std::min<size_t>(freeb / 2, (size_t)8 << 30);
    r.batch = std::max<int64_t>(1, std::min<int64_t>(G, (int64_t)(budget / row_bytes)));
    if (ctx->tune[DEVIL_TUNE_BATCH_GENES] > 0)
      r.batch = std::min<int64_t>(r.batch, ctx->tune[DEVIL_TUNE_BATCH_GENES]);
    DV_ALLOC(r.cbuf, (size_t)ld_pad * es * r.batch);
    if (r.pad_host) DV_ALLOC(r.stage, (size_t)pr->counts_ld * es * r.batch);
  } else if (r.mixed && es == 4 && ctx->tune[DEVIL_TUNE_PAD_COPY] &&
             !counts_vec4(dv::Counts{pr->counts, pr->counts_dtype, pr->counts_ld})) {
    const size_t bytes = (size_t)G * ld_pad * es;
    if (bytes > ctx->pad_size) {
      size_t freeb = 0, totalb = 0;
      DV_HIP(hipMemGetInfo(&freeb, &totalb));
      if (bytes + ctx->pad_size <= freeb / 2) {
        if (ctx->pad_p) DV_HIP(hipFree(ctx->pad_p));
        ctx->pad_p = nullptr;
        ctx->pad_size = 0;
        if (hipMalloc(&ctx->pad_p, bytes) == hipSuccess) ctx->pad_size = bytes;
        else ctx->pad_p = nullptr;
      }
    }
    if (bytes <= ctx->pad_size) {
      DV_HIP(dv::launch_pad_rows(s, pr->counts, pr->counts_ld, G, n, ctx->pad_p, ld_pad));
      r.dev_counts = ctx->pad_p;
      r.dev_ld = ld_pad;
    }
  }
  ctx->aligned_copy = r.dev_counts != pr->counts;
  if ((rc = build_chunks(ctx, n, C > 0 ? r.ends_h.data() : nullptr, C, r.chunks))) return rc;
  DV_ALLOC(r.post, sizeof(double) * post_part_doubles(r.tl.td, r.chunks, r.batch));
  r.eps = (double)(float)pr->tolerance;  // src/beta.cpp:25 float eps
  if ((rc = build_init_tile(ctx, r.d, r.dd, r.tl.td, pr->disp_init ? 0 : 1, r.init_mode, r.ih)))
    return rc;
  r.run_init = r.ih.it.want_alpha || r.init_mode != dv::kInitNone;
  return DEVIL_OK;
}

// Gene batch [g0, g0 + Gb): its counts (uploaded when host-resident), the init
// pass (alpha0, beta0, k) and, on the first batch of an untuned shape, the
// mixed-kernel autotuner.
int run_batch_init(devil_ctx* ctx, ProbRun& r, int64_t g0, StageTimer& tm) {
  const devil_problem* pr = r.pr;
  hipStream_t s = ctx->stream;
  const int64_t Gb = std::min(r.batch, r.G - g0);
  const size_t es = r.es;
  r.g0 = g0;
  r.Gb = Gb;
  dv::Counts& c = r.c;
  c.dtype = pr->counts_dtype;
  if (r.mem == DEVIL_MEM_HOST) {
    const char* src = static_cast<const char*>(pr->counts) + (size_t)g0 * pr->counts_ld * es;
    if (!r.pad_host) {
      DV_HIP(hipMemcpyAsync(r.cbuf.p, src, (size_t)Gb * r.ld_pad * es, hipMemcpyHostToDevice,
                            s));
    } else {
      DV_HIP(hipMemcpyAsync(r.stage.p, src, (size_t)Gb * pr->counts_ld * es,
                            hipMemcpyHostToDevice, s));
      DV_HIP(dv::launch_pad_rows(s, r.stage.p, pr->counts_ld, Gb, r.n, r.cbuf.p, r.ld_pad));
    }
    c.ptr = r.cbuf.p;
    c.ld = r.ld_pad;
  } else {
    c.ptr = static_cast<const char*>(r.dev_counts) + (size_t)g0 * r.dev_ld * es;
    c.ld = r.dev_ld;
  }
  double* a0 = r.alpha0.as<double>() + g0;
  double* b0 = r.beta.as<double>() + g0 * r.p;
  hipEvent_t ev = tm.begin(DEVIL_STAGE_INIT);
  if (r.run_init)
    DV_HIP(dv::launch_init_tiled(s, r.tl.td, r.ih.it, r.chunks.ct, c, Gb, r.n, r.p,
                                 r.post.as<double>(), a0, b0));
  DV_HIP(dv::launch_k_from_alpha(s, a0, Gb, r.kvec.as<double>() + g0));
  tm.end(ev);
  int rc;
  if (r.autotune) {
    if ((rc = autotune_mixed(ctx, r.dd, r.n, r.p, c, Gb, b0, r.kvec.as<double>() + g0, r.mk,
                             r.cc)))
      return rc;
    ctx->tuned[r.key] = r.mk | (int)(r.cc << 8);
    r.autotune = false;
    r.tl = TileHost();
    if ((rc = make_tile(ctx, r.dd.x, r.dd.off.as<double>(), r.dd.eo.as<double>(), r.n, r.p, r.mk,
                        r.tl, r.cc)))
      return rc;
  }
  ctx->last_mk = r.mixed ? r.mk : kMixedNone;
  ctx->last_cc = r.cc;
  return DEVIL_OK;
}

int run_batch_irls(devil_ctx* ctx, ProbRun& r, StageTimer& tm) {
  hipEvent_t ev = tm.begin(DEVIL_STAGE_IRLS);
  int rc;
  if ((rc = run_irls_tiled(ctx, r.tl, r.c, r.Gb, r.p, r.beta.as<double>() + r.g0 * r.p,
                           r.kvec.as<double>() + r.g0, r.pr->max_iter, r.eps,
                           r.iters.as<int32_t>() + r.g0, r.status.as<int32_t>() + r.g0)))
    return rc;
  tm.end(ev);
  return DEVIL_OK;
}

// Overdispersion, bread (+ meat and sandwich) of the current gene batch.
int run_batch_post(devil_ctx* ctx, ProbRun& r, StageTimer& tm) {
  const devil_problem* pr = r.pr;
  hipStream_t s = ctx->stream;
  const int64_t g0 = r.g0, Gb = r.Gb, n = r.n;
  const int p = r.p, C = r.C;
  const dv::Counts& c = r.c;
  double* b0 = r.beta.as<double>() + g0 * p;
  double* th = r.theta.as<double>() + g0;
  int32_t* st = r.status.as<int32_t>() + g0;
  int32_t* tit = r.thit.as<int32_t>() + g0;
  int rc;
  hipEvent_t ev = tm.begin(DEVIL_STAGE_DISPERSION);
  if (pr->od_mode == DEVIL_OD_MOM) {
    DV_HIP(dv::launch_mom_tiled(s, r.tl.td, r.chunks.ct, c, Gb, n, p, b0, r.post.as<double>(),
                                th, st));
    DV_HIP(dv::launch_fill_i32(s, tit, Gb, 0));
  } else if (pr->od_mode == DEVIL_OD_MLE) {
    if ((rc = run_mle_tiled(ctx, r.tl.td, c, Gb, n, p, b0, pr->max_iter, pr->tolerance, th, tit,
                            st)))
      return rc;
  } else if (pr->od_mode == DEVIL_OD_NEW) {
    if ((rc = run_newod_tiled(ctx, r.tl.td, c, Gb, n, p, b0, pr->max_iter, pr->tolerance, th,
                              tit, st)))
      return rc;
  } else {
    DV_HIP(dv::launch_fill_f64(s, th, Gb, 0.0));
    DV_HIP(dv::launch_fill_i32(s, tit, Gb, 0));
  }
  tm.end(ev);
  ev = tm.begin(DEVIL_STAGE_BREAD_MEAT);
  // mixed precision: the bread on the bf16 tables (FP32 element math, FP64
  // sums and inverse) and, with patients, each patient's score sums from the
  // same pass (the HC0 meat stays on the FP64 pass)
  bool bread_done = false;
  if (r.mixed && bf16_layout(r.tl.mixed) && counts_ws(c) && C >= 0) {
    if ((rc = bread_mixed(ctx, r.tl, r.chunks, c, Gb, n, p, r.want_sw ? C : 0, b0, th,
                          r.bread.as<double>() + g0 * p * p,
                          r.want_sw ? r.sandwich.as<double>() + g0 * p * p : nullptr, st,
                          bread_done)))
      return rc;
  }
  if (!bread_done)
    DV_HIP(dv::launch_bread_tiled(s, r.tl.td, r.chunks.ct, c, Gb, n, p, b0, th, C,
                                  r.post.as<double>(), r.bread.as<double>() + g0 * p * p,
                                  r.want_sw ? r.sandwich.as<double>() + g0 * p * p : nullptr,
                                  st));
  tm.end(ev);
  return DEVIL_OK;
}

// Outputs in the caller's layout (beta back to R's column-major G x p); the
// stream is synchronised on return.
int run_outputs(devil_ctx* ctx, ProbRun& r) {
  const devil_problem* pr = r.pr;
  devil_outputs* out = r.out;
  hipStream_t s = ctx->stream;
  const int64_t G = r.G;
  const int p = r.p;
  DevBuf bcol;
  DV_ALLOC(bcol, sizeof(double) * G * p);
  DV_HIP(dv::launch_transpose_gp(s, r.beta.as<double>(), G, p, 0, bcol.as<double>()));
  const hipMemcpyKind kind =
      pr->mem == DEVIL_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
  DV_HIP(hipMemcpyAsync(out->beta, bcol.p, sizeof(double) * G * p, kind, s));
  DV_HIP(hipMemcpyAsync(out->theta, r.theta.p, sizeof(double) * G, kind, s));
  DV_HIP(hipMemcpyAsync(out->beta_iters, r.iters.p, sizeof(int32_t) * G, kind, s));
  DV_HIP(hipMemcpyAsync(out->theta_iters, r.thit.p, sizeof(int32_t) * G, kind, s));
  DV_HIP(hipMemcpyAsync(out->bread, r.bread.p, sizeof(double) * G * p * p, kind, s));
  if (r.want_sw)
    DV_HIP(hipMemcpyAsync(out->sandwich, r.sandwich.p, sizeof(double) * G * p * p, kind, s));
  if (out->status)
    DV_HIP(hipMemcpyAsync(out->status, r.status.p, sizeof(int32_t) * G, kind, s));
  DV_HIP(hipStreamSynchronize(s));
  return DEVIL_OK;
}

void reset_fit_stats(devil_ctx* ctx) {
  for (int k = 0; k < DEVIL_STAGE_TEST_DE; ++k) ctx->stage_ms[k] = 0.0;
  for (int k = 0; k < DEVIL_N_KERNELS; ++k) {
    ctx->k_n[k] = 0;
    ctx->k_ms[k] = ctx->k_bytes[k] = 0.0;
  }
  ctx->replayed = 0;
  ctx->polished = 0;
}

// ---------------------------------------------------------------------------
// Joint IRLS of many problems (devil_fit with n_problems > 1): every round is
// one accumulate launch per (design width, kernel) and one update launch per
// design width for all problems, instead of a pair of launches per problem.
// A problem's share of a round is a JointItem (its round shape, lists,
// outputs and ranges of the two grids); the items of a round travel to the
// device in one copy.  Rounds are queued one ahead of the count reads, as in
// irls_rounds.  The sequence per problem is run_irls_tiled's: mixed (or
// FP64) rounds, the FP64 refit of the genes the mixed rounds set aside, the
// FP64 last step of small coefficients.
struct JointMember {
  ProbRun* r = nullptr;
  bool mixed = false;
  int kind = dv::kJointFp64;  // this phase's accumulate kernel
  dv::TileDesign tfull{};     // this phase's full design
  int p = 0;
  size_t part_cap = 0;
  double row_bytes = 0.0;
  DevBuf act0, act1, part, rep, beta0, bprev;
  int32_t* cur = nullptr;     // this phase's active list and its spare
  int32_t* nxt = nullptr;
  int nA = 0;                 // genes of the list at the phase's start
  double band = 0.0;
  int polish = 0;
  const double* abeta = nullptr;  // beta the accumulate reads
};

int joint_host_setup(devil_ctx* ctx, size_t bytes) {
  if (ctx->joint_host_bytes >= bytes) return DEVIL_OK;
  if (ctx->joint_host) DV_HIP(hipHostFree(ctx->joint_host));
  ctx->joint_host = nullptr;
  ctx->joint_host_dev = nullptr;
  ctx->joint_host_bytes = 0;
  DV_HIP(hipHostMalloc(&ctx->joint_host, bytes, hipHostMallocDefault));
  DV_HIP(hipHostGetDevicePointer(&ctx->joint_host_dev, ctx->joint_host, 0));
  ctx->joint_host_bytes = bytes;
  return DEVIL_OK;
}

// Rounds over the members until every list is empty (at most max_rounds).
// cnt: 2 M device counters (A = the current lists' counts, B = the next);
// nrep: M device counters of the mixed members' set-aside genes.
int joint_rounds(devil_ctx* ctx, std::vector<JointMember*>& ms, int max_rounds, int32_t* cnt,
                 int32_t* nrep) {
  hipStream_t s = ctx->stream;
  const int M = (int)ms.size();
  if (M == 0) return DEVIL_OK;
  int rc;
  if ((rc = round_sync_setup(ctx))) return rc;
  // pinned host: 2 M count slots, then 3 staging slots of M items
  const size_t slot_items = (size_t)M;
  const size_t count_bytes = ((sizeof(int32_t) * 2 * M + 255) / 256) * 256;
  const size_t items_bytes = sizeof(dv::JointItem) * slot_items;
  if ((rc = joint_host_setup(ctx, count_bytes + 3 * items_bytes))) return rc;
  volatile int32_t* slot = static_cast<int32_t*>(ctx->joint_host);
  dv::JointItem* stage = reinterpret_cast<dv::JointItem*>(static_cast<char*>(ctx->joint_host) +
                                                          count_bytes);
  DevBuf ditems;
  DV_ALLOC(ditems, 3 * items_bytes);
  int32_t* A = cnt;
  int32_t* B = cnt + M;
  // round k's list counts per member; count[k][m]
  std::vector<std::vector<int>> count;
  count.emplace_back(M);
  for (int m = 0; m < M; ++m) count[0][m] = ms[m]->nA;
  struct Rec {
    long idx;
    std::vector<int> members;
  };
  std::vector<std::vector<Rec>> recs;  // per round, per accumulate launch
  std::vector<int32_t*> cur(M), nxt(M);
  for (int m = 0; m < M; ++m) {
    cur[m] = ms[m]->cur;
    nxt[m] = ms[m]->nxt;
  }
  // order of the items: by design width, ws kernel before FP64 within one
  std::vector<int> order(M);
  for (int m = 0; m < M; ++m) order[m] = m;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    if (ms[a]->p != ms[b]->p) return ms[a]->p < ms[b]->p;
    return ms[a]->kind > ms[b]->kind;
  });
  int k = 0;
  for (; k < max_rounds; ++k) {
    if (k >= 2) {
      DV_HIP(hipEventSynchronize(ctx->round_ev[k & 1]));
      std::vector<int> ck(M);
      for (int m = 0; m < M; ++m) ck[m] = (int)slot[(k & 1) * M + m];
      count.push_back(ck);  // round k - 2's update: round k - 1's lists
    }
    const std::vector<int>& bound = k == 0 ? count[0] : count[k - 1];
    int live = 0;
    for (int m = 0; m < M; ++m) live += bound[m] > 0 ? 1 : 0;
    if (live == 0) break;
    // the round's items, grouped by width (and kernel within a width)
    dv::JointItem* st = stage + (size_t)(k % 3) * slot_items;
    dv::JointItem* dev = reinterpret_cast<dv::JointItem*>(ditems.as<char>() +
                                                         (size_t)(k % 3) * items_bytes);
    struct Launch {
      int p, kind, first, n, total;
    };
    std::vector<Launch> acc, upd;
    std::vector<int> item_member;
    int ni = 0;
    for (int oi = 0; oi < M; ++oi) {
      const int m = order[oi];
      if (bound[m] <= 0) continue;
      JointMember& J = *ms[m];
      ProbRun& r = *J.r;
      const int tpw = J.kind == dv::kJointWs ? dv::ws_tiles_per_workgroup(J.p, 1) : 4;
      dv::JointItem& it = st[ni];
      std::memset(&it, 0, sizeof(it));
      it.td = round_shape(J.tfull, bound[m], tpw, J.part_cap);
      it.c = r.c;
      it.active = cur[m];
      it.nactive = A + m;
      it.beta = const_cast<double*>(J.abeta);
      it.kvec = r.kvec.as<double>();
      it.part = J.part.as<double>();
      it.ubeta = r.beta.as<double>();
      it.iters = r.iters.as<int32_t>();
      it.status = r.status.as<int32_t>();
      it.next = nxt[m];
      it.nnext = B + m;
      it.band = J.band;
      it.recheck = J.band > 0.0 ? J.rep.as<int32_t>() : nullptr;
      it.nrecheck = J.band > 0.0 ? nrep + m : nullptr;
      it.bprev = (J.band > 0.0 || J.polish) ? J.bprev.as<double>() : nullptr;
      it.eps = r.eps;
      it.max_iter = r.pr->max_iter;
      it.polish = J.polish;
      const int ntile = (bound[m] + 15) / 16;
      const int groups = (ntile + tpw - 1) / tpw;
      it.nwg = groups * it.td.nchunk;
      it.nub = bound[m];
      // accumulate launch of this (width, kernel)
      if (acc.empty() || acc.back().p != J.p || acc.back().kind != J.kind)
        acc.push_back({J.p, J.kind, ni, 0, 0});
      Launch& a = acc.back();
      it.wg0 = a.total;
      a.total += ((it.nwg + 7) / 8) * 8;
      a.n += 1;
      if (upd.empty() || upd.back().p != J.p) upd.push_back({J.p, 0, ni, 0, 0});
      Launch& u = upd.back();
      it.ub0 = u.total;
      u.total += it.nub;
      u.n += 1;
      item_member.push_back(m);
      ++ni;
    }
    DV_HIP(hipMemcpyAsync(dev, st, sizeof(dv::JointItem) * ni, hipMemcpyHostToDevice, s));
    DV_HIP(dv::launch_fill_i32(s, B, M, 0));
    std::vector<Rec> rk;
    for (const Launch& a : acc) {
      const size_t nrec = ctx->launches.size();
      {
        LaunchTimer lt(ctx, a.kind == dv::kJointWs ? DEVIL_KERNEL_IRLS_MIXED
                                                   : DEVIL_KERNEL_IRLS_FP64,
                       0.0);
        DV_HIP(dv::launch_irls_accum_joint(s, a.kind, a.p, dev + a.first, a.n, a.total));
      }
      Rec rec{ctx->launches.size() > nrec ? (long)nrec : -1, {}};
      for (int q = a.first; q < a.first + a.n; ++q) rec.members.push_back(item_member[q]);
      rk.push_back(rec);
    }
    recs.push_back(rk);
    for (const Launch& u : upd)
      DV_HIP(dv::launch_irls_update_joint(s, u.p, dev + u.first, u.n, u.total));
    DV_HIP(dv::launch_copy_counts(s, B, M,
                                  static_cast<int32_t*>(ctx->joint_host_dev) + (k & 1) * M));
    DV_HIP(hipEventRecord(ctx->round_ev[k & 1], s));
    std::swap(A, B);
    for (int m = 0; m < M; ++m) std::swap(cur[m], nxt[m]);
  }
  DV_HIP(hipStreamSynchronize(s));
  for (int j = std::max(0, k - 2); j < k; ++j)
    if ((int)count.size() == j + 1) {
      std::vector<int> ck(M);
      for (int m = 0; m < M; ++m) ck[m] = (int)slot[(j & 1) * M + m];
      count.push_back(ck);
    }
  // exact bytes per launch; launches whose lists had all emptied are dropped
  for (size_t j = 0; j < recs.size(); ++j)
    for (const Rec& rec : recs[j]) {
      if (rec.idx < 0) continue;
      double bytes = 0.0;
      for (int m : rec.members) bytes += ms[m]->row_bytes * (double)count[j][m];
      ctx->launches[rec.idx].bytes = bytes;
      if (bytes == 0.0) ctx->launches[rec.idx].kind = -1;
    }
  collect_launches(ctx);
  return DEVIL_OK;
}

// Whether a problem's IRLS can join: one gene batch (device counts, or host
// counts in one batch), the FP64 kernels or the LDS-staged mixed kernel.
bool joint_eligible(const ProbRun& r) {
  if (r.G == 0 || r.batch < r.G || r.pr->max_iter <= 0) return false;
  if (!r.mixed) return true;
  return r.tl.mixed == kMixedBf16Ws && counts_ws(r.c);
}

int run_irls_joint(devil_ctx* ctx, std::vector<ProbRun*>& runs) {
  hipStream_t s = ctx->stream;
  const int M = (int)runs.size();
  if (M == 0) return DEVIL_OK;
  std::vector<std::unique_ptr<JointMember>> mem(M);
  std::vector<JointMember*> all;
  for (int m = 0; m < M; ++m) {
    mem[m].reset(new (std::nothrow) JointMember());
    if (!mem[m]) return fail(ctx, DEVIL_E_NOMEM, "joint member");
    JointMember& J = *mem[m];
    ProbRun& r = *runs[m];
    J.r = &r;
    J.mixed = r.mixed;
    J.p = r.p;
    const int64_t G = r.G;
    J.row_bytes = (double)count_elem_size(r.c.dtype) * (double)r.n;
    const int64_t ntile = (G + 15) / 16;
    size_t per_tile = (size_t)r.tl.td.nchunk * 16 * r.tl.td.ncp;
    if (J.mixed) per_tile = std::max(per_tile, (size_t)r.tl.tdm.nchunk * 16 * r.tl.tdm.ncp);
    J.part_cap = (size_t)ntile * per_tile;
    DV_ALLOC(J.act0, sizeof(int32_t) * G);
    DV_ALLOC(J.act1, sizeof(int32_t) * G);
    DV_ALLOC(J.part, sizeof(double) * J.part_cap);
    DV_HIP(dv::launch_fill_i32(s, r.iters.as<int32_t>(), G, 0));
    DV_HIP(dv::launch_fill_i32(s, r.status.as<int32_t>(), G, 0));
    DV_HIP(dv::launch_iota(s, J.act0.as<int32_t>(), (int)G));
    if (J.mixed) {
      DV_ALLOC(J.rep, sizeof(int32_t) * G);
      DV_ALLOC(J.beta0, sizeof(double) * G * r.p);
      DV_ALLOC(J.bprev, sizeof(double) * G * r.p);
      DV_HIP(hipMemcpyAsync(J.beta0.p, r.beta.p, sizeof(double) * G * r.p,
                            hipMemcpyDeviceToDevice, s));
      DV_HIP(dv::launch_fill_f64(s, J.bprev.as<double>(), G * r.p, std::nan("")));
    }
    J.kind = J.mixed ? dv::kJointWs : dv::kJointFp64;
    J.tfull = J.mixed ? r.tl.tdm : r.tl.td;
    J.cur = J.act0.as<int32_t>();
    J.nxt = J.act1.as<int32_t>();
    J.nA = (int)G;
    J.band = J.mixed ? kMixedBand : 0.0;
    J.abeta = r.beta.as<double>();
    all.push_back(&J);
  }
  // counters: 2 M list counts, M set-aside counts
  DevBuf cnt;
  DV_ALLOC(cnt, sizeof(int32_t) * 3 * M);
  int32_t* nrep = cnt.as<int32_t>() + 2 * M;
  std::vector<int32_t> h(M);
  for (int m = 0; m < M; ++m) h[m] = (int32_t)all[m]->nA;
  DV_HIP(hipMemcpyAsync(cnt.p, h.data(), sizeof(int32_t) * M, hipMemcpyHostToDevice, s));
  DV_HIP(dv::launch_fill_i32(s, nrep, M, 0));
  int max_rounds = 0;
  for (auto* J : all) max_rounds = std::max(max_rounds, J->r->pr->max_iter);
  int rc;
  if ((rc = joint_rounds(ctx, all, max_rounds, cnt.as<int32_t>(), nrep))) return rc;

  // FP64 refit of the genes the mixed rounds set aside, from their starting beta
  std::vector<int32_t> nr(M, 0);
  DV_HIP(hipMemcpyAsync(nr.data(), nrep, sizeof(int32_t) * M, hipMemcpyDeviceToHost, s));
  DV_HIP(hipStreamSynchronize(s));
  std::vector<JointMember*> refit;
  for (int m = 0; m < M; ++m) {
    JointMember& J = *all[m];
    if (!J.mixed || nr[m] <= 0) continue;
    ProbRun& r = *J.r;
    ctx->replayed += nr[m];
    DV_HIP(dv::launch_restart_genes(s, J.rep.as<int32_t>(), nr[m], r.p, J.beta0.as<double>(),
                                    r.beta.as<double>(), r.iters.as<int32_t>(),
                                    r.status.as<int32_t>(), J.bprev.as<double>()));
    J.kind = dv::kJointFp64;
    J.tfull = r.tl.td;
    J.cur = J.rep.as<int32_t>();
    J.nxt = J.act1.as<int32_t>();
    J.nA = nr[m];
    J.band = 0.0;
    refit.push_back(&J);
  }
  if (!refit.empty()) {
    const int R = (int)refit.size();
    for (int q = 0; q < R; ++q) h[q] = (int32_t)refit[q]->nA;
    DV_HIP(hipMemcpyAsync(cnt.p, h.data(), sizeof(int32_t) * R, hipMemcpyHostToDevice, s));
    int mr = 0;
    for (auto* J : refit) mr = std::max(mr, J->r->pr->max_iter);
    if ((rc = joint_rounds(ctx, refit, mr, cnt.as<int32_t>(), nrep))) return rc;
  }

  // The mixed trajectory leaves ~5e-7 absolute on beta: genes with a small
  // coefficient get their last step redone in FP64 from the beta it started at
  std::vector<JointMember*> mixed;
  for (auto* J : all)
    if (J->mixed) mixed.push_back(J);
  if (mixed.empty()) return DEVIL_OK;
  const int X = (int)mixed.size();
  DV_HIP(dv::launch_fill_i32(s, cnt.as<int32_t>(), X, 0));
  for (int q = 0; q < X; ++q) {
    JointMember& J = *mixed[q];
    ProbRun& r = *J.r;
    DV_HIP(dv::launch_polish_select(s, r.G, r.p, r.beta.as<double>(), J.bprev.as<double>(),
                                    r.status.as<int32_t>(), kPolishBelow, J.act0.as<int32_t>(),
                                    cnt.as<int32_t>() + q));
  }
  std::vector<int32_t> np(X, 0);
  DV_HIP(hipMemcpyAsync(np.data(), cnt.p, sizeof(int32_t) * X, hipMemcpyDeviceToHost, s));
  DV_HIP(hipStreamSynchronize(s));
  std::vector<JointMember*> pol;
  for (int q = 0; q < X; ++q) {
    JointMember& J = *mixed[q];
    if (np[q] <= 0) continue;
    ctx->polished += np[q];
    J.kind = dv::kJointFp64;
    J.tfull = J.r->tl.td;
    J.cur = J.act0.as<int32_t>();
    J.nxt = J.act1.as<int32_t>();
    J.nA = np[q];
    J.band = 0.0;
    J.polish = 1;
    J.abeta = J.bprev.as<double>();
    pol.push_back(&J);
  }
  if (pol.empty()) return DEVIL_OK;
  const int Pn = (int)pol.size();
  for (int q = 0; q < Pn; ++q) h[q] = (int32_t)pol[q]->nA;
  DV_HIP(hipMemcpyAsync(cnt.p, h.data(), sizeof(int32_t) * Pn, hipMemcpyHostToDevice, s));
  return joint_rounds(ctx, pol, 1, cnt.as<int32_t>(), nrep);
}

int fit_one(devil_ctx* ctx, const devil_problem* pr, devil_outputs* out) {
  ProbRun r;
  int rc;
  if ((rc = run_setup(ctx, pr, out, r))) return rc;
  if (r.G == 0) return DEVIL_OK;
  reset_fit_stats(ctx);
  StageTimer tm(ctx);
  for (int64_t g0 = 0; g0 < r.G; g0 += r.batch) {
    if ((rc = run_batch_init(ctx, r, g0, tm))) return rc;
    if ((rc = run_batch_irls(ctx, r, tm))) return rc;
    if ((rc = run_batch_post(ctx, r, tm))) return rc;
  }
  if ((rc = run_outputs(ctx, r))) return rc;
  tm.collect();
  return DEVIL_OK;
}

// Per-problem figures of a sub-context folded into its sums (each phase of a
// problem starts from reset_fit_stats).
void fold_sub_stats(devil_ctx* sub, const double* st_before) {
  sub->replay_sum += sub->replayed;
  sub->polish_sum += sub->polished;
  for (int k = 0; k < DEVIL_N_STAGES; ++k) sub->stage_ms[k] += st_before[k];
  for (int k = 0; k < DEVIL_N_KERNELS; ++k) {
    sub->k_sum_n[k] += sub->k_n[k];
    sub->k_sum_ms[k] += sub->k_ms[k];
    sub->k_sum_bytes[k] += sub->k_bytes[k];
  }
}

// devil_fit of several problems with joint IRLS rounds.  The problems are
// cut into units of equal design width (at most about n_problems / 2S
// problems each, largest first); S host threads, one sub-context and stream
// each, take units in turn.  Per unit: set-up and init pass of its problems
// (a problem that cannot join -- several gene batches, a mixed kernel other
// than the LDS-staged one -- is fitted whole there), the joint IRLS of the
// others, then their overdispersion, bread and outputs.  A unit's state lives
// in its thread's sub-context only.
int fit_many_joint(devil_ctx* ctx, const devil_problem* probs, devil_outputs* outs,
                   int n_problems, int S, const std::vector<int>& order) {
  // units: problems of one width, in decreasing size, cut to at most U each
  const int U = ctx->tune[DEVIL_TUNE_JOINT] > 1
                    ? ctx->tune[DEVIL_TUNE_JOINT]
                    : std::max(1, (n_problems + 2 * S - 1) / (2 * S));
  std::vector<std::vector<int>> units;
  {
    std::map<int, std::vector<int>> byp;
    for (int j : order) byp[probs[j].p].push_back(j);
    for (auto& kv : byp)
      for (size_t a = 0; a < kv.second.size(); a += U)
        units.emplace_back(kv.second.begin() + a,
                           kv.second.begin() + std::min(kv.second.size(), a + U));
    std::stable_sort(units.begin(), units.end(), [&](const std::vector<int>& a,
                                                     const std::vector<int>& b) {
      auto work = [&](const std::vector<int>& v) {
        double w = 0.0;
        for (int j : v) w += (double)probs[j].G * (double)probs[j].n;
        return w;
      };
      return work(a) > work(b);
    });
  }
  std::vector<int> rcs(n_problems, DEVIL_OK);
  std::vector<std::string> errs(n_problems);
  std::atomic<int> next{0};
  std::atomic<int64_t> joined_total{0};
  for (int t = 0; t < S; ++t) {
    devil_ctx* sub = ctx->subs[t];
    std::memcpy(sub->tune, ctx->tune, sizeof(ctx->tune));
    // the joint mixed rounds run the LDS-staged kernel (the autotuner then
    // picks its chunk size only)
    if (sub->tune[DEVIL_TUNE_MIXED_KERNEL] == DEVIL_MIXED_AUTO)
      sub->tune[DEVIL_TUNE_MIXED_KERNEL] = DEVIL_MIXED_BF16_WS;
    sub->tuned = ctx->tune[DEVIL_TUNE_MIXED_KERNEL] == DEVIL_MIXED_AUTO ? ctx->tuned_joint
                                                                          : ctx->tuned;
    sub->timing = ctx->timing;
    for (int k = 0; k < DEVIL_N_STAGES; ++k) sub->stage_ms[k] = 0.0;
  }
  std::vector<std::thread> th;
  for (int t = 0; t < S; ++t) {
    devil_ctx* sub = ctx->subs[t];
    th.emplace_back([&, sub] {
      for (int u; (u = next.fetch_add(1)) < (int)units.size();) {
        const std::vector<int>& unit = units[u];
        double st[DEVIL_N_STAGES];
        for (int k = 0; k < DEVIL_N_STAGES; ++k) st[k] = sub->stage_ms[k];
        reset_fit_stats(sub);
        StageTimer tm(sub);
        std::vector<std::unique_ptr<ProbRun>> runs;
        std::vector<int> jids;
        std::vector<ProbRun*> jr;
        int rc = DEVIL_OK, bad = -1;
        for (int j : unit) {
          sub->err.clear();
          // a unit's problems stay alive together: only its first (largest)
          // may read the sub-context's padded copy
          sub->tune[DEVIL_TUNE_PAD_COPY] = j == unit.front() ? ctx->tune[DEVIL_TUNE_PAD_COPY] : 0;
          std::unique_ptr<ProbRun> r(new (std::nothrow) ProbRun());
          if (!r) rc = fail(sub, DEVIL_E_NOMEM, "problem state");
          else rc = run_setup(sub, &probs[j], &outs[j], *r);
          if (rc == DEVIL_OK && r->G > 0) {
            if (r->batch < r->G) {
              for (int64_t g0 = 0; g0 < r->G && rc == DEVIL_OK; g0 += r->batch) {
                if ((rc = run_batch_init(sub, *r, g0, tm))) break;
                if ((rc = run_batch_irls(sub, *r, tm))) break;
                rc = run_batch_post(sub, *r, tm);
              }
              if (rc == DEVIL_OK) rc = run_outputs(sub, *r);
            } else if ((rc = run_batch_init(sub, *r, 0, tm)) == DEVIL_OK) {
              if (joint_eligible(*r)) {
                jids.push_back(j);
                jr.push_back(r.get());
                runs.push_back(std::move(r));
              } else if ((rc = run_batch_irls(sub, *r, tm)) == DEVIL_OK &&
                         (rc = run_batch_post(sub, *r, tm)) == DEVIL_OK) {
                rc = run_outputs(sub, *r);
              }
            }
          }
          if (rc != DEVIL_OK) {
            bad = j;
            break;
          }
        }
        if (rc == DEVIL_OK && !jr.empty()) {
          joined_total += (int64_t)jr.size();
          hipEvent_t ev = tm.begin(DEVIL_STAGE_IRLS);
          sub->err.clear();
          rc = run_irls_joint(sub, jr);
          tm.end(ev);
          if (rc != DEVIL_OK) bad = jids.front();
          for (size_t q = 0; q < jr.size() && rc == DEVIL_OK; ++q) {
            sub->err.clear();
            if ((rc = run_batch_post(sub, *jr[q], tm)) == DEVIL_OK) rc = run_outputs(sub, *jr[q]);
            if (rc != DEVIL_OK) bad = jids[q];
          }
        }
        if (rc == DEVIL_OK && hipStreamSynchronize(sub->stream) != hipSuccess)
          rc = fail(sub, DEVIL_E_HIP, "hipStreamSynchronize");
        if (rc == DEVIL_OK) tm.collect();
        fold_sub_stats(sub, st);
        runs.clear();
        if (rc != DEVIL_OK) {
          if (bad < 0) bad = unit.front();
          rcs[bad] = rc;
          errs[bad] = sub->err;
          return;
        }
      }
    });
  }
  for (auto& x : th) x.join();
  ctx->joined = joined_total.load();
  for (int j = 0; j < n_problems; ++j)
    if (rcs[j] != DEVIL_OK) {
      ctx->err = "problem " + std::to_string(j) + ": " + errs[j];
      return rcs[j];
    }
  return DEVIL_OK;
}

}  // namespace

extern "C" {

// Several problems: each runs as a whole fit on its own stream (one host
// thread per stream, up to DEVIL_TUNE_CONCURRENCY at a time), so the kernels
// of different problems overlap on the GPU; larger problems start first.
int devil_fit(devil_ctx* ctx, const devil_problem* probs, devil_outputs* outs,
              int32_t n_problems) {
  int rc;
  if ((rc = set_device(ctx))) return rc;
  if (n_problems < 0 || (n_problems > 0 && (!probs || !outs)))
    return fail(ctx, DEVIL_E_ARG, "NULL problems or outputs");
  if (n_problems == 0) return DEVIL_OK;
  for (int k = 0; k < DEVIL_STAGE_TEST_DE; ++k) ctx->stage_ms[k] = 0.0;
  for (int k = 0; k < DEVIL_N_KERNELS; ++k) {
    ctx->k_n[k] = 0;
    ctx->k_ms[k] = ctx->k_bytes[k] = 0.0;
  }
  ctx->replayed = 0;
  ctx->polished = 0;
  ctx->joined = 0;
  if (n_problems == 1) return fit_one(ctx, probs, outs);
  const int S = std::max(1, std::min<int>(n_problems, ctx->tune[DEVIL_TUNE_CONCURRENCY]));
  while ((int)ctx->subs.size() < S) {
    devil_ctx* sub = new (std::nothrow) devil_ctx();
    if (!sub) return fail(ctx, DEVIL_E_NOMEM, "sub-context");
    sub->device = ctx->device;
    if (hipStreamCreateWithFlags(&sub->stream, hipStreamNonBlocking) != hipSuccess) {
      delete sub;
      return fail(ctx, DEVIL_E_HIP, "hipStreamCreate");
    }
    ctx->subs.push_back(sub);
  }
  std::vector<int> order(n_problems);
  for (int j = 0; j < n_problems; ++j) order[j] = j;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    return probs[a].G * probs[a].n > probs[b].G * probs[b].n;
  });
  if (ctx->tune[DEVIL_TUNE_JOINT]) {
    const int jrc = fit_many_joint(ctx, probs, outs, n_problems, S, order);
    for (int t = 0; t < S; ++t) {
      devil_ctx* sub = ctx->subs[t];
      for (int k = 0; k < DEVIL_STAGE_TEST_DE; ++k) ctx->stage_ms[k] += sub->stage_ms[k];
      for (int k = 0; k < DEVIL_N_KERNELS; ++k) {
        ctx->k_n[k] += sub->k_sum_n[k];
        ctx->k_ms[k] += sub->k_sum_ms[k];
        ctx->k_bytes[k] += sub->k_sum_bytes[k];
        sub->k_sum_n[k] = 0;
        sub->k_sum_ms[k] = sub->k_sum_bytes[k] = 0.0;
      }
      // choices made with the kernel fixed to the joint rounds' stay apart
      auto& into = ctx->tune[DEVIL_TUNE_MIXED_KERNEL] == DEVIL_MIXED_AUTO ? ctx->tuned_joint
                                                                          : ctx->tuned;
      for (auto& kv : sub->tuned) into[kv.first] = kv.second;
      ctx->replayed += sub->replay_sum;
      ctx->polished += sub->polish_sum;
      sub->replay_sum = 0;
      sub->polish_sum = 0;
      if (sub->pad_p) (void)hipFree(sub->pad_p);
      sub->pad_p = nullptr;
      sub->pad_size = 0;
    }
    return jrc;
  }
  std::atomic<int> next{0};
  std::vector<int> rcs(n_problems, DEVIL_OK);
  std::vector<std::string> errs(n_problems);
  std::vector<std::thread> th;
  for (int t = 0; t < S; ++t) {
    devil_ctx* sub = ctx->subs[t];
    std::memcpy(sub->tune, ctx->tune, sizeof(ctx->tune));
    sub->tuned = ctx->tuned;
    sub->timing = ctx->timing;
    for (int k = 0; k < DEVIL_N_STAGES; ++k) sub->stage_ms[k] = 0.0;
    th.emplace_back([&, sub] {
      for (int q; (q = next.fetch_add(1)) < n_problems;) {
        const int j = order[q];
        for (int k = 0; k < DEVIL_N_KERNELS; ++k) {
          sub->k_n[k] = 0;
          sub->k_ms[k] = sub->k_bytes[k] = 0.0;
        }
        double st[DEVIL_N_STAGES] = {0, 0, 0, 0, 0};
        for (int k = 0; k < DEVIL_N_STAGES; ++k) st[k] = sub->stage_ms[k];
        sub->err.clear();
        rcs[j] = fit_one(sub, &probs[j], &outs[j]);
        if (rcs[j] != DEVIL_OK) errs[j] = sub->err;
        sub->replay_sum += sub->replayed;
        sub->polish_sum += sub->polished;
        // fit_one resets the sub-context's stage and kernel figures: keep sums
        for (int k = 0; k < DEVIL_N_STAGES; ++k) sub->stage_ms[k] += st[k];
        for (int k = 0; k < DEVIL_N_KERNELS; ++k) {
          sub->k_sum_n[k] += sub->k_n[k];
          sub->k_sum_ms[k] += sub->k_ms[k];
          sub->k_sum_bytes[k] += sub->k_bytes[k];
        }
        if (rcs[j] != DEVIL_OK) return;
      }
    });
  }
  for (auto& x : th) x.join();
  for (int t = 0; t < S; ++t) {
    devil_ctx* sub = ctx->subs[t];
    for (int k = 0; k < DEVIL_STAGE_TEST_DE; ++k) ctx->stage_ms[k] += sub->stage_ms[k];
    for (int k = 0; k < DEVIL_N_KERNELS; ++k) {
      ctx->k_n[k] += sub->k_sum_n[k];
      ctx->k_ms[k] += sub->k_sum_ms[k];
      ctx->k_bytes[k] += sub->k_sum_bytes[k];
      sub->k_sum_n[k] = 0;
      sub->k_sum_ms[k] = sub->k_sum_bytes[k] = 0.0;
    }
    for (auto& kv : sub->tuned) ctx->tuned[kv.first] = kv.second;
    ctx->replayed += sub->replay_sum;
    ctx->polished += sub->polish_sum;
    sub->replay_sum = 0;
    sub->polish_sum = 0;
    // the streams' aligned copies are not kept (one per stream would add up)
    if (sub->pad_p) (void)hipFree(sub->pad_p);
    sub->pad_p = nullptr;
    sub->pad_size = 0;
  }
  // the first failing problem (in problem order) names the error
  for (int j = 0; j < n_problems; ++j)
    if (rcs[j] != DEVIL_OK) {
      ctx->err = "problem " + std::to_string(j) + ": " + errs[j];
      return rcs[j];
    }
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
  if (key == DEVIL_TUNE_MIXED_KERNEL && value != DEVIL_MIXED_F32 && value != DEVIL_MIXED_BF16 &&
      value != DEVIL_MIXED_AUTO && value != DEVIL_MIXED_BF16_WS && value != DEVIL_MIXED_BF16_WS2)
    return fail(ctx, DEVIL_E_ARG, "unknown mixed kernel");
  if (key == DEVIL_TUNE_CHUNK_CELLS && value != 0 && (value < 256 || value > (1 << 20)))
    return fail(ctx, DEVIL_E_ARG, "chunk cells must be 0 (autotune) or in [256, 2^20]");
  if (key == DEVIL_TUNE_CONCURRENCY && (value < 1 || value > 16))
    return fail(ctx, DEVIL_E_ARG, "concurrency must be in [1, 16]");
  if (key == DEVIL_TUNE_MLE_DRIVER && value != DEVIL_MLE_NLMINB && value != DEVIL_MLE_ROOT)
    return fail(ctx, DEVIL_E_ARG, "unknown MLE driver");
  if (key == DEVIL_TUNE_PAD_COPY && value != 0 && value != 1)
    return fail(ctx, DEVIL_E_ARG, "pad copy must be 0 or 1");
  if (key == DEVIL_TUNE_GRAPH && value != 0 && value != 1)
    return fail(ctx, DEVIL_E_ARG, "graph must be 0 or 1");
  if (key == DEVIL_TUNE_BATCH_GENES && (value < 0 || value > INT32_MAX))
    return fail(ctx, DEVIL_E_ARG, "batch genes must be 0 (no cap) or positive");
  if (key == DEVIL_TUNE_JOINT && (value < 0 || value > 4096))
    return fail(ctx, DEVIL_E_ARG, "joint must be 0 (off), 1 (auto) or a unit size <= 4096");
  if (key == DEVIL_TUNE_PAD_COPY && value == 0 && ctx->pad_p) {
    (void)hipSetDevice(ctx->device);
    DV_HIP(hipFree(ctx->pad_p));
    ctx->pad_p = nullptr;
    ctx->pad_size = 0;
  }
  ctx->tune[key] = (int)value;
  return DEVIL_OK;
}

int devil_fit_info(const devil_ctx* ctx, int32_t key, int64_t* value) {
  if (!ctx || !value) return DEVIL_E_ARG;
  switch (key) {
    case DEVIL_INFO_REPLAYED: *value = ctx->replayed; return DEVIL_OK;
    case DEVIL_INFO_MIXED_KERNEL:
      *value = ctx->last_mk == kMixedBf16     ? DEVIL_MIXED_BF16
               : ctx->last_mk == kMixedBf16Ws ? DEVIL_MIXED_BF16_WS
               : ctx->last_mk == kMixedBf16Ws2 ? DEVIL_MIXED_BF16_WS2
               : ctx->last_mk == kMixedF32    ? DEVIL_MIXED_F32
                                              : -1;
      return DEVIL_OK;
    case DEVIL_INFO_CHUNK_CELLS: *value = ctx->last_cc; return DEVIL_OK;
    case DEVIL_INFO_POLISHED: *value = ctx->polished; return DEVIL_OK;
    case DEVIL_INFO_ALIGNED_COPY: *value = ctx->aligned_copy; return DEVIL_OK;
    case DEVIL_INFO_JOINED: *value = ctx->joined; return DEVIL_OK;
    default: return DEVIL_E_ARG;
  }
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

int devil_size_factors(devil_ctx* ctx, const void* counts, int32_t dtype, int64_t G,
                       int64_t n, int64_t ld, int32_t method, double* sf) {
  int rc;
  if ((rc = set_device(ctx))) return rc;
  if ((rc = check_counts_args(ctx, counts, dtype, G, n, ld))) return rc;
  if (n > 0 && !sf) return fail(ctx, DEVIL_E_ARG, "NULL pointer");
  if (method == DEVIL_SF_EDGER)
    return fail(ctx, DEVIL_E_UNSUPPORTED,
                "the \"edgeR\" method (TMM with singleton pairing) needs the edgeR package; "
                "use normed_sum or psinorm");
  if (method != DEVIL_SF_NORMED_SUM && method != DEVIL_SF_PSINORM)
    return fail(ctx, DEVIL_E_ARG, "unknown size factor method");
  if (n == 0) return DEVIL_OK;
  hipStream_t s = ctx->stream;
  if (G <= 1) {  // R/size_factors.R:38-42: unit factors
    DV_HIP(dv::launch_fill_f64(s, sf, n, 1.0));
    DV_HIP(hipStreamSynchronize(s));
    return DEVIL_OK;
  }
  // enough (cell block, gene slice) workgroups to fill the GPU
  const int64_t cell_blocks = (n + 1023) / 1024;
  const int slices = (int)std::max<int64_t>(
      1, std::min<int64_t>({(4096 + cell_blocks - 1) / cell_blocks, G, 256}));
  DevBuf part, dmin, logsf, flag;
  DV_ALLOC(part, sizeof(double) * (size_t)slices * n);
  if (method == DEVIL_SF_PSINORM) DV_ALLOC(dmin, sizeof(double) * (size_t)slices * n);
  DV_ALLOC(logsf, sizeof(double) * n);
  DV_ALLOC(flag, sizeof(int32_t));
  DV_HIP(dv::launch_fill_i32(s, flag.as<int32_t>(), 1, 0));
  dv::Counts c{counts, dtype, ld};
  DV_HIP(dv::launch_size_factors(s, c, G, n, method, slices, part.as<int64_t>(),
                                 part.as<double>(), dmin.as<double>(), sf, logsf.as<double>(),
                                 flag.as<int32_t>()));
  int32_t zero = 0;
  DV_HIP(hipMemcpyAsync(&zero, flag.p, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  DV_HIP(hipStreamSynchronize(s));
  if (zero)
    return fail(ctx, DEVIL_E_ARG,
                "At least one column (sample) contains all zeros, unable to compute size "
                "factor. Please filter out empty samples from input matrix.");
  return DEVIL_OK;
}

int devil_group_cells(devil_ctx* ctx, const void* counts, int32_t dtype, int64_t G, int64_t n,
                      int64_t ld, const int64_t* order, void* out, int64_t out_ld) {
  int rc;
  if ((rc = set_device(ctx))) return rc;
  if ((rc = check_counts_args(ctx, counts, dtype, G, n, ld))) return rc;
  if (G > 0 && n > 0 && (!order || !out)) return fail(ctx, DEVIL_E_ARG, "NULL pointer");
  if (out_ld < n) return fail(ctx, DEVIL_E_ARG, "out_ld < n");
  if (G == 0 || n == 0) return DEVIL_OK;
  hipStream_t s = ctx->stream;
  DevBuf flag;
  DV_ALLOC(flag, sizeof(int32_t));
  DV_HIP(dv::launch_fill_i32(s, flag.as<int32_t>(), 1, 0));
  DV_HIP(dv::launch_check_order(s, order, n, flag.as<int32_t>()));
  int32_t bad = 0;
  DV_HIP(hipMemcpyAsync(&bad, flag.p, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  DV_HIP(hipStreamSynchronize(s));
  if (bad) return fail(ctx, DEVIL_E_ARG, "order holds a cell index outside [0, n)");
  dv::Counts c{counts, dtype, ld};
  DV_HIP(dv::launch_group_cells(s, c, G, n, order, out, out_ld));
  DV_HIP(hipStreamSynchronize(s));
  return DEVIL_OK;
}

int devil_csc_to_dense(devil_ctx* ctx, const int64_t* colptr, const int32_t* rowind,
                       const double* values, int64_t G, int64_t n, int32_t out_dtype,
                       void* out, int64_t out_ld) {
  int rc;
  if ((rc = set_device(ctx))) return rc;
  if (G < 0 || n < 0) return fail(ctx, DEVIL_E_ARG, "negative dimension");
  if (out_dtype != DEVIL_COUNTS_I32 && out_dtype != DEVIL_COUNTS_F32 &&
      out_dtype != DEVIL_COUNTS_F64)
    return fail(ctx, DEVIL_E_ARG, "unknown output dtype");
  if (G == 0 || n == 0) return DEVIL_OK;
  if (!colptr || !out) return fail(ctx, DEVIL_E_ARG, "NULL pointer");
  if (out_ld < n) return fail(ctx, DEVIL_E_ARG, "out_ld < n");
  hipStream_t s = ctx->stream;
  int64_t nnz = 0;
  DV_HIP(hipMemcpyAsync(&nnz, colptr + n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  DV_HIP(hipStreamSynchronize(s));
  if (nnz < 0) return fail(ctx, DEVIL_E_ARG, "colptr[n] < 0");
  if (nnz > 0 && (!rowind || !values)) return fail(ctx, DEVIL_E_ARG, "NULL pointer");
  DevBuf flag;
  DV_ALLOC(flag, sizeof(int32_t));
  DV_HIP(dv::launch_fill_i32(s, flag.as<int32_t>(), 1, 0));
  DV_HIP(dv::launch_check_csc(s, colptr, rowind, G, n, nnz, flag.as<int32_t>()));
  int32_t bad = 0;
  DV_HIP(hipMemcpyAsync(&bad, flag.p, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  DV_HIP(hipStreamSynchronize(s));
  if (bad & 1) return fail(ctx, DEVIL_E_ARG, "colptr not non-decreasing from 0 to nnz");
  if (bad & 2) return fail(ctx, DEVIL_E_ARG, "a row index lies outside [0, G)");
  DV_HIP(dv::launch_csc_to_dense(s, colptr, rowind, values, G, n, out_dtype, out, out_ld));
  DV_HIP(hipStreamSynchronize(s));
  return DEVIL_OK;
}

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
  if (G == 0) return DEVIL_OK;
  Design d;
  d.n = n;
  d.p = 1;
  d.moi = 1.0 / (se / (double)n);
  d.eo.resize(n);
  for (int64_t i = 0; i < n; ++i) d.eo[i] = std::exp(off[i]);
  dv::TileDesign td{};
  td.n = n;
  td.nblk = (n + 15) / 16;
  DeviceDesign dd;
  if ((rc = upload(ctx, dd.eo, d.eo))) return rc;
  InitHost ih;
  if ((rc = build_init_tile(ctx, d, dd, td, 1, dv::kInitNone, ih))) return rc;
  ChunkHost chunks;
  if ((rc = build_chunks(ctx, n, nullptr, 0, chunks))) return rc;
  DevBuf part;
  DV_ALLOC(part, sizeof(double) * (size_t)((G + 15) / 16) * chunks.ct.nchunk * 16 * 20);
  dv::Counts c{counts, dtype, ld};
  DV_HIP(dv::launch_init_tiled(ctx->stream, td, ih.it, chunks.ct, c, G, n, 1,
                               part.as<double>(), alpha0, nullptr));
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
  DeviceDesign dd;
  if ((rc = prepare_design(ctx, X, offset, n, p, DEVIL_MEM_DEVICE, init_mode, d, dd))) return rc;
  if (G == 0) return DEVIL_OK;
  dv::TileDesign td{};
  td.n = n;
  td.nblk = (n + 15) / 16;
  InitHost ih;
  if ((rc = build_init_tile(ctx, d, dd, td, 0, init_mode, ih))) return rc;
  ChunkHost chunks;
  if ((rc = build_chunks(ctx, n, nullptr, 0, chunks))) return rc;
  DevBuf tmp, part;
  DV_ALLOC(tmp, sizeof(double) * G * p);
  DV_ALLOC(part, sizeof(double) * (size_t)((G + 15) / 16) * chunks.ct.nchunk * 16 * 20);
  dv::Counts c{counts, dtype, ld};
  DV_HIP(dv::launch_init_tiled(ctx->stream, td, ih.it, chunks.ct, c, G, n, p, part.as<double>(),
                               nullptr, tmp.as<double>()));
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
  TileHost th;
  if ((rc = tile_from_device(ctx, X, offset, n, p, th,
                             precision == DEVIL_PREC_MIXED ? mixed_kernel(ctx) : kMixedNone)))
    return rc;
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
  TileHost th;
  if ((rc = tile_from_device(ctx, X, offset, n, p, th))) return rc;
  ChunkHost chunks;
  if ((rc = build_chunks(ctx, n, nullptr, 0, chunks))) return rc;
  DevBuf bg, post;
  DV_ALLOC(bg, sizeof(double) * G * p);
  DV_ALLOC(post, sizeof(double) * post_part_doubles(th.td, chunks, G));
  DV_HIP(dv::launch_transpose_gp(s, beta, G, p, 1, bg.as<double>()));
  dv::Counts c{counts, dtype, ld};
  DV_HIP(dv::launch_mom_tiled(s, th.td, chunks.ct, c, G, n, p, bg.as<double>(),
                              post.as<double>(), theta, nullptr));
  DV_HIP(hipStreamSynchronize(s));
  return DEVIL_OK;
}

int devil_fit_dispersion_mle(devil_ctx* ctx, const void* counts, int32_t dtype, int64_t G,
                             int64_t n, int64_t ld, const double* X, int32_t p,
                             const double* beta, const double* offset, int32_t max_iter,
                             double tolerance, double* theta, int32_t* iters) {
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
  TileHost th;
  if ((rc = tile_from_device(ctx, X, offset, n, p, th))) return rc;
  if ((rc = run_mle_tiled(ctx, th.td, c, G, n, p, bg.as<double>(), max_iter, tolerance, theta,
                          iters, st.as<int32_t>())))
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
  std::vector<int32_t> ends_h(n_clusters);
  if (n_clusters > 0)
    DV_HIP(hipMemcpy(ends_h.data(), cluster_ends, sizeof(int32_t) * n_clusters,
                     hipMemcpyDeviceToHost));
  TileHost th;
  if ((rc = tile_from_device(ctx, X, offset, n, p, th))) return rc;
  ChunkHost chunks;
  if ((rc = build_chunks(ctx, n, ends_h.data(), n_clusters, chunks))) return rc;
  DevBuf bg, post;
  DV_ALLOC(bg, sizeof(double) * G * p);
  DV_ALLOC(post, sizeof(double) * post_part_doubles(th.td, chunks, G));
  DV_HIP(dv::launch_transpose_gp(s, beta, G, p, 1, bg.as<double>()));
  dv::Counts c{counts, dtype, ld};
  DV_HIP(dv::launch_bread_tiled(s, th.td, chunks.ct, c, G, n, p, bg.as<double>(), theta,
                                n_clusters, post.as<double>(), bread,
                                n_clusters > 0 ? sandwich : nullptr, status));
  DV_HIP(hipStreamSynchronize(s));
  return DEVIL_OK;
}

int devil_sandwich_hc0(devil_ctx* ctx, const void* counts, int32_t dtype, int64_t G, int64_t n,
                       int64_t ld, const double* X, int32_t p, const double* beta,
                       const double* theta, const double* offset, double* bread,
                       double* sandwich, int32_t* status) {
  int rc;
  if ((rc = set_device(ctx))) return rc;
  if ((rc = check_p(ctx, p))) return rc;
  if ((rc = check_counts_args(ctx, counts, dtype, G, n, ld))) return rc;
  if (!X || !offset || (G > 0 && (!beta || !theta || !bread || !sandwich)))
    return fail(ctx, DEVIL_E_ARG, "NULL pointer");
  if (G == 0) return DEVIL_OK;
  hipStream_t s = ctx->stream;
  TileHost th;
  if ((rc = tile_from_device(ctx, X, offset, n, p, th))) return rc;
  ChunkHost chunks;
  if ((rc = build_chunks(ctx, n, nullptr, 0, chunks))) return rc;
  DevBuf bg, post;
  DV_ALLOC(bg, sizeof(double) * G * p);
  DV_ALLOC(post, sizeof(double) * post_part_doubles(th.td, chunks, G, true));
  DV_HIP(dv::launch_transpose_gp(s, beta, G, p, 1, bg.as<double>()));
  dv::Counts c{counts, dtype, ld};
  DV_HIP(dv::launch_bread_tiled(s, th.td, chunks.ct, c, G, n, p, bg.as<double>(), theta, -1,
                                post.as<double>(), bread, sandwich, status));
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
  DevBuf bg, bh;
  DV_ALLOC(bg, sizeof(double) * G * p);
  const size_t bh_bytes = dv::bh_scratch_bytes(G);
  DV_ALLOC(bh, bh_bytes);
  ctx->stage_ms[DEVIL_STAGE_TEST_DE] = 0.0;
  StageTimer tm(ctx);
  hipEvent_t ev = tm.begin(DEVIL_STAGE_TEST_DE);
  DV_HIP(dv::launch_transpose_gp(s, beta, G, p, 1, bg.as<double>()));
  DV_HIP(dv::launch_test_de(s, bg.as<double>(), bread, sandwich, G, p, (double)n - 2.0,
                            contrast, max_lfc, pval, log_pval, lfc, bh.p, bh_bytes,
                            adj_pval));
  tm.end(ev);
  DV_HIP(hipStreamSynchronize(s));
  tm.collect();
  return DEVIL_OK;
}

int devil_p_adjust(devil_ctx* ctx, const double* pval, int64_t G, int32_t method,
                   double* adj) {
  int rc;
  if ((rc = set_device(ctx))) return rc;
  if (G < 0 || (G > 0 && (!pval || !adj))) return fail(ctx, DEVIL_E_ARG, "NULL pointer or G < 0");
  if (method < DEVIL_PADJ_BH || method > DEVIL_PADJ_NONE)
    return fail(ctx, DEVIL_E_ARG, "unknown p-value adjustment method " + std::to_string(method));
  if (G > INT32_MAX) return fail(ctx, DEVIL_E_ARG, "G above 2^31 - 1");
  if (G == 0) return DEVIL_OK;
  hipStream_t s = ctx->stream;
  DevBuf bh;
  const size_t bh_bytes = dv::bh_scratch_bytes(G);
  DV_ALLOC(bh, bh_bytes);
  dv::PadjBufs b{};
  DV_HIP(dv::launch_padj_sort(s, pval, G, bh.p, bh_bytes, b));
  int32_t m = 0;
  DV_HIP(hipMemcpyAsync(&m, b.m, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  DV_HIP(hipStreamSynchronize(s));
  // BY's q = sum(1 / (1:n)), accumulated in long double as R's sum does
  double q = 1.0;
  if (method == DEVIL_PADJ_BY) {
    long double acc = 0.0L;
    for (int32_t k = 1; k <= m; ++k) acc += (long double)(1.0 / (double)k);
    q = (double)acc;
  }
  DV_HIP(dv::launch_padj_finish(s, pval, G, m, method, q, b, adj));
  DV_HIP(hipStreamSynchronize(s));
  return DEVIL_OK;
}

// devil_test_de with its scratch (the gene-major beta and the BH sort
// buffers) in the device's fixed scratch array: no allocation, no events.
int devil_test_de_noalloc(devil_ctx* ctx, const double* beta, const double* bread,
                          const double* sandwich, int64_t G, int32_t p, int64_t n,
                          const double* contrast, double max_lfc, double* pval,
                          double* log_pval, double* adj_pval, double* lfc) {
  int rc;
  if ((rc = set_device(ctx))) return rc;
  if ((rc = check_p(ctx, p))) return rc;
  if (G < 0 || (G > 0 && (!beta || !bread || !contrast || !pval || !adj_pval || !lfc)))
    return fail(ctx, DEVIL_E_ARG, "NULL pointer or G < 0");
  if (G == 0) return DEVIL_OK;
  if (!ctx->scratch) return fail(ctx, DEVIL_E_HIP, "no scratch array on this device");
  const size_t bg_bytes = ((sizeof(double) * (size_t)G * (size_t)p + 255) / 256) * 256;
  const size_t bh_bytes = dv::bh_scratch_bytes(G);
  if (bg_bytes + bh_bytes > kNoAllocScratchBytes)
    return fail(ctx, DEVIL_E_ARG,
                "G = " + std::to_string(G) + " needs " + std::to_string(bg_bytes + bh_bytes) +
                    " scratch bytes; devil_test_de_noalloc holds " +
                    std::to_string(kNoAllocScratchBytes));
  hipStream_t s = ctx->stream;
  unsigned char* base = static_cast<unsigned char*>(ctx->scratch);
  double* bg = reinterpret_cast<double*>(base);
  DV_HIP(dv::launch_transpose_gp(s, beta, G, p, 1, bg));
  DV_HIP(dv::launch_test_de(s, bg, bread, sandwich, G, p, (double)n - 2.0, contrast, max_lfc,
                            pval, log_pval, lfc, base + bg_bytes, bh_bytes, adj_pval));
  DV_HIP(hipStreamSynchronize(s));
  return DEVIL_OK;
}

}  // extern "C"
