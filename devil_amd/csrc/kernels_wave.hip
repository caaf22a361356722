// gfx950 kernels, one 64-lane wavefront per gene.
//
// Every kernel streams a gene's count row with coalesced loads (lane l reads
// cells l, l+64, ...), accumulates its per-lane partial sums in registers,
// and combines them with a fixed butterfly (dv::wave_sum), so results are
// bitwise reproducible run to run.  The p x p algebra then runs redundantly
// in every lane, which keeps the whole gene inside one wavefront: no LDS, no
// barriers, no atomics.  Grids are one 64-thread block per gene; the
// dispatcher balances genes of unequal cost.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <hipcub/hipcub.hpp>

#include "../../include/devil_hip.h"
#include "device_common.hpp"
#include "kernels.hpp"

namespace dv {

// ---------------------------------------------------------------------------
// init pass: estimate_dispersion (R/overdispersion.R:129-135) and beta_0
// (R/beta.R:13-20, R/main.R:404-406, src/beta_init_univariate.cpp:13-95)
// ---------------------------------------------------------------------------
template <int P>
__global__ __launch_bounds__(64) void k_init(Counts c, int64_t G, int64_t n, InitParams ip,
                                             double* __restrict__ alpha0,
                                             double* __restrict__ beta0) {
  const int64_t g = blockIdx.x;
  if (g >= G) return;
  const int lane = threadIdx.x;
  const void* row = count_row(c.ptr, c.dtype, g, c.ld);
  double s1 = 0.0;
  double w[P];
#pragma unroll
  for (int j = 0; j < P; ++j) w[j] = 0.0;
  for (int64_t i = lane; i < n; i += kWave) {
    const double y = load_count(row, c.dtype, i);
    s1 += y;
    if (ip.init_mode == kInitQR) {
      const double z = log1p(y / ip.eo[i]);
#pragma unroll
      for (int j = 0; j < P; ++j) w[j] += ip.Q[i + j * n] * z;
    } else if (ip.init_mode == kInitUnivariate) {
      const double lr = log(y + 0.5) - ip.logsf[i];
#pragma unroll
      for (int j = 0; j < P; ++j) w[j] += (ip.X[i + j * n] - ip.xmean[j]) * lr;
    }
  }
  s1 = wave_sum(s1);
  const double mean = s1 / (double)n;
  if (ip.want_alpha) {
    double v = 0.0;
    for (int64_t i = lane; i < n; i += kWave) {
      const double d = load_count(row, c.dtype, i) - mean;
      v += d * d;
    }
    v = wave_sum(v) / (double)(n - 1);
    const double d = (v - ip.moi * mean) / (mean * mean);
    if (lane == 0) alpha0[g] = (isnan(d) || d < 0.0) ? 0.01 : d;
  }
  if (ip.init_mode == kInitNone) return;
#pragma unroll
  for (int j = 0; j < P; ++j) w[j] = wave_sum(w[j]);
  double b[P];
  if (ip.init_mode == kInitQR) {
    // solve(R, Q'z) by back substitution
#pragma unroll
    for (int a = P - 1; a >= 0; --a) {
      double s = w[a];
#pragma unroll
      for (int k = a + 1; k < P; ++k) s -= ip.R[a + k * P] * b[k];
      b[a] = s / ip.R[a + a * P];
    }
  } else if (ip.init_mode == kInitRough) {
    b[0] = log1p(mean);
#pragma unroll
    for (int j = 1; j < P; ++j) b[j] = 0.0;
  } else {
#pragma unroll
    for (int j = 0; j < P; ++j) b[j] = w[j] / ip.denom[j];
    if (ip.intercept_first) b[0] = log((s1 + 1e-8) / (ip.sum_sf + 1e-8));
  }
  if (lane == 0) {
#pragma unroll
    for (int j = 0; j < P; ++j) beta0[g * P + j] = b[j];
  }
}

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
// IRLS, src/beta.cpp:37-47:
//   w = exp(-X beta - off); m = (k + y) / (1 + k w)
//   delta = (k X' diag(m w) X)^-1 (k X' (m w - 1)); beta += delta
//   stop when max|delta| < eps or delta[0] is NaN, or at max_iter.
// ---------------------------------------------------------------------------
template <int P>
__global__ __launch_bounds__(64) void k_irls(Counts c, int64_t G, int64_t n,
                                             const double* __restrict__ X,
                                             const double* __restrict__ off,
                                             const double* __restrict__ kvec,
                                             double* __restrict__ beta, int max_iter,
                                             double eps, int32_t* __restrict__ iters,
                                             int32_t* __restrict__ status) {
  constexpr int NT = P * (P + 1) / 2;
  const int64_t g = blockIdx.x;
  if (g >= G) return;
  const int lane = threadIdx.x;
  const void* row = count_row(c.ptr, c.dtype, g, c.ld);
  const double k = kvec[g];
  double b[P];
#pragma unroll
  for (int j = 0; j < P; ++j) b[j] = beta[g * P + j];
  int iter = 0, st = 0;
  bool conv = false;
  while (!conv && iter < max_iter) {
    double A[NT], r[P];
#pragma unroll
    for (int t = 0; t < NT; ++t) A[t] = 0.0;
#pragma unroll
    for (int j = 0; j < P; ++j) r[j] = 0.0;
    for (int64_t i = lane; i < n; i += kWave) {
      double xi[P];
      double xb = 0.0;
#pragma unroll
      for (int j = 0; j < P; ++j) {
        xi[j] = X[i + j * n];
        xb += xi[j] * b[j];
      }
      const double wq = exp(-xb - off[i]);
      const double y = load_count(row, c.dtype, i);
      const double m = (k + y) / (1.0 + k * wq);
      const double q = m * wq;
      const double qm1 = q - 1.0;
#pragma unroll
      for (int a = 0; a < P; ++a) {
        r[a] += xi[a] * qm1;
        const double xq = xi[a] * q;
#pragma unroll
        for (int bb = a; bb < P; ++bb) A[tri<P>(a, bb)] += xq * xi[bb];
      }
    }
    double M[P][P], v[P];
#pragma unroll
    for (int a = 0; a < P; ++a) {
      v[a] = k * wave_sum(r[a]);
#pragma unroll
      for (int bb = a; bb < P; ++bb) {
        const double s = k * wave_sum(A[tri<P>(a, bb)]);
        M[a][bb] = s;
        M[bb][a] = s;
      }
    }
    int perm[P];
    if (lu_factor<P>(M, perm)) st |= DEVIL_GENE_SINGULAR;
    lu_solve<P>(M, perm, v);
    double mx = 0.0;
    bool has_nan = false;
#pragma unroll
    for (int j = 0; j < P; ++j) {
      b[j] += v[j];
      const double d = fabs(v[j]);
      if (isnan(d)) has_nan = true;
      else mx = fmax(mx, d);
    }
    conv = !has_nan && mx < eps;
    ++iter;
    if (isnan(v[0])) {
      conv = true;
      st |= DEVIL_GENE_NAN;
    }
  }
  if (!conv) st |= DEVIL_GENE_MAXITER;
  if (lane == 0) {
#pragma unroll
    for (int j = 0; j < P; ++j) beta[g * P + j] = b[j];
    iters[g] = iter;
    if (status) status[g] = st;
  }
}

// ---------------------------------------------------------------------------
// MOM dispersion, src/mom_overdispersion.cpp:24-53:
//   mu = sf * exp(x'beta); theta = max(0, n/(n-p) * sum((y-mu)^2 - mu) / sum(mu^2))
// ---------------------------------------------------------------------------
template <int P>
__global__ __launch_bounds__(64) void k_mom(Counts c, int64_t G, int64_t n,
                                            const double* __restrict__ X,
                                            const double* __restrict__ beta,
                                            const double* __restrict__ eo,
                                            double* __restrict__ theta,
                                            int32_t* __restrict__ status) {
  const int64_t g = blockIdx.x;
  if (g >= G) return;
  const int lane = threadIdx.x;
  const void* row = count_row(c.ptr, c.dtype, g, c.ld);
  double b[P];
#pragma unroll
  for (int j = 0; j < P; ++j) b[j] = beta[g * P + j];
  double num = 0.0, den = 0.0;
  for (int64_t i = lane; i < n; i += kWave) {
    double eta = 0.0;
#pragma unroll
    for (int j = 0; j < P; ++j) eta += X[i + j * n] * b[j];
    const double mu = eo[i] * exp(eta);
    const double d = load_count(row, c.dtype, i) - mu;
    num += d * d - mu;
    den += mu * mu;
  }
  num = wave_sum(num);
  den = wave_sum(den);
  const double corr = (double)n / (double)(n - P);
  double th = 0.0;
  if (den > 0.0) {
    th = corr * num / den;
    if (th < 0.0) th = 0.0;
  }
  if (lane == 0) {
    theta[g] = th;
    if (status && !(den > 0.0)) status[g] |= DEVIL_GENE_MOM_DEN0;
  }
}

// ---------------------------------------------------------------------------
// bread (src/variance.cpp:22-47) and clustered meat (:186-228) in one pass:
//   mu = sf * exp(x'beta); H = sum x x' (y a + 1) mu / (1 + a mu)^2; bread = H^-1
//   S_c = sum_{i in c} x (y - mu) / (1 + mu a); meat = adj/n sum_c S_c S_c'
//   sandwich = bread meat bread n (R/main.R:491)
// ---------------------------------------------------------------------------
template <int P>
__global__ __launch_bounds__(64) void k_bread_meat(
    Counts c, int64_t G, int64_t n, const double* __restrict__ X,
    const double* __restrict__ beta, const double* __restrict__ theta,
    const double* __restrict__ eo, const int32_t* __restrict__ ends, int C,
    double* __restrict__ bread, double* __restrict__ sandwich, int32_t* __restrict__ status) {
  constexpr int NT = P * (P + 1) / 2;
  const int64_t g = blockIdx.x;
  if (g >= G) return;
  const int lane = threadIdx.x;
  const void* row = count_row(c.ptr, c.dtype, g, c.ld);
  double b[P];
#pragma unroll
  for (int j = 0; j < P; ++j) b[j] = beta[g * P + j];
  const double alpha = theta[g];
  double H[NT], Mt[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    H[t] = 0.0;
    Mt[t] = 0.0;
  }
  const int nblk = (C > 0) ? C : 1;
  int64_t start = 0;
  for (int cl = 0; cl < nblk; ++cl) {
    const int64_t end = (C > 0) ? (int64_t)ends[cl] : n;
    double S[P];
#pragma unroll
    for (int j = 0; j < P; ++j) S[j] = 0.0;
    for (int64_t i = start + lane; i < end; i += kWave) {
      double xi[P];
      double eta = 0.0;
#pragma unroll
      for (int j = 0; j < P; ++j) {
        xi[j] = X[i + j * n];
        eta += xi[j] * b[j];
      }
      const double mu = eo[i] * exp(eta);
      const double y = load_count(row, c.dtype, i);
      const double denom = 1.0 + alpha * mu;
      const double s = (y * alpha + 1.0) * mu / (denom * denom);
      const double wr = (y - mu) / (1.0 + mu * alpha);
#pragma unroll
      for (int a = 0; a < P; ++a) {
        S[a] += xi[a] * wr;
        const double xs = xi[a] * s;
#pragma unroll
        for (int bb = a; bb < P; ++bb) H[tri<P>(a, bb)] += xs * xi[bb];
      }
    }
    if (C > 0) {
#pragma unroll
      for (int j = 0; j < P; ++j) S[j] = wave_sum(S[j]);
#pragma unroll
      for (int a = 0; a < P; ++a)
#pragma unroll
        for (int bb = a; bb < P; ++bb) Mt[tri<P>(a, bb)] += S[a] * S[bb];
    }
    start = end;
  }
  double Hm[P][P], Bi[P][P];
#pragma unroll
  for (int a = 0; a < P; ++a)
#pragma unroll
    for (int bb = a; bb < P; ++bb) {
      const double s = wave_sum(H[tri<P>(a, bb)]);
      Hm[a][bb] = s;
      Hm[bb][a] = s;
    }
  const int sing = lu_inverse<P>(Hm, Bi);
  if (lane == 0) {
    double* out = bread + g * P * P;
#pragma unroll
    for (int a = 0; a < P; ++a)
#pragma unroll
      for (int bb = 0; bb < P; ++bb) out[a + bb * P] = Bi[a][bb];
    if (status && sing) status[g] |= DEVIL_GENE_SINGULAR;
  }
  if (C > 0 && sandwich) {
    const double adj = (C > 1) ? (double)C / (C - 1.0) : 1.0;
    const double sc = adj / (double)n;
    double Me[P][P], T1[P][P];
#pragma unroll
    for (int a = 0; a < P; ++a)
#pragma unroll
      for (int bb = a; bb < P; ++bb) {
        const double v = sc * Mt[tri<P>(a, bb)];
        Me[a][bb] = v;
        Me[bb][a] = v;
      }
#pragma unroll
    for (int a = 0; a < P; ++a)
#pragma unroll
      for (int bb = 0; bb < P; ++bb) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < P; ++k) s += Bi[a][k] * Me[k][bb];
        T1[a][bb] = s;
      }
    if (lane == 0) {
      double* out = sandwich + g * P * P;
#pragma unroll
      for (int a = 0; a < P; ++a)
#pragma unroll
        for (int bb = 0; bb < P; ++bb) {
          double s = 0.0;
#pragma unroll
          for (int k = 0; k < P; ++k) s += T1[a][k] * Bi[k][bb];
          out[a + bb * P] = s * (double)n;
        }
    }
  }
}

// ---------------------------------------------------------------------------
// MLE overdispersion with Cox-Reid adjustment.
// Score and its derivative in log(theta): src/overdispersion.cpp:95-178 and
// :186-242, evaluated without their cancellations (the same quantities):
//   sum psi(y + 1/theta) - n psi(1/theta)   ->  sum_i dpsi_diff(1/theta, y_i)
//   sum psi1(y + 1/theta) - n psi1(1/theta) ->  sum_i dpsi1_diff(1/theta, y_i)
//   log(1 + mu theta) - mu/(mu + 1/theta)    ->  nb_f(mu theta)
// The max_y * 1e6 < 1/theta shortcut, the min() clamp and the Cox-Reid terms
// are as in the reference.  Zero counts contribute 0 to the difference sums.
// Driver: R/overdispersion.R:25-46 special cases and start value, then the
// safeguarded Newton iteration on score(log theta) = 0:
//   bracket [log 1e-8, log 1e16]; x' = x - s/d if d < 0, else x +- 1 towards
//   sign(s); bisect when x' leaves the open bracket; stop at |x'-x| < 1e-9.
// ---------------------------------------------------------------------------
struct MlePass {
  double score, dscore;
};

template <int P>
__device__ MlePass mle_pass(const void* row, int dtype, int64_t n, const double* X,
                            const double* b, const double* off, double x, double sum_y,
                            double sum_prod_y, double max_y, bool want_d) {
  constexpr int NT = P * (P + 1) / 2;
  const int lane = threadIdx.x;
  const double theta = exp(x);
  const double tn1 = 1.0 / theta;
  double B0[NT], B1[NT], B2[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    B0[t] = 0.0;
    B1[t] = 0.0;
    B2[t] = 0.0;
  }
  double sc = 0.0, l1 = 0.0, l2 = 0.0, dg = 0.0, tg = 0.0;
  for (int64_t i = lane; i < n; i += kWave) {
    double xi[P];
    double eta = 0.0;
#pragma unroll
    for (int j = 0; j < P; ++j) {
      xi[j] = X[i + j * n];
      eta += xi[j] * b[j];
    }
    double mu = exp(eta + off[i]);
    if (mu == 0.0) mu = 1e-6;
    const double y = load_count(row, dtype, i);
    const double mt = mu * theta;
    sc += nb_f(mt) + y / (mu + tn1);
    if (y != 0.0) dg += dpsi_diff(tn1, y);
    const double w = 1.0 / (1.0 / mu + theta);
    const double dw = -1.0 * w * w;
    if (want_d) {
      if (y != 0.0) tg += dpsi1_diff(tn1, y);
      l1 += log1p(mt) + (y - mu) / (mu + tn1);
      l2 += (mu * mu * theta + y) / (1 + mt) / (1 + mt);
    }
    const double d2w = -2 * dw * w;
#pragma unroll
    for (int a = 0; a < P; ++a) {
#pragma unroll
      for (int bb = a; bb < P; ++bb) {
        const double xx = xi[a] * xi[bb];
        B0[tri<P>(a, bb)] += xx * w;
        B1[tri<P>(a, bb)] += xx * dw;
        if (want_d) B2[tri<P>(a, bb)] += xx * d2w;
      }
    }
  }
  sc = wave_sum(sc);
  dg = wave_sum(dg);
  double Bm[P][P], Dm[P][P], D2m[P][P], Bi[P][P];
#pragma unroll
  for (int a = 0; a < P; ++a)
#pragma unroll
    for (int bb = a; bb < P; ++bb) {
      const double v0 = wave_sum(B0[tri<P>(a, bb)]) + (a == bb ? 1e-6 : 0.0);
      const double v1 = wave_sum(B1[tri<P>(a, bb)]);
      Bm[a][bb] = v0;
      Bm[bb][a] = v0;
      Dm[a][bb] = v1;
      Dm[bb][a] = v1;
    }
  chol_inverse<P>(Bm, Bi);
  double tr_bdb = 0.0;
#pragma unroll
  for (int a = 0; a < P; ++a)
#pragma unroll
    for (int bb = 0; bb < P; ++bb) tr_bdb += Bi[a][bb] * Dm[bb][a];
  const double cr_s = -0.5 * tr_bdb * 0.99;
  const double corr = tn1 > 1e5 ? sum_prod_y / (2 * tn1) : 0.0;
  double dterm;
  if (max_y * 1e6 < tn1) {
    dterm = sum_y - corr;
  } else {
    dterm = fmin(dg * tn1, sum_y - corr);
  }
  MlePass r;
  r.score = sc * tn1 - dterm + cr_s * theta;
  r.dscore = 0.0;
  if (want_d) {
    tg = wave_sum(tg);
    l1 = wave_sum(l1);
    l2 = wave_sum(l2);
#pragma unroll
    for (int a = 0; a < P; ++a)
#pragma unroll
      for (int bb = a; bb < P; ++bb) {
        const double v2 = wave_sum(B2[tri<P>(a, bb)]);
        D2m[a][bb] = v2;
        D2m[bb][a] = v2;
      }
    double DiDb[P][P];
#pragma unroll
    for (int a = 0; a < P; ++a)
#pragma unroll
      for (int bb = 0; bb < P; ++bb) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < P; ++k) s += Bi[a][k] * Dm[k][bb];
        DiDb[a][bb] = s;
      }
    double ddetb = 0.0, tr_sq = 0.0, tr_d2 = 0.0;
#pragma unroll
    for (int a = 0; a < P; ++a) {
      ddetb += DiDb[a][a];
#pragma unroll
      for (int k = 0; k < P; ++k) {
        tr_sq += DiDb[a][k] * DiDb[k][a];
        tr_d2 += Bi[a][k] * D2m[k][a];
      }
    }
    const double d2detb = ddetb * ddetb - tr_sq + tr_d2;
    const double cr_t = (0.5 * ddetb * ddetb - 0.5 * d2detb) * 0.99;
    const double cr_t2 = -0.5 * ddetb * 0.99;
    const double tn2 = 1.0 / (theta * theta);
    const double tgt = tn2 * tg;
    const double ll = -2 * tn1 * (l1 - dg) + (l2 + tgt);
    r.dscore = ll + cr_t * (theta * theta) + (l1 - dg) * tn1 + cr_t2 * theta;
  }
  return r;
}

template <int P>
__global__ __launch_bounds__(64) void k_mle(Counts c, int64_t G, int64_t n,
                                            const double* __restrict__ X,
                                            const double* __restrict__ beta,
                                            const double* __restrict__ off, int max_iter,
                                            double* __restrict__ theta,
                                            int32_t* __restrict__ iters,
                                            int32_t* __restrict__ status) {
  const int64_t g = blockIdx.x;
  if (g >= G) return;
  const int lane = threadIdx.x;
  const void* row = count_row(c.ptr, c.dtype, g, c.ld);
  double b[P];
#pragma unroll
  for (int j = 0; j < P; ++j) b[j] = beta[g * P + j];
  double s1 = 0.0, spy = 0.0, my = 0.0, nz = 0.0;
  for (int64_t i = lane; i < n; i += kWave) {
    const double y = load_count(row, c.dtype, i);
    s1 += y;
    spy += (y - 1) * y;
    my = fmax(my, y);
    nz += (y == 0.0) ? 1.0 : 0.0;
  }
  s1 = wave_sum(s1);
  spy = wave_sum(spy);
  my = wave_max(my);
  nz = wave_sum(nz);
  double th = 0.0;
  int it = 0;
  int st = 0;
  if (nz == (double)n) {
    st = DEVIL_GENE_MLE_ZERO;  // all(y == 0) -> 0 (R/overdispersion.R:28-30)
  } else {
    const double ybar = s1 / (double)n;
    double v = 0.0;
    for (int64_t i = lane; i < n; i += kWave) {
      const double d = load_count(row, c.dtype, i) - ybar;
      v += d * d;
    }
    v = wave_sum(v) / (double)(n - 1);
    double start = (v - ybar) / (ybar * ybar);
    if (isnan(start) || start <= 0) start = 0.5;
    const MlePass fl = mle_pass<P>(row, c.dtype, n, X, b, off, log(1e-8), s1, spy, my, false);
    it = 1;
    if (fl.score < 0) {
      st = DEVIL_GENE_MLE_ZERO;
    } else {
      double lo = log(1e-8), hi = log(1e16);
      double x = log(start);
      x = fmin(fmax(x, lo), hi);
      for (int k = 0; k < max_iter; ++k) {
        const MlePass r = mle_pass<P>(row, c.dtype, n, X, b, off, x, s1, spy, my, true);
        ++it;
        const double s = r.score, d = r.dscore;
        if (!isfinite(s) || !isfinite(d)) break;
        if (s == 0.0) break;
        if (s > 0) lo = fmax(lo, x);
        else hi = fmin(hi, x);
        double xn = (d < 0) ? x - s / d : (s > 0 ? x + 1.0 : x - 1.0);
        if (!(xn > lo && xn < hi)) xn = 0.5 * (lo + hi);
        const double step = xn - x;
        x = xn;
        if (fabs(step) < 1e-9) break;
      }
      th = exp(x);
    }
  }
  if (lane == 0) {
    theta[g] = th;
    iters[g] = it;
    if (status) status[g] |= st;
  }
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
#define DV_DISPATCH_P(p, KERNEL, GRID, BLOCK, STREAM, ...)                         \
  switch (p) {                                                                    \
    case 1: hipLaunchKernelGGL(KERNEL<1>, GRID, BLOCK, 0, STREAM, __VA_ARGS__); break; \
    case 2: hipLaunchKernelGGL(KERNEL<2>, GRID, BLOCK, 0, STREAM, __VA_ARGS__); break; \
    case 3: hipLaunchKernelGGL(KERNEL<3>, GRID, BLOCK, 0, STREAM, __VA_ARGS__); break; \
    case 4: hipLaunchKernelGGL(KERNEL<4>, GRID, BLOCK, 0, STREAM, __VA_ARGS__); break; \
    case 5: hipLaunchKernelGGL(KERNEL<5>, GRID, BLOCK, 0, STREAM, __VA_ARGS__); break; \
    case 6: hipLaunchKernelGGL(KERNEL<6>, GRID, BLOCK, 0, STREAM, __VA_ARGS__); break; \
    case 7: hipLaunchKernelGGL(KERNEL<7>, GRID, BLOCK, 0, STREAM, __VA_ARGS__); break; \
    case 8: hipLaunchKernelGGL(KERNEL<8>, GRID, BLOCK, 0, STREAM, __VA_ARGS__); break; \
    case 9: hipLaunchKernelGGL(KERNEL<9>, GRID, BLOCK, 0, STREAM, __VA_ARGS__); break; \
    case 10: hipLaunchKernelGGL(KERNEL<10>, GRID, BLOCK, 0, STREAM, __VA_ARGS__); break; \
    default: return hipErrorInvalidValue;                                         \
  }

static inline dim3 gene_grid(int64_t G) { return dim3((unsigned)G); }

hipError_t launch_init(hipStream_t s, Counts c, int64_t G, int64_t n, int p,
                       const InitParams& ip, double* alpha0, double* beta0) {
  if (G <= 0) return hipSuccess;
  DV_DISPATCH_P(p, k_init, gene_grid(G), dim3(64), s, c, G, n, ip, alpha0, beta0);
  return hipGetLastError();
}

hipError_t launch_k_from_alpha(hipStream_t s, const double* alpha0, int64_t G, double* k) {
  if (G <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_k_from_alpha, dim3((unsigned)((G + 255) / 256)), dim3(256), 0, s,
                     alpha0, G, k);
  return hipGetLastError();
}

hipError_t launch_irls_fp64(hipStream_t s, Counts c, int64_t G, int64_t n, int p,
                            const double* X, const double* off, const double* k,
                            double* beta, int max_iter, double eps, int32_t* iters,
                            int32_t* status) {
  if (G <= 0) return hipSuccess;
  DV_DISPATCH_P(p, k_irls, gene_grid(G), dim3(64), s, c, G, n, X, off, k, beta, max_iter,
                eps, iters, status);
  return hipGetLastError();
}

hipError_t launch_mom(hipStream_t s, Counts c, int64_t G, int64_t n, int p,
                      const double* X, const double* beta, const double* eo,
                      double* theta, int32_t* status) {
  if (G <= 0) return hipSuccess;
  DV_DISPATCH_P(p, k_mom, gene_grid(G), dim3(64), s, c, G, n, X, beta, eo, theta, status);
  return hipGetLastError();
}

hipError_t launch_bread_meat(hipStream_t s, Counts c, int64_t G, int64_t n, int p,
                             const double* X, const double* beta, const double* theta,
                             const double* eo, const int32_t* ends, int C, double* bread,
                             double* sandwich, int32_t* status) {
  if (G <= 0) return hipSuccess;
  DV_DISPATCH_P(p, k_bread_meat, gene_grid(G), dim3(64), s, c, G, n, X, beta, theta, eo,
                ends, C, bread, sandwich, status);
  return hipGetLastError();
}

hipError_t launch_mle(hipStream_t s, Counts c, int64_t G, int64_t n, int p,
                      const double* X, const double* beta, const double* off,
                      int max_iter, double* theta, int32_t* iters, int32_t* status) {
  if (G <= 0) return hipSuccess;
  DV_DISPATCH_P(p, k_mle, gene_grid(G), dim3(64), s, c, G, n, X, beta, off, max_iter,
                theta, iters, status);
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
