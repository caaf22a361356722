// Device helpers shared by the gfx950 kernels: wavefront reductions, count
// loads, register-resident p x p LU (partial pivoting) and the special
// functions the dispersion and test kernels need.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dv {

constexpr int kWave = 64;

// Butterfly sum over the 64 lanes; every lane ends with the same bits
// (IEEE addition is commutative), so the result is deterministic.
// This thread's wavefront index in its workgroup as a wave-uniform value, so
// that branches on a wave's tiles compile to scalar branches rather than
// exec-mask regions (blocks are whole wavefronts of 64 threads).
__device__ __forceinline__ int wave_index() {
  return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v = fmax(v, __shfl_xor(v, o, kWave));
  return v;
}

// Count element i of a gene row stored as int32 (0), float (1) or double (2).
__device__ __forceinline__ double load_count(const void* row, int dtype, int64_t i) {
  if (dtype == 0) return (double)static_cast<const int32_t*>(row)[i];
  if (dtype == 1) return (double)static_cast<const float*>(row)[i];
  return static_cast<const double*>(row)[i];
}

__device__ __forceinline__ const void* count_row(const void* base, int dtype, int64_t g,
                                                 int64_t ld) {
  const size_t es = dtype == 0 ? 4 : dtype == 1 ? 4 : 8;
  return static_cast<const char*>(base) + (size_t)g * (size_t)ld * es;
}

// Packed upper triangle index of (a, b), a <= b, for a P x P symmetric matrix.
template <int P>
__host__ __device__ constexpr int tri(int a, int b) {
  return a * P - a * (a - 1) / 2 + (b - a);
}

// LU factorisation with partial pivoting in registers.  The pivot is the
// first entry of largest magnitude (LAPACK idamax / Eigen PartialPivLU);
// an exactly zero pivot column is left unscaled and reported, as Eigen does.
template <int P>
__device__ __forceinline__ int lu_factor(double (&A)[P][P], int (&perm)[P]) {
  int sing = 0;
#pragma unroll
  for (int k = 0; k < P; ++k) {
    int r = k;
    double best = fabs(A[k][k]);
#pragma unroll
    for (int i = k + 1; i < P; ++i) {
      const double v = fabs(A[i][k]);
      if (v > best) {
        best = v;
        r = i;
      }
    }
    perm[k] = r;
#pragma unroll
    for (int i = k + 1; i < P; ++i) {
      if (i == r) {
#pragma unroll
        for (int j = 0; j < P; ++j) {
          const double t = A[k][j];
          A[k][j] = A[i][j];
          A[i][j] = t;
        }
      }
    }
    if (best == 0.0) {
      sing = 1;
    } else {
      const double piv = A[k][k];
#pragma unroll
      for (int i = k + 1; i < P; ++i) A[i][k] = A[i][k] / piv;
    }
#pragma unroll
    for (int i = k + 1; i < P; ++i)
#pragma unroll
      for (int j = k + 1; j < P; ++j) A[i][j] -= A[i][k] * A[k][j];
  }
  return sing;
}

template <int P>
__device__ __forceinline__ void lu_solve(const double (&LU)[P][P], const int (&perm)[P],
                                         double (&v)[P]) {
#pragma unroll
  for (int k = 0; k < P; ++k) {
#pragma unroll
    for (int i = k + 1; i < P; ++i) {
      if (perm[k] == i) {
        const double t = v[k];
        v[k] = v[i];
        v[i] = t;
      }
    }
  }
#pragma unroll
  for (int i = 0; i < P; ++i)
#pragma unroll
    for (int k = 0; k < i; ++k) v[i] -= LU[i][k] * v[k];
#pragma unroll
  for (int i = P - 1; i >= 0; --i) {
#pragma unroll
    for (int k = i + 1; k < P; ++k) v[i] -= LU[i][k] * v[k];
    v[i] = v[i] / LU[i][i];
  }
}

// Inverse through the LU factors (Eigen's PartialPivLU::inverse).
template <int P>
__device__ __forceinline__ int lu_inverse(double (&A)[P][P], double (&Ainv)[P][P]) {
  int perm[P];
  const int sing = lu_factor<P>(A, perm);
#pragma unroll
  for (int c = 0; c < P; ++c) {
    double v[P];
#pragma unroll
    for (int i = 0; i < P; ++i) v[i] = (i == c) ? 1.0 : 0.0;
    lu_solve<P>(A, perm, v);
#pragma unroll
    for (int i = 0; i < P; ++i) Ainv[i][c] = v[i];
  }
  return sing;
}

// Inverse of a symmetric positive definite matrix by Cholesky (armadillo
// inv_sympd).  Returns 1 and NaNs when a pivot is not positive.
template <int P>
__device__ __forceinline__ int chol_inverse(const double (&A)[P][P], double (&Ainv)[P][P]) {
  double L[P][P];
  int bad = 0;
#pragma unroll
  for (int j = 0; j < P; ++j) {
#pragma unroll
    for (int i = 0; i < P; ++i) L[i][j] = 0.0;
  }
#pragma unroll
  for (int j = 0; j < P; ++j) {
    double s = A[j][j];
#pragma unroll
    for (int k = 0; k < j; ++k) s -= L[j][k] * L[j][k];
    if (!(s > 0.0)) bad = 1;
    const double d = sqrt(s);
    L[j][j] = d;
#pragma unroll
    for (int i = j + 1; i < P; ++i) {
      double t = A[i][j];
#pragma unroll
      for (int k = 0; k < j; ++k) t -= L[i][k] * L[j][k];
      L[i][j] = t / d;
    }
  }
#pragma unroll
  for (int c = 0; c < P; ++c) {
    double v[P];
#pragma unroll
    for (int i = 0; i < P; ++i) v[i] = (i == c) ? 1.0 : 0.0;
#pragma unroll
    for (int i = 0; i < P; ++i) {
#pragma unroll
      for (int k = 0; k < i; ++k) v[i] -= L[i][k] * v[k];
      v[i] = v[i] / L[i][i];
    }
#pragma unroll
    for (int i = P - 1; i >= 0; --i) {
#pragma unroll
      for (int k = i + 1; k < P; ++k) v[i] -= L[k][i] * v[k];
      v[i] = v[i] / L[i][i];
    }
#pragma unroll
    for (int i = 0; i < P; ++i) Ainv[i][c] = bad ? __builtin_nan("") : v[i];
  }
  return bad;
}

// ---------------------------------------------------------------------------
// special functions (x > 0)
// ---------------------------------------------------------------------------

// digamma: shift to x >= 10, then ln x - 1/(2x) - sum_k B_2k / (2k x^2k)
// (Abramowitz & Stegun 6.3.18), seven Bernoulli terms.
__device__ inline double digamma(double x) {
  if (!(x > 0.0)) return __builtin_nan("");
  double r = 0.0;
  while (x < 10.0) {
    r -= 1.0 / x;
    x += 1.0;
  }
  const double f = 1.0 / (x * x);
  const double t = f * (1.0 / 12 - f * (1.0 / 120 - f * (1.0 / 252 - f * (1.0 / 240 -
                   f * (1.0 / 132 - f * (691.0 / 32760 - f * (1.0 / 12)))))));
  return r + log(x) - 0.5 / x - t;
}

// trigamma: shift to x >= 10, then 1/x + 1/(2x^2) + sum_k B_2k / x^(2k+1)
// (A&S 6.4.12).
__device__ inline double trigamma(double x) {
  if (!(x > 0.0)) return __builtin_nan("");
  double r = 0.0;
  while (x < 10.0) {
    r += 1.0 / (x * x);
    x += 1.0;
  }
  const double t = 1.0 / x, f = t * t;
  return r + t + 0.5 * f +
         t * f * (1.0 / 6 - f * (1.0 / 30 - f * (1.0 / 42 - f * (1.0 / 30 -
         f * (5.0 / 66 - f * (691.0 / 2730 - f * (7.0 / 6)))))));
}

// psi(a + y) - psi(a) for a > 0, y >= 0, without the cancellation of the
// direct difference: a finite sum for small integer y, otherwise a shift of a
// to >= 10 with exact terms y/(a(a+y)) and the term-by-term difference of
// the asymptotic series,
//   log1p(y/a) + y/(2a(a+y)) + sum_k c_k a^-2k (1 - (1+y/a)^-2k).
__device__ inline double dpsi_diff(double a, double y) {
  if (y == 0.0) return 0.0;
  if (y <= 16.0 && y == floor(y)) {
    double s = 0.0;
    for (int j = 0; j < (int)y; ++j) s += 1.0 / (a + j);
    return s;
  }
  double s = 0.0;
  while (a < 10.0) {
    s += y / (a * (a + y));
    a += 1.0;
  }
  const double L = log1p(y / a);
  const double ia2 = 1.0 / (a * a);
  const double q = exp(-2.0 * L);
  const double c[7] = {1.0 / 12, -1.0 / 120, 1.0 / 252, -1.0 / 240, 1.0 / 132,
                       -691.0 / 32760, 1.0 / 12};
  double corr = c[0] * ia2 * (-expm1(-2.0 * L));
  double ak = ia2, qk = q;
#pragma unroll
  for (int k = 1; k < 7; ++k) {
    ak *= ia2;
    qk *= q;
    corr += c[k] * ak * (1.0 - qk);
  }
  return s + L + y / (2.0 * a * (a + y)) + corr;
}

// psi1(a + y) - psi1(a), same construction with
// psi1(x) ~ 1/x + 1/(2x^2) + sum_k B_2k x^-(2k+1).
__device__ inline double dpsi1_diff(double a, double y) {
  if (y == 0.0) return 0.0;
  if (y <= 16.0 && y == floor(y)) {
    double s = 0.0;
    for (int j = 0; j < (int)y; ++j) s -= 1.0 / ((a + j) * (a + j));
    return s;
  }
  double s = 0.0;
  while (a < 10.0) {
    s -= y * (2.0 * a + y) / (a * a * (a + y) * (a + y));
    a += 1.0;
  }
  const double L = log1p(y / a);
  const double ia = 1.0 / a;
  const double q = exp(-2.0 * L);
  const double B[7] = {1.0 / 6, -1.0 / 30, 1.0 / 42, -1.0 / 30, 5.0 / 66, -691.0 / 2730,
                       7.0 / 6};
  double d = -y / (a * (a + y)) - 0.5 * y * (2.0 * a + y) / (a * a * (a + y) * (a + y));
  double am = ia * ia * ia, qm = q * exp(-L);
  d -= B[0] * am * (-expm1(-3.0 * L));
#pragma unroll
  for (int k = 1; k < 7; ++k) {
    am *= ia * ia;
    qm *= q;
    d -= B[k] * am * (1.0 - qm);
  }
  return s + d;
}

// f(x) = log(1 + x) - x/(1 + x) for x >= 0; below 1e-2 its alternating series
// sum_{k>=2} (-1)^k (k-1)/k x^k (Horner, 11 terms).
__device__ inline double nb_f(double x) {
  if (x < 1e-2) {
    double s = 11.0 / 12;
#pragma unroll
    for (int k = 11; k >= 2; --k) s = (double)(k - 1) / k - x * s;
    return x * x * s;
  }
  return log1p(x) - x / (1.0 + x);
}

// Stirling remainder of lgamma for x >= 10.
__device__ inline double lgamma_corr(double x) {
  const double f = 1.0 / (x * x);
  return (1.0 / 12 - f * (1.0 / 360 - f * (1.0 / 1260 - f * (1.0 / 1680 -
          f * (1.0 / 1188 - f * (691.0 / 360360 - f * (1.0 / 156))))))) / x;
}

// log B(a, b) with the large-argument forms that avoid cancellation.
__device__ inline double lbeta(double a, double b) {
  const double p = fmin(a, b), q = fmax(a, b);
  const double half_log_2pi = 0.91893853320467274178;
  if (p >= 10.0) {
    const double corr = lgamma_corr(p) + lgamma_corr(q) - lgamma_corr(p + q);
    return -0.5 * log(q) + half_log_2pi + corr + (p - 0.5) * log(p / (p + q)) +
           q * log1p(-p / (p + q));
  } else if (q >= 10.0) {
    const double corr = lgamma_corr(q) - lgamma_corr(p + q);
    return lgamma(p) + corr + p - p * log(p + q) + (q - 0.5) * log1p(-p / (p + q));
  }
  return lgamma(p) + lgamma(q) - lgamma(p + q);
}

// Continued fraction of the regularised incomplete beta (modified Lentz).
__device__ inline double betacf(double a, double b, double x) {
  const double tiny = 1e-300, epsr = 1e-16;
  const double qab = a + b, qap = a + 1.0, qam = a - 1.0;
  double c = 1.0, d = 1.0 - qab * x / qap;
  if (fabs(d) < tiny) d = tiny;
  d = 1.0 / d;
  double h = d;
  for (int m = 1; m <= 100000; ++m) {
    const double m2 = 2.0 * m;
    double aa = m * (b - m) * x / ((qam + m2) * (a + m2));
    d = 1.0 + aa * d;
    if (fabs(d) < tiny) d = tiny;
    c = 1.0 + aa / c;
    if (fabs(c) < tiny) c = tiny;
    d = 1.0 / d;
    h *= d * c;
    aa = -(a + m) * (qab + m) * x / ((a + m2) * (qap + m2));
    d = 1.0 + aa * d;
    if (fabs(d) < tiny) d = tiny;
    c = 1.0 + aa / c;
    if (fabs(c) < tiny) c = tiny;
    d = 1.0 / d;
    const double del = d * c;
    h *= del;
    if (fabs(del - 1.0) < epsr) break;
  }
  return h;
}

// log of 2*P(T_df > |t|) = log I_{df/(df+t^2)}(df/2, 1/2); exact in log space
// where the probability underflows.
__device__ inline double log_pt2_upper(double t, double df) {
  if (isnan(t) || isnan(df) || !(df > 0.0)) return __builtin_nan("");
  t = fabs(t);
  if (isinf(t)) return -__builtin_huge_val();
  if (t == 0.0) return 0.0;
  const double lt = log(t), ldf = log(df);
  const double l_sum = (t * t > df) ? 2.0 * lt + log1p(df / t / t) : ldf + log1p(t / df * t);
  const double lx = ldf - l_sum, lomx = 2.0 * lt - l_sum;
  const double x = exp(lx), omx = exp(lomx);
  const double a = 0.5 * df, b = 0.5;
  const double lfront = a * lx + b * lomx - lbeta(a, b);
  if (x < (a + 1.0) / (a + b + 2.0)) return lfront + log(betacf(a, b, x)) - log(a);
  const double lc = lfront + log(betacf(b, a, omx)) - log(b);
  return log1p(-exp(lc));
}

// ---------------------------------------------------------------------------
// MFMA tile helpers
// ---------------------------------------------------------------------------
typedef double d4 __attribute__((ext_vector_type(4)));
typedef int i4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ d4 mfma16x16x4(double a, double b, d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// exp(x) = 2^n e^r with n = rint(x / ln 2) and r = x - n ln 2 (Cody-Waite,
// two constants), e^r by its degree-12 Taylor polynomial (|r| <= 0.347,
// truncation < 2.5e-16 relative), scaled with v_ldexp_f64 which saturates to
// inf / 0 outside the double range.
__device__ __forceinline__ double exp_f64(double x) {
  const double n = rint(x * 1.4426950408889634);
  double r = fma(-n, 6.93147180369123816490e-01, x);
  r = fma(-n, 1.90821492927058770002e-10, r);
  double p = 1.0 / 479001600.0;
  p = fma(p, r, 1.0 / 39916800.0);
  p = fma(p, r, 1.0 / 3628800.0);
  p = fma(p, r, 1.0 / 362880.0);
  p = fma(p, r, 1.0 / 40320.0);
  p = fma(p, r, 1.0 / 5040.0);
  p = fma(p, r, 1.0 / 720.0);
  p = fma(p, r, 1.0 / 120.0);
  p = fma(p, r, 1.0 / 24.0);
  p = fma(p, r, 1.0 / 6.0);
  p = fma(p, r, 0.5);
  p = fma(p, r, 1.0);
  p = fma(p, r, 1.0);
  return ldexp(p, (int)n);
}

// 1/d from v_rcp_f64 refined by two Newton steps (about 1 ulp).
__device__ __forceinline__ double rcp_f64(double d) {
  double r = __builtin_amdgcn_rcp(d);
  double e = fma(-d, r, 1.0);
  r = fma(r, e, r);
  e = fma(-d, r, 1.0);
  return fma(r, e, r);
}

// 1/d from v_rcp_f64 (relative error below 4.6e-8 on gfx950, measured) and
// one Newton step: relative error below 2.2e-15.
__device__ __forceinline__ double rcp1_f64(double d) {
  const double r = __builtin_amdgcn_rcp(d);
  return fma(r, fma(-d, r, 1.0), r);
}

// log(1 + x) for x >= 0 (and NaN / inf passed through).  u = 1 + x is split
// as 2^e m with m in [sqrt(1/2), sqrt(2)); log m = 2 atanh(s), s = (m-1)/(m+1),
// |s| <= 0.1716, by its odd series to s^21 (truncation below 1e-17); the
// rounding of 1 + x is restored by (x - (u - 1)) / u.
__device__ __forceinline__ double log1p_f64(double x) {
  const double u = 1.0 + x;
  const double c = (x - (u - 1.0)) * rcp1_f64(u);
  int e = __builtin_amdgcn_frexp_exp(u);
  double m = __builtin_amdgcn_frexp_mant(u);  // [0.5, 1)
  if (m < 0.70710678118654752440) {
    m *= 2.0;
    e -= 1;
  }
  const double s = (m - 1.0) * rcp_f64(m + 1.0);
  const double z = s * s;
  double p = 1.0 / 21;
  p = fma(p, z, 1.0 / 19);
  p = fma(p, z, 1.0 / 17);
  p = fma(p, z, 1.0 / 15);
  p = fma(p, z, 1.0 / 13);
  p = fma(p, z, 1.0 / 11);
  p = fma(p, z, 1.0 / 9);
  p = fma(p, z, 1.0 / 7);
  p = fma(p, z, 1.0 / 5);
  p = fma(p, z, 1.0 / 3);
  const double lm = fma(2.0 * s * z, p, 2.0 * s);
  const double ed = (double)e;
  const double r = fma(ed, 6.93147180369123816490e-01, fma(ed, 1.90821492927058770002e-10, lm));
  return isfinite(u) ? r + c : u;
}

// Table-driven exponential (P. T. P. Tang, "Table-driven implementation of
// the exponential function in IEEE floating-point arithmetic", ACM TOMS
// 15(2), 1989): x = (64 m + j) ln2/64 + r with |r| <= ln2/128, exp x =
// 2^m 2^(j/64) e^r, e^r - 1 by its Taylor series to r^6 (truncation below
// 3e-20); 2^(j/64) from the table, scaled with v_ldexp_f64 (inf / 0 outside
// the double range).  The index j = n mod 64 stays in [0, 64) for any x.
// Entries are 2^(j/64) correctly rounded (computed to 60 digits).
static __device__ constexpr double kExp2Tab[64] = {
    0x1.0000000000000p+0, 0x1.02c9a3e778061p+0, 0x1.059b0d3158574p+0,
    0x1.0874518759bc8p+0, 0x1.0b5586cf9890fp+0, 0x1.0e3ec32d3d1a2p+0,
    0x1.11301d0125b51p+0, 0x1.1429aaea92de0p+0, 0x1.172b83c7d517bp+0,
    0x1.1a35beb6fcb75p+0, 0x1.1d4873168b9aap+0, 0x1.2063b88628cd6p+0,
    0x1.2387a6e756238p+0, 0x1.26b4565e27cddp+0, 0x1.29e9df51fdee1p+0,
    0x1.2d285a6e4030bp+0, 0x1.306fe0a31b715p+0, 0x1.33c08b26416ffp+0,
    0x1.371a7373aa9cbp+0, 0x1.3a7db34e59ff7p+0, 0x1.3dea64c123422p+0,
    0x1.4160a21f72e2ap+0, 0x1.44e086061892dp+0, 0x1.486a2b5c13cd0p+0,
    0x1.4bfdad5362a27p+0, 0x1.4f9b2769d2ca7p+0, 0x1.5342b569d4f82p+0,
    0x1.56f4736b527dap+0, 0x1.5ab07dd485429p+0, 0x1.5e76f15ad2148p+0,
    0x1.6247eb03a5585p+0, 0x1.6623882552225p+0, 0x1.6a09e667f3bcdp+0,
    0x1.6dfb23c651a2fp+0, 0x1.71f75e8ec5f74p+0, 0x1.75feb564267c9p+0,
    0x1.7a11473eb0187p+0, 0x1.7e2f336cf4e62p+0, 0x1.82589994cce13p+0,
    0x1.868d99b4492edp+0, 0x1.8ace5422aa0dbp+0, 0x1.8f1ae99157736p+0,
    0x1.93737b0cdc5e5p+0, 0x1.97d829fde4e50p+0, 0x1.9c49182a3f090p+0,
    0x1.a0c667b5de565p+0, 0x1.a5503b23e255dp+0, 0x1.a9e6b5579fdbfp+0,
    0x1.ae89f995ad3adp+0, 0x1.b33a2b84f15fbp+0, 0x1.b7f76f2fb5e47p+0,
    0x1.bcc1e904bc1d2p+0, 0x1.c199bdd85529cp+0, 0x1.c67f12e57d14bp+0,
    0x1.cb720dcef9069p+0, 0x1.d072d4a07897cp+0, 0x1.d5818dcfba487p+0,
    0x1.da9e603db3285p+0, 0x1.dfc97337b9b5fp+0, 0x1.e502ee78b3ff6p+0,
    0x1.ea4afa2a490dap+0, 0x1.efa1bee615a27p+0, 0x1.f50765b6e4540p+0,
    0x1.fa7c1819e90d8p+0,
};

// threads [0, nthreads) of the workgroup copy the table to tab (LDS); the
// caller synchronises
__device__ __forceinline__ void exp_tab_fill(double* tab, int tid, int nthreads) {
  for (int j = tid; j < 64; j += nthreads) tab[j] = kExp2Tab[j];
}

__device__ __forceinline__ double exp_tab(double x, const double* tab) {
  const double nd = rint(x * 92.332482616893656877);  // 64 / ln 2
  double r = fma(-nd, 6.93147180369123816490e-01 / 64, x);
  r = fma(-nd, 1.90821492927058770002e-10 / 64, r);
  const int n = (int)nd;
  double q = 1.0 / 720;
  q = fma(q, r, 1.0 / 120);
  q = fma(q, r, 1.0 / 24);
  q = fma(q, r, 1.0 / 6);
  q = fma(q, r, 0.5);
  const double p = fma(r * r, q, r);  // e^r - 1
  const double t = tab[n & 63];
  return __builtin_amdgcn_ldexp(fma(t, p, t), n >> 6);
}

// Table-driven logarithm (P. T. P. Tang, "Table-driven implementation of the
// logarithm function in IEEE floating-point arithmetic", ACM TOMS 16(4),
// 1990): u = 2^e m, m in [1, 2), F = 1 + j/128 the nearest breakpoint,
// log u = e log 2 + log F + log1p(r) with r = (m - F) / F, |r| <= 2^-8, and
// log1p(r) by its Taylor series to r^7 (truncation below 2^-67).  Entry j of
// the table holds 1/F and log F; a workgroup fills it once with log_tab_fill.
struct LogTabEntry {
  double inv;  // 1 / (1 + j/128), correctly rounded
  double ln;   // log(1 + j/128) (log1p_f64)
};
constexpr int kLogTab = 129;

// threads [0, nthreads) of the workgroup fill tab; the caller synchronises
__device__ __forceinline__ void log_tab_fill(LogTabEntry* tab, int tid, int nthreads) {
  for (int j = tid; j < kLogTab; j += nthreads) {
    const double f = (double)j * (1.0 / 128);
    tab[j].inv = 1.0 / (1.0 + f);
    tab[j].ln = log1p_f64(f);
  }
}

// log(u) + c / u for u > 0 finite, where c is a correction much smaller than
// u (the rounding error of u when u = 1 + x); other u as log: -inf at 0,
// NaN below, u itself for inf / NaN.
__device__ __forceinline__ double log_tab(double u, double c, const LogTabEntry* tab) {
  int e = __builtin_amdgcn_frexp_exp(u) - 1;
  const double m = 2.0 * __builtin_amdgcn_frexp_mant(u);  // [1, 2)
  const double jd = __builtin_rint(fma(m, 128.0, -128.0));
  // jd is in [0, 128] for finite u > 0; the clamp keeps other u in the table
  const LogTabEntry t = tab[min(max((int)jd, 0), kLogTab - 1)];
  const double f = m - fma(jd, 1.0 / 128, 1.0);  // exact
  const double r = f * t.inv;
  double q = 1.0 / 7;
  q = fma(q, r, -1.0 / 6);
  q = fma(q, r, 1.0 / 5);
  q = fma(q, r, -1.0 / 4);
  q = fma(q, r, 1.0 / 3);
  q = fma(q, r, -0.5);
  const double lp = fma(r * r, q, r);
  const double ed = (double)e;
  const double hi = fma(ed, 6.93147180369123816490e-01, t.ln);
  const double lo = fma(ed, 1.90821492927058770002e-10, __builtin_amdgcn_ldexp(c * t.inv, -e));
  const double res = hi + (lp + lo);
  if (u > 0.0 && u < __builtin_inf()) return res;
  return u == 0.0 ? -__builtin_inf() : (u > 0.0 || u != u ? u : __builtin_nan(""));
}

// log(1 + x) through log_tab: u = 1 + x rounded, c = x - (u - 1) its rounding
// error (exact for 0 <= x < 2^53).
__device__ __forceinline__ double log1p_tab(double x, const LogTabEntry* tab) {
  const double u = 1.0 + x;
  return log_tab(u, x - (u - 1.0), tab);
}

}  // namespace dv
