// The "new" / "I" overdispersion of fit_overdispersion_cppp
// (src/new_overdispersion.cpp:195-481) for one gene, in reverse
// communication: each call of newod_advance brings F_k at the point the
// driver asked for (NewodState::pending) and returns either the next point
// or the result.  One round of the MLE-style kernels evaluates every active
// gene's pending point (the per-cell sums of F_k); this driver keeps the
// reference's control flow: the moment start, one Newton step with a
// forward-difference derivative, the fixed-point step, Aitken acceleration
// with backtracking, the secant and bracket fallbacks and Brent's method,
// with the reference's constants.
//
// F_k (:270-291) from the per-cell sums:
//   C1 = sum (X beta) - (sum digamma(k + y) - sum log1p(k w)) + sum (k + y) w / (1 + k w),
//   w = exp(-(X beta + off)), then exp(logH_p - logH_m) by H_log_gh_pmhalf
//   (:141-158): Laplace modes of h(x) = p log x + n (x log x - lgamma x) - C1 x
//   for p = +-1/2 and a 16-point Gauss-Hermite rule around each (:24-123).
// Where the reference stops with an error (C1 <= n, a mode whose curvature is
// not positive, a quadrature that underflows, f not finite at the start) the
// gene ends with err != 0.
#pragma once
#include <hip/hip_runtime.h>

#include "device_common.hpp"

namespace dv {

// Gauss-Hermite rule of order 16 (weight exp(-x^2)), the nodes and weights
// of the Jacobi-matrix eigenproblem (Golub-Welsch) the reference solves
__device__ __constant__ const double kGh16X[16] = {
    -4.688738939305819, -3.869447904860123, -3.176999161979956, -2.5462021578474814,
    -1.9517879909162539, -1.3802585391988809, -0.8229514491446559, -0.27348104613815244,
    0.27348104613815244, 0.8229514491446559, 1.3802585391988809, 1.9517879909162539,
    2.5462021578474814, 3.176999161979956, 3.869447904860123, 4.688738939305819};
__device__ __constant__ const double kGh16W[16] = {
    2.6548074740111673e-10, 2.3209808448652032e-07, 2.7118600925378892e-05,
    0.0009322840086241807, 0.012880311535509989, 0.08381004139898583, 0.2806474585285337,
    0.5079294790166137, 0.5079294790166137, 0.2806474585285337, 0.08381004139898583,
    0.012880311535509989, 0.0009322840086241807, 2.7118600925378892e-05,
    2.3209808448652032e-07, 2.6548074740111673e-10};

__device__ inline double nod_g(double s, double t, double x) {
  return s * (x * log(x) - lgamma(x)) - t * x;
}

// find_mode (:70-93); returns false where the reference stops
__device__ inline bool nod_find_mode(double p, double s, double t, double tol, double x0,
                                     double& mu0, double& sigma0, double& h_mu0) {
  double x = fmax(1e-9, x0);
  for (int i = 0; i < 16; ++i) {
    const double g1v = p / x + (s * (log(x) + 1.0 - digamma(x)) - t);
    if (fabs(g1v) < tol) break;
    const double g2v = -p / (x * x) + s * (1.0 / x - trigamma(x));
    double xn = x - g1v / g2v;
    if (!(xn > 0.0) || !isfinite(xn)) xn = 0.5 * x;
    if (fabs(xn - x) < tol * (1.0 + fabs(x))) {
      x = xn;
      break;
    }
    x = xn;
  }
  const double curv = -(-p / (x * x) + s * (1.0 / x - trigamma(x)));
  if (!(curv > 0.0) || !isfinite(curv)) return false;
  mu0 = fmax(1e-12, x);
  sigma0 = sqrt(1.0 / curv);
  h_mu0 = p * log(mu0) + nod_g(s, t, mu0);
  return true;
}

// F_k from C1 (H_log_gh_pmhalf with s = n, t = C1; NaN where the reference
// stops), by one wavefront: C1 is the same in every lane and every lane gets
// F.  Both modes are found on every lane; then log_integral_at (:95-123)
// takes one Gauss-Hermite node per lane (lanes 0-15 the p = +1/2 integral,
// 16-31 the p = -1/2 one), with the two-pass log-sum-exp reduced within each
// 16-lane group.
__device__ inline double nod_F_from_C1_wave(double n, double C1, double tol, int& err) {
  if (!(C1 > n)) {
    err = 1;
    return __builtin_nan("");
  }
  double m1, s1, h1, m2, s2, h2;
  const double x0 = fmax(1e-3, (0.5 + n) / (C1 - n));
  if (!nod_find_mode(0.5, n, C1, tol, x0, m1, s1, h1) ||
      !nod_find_mode(-0.5, n, C1, tol, m1, m2, s2, h2)) {
    err = 2;
    return __builtin_nan("");
  }
  const int lane = threadIdx.x & 63;
  const int node = lane & 15;
  const bool second = (lane & 16) != 0;
  const double p = second ? -0.5 : 0.5;
  const double mu0 = second ? m2 : m1, sigma0 = second ? s2 : s1, h_mu0 = second ? h2 : h1;
  const double scale = sigma0 * sqrt(2.0);
  const double u = kGh16X[node], xx = mu0 + scale * u;
  double lt = -INFINITY;
  if (lane < 32 && xx > 0.0 && isfinite(xx))
    lt = log(kGh16W[node]) + ((p * log(xx) + nod_g(n, C1, xx)) - h_mu0 + u * u);
  double mx = lt;
#pragma unroll
  for (int o = 8; o >= 1; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
  double acc = lt > -INFINITY ? exp(lt - mx) : 0.0;
#pragma unroll
  for (int o = 8; o >= 1; o >>= 1) acc += __shfl_xor(acc, o, 64);
  const double li = h_mu0 + log(scale) + (mx + log(acc));
  const double acc1 = __shfl(acc, 0, 64), acc2 = __shfl(acc, 16, 64);
  const double lp = __shfl(li, 0, 64), lm = __shfl(li, 16, 64);
  if (!(acc1 > 0.0) || !(acc2 > 0.0)) {
    err = 2;
    return __builtin_nan("");
  }
  return exp(lp - lm);
}

struct NewodState {
  double k, k0, k1, k2, f0, f1, f2, fk, k_acc, k_prop, f_prop, k_newton, h, k_sec;
  double ba, bb, bc, bfa, bfb, bfc, bd, be, f_center, lo, hi, flo, fhi, blo, bhi;
  double pending;  // k of the requested evaluation
  int32_t pc, ret, err, converged, it, bad_accel, bt, expand, bit, evals, pad0, pad1;
};

enum { kNodDone = -1 };

// Call with F = F_k at st.pending (ignored on the first call, st.pc = 0).
// Returns 0 with the next point in st.pending, 1 when done: kappa = 1 / st.k
// (NaN when st.err != 0), kappa_iters = st.it + 1, st.converged.
__device__ inline int newod_advance(NewodState& st, double F, int max_iter, double eps) {
  const double k_min = 1e-4, k_max = 1e4, jump_factor = 10.0;
  // f_of_k from F at st.pending
#define NOD_F_OF(F_) (log(F_) - log(st.pending))
#define NOD_EVAL(kk, LABEL)  \
  do {                       \
    st.pending = (kk);       \
    st.pc = LABEL;           \
    return 0;                \
    case LABEL:;             \
  } while (0)
  if (st.pc > 0) {
    st.evals += 1;
    st.fk = F;
    if (!isfinite(F) && st.err) {  // the evaluation hit a reference stop
      st.k = __builtin_nan("");
      st.pc = kNodDone;
      return 1;
    }
  }
  switch (st.pc) {
    case 0:
      st.evals = 0;
      st.it = 0;
      st.converged = 0;
      st.bad_accel = 0;
      st.k0 = fmax(k_min, fmin(k_max, st.k));
      NOD_EVAL(st.k0, 1);
      st.f0 = NOD_F_OF(st.fk);
      if (!isfinite(st.f0)) {
        st.err = 3;
        st.k = __builtin_nan("");
        st.pc = kNodDone;
        return 1;
      }
      // one Newton step with a forward-difference derivative
      st.h = fmax(1e-8, 1e-6 * st.k0);
      NOD_EVAL(st.k0 + st.h, 2);
      {
        const double fp0 = (NOD_F_OF(st.fk) - st.f0) / st.h;
        st.k_newton = -1.0;
        if (isfinite(fp0) && fabs(fp0) > 1e-10)
          st.k_newton = fmax(k_min, fmin(k_max, st.k0 - st.f0 / fp0));
      }
      if (st.k_newton > 0.0) {
        NOD_EVAL(st.k_newton, 3);
        {
          const double fn = NOD_F_OF(st.fk);
          if (isfinite(fn) && fabs(fn) < fabs(st.f0)) {
            st.k0 = st.k_newton;
            st.f0 = fn;
          }
        }
      }
      NOD_EVAL(st.k0, 4);  // k1 = F_k(k0)
      st.k1 = fmin(st.fk, jump_factor * fmax(st.k0, 1.0));
      st.k1 = fmax(k_min, fmin(k_max, st.k1));
      NOD_EVAL(st.k1, 5);
      st.f1 = NOD_F_OF(st.fk);
      if (st.f0 * st.f1 < 0) {
        st.ba = fmin(st.k0, st.k1);
        st.bb = fmax(st.k0, st.k1);
        st.bfa = st.f0;  // as the reference passes them (:404)
        st.bfb = st.f1;
        st.ret = 100;
        goto brent;
      }
      for (st.it = 0; st.it < max_iter; ++st.it) {
        NOD_EVAL(st.k1, 6);  // k2 = F_k(k1)
        {
          const double adaptive_jump = jump_factor / (1.0 + 0.1 * st.it);
          st.k2 = fmin(st.fk, adaptive_jump * fmax(st.k1, 1.0));
          st.k2 = fmax(k_min, fmin(k_max, st.k2));
        }
        NOD_EVAL(st.k2, 7);
        st.f2 = NOD_F_OF(st.fk);
        {
          const double adaptive_jump = jump_factor / (1.0 + 0.1 * st.it);
          const double d1 = st.k1 - st.k0, d2 = st.k2 - st.k1, denom = d2 - d1;
          double k_acc = (fabs(denom) > 1e-14) ? st.k0 - (d1 * d1) / denom : st.k2;
          if (!(k_acc > 0) || !isfinite(k_acc)) k_acc = st.k2;
          k_acc = fmin(k_acc, adaptive_jump * fmax(st.k2, 1.0));
          st.k_acc = fmax(k_min, fmin(k_max, k_acc));
        }
        st.k_prop = st.k_acc;
        NOD_EVAL(st.k_prop, 8);
        st.f_prop = NOD_F_OF(st.fk);
        for (st.bt = 0; (fabs(st.f_prop) >= fabs(st.f1) || !isfinite(st.f_prop)) && st.bt < 8;
             ++st.bt) {
          st.k_prop = 0.5 * (st.k_prop + st.k1);
          NOD_EVAL(st.k_prop, 9);
          st.f_prop = NOD_F_OF(st.fk);
        }
        if (!(fabs(st.f_prop) < fabs(st.f1))) {
          st.bad_accel++;
          if (st.bad_accel == 1 && fabs(st.f1 - st.f0) > 1e-14) {
            st.k_sec = st.k1 - st.f1 * (st.k1 - st.k0) / (st.f1 - st.f0);
            st.k_sec = fmax(k_min, fmin(k_max, st.k_sec));
            NOD_EVAL(st.k_sec, 10);
            {
              const double fs = NOD_F_OF(st.fk);
              if (isfinite(fs) && fabs(fs) < fabs(st.f1)) {
                st.k_prop = st.k_sec;
                st.f_prop = fs;
                st.bad_accel = 0;
              }
            }
          }
          if (st.bad_accel >= 1) {
            if (st.f1 * st.f2 < 0) {
              st.ba = fmin(st.k1, st.k2);
              st.bb = fmax(st.k1, st.k2);
              st.bfa = st.f1;
              st.bfb = st.f2;
              st.ret = 100;
              goto brent;
            }
          }
          if (st.bad_accel >= 1) {
            // find_bracket (:300-316) around k1
            st.blo = -1.0;
            st.bhi = -1.0;
            NOD_EVAL(st.k1, 11);
            st.f_center = NOD_F_OF(st.fk);
            for (st.expand = 1; st.expand <= 10; ++st.expand) {
              {
                const double factor = pow(2.0, (double)st.expand);
                st.lo = fmax(k_min, st.k1 / factor);
                st.hi = fmin(k_max, st.k1 * factor);
              }
              NOD_EVAL(st.lo, 12);
              st.flo = NOD_F_OF(st.fk);
              if (st.flo * st.f_center < 0) {
                st.blo = st.lo;
                st.bhi = st.k1;
                break;
              }
              NOD_EVAL(st.hi, 13);
              st.fhi = NOD_F_OF(st.fk);
              if (st.f_center * st.fhi < 0) {
                st.blo = st.k1;
                st.bhi = st.hi;
                break;
              }
              if (st.flo * st.fhi < 0) {
                st.blo = st.lo;
                st.bhi = st.hi;
                break;
              }
            }
            if (st.blo > 0) {
              NOD_EVAL(st.blo, 14);
              st.bfa = NOD_F_OF(st.fk);
              NOD_EVAL(st.bhi, 15);
              st.bfb = NOD_F_OF(st.fk);
              st.ba = st.blo;
              st.bb = st.bhi;
              st.ret = 100;
              goto brent;
            }
          }
          st.k_prop = st.k2;
          st.f_prop = st.f2;
        } else {
          st.bad_accel = 0;
        }
        if (fabs(st.k_prop - st.k1) < eps * (1.0 + st.k1)) {
          st.k = st.k_prop;
          st.converged = 1;
          st.pc = kNodDone;
          return 1;
        }
        st.k0 = st.k1;
        st.f0 = st.f1;
        st.k1 = st.k_prop;
        st.f1 = st.f_prop;
      }
      st.k = fmax(k_min, st.k1);
      st.pc = kNodDone;
      return 1;

    brent:
      // brent (:319-348) on f_of_k from (ba, fa), (bb, fb)
      st.bc = st.ba;
      st.bfc = st.bfa;
      st.bd = st.bb - st.ba;
      st.be = st.bd;
      for (st.bit = 0; st.bit < 60; ++st.bit) {
        if (fabs(st.bfc) < fabs(st.bfb)) {
          st.ba = st.bb;
          st.bb = st.bc;
          st.bc = st.ba;
          st.bfa = st.bfb;
          st.bfb = st.bfc;
          st.bfc = st.bfa;
        }
        {
          const double tol1 = 2.0 * 2.220446049250313e-16 * fabs(st.bb) + 0.5 * eps;
          const double xm = 0.5 * (st.bc - st.bb);
          if (fabs(xm) <= tol1 || st.bfb == 0.0) break;
          if (fabs(st.be) >= tol1 && fabs(st.bfa) > fabs(st.bfb)) {
            const double s = st.bfb / st.bfa;
            double p, q;
            if (st.ba == st.bc) {
              p = 2.0 * xm * s;
              q = 1.0 - s;
            } else {
              const double r = st.bfb / st.bfc, t = st.bfa / st.bfc;
              p = s * (2.0 * xm * t * (t - r) - (st.bb - st.ba) * (r - 1.0));
              q = (t - 1.0) * (r - 1.0) * (s - 1.0);
            }
            if (p > 0) q = -q;
            p = fabs(p);
            if (2.0 * p < fmin(3.0 * xm * q - fabs(tol1 * q), fabs(st.be * q))) {
              st.be = st.bd;
              st.bd = p / q;
            } else {
              st.bd = xm;
              st.be = st.bd;
            }
          } else {
            st.bd = xm;
            st.be = st.bd;
          }
          st.ba = st.bb;
          st.bfa = st.bfb;
          st.bb += (fabs(st.bd) > tol1 ? st.bd : (xm > 0 ? tol1 : -tol1));
        }
        NOD_EVAL(st.bb, 16);
        st.bfb = NOD_F_OF(st.fk);
        if ((st.bfb > 0 && st.bfc > 0) || (st.bfb < 0 && st.bfc < 0)) {
          st.bc = st.ba;
          st.bfc = st.bfa;
          st.be = st.bd = st.bb - st.ba;
        }
      }
      st.k = st.bb;
      st.converged = 1;
      st.pc = kNodDone;
      return 1;
    default:
      st.pc = kNodDone;
      return 1;
  }
#undef NOD_EVAL
#undef NOD_F_OF
}

}  // namespace dv
