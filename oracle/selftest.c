/*
 * selftest.c -- drives every oracle entry point on small synthetic inputs
 * (TEST INFRASTRUCTURE ONLY, like the oracle itself).  Built with
 * AddressSanitizer and UndefinedBehaviorSanitizer by `make -C oracle
 * sanitize`, so an out-of-bounds access, use after free, leak, signed
 * overflow or invalid shift anywhere on these paths aborts the run.  The
 * checks here are sanity checks (finite outputs, ranges); parity itself is
 * pinned by the Python tests.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "devil_oracle.h"

static uint64_t rng_state = 0x9e3779b97f4a7c15ull;

static double unif(void) { /* xorshift64*, (0, 1) */
  rng_state ^= rng_state >> 12;
  rng_state ^= rng_state << 25;
  rng_state ^= rng_state >> 27;
  return ((double)((rng_state * 2685821657736338717ull) >> 11) + 0.5) / 9007199254740992.0;
}

static double gamma_draw(double shape) { /* Marsaglia-Tsang, shape >= 1 via boost */
  if (shape < 1.0) return gamma_draw(shape + 1.0) * pow(unif(), 1.0 / shape);
  const double d = shape - 1.0 / 3.0, c = 1.0 / sqrt(9.0 * d);
  for (;;) {
    double x, v;
    do {
      const double u1 = unif(), u2 = unif();
      x = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
      v = 1.0 + c * x;
    } while (v <= 0.0);
    v = v * v * v;
    const double u = unif();
    if (log(u) < 0.5 * x * x + d - d * v + d * log(v)) return d * v;
  }
}

static double poisson_draw(double lam) {
  if (lam > 50.0) { /* normal approximation, rounded, for large means */
    const double u1 = unif(), u2 = unif();
    const double z = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    const double v = floor(lam + sqrt(lam) * z + 0.5);
    return v < 0.0 ? 0.0 : v;
  }
  const double L = exp(-lam);
  double k = 0.0, prod = unif();
  while (prod > L) {
    k += 1.0;
    prod *= unif();
  }
  return k;
}

static int fails = 0;
#define CHECK(cond, ...)                  \
  do {                                    \
    if (!(cond)) {                        \
      fprintf(stderr, "FAIL: " __VA_ARGS__); \
      fprintf(stderr, "\n");              \
      ++fails;                            \
    }                                     \
  } while (0)

static void fit_case(int64_t G, int64_t n, int p, int od, int init, int clustered, int driver) {
  double *Y = malloc(sizeof(double) * G * n), *X = malloc(sizeof(double) * n * p);
  double *off = malloc(sizeof(double) * n);
  for (int64_t i = 0; i < n; ++i) {
    X[i] = 1.0;
    for (int j = 1; j < p; ++j) X[i + j * n] = (j == 1) ? (double)(i & 1) : unif() - 0.5;
    off[i] = log(0.5 + unif());
  }
  for (int64_t g = 0; g < G; ++g) {
    const double mu0 = exp(2.0 * unif() - 0.5), theta = 0.1 + unif();
    for (int64_t i = 0; i < n; ++i) {
      const double mu = mu0 * exp(off[i]) * (1.0 + 0.5 * X[i + (p > 1 ? n : 0)]);
      Y[g * n + i] = (g == 0) ? 0.0 : poisson_draw(gamma_draw(1.0 / theta) * theta * mu);
    }
  }
  int32_t ends[4] = {(int32_t)(n / 4), (int32_t)(n / 2), (int32_t)(3 * n / 4), (int32_t)n};
  orc_fit_args a;
  memset(&a, 0, sizeof a);
  a.G = G;
  a.n = n;
  a.ld = n;
  a.p = p;
  a.Y = Y;
  a.X = X;
  a.offset = off;
  a.cluster_ends = clustered ? ends : NULL;
  a.n_clusters = clustered ? 4 : 0;
  a.init_mode = init;
  a.od_mode = od;
  a.max_iter = 200;
  a.tolerance = 1e-3;
  a.n_threads = 2;
  a.mle_driver = driver;
  orc_fit_out o;
  o.beta = malloc(sizeof(double) * G * p);
  o.theta = malloc(sizeof(double) * G);
  o.beta_iters = malloc(sizeof(int32_t) * G);
  o.theta_iters = malloc(sizeof(int32_t) * G);
  o.bread = malloc(sizeof(double) * G * p * p);
  o.sandwich = clustered ? malloc(sizeof(double) * G * p * p) : NULL;
  o.alpha0 = malloc(sizeof(double) * G);
  o.beta0 = malloc(sizeof(double) * G * p);
  const int rc = orc_cpu_fit(&a, &o);
  CHECK(rc == 0, "cpu_fit rc %d (od %d init %d)", rc, od, init);
  for (int64_t g = 1; g < G; ++g) {
    CHECK(isfinite(o.beta[g]), "beta not finite (gene %ld od %d)", (long)g, od);
    CHECK(o.beta_iters[g] >= 1 && o.beta_iters[g] <= 200, "beta iters %d", o.beta_iters[g]);
    if (od != ORC_OD_NEW) CHECK(o.theta[g] >= 0.0, "theta %g (od %d)", o.theta[g], od);
  }
  /* the Wald test and every p-value adjustment on the fit */
  double *c = calloc((size_t)p, sizeof(double)), *pv = malloc(sizeof(double) * G);
  double *lp = malloc(sizeof(double) * G), *adj = malloc(sizeof(double) * G);
  double *lfc = malloc(sizeof(double) * G);
  c[p > 1 ? 1 : 0] = 1.0;
  orc_test_de(o.beta, o.bread, o.sandwich, G, p, n, c, 10.0, pv, lp, adj, lfc);
  for (int m = ORC_PADJ_BH; m <= ORC_PADJ_NONE; ++m) {
    orc_p_adjust(pv, G, m, adj);
    for (int64_t g = 0; g < G; ++g)
      CHECK(isnan(pv[g]) || (adj[g] >= pv[g] * (1 - 1e-12) && adj[g] <= 1.0),
            "p.adjust %d gene %ld: %g from %g", m, (long)g, adj[g], pv[g]);
  }
  double *sf = malloc(sizeof(double) * n);
  for (int method = 0; method < 2; ++method) {
    orc_size_factors(Y, G, n, n, method, sf);
    for (int64_t i = 0; i < n; ++i) CHECK(sf[i] > 0.0, "size factor %g", sf[i]);
  }
  double *meat = malloc(sizeof(double) * p * p), b1[16], *sfe = malloc(sizeof(double) * n);
  for (int j = 0; j < p; ++j) b1[j] = o.beta[1 + j * G];
  for (int64_t i = 0; i < n; ++i) sfe[i] = exp(off[i]);
  orc_meat_hc0(X, Y + n, b1, o.theta[1], sfe, n, p, meat);
  free(sfe);
  for (int q = 0; q < p * p; ++q)
    CHECK(!isfinite(o.theta[1]) || isfinite(meat[q]), "HC0 meat (od %d, n %ld, theta %g)", od,
          (long)n, o.theta[1]);
  free(meat);
  free(sf);
  free(c);
  free(pv);
  free(lp);
  free(adj);
  free(lfc);
  free(o.beta);
  free(o.theta);
  free(o.beta_iters);
  free(o.theta_iters);
  free(o.bread);
  free(o.sandwich);
  free(o.alpha0);
  free(o.beta0);
  free(Y);
  free(X);
  free(off);
}

int main(void) {
  /* special functions over their ranges */
  for (double x = 1e-6; x < 1e7; x *= 1.7) {
    CHECK(isfinite(orc_digamma(x)) && isfinite(orc_trigamma(x)), "psi at %g", x);
    CHECK(isfinite(orc_lgamma_diff(x, 3.0)) && isfinite(orc_dpsi_diff(x, 7.0)) &&
              isfinite(orc_dpsi1_diff(x, 2.0)),
          "differences at %g", x);
  }
  for (double t = 0.01; t < 60.0; t *= 1.5) {
    double lp;
    const double pv = orc_pt2_upper(t, 20.0, &lp);
    CHECK(pv >= 0.0 && pv <= 1.0 && lp <= 0.0, "pt at %g", t);
  }
  double gx[16], gw[16];
  orc_gh16(gx, gw);
  double lp, lm;
  CHECK(orc_h_log_gh_pmhalf(200.0, 230.0, 1e-10, &lp, &lm) == 1 && isfinite(lp - lm),
        "H_log_gh");
  CHECK(orc_h_log_gh_pmhalf(10.0, 5.0, 1e-10, &lp, &lm) == 0, "H_log_gh with t <= s");
  /* every fit mode, init mode and the clustered sandwich */
  fit_case(24, 300, 2, ORC_OD_MOM, ORC_INIT_QR, 1, ORC_MLE_NLMINB);
  fit_case(16, 400, 4, ORC_OD_MLE, ORC_INIT_QR, 0, ORC_MLE_NLMINB);
  fit_case(12, 300, 3, ORC_OD_MLE, ORC_INIT_ROUGH, 1, ORC_MLE_ROOT);
  fit_case(12, 500, 2, ORC_OD_NEW, ORC_INIT_UNIVARIATE, 0, ORC_MLE_NLMINB);
  fit_case(10, 200, 5, ORC_OD_NONE, ORC_INIT_QR, 1, ORC_MLE_NLMINB);
  fit_case(3, 8, 2, ORC_OD_NEW, ORC_INIT_QR, 0, ORC_MLE_NLMINB);
  fit_case(8, 1, 1, ORC_OD_MOM, ORC_INIT_QR, 0, ORC_MLE_NLMINB);
  if (fails) {
    fprintf(stderr, "%d check(s) failed\n", fails);
    return 1;
  }
  printf("oracle selftest passed\n");
  return 0;
}
