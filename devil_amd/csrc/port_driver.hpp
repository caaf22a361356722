// One-parameter restatement of stats::nlminb's PORT driver (DMNHB: Newton
// steps with the analytic Hessian in a trust region, box bounds) as
// fit_dispersion calls it (R/overdispersion.R:53-73): objective -loglik,
// gradient -score, Hessian -dscore in log(theta), bounds [log 1e-16,
// log 1e16], rel.tol = tolerance, iter.max = max_iter, every other control
// at R's default (eval.max 200, abs.tol 0, x.tol 1.5e-8, xf.tol 2.2e-14,
// step.min = step.max = 1, sing.tol = rel.tol).
//
// The driver runs in reverse communication, as PORT itself does: each call
// of port_advance hands it f, g and h at the point it asked for last
// (PortState::xe) and gets back either the next point to evaluate or the
// result.  One IRLS-style round of the MLE kernels evaluates every active
// gene at its xe.
//
// Per iteration from the accepted point x0: the Newton step -g/h when h > 0
// and |g/h| <= (1 + PHMXFC) * radius, else the step of length radius
// downhill (the locally constrained step in one dimension), cut at the
// bounds.  The candidate is assessed as PORT's DA7SST documents it: accepted
// when f fell by more than TUNER2 of the predicted reduction (radius scaled
// by Fletcher's factors, or kept when the decrease was good), else the step
// is recomputed with a smaller radius; a very good non-Newton step is retried
// with a larger radius and the better of the two kept.  Convergence tests
// after each assessment: relative function (f0 - f <= 2 preduc and the
// Newton step's predicted reduction <= rel.tol |f0|), x (a Newton step with
// |dx| / (|x| + |x0|) <= x.tol), singular (the predicted reduction of a step
// of length step.min <= sing.tol |f0| when no Newton step is taken) and
// false convergence (no decrease and |dx| / (|x| + |x0|) <= xf.tol).
#pragma once
#include <hip/hip_runtime.h>

namespace dv {

struct PortState {
  double x, f, g, h;      // accepted point
  double x0, f0, g0, h0;  // start of the current iteration
  double radius, radfac, last_step;
  double flstgd, xlstgd, glstgd, hlstgd, plstgd, gthstg, dstsav;  // saved good step
  double xe;              // point whose f, g, h the next call brings
  double s, preduc, nreduc, dst0, dstnrm, gtstep;  // the candidate step
  int32_t niter, nf, radinc, restor, irc_in, newton, acc_irc, phase, code, pad;
};

enum { kPortStart = 0, kPortTrial = 1, kPortDone = 2 };

struct PortParams {
  double lo, up, rfctol, sctol;
  int iter_max;
};

constexpr double kPtLmax0 = 1.0, kPtLmaxs = 1.0, kPtXctol = 1.5e-8, kPtXftol = 2.2e-14;
constexpr double kPtTuner1 = 0.1, kPtTuner2 = 1e-4, kPtTuner3 = 0.75, kPtTuner4 = 0.5,
                 kPtTuner5 = 0.75;
constexpr double kPtIncfac = 2.0, kPtDecfac = 0.5, kPtRdfcmn = 0.1, kPtRdfcmx = 4.0,
                 kPtPhmxfc = 0.1;
constexpr int kPtEvalMax = 200;

// trust-region step from x0 (g, h) with radius rad, cut at [lo, up]
__device__ inline void port_step(PortState& st, double rad, const PortParams& pp) {
  const double g = st.g0, h = st.h0, x0 = st.x0;
  int newton = 0;
  double s;
  if (h > 0.0) {
    const double sn = -g / h;
    st.dst0 = fabs(sn);
    st.nreduc = 0.5 * g * g / h;
    if (st.dst0 <= (1.0 + kPtPhmxfc) * rad) {
      s = sn;
      newton = 1;
    } else {
      s = g > 0 ? -rad : rad;
    }
  } else {
    st.dst0 = -1.0;
    st.nreduc = 0.0;
    s = g > 0 ? -rad : (g < 0 ? rad : 0.0);
  }
  if (x0 + s > pp.up) {
    s = pp.up - x0;
    newton = 0;
  } else if (x0 + s < pp.lo) {
    s = pp.lo - x0;
    newton = 0;
  }
  st.s = s;
  st.newton = newton;
  st.dstnrm = fabs(s);
  st.gtstep = g * s;
  st.preduc = newton ? st.nreduc : -(g * s + 0.5 * h * s * s);
}

// predicted reduction of the step of length step.min from x0
__device__ inline double port_preduc_lmaxs(const PortState& st, const PortParams& pp) {
  PortState t = st;
  port_step(t, kPtLmaxs, pp);
  return t.preduc;
}

// One call per evaluation: f, g, h at the point st.xe asked for last.
// Returns 0 with the next point in st.xe, or 1 when the driver has stopped
// (st.x is nlminb's par, st.code its PORT return code: 3 x-, 4 relative
// function, 5 both, 7 singular, 8 false convergence, 9 evaluation limit, 10
// iteration limit, 63 objective not finite at the start).
__device__ inline int port_advance(PortState& st, double fe, double ge, double he,
                                   const PortParams& pp) {
  enum { kBegin, kCandidate, kAssess };
  int action;
  double ft = fe, gt = ge, ht = he;
  if (st.phase == kPortStart) {
    st.x = st.xe;
    st.nf = 1;
    st.niter = 0;
    st.code = 0;
    if (!isfinite(fe)) {
      st.code = 63;
      st.phase = kPortDone;
      return 1;
    }
    st.f = fe;
    st.g = ge;
    st.h = he;
    st.radfac = 1.0;
    st.last_step = 0.0;
    // the first step may be (1 + PHMXFC) times the radius: keep it <= step.max
    st.radius = kPtLmax0 / (1.0 + kPtPhmxfc);
    action = kBegin;
  } else {
    st.nf += 1;
    action = kAssess;
  }
  for (;;) {
    if (action == kBegin) {
      if (st.niter >= pp.iter_max) {
        st.code = 10;
        st.phase = kPortDone;
        return 1;
      }
      st.niter += 1;
      st.x0 = st.x;
      st.f0 = st.f;
      st.g0 = st.g;
      st.h0 = st.h;
      if (st.niter > 1) st.radius = st.radfac * st.last_step;
      st.irc_in = 4;
      st.radinc = 0;
      st.restor = 0;
      st.flstgd = st.f0;
      st.xlstgd = st.x0;
      st.glstgd = st.g0;
      st.hlstgd = st.h0;
      st.plstgd = st.gthstg = st.dstsav = 0.0;
      action = kCandidate;
      continue;
    }
    if (action == kCandidate) {
      port_step(st, st.radius, pp);
      if (st.dstnrm > 0.0) {
        if (st.nf >= kPtEvalMax) {
          st.code = 9;
          st.phase = kPortDone;
          return 1;
        }
        st.xe = st.x0 + st.s;
        st.phase = kPortTrial;
        return 0;
      }
      ft = st.f0;  // zero step: assessed at x0 without an evaluation
      gt = st.g0;
      ht = st.h0;
      action = kAssess;
      continue;
    }
    // ---- assess the candidate (DA7SST) ----
    const double x0 = st.x0, f0 = st.f0;
    const double xt = x0 + st.s;
    if (!isfinite(ft)) {  // objective not computable: halve the radius
      st.radius = kPtDecfac * st.dstnrm;
      st.radinc -= 1;
      st.irc_in = 5;
      action = kCandidate;
      continue;
    }
    const double den = fabs(xt) + fabs(x0);
    const double reldx = den > 0.0 ? st.dstnrm / den : 0.0;
    double fc = ft, xc = xt, gc = gt, hc = ht;
    double preduc = st.preduc, gtstep = st.gtstep, dstnrm = st.dstnrm;
    int goodx = 1, irc = 0;
    double rfac1 = 1.0, radf = 1.0;
    if (st.irc_in == 5 && !(fc < st.flstgd) && st.flstgd < f0) {
      // the recomputed step is worse than the saved good step: restore it
      rfac1 = dstnrm / st.dstsav;
      fc = st.flstgd;
      xc = st.xlstgd;
      gc = st.glstgd;
      hc = st.hlstgd;
      preduc = st.plstgd;
      gtstep = st.gthstg;
      dstnrm = st.dstsav;
      goodx = 0;
      st.restor = 1;
    }
    const double fdif = f0 - fc;
    bool save = false;
    if (!(fdif > kPtTuner2 * preduc) && st.radinc <= 0) {
      // no or only a trivial decrease: recompute with a smaller radius
      if (!(fc < f0)) {
        st.flstgd = fc;
        fc = f0;
        xc = x0;
        gc = st.g0;
        hc = st.h0;
        st.restor = 1;
      }
      irc = 5;
      st.radinc -= 1;
      const double emax = gtstep + fdif;
      radf = 0.5 * rfac1;
      if (emax < gtstep) radf = rfac1 * fmax(kPtRdfcmn, 0.5 * gtstep / emax);
      if (reldx <= kPtXftol) irc = 12;
      else save = fc < f0;
    } else if (!(fdif > preduc * kPtTuner1)) {
      // decrease much less than predicted: accept with a smaller radius
      irc = 4;
      const double emax = gtstep + fdif;
      radf = 0.5 * rfac1;
      if (emax < gtstep) radf = rfac1 * fmax(kPtRdfcmn, 0.5 * gtstep / emax);
      if (reldx <= kPtXftol) irc = 12;
    } else if (fdif < -kPtTuner3 * gtstep || st.radinc < 0 || st.restor == 1) {
      radf = 1.0;  // good decrease: radius unchanged
      irc = 3;
    } else {
      // very good decrease: a longer step looks worthwhile
      radf = kPtRdfcmx;
      if (fdif < (0.5 / radf - 1.0) * gtstep)
        radf = fmax(kPtIncfac, 0.5 * gtstep / (gtstep + fdif));
      irc = 4;
      if (!st.newton &&
          !(st.dst0 >= 0.0 && (st.dst0 < 2.0 * dstnrm || st.nreduc < 1.2 * fdif))) {
        irc = 5;
        st.radinc += 1;
        save = true;
      }
    }
    if (save) {
      st.flstgd = fc;
      st.xlstgd = xc;
      st.glstgd = gc;
      st.hlstgd = hc;
      if (st.restor != 1) st.restor = 2;
      st.dstsav = dstnrm;
      st.plstgd = preduc;
      st.gthstg = gtstep;
    }
    if (st.restor == 1 && st.flstgd < f0) {
      st.restor = 3;
      xc = st.xlstgd;
      fc = st.flstgd;
      gc = st.glstgd;
      hc = st.hlstgd;
    }
    // convergence tests
    if (!(0.5 * fdif > preduc)) {
      const double emax = pp.rfctol * fabs(f0), emaxs = pp.sctol * fabs(f0);
      if (preduc <= emaxs && (dstnrm > kPtLmaxs || st.newton)) irc = 11;
      if (st.dst0 >= 0.0) {
        int i = 0;
        if ((st.nreduc > 0.0 && st.nreduc <= emax) || (st.nreduc == 0.0 && preduc == 0.0))
          i = 2;
        if (st.newton && reldx <= kPtXctol && goodx) i += 1;
        if (i > 0) irc = i + 6;
      }
      if (!(irc > 5 && irc != 12) && !st.newton) {
        int test;
        if (dstnrm > kPtLmaxs) {
          const double xm = kPtLmaxs / dstnrm;
          test = !(0.5 * dstnrm <= kPtLmaxs) && !(xm * (2.0 - xm) * preduc >= emaxs);
        } else {
          test = !(preduc >= emaxs) && (st.dst0 <= 0.0 || !(0.5 * st.dst0 <= kPtLmaxs));
        }
        if (test && port_preduc_lmaxs(st, pp) <= emaxs) irc = 11;
      }
    }
    if (irc >= 7 && irc <= 12) {
      const bool better = fc < f0;
      st.x = better ? xc : x0;
      st.f = better ? fc : f0;
      st.code = irc - 4;
      st.phase = kPortDone;
      return 1;
    }
    if (irc == 3 || irc == 4) {
      st.x = xc;
      st.f = fc;
      st.g = gc;
      st.h = hc;
      st.radfac = radf;
      st.acc_irc = irc;
      const double s = xc - x0;
      if (irc == 3) {  // gradient tests
        const double temp1 = st.h0 * s + st.g0 - gc;
        if (fabs(temp1) > kPtTuner4 * fabs(st.g0) && gc * s < kPtTuner5 * (st.g0 * s))
          st.radfac = kPtIncfac;
      }
      st.last_step = fabs(s);
      action = kBegin;
      continue;
    }
    // irc 5: recompute from x0 with the radius scaled
    st.radius = radf * dstnrm;
    st.irc_in = 5;
    action = kCandidate;
  }
}

}  // namespace dv
