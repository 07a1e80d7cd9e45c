// sbplx.hpp -- one-variable LN_SBPLX lambda search, resumable.
//
// Behaviour of PGURE<T>::Optimize (src/pgure.hpp:196-237): NLopt 2.6.2
// LN_SBPLX on lambda with lower bound 0, upper bound `ub`, relative f
// tolerance `tol`, absolute x tolerance 1e-12, evaluation cap `maxeval` and
// initial step sqrt(start); the value returned is the last lambda evaluated
// (the objective stores x[0] into the member lambda at src/pgure.hpp:128).
// For one variable Subplex (Rowan 1990, as implemented in NLopt's sbplx.c)
// is a sequence of Nelder-Mead runs (NLopt nldrmd: reflection 1, expansion 2,
// contraction 1/2, shrink 1/2) each stopped when the simplex has shrunk to
// psi = 1/4 of its initial size; between runs the step is rescaled by psi
// with the sign of the last move, and the search stops on the relative f
// tolerance or the x tolerance.
//
// Resumable: `run()` replays the search from the start against the values
// recorded so far and stops with `kNeed` at the first evaluation it has no
// value for (every call site returns at once, before any bookkeeping).  The
// caller evaluates that lambda (on the GPU, batched over many windows) and
// calls `run()` again.  Replay is deterministic, so the i-th evaluation of
// every replay is the same lambda.
#pragma once

#include <cmath>
#include <cstdint>
#include <limits>
#include <vector>

namespace pgsvt {

class Sbplx1D {
 public:
  enum Result { kNeed = 0, kFtol = 3, kXtol = 4, kMaxEval = 5, kFailure = -1, kInvalid = -2 };

  Sbplx1D(double start, double lb, double ub, double step, double ftol_rel, double xtol_abs, int maxeval)
      : start_(start), lb_(lb), ub_(ub), step_(step), ftol_(ftol_rel), xtol_(xtol_abs), maxeval_(maxeval) {}

  // Continue the search.  Returns kNeed with pending() set, or a final code.
  int run() {
    cursor_ = 0;
    need_ = false;
    const int rc = minimize();
    return need_ ? kNeed : rc;
  }
  double pending() const { return pending_; }
  void supply(double f) { values_.push_back(f); }
  int evaluations() const { return (int)values_.size(); }
  double last_x() const { return xs_.empty() ? start_ : xs_.back(); }
  double best_x() const { return best_x_; }
  double best_f() const { return best_f_; }

 private:
  double start_, lb_, ub_, step_, ftol_, xtol_;
  bool need_ = false;  // the replay reached an evaluation with no value yet
  int maxeval_;
  std::vector<double> values_, xs_;
  size_t cursor_ = 0;
  double pending_ = 0, best_x_ = 0, best_f_ = 0;
  int nevals_ = 0;

  double f(double x) {
    if (cursor_ < values_.size()) {
      if (cursor_ >= xs_.size()) xs_.push_back(x);
      return values_[cursor_++];
    }
    if (xs_.size() == values_.size()) xs_.push_back(x);
    pending_ = x;
    need_ = true;
    return std::numeric_limits<double>::quiet_NaN();
  }

  static bool relstop(double vold, double vnew, double reltol, double abstol) {
    if (std::isinf(vold)) return false;
    return std::fabs(vnew - vold) < abstol ||
           std::fabs(vnew - vold) < reltol * (std::fabs(vnew) + std::fabs(vold)) * 0.5 ||
           (reltol > 0 && vnew == vold);
  }
  static bool close_(double a, double b) { return std::fabs(a - b) <= 1e-13 * (std::fabs(a) + std::fabs(b)); }
  bool reflect(double& xnew, double c, double scale, double xold) const {
    double nx = c + scale * (c - xold);
    if (nx < lb_) nx = lb_;
    if (nx > ub_) nx = ub_;
    const bool eqc = close_(nx, c), eqo = close_(nx, xold);
    xnew = nx;
    return !(eqc || eqo);
  }
  // bookkeeping of one evaluation: best point (ties go to the newer point)
  // and the evaluation cap
  bool count(double xc, double fc, double& x, double& minf) {
    nevals_++;
    if (fc <= minf) {
      minf = fc;
      x = xc;
    }
    return maxeval_ > 0 && nevals_ >= maxeval_;
  }

  int nelder_mead(double& x, double& minf, double xstep, double psi) {
    double pf[2], px[2];
    double init_diam = 0;
    px[0] = x;
    pf[0] = minf;
    px[1] = x + xstep;
    if (px[1] > ub_) px[1] = (ub_ - x > std::fabs(xstep) * 0.1) ? ub_ : x - std::fabs(xstep);
    if (px[1] < lb_) {
      if (x - lb_ > std::fabs(xstep) * 0.1) {
        px[1] = lb_;
      } else {
        px[1] = x + std::fabs(xstep);
        if (px[1] > ub_) px[1] = 0.5 * ((ub_ - x > x - lb_ ? ub_ : lb_) + x);
      }
    }
    if (close_(px[1], x)) return kFailure;
    pf[1] = f(px[1]);
    if (need_) return kNeed;
    if (count(px[1], pf[1], x, minf)) return kMaxEval;
    for (;;) {
      const int lo = (pf[1] < pf[0]) ? 1 : 0, hi = 1 - lo;
      const double fl = pf[lo], xl = px[lo];
      double fh = pf[hi];
      const double xh = px[hi];
      if (init_diam == 0) init_diam = std::fabs(xl - xh);
      const double c = xl;
      if (std::fabs(xl - xh) < psi * init_diam) return kXtol;
      double xr;
      if (!reflect(xr, c, 1.0, xh)) return kXtol;
      const double fr = f(xr);
      if (need_) return kNeed;
      if (count(xr, fr, x, minf)) return kMaxEval;
      if (fr < fl) {
        double xe;
        const bool ok = reflect(xe, c, 2.0, xh);
        px[hi] = xe;
        if (!ok) return kXtol;
        fh = f(xe);
        if (need_) return kNeed;
        if (count(xe, fh, x, minf)) return kMaxEval;
        if (fh >= fr) {
          fh = fr;
          px[hi] = xr;
        }
      } else {
        double xc;
        if (!reflect(xc, c, fh <= fr ? -0.5 : 0.5, xh)) return kXtol;
        const double fc = f(xc);
        if (need_) return kNeed;
        if (count(xc, fc, x, minf)) return kMaxEval;
        if (fc < fr && fc < fh) {
          px[hi] = xc;
          fh = fc;
        } else {  // shrink towards the best point and restart
          px[hi] = xl + 0.5 * (px[hi] - xl);
          pf[hi] = f(px[hi]);
          if (need_) return kNeed;
          if (count(px[hi], pf[hi], x, minf)) return kMaxEval;
          continue;
        }
      }
      pf[hi] = fh;
    }
  }

  int minimize() {
    const double psi = 0.25;
    nevals_ = 0;
    double x = start_, minf = 0;
    if (lb_ > ub_ || x < lb_ || x > ub_ || step_ == 0) return kInvalid;
    minf = f(x);
    if (need_) return kNeed;
    nevals_++;
    best_x_ = x;
    best_f_ = minf;
    if (maxeval_ > 0 && nevals_ >= maxeval_) return kMaxEval;
    double xstep = step_;
    for (;;) {
      const double xprev = x, fprev = minf;
      int ret = nelder_mead(x, minf, xstep, psi);
      if (ret == kNeed) return kNeed;
      best_x_ = x;
      best_f_ = minf;
      if (ret == kFailure) return kXtol;
      if (ret != kXtol) return ret;
      if (relstop(fprev, minf, ftol_, 0.0)) return kFtol;
      if (relstop(xprev, x, 0.0, xtol_)) {
        if (!(std::fabs(xstep) * psi > xtol_)) return kXtol;
      }
      const double dx = x - xprev;
      xstep = (dx == 0) ? -(xstep * psi) : std::copysign(xstep * psi, dx);
    }
  }
};

}  // namespace pgsvt
