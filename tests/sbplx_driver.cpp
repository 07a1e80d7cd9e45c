// Test driver: runs the library's resumable SBPLX (header
// pgure-svt_amd/csrc/sbplx.hpp) to completion against a caller-supplied
// objective, so its evaluation sequence can be compared with the oracle's.
#include "../pgure-svt_amd/csrc/sbplx.hpp"

extern "C" int product_sbplx(double (*f)(double, void*), void* ctx, double start, double lb, double ub, double step,
                             double ftol, double xtol_abs, int maxeval, double* x_last, int* nevals) {
  pgsvt::Sbplx1D o(start, lb, ub, step, ftol, xtol_abs, maxeval);
  int rc;
  while ((rc = o.run()) == pgsvt::Sbplx1D::kNeed) o.supply(f(o.pending(), ctx));
  *x_last = o.last_x();
  *nevals = o.evaluations();
  return rc;
}
