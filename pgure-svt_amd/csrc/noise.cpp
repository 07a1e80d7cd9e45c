// noise.cpp -- host Poisson-Gaussian noise estimator (src/noise.hpp).
//
// Same arithmetic, in the same order, as the reference: quadtree with the
// 2.5% F-test, Huber IRLS robust mean, MAD robust variance of the
// Laplacian-filtered leaf, robust WLS line fit of variance against mean.
// Four changes make it fast without changing a single rounding:
//   * the robust mean's interquartile distance is taken on r = A - m; since
//     rounding is monotone, sorted(r)[k] = fl(sorted(A)[k] - m), so the two
//     order statistics of A it needs are selected once per leaf instead of
//     sorting r in every IRLS iteration;
//   * medians and order statistics are selected, not sorted (a monotone
//     histogram narrows large leaves to one bucket first);
//   * each robust-mean iteration is one pass over the leaf (the reference's
//     two passes fused, every sum in its own order);
//   * repeated quadtree nodes are evaluated once.
#include "noise.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <string>

namespace pgsvt {
namespace {

typedef double v2d __attribute__((vector_size(16)));
typedef int64_t v2l __attribute__((vector_size(16)));
inline v2d vabs(v2d x) {
  const v2l mask = {INT64_MAX, INT64_MAX};
  return (v2d)((v2l)x & mask);
}

// Armadillo accu order: even indices, odd indices, then their sum.
template <class F>
inline double accu2f(size_t n, F f) {
  double a1 = 0.0, a2 = 0.0;
  size_t i = 0;
  for (; i + 1 < n; i += 2) {
    a1 += f(i);
    a2 += f(i + 1);
  }
  if (i < n) a1 += f(i);
  return a1 + a2;
}

// Order statistics of ranks ks[0] > ks[1] > ... (nk of them, 0-based) of
// v[0 .. n), exact.  Large inputs go through a histogram: the bucket index
// floor((x - lo) * scale) never decreases with x (every step rounds
// monotonically), so the element of rank k is the element of rank
// k - (count of the lower buckets) within its bucket, selected there.  Small
// inputs (and degenerate ranges) use nth_element on a copy; after selecting
// rank k, every smaller rank lies in the first k elements (the next rank
// down is their maximum).
void order_stats(const double* v, size_t n, const size_t* ks, int nk, double* out) {
  auto small = [&] {
    thread_local std::vector<double> c;
    c.assign(v, v + n);
    size_t end = n;
    for (int j = 0; j < nk; ++j) {
      if (ks[j] + 1 == end)
        out[j] = *std::max_element(c.begin(), c.begin() + end);
      else {
        std::nth_element(c.begin(), c.begin() + ks[j], c.begin() + end);
        out[j] = c[ks[j]];
      }
      end = ks[j];
    }
  };
  if (n < 4096) return small();
  double lo = v[0], hi = v[0];
  for (size_t i = 1; i < n; ++i) {
    lo = std::min(lo, v[i]);
    hi = std::max(hi, v[i]);
  }
  constexpr uint32_t B = 4096;
  const double scale = (double)(B - 1) / (hi - lo);
  if (!(hi > lo) || !std::isfinite(scale)) return small();
  auto bucket = [&](double x) { return std::min((uint32_t)((x - lo) * scale), B - 1); };
  std::vector<uint32_t> cnt(B, 0);
  for (size_t i = 0; i < n; ++i) ++cnt[bucket(v[i])];
  uint32_t bj[2];
  size_t below[2];
  for (int j = 0; j < nk; ++j) {
    size_t c = 0;
    uint32_t b = 0;
    while (c + cnt[b] <= ks[j]) c += cnt[b++];
    bj[j] = b;
    below[j] = c;
  }
  std::vector<double> t[2];
  for (int j = 0; j < nk; ++j) t[j].reserve(cnt[bj[j]]);
  for (size_t i = 0; i < n; ++i) {
    const uint32_t b = bucket(v[i]);
    for (int j = 0; j < nk; ++j)
      if (b == bj[j]) t[j].push_back(v[i]);
  }
  for (int j = 0; j < nk; ++j) {
    const size_t r = ks[j] - below[j];
    std::nth_element(t[j].begin(), t[j].begin() + r, t[j].end());
    out[j] = t[j][r];
  }
}

// arma::median: even counts average the upper and lower middle as
// upper + (lower - upper) / 2
double median_of(const std::vector<double>& v) {
  const size_t n = v.size(), half = n / 2;
  if (n % 2 == 0) {
    const size_t ks[2] = {half, half - 1};
    double o[2];
    order_stats(v.data(), n, ks, 2, o);
    return o[0] + (o[1] - o[0]) / 2.0;
  }
  double o;
  order_stats(v.data(), n, &half, 1, &o);
  return o;
}

uint32_t iqd_m(uint32_t N) { return (uint32_t)std::floor((std::floor((N + 1) / 2) + 1) / 2); }

// interquartile distance of an unsorted vector (two order statistics)
double iqd_select(const std::vector<double>& v) {
  const uint32_t N = (uint32_t)v.size(), m = iqd_m(N);
  const size_t ks[2] = {N - m - 1, m - 1};
  double o[2];
  order_stats(v.data(), N, ks, 2, o);
  return o[0] - o[1];
}

inline void huber(const double* x, double* w, size_t n) {
  const double p = 0.75;
  for (size_t k = 0; k < n; ++k) {
    const double a = std::fabs(x[k]);
    w[k] = (a < p) ? 1. : p / a;
  }
}

// RobustMeanEstimate (src/noise.hpp:238-271).  The interquartile distance of
// r = A - m only reads two order statistics of r, and sorted(r)[k] =
// fl(sorted(A)[k] - m), so the two order statistics of A are selected once.
double robust_mean(const std::vector<double>& A) {
  const size_t N = A.size();
  const uint32_t mq = iqd_m((uint32_t)N);
  double a_hi, a_lo;
  {
    const size_t ks[2] = {N - mq - 1, mq - 1};
    double o[2];
    order_stats(A.data(), N, ks, 2, o);
    a_hi = o[0];
    a_lo = o[1];
  }
  const double tol = 1e-6, eps = 1e-12;
  double m = 0, m0 = 1e12;
  // sum(w A) and sum(w), each with its own even/odd accumulator pair in
  // index order; the first weights are all 1
  double wa1 = 0, wa2 = 0, w1 = 0, w2 = 0;
  size_t k = 0;
  for (; k + 1 < N; k += 2) {
    wa1 += A[k];
    wa2 += A[k + 1];
    w1 += 1.;
    w2 += 1.;
  }
  if (k < N) {
    wa1 += A[k];
    w1 += 1.;
  }
  // One pass per iteration: sum |A - m| of this iteration, the next
  // weights and, from them, the next iteration's sum(w A) and sum(w).
  // Every sum keeps its accumulators and order, so fusing the reference's
  // two passes changes no rounding (the sums of the last weights go
  // unused when the loop ends).
  for (uint32_t it = 0; it < 10000; ++it) {
    m = wa1 + wa2;
    const double aux = w1 + w2;
    m = (std::fabs(aux) < eps) ? m0 : m / aux;
    const double d = ((a_hi - m) - (a_lo - m)) + eps;
    const double inv = 1. / d;
    // two elements at a time: lane 0 carries the even-index sums, lane 1
    // the odd ones, as the scalar accumulator pairs do; the Huber weight is
    // a select (0.75 / a is exact either way), which keeps the outliers,
    // about a third of a leaf, off the branch predictor
    const v2d vm = {m, m}, vinv = {inv, inv}, vp = {0.75, 0.75}, one = {1., 1.};
    v2d ve = {0., 0.}, vwa = {0., 0.}, vws = {0., 0.};
    for (k = 0; k + 1 < N; k += 2) {
      v2d av;
      std::memcpy(&av, &A[k], sizeof av);
      const v2d r = av - vm;
      const v2d a = vabs(r * vinv);
      const v2d wq = (a < vp) ? one : vp / a;
      ve += vabs(r);
      vwa += wq * av;
      vws += wq;
    }
    double e1 = ve[0], e2 = ve[1];
    wa1 = vwa[0];
    wa2 = vwa[1];
    w1 = vws[0];
    w2 = vws[1];
    if (k < N) {
      const double r = A[k] - m;
      const double a = std::fabs(r * inv);
      const double wq = (a < 0.75) ? 1. : 0.75 / a;
      e1 += std::fabs(r);
      wa1 += wq * A[k];
      w1 += wq;
    }
    const double e = (e1 + e2) / (double)N;
    if (std::fabs(m0 - m) < tol || e < tol) break;
    m0 = m;
  }
  return m;
}

// RobustVarianceEstimate (src/noise.hpp:232-236)
double robust_var(std::vector<double> a) {
  const double med = median_of(a);
  for (double& v : a) v = std::fabs(v - med);
  const double sig = 1.4826 * median_of(a);
  return sig * sig;
}

const uint32_t kDof[12] = {2, 4, 8, 16, 32, 64, 128, 256, 512, 1024, 2048, 4096};
const double kF0025[12] = {15.4392, 2.86209, 1.64602, 1.27893, 1.13046, 1.06318,
                           1.03110, 1.01543, 1.00769, 1.00384, 1.00192, 1.00096};

// SplitBlockQ (src/noise.hpp:182-221) on the N x N block at (i0, j0)
bool split_block(const double* F, uint32_t n, uint32_t i0, uint32_t j0, uint32_t N) {
  if (N <= 8) return false;
  const double sq = std::sqrt((double)(5 * 5 + 5));
  const uint32_t R = N * N;
  const double OoR = 1.0 / R, OoRm1 = 1.0 / (R - 1);
  std::vector<double> a(R), rs(R);
  // element k = y + N x (column-major), residual of the 5-point stencil
  // with the reference's wrap-around neighbours; rows 1 .. N - 2 two at a
  // time (the same operations per lane), the wrapped rows 0 and N - 1 alone
  const v2d v5 = {5., 5.}, vsq = {sq, sq};
  for (uint32_t x = 0; x < N; ++x) {
    const uint32_t xp = (x + 1 == N) ? 1 : x + 1, xm = (x == 0) ? N - 2 : x - 1;
    const double* C = &F[i0 + (size_t)n * (j0 + x)];
    const double* Cp = &F[i0 + (size_t)n * (j0 + xp)];
    const double* Cm = &F[i0 + (size_t)n * (j0 + xm)];
    double* ak = &a[(size_t)N * x];
    double* rk = &rs[(size_t)N * x];
    auto one = [&](uint32_t y) {
      const uint32_t yp = (y + 1 == N) ? 1 : y + 1, ym = (y == 0) ? N - 2 : y - 1;
      ak[y] = C[y];
      rk[y] = (5u * C[y] - (C[yp] + C[ym] + Cm[y] + Cp[y])) / sq;
    };
    one(0);
    uint32_t y = 1;
    for (; y + 2 < N; y += 2) {
      v2d c, cu, cd, cm, cp;
      std::memcpy(&c, &C[y], sizeof c);
      std::memcpy(&cu, &C[y + 1], sizeof cu);
      std::memcpy(&cd, &C[y - 1], sizeof cd);
      std::memcpy(&cm, &Cm[y], sizeof cm);
      std::memcpy(&cp, &Cp[y], sizeof cp);
      const v2d r = (v5 * c - (cu + cd + cm + cp)) / vsq;
      std::memcpy(&ak[y], &c, sizeof c);
      std::memcpy(&rk[y], &r, sizeof r);
    }
    for (; y < N; ++y) one(y);
  }
  // four sums, each with its own even/odd accumulator pair (as accu):
  // lane 0 the even elements, lane 1 the odd ones
  v2d vz = {0., 0.}, vr = {0., 0.};
  size_t k = 0;
  for (; k + 1 < R; k += 2) {
    v2d av, rv;
    std::memcpy(&av, &a[k], sizeof av);
    std::memcpy(&rv, &rs[k], sizeof rv);
    vz += av;
    vr += rv;
  }
  double z1 = vz[0], z2 = vz[1], r1 = vr[0], r2 = vr[1];
  if (R % 2) {
    z1 += a[R - 1];
    r1 += rs[R - 1];
  }
  const double accuZ = (z1 + z2) * OoR, accuR = (r1 + r2) * OoR;
  const v2d mz = {accuZ, accuZ}, mr = {accuR, accuR};
  v2d vs = {0., 0.}, ve = {0., 0.};
  for (k = 0; k + 1 < R; k += 2) {
    v2d av, rv;
    std::memcpy(&av, &a[k], sizeof av);
    std::memcpy(&rv, &rs[k], sizeof rv);
    const v2d t = av - mz, u = rv - mr;
    vs += t * t;
    ve += u * u;
  }
  double s1 = vs[0], s2 = vs[1], e1 = ve[0], e2 = ve[1];
  if (R % 2) {
    const double t1 = a[R - 1] - accuZ, u1 = rs[R - 1] - accuR;
    s1 += t1 * t1;
    e1 += u1 * u1;
  }
  const double Sz = (s1 + s2) * OoRm1, Se = (e1 + e2) * OoRm1;
  const double stat = (Sz > Se) ? Sz / Se : Se / Sz;
  for (int k = 0; k < 12; ++k)
    if (kDof[k] == N) return stat > kF0025[k];
  return false;
}

// QuadTree (src/noise.hpp:419-458): after a split the recursion starts at the
// previous last node, not at the first new child.
void quadtree(const SplitTest& split, std::vector<std::array<uint32_t, 3>>& tree, std::vector<uint32_t>& del,
              uint32_t part) {
  const auto node = tree[part];
  const uint32_t i = node[0], j = node[1];
  uint32_t s = node[2];
  if (!split(i, j, s)) return;
  s /= 2;
  const uint32_t last = (uint32_t)tree.size() - 1;
  tree.push_back({i, j, s});
  tree.push_back({i + s, j, s});
  tree.push_back({i, j + s, s});
  tree.push_back({i + s, j + s, s});
  del.push_back(part);
  for (uint32_t it = last; it < last + 4; ++it) quadtree(split, tree, del, it);
}

// ConvolveFIR (src/noise.hpp:395-417) values, in any order (only their
// median is used)
std::vector<double> laplacian_values(const std::vector<double>& in, uint32_t N) {
  std::vector<double> out((size_t)N * N);
  auto I = [&](uint32_t r, uint32_t c) { return in[r + (size_t)N * c]; };
  const double nl_edge = (1.0 * 0.125) * -1.0, nl_mid = -1.0 * -1.0;
  const double nl[9] = {nl_edge, nl_edge, nl_edge, nl_edge, nl_mid, nl_edge, nl_edge, nl_edge, nl_edge};
  for (uint32_t x = 0; x < N; ++x) {
    for (uint32_t y = 0; y < N; ++y) {
      const uint32_t xp = (x + 1 == N) ? 1 : x + 1, yp = (y + 1 == N) ? 1 : y + 1;
      const uint32_t xm = (x == 0) ? N - 2 : x - 1, ym = (y == 0) ? N - 2 : y - 1;
      const double nb[9] = {I(xm, ym), I(xm, y), I(xm, yp), I(x, ym), I(x, y), I(x, yp), I(xp, ym), I(xp, y), I(xp, yp)};
      out[y + (size_t)N * x] = accu2f(9, [&](size_t k) { return nb[k] * nl[k]; });
    }
  }
  return out;
}

// WLSFit (src/noise.hpp:328-383)
void wls_fit(const std::vector<double>& x, const std::vector<double>& y, double& pa, double& pb) {
  const size_t N = x.size();
  const double tol = 1e-6, eps = 1e-12;
  double a0 = 1e12, b0 = 1e12, p0 = 0, p1 = 0;
  std::vector<double> w(N, 1.0), w2(N), r(N);
  for (uint32_t it = 0; it < 10000; ++it) {
    for (size_t k = 0; k < N; ++k) w2[k] = w[k] * w[k];
    const double sw2 = accu2f(N, [&](size_t k) { return w2[k]; });
    const double sw2x = accu2f(N, [&](size_t k) { return w2[k] * x[k]; });
    const double sw2y = accu2f(N, [&](size_t k) { return w2[k] * y[k]; });
    p0 = sw2 * accu2f(N, [&](size_t k) { return w2[k] * (x[k] * y[k]); }) - sw2x * sw2y;
    const double aux = sw2 * accu2f(N, [&](size_t k) { return w2[k] * (x[k] * x[k]); }) - sw2x * sw2x;
    p0 = (std::fabs(aux) < eps) ? a0 : p0 / aux;
    p1 = sw2y - p0 * sw2x;
    p1 = (std::fabs(aux) < eps) ? b0 : p1 / sw2;
    for (size_t k = 0; k < N; ++k) r[k] = y[k] - (x[k] * p0 + p1);
    const double e = accu2f(N, [&](size_t k) { return std::fabs(r[k]); }) / (double)N;
    if ((std::fabs(a0 - p0) < tol && std::fabs(b0 - p1) < tol) || e < tol) break;
    a0 = p0;
    b0 = p1;
    const double d = iqd_select(r) + eps;
    for (size_t k = 0; k < N; ++k) r[k] /= d;
    huber(r.data(), w.data(), N);
  }
  pa = p0;
  pb = p1;
}

// ComputeMode (src/noise.hpp:273-301)
double compute_mode(const std::vector<double>& A) {
  const size_t N = A.size();
  const double M = *std::max_element(A.begin(), A.end());
  const double dyn = 1. * N;
  std::vector<double> a(N);
  for (size_t k = 0; k < N; ++k) a[k] = std::round(A[k] * dyn / M);
  uint32_t maxCount = 0;
  double maxValue = 0.;
  for (size_t i = 0; i < N; ++i) {
    uint32_t count = 0;
    for (size_t j = 0; j < N; ++j) count += (a[j] == a[i]);
    if (count > maxCount) {
      maxCount = count;
      maxValue = a[i];
    }
  }
  return maxValue * (M / dyn);
}

}  // namespace

double split_threshold(uint32_t N) {
  if (N <= 8) return -1.0;
  for (int k = 0; k < 12; ++k)
    if (kDof[k] == N) return kF0025[k];
  return -1.0;
}

std::vector<std::array<uint32_t, 3>> quadtree_leaves(uint32_t n, const SplitTest& split) {
  std::vector<std::array<uint32_t, 3>> tree(1, {0u, 0u, n});
  std::vector<uint32_t> del;
  quadtree(split, tree, del, 0);
  std::sort(del.begin(), del.end());
  del.erase(std::unique(del.begin(), del.end()), del.end());
  if (!del.empty())
    for (size_t k = del.size() - 1; k > 0; --k) tree.erase(tree.begin() + del[k]);  // del[0] is never shed
  return tree;
}

FrameLeaves frame_leaves(const double* F, uint32_t n) {
  const std::vector<std::array<uint32_t, 3>> tree =
      quadtree_leaves(n, [&](uint32_t i, uint32_t j, uint32_t N) { return split_block(F, n, i, j, N); });
  FrameLeaves out;
  out.mean.reserve(tree.size());
  out.var.reserve(tree.size());
  // the reference's tree holds repeated nodes; identical blocks give
  // identical statistics, so each distinct block is evaluated once
  std::vector<std::array<uint32_t, 3>> seen;
  std::vector<std::pair<double, double>> stat;
  for (const auto& nd : tree) {
    size_t hit = seen.size();
    for (size_t k = 0; k < seen.size(); ++k)
      if (seen[k] == nd) hit = k;
    if (hit == seen.size()) {
      const uint32_t x = nd[0], y = nd[1], s = nd[2];
      std::vector<double> col((size_t)s * s);
      for (uint32_t c = 0; c < s; ++c)
        std::memcpy(&col[(size_t)s * c], &F[x + (size_t)n * (y + c)], s * sizeof(double));
      seen.push_back(nd);
      stat.emplace_back(robust_mean(col), robust_var(laplacian_values(col, s)));
    }
    out.mean.push_back(stat[hit].first);
    out.var.push_back(stat[hit].second);
  }
  return out;
}

void estimate_from_leaves(const std::vector<std::shared_ptr<const FrameLeaves>>& frames, uint32_t n,
                          uint32_t method, double& alpha, double& mu, double& sigma) {
  std::vector<double> means, vars;
  for (const auto& lv : frames) {
    for (double v : lv->mean)
      if (v >= 0.) means.push_back(v);
    for (double v : lv->var)
      if (v >= 0.) vars.push_back(v);
  }
  std::vector<size_t> idx(means.size());
  for (size_t k = 0; k < idx.size(); ++k) idx[k] = k;
  std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) { return means[a] < means[b]; });
  std::vector<double> rm(idx.size()), rv(idx.size());
  for (size_t k = 0; k < idx.size(); ++k) {
    rm[k] = means[idx[k]];
    rv[k] = vars[idx[k]];
  }
  double a = 0, b = 0;
  wls_fit(rm, rv, a, b);
  alpha = (alpha >= 0.) ? alpha : a;
  switch (method) {
    case 1: {
      const int L = (int)std::floor(1. * ((uint32_t)(n * n) / (uint32_t)means.size()));
      const uint32_t ie = (uint32_t)std::round(0.05 * L);
      if ((size_t)ie + 1 > rm.size())
        throw NoiseError("noise method 1 needs " + std::to_string((size_t)ie + 1) + " quadtree leaves, the window has " +
                         std::to_string(rm.size()));
      std::vector<double> ra(rm.begin(), rm.begin() + ie + 1), rb(rv.begin(), rv.begin() + ie + 1);
      mu = (mu >= 0.) ? mu : compute_mode(ra);
      const double dSi = compute_mode(rb);
      sigma = (sigma >= 0.) ? sigma : std::sqrt(dSi);
      break;
    }
    case 2: {
      mu = (mu >= 0.) ? mu : compute_mode(rm);
      const double dSi = compute_mode(rv);
      sigma = (sigma >= 0.) ? sigma : std::sqrt(std::max(dSi, std::max(b + a * dSi, 0.)));
      break;
    }
    case 3: {
      mu = (mu >= 0.) ? mu : compute_mode(rm);
      sigma = (sigma >= 0.) ? sigma : std::sqrt(std::fabs(b + a * mu));
      break;
    }
    default: {
      mu = (mu >= 0.) ? mu : rm[0];
      sigma = (sigma >= 0.) ? sigma : std::sqrt(std::fabs(b + a * mu));
      break;
    }
  }
}

void estimate_noise(const double* u, uint32_t n, uint32_t nt, uint32_t method, double& alpha, double& mu,
                    double& sigma) {
  std::vector<std::shared_ptr<const FrameLeaves>> frames;
  for (uint32_t k = 0; k < nt; ++k)
    frames.push_back(std::make_shared<const FrameLeaves>(frame_leaves(u + (size_t)k * n * n, n)));
  estimate_from_leaves(frames, n, method, alpha, mu, sigma);
}

}  // namespace pgsvt
