// pgsvt_oracle.cpp -- CPU FP64 restatement of the reference PGURE-SVT path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker the GPU library is
// compared against and the CPU baseline that bench.py times.  Nothing in the
// product (pgure-svt_amd/) links, loads or calls it.
//
// It restates, in plain C++ without Armadillo or NLopt, the behaviour of
//   src/pguresvt.hpp  (driver, window selection, normalisation)
//   src/arps.hpp      (ARPS block matching)
//   src/svt.hpp       (patch selection, Casorati gather, SVD, threshold,
//                      overlap-add)
//   src/pgure.hpp     (perturbations, PGURE risk, lambda optimiser)
//   src/noise.hpp     (Poisson-Gaussian noise estimator)
//   src/utils.hpp     (SoftThreshold, the std::thread window scheduler)
//   src/pcg/pcg_random.hpp (pcg64 = setseq_xsl_rr_128_64)
// including the quirks that decide numerical parity:
//   - mu and sigma swapped at the PGURE constructor call (pguresvt.hpp:133)
//   - delta1 is an integer cube, so `delta1 * eps1` truncates eps1 (=1e-4)
//     to the integer 0 and U1 == U (pgure.hpp:80, 156)
//   - with motion estimation off only the reference frame holds real patch
//     positions; every other frame uses (0,0) (arps.hpp:42,60-64)
//   - patch index set i*nyMbs+j for the overlap grid (svt.hpp:61-92)
//   - the optimiser returns the last evaluated lambda (pgure.hpp:128,236)
//   - the quadtree revisits the previous last node and never sheds node
//     dele(0) (noise.hpp:66-72, 433-455)
// Summation order follows Armadillo's accu (two interleaved partial sums,
// even and odd indices), medians follow op_median (nth_element, upper middle
// + (lower middle - upper middle)/2 for even counts).
//
// The lambda optimiser is a restatement of NLopt 2.6.2 LN_SBPLX for one
// variable (Rowan's subplex with NLopt's nldrmd Nelder-Mead inner solver),
// written from the published algorithm; NLopt is absent here, so its iterate
// sequence against the real library is unpinned.
//
// SVD: FP64 one-sided Jacobi (default) or LAPACK dgesdd loaded at run time
// (the Armadillo svd_econ "dc" path), selectable with oracle_set_svd().

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <random>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>
#include <dlfcn.h>

#include "../include/pgsvt.h"

namespace orc {

// ---------------------------------------------------------------- helpers --

// Armadillo accu / arrayops::accumulate order: even indices into acc1, odd
// into acc2, trailing odd element into acc1, result acc1 + acc2.
template <class F>
static inline double accu2f(size_t n, F f) {
  double a1 = 0.0, a2 = 0.0;
  size_t i = 0;
  for (; i + 1 < n; i += 2) {
    a1 += f(i);
    a2 += f(i + 1);
  }
  if (i < n) a1 += f(i);
  return a1 + a2;
}
static inline double accu2(const double* a, size_t n) {
  return accu2f(n, [a](size_t i) { return a[i]; });
}

// arma::median of a vector (op_median::direct_median + op_mean::robust_mean).
static double arma_median(std::vector<double> v) {
  const size_t n = v.size(), half = n / 2;
  std::nth_element(v.begin(), v.begin() + half, v.end());
  if (n % 2 == 0) {
    const double v1 = v[half];
    const double v2 = *std::max_element(v.begin(), v.begin() + half);
    return v1 + (v2 - v1) / 2.0;
  }
  return v[half];
}

// ------------------------------------------------------------------ pcg64 --
// setseq_xsl_rr_128_64 with the default stream (src/pcg/pcg_random.hpp:
// 166-169 constants, 501-503 seeding, 445-451 generation, 1085-1113 output).
typedef unsigned __int128 u128;
static const u128 kPcgMult =
    ((u128)2549297995355413924ULL << 64) | (u128)4865540595714422341ULL;
static const u128 kPcgInc =
    ((u128)6364136223846793005ULL << 64) | (u128)1442695040888963407ULL;

struct Pcg64 {
  typedef uint64_t result_type;
  u128 state;
  explicit Pcg64(uint64_t seed) { state = ((u128)seed + kPcgInc) * kPcgMult + kPcgInc; }
  static constexpr uint64_t min() { return 0; }
  static constexpr uint64_t max() { return ~(uint64_t)0; }
  uint64_t operator()() {
    state = state * kPcgMult + kPcgInc;  // output_previous == false for 128 bit
    const unsigned rot = (unsigned)(state >> 122) & 63u;
    const uint64_t x = (uint64_t)(state ^ (state >> 64));
    return rot ? (x >> rot) | (x << (64 - rot)) : x;
  }
};

// ------------------------------------------------------------------ errors --
static thread_local std::string g_err;
static int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

// ------------------------------------------------------------ SVD backends --
typedef void (*dgesdd_t)(const char*, const int*, const int*, double*, const int*,
                         double*, double*, const int*, double*, const int*, double*,
                         const int*, int*, int*);
static std::atomic<int> g_svd_backend{0};  // 0 jacobi, 1 lapack
static dgesdd_t g_dgesdd = nullptr;
static std::string g_lapack_name = "none";

static bool load_lapack() {
  if (g_dgesdd) return true;
  const char* libs[][2] = {
      {"/opt/conda/lib/libmkl_rt.so.1", "dgesdd_"},
      {"libmkl_rt.so.1", "dgesdd_"},
      {"/usr/local/lib/python3.10/dist-packages/scipy.libs/libscipy_openblas-68440149.so",
       "scipy_dgesdd_"},
      {"liblapack.so.3", "dgesdd_"},
  };
  for (auto& l : libs) {
    void* h = dlopen(l[0], RTLD_NOW | RTLD_LOCAL);
    if (!h) continue;
    void* s = dlsym(h, l[1]);
    if (s) {
      g_dgesdd = (dgesdd_t)s;
      g_lapack_name = std::string(l[0]) + ":" + l[1];
      if (std::string(l[0]).find("mkl") != std::string::npos) {
        typedef void (*setn_t)(int);
        if (auto f = (setn_t)dlsym(h, "MKL_Set_Num_Threads")) f(1);
      }
      return true;
    }
  }
  return false;
}

// Thin SVD of the m x n column-major matrix A (m >= n): U m x n, s n, V n x n.
// Only U diag(g(s)) V^T is ever consumed, so signs / order of degenerate
// subspaces do not matter.
static void svd_jacobi(int m, int n, const double* A, double* U, double* s, double* V) {
  std::vector<double> X(A, A + (size_t)m * n);
  for (int i = 0; i < n * n; ++i) V[i] = 0.0;
  for (int i = 0; i < n; ++i) V[i + n * i] = 1.0;
  for (int sweep = 0; sweep < 80; ++sweep) {
    bool rotated = false;
    for (int i = 0; i < n - 1; ++i) {
      for (int j = i + 1; j < n; ++j) {
        double a = 0, b = 0, c = 0;
        double* xi = &X[(size_t)m * i];
        double* xj = &X[(size_t)m * j];
        for (int r = 0; r < m; ++r) {
          a += xi[r] * xi[r];
          b += xj[r] * xj[r];
          c += xi[r] * xj[r];
        }
        if (std::fabs(c) <= 1e-15 * std::sqrt(a * b) || c == 0.0) continue;
        rotated = true;
        const double zeta = (b - a) / (2.0 * c);
        const double t = std::copysign(1.0, zeta) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
        const double cs = 1.0 / std::sqrt(1.0 + t * t), sn = cs * t;
        for (int r = 0; r < m; ++r) {
          const double p = xi[r], q = xj[r];
          xi[r] = cs * p - sn * q;
          xj[r] = sn * p + cs * q;
        }
        double* vi = &V[(size_t)n * i];
        double* vj = &V[(size_t)n * j];
        for (int r = 0; r < n; ++r) {
          const double p = vi[r], q = vj[r];
          vi[r] = cs * p - sn * q;
          vj[r] = sn * p + cs * q;
        }
      }
    }
    if (!rotated) break;
  }
  for (int i = 0; i < n; ++i) {
    double a = 0;
    for (int r = 0; r < m; ++r) a += X[(size_t)m * i + r] * X[(size_t)m * i + r];
    s[i] = std::sqrt(a);
    for (int r = 0; r < m; ++r) U[(size_t)m * i + r] = s[i] > 0 ? X[(size_t)m * i + r] / s[i] : 0.0;
  }
  // descending order, as LAPACK returns it
  std::vector<int> ord(n);
  for (int i = 0; i < n; ++i) ord[i] = i;
  std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return s[x] > s[y]; });
  std::vector<double> U2(U, U + (size_t)m * n), V2(V, V + (size_t)n * n), s2(s, s + n);
  for (int k = 0; k < n; ++k) {
    s[k] = s2[ord[k]];
    std::memcpy(U + (size_t)m * k, &U2[(size_t)m * ord[k]], sizeof(double) * m);
    std::memcpy(V + (size_t)n * k, &V2[(size_t)n * ord[k]], sizeof(double) * n);
  }
}

static bool svd_lapack(int m, int n, const double* A, double* U, double* s, double* V) {
  std::vector<double> a(A, A + (size_t)m * n), vt((size_t)n * n);
  const char jobz = 'S';
  int lda = m, ldu = m, ldvt = n, info = 0, lwork = -1;
  std::vector<int> iwork(8 * n);
  double wq = 0;
  g_dgesdd(&jobz, &m, &n, a.data(), &lda, s, U, &ldu, vt.data(), &ldvt, &wq, &lwork, iwork.data(), &info);
  lwork = (int)wq + 1;
  std::vector<double> work(lwork);
  g_dgesdd(&jobz, &m, &n, a.data(), &lda, s, U, &ldu, vt.data(), &ldvt, work.data(), &lwork, iwork.data(), &info);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) V[i + n * j] = vt[j + n * i];
  return info == 0;
}

static void svd_econ(int m, int n, const double* A, double* U, double* s, double* V) {
  if (g_svd_backend.load() == 1 && g_dgesdd) {
    if (svd_lapack(m, n, A, U, s, V)) return;
  }
  svd_jacobi(m, n, A, U, s, V);
}

// ---------------------------------------------------------- cube helpers --
struct Cube {
  uint32_t nr = 0, nc = 0, ns = 0;
  std::vector<double> d;
  Cube() {}
  Cube(uint32_t r, uint32_t c, uint32_t s) : nr(r), nc(c), ns(s), d((size_t)r * c * s, 0.0) {}
  double& at(uint32_t r, uint32_t c, uint32_t s) { return d[r + (size_t)nr * (c + (size_t)nc * s)]; }
  double at(uint32_t r, uint32_t c, uint32_t s) const { return d[r + (size_t)nr * (c + (size_t)nc * s)]; }
  size_t n() const { return d.size(); }
};

// ----------------------------------------------------- noise estimator (A3) --
// src/noise.hpp:16-460
struct NoiseEstimator {
  uint32_t method, size = 8, weightType = 0;
  double lap[9];  // column-major 3x3 laplacian
  static constexpr uint32_t kDof[12] = {2, 4, 8, 16, 32, 64, 128, 256, 512, 1024, 2048, 4096};
  static constexpr double kF0025[12] = {15.4392, 2.86209, 1.64602, 1.27893, 1.13046, 1.06318,
                                        1.03110, 1.01543, 1.00769, 1.00384, 1.00192, 1.00096};
  std::vector<std::array<uint32_t, 3>> tree;
  std::vector<uint32_t> del;

  explicit NoiseEstimator(uint32_t m) : method(m) {
    for (int i = 0; i < 9; ++i) lap[i] = 1.0 * 0.125;
    lap[4] = -1;
  }

  // noise.hpp:182-221; A is the s x s block at (i, j) of frame F (nr rows)
  bool SplitBlockQ(const double* F, uint32_t nr, uint32_t i0, uint32_t j0, uint32_t N) {
    if (N <= size) return false;
    const uint32_t l = 5;
    std::vector<double> A((size_t)N * N), res((size_t)N * N);
    for (uint32_t x = 0; x < N; ++x)
      for (uint32_t y = 0; y < N; ++y) A[y + (size_t)N * x] = F[(i0 + y) + (size_t)nr * (j0 + x)];
    for (uint32_t x = 0; x < N; ++x) {
      for (uint32_t y = 0; y < N; ++y) {
        const uint32_t xp = (x + 1 == N) ? 1 : x + 1;
        const uint32_t yp = (y + 1 == N) ? 1 : y + 1;
        const uint32_t xm = (x == 0) ? N - 2 : x - 1;
        const uint32_t ym = (y == 0) ? N - 2 : y - 1;
        res[y + (size_t)N * x] = l * A[y + (size_t)N * x] -
                                 (A[yp + (size_t)N * x] + A[ym + (size_t)N * x] +
                                  A[y + (size_t)N * xm] + A[y + (size_t)N * xp]);
      }
    }
    const double sq = std::sqrt((double)(l * l + l));
    for (auto& r : res) r /= sq;
    const uint32_t R = N * N;
    const double OoR = 1.0 / R, OoRm1 = 1.0 / (R - 1);
    const double accuZ = accu2(A.data(), R) * OoR;
    const double Sz = accu2f(R, [&](size_t k) { double t = A[k] - accuZ; return t * t; }) * OoRm1;
    const double accuR = accu2(res.data(), R) * OoR;
    const double Se = accu2f(R, [&](size_t k) { double t = res[k] - accuR; return t * t; }) * OoRm1;
    const double stat = (Sz > Se) ? Sz / Se : Se / Sz;
    for (int k = 0; k < 12; ++k)
      if (kDof[k] == N) return stat > kF0025[k];
    return false;  // not reachable for power-of-two frames <= 4096
  }

  // noise.hpp:419-458 (including the revisit of the previous last node)
  void QuadTree(const double* F, uint32_t nr, uint32_t part) {
    const auto node = tree[part];
    const uint32_t i = node[0], j = node[1];
    uint32_t s = node[2];
    if (!SplitBlockQ(F, nr, i, j, s)) return;
    s /= 2;
    const uint32_t n = (uint32_t)tree.size() - 1;
    tree.push_back({i, j, s});
    tree.push_back({i + s, j, s});
    tree.push_back({i, j + s, s});
    tree.push_back({i + s, j + s, s});
    del.push_back(part);
    uint32_t it = n;
    do {
      QuadTree(F, nr, it);
      it++;
    } while (it < n + 4);
  }

  static double IQD(const std::vector<double>& a) {  // noise.hpp:223-230
    std::vector<double> s(a);
    std::sort(s.begin(), s.end());
    const uint32_t N = (uint32_t)s.size();
    const uint32_t m = (uint32_t)std::floor((std::floor((N + 1) / 2) + 1) / 2);
    return s[N - m - 1] - s[m - 1];
  }

  static double RobustVar(const std::vector<double>& a) {  // noise.hpp:232-236
    const double med = arma_median(a);
    std::vector<double> d(a.size());
    for (size_t k = 0; k < a.size(); ++k) d[k] = std::fabs(a[k] - med);
    const double sig = 1.4826 * arma_median(d);
    return sig * sig;
  }

  void Weight(const std::vector<double>& x, std::vector<double>& w) const {  // noise.hpp:303-326
    if (weightType == 0) {
      const double p = 0.75;
      for (size_t k = 0; k < x.size(); ++k) w[k] = (std::fabs(x[k]) < p) ? 1. : p / std::fabs(x[k]);
    } else {
      const double p = 3.5, pp = 12.25;
      for (size_t k = 0; k < x.size(); ++k)
        w[k] = (std::fabs(x[k]) > p) ? 0. : (pp - x[k] * x[k]) * (pp - x[k] * x[k]) / (pp * pp);
    }
  }

  double RobustMean(const std::vector<double>& A) {  // noise.hpp:238-271
    const uint32_t I = 10000;
    const size_t N = A.size();
    double e = 0, tol = 1e-6, d = 1, m = 0, m0 = 1e12, eps = 1e-12, aux = 0;
    std::vector<double> w(N, 1.0), r(N);
    for (uint32_t it = 0; it < I; ++it) {
      for (size_t k = 0; k < N; ++k) r[k] = w[k] * A[k];
      m = accu2(r.data(), N);
      aux = accu2(w.data(), N);
      m = (std::fabs(aux) < eps) ? m0 : m / aux;
      for (size_t k = 0; k < N; ++k) r[k] = A[k] - m;
      e = accu2f(N, [&](size_t k) { return std::fabs(r[k]); }) / (double)N;
      if (std::fabs(m0 - m) < tol || e < tol) break;
      m0 = m;
      d = IQD(r) + eps;
      const double inv = 1. / d;
      for (size_t k = 0; k < N; ++k) r[k] *= inv;
      Weight(r, w);
    }
    return m;
  }

  double ComputeMode(const std::vector<double>& A) {  // noise.hpp:273-301
    uint32_t maxCount = 0;
    double maxValue = 0.;
    double M = *std::max_element(A.begin(), A.end());
    const size_t N = A.size();
    const double dyn = 1. * N;
    std::vector<double> a(N);
    for (size_t k = 0; k < N; ++k) a[k] = std::round(A[k] * dyn / M);
    for (size_t i = 0; i < N; ++i) {
      uint32_t count = 0;
      for (size_t j = 0; j < N; ++j)
        if (a[j] == a[i]) count++;
      if (count > maxCount) {
        maxCount = count;
        maxValue = a[i];
      }
    }
    return maxValue * (M / dyn);
  }

  void WLSFit(const std::vector<double>& x, const std::vector<double>& y, double& pa, double& pb) {  // noise.hpp:328-383
    const uint32_t I = 10000;
    const size_t N = x.size();
    double e = 0, tol = 1e-6, d = 1, a0 = 1e12, b0 = 1e12, eps = 1e-12, aux = 0;
    std::vector<double> w(N, 1.0), w2(N), xy(N), r(N);
    double p0 = 0, p1 = 0;
    for (uint32_t it = 0; it < I; ++it) {
      for (size_t k = 0; k < N; ++k) w2[k] = w[k] * w[k];
      const double sw2 = accu2(w2.data(), N);
      const double sw2x = accu2f(N, [&](size_t k) { return w2[k] * x[k]; });
      const double sw2y = accu2f(N, [&](size_t k) { return w2[k] * y[k]; });
      for (size_t k = 0; k < N; ++k) xy[k] = w2[k] * (x[k] * y[k]);
      p0 = sw2 * accu2(xy.data(), N) - sw2x * sw2y;
      for (size_t k = 0; k < N; ++k) xy[k] = w2[k] * (x[k] * x[k]);
      aux = sw2 * accu2(xy.data(), N) - sw2x * sw2x;
      p0 = (std::fabs(aux) < eps) ? a0 : p0 / aux;
      p1 = sw2y - p0 * sw2x;
      p1 = (std::fabs(aux) < eps) ? b0 : p1 / sw2;
      for (size_t k = 0; k < N; ++k) r[k] = y[k] - (x[k] * p0 + p1);
      e = accu2f(N, [&](size_t k) { return std::fabs(r[k]); }) / (double)N;
      if ((std::fabs(a0 - p0) < tol && std::fabs(b0 - p1) < tol) || e < tol) break;
      a0 = p0;
      b0 = p1;
      d = IQD(r) + eps;
      for (size_t k = 0; k < N; ++k) r[k] /= d;
      Weight(r, w);
    }
    pa = p0;
    pb = p1;
  }

  // noise.hpp:395-417 (the output is written transposed, as in the reference)
  static std::vector<double> ConvolveFIR(const std::vector<double>& in, uint32_t N, const double* lap) {
    std::vector<double> out((size_t)N * N, 0.0);
    auto I = [&](uint32_t r, uint32_t c) { return in[r + (size_t)N * c]; };
    double nl[9];
    for (int k = 0; k < 9; ++k) nl[k] = lap[k] * -1.0;
    for (uint32_t x = 0; x < N; ++x) {
      for (uint32_t y = 0; y < N; ++y) {
        const uint32_t xp = (x + 1 == N) ? 1 : x + 1;
        const uint32_t yp = (y + 1 == N) ? 1 : y + 1;
        const uint32_t xm = (x == 0) ? N - 2 : x - 1;
        const uint32_t ym = (y == 0) ? N - 2 : y - 1;
        // neighbours(row, col), column-major linear order
        const double nb[9] = {I(xm, ym), I(xm, y), I(xm, yp), I(x, ym), I(x, y),
                              I(x, yp),  I(xp, ym), I(xp, y), I(xp, yp)};
        out[y + (size_t)N * x] = accu2f(9, [&](size_t k) { return nb[k] * nl[k]; });
      }
    }
    return out;
  }

  // noise.hpp:35-153
  void Estimate(const Cube& u, double& alpha_, double& mu_, double& sigma_) {
    double alpha = alpha_, mu = mu_, sigma = sigma_, dSi = 0.0;
    const uint32_t Nx = u.nc, Ny = u.nr, T = u.ns;
    const size_t maxV = (size_t)Nx * Ny;
    std::vector<double> means(T * maxV, -1.0), vars(T * maxV, -1.0);
    for (uint32_t i = 0; i < T; ++i) {
      const double* F = &u.d[(size_t)u.nr * u.nc * i];
      tree.assign(1, {0u, 0u, Nx});
      del.clear();
      QuadTree(F, u.nr, 0);
      std::vector<std::array<uint32_t, 3>> t2 = tree;
      std::vector<uint32_t> dl = del;
      std::sort(dl.begin(), dl.end());
      dl.erase(std::unique(dl.begin(), dl.end()), dl.end());
      if (!dl.empty())
        for (size_t k = dl.size() - 1; k > 0; --k) t2.erase(t2.begin() + dl[k]);
      for (size_t n = 0; n < t2.size(); ++n) {
        const uint32_t x = t2[n][0], y = t2[n][1], s = t2[n][2];
        std::vector<double> col((size_t)s * s);
        for (uint32_t c = 0; c < s; ++c)
          for (uint32_t r = 0; r < s; ++r) col[r + (size_t)s * c] = F[(x + r) + (size_t)u.nr * (y + c)];
        const double meanEst = RobustMean(col);
        std::vector<double> conv = ConvolveFIR(col, s, lap);
        means[i * maxV + n] = meanEst;
        vars[i * maxV + n] = RobustVar(conv);
      }
    }
    std::vector<double> m2, v2;
    for (double v : means)
      if (v >= 0.) m2.push_back(v);
    for (double v : vars)
      if (v >= 0.) v2.push_back(v);
    std::vector<size_t> idx(m2.size());
    for (size_t k = 0; k < idx.size(); ++k) idx[k] = k;
    std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) { return m2[a] < m2[b]; });
    std::vector<double> rm(idx.size()), rv(idx.size());
    for (size_t k = 0; k < idx.size(); ++k) {
      rm[k] = m2[idx[k]];
      rv[k] = v2[idx[k]];
    }
    double a = 0, b = 0;
    WLSFit(rm, rv, a, b);
    alpha = (alpha >= 0.) ? alpha : a;
    switch (method) {
      case 1: {
        const int L = (int)std::floor(1. * (Nx * Ny / m2.size()));
        const uint32_t ie = (uint32_t)std::round(0.05 * L);
        // RestrictArray(a, 0, ie) is bounds-checked in the reference and
        // aborts when the window has fewer leaves (noise.hpp:384-392)
        if ((size_t)ie + 1 > rm.size()) throw std::runtime_error("noise method 1: too few quadtree leaves");
        std::vector<double> ra(rm.begin(), rm.begin() + ie + 1), rb(rv.begin(), rv.begin() + ie + 1);
        mu = (mu >= 0.) ? mu : ComputeMode(ra);
        dSi = ComputeMode(rb);
        sigma = (sigma >= 0.) ? sigma : std::sqrt(dSi);
        break;
      }
      case 2: {
        mu = (mu >= 0.) ? mu : ComputeMode(rm);
        dSi = ComputeMode(rv);
        sigma = (sigma >= 0.) ? sigma : std::sqrt(std::max(dSi, std::max(b + a * dSi, 0.)));
        break;
      }
      case 3: {
        mu = (mu >= 0.) ? mu : ComputeMode(rm);
        sigma = (sigma >= 0.) ? sigma : std::sqrt(std::fabs(b + a * mu));
        break;
      }
      default: {
        mu = (mu >= 0.) ? mu : rm[0];
        sigma = (sigma >= 0.) ? sigma : std::sqrt(std::fabs(b + a * mu));
        break;
      }
    }
    alpha_ = alpha;
    mu_ = mu;
    sigma_ = sigma;
  }
};
constexpr uint32_t NoiseEstimator::kDof[12];
constexpr double NoiseEstimator::kF0025[12];

// ------------------------------------------------------ ARPS (A4, A5) ------
// src/arps.hpp:17-376.  patches is int [Nw][P_all][2] (the icube(2,P,Nw)).
struct Motion {
  const Cube& A;
  uint32_t bs, timeIter, timeWindow, mw, nImages;
  uint32_t Nx, Ny, Nt, nxMbs, nyMbs, vecSize;
  double OoBs2;
  std::vector<int32_t> patches, motions;  // [slot][it][2]

  Motion(const Cube& A_, uint32_t bs_, uint32_t t, uint32_t tw, uint32_t mw_, uint32_t ni)
      : A(A_), bs(bs_), timeIter(t), timeWindow(tw), mw(mw_), nImages(ni) {
    Nx = A.nr;
    Ny = A.nc;
    Nt = A.ns;
    nxMbs = Nx - bs;
    nyMbs = Ny - bs;
    OoBs2 = 1.0 / (bs * bs);
    vecSize = (1 + nxMbs) * (1 + nyMbs);
    patches.assign((size_t)2 * vecSize * (2 * tw + 1), 0);
    motions.assign((size_t)2 * vecSize * (2 * tw), 0);
  }
  int32_t& P(int c, size_t it, size_t k) { return patches[c + 2 * (it + (size_t)vecSize * k)]; }
  int32_t& M(int c, size_t it, size_t k) { return motions[c + 2 * (it + (size_t)vecSize * k)]; }

  double Cost(const double* ref, int r0, int c0, int frame) const {
    // accu(square(ref - cand)) * (1/bs^2); block linear order is column-major
    const uint32_t n = bs * bs;
    return accu2f(n, [&](size_t q) {
             const uint32_t a = (uint32_t)q % bs, b = (uint32_t)q / bs;
             const double d = ref[q] - A.at(r0 + a, c0 + b, frame);
             return d * d;
           }) *
           OoBs2;
  }

  void ARPS(int curFrame, int i1, int i2, int i3) {
    (void)curFrame;  // only used by the disabled predictive cost (pMotion < 0)
    const int W = 2 * (int)mw + 1;
    std::vector<uint8_t> check((size_t)W * W);
    std::vector<double> ref(bs * bs);
    for (size_t it = 0; it < vecSize; ++it) {
      double costs[6];
      for (double& c : costs) c = 1e9;
      std::fill(check.begin(), check.end(), 0);
      int LDSP[6][2] = {{0, -1}, {-1, 0}, {0, 0}, {1, 0}, {0, 1}, {0, 0}};
      const int SDSP[5][2] = {{0, -1}, {-1, 0}, {0, 0}, {1, 0}, {0, 1}};
      const int i = (int)(it % (1 + nxMbs));
      const int j = (int)(it / (1 + nyMbs));
      int x = j, y = i;
      for (uint32_t q = 0; q < bs * bs; ++q) ref[q] = A.at(i + q % bs, j + q / bs, i1);
      costs[2] = Cost(ref.data(), i, j, i2);
      check[mw + (size_t)W * mw] = 1;  // checkMat(mw, mw)
      uint32_t maxIdx;
      int step;
      if (j == 0) {
        step = 2;
        maxIdx = 5;
      } else {
        const int yT = std::abs(M(0, it, i3)), xT = std::abs(M(1, it, i3));
        step = (xT <= yT) ? yT : xT;
        if ((yT == 0 && xT == step) || (xT == 0 && yT == step)) {
          maxIdx = 5;
        } else {
          maxIdx = 6;
          LDSP[5][0] = M(1, it, i3);
          LDSP[5][1] = M(0, it, i3);
        }
      }
      LDSP[0][0] = 0; LDSP[0][1] = -step;
      LDSP[1][0] = -step; LDSP[1][1] = 0;
      LDSP[2][0] = 0; LDSP[2][1] = 0;
      LDSP[3][0] = step; LDSP[3][1] = 0;
      LDSP[4][0] = 0; LDSP[4][1] = step;
      for (uint32_t k = 0; k < maxIdx; ++k) {
        const int rv = y + LDSP[k][1], rh = x + LDSP[k][0];
        const bool skip = (k == 2) || (step == 0) || rh < 0 || rv < 0 || (rh + (int)bs - 1) >= (int)Ny ||
                          (rv + (int)bs - 1) >= (int)Nx;
        if (!skip) {
          costs[k] = Cost(ref.data(), rv, rh, i2);
          check[(LDSP[k][1] + mw) + (size_t)W * (LDSP[k][0] + mw)] = 1;
        }
      }
      int pt = 0;
      for (int k = 1; k < 6; ++k)
        if (costs[k] < costs[pt]) pt = k;
      x += LDSP[pt][0];
      y += LDSP[pt][1];
      double cost = costs[pt];
      for (double& c : costs) c = 1e9;
      costs[2] = cost;
      uint32_t nSDSP = 0;
      bool done = false;
      do {
        for (int k = 0; k < 5; ++k) {
          const int rv = y + SDSP[k][1], rh = x + SDSP[k][0];
          bool skip = (k == 2) || rh < 0 || rv < 0 || (rh + (int)bs - 1) >= (int)Ny ||
                      (rv + (int)bs - 1) >= (int)Nx || rh < (j - (int)mw) || rh > (j + (int)mw) ||
                      rv < (i - (int)mw) || rv > (i + (int)mw);
          if (!skip) skip = check[(y - i + SDSP[k][1] + mw) + (size_t)W * (x - j + SDSP[k][0] + mw)] == 1;
          if (!skip) {
            costs[k] = Cost(ref.data(), rv, rh, i2);
            check[(y - i + SDSP[k][1] + mw) + (size_t)W * (x - j + SDSP[k][0] + mw)] = 1;
          }
        }
        pt = 0;
        for (int k = 1; k < 6; ++k)
          if (costs[k] < costs[pt]) pt = k;
        cost = costs[pt];
        if (pt == 2 || nSDSP >= 1000000u) {
          done = true;
        } else {
          x += SDSP[pt][0];
          y += SDSP[pt][1];
          for (double& c : costs) c = 1e9;
          costs[2] = cost;
        }
        nSDSP++;
      } while (!done);
      M(0, it, i3) = y - i;
      M(1, it, i3) = x - j;
      P(0, it, i2) = y;
      P(1, it, i2) = x;
    }
  }

  void Estimate(bool on) {  // arps.hpp:52-134
    auto fill_ref = [&](uint32_t k) {
      for (size_t i = 0; i < vecSize; ++i) {
        P(0, i, k) = (int32_t)(i % (1 + nyMbs));
        P(1, i, k) = (int32_t)(i / (1 + nxMbs));
      }
    };
    if (timeIter < timeWindow) {
      const uint32_t loopEnd = Nt - timeIter - 1;
      fill_ref(timeIter);
      if (on) {
        for (uint32_t i = 0; i < loopEnd; ++i) ARPS(i, timeIter + i, timeIter + i + 1, timeIter + i);
        for (uint32_t i = 0; i < timeIter; ++i) {
          const int n = -1 * (int)(i + 1);
          ARPS(n, timeIter + n + 1, timeIter + n, timeIter + n + 1);
        }
      }
    } else if (timeIter >= nImages - timeWindow) {
      const uint32_t endFrame = timeIter - (nImages - Nt);
      const uint32_t loopEnd = 2 * timeWindow - endFrame;
      fill_ref(endFrame);
      if (on) {
        for (uint32_t i = 0; i < loopEnd; ++i) ARPS(i, endFrame + i, endFrame + i + 1, endFrame + i);
        for (uint32_t i = 0; i < endFrame; ++i) {
          const int n = -1 * (int)(i + 1);
          if (2 * timeWindow == endFrame)
            ARPS(n, endFrame + n + 1, endFrame + n, endFrame + n);
          else
            ARPS(n, endFrame + n + 1, endFrame + n, endFrame + n + 1);
        }
      }
    } else {
      fill_ref(timeWindow);
      if (on) {
        for (uint32_t i = 0; i < timeWindow; ++i) ARPS(i, timeWindow + i, timeWindow + i + 1, timeWindow + i);
        for (uint32_t i = 0; i < timeWindow; ++i) {
          const int n = -1 * (int)(i + 1);
          ARPS(n, timeWindow + n + 1, timeWindow + n, timeWindow + n + 1);
        }
      }
    }
  }
};

// -------------------------------------------------- SVT (A6-A10) ----------
// src/svt.hpp:20-183
static std::vector<uint32_t> patch_indices(uint32_t Nx, uint32_t Ny, uint32_t bs, uint32_t ov) {
  const uint32_t nxMbs = Nx - bs, nyMbs = Ny - bs, nxD = nxMbs / ov, nyD = nyMbs / ov;
  std::vector<uint32_t> j;
  for (size_t i = 0; i < 1 + nyMbs; i += ov)
    for (size_t k = 0; k < 1 + nxMbs; k += ov) j.push_back((uint32_t)(i * nyMbs + k));
  std::vector<uint32_t> bottom(1 + nyD), right(1 + nxD);
  for (size_t i = 0; i < 1 + nyMbs; i += ov) bottom[i / ov] = (uint32_t)((nyMbs + 1) * i + nxMbs);
  for (size_t i = 0; i < 1 + nxMbs; i += ov) right[i / ov] = (uint32_t)((nyMbs + 1) * nxMbs + i);
  j.insert(j.end(), right.begin(), right.end());
  j.insert(j.end(), bottom.begin(), bottom.end());
  std::sort(j.begin(), j.end());
  j.erase(std::unique(j.begin(), j.end()), j.end());
  return j;
}

struct SVT {
  const std::vector<int32_t>& patches;  // [Nw][P_all][2]
  uint32_t Nx, Ny, Nt, bs, ov, P_all;
  bool expW;
  std::vector<uint32_t> actual;
  uint32_t m, n;
  std::vector<double> U, S, V;  // per patch: U m x n, S n, V n x n

  SVT(const std::vector<int32_t>& p, uint32_t nx, uint32_t ny, uint32_t nt, uint32_t b, uint32_t o, bool e)
      : patches(p), Nx(nx), Ny(ny), Nt(nt), bs(b), ov(o), expW(e) {
    P_all = (nx - b + 1) * (ny - b + 1);
    m = b * b;
    n = nt;
  }
  int pos(int c, uint32_t idx, uint32_t k) const { return patches[c + 2 * (idx + (size_t)P_all * k)]; }

  void Decompose(const Cube& u) {  // svt.hpp:58-118
    actual = patch_indices(Nx, Ny, bs, ov);
    const size_t P = actual.size();
    U.assign(P * m * n, 0.0);
    S.assign(P * n, 0.0);
    V.assign(P * n * n, 0.0);
    std::vector<double> block((size_t)m * n);
    for (size_t it = 0; it < P; ++it) {
      for (uint32_t k = 0; k < n; ++k) {
        const int r0 = pos(0, actual[it], k), c0 = pos(1, actual[it], k);
        for (uint32_t q = 0; q < m; ++q) block[q + (size_t)m * k] = u.at(r0 + q % bs, c0 + q / bs, k);
      }
      svd_econ((int)m, (int)n, block.data(), &U[it * m * n], &S[it * n], &V[it * n * n]);
    }
  }

  Cube Reconstruct(double lambda) const {  // svt.hpp:121-167
    Cube v(Nx, Ny, Nt), w(Nx, Ny, Nt);
    const size_t P = actual.size();
    std::vector<double> st(n), B((size_t)m * n);
    for (size_t it = 0; it < P; ++it) {
      const double* Ub = &U[it * m * n];
      const double* Sb = &S[it * n];
      const double* Vb = &V[it * n * n];
      if (expW) {
        double smax = Sb[0];
        for (uint32_t k = 1; k < n; ++k) smax = std::max(smax, Sb[k]);
        const double kk = -0.5 * lambda;
        for (uint32_t k = 0; k < n; ++k) {
          const double wv = std::fabs(smax * std::exp(kk * (Sb[k] * Sb[k])));
          const double sg = (Sb[k] > 0) - (Sb[k] < 0);
          st[k] = sg * std::max(std::fabs(Sb[k]) - wv, 0.0);
        }
      } else {
        for (uint32_t k = 0; k < n; ++k) {
          const double sg = (Sb[k] > 0) - (Sb[k] < 0);
          st[k] = sg * std::max(std::fabs(Sb[k]) - lambda, 0.0);
        }
      }
      for (uint32_t c = 0; c < n; ++c)
        for (uint32_t r = 0; r < m; ++r) {
          double acc = 0;
          for (uint32_t l = 0; l < n; ++l) acc += Ub[r + (size_t)m * l] * st[l] * Vb[c + (size_t)n * l];
          B[r + (size_t)m * c] = acc;
        }
      for (uint32_t k = 0; k < n; ++k) {
        const int r0 = pos(0, actual[it], k), c0 = pos(1, actual[it], k);
        for (uint32_t q = 0; q < m; ++q) {
          v.at(r0 + q % bs, c0 + q / bs, k) += B[q + (size_t)m * k];
          w.at(r0 + q % bs, c0 + q / bs, k) += 1.0;
        }
      }
    }
    for (size_t k = 0; k < v.n(); ++k) {
      v.d[k] /= w.d[k];
      if (!std::isfinite(v.d[k])) v.d[k] = 0.0;
    }
    return v;
  }
};

// --------------------------------------------- PGURE (A11-A13) -----------
// src/pgure.hpp:20-239
struct PGURE {
  const Cube& U;
  double alpha, mu, sigma, sigmasq, OoN, eps1 = 0, eps2 = 0;
  bool optimize;
  double lambda = 0;
  SVT svt0, svt1, svt2p, svt2m;
  std::vector<int64_t> delta1;
  std::vector<double> delta2;

  PGURE(const Cube& u, const std::vector<int32_t>& patches, double a, double m, double s, uint32_t bs, uint32_t ov,
        int64_t seed, bool expW, bool opt)
      : U(u), alpha(a), mu(m), sigma(s), optimize(opt),
        svt0(patches, u.nr, u.nc, u.ns, bs, ov, expW), svt1(patches, u.nr, u.nc, u.ns, bs, ov, expW),
        svt2p(patches, u.nr, u.nc, u.ns, bs, ov, expW), svt2m(patches, u.nr, u.nc, u.ns, bs, ov, expW) {
    OoN = 1.0 / (u.nr * u.nc * u.ns);
    sigmasq = sigma * sigma;
    uint64_t sd = seed >= 0 ? (uint64_t)seed : (uint64_t)std::random_device{}() << 32 ^ std::random_device{}();
    Pcg64 rng(sd);
    svt0.Decompose(U);
    if (optimize) {
      double umax = U.d[0];
      for (double v : U.d) umax = std::max(umax, v);
      eps1 = umax * 0.0001;
      eps2 = 100 * eps1;
      const size_t N = U.n();
      delta1.resize(N);
      delta2.resize(N);
      const double kappa = 1.;
      const double vP = 0.5 + 0.5 * kappa / std::sqrt(kappa * kappa + 4);
      const double vQ = 1 - vP;
      const double vQvP = std::sqrt(vQ / vP), vPvQ = std::sqrt(vP / vQ);
      std::bernoulli_distribution b1(0.5), b2(vP);
      for (size_t k = 0; k < N; ++k) delta1[k] = b1(rng) ? -1 : 1;
      for (size_t k = 0; k < N; ++k) delta2[k] = b2(rng) ? -1 * vQvP : vPvQ;
      // delta1 is an integer cube: the scalar eps1 is converted to the cube's
      // integer element type before the product (pgure.hpp:80).
      const int64_t eps1_int = (int64_t)eps1;
      Cube U1(U.nr, U.nc, U.ns), U2p(U.nr, U.nc, U.ns), U2m(U.nr, U.nc, U.ns);
      for (size_t k = 0; k < N; ++k) {
        U1.d[k] = U.d[k] + (double)(delta1[k] * eps1_int);
        U2p.d[k] = U.d[k] + delta2[k] * eps2;
        U2m.d[k] = U.d[k] - delta2[k] * eps2;
      }
      svt1.Decompose(U1);
      svt2p.Decompose(U2p);
      svt2m.Decompose(U2m);
    }
  }

  Cube Reconstruct(double l) const { return svt0.Reconstruct(l); }

  double Risk(double l) {  // pgure.hpp:120-137
    lambda = l;
    const Cube Uhat = svt0.Reconstruct(l), U1 = svt1.Reconstruct(l), U2p = svt2p.Reconstruct(l),
               U2m = svt2m.Reconstruct(l);
    const size_t N = U.n();
    const double* u = U.d.data();
    const double t1 = accu2f(N, [&](size_t k) { const double d = std::fabs(Uhat.d[k] - u[k]); return d * d; });
    const double t2 = accu2(u, N);
    const double t3 = accu2f(N, [&](size_t k) {
      return (double)delta1[k] * (alpha * u[k] - alpha * mu + sigmasq) * (U1.d[k] - Uhat.d[k]);
    });
    const double t4 = accu2f(N, [&](size_t k) { return delta2[k] * (U2p.d[k] - 2 * Uhat.d[k] + U2m.d[k]); });
    const double t5 = accu2(Uhat.d.data(), N);
    return OoN * (t1 - (alpha + mu) * t2 + (2 / eps1 * t3) - (2 * sigmasq * alpha / (eps2 * eps2) * t4) +
                  (2 * mu * t5) + mu) -
           sigmasq;
  }
};

// ------------------------------------- LN_SBPLX restatement (A14) --------
namespace sbplx {
enum { SUCCESS = 1, FTOL = 3, XTOL = 4, MAXEVAL = 5, FAILURE = -1, INVALID = -2 };
struct Stop {
  double ftol_rel = 0, ftol_abs = 0, xtol_rel = 0, xtol_abs = 0;
  int maxeval = 0, nevals = 0;
};
static bool relstop(double vold, double vnew, double reltol, double abstol) {
  if (std::isinf(vold)) return false;
  return std::fabs(vnew - vold) < abstol || std::fabs(vnew - vold) < reltol * (std::fabs(vnew) + std::fabs(vold)) * 0.5 ||
         (reltol > 0 && vnew == vold);
}
static bool close_(double a, double b) { return std::fabs(a - b) <= 1e-13 * (std::fabs(a) + std::fabs(b)); }
static bool reflect(double& xnew, double c, double scale, double xold, double lb, double ub) {
  double nx = c + scale * (c - xold);
  if (nx < lb) nx = lb;
  if (nx > ub) nx = ub;
  const bool eqc = close_(nx, c), eqo = close_(nx, xold);
  xnew = nx;
  return !(eqc || eqo);
}
typedef std::function<double(double)> Fn;

// Nelder-Mead in one dimension with the subplex psi stopping rule.
static int nldrmd(const Fn& f, double lb, double ub, double& x, double& minf, double xstep, Stop& st, double psi) {
  double pf[2], px[2];
  double init_diam = 0;
  px[0] = x;
  pf[0] = minf;
  px[1] = x + xstep;
  if (px[1] > ub) px[1] = (ub - x > std::fabs(xstep) * 0.1) ? ub : x - std::fabs(xstep);
  if (px[1] < lb) {
    if (x - lb > std::fabs(xstep) * 0.1) {
      px[1] = lb;
    } else {
      px[1] = x + std::fabs(xstep);
      if (px[1] > ub) px[1] = 0.5 * ((ub - x > x - lb ? ub : lb) + x);
    }
  }
  if (close_(px[1], x)) return FAILURE;
#define ORC_CHECK_EVAL(xc, fc)              \
  do {                                      \
    st.nevals++;                            \
    if ((fc) <= minf) {                     \
      minf = (fc);                          \
      x = (xc);                             \
    }                                       \
    if (st.maxeval > 0 && st.nevals >= st.maxeval) return MAXEVAL; \
  } while (0)
  pf[1] = f(px[1]);
  ORC_CHECK_EVAL(px[1], pf[1]);
  for (;;) {  // restart
    for (;;) {
      const int lo = (pf[1] < pf[0]) ? 1 : 0, hi = 1 - lo;
      const double fl = pf[lo], xl = px[lo];
      double fh = pf[hi];
      const double xh = px[hi];
      if (init_diam == 0) init_diam = std::fabs(xl - xh);
      const double c = xl * 1.0;
      if (std::fabs(xl - xh) < psi * init_diam) return XTOL;
      double xr;
      if (!reflect(xr, c, 1.0, xh, lb, ub)) return XTOL;
      const double fr = f(xr);
      ORC_CHECK_EVAL(xr, fr);
      if (fr < fl) {
        double xe;
        const bool ok = reflect(xe, c, 2.0, xh, lb, ub);
        px[hi] = xe;
        if (!ok) return XTOL;
        fh = f(xe);
        ORC_CHECK_EVAL(xe, fh);
        if (fh >= fr) {
          fh = fr;
          px[hi] = xr;
        }
      } else {
        double xc;
        if (!reflect(xc, c, fh <= fr ? -0.5 : 0.5, xh, lb, ub)) return XTOL;
        const double fc = f(xc);
        ORC_CHECK_EVAL(xc, fc);
        if (fc < fr && fc < fh) {
          px[hi] = xc;
          fh = fc;
        } else {
          px[hi] = xl + 0.5 * (px[hi] - xl);
          pf[hi] = f(px[hi]);
          ORC_CHECK_EVAL(px[hi], pf[hi]);
          break;  // goto restart
        }
      }
      pf[hi] = fh;
    }
  }
#undef ORC_CHECK_EVAL
}

// Subplex outer loop for one variable (NLopt sbplx_minimize, n = 1).
static int minimize(const Fn& f, double lb, double ub, double& x, double& minf, double xstep0, Stop& st) {
  const double psi = 0.25;
  if (lb > ub || x < lb || x > ub) return INVALID;
  if (xstep0 == 0) return INVALID;
  minf = f(x);
  st.nevals++;
  if (st.maxeval > 0 && st.nevals >= st.maxeval) return MAXEVAL;
  double xstep = xstep0, dx = 0;
  for (;;) {
    const double xprev = x, fprev = minf;
    double xs = x;
    int ret = nldrmd(f, lb, ub, xs, minf, xstep, st, psi);
    x = xs;
    if (ret == FAILURE) return XTOL;
    if (ret != XTOL) return ret;
    if (relstop(fprev, minf, st.ftol_rel, st.ftol_abs)) return FTOL;
    if (relstop(xprev, x, st.xtol_rel, st.xtol_abs)) {
      if (!(std::fabs(xstep) * psi > st.xtol_abs && std::fabs(xstep) * psi > st.xtol_rel * std::fabs(x))) return XTOL;
    }
    dx = x - xprev;
    const double scale = psi;  // nsubs == 1
    xstep = (dx == 0) ? -(xstep * scale) : std::copysign(xstep * scale, dx);
  }
}
}  // namespace sbplx

// ------------------------------------------------------ driver (A1, A2) --
struct Geometry {
  uint32_t Nt, fw, Nw;
};
static Geometry geometry(const pgsvt_params& p) {
  Geometry g;
  g.Nt = (p.patch_size * p.patch_size < p.trajectory_length) ? p.patch_size * p.patch_size - 1 : p.trajectory_length;
  g.fw = (uint32_t)std::floor(g.Nt / 2);
  g.Nw = 2 * g.fw + 1;
  return g;
}

static double elem(const void* X, int dtype, size_t k) {
  switch (dtype) {
    case PGSVT_U8: return (double)((const uint8_t*)X)[k];
    case PGSVT_U16: return (double)((const uint16_t*)X)[k];
    case PGSVT_F32: return (double)((const float*)X)[k];
    default: return ((const double*)X)[k];
  }
}

struct Window {
  uint32_t s0, ref;
  double uMax, wMax;
  Cube u, w;  // w is the (median-filtered) motion-estimation input
};

// Median pre-filter of one frame (src/pguresvt.hpp:67-84 with
// src/medfilter.hpp:247-538): the frame is converted to uint16 (negative ->
// 0, otherwise truncation, as Armadillo conv_to does for unsigned targets)
// and each pixel becomes the median of the (2r+1)^2 window with edge pixels
// replicated.  ctmf's single stripe (its stripe count collapses to one for
// any r >= 4 through unsigned wrap-around) walks exactly this window.
static void median_frame(const double* in, double* out, uint32_t nr, uint32_t nc, int r) {
  std::vector<uint16_t> f((size_t)nr * nc);
  for (size_t k = 0; k < f.size(); ++k) {
    const double v = in[k];
    f[k] = v < 0 ? (uint16_t)0 : (uint16_t)(int64_t)v;
  }
  std::vector<uint16_t> win;
  win.reserve((size_t)(2 * r + 1) * (2 * r + 1));
  for (int c = 0; c < (int)nc; ++c)
    for (int rr = 0; rr < (int)nr; ++rr) {
      win.clear();
      for (int dc = -r; dc <= r; ++dc)
        for (int dr = -r; dr <= r; ++dr) {
          const int y = std::min(std::max(rr + dr, 0), (int)nr - 1);
          const int x = std::min(std::max(c + dc, 0), (int)nc - 1);
          win.push_back(f[y + (size_t)nr * x]);
        }
      const size_t h = win.size() / 2;
      std::nth_element(win.begin(), win.begin() + h, win.end());
      out[rr + (size_t)nr * c] = (double)win[h];
    }
}

// src/pguresvt.hpp:100-120 (median filter off: Z == X, so w == u)
static Window make_window(const void* X, int dtype, uint32_t nr, uint32_t nc, uint32_t nf, const Geometry& g,
                          uint32_t t, const std::vector<double>* Z = nullptr) {
  Window w;
  if (t < g.fw) {
    w.s0 = 0;
    w.ref = t;
  } else if (t >= nf - g.fw) {
    w.s0 = nf - 2 * g.fw - 1;
    w.ref = t - (nf - g.Nt);
  } else {
    w.s0 = t - g.fw;
    w.ref = g.fw;
  }
  w.u = Cube(nr, nc, g.Nw);
  const size_t fs = (size_t)nr * nc;
  for (size_t k = 0; k < w.u.n(); ++k) w.u.d[k] = elem(X, dtype, (size_t)w.s0 * fs + k);
  double mx = w.u.d[0];
  for (double v : w.u.d) mx = std::max(mx, v);  // arma max
  w.uMax = mx;
  if (Z) {
    w.w = Cube(nr, nc, g.Nw);
    std::memcpy(w.w.d.data(), Z->data() + (size_t)w.s0 * fs, w.w.n() * sizeof(double));
    double zm = w.w.d[0];
    for (double v : w.w.d) zm = std::max(zm, v);
    w.wMax = zm;
    for (double& v : w.w.d) v /= zm;
  }
  for (double& v : w.u.d) v /= mx;
  if (!Z) {
    w.wMax = mx;
    w.w = w.u;
  }
  return w;
}

static std::vector<double> median_sequence(const void* X, int dtype, uint32_t nr, uint32_t nc, uint32_t nf, int r,
                                           int64_t nJobs);

static int validate(const void* X, int dtype, uint32_t nr, uint32_t nc, uint32_t nf, const pgsvt_params* p) {
  if (!X || !p) return fail(PGSVT_E_ARG, "null pointer");
  if (dtype < 0 || dtype > 3) return fail(PGSVT_E_ARG, "invalid dtype");
  if (nr != nc) return fail(PGSVT_E_ARG, "frames must be square (the reference patch tables assume n_rows == n_cols)");
  if (p->patch_size < 2 || p->patch_size > nr) return fail(PGSVT_E_ARG, "invalid patch_size");
  if (p->patch_overlap < 1 || p->patch_overlap > p->patch_size) return fail(PGSVT_E_ARG, "invalid patch_overlap");
  if (p->trajectory_length < 1 || p->trajectory_length % 2 == 0) return fail(PGSVT_E_ARG, "invalid trajectory_length");
  const Geometry g = geometry(*p);
  if (g.Nt < 1 || nf < g.Nw) return fail(PGSVT_E_ARG, "fewer frames than the temporal window");
  if (p->motion_estimation && (p->motion_window < 2 || p->motion_window % 2 == 0))
    return fail(PGSVT_E_ARG, "invalid motion_window");
  const size_t N = (size_t)nr * nc * nf;
  for (size_t k = 0; k < N; ++k) {
    const double v = elem(X, dtype, k);
    if (!std::isfinite(v)) return fail(PGSVT_E_NUMERIC, "non-finite input");
    if (v < 0) return fail(PGSVT_E_ARG, "negative input");
  }
  return 0;
}

static std::vector<int32_t> window_patches(const Window& w, const pgsvt_params& p, const Geometry& g, uint32_t t,
                                           uint32_t nf) {
  Motion me(w.w, p.patch_size, t, g.fw, p.motion_window, nf);
  me.Estimate(p.motion_estimation != 0);
  return me.patches;
}

struct WindowResult {
  double lambda, alpha, mu, sigma;
  int status;
};

// one iteration of the pgureFunc lambda (src/pguresvt.hpp:90-167)
static WindowResult run_window(const void* X, int dtype, uint32_t nr, uint32_t nc, uint32_t nf, const pgsvt_params& p,
                               uint32_t t, double* Y, const std::vector<double>* Z) {
  const Geometry g = geometry(p);
  WindowResult res{};
  double lambda = (p.lambda1 >= 0.0) ? p.lambda1 : -1.0;
  double alpha = (p.noise_alpha >= 0.0) ? p.noise_alpha : -1.0;
  double mu = (p.noise_mu >= 0.0) ? p.noise_mu : -1.0;
  double sigma = (p.noise_sigma >= 0.0) ? p.noise_sigma : -1.0;
  Window w = make_window(X, dtype, nr, nc, nf, g, t, Z);
  if (p.optimize_pgure) {
    NoiseEstimator ne(p.noise_method);
    try {
      ne.Estimate(w.u, alpha, mu, sigma);
    } catch (const std::runtime_error&) {  // the reference aborts here
      res.status = PGSVT_E_NUMERIC;
      return res;
    }
  }
  std::vector<int32_t> patches = window_patches(w, p, g, t, nf);
  PGURE opt(w.u, patches, alpha, sigma, mu, p.patch_size, p.patch_overlap, p.random_seed,
            p.exponential_weighting != 0, p.optimize_pgure != 0);
  if (p.optimize_pgure) {
    const double OoNxNyNt = 1.0 / (nc * nr * g.Nt);
    double start = (lambda >= 0.0) ? lambda : accu2(w.u.d.data(), w.u.n()) * OoNxNyNt;
    start = std::max(0.0, start);
    const double ub = std::max(100.0, start);
    sbplx::Stop st;
    st.ftol_rel = p.tol;
    st.xtol_abs = 1e-12;
    st.maxeval = (int)p.max_iter;
    double x = start, minf = 0;
    const int rc = sbplx::minimize([&](double l) { return opt.Risk(l); }, 0.0, ub, x, minf, std::sqrt(start), st);
    if (rc < 0) res.status = PGSVT_E_NUMERIC;
    lambda = opt.lambda;
  }
  Cube v = opt.Reconstruct(lambda);
  for (double& d : v.d) d *= w.uMax;
  res.lambda = lambda;
  res.alpha = alpha;
  res.mu = mu;
  res.sigma = sigma;
  uint32_t slice;
  if (t < g.fw)
    slice = t;
  else if (t >= nf - g.fw)
    slice = t - (nf - g.Nt);
  else
    slice = g.fw;
  const size_t fs = (size_t)nr * nc;
  std::memcpy(Y + (size_t)t * fs, &v.d[(size_t)slice * fs], fs * sizeof(double));
  return res;
}

// src/utils.hpp:108-168
template <class F>
static void parallel(F func, uint32_t first, uint32_t last, int64_t nJobs) {
  const uint32_t total = (nJobs > 0) ? (uint32_t)nJobs : std::thread::hardware_concurrency();
  if (nJobs == 0 || total <= 1 || (last - first) <= 1) {
    for (uint32_t a = first; a != last; ++a) func(a);
    return;
  }
  std::vector<std::thread> th;
  if (last - first <= total) {
    for (uint32_t a = first; a != last; ++a) th.emplace_back([&func, a] { func(a); });
    for (auto& x : th) x.join();
    return;
  }
  auto slice = [&func](uint32_t a, uint32_t b) {
    while (a < b) func(a++);
  };
  const uint64_t per = (last - first + total - 1) / total;
  for (uint32_t k = 0; k + 1 < total; ++k) {
    uint32_t a = (uint32_t)std::min<uint64_t>(per * k + first, last);
    uint32_t b = (uint32_t)std::min<uint64_t>((uint64_t)a + per, last);
    th.emplace_back(slice, a, b);
  }
  slice((uint32_t)std::min<uint64_t>(per * (total - 1), last), last);
  for (auto& x : th) x.join();
}

static std::vector<double> median_sequence(const void* X, int dtype, uint32_t nr, uint32_t nc, uint32_t nf, int r,
                                           int64_t nJobs) {
  const size_t fs = (size_t)nr * nc;
  std::vector<double> Z(fs * nf);
  parallel(
      [&](uint32_t f) {
        std::vector<double> in(fs);
        for (size_t k = 0; k < fs; ++k) in[k] = elem(X, dtype, f * fs + k);
        median_frame(in.data(), &Z[f * fs], nr, nc, r);
      },
      0, nf, nJobs);
  return Z;
}

static std::unique_ptr<std::vector<double>> maybe_median(const void* X, int dtype, uint32_t nr, uint32_t nc,
                                                         uint32_t nf, const pgsvt_params& p) {
  if (p.motion_filter <= 0) return nullptr;
  return std::unique_ptr<std::vector<double>>(
      new std::vector<double>(median_sequence(X, dtype, nr, nc, nf, (int)p.motion_filter, p.n_jobs)));
}

}  // namespace orc

// ================================================================ C API ====
using namespace orc;

// ------------------------------------------ hot-pixel filter (front end) --
// HotPixelFilter (src/hotpixel.hpp:19-64) on uint16 frames, in place.
// Armadillo's median of integers: odd n -> element n/2 of the sorted values;
// even n -> upper + (lower - upper) / 2 in int arithmetic (rounded towards
// the upper middle).  The frame median is the median of the column medians
// (arma::median applied twice), likewise the MAD of |x - median| (times
// 1 / 0.6745).  Stated deviation from the reference: |x - median| is the
// exact distance; the reference's uint16 subtraction wraps below the median.
static uint32_t int_median(std::vector<uint32_t> v) {
  const size_t n = v.size(), h = n / 2;
  std::sort(v.begin(), v.end());
  if (n % 2) return v[h];
  return (uint32_t)((int)v[h] + ((int)v[h - 1] - (int)v[h]) / 2);
}

static void hot_pixel_frame(uint16_t* F, uint32_t nr, uint32_t nc, double threshold) {
  std::vector<uint32_t> col(nr), meds(nc);
  for (uint32_t c = 0; c < nc; ++c) {
    for (uint32_t r = 0; r < nr; ++r) col[r] = F[r + (size_t)nr * c];
    meds[c] = int_median(col);
  }
  const uint32_t med = int_median(meds);
  auto dist = [&](uint32_t x) { return x > med ? x - med : med - x; };
  for (uint32_t c = 0; c < nc; ++c) {
    for (uint32_t r = 0; r < nr; ++r) col[r] = dist(F[r + (size_t)nr * c]);
    meds[c] = int_median(col);
  }
  const double mad = (double)int_median(meds) * (1.0 / 0.6745);
  std::vector<size_t> outliers;  // found before any replacement (arma::find)
  for (size_t i = 0; i < (size_t)nr * nc; ++i)
    if ((double)dist(F[i]) > threshold * mad) outliers.push_back(i);
  for (size_t i : outliers) {  // replaced in index order, in place
    const uint32_t r = (uint32_t)(i % nr), c = (uint32_t)(i / nr);
    if (r > 0 && r < nr - 1 && c > 0 && c < nc - 1) {
      double w[8];
      int q = 0;
      for (int dc = -1; dc <= 1; ++dc)  // the 8 neighbours (their order does not matter: sorted next)
        for (int dr = -1; dr <= 1; ++dr)
          if (dr || dc) w[q++] = (double)F[(r + dr) + (size_t)nr * (c + dc)];
      std::sort(w, w + 8);
      F[i] = (uint16_t)(0.5 * (w[3] + w[4]));
    } else {
      F[i] = (uint16_t)med;
    }
  }
}

extern "C" {

const char* oracle_last_error(void) { return g_err.c_str(); }

int oracle_set_svd(int backend) {
  if (backend == 1) {
    if (!load_lapack()) return fail(PGSVT_E_ARG, "no LAPACK dgesdd found");
    g_svd_backend = 1;
    return 0;
  }
  g_svd_backend = 0;
  return 0;
}

const char* oracle_svd_name(void) {
  static std::string s;
  s = (g_svd_backend.load() == 1) ? "dgesdd(" + g_lapack_name + ")" : "jacobi-fp64";
  return s.c_str();
}

int oracle_denoise(const void* X, int dtype, uint32_t nr, uint32_t nc, uint32_t nf, const pgsvt_params* p, double* Y,
                   double* est) {
  if (int rc = validate(X, dtype, nr, nc, nf, p)) return rc;
  auto Z = maybe_median(X, dtype, nr, nc, nf, *p);
  std::vector<WindowResult> res(nf);
  parallel([&](uint32_t t) { res[t] = run_window(X, dtype, nr, nc, nf, *p, t, Y, Z.get()); }, 0, nf, p->n_jobs);
  for (uint32_t t = 0; t < nf; ++t) {
    est[t + (size_t)nf * 0] = res[t].lambda;
    est[t + (size_t)nf * 1] = res[t].alpha;
    est[t + (size_t)nf * 2] = res[t].mu;
    est[t + (size_t)nf * 3] = res[t].sigma;
  }
  for (uint32_t t = 0; t < nf; ++t)
    if (res[t].status)
      return fail(res[t].status, "window " + std::to_string(t) + ": noise estimate or optimiser failure");
  return 0;
}

// Denoise only the output frames [t0, t1) (CPU-baseline sampling).
int oracle_denoise_range(const void* X, int dtype, uint32_t nr, uint32_t nc, uint32_t nf, const pgsvt_params* p,
                         uint32_t t0, uint32_t t1, double* Y, double* est) {
  if (int rc = validate(X, dtype, nr, nc, nf, p)) return rc;
  if (t0 >= t1 || t1 > nf) return fail(PGSVT_E_ARG, "bad range");
  auto Z = maybe_median(X, dtype, nr, nc, nf, *p);
  std::vector<WindowResult> res(nf);
  parallel([&](uint32_t t) { res[t] = run_window(X, dtype, nr, nc, nf, *p, t, Y, Z.get()); }, t0, t1, p->n_jobs);
  for (uint32_t t = t0; t < t1; ++t) {
    est[t + (size_t)nf * 0] = res[t].lambda;
    est[t + (size_t)nf * 1] = res[t].alpha;
    est[t + (size_t)nf * 2] = res[t].mu;
    est[t + (size_t)nf * 3] = res[t].sigma;
  }
  return 0;
}

int oracle_window_trajectories(const void* X, int dtype, uint32_t nr, uint32_t nc, uint32_t nf,
                               const pgsvt_params* p, uint32_t t, int32_t* patches) {
  if (int rc = validate(X, dtype, nr, nc, nf, p)) return rc;
  const Geometry g = geometry(*p);
  auto Z = maybe_median(X, dtype, nr, nc, nf, *p);
  Window w = make_window(X, dtype, nr, nc, nf, g, t, Z.get());
  std::vector<int32_t> pt = window_patches(w, *p, g, t, nf);
  std::memcpy(patches, pt.data(), pt.size() * sizeof(int32_t));
  return 0;
}

int oracle_window_risk(const void* X, int dtype, uint32_t nr, uint32_t nc, uint32_t nf, const pgsvt_params* p,
                       uint32_t t, const double* noise, const double* lambdas, uint32_t n, double* risk) {
  if (int rc = validate(X, dtype, nr, nc, nf, p)) return rc;
  const Geometry g = geometry(*p);
  auto Z = maybe_median(X, dtype, nr, nc, nf, *p);
  Window w = make_window(X, dtype, nr, nc, nf, g, t, Z.get());
  std::vector<int32_t> pt = window_patches(w, *p, g, t, nf);
  PGURE opt(w.u, pt, noise[0], noise[2], noise[1], p->patch_size, p->patch_overlap, p->random_seed,
            p->exponential_weighting != 0, true);
  for (uint32_t k = 0; k < n; ++k) risk[k] = opt.Risk(lambdas[k]);
  return 0;
}

int oracle_window_reconstruct(const void* X, int dtype, uint32_t nr, uint32_t nc, uint32_t nf,
                              const pgsvt_params* p, uint32_t t, double lambda, double* vol) {
  if (int rc = validate(X, dtype, nr, nc, nf, p)) return rc;
  const Geometry g = geometry(*p);
  auto Z = maybe_median(X, dtype, nr, nc, nf, *p);
  Window w = make_window(X, dtype, nr, nc, nf, g, t, Z.get());
  std::vector<int32_t> pt = window_patches(w, *p, g, t, nf);
  SVT s(pt, w.u.nr, w.u.nc, w.u.ns, p->patch_size, p->patch_overlap, p->exponential_weighting != 0);
  s.Decompose(w.u);
  Cube v = s.Reconstruct(lambda);
  std::memcpy(vol, v.d.data(), v.n() * sizeof(double));
  return 0;
}

int oracle_window_singular_values(const void* X, int dtype, uint32_t nr, uint32_t nc, uint32_t nf,
                                  const pgsvt_params* p, uint32_t t, double* sv, uint32_t* n_patches) {
  if (int rc = validate(X, dtype, nr, nc, nf, p)) return rc;
  const Geometry g = geometry(*p);
  auto Z = maybe_median(X, dtype, nr, nc, nf, *p);
  Window w = make_window(X, dtype, nr, nc, nf, g, t, Z.get());
  std::vector<int32_t> pt = window_patches(w, *p, g, t, nf);
  SVT s(pt, w.u.nr, w.u.nc, w.u.ns, p->patch_size, p->patch_overlap, p->exponential_weighting != 0);
  s.Decompose(w.u);
  std::memcpy(sv, s.S.data(), s.S.size() * sizeof(double));
  *n_patches = (uint32_t)s.actual.size();
  return 0;
}

int oracle_noise_estimate(const double* u, uint32_t nr, uint32_t nc, uint32_t nt, uint32_t method, double* alpha,
                          double* mu, double* sigma) {
  Cube c(nr, nc, nt);
  std::memcpy(c.d.data(), u, c.n() * sizeof(double));
  NoiseEstimator ne(method);
  try {
    ne.Estimate(c, *alpha, *mu, *sigma);
  } catch (const std::runtime_error& e) {
    return fail(PGSVT_E_NUMERIC, e.what());
  }
  return 0;
}

// The estimator's pieces on their own (src/noise.hpp), so that tests can
// check each against a closed form.
int oracle_noise_robust_mean(const double* a, uint32_t n, double* out) {  // noise.hpp:238-271
  NoiseEstimator ne(4);
  *out = ne.RobustMean(std::vector<double>(a, a + n));
  return 0;
}

int oracle_noise_robust_var(const double* a, uint32_t n, double* out) {  // noise.hpp:232-236
  *out = NoiseEstimator::RobustVar(std::vector<double>(a, a + n));
  return 0;
}

int oracle_noise_mode(const double* a, uint32_t n, double* out) {  // noise.hpp:273-301
  NoiseEstimator ne(4);
  *out = ne.ComputeMode(std::vector<double>(a, a + n));
  return 0;
}

int oracle_noise_wls(const double* x, const double* y, uint32_t n, double* a, double* b) {  // noise.hpp:328-383
  NoiseEstimator ne(4);
  ne.WLSFit(std::vector<double>(x, x + n), std::vector<double>(y, y + n), *a, *b);
  return 0;
}

// F-test split decision for one N x N column-major block (noise.hpp:182-221)
int oracle_noise_split(const double* block, uint32_t N, int32_t* split) {
  NoiseEstimator ne(4);
  *split = ne.SplitBlockQ(block, N, 0, 0, N) ? 1 : 0;
  return 0;
}

// Quadtree nodes of one N x N frame after the parents are shed
// (noise.hpp:55-72, 419-458): (row, col, size) triples in tree order;
// returns the node count, writes at most cap triples.
uint32_t oracle_noise_tree(const double* frame, uint32_t N, uint32_t* nodes, uint32_t cap) {
  NoiseEstimator ne(4);
  ne.tree.assign(1, {0u, 0u, N});
  ne.del.clear();
  ne.QuadTree(frame, N, 0);
  std::vector<std::array<uint32_t, 3>> t2 = ne.tree;
  std::vector<uint32_t> dl = ne.del;
  std::sort(dl.begin(), dl.end());
  dl.erase(std::unique(dl.begin(), dl.end()), dl.end());
  if (!dl.empty())
    for (size_t k = dl.size() - 1; k > 0; --k) t2.erase(t2.begin() + dl[k]);
  for (size_t k = 0; k < t2.size() && k < cap; ++k)
    for (int q = 0; q < 3; ++q) nodes[3 * k + q] = t2[k][q];
  return (uint32_t)t2.size();
}

int oracle_median_filter(const void* X, int dtype, uint32_t nr, uint32_t nc, uint32_t nf, int r, double* Z) {
  std::vector<double> z = median_sequence(X, dtype, nr, nc, nf, r, 1);
  std::memcpy(Z, z.data(), z.size() * sizeof(double));
  return 0;
}

int oracle_pcg64(uint64_t seed, uint64_t n, uint64_t* out) {
  Pcg64 r(seed);
  for (uint64_t k = 0; k < n; ++k) out[k] = r();
  return 0;
}

int oracle_perturbations(int64_t seed, uint64_t n, int8_t* d1, double* d2) {
  Pcg64 rng((uint64_t)seed);
  const double vP = 0.5 + 0.5 * 1. / std::sqrt(1. * 1. + 4);
  const double vQ = 1 - vP, vQvP = std::sqrt(vQ / vP), vPvQ = std::sqrt(vP / vQ);
  std::bernoulli_distribution b1(0.5), b2(vP);
  for (uint64_t k = 0; k < n; ++k) d1[k] = b1(rng) ? -1 : 1;
  for (uint64_t k = 0; k < n; ++k) d2[k] = b2(rng) ? -1 * vQvP : vPvQ;
  return 0;
}

uint32_t oracle_patch_indices(uint32_t N, uint32_t bs, uint32_t step, uint32_t* out) {
  std::vector<uint32_t> v = patch_indices(N, N, bs, step);
  if (out) std::memcpy(out, v.data(), v.size() * sizeof(uint32_t));
  return (uint32_t)v.size();
}

// One-variable LN_SBPLX restatement on a caller-supplied objective.
int oracle_sbplx(double (*f)(double, void*), void* ctx, double start, double lb, double ub, double step, double ftol,
                 double xtol_abs, int maxeval, double* x_best, double* f_best, double* x_last, int* nevals) {
  sbplx::Stop st;
  st.ftol_rel = ftol;
  st.xtol_abs = xtol_abs;
  st.maxeval = maxeval;
  double last = start;
  double x = start, minf = 0;
  const int rc = sbplx::minimize(
      [&](double l) {
        last = l;
        return f(l, ctx);
      },
      lb, ub, x, minf, step, st);
  *x_best = x;
  *f_best = minf;
  *x_last = last;
  *nevals = st.nevals;
  return rc;
}

int oracle_hotpixel(uint16_t* seq, uint32_t nr, uint32_t nc, uint32_t nf, double threshold) {
  for (uint32_t k = 0; k < nf; ++k) hot_pixel_frame(seq + (size_t)k * nr * nc, nr, nc, threshold);
  return 0;
}

}  // extern "C"
