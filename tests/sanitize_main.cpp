// Host code under AddressSanitizer and UndefinedBehaviorSanitizer: the CPU
// FP64 oracle, the product's host noise estimator, TIFF reader/writer and
// lambda search, built with g++ -fsanitize=address,undefined (no GPU code
// is involved) and exercised on small inputs, error paths included.
// tests/test_sanitizers.py builds and runs it; any sanitizer report fails
// the run.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../include/pgsvt.h"
#include "../pgure-svt_amd/csrc/noise.hpp"
#include "../pgure-svt_amd/csrc/sbplx.hpp"
#include "../pgure-svt_amd/csrc/tiff.hpp"

extern "C" {
int oracle_denoise(const void* X, int dtype, uint32_t nr, uint32_t nc, uint32_t nf, const pgsvt_params* p, double* Y,
                   double* est);
int oracle_window_risk(const void* X, int dtype, uint32_t nr, uint32_t nc, uint32_t nf, const pgsvt_params* p,
                       uint32_t t, const double* noise, const double* lambdas, uint32_t n, double* risk);
int oracle_noise_estimate(const double* u, uint32_t nr, uint32_t nc, uint32_t nt, uint32_t method, double* alpha,
                          double* mu, double* sigma);
uint32_t oracle_noise_tree(const double* frame, uint32_t N, uint32_t* nodes, uint32_t cap);
}

static int failures = 0;
#define CHECK(c)                                                    \
  do {                                                              \
    if (!(c)) {                                                     \
      std::fprintf(stderr, "CHECK failed at line %d: %s\n", __LINE__, #c); \
      ++failures;                                                   \
    }                                                               \
  } while (0)

// drifting bright squares on a background, with a deterministic LCG noise
static std::vector<float> sequence(uint32_t n, uint32_t nf) {
  std::vector<float> X((size_t)n * n * nf);
  uint64_t s = 12345;
  for (uint32_t k = 0; k < nf; ++k)
    for (uint32_t c = 0; c < n; ++c)
      for (uint32_t r = 0; r < n; ++r) {
        s = s * 6364136223846793005ULL + 1442695040888963407ULL;
        const double u = (double)(s >> 11) * 0x1.0p-53;
        const bool in = ((r + k) % n) < n / 3 && c > n / 4 && c < n / 2;
        X[r + (size_t)n * (c + (size_t)n * k)] = (float)((in ? 80.0 : 10.0) + 12.0 * (u - 0.5));
      }
  return X;
}

// the defaults of pguresvt_{u8,u16,f,d} (pguresvt/_pguresvt.pyx:151-168)
static pgsvt_params params() {
  pgsvt_params p;
  p.trajectory_length = 15;
  p.patch_size = 4;
  p.patch_overlap = 1;
  p.motion_window = 7;
  p.motion_filter = 5;
  p.noise_method = 4;
  p.max_iter = 500;
  p.optimize_pgure = 1;
  p.exponential_weighting = 1;
  p.motion_estimation = 1;
  p.noise_alpha = p.noise_mu = p.noise_sigma = -1.0;
  p.tol = 1e-7;
  p.n_jobs = 2;
  p.random_seed = 7;
  p.lambda1 = -1.0;
  return p;
}

int main() {
  const uint32_t n = 32, nf = 16;
  const std::vector<float> X = sequence(n, nf);
  std::vector<double> Y((size_t)n * n * nf), est(4 * nf);
  // the full reference driver: median filter, ARPS, PGURE search
  pgsvt_params p = params();
  p.motion_filter = 3;
  CHECK(oracle_denoise(X.data(), PGSVT_F32, n, n, nf, &p, Y.data(), est.data()) == 0);
  for (double v : Y) CHECK(std::isfinite(v));
  // motion off, fixed lambda, plain thresholding, patch step 2
  p = params();
  p.motion_estimation = 0;
  p.optimize_pgure = 0;
  p.exponential_weighting = 0;
  p.lambda1 = 0.05;
  p.patch_overlap = 2;
  CHECK(oracle_denoise(X.data(), PGSVT_F32, n, n, nf, &p, Y.data(), est.data()) == 0);
  // a window risk at several lambdas
  p = params();
  const double noise[3] = {0.02, 0.03, 0.05}, lams[3] = {0.5, 5.0, 50.0};
  double risk[3];
  CHECK(oracle_window_risk(X.data(), PGSVT_F32, n, n, nf, &p, 9, noise, lams, 3, risk) == 0);
  // the noise estimator, product against oracle, all methods
  std::vector<double> u((size_t)n * n * 15);
  double mx = 0;
  for (size_t i = 0; i < u.size(); ++i) mx = std::fmax(mx, (double)X[i]);
  for (size_t i = 0; i < u.size(); ++i) u[i] = X[i] / mx;
  for (uint32_t m = 1; m <= 4; ++m) {
    double a0 = -1, m0 = -1, s0 = -1, a1 = -1, m1 = -1, s1 = -1;
    const int orc = oracle_noise_estimate(u.data(), n, n, 15, m, &a0, &m0, &s0);
    bool threw = false;
    try {
      pgsvt::estimate_noise(u.data(), n, 15, m, a1, m1, s1);
    } catch (const pgsvt::NoiseError&) {
      threw = true;
    }
    CHECK(threw == (orc != 0));
    if (!threw) CHECK(a0 == a1 && m0 == m1 && s0 == s1);
  }
  // method 1 without enough leaves (an all-zero 256 x 256 window has one
  // per frame, 15 against 219): an error, not an out-of-range read
  const uint32_t nz = 256;
  std::vector<double> zero((size_t)nz * nz * 15, 0.0);
  double a = -1, m = -1, s = -1;
  CHECK(oracle_noise_estimate(zero.data(), nz, nz, 15, 1, &a, &m, &s) != 0);
  bool threw = false;
  try {
    pgsvt::estimate_noise(zero.data(), nz, 15, 1, a, m, s);
  } catch (const pgsvt::NoiseError&) {
    threw = true;
  }
  CHECK(threw);
  std::vector<uint32_t> nodes(3 * 64);
  CHECK(oracle_noise_tree(u.data(), n, nodes.data(), 64) >= 1);
  // TIFF round trip and malformed files
  const char* path = "/tmp/pgsvt_sanitize.tif";
  std::vector<uint16_t> cube((size_t)n * n * 3);
  for (size_t i = 0; i < cube.size(); ++i) cube[i] = (uint16_t)(i * 37);
  pgsvt::tiff_write_u16(path, cube.data(), n, n, 3);
  const pgsvt::TiffInfo ti = pgsvt::tiff_info(path);
  CHECK(ti.width == n && ti.height == n && ti.pages == 3);
  std::vector<uint16_t> back(cube.size());
  pgsvt::tiff_read_u16(path, 0, 3, back.data());
  CHECK(back == cube);
  {
    FILE* f = std::fopen(path, "rb");
    std::vector<uint8_t> buf(1 << 16);
    const size_t len = std::fread(buf.data(), 1, buf.size(), f);
    std::fclose(f);
    buf.resize(len);
    // every byte of the first IFD set to 0xff in turn: each is an error or
    // a readable file, never a sanitizer report
    uint32_t ifd = buf[4] | buf[5] << 8 | buf[6] << 16 | (uint32_t)buf[7] << 24;
    for (uint32_t at = ifd; at < ifd + 2 + 12 * 13 && at < buf.size(); ++at) {
      std::vector<uint8_t> bad = buf;
      bad[at] = 0xff;
      FILE* g = std::fopen(path, "wb");
      std::fwrite(bad.data(), 1, bad.size(), g);
      std::fclose(g);
      try {
        const pgsvt::TiffInfo bi = pgsvt::tiff_info(path);
        if (bi.pages >= 1 && (size_t)bi.width * bi.height <= cube.size() / 3) {
          std::vector<uint16_t> o((size_t)bi.width * bi.height);
          pgsvt::tiff_read_u16(path, 0, 1, o.data());
        }
      } catch (const std::runtime_error&) {
      }
    }
  }
  std::remove(path);
  // the lambda search on a quadratic
  pgsvt::Sbplx1D opt(0.7, 0.0, 100.0, std::sqrt(0.7), 1e-7, 1e-12, 500);
  int rc;
  while ((rc = opt.run()) == pgsvt::Sbplx1D::kNeed) opt.supply((opt.pending() - 3.3) * (opt.pending() - 3.3) + 1.0);
  CHECK(rc > 0 && std::fabs(opt.best_x() - 3.3) < 1e-3);
  if (failures) return 1;
  std::printf("sanitized host run ok\n");
  return 0;
}
