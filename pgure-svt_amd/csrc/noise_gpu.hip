// noise_gpu.hip -- quadtree leaf statistics of the noise estimator on the GPU
// (src/noise.hpp:55-100, 173-271, 395-458), for the learned-lambda path.
//
// Split tests: one workgroup per candidate block of a frame (every block of
// the complete quadtree whose size has an F-table entry), two-pass sums of
// the values and of their 5-point residuals; the host replays the
// reference's recursion with the decisions.  Leaf statistics: one workgroup
// per distinct leaf; leaves too large for its registers (every frame's root
// block among them) have their values and Laplacians formed once, spread
// over the GPU (k_leaf_prep), and the workgroup streams them.  The robust variance (median absolute deviation of the
// 9-point Laplacian) is bit-identical to the host estimator: every Laplacian
// value is formed per element in the host's order, and the medians are
// exact order statistics, selected by a radix walk over the values' bit
// patterns.  The robust mean's IRLS sums are workgroup reductions, whose
// order differs from the host's even/odd accumulators: the mean equals the
// host's to within their rounding, and so may the split tests' sums, which
// changes a decision only when a block's F statistic lies within that
// rounding of its threshold.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "kernels.hpp"

namespace pgsvt {
namespace {

// sum over the 256 threads of the workgroup, returned to every thread
// (red: 4 doubles of LDS, free again when the function returns to all)
__device__ double wg_sum(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  __syncthreads();  // the previous sum's readers are done with red
  if ((threadIdx.x & 63) == 0) red[threadIdx.x / 64] = v;
  __syncthreads();
  return ((red[0] + red[1]) + red[2]) + red[3];
}

// three sums at once (red: 12 doubles)
__device__ void wg_sum3(double& a, double& b, double& c, double* red) {
  for (int o = 32; o > 0; o >>= 1) {
    a += __shfl_xor(a, o);
    b += __shfl_xor(b, o);
    c += __shfl_xor(c, o);
  }
  __syncthreads();
  if ((threadIdx.x & 63) == 0) {
    red[threadIdx.x / 64] = a;
    red[4 + threadIdx.x / 64] = b;
    red[8 + threadIdx.x / 64] = c;
  }
  __syncthreads();
  a = ((red[0] + red[1]) + red[2]) + red[3];
  b = ((red[4] + red[5]) + red[6]) + red[7];
  c = ((red[8] + red[9]) + red[10]) + red[11];
}

// totally ordered 64-bit key of a double (negative values reversed)
__device__ __forceinline__ uint64_t order_key(double x) {
  const uint64_t u = __builtin_bit_cast(uint64_t, x);
  return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
__device__ __forceinline__ double from_key(uint64_t k) {
  return __builtin_bit_cast(double, (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k);
}

// LDS of the selection: two 256-bin histograms, the wave totals of their
// prefix sums, the chosen prefixes and remaining ranks
struct SelectLds {
  uint32_t hist[2][256];
  uint32_t wtot[2][4];
  uint64_t prefix[2];
  uint32_t k[2];
};

// inclusive prefix sum of v over the 256 threads (t: this thread's value)
__device__ __forceinline__ uint32_t wg_incl_scan(uint32_t v, uint32_t* wtot) {
  const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(v, o, 64);
    v += lane >= (uint32_t)o ? y : 0u;
  }
  if (lane == 63) wtot[w] = v;
  __syncthreads();
  for (uint32_t q = 0; q < w; ++q) v += wtot[q];
  return v;
}

// The elements of ranks k0 and k1 (0-based, ascending) of the values each
// thread visits (each(f) calls f(value) for the thread's elements): eight
// passes, each a 256-bin histogram of the next key byte among the values
// that match the bytes chosen so far, for both ranks at once.
template <class Each>
__device__ void wg_select2(uint32_t k0, uint32_t k1, Each each, SelectLds& L, double& out0, double& out1) {
  uint64_t prefix[2] = {0, 0}, mask = 0;
  uint32_t kk[2] = {k0, k1};
  for (int shift = 56; shift >= 0; shift -= 8) {
    L.hist[0][threadIdx.x] = 0u;
    L.hist[1][threadIdx.x] = 0u;
    __syncthreads();
    each([&](double x) {
      const uint64_t key = order_key(x);
      const uint32_t bin = (uint32_t)(key >> shift) & 255u;
      if ((key & mask) == prefix[0]) atomicAdd(&L.hist[0][bin], 1u);
      if ((key & mask) == prefix[1]) atomicAdd(&L.hist[1][bin], 1u);
    });
    __syncthreads();
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const uint32_t c = L.hist[h][threadIdx.x];
      const uint32_t incl = wg_incl_scan(c, L.wtot[h]), excl = incl - c;
      if (excl <= kk[h] && kk[h] < incl) {  // exactly one thread: the bin of rank k
        L.prefix[h] = prefix[h] | ((uint64_t)threadIdx.x << shift);
        L.k[h] = kk[h] - excl;
      }
      __syncthreads();
    }
    prefix[0] = L.prefix[0];
    prefix[1] = L.prefix[1];
    kk[0] = L.k[0];
    kk[1] = L.k[1];
    mask |= 0xffull << shift;
    __syncthreads();  // L is rewritten by the next pass
  }
  out0 = from_key(prefix[0]);
  out1 = from_key(prefix[1]);
}

// SplitBlockQ (src/noise.hpp:182-221) statistics of block (i0, j0, N) of
// frame `frame`: the F statistic max(Sz, Se) / min(Sz, Se) of the values
// and of their 5-point residuals (wrap-around within the block), element
// k = y + N x.
__global__ __launch_bounds__(256) void k_split_stats(const double* __restrict__ X, uint32_t n, double scale,
                                                     const uint32_t* __restrict__ frames,
                                                     const uint32_t* __restrict__ blocks, uint32_t nblocks,
                                                     double* __restrict__ stat) {
#pragma clang fp contract(off)
  __shared__ double red[4];
  const uint32_t bi = blockIdx.x, fi = blockIdx.y;
  const uint32_t i0 = blocks[3 * bi], j0 = blocks[3 * bi + 1], N = blocks[3 * bi + 2];
  const double* F = X + (size_t)frames[fi] * n * n;
  const uint32_t R = N * N;
  const double sq = sqrt((double)(5 * 5 + 5));
  auto val = [&](uint32_t y, uint32_t x) { return F[(i0 + y) + (size_t)n * (j0 + x)] / scale; };
  auto res = [&](uint32_t y, uint32_t x) {
    const uint32_t yp = (y + 1 == N) ? 1 : y + 1, ym = (y == 0) ? N - 2 : y - 1;
    const uint32_t xp = (x + 1 == N) ? 1 : x + 1, xm = (x == 0) ? N - 2 : x - 1;
    return (5.0 * val(y, x) - (((val(yp, x) + val(ym, x)) + val(y, xm)) + val(y, xp))) / sq;
  };
  double sa = 0.0, sr = 0.0;
  for (uint32_t e = threadIdx.x; e < R; e += 256) {
    sa += val(e % N, e / N);
    sr += res(e % N, e / N);
  }
  const double OoR = 1.0 / R, OoRm1 = 1.0 / (R - 1);
  const double accuZ = wg_sum(sa, red) * OoR;
  const double accuR = wg_sum(sr, red) * OoR;
  double ss = 0.0, se = 0.0;
  for (uint32_t e = threadIdx.x; e < R; e += 256) {
    const double t = val(e % N, e / N) - accuZ, u = res(e % N, e / N) - accuR;
    ss += t * t;
    se += u * u;
  }
  const double Sz = wg_sum(ss, red) * OoRm1;
  const double Se = wg_sum(se, red) * OoRm1;
  if (threadIdx.x == 0) stat[(size_t)fi * nblocks + bi] = (Sz > Se) ? Sz / Se : Se / Sz;
}

// values per thread held in registers: leaves up to 64 x 64
constexpr uint32_t kLeafCache = 16;

// Leaf element e = r + s c (column-major) of leaf (frame, i0, j0, s): its
// value X(i0 + r, j0 + c) / scale and its 9-point Laplacian from the
// wrapped 3 x 3 neighbourhood within the leaf, in the host's order (even
// terms, odd terms, then their sum).
struct LeafView {
  const double* F;
  uint32_t n, i0, j0, s;
  double scale;
  __device__ LeafView(const double* X, uint32_t n_, double scale_, const uint32_t* Lf)
      : F(X + (size_t)Lf[0] * n_ * n_), n(n_), i0(Lf[1]), j0(Lf[2]), s(Lf[3]), scale(scale_) {}
  __device__ double val(uint32_t e) const { return F[(i0 + e % s) + (size_t)n * (j0 + e / s)] / scale; }
  __device__ double lap(uint32_t e) const {
#pragma clang fp contract(off)
    const double ne = (1.0 * 0.125) * -1.0, nm = -1.0 * -1.0;
    const uint32_t x = e / s, y = e % s;
    const uint32_t xp = (x + 1 == s) ? 1 : x + 1, xm = (x == 0) ? s - 2 : x - 1;
    const uint32_t yp = (y + 1 == s) ? 1 : y + 1, ym = (y == 0) ? s - 2 : y - 1;
    auto I = [&](uint32_t r, uint32_t c) { return val(r + s * c); };
    double a1 = 0.0, a2 = 0.0;
    a1 += I(xm, ym) * ne;
    a2 += I(xm, y) * ne;
    a1 += I(xm, yp) * ne;
    a2 += I(x, ym) * ne;
    a1 += I(x, y) * nm;
    a2 += I(x, yp) * ne;
    a1 += I(xp, ym) * ne;
    a2 += I(xp, y) * ne;
    a1 += I(xp, yp) * ne;
    return a1 + a2;
  }
};

// Values and Laplacians of the leaves too large for the register cache,
// computed once: leaf big[q] into vals / laps from element boff[big[q]].
__global__ __launch_bounds__(256) void k_leaf_prep(const double* __restrict__ X, uint32_t n, double scale,
                                                   const uint32_t* __restrict__ leaves,
                                                   const uint32_t* __restrict__ big,
                                                   const uint64_t* __restrict__ boff, double* __restrict__ vals,
                                                   double* __restrict__ laps) {
#pragma clang fp contract(off)
  const uint32_t leaf = big[blockIdx.y];
  const LeafView v(X, n, scale, leaves + 4 * (size_t)leaf);
  const uint32_t N = v.s * v.s;
  const uint64_t o = boff[leaf];
  for (uint32_t e = blockIdx.x * 256 + threadIdx.x; e < N; e += gridDim.x * 256) {
    vals[o + e] = v.val(e);
    laps[o + e] = v.lap(e);
  }
}

// RobustMeanEstimate (src/noise.hpp:238-271) and RobustVarianceEstimate of
// the 9-point Laplacian (src/noise.hpp:232-236, 395-417) of leaf
// (frame, i0, j0, s): out [leaf][2] = (mean, variance).  Leaves of at most
// 256 kLeafCache elements hold their values and Laplacians in registers
// (element threadIdx.x + 256 i in slot i); larger ones read them from vals /
// laps at boff[leaf] (k_leaf_prep), or recompute them from the frame at
// every pass where boff[leaf] is kNoOffset.
constexpr uint64_t kNoOffset = ~0ull;
__global__ __launch_bounds__(256) void k_leaf_stats(const double* __restrict__ X, uint32_t n, double scale,
                                                    const uint32_t* __restrict__ leaves,
                                                    const uint64_t* __restrict__ boff,
                                                    const double* __restrict__ vals,
                                                    const double* __restrict__ laps, double* __restrict__ out) {
#pragma clang fp contract(off)
  __shared__ double red[12];
  __shared__ SelectLds sel;
  const LeafView lv(X, n, scale, leaves + 4 * (size_t)blockIdx.x);
  const uint32_t N = lv.s * lv.s;
  const bool cached = N <= 256 * kLeafCache;
  const uint64_t o = cached ? kNoOffset : boff[blockIdx.x];
  const double* V = vals + (o == kNoOffset ? 0 : o);
  const double* Lp = laps + (o == kNoOffset ? 0 : o);
  double cv[kLeafCache], cl[kLeafCache];
  if (cached) {
#pragma unroll
    for (uint32_t i = 0; i < kLeafCache; ++i) {
      const uint32_t e = threadIdx.x + 256 * i;
      cv[i] = e < N ? lv.val(e) : 0.0;
      cl[i] = e < N ? lv.lap(e) : 0.0;
    }
  }
  // visit the thread's values (or Laplacians, or a function g of them)
  auto each_of = [&](bool of_lap, auto g) {
    return [&, of_lap, g](auto f) {
      if (cached) {
#pragma unroll
        for (uint32_t i = 0; i < kLeafCache; ++i)
          if (threadIdx.x + 256 * i < N) f(g(of_lap ? cl[i] : cv[i]));
      } else if (o != kNoOffset) {
        const double* P = of_lap ? Lp : V;
        for (uint32_t e = threadIdx.x; e < N; e += 256) f(g(P[e]));
      } else {
        for (uint32_t e = threadIdx.x; e < N; e += 256) f(g(of_lap ? lv.lap(e) : lv.val(e)));
      }
    };
  };
  auto ident = [](double x) { return x; };

  // robust mean: Huber IRLS from the mean, scale from the interquartile
  // distance (its two order statistics of A fixed: sorted(A - m) is
  // sorted(A) - m, rounding being monotone)
  const uint32_t mq = ((N + 1) / 2 + 1) / 2;
  double a_hi, a_lo;
  wg_select2(N - mq - 1, mq - 1, each_of(false, ident), sel, a_hi, a_lo);
  const double tol = 1e-6, eps = 1e-12;
  double m = 0, m0 = 1e12;
  double wa = 0.0, ws = 0.0, e0 = 0.0;
  each_of(false, ident)([&](double A) { wa += A; });
  wg_sum3(wa, ws, e0, red);
  ws = (double)N;
  for (uint32_t it = 0; it < 10000; ++it) {
    m = (fabs(ws) < eps) ? m0 : wa / ws;
    const double d = ((a_hi - m) - (a_lo - m)) + eps;
    const double inv = 1. / d;
    double e1 = 0.0, wa1 = 0.0, ws1 = 0.0;
    each_of(false, ident)([&](double A) {
      const double r = A - m, a = fabs(r * inv);
      const double wq = (a < 0.75) ? 1. : 0.75 / a;
      e1 += fabs(r);
      wa1 += wq * A;
      ws1 += wq;
    });
    wg_sum3(e1, wa1, ws1, red);
    const double err = e1 / (double)N;
    wa = wa1;
    ws = ws1;
    if (fabs(m0 - m) < tol || err < tol) break;
    m0 = m;
  }

  // robust variance: (1.4826 median |lap - median(lap)|)^2
  auto median = [&](auto each) {
    const uint32_t half = N / 2;
    double o0, o1;
    wg_select2(half, N % 2 == 0 ? half - 1 : half, each, sel, o0, o1);
    return (N % 2 == 0) ? o0 + (o1 - o0) / 2.0 : o0;
  };
  const double med = median(each_of(true, ident));
  const double mad = median(each_of(true, [med](double x) { return fabs(x - med); }));
  const double sig = 1.4826 * mad;
  if (threadIdx.x == 0) {
    out[2 * (size_t)blockIdx.x] = m;
    out[2 * (size_t)blockIdx.x + 1] = sig * sig;
  }
}

}  // namespace

hipError_t launch_split_stats(const double* X, uint32_t n, double scale, const uint32_t* frames, uint32_t nf,
                              const uint32_t* blocks, uint32_t nblocks, double* stat, hipStream_t s) {
  if (nblocks == 0 || nf == 0) return hipSuccess;
  hipLaunchKernelGGL(k_split_stats, dim3(nblocks, nf), dim3(256), 0, s, X, n, scale, frames, blocks, nblocks, stat);
  return hipGetLastError();
}

uint32_t leaf_cache_elements() { return 256 * kLeafCache; }

hipError_t launch_leaf_stats(const double* X, uint32_t n, double scale, const uint32_t* leaves, uint32_t nleaves,
                             const uint32_t* big, uint32_t nbig, const uint64_t* boff, double* vals, double* laps,
                             double* out, hipStream_t s) {
  if (nleaves == 0) return hipSuccess;
  if (nbig > 0) {
    // elements per leaf above 256 kLeafCache: at least 17 chunks of 256
    const uint32_t chunks = std::min<uint32_t>((n * n + 255) / 256, 64);
    hipLaunchKernelGGL(k_leaf_prep, dim3(chunks, nbig), dim3(256), 0, s, X, n, scale, leaves, big, boff, vals, laps);
  }
  hipLaunchKernelGGL(k_leaf_stats, dim3(nleaves), dim3(256), 0, s, X, n, scale, leaves, boff, vals, laps, out);
  return hipGetLastError();
}

}  // namespace pgsvt
