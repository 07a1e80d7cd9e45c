// hotpixel.hip -- hot-pixel (outlier) filter of the command-line front end
// (HotPixelFilter, src/hotpixel.hpp:19-64), one workgroup per uint16 frame.
//
// Per frame, as the reference:
//   med  = median of the column medians of the frame (arma::median twice;
//          integer median: odd n -> element n/2, even n -> upper +
//          (lower - upper) / 2 in int arithmetic, i.e. rounded towards the
//          upper middle element);
//   mad  = median of the column medians of |x - med|, times 1 / 0.6745;
//   pixels with |x - med| > threshold * mad are replaced, in column-major
//   index order and in place: inside the frame by the mean of the 4th and
//   5th of the 8 sorted neighbours (truncated to uint16), on the border by
//   med (truncated).
// Deviation: the reference forms |x - med| in uint16 arithmetic, which
// wraps for x < med, and compares against threshold * mad converted to
// uint16, which is undefined for values above 65535; here |x - med| is the
// exact integer distance and the comparison is done in double.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"

namespace pgsvt {
namespace {

// k-th smallest (0-based) of n uint16 values get(i), by one wave: a radix
// select over the high then the low byte with a 256-bin LDS histogram.
template <class Get>
__device__ uint32_t wave_select(Get get, uint32_t n, uint32_t k, uint32_t* hist) {
  const uint32_t lane = threadIdx.x % 64;
  uint32_t prefix = 0;  // selected high byte, then the full value
  for (int pass = 0; pass < 2; ++pass) {
    for (uint32_t b = lane; b < 256; b += 64) hist[b] = 0;
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    for (uint32_t i = lane; i < n; i += 64) {
      const uint32_t v = get(i);
      if (pass == 0) atomicAdd(&hist[v >> 8], 1u);
      else if ((v >> 8) == prefix) atomicAdd(&hist[v & 255], 1u);
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    // bin holding rank k: lane-local sums of 4 bins, then a wave scan
    uint32_t c4 = 0;
    for (int j = 0; j < 4; ++j) c4 += hist[4 * lane + j];
    uint32_t incl = c4;
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t t = __shfl_up(incl, o);
      if (lane >= (uint32_t)o) incl += t;
    }
    const uint32_t excl = incl - c4;
    const bool mine = excl <= k && k < incl;
    const uint64_t who = __ballot(mine);
    const uint32_t src = (uint32_t)__builtin_ctzll(who);
    uint32_t bin = 0, below = 0;
    if (mine) {
      uint32_t acc = excl;
      for (int j = 0; j < 4; ++j) {
        const uint32_t h = hist[4 * lane + j];
        if (k < acc + h) {
          bin = 4 * lane + j;
          below = acc;
          break;
        }
        acc += h;
      }
    }
    bin = __shfl(bin, src);
    below = __shfl(below, src);
    k -= below;
    prefix = (pass == 0) ? bin : (prefix << 8 | bin);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
  }
  return prefix;
}

// Armadillo's median of n integers
template <class Get>
__device__ uint32_t wave_median(Get get, uint32_t n, uint32_t* hist) {
  const uint32_t hi = wave_select(get, n, n / 2, hist);
  if (n % 2) return hi;
  const uint32_t lo = wave_select(get, n, n / 2 - 1, hist);
  return (uint32_t)((int)hi + ((int)lo - (int)hi) / 2);
}

__global__ __launch_bounds__(256) void k_hotpixel(uint16_t* __restrict__ seq, uint32_t nr, uint32_t nc,
                                                  double threshold, uint32_t* __restrict__ colmed_g) {
  __shared__ uint32_t hist[4][256];
  __shared__ uint32_t smed, smad;
  __shared__ uint8_t flag[256];
  uint16_t* F = seq + (size_t)blockIdx.x * nr * nc;
  uint32_t* cm = colmed_g + (size_t)blockIdx.x * nc;  // column medians of this frame
  const uint32_t w = threadIdx.x / 64;
  // frame median: median of the column medians
  for (uint32_t c = w; c < nc; c += 4)
    cm[c] = wave_median([&](uint32_t i) { return (uint32_t)F[i + (size_t)nr * c]; }, nr, hist[w]);
  __syncthreads();
  if (w == 0) {
    const uint32_t m = wave_median([&](uint32_t i) { return cm[i]; }, nc, hist[0]);
    if (threadIdx.x == 0) smed = m;
  }
  __syncthreads();
  const uint32_t med = smed;
  auto dist = [&](uint32_t x) { return x > med ? x - med : med - x; };
  for (uint32_t c = w; c < nc; c += 4)
    cm[c] = wave_median([&](uint32_t i) { return dist(F[i + (size_t)nr * c]); }, nr, hist[w]);
  __syncthreads();
  if (w == 0) {
    const uint32_t m = wave_median([&](uint32_t i) { return cm[i]; }, nc, hist[0]);
    if (threadIdx.x == 0) smad = m;
  }
  __syncthreads();
  const double lim = threshold * ((double)smad * (1.0 / 0.6745));
  // chunks of 256 pixels in index order: flag, then replace one by one
  const uint32_t npix = nr * nc;
  for (uint32_t base = 0; base < npix; base += 256) {
    const uint32_t i = base + threadIdx.x;
    flag[threadIdx.x] = (i < npix && (double)dist(F[i]) > lim) ? 1 : 0;
    __syncthreads();
    if (threadIdx.x == 0) {
      for (uint32_t t = 0; t < 256 && base + t < npix; ++t) {
        if (!flag[t]) continue;
        const uint32_t idx = base + t, r = idx % nr, c = idx / nr;
        if (r > 0 && r < nr - 1 && c > 0 && c < nc - 1) {
          double win[8];
          int q = 0;
          for (int dc = -1; dc <= 1; ++dc)
            for (int dr = -1; dr <= 1; ++dr)
              if (dr || dc) win[q++] = (double)F[(r + dr) + (size_t)nr * (c + dc)];
          for (int a = 1; a < 8; ++a)  // insertion sort of 8 values
            for (int b = a; b > 0 && win[b - 1] > win[b]; --b) {
              const double tmp = win[b];
              win[b] = win[b - 1];
              win[b - 1] = tmp;
            }
          F[idx] = (uint16_t)(0.5 * (win[3] + win[4]));
        } else {
          F[idx] = (uint16_t)med;
        }
      }
    }
    __syncthreads();
  }
}

}  // namespace

hipError_t launch_hotpixel(uint16_t* seq, uint32_t nr, uint32_t nc, uint32_t nf, double threshold,
                           uint32_t* scratch, hipStream_t s) {
  hipLaunchKernelGGL(k_hotpixel, dim3(nf), dim3(256), 0, s, seq, nr, nc, threshold, scratch);
  return hipGetLastError();
}

}  // namespace pgsvt
