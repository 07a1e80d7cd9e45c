// kernels.hip -- gfx950 kernels of the PGURE-SVT patchwise SVT path.
//
// Layouts (all column-major "cubes" as in the reference, element (r, c, f) at
// r + nr*(c + nc*f)):
//   X, Z     double [nf][nc][nr]            input / median-filtered input
//   u64,w64  double [nb][vol]               normalised window (risk, ARPS)
//   u32      float  [nb][nvar][vol]         variant windows (SVD, reconstruct)
//   pos      int32  [nb][nw][p_all]         patch position, row | col << 16
//   W        int32  [nb][vol]               overlap count per voxel
//   factors  float  [nb][nvar][p_sel][fstride]  V^T (n x sr), W^T (n x m), s (n)
//   acc      int64  [nb][nvar][vol]         fixed-point overlap-add sums
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>
#include <type_traits>

#include "common.hpp"
#include "kernels.hpp"

namespace pgsvt {

// ------------------------------------------------------------ utilities --
__global__ void k_to_double(const void* __restrict__ raw, int dtype, double* __restrict__ out, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    double v;
    switch (dtype) {
      case 0: v = (double)((const uint8_t*)raw)[i]; break;
      case 1: v = (double)((const uint16_t*)raw)[i]; break;
      case 2: v = (double)((const float*)raw)[i]; break;
      default: v = ((const double*)raw)[i]; break;
    }
    out[i] = v;
  }
}

hipError_t launch_to_double(const void* raw, int dtype, double* out, size_t n, hipStream_t s) {
  const unsigned grid = (unsigned)std::min<size_t>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(k_to_double, dim3(grid), dim3(256), 0, s, raw, dtype, out, n);
  return hipGetLastError();
}

// max of every frame (order-free, exact)
__global__ void k_frame_max(const double* __restrict__ X, uint32_t fs, double* __restrict__ out) {
  __shared__ double sm[256];
  const double* F = X + (size_t)blockIdx.x * fs;
  double m = -__builtin_inf();
  for (uint32_t i = threadIdx.x; i < fs; i += blockDim.x) m = fmax(m, F[i]);
  sm[threadIdx.x] = m;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = sm[0];
}

hipError_t launch_frame_max(const double* X, uint32_t fs, uint32_t nf, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_frame_max, dim3(nf), dim3(256), 0, s, X, fs, out);
  return hipGetLastError();
}

// ------------------------------------------------------ median pre-filter --
// src/pguresvt.hpp:67-84, src/medfilter.hpp:247-538: each pixel becomes the
// median of the (2r+1)^2 window of the frame converted to uint16 (negative ->
// 0, otherwise truncated), edge pixels replicated.  The window is held in
// registers and the median found by a 16-step radix select.
__device__ __forceinline__ uint32_t to_u16(double v) { return v < 0.0 ? 0u : (uint32_t)(uint16_t)(int64_t)v; }

template <int R>
__global__ __launch_bounds__(256) void k_median(const double* __restrict__ X, double* __restrict__ Z, uint32_t nr,
                                                uint32_t nc) {
  constexpr int K = (2 * R + 1) * (2 * R + 1);
  const uint32_t fs = nr * nc;
  const uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x;
  if (pix >= fs) return;
  const double* F = X + (size_t)blockIdx.y * fs;
  const int r0 = (int)(pix % nr), c0 = (int)(pix / nr);
  uint32_t v[K];
#pragma unroll
  for (int dc = -R; dc <= R; ++dc) {
    const int c = min(max(c0 + dc, 0), (int)nc - 1);
#pragma unroll
    for (int dr = -R; dr <= R; ++dr) {
      const int r = min(max(r0 + dr, 0), (int)nr - 1);
      v[(dc + R) * (2 * R + 1) + (dr + R)] = to_u16(F[r + (size_t)nr * c]);
    }
  }
  uint32_t k = K / 2, prefix = 0, mask = 0;
#pragma unroll 1
  for (int bit = 15; bit >= 0; --bit) {
    const uint32_t b = 1u << bit;
    uint32_t cnt = 0;
#pragma unroll
    for (int q = 0; q < K; ++q) cnt += ((v[q] & (mask | b)) == prefix) ? 1u : 0u;
    if (k >= cnt) {
      k -= cnt;
      prefix |= b;
    }
    mask |= b;
  }
  Z[(size_t)blockIdx.y * fs + pix] = (double)prefix;
}

__global__ void k_median_generic(const double* __restrict__ X, double* __restrict__ Z, uint32_t nr, uint32_t nc,
                                 int R) {
  const uint32_t fs = nr * nc;
  const uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x;
  if (pix >= fs) return;
  const double* F = X + (size_t)blockIdx.y * fs;
  const int r0 = (int)(pix % nr), c0 = (int)(pix / nr);
  const uint32_t K = (2 * R + 1) * (2 * R + 1);
  uint32_t k = K / 2, prefix = 0, mask = 0;
  for (int bit = 15; bit >= 0; --bit) {
    const uint32_t b = 1u << bit;
    uint32_t cnt = 0;
    for (int dc = -R; dc <= R; ++dc) {
      const int c = min(max(c0 + dc, 0), (int)nc - 1);
      for (int dr = -R; dr <= R; ++dr) {
        const int r = min(max(r0 + dr, 0), (int)nr - 1);
        cnt += ((to_u16(F[r + (size_t)nr * c]) & (mask | b)) == prefix) ? 1u : 0u;
      }
    }
    if (k >= cnt) {
      k -= cnt;
      prefix |= b;
    }
    mask |= b;
  }
  Z[(size_t)blockIdx.y * fs + pix] = (double)prefix;
}

hipError_t launch_median(const double* X, double* Z, uint32_t nr, uint32_t nc, uint32_t nf, int r, hipStream_t s) {
  const dim3 grid((nr * nc + 255) / 256, nf);
  switch (r) {
    case 1: hipLaunchKernelGGL(k_median<1>, grid, dim3(256), 0, s, X, Z, nr, nc); break;
    case 2: hipLaunchKernelGGL(k_median<2>, grid, dim3(256), 0, s, X, Z, nr, nc); break;
    case 3: hipLaunchKernelGGL(k_median<3>, grid, dim3(256), 0, s, X, Z, nr, nc); break;
    case 4: hipLaunchKernelGGL(k_median<4>, grid, dim3(256), 0, s, X, Z, nr, nc); break;
    case 5: hipLaunchKernelGGL(k_median<5>, grid, dim3(256), 0, s, X, Z, nr, nc); break;
    default: hipLaunchKernelGGL(k_median_generic, grid, dim3(256), 0, s, X, Z, nr, nc, r); break;
  }
  return hipGetLastError();
}

// ------------------------------------------------- delta2 perturbations --
// src/pgure.hpp:167-186: the first n_vox draws of pcg64(seed) make delta1
// (not needed: its product with eps1 is zero), the next n_vox draws make
// delta2 through bernoulli(vP).  Each thread jumps ahead to its chunk.
__global__ void k_delta2(U128 s0, uint64_t nvox, double vP, uint8_t* __restrict__ out) {
  constexpr uint64_t CH = 64;
  const uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t v0 = c * CH;
  if (v0 >= nvox) return;
  U128 st = pcg_advance(s0, nvox + v0);  // draw j uses the state after j+1 steps
  const uint64_t v1 = min(v0 + CH, nvox);
  for (uint64_t v = v0; v < v1; ++v) {
    st = pcg_step(st);
    out[v] = bernoulli_draw(pcg_output(st), vP) ? 1 : 0;
  }
}

hipError_t launch_delta2(int64_t seed, uint64_t nvox, uint8_t* out, hipStream_t s) {
  const double vP = 0.5 + 0.5 * 1. / std::sqrt(1. * 1. + 4);
  const uint64_t chunks = (nvox + 63) / 64;
  hipLaunchKernelGGL(k_delta2, dim3((unsigned)((chunks + 255) / 256)), dim3(256), 0, s, pcg_seed((uint64_t)seed),
                     nvox, vP, out);
  return hipGetLastError();
}

// ------------------------------------------------------ window volumes --
// src/pguresvt.hpp:100-120 (u = window / max, w likewise for ARPS) and
// src/pgure.hpp:80-82 (U2p = U + delta2*eps2, U2m = U - delta2*eps2).
template <class T>
__global__ void k_prepare(const double* __restrict__ X, const double* __restrict__ Z, const WinMeta* __restrict__ meta,
                          const uint8_t* __restrict__ d2, uint32_t fs, uint32_t vol, uint32_t nvar, double eps2,
                          double d2neg, double d2pos, double* __restrict__ u64, double* __restrict__ w64,
                          T* __restrict__ u32) {
  const uint32_t b = blockIdx.y;
  const WinMeta m = meta[b];
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < vol; v += gridDim.x * blockDim.x) {
    const size_t src = (size_t)m.s0 * fs + v;
    const double u = X[src] / m.uMax;
    u64[(size_t)b * vol + v] = u;
    if (Z) w64[(size_t)b * vol + v] = Z[src] / m.wMax;
    T* o = u32 + (size_t)b * nvar * vol + v;
    o[0] = (T)u;
    if (nvar == 3) {
      const double dl = d2[v] ? d2neg : d2pos;
      o[vol] = (T)(u + dl * eps2);
      o[2 * (size_t)vol] = (T)(u - dl * eps2);
    }
  }
}

hipError_t launch_prepare(const double* X, const double* Z, const WinMeta* meta, uint32_t nb, const uint8_t* d2,
                          const Shape& sh, uint32_t nvar, double eps2, double d2neg, double d2pos, double* u64,
                          double* w64, void* uvar, hipStream_t s) {
  const dim3 grid(std::min<uint32_t>((sh.vol + 255) / 256, 1024), nb);
  if (sh.fp64)
    hipLaunchKernelGGL(k_prepare<double>, grid, dim3(256), 0, s, X, Z, meta, d2, sh.fs, sh.vol, nvar, eps2, d2neg,
                       d2pos, u64, w64, (double*)uvar);
  else
    hipLaunchKernelGGL(k_prepare<float>, grid, dim3(256), 0, s, X, Z, meta, d2, sh.fs, sh.vol, nvar, eps2, d2neg,
                       d2pos, u64, w64, (float*)uvar);
  return hipGetLastError();
}

// reference-frame positions: patches(0,i,aref) = i % (1+nyMbs),
// patches(1,i,aref) = i / (1+nxMbs) (src/arps.hpp:60-64, 85-89, 114-118)
__global__ void k_fill_ref(const WinMeta* __restrict__ meta, uint32_t p_all, uint32_t n1, int32_t* __restrict__ pos,
                           uint32_t nw) {
  const uint32_t b = blockIdx.y;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= p_all) return;
  pos[((size_t)b * nw + meta[b].aref) * p_all + i] = (int32_t)((i % n1) | ((i / n1) << 16));
}

hipError_t launch_fill_ref(const WinMeta* meta, uint32_t nb, const Shape& sh, int32_t* pos, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_ref, dim3((sh.p_all + 255) / 256, nb), dim3(256), 0, s, meta, sh.p_all, sh.n1, pos,
                     sh.nw);
  return hipGetLastError();
}

// ---------------------------------------------------------------- ARPS --
// src/arps.hpp:145-375.  One thread runs one block search (or a dependent
// pair of them).  Costs are FP64 sums of squared differences in Armadillo's
// accu order with contraction disabled, so they round exactly as the
// reference does; ties go to the first index.  The visited map (checkMat)
// lives in LDS, one bit per candidate offset, laid out word-major so that
// consecutive threads hit consecutive banks.
// Search windows up to kArpsMaxMw (2 * 15 + 1 offsets a side) run 128
// threads per block; wider ones, up to kArpsMaxMwWide, 64 threads with a
// larger visited map each.
constexpr int kArpsMaxMw = 15, kArpsMaxMwWide = 31;
template <int MW>
struct ArpsCfg {
  static constexpr int kWords = ((2 * MW + 1) * (2 * MW + 1) + 31) / 32;
  static constexpr int kThreads = MW <= 15 ? 128 : 64;
};

// The reference block: held in registers up to 8 x 8, read from the source
// frame (a cached, unchanging window) above that.
template <int BS, bool REG = (BS <= 8)>
struct ArpsRef {
  double v[BS * BS];
  __device__ void load(const double* S, uint32_t nr) {
#pragma unroll
    for (int q = 0; q < BS * BS; ++q) v[q] = S[(q % BS) + (size_t)nr * (q / BS)];
  }
  __device__ double operator[](int q) const { return v[q]; }
};
template <int BS>
struct ArpsRef<BS, false> {
  const double* S;
  uint32_t nr;
  __device__ void load(const double* S_, uint32_t nr_) {
    S = S_;
    nr = nr_;
  }
  __device__ double operator[](int q) const { return S[(q % BS) + (size_t)nr * (q / BS)]; }
};

template <int BS>
__device__ __forceinline__ double arps_cost(const ArpsRef<BS>& ref, const double* __restrict__ F, uint32_t nr,
                                            int r0, int c0, double inv) {
#pragma clang fp contract(off)
  double a1 = 0.0, a2 = 0.0;
  constexpr int n = BS * BS;
  const double* base = F + r0 + (size_t)nr * c0;
  int q = 0;
#pragma unroll
  for (; q + 1 < n; q += 2) {
    const double d1 = ref[q] - base[(q % BS) + (size_t)nr * (q / BS)];
    a1 += d1 * d1;
    const double d2 = ref[q + 1] - base[((q + 1) % BS) + (size_t)nr * ((q + 1) / BS)];
    a2 += d2 * d2;
  }
  if (q < n) {
    const double d1 = ref[q] - base[(q % BS) + (size_t)nr * (q / BS)];
    a1 += d1 * d1;
  }
  return (a1 + a2) * inv;
}

struct Visit {
  uint32_t* w;
  int stride, W;
  __device__ void clear(int words) {
    for (int k = 0; k < words; ++k) w[k * stride] = 0u;
  }
  __device__ bool get(int r, int c) const {
    const int b = r + W * c;
    return (w[(b >> 5) * stride] >> (b & 31)) & 1u;
  }
  __device__ void set(int r, int c) {
    const int b = r + W * c;
    w[(b >> 5) * stride] |= 1u << (b & 31);
  }
};

template <int BS>
__device__ void arps_search(const double* __restrict__ win, uint32_t nr, uint32_t nc, uint32_t fs, int n1, int it,
                            int src, int dst, int pmy, int pmx, int mw, Visit& vis, int& oy, int& ox) {
  const int i = it % n1, j = it / n1;
  const int Nx = (int)nr, Ny = (int)nc;
  const double inv = 1.0 / (BS * BS);
  ArpsRef<BS> ref;
  ref.load(win + (size_t)src * fs + i + (size_t)nr * j, nr);
  const double* D = win + (size_t)dst * fs;
  const int Wd = 2 * mw + 1;
  vis.clear((Wd * Wd + 31) / 32);
  int x = j, y = i;
  double cost = arps_cost<BS>(ref, D, nr, i, j, inv);
  vis.set(mw, mw);
  int step, maxIdx;
  if (j == 0) {
    step = 2;
    maxIdx = 5;
  } else {
    const int yT = abs(pmy), xT = abs(pmx);
    step = (xT <= yT) ? yT : xT;
    maxIdx = ((yT == 0 && xT == step) || (xT == 0 && yT == step)) ? 5 : 6;
  }
  // LDSP (src/arps.hpp:206-294): points (hor, ver) in index order
  const int lh[6] = {0, -step, 0, step, 0, pmx};
  const int lv[6] = {-step, 0, 0, 0, step, pmy};
  double best = 1e9;
  int bh = lh[0], bv = lv[0];
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    double ck = 1e9;
    if (k == 2) {
      ck = cost;
    } else if (k < maxIdx) {
      const int rv = y + lv[k], rh = x + lh[k];
      const bool skip = (step == 0) || rh < 0 || rv < 0 || (rh + BS - 1) >= Ny || (rv + BS - 1) >= Nx;
      if (!skip) {
        ck = arps_cost<BS>(ref, D, nr, rv, rh, inv);
        const int vr = lv[k] + mw, vc = lh[k] + mw;
        if (vr >= 0 && vr < Wd && vc >= 0 && vc < Wd) vis.set(vr, vc);
      }
    }
    if (k == 0 || ck < best) {
      best = ck;
      bh = lh[k];
      bv = lv[k];
    }
  }
  x += bh;
  y += bv;
  cost = best;
  // SDSP (src/arps.hpp:299-368); it can take at most Wd*Wd moves, since every
  // move goes to a freshly evaluated offset inside the search window.
  const int sh_[5] = {0, -1, 0, 1, 0};
  const int sv_[5] = {-1, 0, 0, 0, 1};
  for (int n = 0; n < Wd * Wd + 2; ++n) {
    double b2 = 1e9;
    int k2 = 0;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      double ck = 1e9;
      if (k == 2) {
        ck = cost;
      } else {
        const int rv = y + sv_[k], rh = x + sh_[k];
        bool skip = rh < 0 || rv < 0 || (rh + BS - 1) >= Ny || (rv + BS - 1) >= Nx || rh < (j - mw) ||
                    rh > (j + mw) || rv < (i - mw) || rv > (i + mw);
        if (!skip) skip = vis.get(y - i + sv_[k] + mw, x - j + sh_[k] + mw);
        if (!skip) {
          ck = arps_cost<BS>(ref, D, nr, rv, rh, inv);
          vis.set(y - i + sv_[k] + mw, x - j + sh_[k] + mw);
        }
      }
      if (k == 0 || ck < b2) {
        b2 = ck;
        k2 = k;
      }
    }
    if (k2 == 2) break;
    x += sh_[k2];
    y += sv_[k2];
    cost = b2;
  }
  oy = y;
  ox = x;
}

// jobs: int16 [nb][max_jobs][4] = (src_a, dst_a, src_b, dst_b); pass b (if
// src_b >= 0) reads the motion pass a left in the same slot.
template <int BS, int MW>
__global__ __launch_bounds__(ArpsCfg<MW>::kThreads) void k_arps(const double* __restrict__ w64,
                                                                const int16_t* __restrict__ jobs, uint32_t max_jobs,
                                                                uint32_t nr, uint32_t nc, uint32_t fs, uint32_t nw,
                                                                uint32_t vol, uint32_t p_all, int n1, int mw,
                                                                int32_t* __restrict__ pos) {
  constexpr int kThreads = ArpsCfg<MW>::kThreads;
  __shared__ uint32_t sv[ArpsCfg<MW>::kWords * kThreads];
  const uint32_t b = blockIdx.z, job = blockIdx.y;
  const int it = blockIdx.x * blockDim.x + threadIdx.x;
  const int16_t* J = jobs + ((size_t)b * max_jobs + job) * 4;
  const int sa = J[0], da = J[1], sb = J[2], db = J[3];
  if (sa < 0 || it >= (int)p_all) return;
  Visit vis{sv + threadIdx.x, kThreads, 2 * mw + 1};
  const double* win = w64 + (size_t)b * vol;
  int y, x;
  arps_search<BS>(win, nr, nc, fs, n1, it, sa, da, 0, 0, mw, vis, y, x);
  pos[((size_t)b * nw + da) * p_all + it] = (int32_t)(y | (x << 16));
  if (sb >= 0) {
    const int i = it % n1, j = it / n1;
    int y2, x2;
    arps_search<BS>(win, nr, nc, fs, n1, it, sb, db, y - i, x - j, mw, vis, y2, x2);
    pos[((size_t)b * nw + db) * p_all + it] = (int32_t)(y2 | (x2 << 16));
  }
}

template <int MW>
hipError_t launch_arps_mw(const double* w64, const int16_t* jobs, uint32_t max_jobs, uint32_t nb, const Shape& sh,
                          int mw, int32_t* pos, hipStream_t s) {
  constexpr int kThreads = ArpsCfg<MW>::kThreads;
  const dim3 grid((sh.p_all + kThreads - 1) / kThreads, max_jobs, nb);
  const dim3 blk(kThreads);
#define PG_ARPS(B)                                                                                              \
  case B:                                                                                                       \
    hipLaunchKernelGGL((k_arps<B, MW>), grid, blk, 0, s, w64, jobs, max_jobs, sh.nr, sh.nc, sh.fs, sh.nw, sh.vol, \
                       sh.p_all, (int)sh.n1, mw, pos);                                                          \
    break;
  switch (sh.bs) {
    PG_ARPS(2) PG_ARPS(3) PG_ARPS(4) PG_ARPS(5) PG_ARPS(6) PG_ARPS(7) PG_ARPS(8) PG_ARPS(16)
    default: return hipErrorInvalidValue;
  }
#undef PG_ARPS
  return hipGetLastError();
}

hipError_t launch_arps(const double* w64, const int16_t* jobs, uint32_t max_jobs, uint32_t nb, const Shape& sh,
                       int mw, int32_t* pos, hipStream_t s) {
  if (mw <= kArpsMaxMw) return launch_arps_mw<kArpsMaxMw>(w64, jobs, max_jobs, nb, sh, mw, pos, s);
  if (mw <= kArpsMaxMwWide) return launch_arps_mw<kArpsMaxMwWide>(w64, jobs, max_jobs, nb, sh, mw, pos, s);
  return hipErrorInvalidValue;
}

// Shrink factor d = g(s)/s of a singular value s of a patch whose largest
// is s0, at lambda (src/svt.hpp:135-139): (s - w)_+ / s with w =
// |s0 exp(-lambda s^2 / 2)| (exponential weighting) or lambda.  One
// expression, never contracted, for every kernel that forms d: the
// evaluation and the gather must agree on each value and on which ranks
// are kept.
template <class T>
__device__ __forceinline__ T shrink_factor(T s, T s0, double lam, int exp_w) {
#pragma clang fp contract(off)
  const T w = exp_w ? fabs(s0 * exp((T)(-0.5 * lam) * (s * s))) : (T)lam;
  return s > T(0) ? fmax(s - w, T(0)) / s : T(0);
}

// --------------------------------------------------- batched Jacobi SVD --
// One-sided (Hestenes) Jacobi on the m x n Casorati matrix of each patch
// (arma::svd_econ at src/svt.hpp:111).  Output per patch, sorted by
// descending singular value: V^T (n x n), W^T = (X V)^T (n x m) and s (n).
// A group of L lanes holds one matrix; lane l owns rows l, l + L, l + 2L ...
// of X (R rows), so the rotation parameters, which every lane of the group
// computes, are shared by R rows.  Column dot products
// are a sum over the lane's rows, then a butterfly across the group (DPP
// inside a 16-lane row, lane swaps for 32/64-lane groups).  Every lane ends
// the butterfly with the bit-identical sum, so all lanes apply the same
// rotation.  Column norms are carried through the sweep with the exact
// Jacobi update (a -= t c, b += t c) and refreshed at the start of every
// sweep, so a pair costs one butterfly.  V is not accumulated: after
// convergence V = X^T W S^-2 (X reloaded), row by row.  Its error in a
// reconstruction d_r w_r v_r^T stays at eps * s_max (that of v_r is
// eps * s_max / s_r, times |w_r| = s_r); ranks with s_r <= eps * s_max get
// v_r = 0, below that level too.
template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xF, 0xF, false));
}

// The empty asm statement on entry makes v an opaque, already rounded
// value.  Without it the compiler contracts the caller's product into the
// first butterfly add, fma(x_i, x_j, neighbour's rounded product), which
// rounds differently in the two lanes of a pair: the lanes then hold
// different sums and apply different rotations (seen as singular values
// off by 1e-2 of s_max on near-degenerate column pairs).
template <int L>
__device__ __forceinline__ float group_sum(float v) {
  asm volatile("" : "+v"(v));
  if (L >= 2) v += dpp_f<0xB1>(v);   // quad_perm [1,0,3,2]
  if (L >= 4) v += dpp_f<0x4E>(v);   // quad_perm [2,3,0,1]
  if (L >= 8) v += dpp_f<0x141>(v);  // row_half_mirror
  if (L >= 16) v += dpp_f<0x140>(v); // row_mirror
  if (L >= 32) v += __shfl_xor(v, 16);
  if (L >= 64) v += __shfl_xor(v, 32);
  return v;
}

// FP64 (the divergence factors and the precision-study path): the same
// DPP steps on both halves of the double; after each step the lanes that
// are paired next hold equal partial sums, so the result is the one of
// the lane-swap (xor) butterfly and bit-identical in every lane
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
  const int2 h = __builtin_bit_cast(int2, v);
  int2 o;
  o.x = __builtin_amdgcn_update_dpp(0, h.x, CTRL, 0xF, 0xF, false);
  o.y = __builtin_amdgcn_update_dpp(0, h.y, CTRL, 0xF, 0xF, false);
  return __builtin_bit_cast(double, o);
}
template <int L>
__device__ __forceinline__ double group_sum(double v) {
  asm volatile("" : "+v"(v));
  if (L >= 2) v += dpp_d<0xB1>(v);   // quad_perm [1,0,3,2]
  if (L >= 4) v += dpp_d<0x4E>(v);   // quad_perm [2,3,0,1]
  if (L >= 8) v += dpp_d<0x141>(v);  // row_half_mirror
  if (L >= 16) v += dpp_d<0x140>(v); // row_mirror
  if (L >= 32) v += __shfl_xor(v, 16);
  if (L >= 64) v += __shfl_xor(v, 32);
  return v;
}

// reciprocal, square root and reciprocal square root: the hardware
// approximations in FP32, IEEE-correct operations in FP64
__device__ __forceinline__ float rcp_(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double rcp_(double x) { return 1.0 / x; }
__device__ __forceinline__ float sqrt_(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ double sqrt_(double x) { return sqrt(x); }
__device__ __forceinline__ float rsq_(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ double rsq_(double x) { return 1.0 / sqrt(x); }

// RT: the window length n = nw is known at run time only (n <= N): the
// matrix is held as N columns, the last N - n zero, and every loop stops at
// n (uniform branches); the stored layout follows n.
template <int BS, int N, int L, class T, bool RT = false>
__global__ __launch_bounds__(256) void k_svd(const T* __restrict__ u32, const int32_t* __restrict__ pos,
                                             const uint32_t* __restrict__ sel, uint32_t nvar, uint32_t nvar_stored,
                                             uint32_t p_sel, uint32_t p_all, uint32_t nr, uint32_t fs, uint32_t nw,
                                             uint32_t vol, uint32_t lim, uint32_t fstride, int max_sweeps, T tol2,
                                             T* __restrict__ factors, uint32_t* __restrict__ stats) {
  constexpr int M = BS * BS, R = (M + L - 1) / L, SRN = (N + 3) & ~3;
  const int n = RT ? (int)nw : N;
  const int SR = RT ? ((n + 3) & ~3) : SRN;
  const uint32_t b = blockIdx.y;
  const int lane = threadIdx.x % L;
  const uint32_t g = (blockIdx.x * blockDim.x + threadIdx.x) / L;
  const uint32_t var = g / p_sel, p = g % p_sel;
  const bool valid = var < nvar;  // uniform per group
  const uint32_t it = valid ? (p_sel == p_all ? p : sel[p]) : 0;  // step 1: identity selection

  const T* U = u32 + ((size_t)b * nvar_stored + (valid ? var : 0)) * vol;
  // the lane's rows of the Casorati matrix X (column k: frame k's patch)
  auto load_x = [&](T (&y)[R][N]) {
#pragma unroll
    for (int k = 0; k < N; ++k) {
      const bool kin = !RT || k < n;
      const int32_t pk = (valid && kin) ? pos[((size_t)b * nw + k) * p_all + it] : 0;
      const uint32_t row = min((uint32_t)(pk & 0xffff), lim), col = min((uint32_t)(pk >> 16), lim);
#pragma unroll
      for (int h = 0; h < R; ++h) {
        const int q = lane + L * h;
        y[h][k] = (valid && kin && q < M) ? U[(size_t)k * fs + (row + q % BS) + (size_t)nr * (col + q / BS)] : T(0);
      }
    }
  };
  T x[R][N];
  load_x(x);

  T a[N];
  int sweep = 0;
  for (; sweep < max_sweeps; ++sweep) {
#pragma unroll
    for (int k = 0; k < N; ++k) {
      T t = x[0][k] * x[0][k];
#pragma unroll
      for (int h = 1; h < R; ++h) t = fma(x[h][k], x[h][k], t);
      a[k] = group_sum<L>(t);
    }
    // sum of t^2 over the sweep: zero exactly when no pair rotated.  A
    // floating-point chain (never reassociated) keeps this in one register.
    T tsum = T(0);
#pragma unroll
    for (int i = 0; i < N - 1; ++i) {
      if (RT && i + 1 >= n) break;
#pragma unroll
      for (int j = i + 1; j < N; ++j) {
        if (RT && j >= n) break;
        // rotation zeroing the (i, j) inner product: zeta = (b - a) / 2c,
        // t = sign(zeta) / (|zeta| + sqrt(1 + zeta^2)), cs = 1/sqrt(1+t^2)
        T cp = x[0][i] * x[0][j];
#pragma unroll
        for (int h = 1; h < R; ++h) cp = fma(x[h][i], x[h][j], cp);
        const T c = group_sum<L>(cp);
        const bool rot = c * c > tol2 * a[i] * a[j];
        // Late in the iteration most pairs rotate in none of the wave's
        // matrices; the update is then the identity (t = 0, cs = 1, sn = 0
        // reproduce every value exactly) and is skipped for the whole wave.
        if (!__any(rot)) continue;
        const T zeta = (a[j] - a[i]) * rcp_(c + c);
        const T den = fabs(zeta) + sqrt_(fma(zeta, zeta, T(1)));
        const T t = rot ? copysign(rcp_(den), zeta) : T(0);
        const T cs = rsq_(fma(t, t, T(1)));
        const T sn = cs * t;
#pragma unroll
        for (int h = 0; h < R; ++h) {
          const T xi = x[h][i], xj = x[h][j];
          x[h][i] = cs * xi - sn * xj;
          x[h][j] = sn * xi + cs * xj;
        }
        a[i] = fma(-t, c, a[i]);
        a[j] = fma(t, c, a[j]);
        tsum = fma(t, t, tsum);
      }
    }
    if (!__any(tsum > T(0))) break;
  }
  if (stats && lane == 0 && valid) atomicMax(stats, (uint32_t)sweep + 1);
  // final column norms (identical in every lane of the group)
#pragma unroll
  for (int k = 0; k < N; ++k) {
    T t = x[0][k] * x[0][k];
#pragma unroll
    for (int h = 1; h < R; ++h) t = fma(x[h][k], x[h][k], t);
    a[k] = group_sum<L>(t);
  }
  if (!valid) return;
  // Store in descending singular-value order, transposed: row r of V^T is
  // the right singular vector of the r-th largest singular value (row
  // stride SR, padded with zeros), row r of W^T is the matching column of
  // W = X V = U S (the rotated Casorati columns), then the sorted values.
  // rank(l) counts the values that sort before column l (ties by index).
  T* F = factors + (((size_t)b * nvar_stored + var) * p_sel + p) * fstride;
  T* FW = F + n * SR;
  T* FS = FW + n * M;
  // ranks (< N <= 32) packed in bytes, four per register: the epilogue
  // keeps X, 1/s^2 and the ranks live, and unpacked ranks would cost the
  // kernel a wave per SIMD
  uint32_t rk8[(N + 3) / 4];
#pragma unroll
  for (int i = 0; i < (N + 3) / 4; ++i) rk8[i] = 0u;
  auto rank_of = [&](int l) -> int { return (int)((rk8[l / 4] >> (8 * (l % 4))) & 0xffu); };
  T inv[N];
  T amax = a[0];
#pragma unroll
  for (int k = 1; k < N; ++k) amax = fmax(amax, a[k]);
  constexpr T eps = std::numeric_limits<T>::epsilon();
#pragma unroll
  for (int l = 0; l < N; ++l) {
    // (RT: the zero columns past n rank after every real one)
    uint32_t rl = 0;
#pragma unroll
    for (int m = 0; m < N; ++m) rl += (a[m] > a[l] || (a[m] == a[l] && m < l)) ? 1u : 0u;
    rk8[l / 4] |= rl << (8 * (l % 4));
    inv[l] = a[l] > eps * eps * amax ? T(1) / a[l] : T(0);
    if (RT && l >= n) continue;
#pragma unroll
    for (int h = 0; h < R; ++h) {
      const int q = lane + L * h;
      if (q < M) FW[rl * M + q] = x[h][l];
    }
    if (lane == 0) FS[rl] = sqrt(a[l]);
  }

  // V^T, column by column: element k of v_l = X^T w_l / s_l^2 from column
  // k of X (reloaded), stored by lane k % L
#pragma unroll
  for (int k = 0; k < SRN; ++k) {
    if (RT && k >= SR) break;
    if (k < n) {
      const int32_t pk = pos[((size_t)b * nw + k) * p_all + it];
      const uint32_t row = min((uint32_t)(pk & 0xffff), lim), col = min((uint32_t)(pk >> 16), lim);
      T xk[R];
#pragma unroll
      for (int h = 0; h < R; ++h) {
        const int q = lane + L * h;
        xk[h] = q < M ? U[(size_t)k * fs + (row + q % BS) + (size_t)nr * (col + q / BS)] : T(0);
      }
#pragma unroll
      for (int l = 0; l < N; ++l) {
        if (RT && l >= n) break;  // uniform: the zero columns
        T pv = xk[0] * x[0][l];
#pragma unroll
        for (int h = 1; h < R; ++h) pv = fma(xk[h], x[h][l], pv);
        const T vk = group_sum<L>(pv) * inv[l];
        const int rl = rank_of(l);
        if (lane == k % L) F[rl * SR + k] = vk;
      }
    } else {
#pragma unroll
      for (int l = 0; l < N; ++l)
        if ((!RT || l < n) && lane == k % L) F[rank_of(l) * SR + k] = T(0);
    }
  }
}

// Static data of the fused evaluation (patches up to 4 x 4), from the U
// variant's factors once the SVD is done: at the patch's bucket place
// j_k = place[b][k][p] of every frame k, its rank-0 row w_0 v_0[k] into W0
// [nb][nw][p_sel][m] and its two largest singular values (s_0, s_1) into
// S01 [nb][nw][p_sel][2].  4 lanes per (patch, frame), frames fastest, so
// a patch's 15 rows read its w_0 from one cache line and each row is one
// line-sized store.  (Inside k_svd these stores would cost it a wave per
// SIMD: its epilogue already holds X, 1/s^2 and the ranks.)  nk = 0: every
// frame; nk > 0: only the frames a gather reads with that frame subset
// (i (nw - 1) / (nk - 1), i < nk) and the window's output frame (mode 1).
template <int BS, class T>
__global__ __launch_bounds__(256) void k_static_rows(const T* __restrict__ factors, const int32_t* __restrict__ place,
                                                     const WinMeta* __restrict__ meta, uint32_t nvar_stored,
                                                     uint32_t p_sel, uint32_t nw, uint32_t nk, uint32_t fstride,
                                                     T* __restrict__ W0, T* __restrict__ S01) {
  constexpr int M = BS * BS, QL = (M + 3) / 4;  // row elements per lane
  const uint32_t b = blockIdx.y;
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t sub = t % 4, row = t / 4;
  const uint32_t nf = nk ? nk + 1 : nw;  // frames written per patch
  const uint32_t p = row / nf, i = row % nf;
  if (p >= p_sel) return;
  const uint32_t k = !nk ? i : (i == nk ? meta[b].ref : (nk > 1 ? i * (nw - 1) / (nk - 1) : nw / 2));
  const uint32_t SR = (nw + 3) & ~3u;
  const T* F = factors + ((size_t)b * nvar_stored * p_sel + p) * fstride;  // the U variant
  const T* FW = F + (size_t)nw * SR;
  const T* FS = FW + (size_t)nw * M;
  const T v0k = F[k];  // V^T row 0, element k
  const int32_t jk = place[((size_t)b * nw + k) * p_sel + p];
  T* o = W0 + (((size_t)b * nw + k) * p_sel + jk) * M;
  if constexpr (M == 16 && sizeof(T) == 4) {  // lane sub: elements 4 sub .. 4 sub + 3
    const float4 w = *reinterpret_cast<const float4*>(FW + 4 * sub);
    *reinterpret_cast<float4*>(o + 4 * sub) = make_float4(w.x * v0k, w.y * v0k, w.z * v0k, w.w * v0k);
  } else {
#pragma unroll
    for (int i = 0; i < QL; ++i) {
      const uint32_t q = sub + 4 * i;
      if (q < (uint32_t)M) o[q] = FW[q] * v0k;
    }
  }
  if (sub == 0) {
    typedef typename std::conditional<sizeof(T) == 4, float2, double2>::type T2;
    T2 sp;
    sp.x = FS[0];
    sp.y = nw > 1 ? FS[1] : T(0);
    reinterpret_cast<T2*>(S01)[((size_t)b * nw + k) * p_sel + jk] = sp;
  }
}

// ---------------------------------------------------- position buckets --
// The overlap-add of src/svt.hpp:148-160 is done as a gather.  For every
// window and frame the selected patches are bucketed by their position in
// that frame (a counting sort, built once per window since positions do
// not depend on lambda).  A voxel then sums, in a fixed order, the
// contributions of the patches in the bs x bs positions that cover it.
__global__ void k_bucket_count(const int32_t* __restrict__ pos, const uint32_t* __restrict__ sel, uint32_t p_sel,
                               uint32_t p_all, uint32_t nw, uint32_t n1, int32_t* __restrict__ cnt) {
  const uint32_t b = blockIdx.z, k = blockIdx.y;
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= p_sel) return;
  const int32_t pk = pos[((size_t)b * nw + k) * p_all + sel[p]];
  const uint32_t r = min((uint32_t)(pk & 0xffff), n1 - 1), c = min((uint32_t)(pk >> 16), n1 - 1);
  atomicAdd(cnt + ((size_t)b * nw + k) * n1 * n1 + r + (size_t)n1 * c, 1);
}

// exclusive scan of one (window, frame) count table into offsets.  Wave w
// owns a contiguous chunk: first every wave sums its chunk (64 lanes on 64
// consecutive counts per step, coalesced), then each wave scans its chunk
// 64 counts per step, starting from the sum of the chunks before it.
template <int W>
__device__ __forceinline__ int32_t wave_incl_scan(int32_t x, uint32_t lane) {
#pragma unroll
  for (int o = 1; o < W; o <<= 1) {
    const int32_t y = __shfl_up(x, o, W);
    x += lane >= (uint32_t)o ? y : 0;
  }
  return x;
}

__global__ __launch_bounds__(1024) void k_bucket_scan(const int32_t* __restrict__ cnt, uint32_t npos,
                                                      int32_t* __restrict__ off) {
  __shared__ int32_t wt[16];
  const size_t tab = (size_t)blockIdx.y * gridDim.x + blockIdx.x;
  const int32_t* C = cnt + tab * npos;
  int32_t* O = off + tab * (npos + 1);
  const uint32_t w = threadIdx.x / 64, lane = threadIdx.x % 64;
  const uint32_t chunk = (npos + 15) / 16;
  const uint32_t lo = min(w * chunk, npos), hi = min(lo + chunk, npos);
  int32_t sum = 0;
  for (uint32_t i = lo + lane; i < hi; i += 64) sum += C[i];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o);
  if (lane == 0) wt[w] = sum;
  __syncthreads();
  int32_t carry = 0;
  for (uint32_t k = 0; k < w; ++k) carry += wt[k];
  for (uint32_t base = lo; base < hi; base += 64) {
    const uint32_t i = base + lane;
    const int32_t x = i < hi ? C[i] : 0;
    const int32_t incl = wave_incl_scan<64>(x, lane);
    if (i < hi) O[i] = carry + incl - x;
    carry += __shfl(incl, 63);
  }
  if (threadIdx.x == 0) {
    int32_t t = 0;
    for (int k = 0; k < 16; ++k) t += wt[k];
    O[npos] = t;
  }
}

__global__ void k_bucket_fill(const int32_t* __restrict__ pos, const uint32_t* __restrict__ sel, uint32_t p_sel,
                              uint32_t p_all, uint32_t nw, uint32_t n1, const int32_t* __restrict__ off,
                              int32_t* __restrict__ fill, int32_t* __restrict__ list) {
  const uint32_t b = blockIdx.z, k = blockIdx.y;
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= p_sel) return;
  const size_t tab = (size_t)b * nw + k;
  const int32_t pk = pos[tab * p_all + sel[p]];
  const uint32_t r = min((uint32_t)(pk & 0xffff), n1 - 1), c = min((uint32_t)(pk >> 16), n1 - 1);
  const size_t q = r + (size_t)n1 * c;
  const int32_t slot = atomicAdd(fill + tab * n1 * n1 + q, 1);
  list[tab * p_sel + off[tab * (n1 * n1 + 1) + q] + slot] = (int32_t)p;
}

// ascending patch order inside every bucket (fixed summation order)
__global__ void k_bucket_sort(const int32_t* __restrict__ off, uint32_t npos, uint32_t p_sel,
                              int32_t* __restrict__ list) {
  const size_t tab = blockIdx.y;
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= npos) return;
  const int32_t* O = off + tab * (npos + 1);
  int32_t* Lst = list + tab * p_sel;
  const int32_t lo = O[q], hi = O[q + 1];
  for (int32_t i = lo + 1; i < hi; ++i) {
    const int32_t v = Lst[i];
    int32_t j = i - 1;
    while (j >= lo && Lst[j] > v) {
      Lst[j + 1] = Lst[j];
      --j;
    }
    Lst[j + 1] = v;
  }
}

// inv[p] = place of patch p in its frame's bucket-sorted order
__global__ void k_bucket_inverse(const int32_t* __restrict__ list, uint32_t p_sel, int32_t* __restrict__ inv) {
  const size_t tab = blockIdx.y;
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= p_sel) return;
  inv[tab * p_sel + list[tab * p_sel + j]] = (int32_t)j;
}

// 32-bit fill (stands in for memset so the launch sequence can be captured
// in a graph)
__global__ void k_fill32(uint32_t* __restrict__ p, uint64_t n, uint32_t v) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// rows of `width` 32-bit words: row r of batch b from src row r * step of
// that batch (batch strides src_batch, dst_batch words)
__global__ void k_copy_rows(const uint32_t* __restrict__ src, uint32_t* __restrict__ dst, uint32_t width,
                            uint32_t step, size_t src_batch, size_t dst_batch) {
  const uint32_t r = blockIdx.x, b = blockIdx.y;
  const uint32_t* S = src + b * src_batch + (size_t)r * step * width;
  uint32_t* D = dst + b * dst_batch + (size_t)r * width;
  for (uint32_t w = threadIdx.x; w < width; w += blockDim.x) D[w] = S[w];
}

hipError_t launch_copy_rows(const void* src, void* dst, uint32_t rows, uint32_t width, uint32_t step,
                            size_t src_batch, size_t dst_batch, uint32_t nb, hipStream_t s) {
  if (rows == 0 || nb == 0) return hipSuccess;
  hipLaunchKernelGGL(k_copy_rows, dim3(rows, nb), dim3(256), 0, s, (const uint32_t*)src, (uint32_t*)dst, width, step,
                     src_batch, dst_batch);
  return hipGetLastError();
}

hipError_t launch_fill32(void* p, uint64_t n, uint32_t v, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const uint32_t blocks = (uint32_t)std::min<uint64_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_fill32, dim3(blocks), dim3(256), 0, s, (uint32_t*)p, n, v);
  return hipGetLastError();
}

hipError_t launch_buckets(const int32_t* pos, const uint32_t* sel, uint32_t nb, const Shape& sh, int32_t* cnt,
                          int32_t* off, int32_t* list, int32_t* inv, hipStream_t s) {
  const uint32_t npos = sh.n1 * sh.n1;
  const size_t tabs = (size_t)nb * sh.nw;
  hipError_t e = launch_fill32(cnt, tabs * npos, 0, s);
  if (e != hipSuccess) return e;
  const dim3 g1((sh.p_sel + 255) / 256, sh.nw, nb);
  hipLaunchKernelGGL(k_bucket_count, g1, dim3(256), 0, s, pos, sel, sh.p_sel, sh.p_all, sh.nw, sh.n1, cnt);
  hipLaunchKernelGGL(k_bucket_scan, dim3(sh.nw, nb), dim3(1024), 0, s, cnt, npos, off);
  e = launch_fill32(cnt, tabs * npos, 0, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_bucket_fill, g1, dim3(256), 0, s, pos, sel, sh.p_sel, sh.p_all, sh.nw, sh.n1, off, cnt, list);
  hipLaunchKernelGGL(k_bucket_sort, dim3((npos + 255) / 256, (unsigned)tabs), dim3(256), 0, s, off, npos, sh.p_sel,
                     list);
  hipLaunchKernelGGL(k_bucket_inverse, dim3((sh.p_sel + 255) / 256, (unsigned)tabs), dim3(256), 0, s, list, sh.p_sel,
                     inv);
  return hipGetLastError();
}

// weights of src/svt.hpp:157-163: W = number of patches covering a voxel
// (from the bucket sizes), stored as 1/W (FP64, 0 where nothing covers it)
// and as delta2/W for the divergence sum, in the reconstruction's type
// (FP32: half the bytes the evaluation reads; its rounding, 6e-8 of each
// term, is far below that of the FP32 reconstructions it multiplies).
template <class D>
__global__ void k_weights(const int32_t* __restrict__ off, const uint8_t* __restrict__ d2, uint32_t nr, uint32_t fs,
                          uint32_t vol, uint32_t n1, uint32_t bs, double d2neg, double d2pos,
                          double* __restrict__ invw, D* __restrict__ dw) {
  const uint32_t b = blockIdx.y;
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= vol) return;
  const uint32_t k = v / fs, pix = v % fs;
  const int r = (int)(pix % nr), c = (int)(pix / nr);
  const int32_t* O = off + ((size_t)b * (vol / fs) + k) * ((size_t)n1 * n1 + 1);
  int32_t w = 0;
  for (int cb = 0; cb < (int)bs; ++cb) {
    const int cc = c - cb;
    if (cc < 0 || cc >= (int)n1) continue;
    for (int ra = 0; ra < (int)bs; ++ra) {
      const int rr = r - ra;
      if (rr < 0 || rr >= (int)n1) continue;
      const size_t q = (size_t)rr + (size_t)n1 * cc;
      w += O[q + 1] - O[q];
    }
  }
  const double iw = w > 0 ? 1.0 / (double)w : 0.0;
  invw[(size_t)b * vol + v] = iw;
  if (dw) dw[(size_t)b * vol + v] = (D)((d2[v] ? d2neg : d2pos) * iw);
}

hipError_t launch_weights(const int32_t* off, const uint8_t* d2, uint32_t nb, const Shape& sh, double d2neg,
                          double d2pos, double* invw, void* dw, hipStream_t s) {
  const dim3 grid((sh.vol + 255) / 256, nb);
  if (sh.fp64)
    hipLaunchKernelGGL(k_weights<double>, grid, dim3(256), 0, s, off, d2, sh.nr, sh.fs, sh.vol, sh.n1, sh.bs, d2neg,
                       d2pos, invw, (double*)dw);
  else
    hipLaunchKernelGGL(k_weights<float>, grid, dim3(256), 0, s, off, d2, sh.nr, sh.fs, sh.vol, sh.n1, sh.bs, d2neg,
                       d2pos, invw, (float*)dw);
  return hipGetLastError();
}

// Lane K of each 16-lane row of the wave, in every lane of that row (DPP
// row_newbcast; FP64 as two 32-bit halves).
template <int K>
__device__ __forceinline__ float row_bcast(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x150 + K, 0xF, 0xF, true));
}
template <int K>
__device__ __forceinline__ double row_bcast(double v) {
  const int2 h = __builtin_bit_cast(int2, v);
  int2 o;
  o.x = __builtin_amdgcn_mov_dpp(h.x, 0x150 + K, 0xF, 0xF, true);
  o.y = __builtin_amdgcn_mov_dpp(h.y, 0x150 + K, 0xF, 0xF, true);
  return __builtin_bit_cast(double, o);
}

// Workgroups of a launch are handed to the 8 XCDs round-robin by linear
// id, and each XCD has its own L2.  xcd_block maps the linear id so that XCD
// x runs one contiguous range of logical (x, y) blocks: neighbouring blocks,
// which read overlapping patches, then meet in the same L2.  A bijection on
// the 2-D grid; returns the logical (x, y).
__device__ __forceinline__ void xcd_block(uint32_t& bx, uint32_t& by) {
  constexpr uint32_t kXcd = 8;
  const uint32_t total = gridDim.x * gridDim.y;
  const uint32_t lin = blockIdx.y * gridDim.x + blockIdx.x;
  const uint32_t x = lin % kXcd, idx = lin / kXcd, q = total / kXcd, r = total % kXcd;
  const uint32_t logical = x * q + min(x, r) + idx;
  bx = logical % gridDim.x;
  by = logical / gridDim.x;
}

// Sum of a per-thread count over the block into *out (a profiling aid: the
// kept ranks of the evaluation); every thread of the block must call it.
__device__ __forceinline__ void block_count(double* out, int v) {
  __shared__ int wsum[4];
  int w = v;
  for (int o = 32; o > 0; o >>= 1) w += __shfl_xor(w, o);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x / 64] = w;
  __syncthreads();
  if (threadIdx.x == 0) *out = (double)(((wsum[0] + wsum[1]) + wsum[2]) + wsum[3]);
}

// --------------------------------------------- threshold + reconstruct --
// One lambda evaluation of the thresholded reconstruction of every patch
// (src/svt.hpp:121-146 with src/utils.hpp:96-106):
//   B = U diag(g(s)) V^T = sum over kept ranks r of (g(s_r)/s_r) w_r v_r^T
// with w_r = X v_r = s_r u_r, stored by the SVD.  Factors are sorted by
// descending s and g(s)/s grows with s, so the kept ranks are a prefix and
// only their rows are read.  Patches above 4 x 4: a group of L = 64 lanes
// owns one patch, lane q one row of B; the U reconstruction's row of every
// frame goes to its
// bucket-sorted place in `B0` ([slot][frame][patch in bucket order][row])
// for the gather.
// nvar_run == 3 (PGURE): the divergence sum sum_v (B1 - 2 B0 + B2)_v
// delta2_v / W_v of src/pgure.hpp:136 comes from the divergence factors of
// the variants U, U + eps2 delta2, U - eps2 delta2 (k_divq) as
// sum_r (d1_r q1_r - 2 d0_r q0_r + d2_r q2_r), in FP64.
// RT: window length n = nw at run time (n <= N).  Patches above 8 x 8 give
// each lane several rows of B (lane q rows q, q + L, ...).
template <int BS, int N, int L, class T, bool RT = false>
__global__ __launch_bounds__(256) void k_eval(
    const T* __restrict__ factors, const WinMeta* __restrict__ meta, const uint32_t* __restrict__ wl,
    uint32_t nvar_stored, uint32_t nvar_run, uint32_t p_sel, uint32_t nw, uint32_t fstride, int exp_w,
    const T* __restrict__ divq, const int32_t* __restrict__ inv, T* __restrict__ B0,
    double* __restrict__ partial, const double* __restrict__ slot_lam, double* __restrict__ kept) {
  constexpr int M = BS * BS, G = 256 / L, SRN = (N + 3) & ~3, RQ = (M + L - 1) / L;
  __shared__ double red[4];
  const int n = RT ? (int)nw : N;
  const int SR = RT ? ((n + 3) & ~3) : SRN;
  uint32_t bx, slot;
  xcd_block(bx, slot);
  const uint32_t b = wl[slot];
  const int grp = threadIdx.x / L, lane = threadIdx.x % L;
  const uint32_t p = bx * G + grp;
  const bool valid = p < p_sel;
  const uint32_t pa = valid ? p : 0;  // addresses of groups past the end: patch 0
  const double lam = slot_lam ? slot_lam[slot] : meta[b].lambda;
  const T* F = factors + ((size_t)b * nvar_stored * p_sel + pa) * fstride;  // the U variant's factors
  const T* FW = F + n * SR;
  const T* FS = FW + n * M;
  const bool pgq = nvar_run == 3;
  // singular values (lane r: s_r) of U, and with PGURE those of the
  // perturbed variants and the divergence factors of all three
  const int lk = min(lane, n - 1);
  const T sl = FS[lk], smax = FS[0];
  T sp[2] = {T(0), T(0)}, spm[2] = {T(0), T(0)}, ql[3] = {T(0), T(0), T(0)};
  if (pgq) {
    // the patch's record: per variant sr singular values, sr factors
    const T* Q = divq + ((size_t)b * p_sel + pa) * nvar_stored * 2 * SR;
#pragma unroll
    for (int v = 0; v < 3; ++v) ql[v] = Q[v * 2 * SR + SR + lk];
#pragma unroll
    for (int v = 0; v < 2; ++v) {
      sp[v] = Q[(v + 1) * 2 * SR + lk];
      spm[v] = Q[(v + 1) * 2 * SR];
    }
  }
  // shrink factor g(s_r)/s_r in lane r < N; kept ranks: 1 + highest lane
  // of the group with g > 0
  auto shrink = [&](T s, T s0) {
    T d = T(0);
    if (valid && lane < n) d = shrink_factor(s, s0, lam, exp_w);
    return d;
  };
  auto kept_ranks = [&](T d) {
    const uint64_t bal = __ballot(d > T(0));
    const uint32_t gbits =
        (uint32_t)((bal >> ((threadIdx.x % 64) / L * L)) & ((L >= 64) ? ~0ull : ((1ull << L) - 1)));
    return gbits ? 32 - __clz(gbits) : 0;
  };
  const T dl = shrink(sl, smax);
  const int nkeep = kept_ranks(dl);
  double s4 = 0.0;
  int kr = nkeep;  // kept ranks of the evaluated variants (profiling)
  if (pgq) {
    // lanes past the kept ranks have d_r = 0
    s4 = -2.0 * (double)dl * (double)ql[0];
#pragma unroll
    for (int v = 0; v < 2; ++v) {
      const T d = shrink(sp[v], spm[v]);
      kr += kept_ranks(d);
      s4 = fma((double)d, (double)ql[v + 1], s4);
    }
  }
  T bq[RQ][N];
#pragma unroll
  for (int h = 0; h < RQ; ++h)
#pragma unroll
    for (int k = 0; k < N; ++k) bq[h][k] = T(0);
  // rank loop, loads of rank r + 1 issued before the products of rank r
  // (FP32: the V^T row as SR/4 float4 loads; FP64: SR scalar loads; with a
  // window length known at run time the loads cover SRN values, within the
  // factor block, and the values past n are never stored)
  constexpr int VW = (sizeof(T) == 4) ? 4 : 1, NV = SRN / VW;
  typedef typename std::conditional<sizeof(T) == 4, float4, double>::type VT;
  T wn[RQ];
  VT vn[NV];
  auto fetch = [&](int r) {
#pragma unroll
    for (int h = 0; h < RQ; ++h) wn[h] = (valid && lane + L * h < M) ? FW[r * M + lane + L * h] : T(0);
    const VT* vr = reinterpret_cast<const VT*>(F + r * SR);
#pragma unroll
    for (int j = 0; j < NV; ++j) vn[j] = vr[j];
  };
  if (nkeep > 0) fetch(0);
  for (int r = 0; r < nkeep; ++r) {
    const T dr = __shfl(dl, r, L);
    T y[RQ];
#pragma unroll
    for (int h = 0; h < RQ; ++h) y[h] = dr * wn[h];
    VT vc[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) vc[j] = vn[j];
    if (r + 1 < nkeep) fetch(r + 1);
#pragma unroll
    for (int h = 0; h < RQ; ++h) {
#pragma unroll
      for (int j = 0; j < NV; ++j) {
        if constexpr (VW == 4) {
          const float4 t = vc[j];
          if (4 * j < N) bq[h][4 * j] = fma(y[h], t.x, bq[h][4 * j]);
          if (4 * j + 1 < N) bq[h][4 * j + 1] = fma(y[h], t.y, bq[h][4 * j + 1]);
          if (4 * j + 2 < N) bq[h][4 * j + 2] = fma(y[h], t.z, bq[h][4 * j + 2]);
          if (4 * j + 3 < N) bq[h][4 * j + 3] = fma(y[h], t.w, bq[h][4 * j + 3]);
        } else {
          if (j < N) bq[h][j] = fma(y[h], (T)vc[j], bq[h][j]);
        }
      }
    }
  }
  if (valid) {
    const int32_t* I = inv + (size_t)b * nw * p_sel + p;
#pragma unroll
    for (int k = 0; k < N; ++k) {
      if (RT && k >= n) break;
      // row of patch p for frame k at its bucket-sorted place
      const size_t at = (((size_t)slot * nw + k) * p_sel + I[(size_t)k * p_sel]) * M;
#pragma unroll
      for (int h = 0; h < RQ; ++h)
        if (lane + L * h < M) B0[at + lane + L * h] = bq[h][k];
    }
  }
  if (!pgq) return;  // uniform over the launch
  for (int o = 32; o > 0; o >>= 1) s4 += __shfl_xor(s4, o);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x / 64] = s4;
  __syncthreads();
  if (threadIdx.x == 0) partial[(size_t)slot * gridDim.x + bx] = ((red[0] + red[1]) + red[2]) + red[3];
  if (kept) block_count(kept + (size_t)slot * gridDim.x + bx, (lane == 0 && valid) ? kr : 0);
}

// Lane K of each 16-lane row, for 32-bit integers (trajectory positions,
// bucket places).
template <int K>
__device__ __forceinline__ int32_t row_bcast(int32_t v) {
  return __builtin_amdgcn_mov_dpp(v, 0x150 + K, 0xF, 0xF, true);
}

// f(std::integral_constant<int, K>) for K = K0 .. K1 - 1, in order
template <int K0, int K1, class F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (K0 < K1) {
    f(std::integral_constant<int, K0>());
    static_for<K0 + 1, K1>(f);
  }
}

// Lane K of the group of L lanes (16: DPP row broadcast, 64: the wave's
// lane K), in every lane of the group.
template <int L, int K, class V>
__device__ __forceinline__ V group_bcast(V v) {
  static_assert(L == 16 || L == 64, "groups of 16 lanes or whole waves");
  if constexpr (L == 16) {
    return row_bcast<K>(v);
  } else if constexpr (sizeof(V) == 8) {
    const int2 h = __builtin_bit_cast(int2, v);
    int2 o;
    o.x = __builtin_amdgcn_readlane(h.x, K);
    o.y = __builtin_amdgcn_readlane(h.y, K);
    return __builtin_bit_cast(V, o);
  } else {
    return __builtin_bit_cast(V, __builtin_amdgcn_readlane(__builtin_bit_cast(int32_t, v), K));
  }
}

// Divergence factors, once per window after the SVD.  The divergence sum of
// src/pgure.hpp:136 is linear in the reconstructions: for
// B = sum_r d_r w_r v_r^T (d_r = g(s_r)/s_r, the only part that depends on
// lambda),
//   sum over the patch's voxels of delta2_v / W_v B_v = sum_r d_r q_r,
//   q_r = w_r^T D v_r,
// D (m x n) the delta2/W volume at the patch's voxels along its trajectory.
// One group of L lanes per (patch, variant): lane q holds row q of D, lane j
// element j of each V^T row; t_q = sum_k D[q][k] v_r[k] and q_r = sum_q
// w_r[q] t_q are formed in FP64 and stored with the variant's singular
// values (zero-padded to sr): the search rounds read s and q of the three
// variants from there.  L = 16: every row is loaded first; L = 64 (larger
// patches, or windows above 16 frames): one rank at a time.  Patches up to
// 4 x 4 (the fused evaluation) store rank-major planes divq
// [nb][nvar][2][sr][p_sel] (s, then q), so that a round reads one coalesced
// value per patch and kept rank; larger patches per-patch records divq
// [nb][p_sel][nvar][2][sr].
// RT: window length n = nw known at run time only (n <= N), as in k_svd.
// Patches above 8 x 8 give each lane several rows of D (lane q rows q,
// q + L, ...).
template <int BS, int N, int L, class T, bool RT = false>
__global__ __launch_bounds__(256) void k_divq(const T* __restrict__ factors, const int32_t* __restrict__ pos,
                                              const uint32_t* __restrict__ sel, const T* __restrict__ dw,
                                              uint32_t nvar, uint32_t p_sel, uint32_t p_all, uint32_t nr,
                                              uint32_t fs, uint32_t nw, uint32_t vol, uint32_t lim,
                                              uint32_t fstride, T* __restrict__ divq) {
  constexpr int M = BS * BS, SRN = (N + 3) & ~3, G = 256 / L, RQ = (M + L - 1) / L;
  static_assert(SRN <= L, "one lane per element of a V^T row");
  const int n = RT ? (int)nw : N;
  const int SR = RT ? ((n + 3) & ~3) : SRN;
  uint32_t bx, b;
  xcd_block(bx, b);
  const int lane = threadIdx.x % L;
  const uint32_t g = bx * G + threadIdx.x / L;  // (variant, patch), patch fastest
  const bool valid = g < nvar * p_sel;
  const uint32_t ga = valid ? g : 0;  // groups past the end read pair 0
  const uint32_t var = ga / p_sel, p = ga % p_sel;
  const uint32_t it = p_sel == p_all ? p : sel[p];
  const T* F = factors + (((size_t)b * nvar + var) * p_sel + p) * fstride;
  const T* FW = F + n * SR;
  const int lk = min(lane, n - 1), lv = min(lane, SR - 1);
  const T sv = FW[n * M + lk];  // s_lane
  // D frame by frame (lane k: frame k's trajectory position)
  const int32_t pk_l = pos[((size_t)b * nw + lk) * p_all + it];
  const T* DW = dw + (size_t)b * vol;
  double d[RQ][N];
  static_for<0, N>([&](auto kc) {
    constexpr int k = decltype(kc)::value;
    const int32_t pkk = group_bcast<L, k>(pk_l);
    const uint32_t r0 = min((uint32_t)(pkk & 0xffff), lim), c0 = min((uint32_t)(pkk >> 16), lim);
#pragma unroll
    for (int h = 0; h < RQ; ++h) {
      const int lw = min(lane + L * h, M - 1);
      d[h][k] = (!RT || k < n) ? (double)DW[(uint32_t)k * fs + (r0 + lw % BS) + nr * (c0 + lw / BS)] : 0.0;
    }
  });
  // q_r of rank r from lane j's V^T element and lane q's W^T elements
  auto factor = [&](T vr, const T (&wr)[RQ]) {
    double t[RQ];
#pragma unroll
    for (int h = 0; h < RQ; ++h) t[h] = 0.0;
    static_for<0, N>([&](auto kc) {
      constexpr int k = decltype(kc)::value;
      if (!RT || k < n) {
        const double vk = (double)group_bcast<L, k>(vr);
#pragma unroll
        for (int h = 0; h < RQ; ++h) t[h] = fma(d[h][k], vk, t[h]);
      }
    });
    double y = 0.0;
#pragma unroll
    for (int h = 0; h < RQ; ++h) y = (lane + L * h < M) ? fma((double)wr[h], t[h], y) : y;
    return group_sum<L>(y);
  };
  auto wrow = [&](int r, T (&wr)[RQ]) {
#pragma unroll
    for (int h = 0; h < RQ; ++h) wr[h] = FW[r * M + min(lane + L * h, M - 1)];
  };
  double qv = 0.0;  // lane r: q_r
  if constexpr (L == 16) {
    T vr[N], wr[N][RQ];
#pragma unroll
    for (int r = 0; r < N; ++r) {
      if (RT && r >= n) break;
      vr[r] = F[r * SR + lv];
      wrow(r, wr[r]);
    }
    static_for<0, N>([&](auto rc) {
      constexpr int r = decltype(rc)::value;
      if (!RT || r < n) {
        const double y = factor(vr[r], wr[r]);
        if (lane == r) qv = y;
      }
    });
  } else {
#pragma unroll 1
    for (int r = 0; r < n; ++r) {
      T wr[RQ];
      wrow(r, wr);
      const double y = factor(F[r * SR + lv], wr);
      if (lane == r) qv = y;
    }
  }
  if (valid && lane < SR) {
    if constexpr (BS <= 4) {
      // rank-major planes [b][var][s | q][rank][patch] for the fused
      // evaluation: a search round reads only the ranks it keeps, one
      // coalesced value per patch
      T* rec = divq + (((size_t)b * nvar + var) * 2 * SR + lane) * p_sel + p;
      rec[0] = lane < n ? sv : T(0);
      rec[(size_t)SR * p_sel] = (T)qv;
    } else {
      T* rec = divq + (((size_t)b * p_sel + p) * nvar + var) * 2 * SR;
      rec[lane] = lane < n ? sv : T(0);
      rec[SR + lane] = (T)qv;
    }
  }
}

// Lambda evaluation for patches up to 4 x 4, per patch (one thread each).
// The reconstruction of a patch is B = sum_r d_r w_r v_r^T (src/svt.hpp:
// 135-146 with src/utils.hpp:96-106), d_r = g(s_r)/s_r, w_r = X v_r = s_r u_r;
// factors are sorted by descending s and g(s)/s grows with s, so the kept
// ranks (d_r > 0) are a prefix.  The gather forms d_0 and whether rank 1 is
// kept itself, from the window's static singular pairs (s_0, s_1) that
// k_svd writes at every bucket place; this thread writes, for a patch that
// keeps rank 1, the factors d_r of its further kept ranks to dr
// [slot][p][r], r >= 1, followed by a 0 when fewer than n ranks are kept.
// The gather forms the patch's rows from d_0 and the static rows (w_0 v_0[k]
// at every bucket place, written by k_svd) and, for further ranks, from dr
// and w_r, v_r of the factors, as it stages them, so the reconstruction
// itself is never written.  With PGURE (qdiv set) the thread also adds the
// patch's divergence sum sum_r (d1_r q1_r - 2 d0_r q0_r + d2_r q2_r)
// (src/pgure.hpp:136, see k_divq) to the FP64 partial of its block.  The
// singular values and divergence factors come from the rank-major planes of
// k_divq, or (qdiv null) the U variant's singular values from its factors.
// RT: window length n = nw at run time (n <= N).  dr null: the divergence
// sums only, e.g. over a subsample of the patches decomposed on their own.
template <int BS, int N, class T, bool RT = false>
__global__ __launch_bounds__(256) void k_round_prep(const T* __restrict__ factors, const T* __restrict__ qdiv,
                                                    const WinMeta* __restrict__ meta,
                                                    const uint32_t* __restrict__ wl, uint32_t nvar_stored,
                                                    uint32_t p_sel, uint32_t nw, uint32_t fstride, int exp_w,
                                                    const double* __restrict__ slot_lam, T* __restrict__ dr,
                                                    double* __restrict__ partial, double* __restrict__ kept) {
  constexpr int M = BS * BS;
  __shared__ double red[4];
  const int n = RT ? (int)nw : N;
  const int SR = (n + 3) & ~3;
  const uint32_t slot = blockIdx.y, b = wl[slot];
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = p < p_sel;
  const uint32_t pa = valid ? p : 0;
  const double lam = slot_lam ? slot_lam[slot] : meta[b].lambda;
  const T* F = factors + ((size_t)b * nvar_stored * p_sel + pa) * fstride;  // the U variant's factors
  const T* FW = F + n * SR;                                                  // W^T rows
  auto sval = [&](int v, int r) -> T {
    return qdiv ? qdiv[(((size_t)b * nvar_stored + v) * 2 * SR + r) * p_sel + pa] : FW[n * M + r];
  };
  auto qval = [&](int v, int r) -> T { return qdiv[(((size_t)b * nvar_stored + v) * 2 * SR + SR + r) * p_sel + pa]; };
  double s4 = 0.0;
  int kr = 0;  // kept ranks of the evaluated variants (profiling)
  const int nvar_run = qdiv ? 3 : 1;
  for (int v = 0; v < nvar_run && valid; ++v) {
    const T s0 = sval(v, 0);
    int r = 0;
    for (; r < n; ++r) {
      const T sr = (r == 0) ? s0 : sval(v, r);
      const T d = shrink_factor(sr, s0, lam, exp_w);
      if (!(d > T(0))) break;
      if (qdiv) s4 = fma((v == 0 ? -2.0 : 1.0) * (double)d, (double)qval(v, r), s4);
      if (v == 0 && dr && r > 0) dr[((size_t)slot * p_sel + p) * nw + r] = d;
    }
    if (v == 0 && dr && r > 1 && r < n) dr[((size_t)slot * p_sel + p) * nw + r] = T(0);  // end of the kept ranks
    kr += r;
  }
  if (!qdiv && !kept) return;  // uniform over the launch
  if (qdiv) {
    for (int o = 32; o > 0; o >>= 1) s4 += __shfl_xor(s4, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x / 64] = s4;
    __syncthreads();
    if (threadIdx.x == 0) partial[(size_t)slot * gridDim.x + blockIdx.x] = ((red[0] + red[1]) + red[2]) + red[3];
  }
  if (kept) block_count(kept + (size_t)slot * gridDim.x + blockIdx.x, kr);
}

// ----------------------------------------------------------- gather --
// Uhat_v = (sum of the contributions covering v) / W_v, contributions taken
// position by position (patch column cb, then row ra) and bucket by bucket
// in ascending patch order, summed in FP64.
//   mode 0: risk sums per block: sum Uhat^2, sum Uhat, sum Uhat U, sum U, sum U^2
//   mode 1: final frame: Y(:,:,out) = uMax * Uhat(:,:,ref)
//   mode 2: whole window volume of slot 0 into out
// One thread per voxel.  For each patch column cb the buckets of positions
// (r-bs+1 .. r, c-cb) are one contiguous run of rows in bucket order, read
// straight from memory (neighbouring voxels share them in cache).
template <int BS, class T>
__global__ __launch_bounds__(256) void k_gather_direct(const T* __restrict__ B0, const int32_t* __restrict__ off,
                                                       const double* __restrict__ invw,
                                                       const double* __restrict__ u64,
                                                       const WinMeta* __restrict__ meta,
                                                       const uint32_t* __restrict__ wl, uint32_t nr, uint32_t fs,
                                                       uint32_t nw, uint32_t vol, uint32_t n1, uint32_t p_sel,
                                                       int mode, double* __restrict__ out,
                                                       double* __restrict__ partial) {
  constexpr int M = BS * BS;
  __shared__ double sm[5][4];
  uint32_t bx, slot;
  xcd_block(bx, slot);
  const uint32_t b = wl[slot];
  const WinMeta m = meta[b];
  const uint32_t span = (mode == 1) ? fs : vol;
  const uint32_t i = bx * blockDim.x + threadIdx.x;
  double a_q = 0, a_h = 0, a_hu = 0, a_u = 0, a_u2 = 0;
  if (i < span) {
    const uint32_t v = (mode == 1) ? m.ref * fs + i : i;
    const uint32_t k = v / fs, pix = v % fs;
    const int r = (int)(pix % nr), c = (int)(pix / nr);
    const int32_t* O = off + ((size_t)b * nw + k) * ((size_t)n1 * n1 + 1);
    const T* Bk = B0 + ((size_t)slot * nw + k) * p_sel * M;
    // Patch columns in chunks of CH (all at once for small patches): bucket
    // bounds of positions (r - ra, c - cb), bnd[.][e] for rows r-BS+1+e
    // (clamped to the frame), e = BS the end of row r; then the first
    // contribution of every bucket in one batch of independent loads, and
    // further ones (buckets holding several patches) in order.
    constexpr int CH = (BS <= 5) ? BS : 1;
    double acc = 0.0;
#pragma unroll
    for (int cb0 = 0; cb0 < BS; cb0 += CH) {
      int32_t bnd[CH][BS + 1];
#pragma unroll
      for (int h = 0; h < CH; ++h) {
        const int pc = c - (cb0 + h);
        const bool in = pc >= 0 && pc < (int)n1;
        const size_t colq = (size_t)n1 * (in ? pc : 0);
#pragma unroll
        for (int e = 0; e <= BS; ++e) {
          const int prc = min(max(r - (BS - 1) + e, 0), (int)n1);
          bnd[h][e] = in ? O[colq + prc] : 0;
        }
      }
      T first[CH][BS];
#pragma unroll
      for (int h = 0; h < CH; ++h)
#pragma unroll
        for (int ra = 0; ra < BS; ++ra) {
          const int e = BS - 1 - ra;  // position r - ra
          const int32_t j = bnd[h][e];
          first[h][ra] = (j < bnd[h][e + 1]) ? Bk[(size_t)j * M + ra + BS * (cb0 + h)] : T(0);
        }
#pragma unroll
      for (int h = 0; h < CH; ++h)
#pragma unroll
        for (int ra = 0; ra < BS; ++ra) {
          const int e = BS - 1 - ra;
          const int32_t j0 = bnd[h][e], j1 = bnd[h][e + 1];
          if (j0 < j1) {
            acc += (double)first[h][ra];
            for (int32_t j = j0 + 1; j < j1; ++j) acc += (double)Bk[(size_t)j * M + ra + BS * (cb0 + h)];
          }
        }
    }
    const double uh = acc * invw[(size_t)b * vol + v];
    if (mode == 0) {
      const double U = u64[(size_t)b * vol + v];
      a_q = uh * uh;
      a_h = uh;
      a_hu = uh * U;
      a_u = U;
      a_u2 = U * U;
    } else if (mode == 1) {
      out[(size_t)m.out * fs + i] = uh * m.uMax;
    } else {
      out[v] = uh;
    }
  }
  if (mode != 0) return;
  // wave reduction (lanes exchange through DPP/permute), then 4 waves
  for (int o = 32; o > 0; o >>= 1) {
    a_q += __shfl_xor(a_q, o);
    a_h += __shfl_xor(a_h, o);
    a_hu += __shfl_xor(a_hu, o);
    a_u += __shfl_xor(a_u, o);
    a_u2 += __shfl_xor(a_u2, o);
  }
  const int w = threadIdx.x / 64;
  if ((threadIdx.x & 63) == 0) {
    sm[0][w] = a_q;
    sm[1][w] = a_h;
    sm[2][w] = a_hu;
    sm[3][w] = a_u;
    sm[4][w] = a_u2;
  }
  __syncthreads();
  if (threadIdx.x < 5)
    partial[((size_t)slot * gridDim.x + bx) * 5 + threadIdx.x] =
        ((sm[threadIdx.x][0] + sm[threadIdx.x][1]) + sm[threadIdx.x][2]) + sm[threadIdx.x][3];
}

// Tiled gather for patches up to 4 x 4, fused with the reconstruction.  A
// block covers a tile of TR x TC voxels of one frame k; the patches that can
// cover the tile sit in one contiguous bucket range per patch column.  The
// block forms those patches' rows for frame k, B[q][k] = d_0 (w_0[q] v_0[k])
// followed by fma(d_r w_r[q], v_r[k], .) over further kept ranks r (d_0 and
// whether rank 1 is kept formed here from the SVD's static singular pairs,
// d_r of further ranks from k_round_prep, w_0 v_0[k] from the SVD's static
// rows, further ranks from the factors; in rank order), and stores them in
// LDS,
// the column ranges packed one after another (TOT rows per tile), one plane
// per patch row (lanes on neighbouring voxels read
// neighbouring words).  Each thread then sums its voxel: first the leading
// patch of each of its bs x bs buckets (one LDS read and a select, no
// branch), then, only in waves where a bucket holds several patches, the
// remaining ones.  The order is fixed, so the sums are reproducible.  A tile
// with more than TOT patch rows (trajectories piling up) forms each
// contribution where it is summed, in the same order.  Blocks of one tile run frame after frame, so that the
// factor rows the tile's patches use in every frame meet in one L2.
//   mode 0: risk sums per block: sum Uhat^2, sum Uhat, sum Uhat U, sum U, sum U^2
//   mode 1: final frame: Y(:,:,out) = uMax * Uhat(:,:,ref)
//   mode 2: whole window volume of slot 0 into out
// 5 waves per SIMD for the FP32 4 x 4 instance (at most 96 VGPRs): its
// latency-bound staging needs the waves more than the registers
// voxel tile of one gather block (rows x columns of a frame); 256 voxels.
// 32 x 8: 32 voxels of a column are contiguous in the column-major frame
// (16 x 16 stages fewer halo rows but measured 10 % slower)
constexpr int kTileRows = 32, kTileCols = 8;
static_assert(kTileRows * kTileCols == 256, "one voxel per thread");

template <int BS, class T>
constexpr int gather_min_waves() {
  return (BS == 4 && sizeof(T) == 4) ? 5 : 1;
}

template <int BS, class T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(gather_min_waves<BS, T>()))) void k_gather_tile(
    const T* __restrict__ factors, const T* __restrict__ dr, const T* __restrict__ S01, const T* __restrict__ W0,
    const int32_t* __restrict__ list, const int32_t* __restrict__ off, const double* __restrict__ invw,
    const double* __restrict__ u64, const WinMeta* __restrict__ meta, const uint32_t* __restrict__ wl,
    const double* __restrict__ slot_lam, int exp_w, uint32_t nr, uint32_t fs, uint32_t nw, uint32_t vol, uint32_t n1,
    uint32_t p_sel, uint32_t fstride, uint32_t nvar_stored, int mode, uint32_t nk, double* __restrict__ out,
    double* __restrict__ partial) {
  // TOT: staged patch rows per tile, the bucket ranges of its NC patch
  // columns packed one after another
  constexpr int M = BS * BS, TR = kTileRows, TC = kTileCols, NC = TC + BS - 1, NRW = TR + BS, TOT = 448;
  constexpr int RS = TOT + 1;                                // plane stride (odd: spreads the planes over banks)
  constexpr int W = (sizeof(T) == 4 && M % 4 == 0) ? 4 : 1;  // row values per thread while staging
  typedef typename std::conditional<W == 4, float4, T>::type ST;
  // sB: staged patch rows, one plane per patch row q (plane q, LDS row j);
  // the pad element of each plane is read (and ignored) by empty buckets.
  // sP: first LDS row of each patch column.
  // sT: bucket offsets of the tile's halo, cell (lc, i) = position (row
  // r0 - (BS-1) + i clamped to [0, n1], column c0 - (BS-1) + lc), all 0 for
  // columns outside the frame (empty buckets).  A voxel's bounds are then
  // cells at fixed offsets from one per-thread base.
  __shared__ T sB[M * RS];
  __shared__ int32_t sT[NC * NRW];
  __shared__ int32_t sP[NC];
  __shared__ double sm[5][4];
  uint32_t bx, slot;
  xcd_block(bx, slot);
  const uint32_t b = wl[slot];
  const WinMeta m = meta[b];
  const double lam = slot_lam ? slot_lam[slot] : m.lambda;
  const uint32_t ntr = (nr + TR - 1) / TR;  // square frames
  // mode 0 covers nk frames of the window: all (nk = nw), or nk spread over
  // it (k_i = i (nw - 1) / (nk - 1)); mode 1 the reference frame
  const uint32_t i = bx % nk;
  const uint32_t k = (mode == 1) ? m.ref : (nk == nw ? i : (nk > 1 ? i * (nw - 1) / (nk - 1) : nw / 2));
  const uint32_t tile = (mode == 1) ? bx : bx / nk;
  const uint32_t tr = tile % ntr, tc = tile / ntr;
  const int r0 = (int)(tr * TR), c0 = (int)(tc * TC);
  const int r = r0 + (int)(threadIdx.x % TR), c = c0 + (int)(threadIdx.x / TR);
  const bool live = r < (int)nr && c < (int)nr;
  const int32_t* O = off + ((size_t)b * nw + k) * ((size_t)n1 * n1 + 1);
  const int32_t* Lk = list + ((size_t)b * nw + k) * p_sel;  // frame k's patches in bucket order
  // bucket place j: the static row w_0 v_0[k] at W0k + j * M, the patch's
  // singular pair (s_0, s_1) at Sk + 2 j; patch id: d_r of further ranks at
  // Ds[id * nw + r] (up to a 0, or r = nw); in the U variant's factors w_r
  // at Ws + id * fstride + r * M (W^T) and v_r[k] at Vk + id * fstride +
  // r * sr (V^T)
  const uint32_t sr = (nw + 3) & ~3u;
  const T* W0k = W0 + ((size_t)b * nw + k) * p_sel * M;
  const T* Sk = S01 + ((size_t)b * nw + k) * p_sel * 2;
  const T* Ds = dr + (size_t)slot * p_sel * nw;
  const T* Vk = factors + (size_t)b * nvar_stored * p_sel * fstride + k;
  const T* Ws = factors + (size_t)b * nvar_stored * p_sel * fstride + nw * sr;
  // the voxel's 1/W and U are independent of the staging: load them first
  const uint32_t pix = (uint32_t)min(r, (int)nr - 1) + nr * (uint32_t)min(c, (int)nr - 1), v = k * fs + pix;
  const double iw = invw[(size_t)b * vol + v];
  const double U = (mode == 0) ? u64[(size_t)b * vol + v] : 0.0;
  {
    constexpr int NO = (NC * NRW + 255) / 256;
    int32_t ov[NO];
#pragma unroll
    for (int s = 0; s < NO; ++s) {
      const int e = min((int)threadIdx.x + 256 * s, NC * NRW - 1);
      const int col = c0 - (BS - 1) + e / NRW, row = min(max(r0 - (BS - 1) + e % NRW, 0), (int)n1);
      const int32_t o = O[(size_t)n1 * min(max(col, 0), (int)n1 - 1) + row];
      ov[s] = (col >= 0 && col < (int)n1) ? o : 0;
    }
#pragma unroll
    for (int s = 0; s < NO; ++s) {
      const int e = (int)threadIdx.x + 256 * s;
      if (e < NC * NRW) sT[e] = ov[s];
    }
  }
  __syncthreads();
  // column ranges, packed: column lc's rows start at LDS row pre[lc]
  int32_t pre[NC + 1];
  pre[0] = 0;
#pragma unroll
  for (int lc = 0; lc < NC; ++lc) pre[lc + 1] = pre[lc] + (sT[lc * NRW + NRW - 1] - sT[lc * NRW]);
  const bool fits = pre[NC] <= TOT;
#pragma unroll
  for (int lc = 0; lc < NC; ++lc)
    if ((int)threadIdx.x == lc) sP[lc] = pre[lc];
  if (fits) {
    // the tile's rows in LDS order (row R of column lc is bucket place
    // sT[lc][0] + R - pre[lc]): thread -> (rows lr0 + RPI s, values q ..
    // q + W - 1), NS slots.  Every row's loads are issued first, then every
    // row is formed.
    constexpr int QW = M / W, RPI = 256 / QW, NS = (TOT + RPI - 1) / RPI;
    const int lr0 = (int)threadIdx.x / QW, q = ((int)threadIdx.x % QW) * W;
    const int tot = pre[NC];
    // Every load of a row is at its bucket place (the patch index, the
    // thread's values of the static row, the singular pair), issued for
    // every row before the first use; further ranks are read as needed.
    // Row value q: d_0 (w_0[q] v_0[k]), then the FMA chain of the
    // reconstruction over the further kept ranks, fma(d_r w_r[q], v_r[k], .).
    int32_t ids[NS];
    ST w0[NS];
    T s0[NS], s1[NS];
    // 4 x 4 patches in FP32: the 4 lanes of a row are one quad; lanes 0, 2
    // and 3 of the quad load the row's patch index, s_0 and s_1 (the loads
    // in flight hold one register per row instead of three).  After the
    // wait the slots go in pairs (s, s + 1): lanes 2 and 3 form d_0 =
    // shrink(s_0, s_0) and shrink(s_1, s_0) of the first row, lanes 0 and 1
    // those of the second (its s_0, s_1 moved down by DPP), in the same
    // instructions, and the quad shares the index and the factors by DPP
    // broadcasts.
    constexpr bool QUAD = (QW == 4 && sizeof(T) == 4);
    uint32_t sc[NS];
    const int sub = (int)threadIdx.x & 3;
#pragma unroll
    for (int sl = 0; sl < NS; ++sl) {
      const int R = lr0 + RPI * sl;
      const bool use = lr0 < RPI && R < tot;
      // the row's column: lc = number of columns after the first that start
      // at or before R, base = pre[lc]
      int lc = 0, base = 0;
#pragma unroll
      for (int cc = 1; cc < NC; ++cc) {
        const bool after = pre[cc] <= R;
        lc += after ? 1 : 0;
        base = after ? pre[cc] : base;
      }
      const int32_t j = use ? sT[lc * NRW] + (R - base) : 0;
      w0[sl] = *reinterpret_cast<const ST*>(W0k + (size_t)j * M + q);
      if constexpr (QUAD) {
        const void* a = sub == 0 ? (const void*)(Lk + j) : (const void*)(Sk + 2 * (size_t)j + (sub - 2));
        sc[sl] = (use && sub != 1) ? *reinterpret_cast<const uint32_t*>(a) : (sub == 0 ? 0xffffffffu : 0u);
      } else {
        ids[sl] = use ? Lk[j] : -1;
        s0[sl] = Sk[2 * (size_t)j];
        s1[sl] = Sk[2 * (size_t)j + 1];
      }
    }
    T dn0 = T(0), dn1 = T(0);  // the second row's factors of a pair (QUAD)
#pragma unroll
    for (int sl = 0; sl < NS; ++sl) {
      {
        T d0, d1;
        if constexpr (QUAD) {  // every lane of the quad is active here
          const int x = (int)sc[sl];
          ids[sl] = __builtin_amdgcn_mov_dpp(x, 0x00, 0xF, 0xF, false);  // quad_perm [0,0,0,0]
          if (sl % 2 == 1) {  // formed with the row before
            d0 = dn0;
            d1 = dn1;
          } else {
            // lanes 2, 3: this row's s_0, s_1; lanes 0, 1: the next row's
            // (quad_perm [2,3,2,3]), or unused values for the last row
            int v = x, sa = __builtin_amdgcn_mov_dpp(x, 0xAA, 0xF, 0xF, false);  // [2,2,2,2]
            if (sl + 1 < NS) {
              const int y = (int)sc[sl + 1];
              const int yv = __builtin_amdgcn_mov_dpp(y, 0xEE, 0xF, 0xF, false);
              const int ya = __builtin_amdgcn_mov_dpp(y, 0xAA, 0xF, 0xF, false);
              v = sub < 2 ? yv : x;
              sa = sub < 2 ? ya : sa;
            }
            const int d = __builtin_bit_cast(
                int, shrink_factor(__builtin_bit_cast(T, v), __builtin_bit_cast(T, sa), lam, exp_w));
            d0 = __builtin_bit_cast(T, __builtin_amdgcn_mov_dpp(d, 0xAA, 0xF, 0xF, false));  // [2,2,2,2]
            d1 = __builtin_bit_cast(T, __builtin_amdgcn_mov_dpp(d, 0xFF, 0xF, 0xF, false));  // [3,3,3,3]
            if (sl + 1 < NS) {
              dn0 = __builtin_bit_cast(T, __builtin_amdgcn_mov_dpp(d, 0x00, 0xF, 0xF, false));
              dn1 = __builtin_bit_cast(T, __builtin_amdgcn_mov_dpp(d, 0x55, 0xF, 0xF, false));  // [1,1,1,1]
            }
          }
        } else {
          d0 = shrink_factor(s0[sl], s0[sl], lam, exp_w);
          d1 = shrink_factor(s1[sl], s0[sl], lam, exp_w);
        }
        if (ids[sl] >= 0) {
          const int32_t id = ids[sl];
          ST row;
          T* ac = reinterpret_cast<T*>(&row);
          const T* ww = reinterpret_cast<const T*>(&w0[sl]);
#pragma unroll
          for (int w = 0; w < W; ++w) ac[w] = (d0 > T(0)) ? d0 * ww[w] : T(0);
          if (d1 > T(0)) {
            for (int rk = 1; rk < (int)nw; ++rk) {
              const T d = Ds[(size_t)id * nw + rk];
              if (!(d > T(0))) break;
              const T vk = Vk[(size_t)id * fstride + (size_t)rk * sr];
              const ST wv = *reinterpret_cast<const ST*>(Ws + (size_t)id * fstride + (size_t)rk * M + q);
              const T* wk = reinterpret_cast<const T*>(&wv);
#pragma unroll
              for (int w = 0; w < W; ++w) ac[w] = fma(d * wk[w], vk, ac[w]);
            }
          }
          const int R = lr0 + RPI * sl;
#pragma unroll
          for (int w = 0; w < W; ++w) sB[(q + w) * RS + R] = ac[w];
        }
      }
    }
  }
  __syncthreads();
  double a_q = 0, a_h = 0, a_hu = 0, a_u = 0, a_u2 = 0;
  if (live) {
    double acc = 0.0;
    // bucket bounds bnd[cb][e] of positions (r - BS + 1 + e, c - cb); e = BS
    // is the end of row r.  cb0: the voxel's column base cell.
    const int tb = ((c - c0) + BS - 1) * NRW + (r - r0);
    int32_t bnd[BS][BS + 1];
    int base[BS];  // bucket index j -> LDS row j + base[cb]
#pragma unroll
    for (int cb = 0; cb < BS; ++cb) {
#pragma unroll
      for (int e = 0; e <= BS; ++e) bnd[cb][e] = sT[tb - cb * NRW + e];
      base[cb] = sP[(c - c0) + BS - 1 - cb] - sT[tb - (r - r0) - cb * NRW];
    }
    if (fits) {
      bool multi = false;
#pragma unroll
      for (int cb = 0; cb < BS; ++cb) {
#pragma unroll
        for (int ra = 0; ra < BS; ++ra) {
          const int e = BS - 1 - ra;  // position r - ra
          const int32_t j0 = bnd[cb][e], n = bnd[cb][e + 1] - j0;
          // an empty bucket reads the first row of the next bucket or the
          // plane's pad element: in bounds, and not added
          const T f = sB[(ra + BS * cb) * RS + j0 + base[cb]];
          acc += (double)((n > 0) ? f : T(0));
          multi = multi || n > 1;
        }
      }
      if (__any(multi)) {
#pragma unroll
        for (int cb = 0; cb < BS; ++cb) {
#pragma unroll
          for (int ra = 0; ra < BS; ++ra) {
            const int e = BS - 1 - ra;
            for (int32_t j = bnd[cb][e] + 1; j < bnd[cb][e + 1]; ++j)
              acc += (double)sB[(ra + BS * cb) * RS + j + base[cb]];
          }
        }
      }
    } else {
#pragma unroll
      for (int cb = 0; cb < BS; ++cb)
#pragma unroll
        for (int ra = 0; ra < BS; ++ra) {
          const int e = BS - 1 - ra, qv = ra + BS * cb;
          for (int32_t j = bnd[cb][e]; j < bnd[cb][e + 1]; ++j) {
            // element qv of the patch's row, formed as in the staging
            const int32_t id = Lk[j];
            const T sa = Sk[2 * (size_t)j], sb = Sk[2 * (size_t)j + 1];
            const T d0 = shrink_factor(sa, sa, lam, exp_w);
            T x = (d0 > T(0)) ? d0 * W0k[(size_t)j * M + qv] : T(0);
            if (shrink_factor(sb, sa, lam, exp_w) > T(0)) {
              for (int rk = 1; rk < (int)nw; ++rk) {
                const T d = Ds[(size_t)id * nw + rk];
                if (!(d > T(0))) break;
                x = fma(d * Ws[(size_t)id * fstride + (size_t)rk * M + qv], Vk[(size_t)id * fstride + (size_t)rk * sr],
                        x);
              }
            }
            acc += (double)x;
          }
        }
    }
    const double uh = acc * iw;
    if (mode == 0) {
      a_q = uh * uh;
      a_h = uh;
      a_hu = uh * U;
      a_u = U;
      a_u2 = U * U;
    } else if (mode == 1) {
      out[(size_t)m.out * fs + pix] = uh * m.uMax;
    } else {
      out[v] = uh;
    }
  }
  if (mode != 0) return;
  for (int o = 32; o > 0; o >>= 1) {
    a_q += __shfl_xor(a_q, o);
    a_h += __shfl_xor(a_h, o);
    a_hu += __shfl_xor(a_hu, o);
    a_u += __shfl_xor(a_u, o);
    a_u2 += __shfl_xor(a_u2, o);
  }
  const int w = threadIdx.x / 64;
  if ((threadIdx.x & 63) == 0) {
    sm[0][w] = a_q;
    sm[1][w] = a_h;
    sm[2][w] = a_hu;
    sm[3][w] = a_u;
    sm[4][w] = a_u2;
  }
  __syncthreads();
  if (threadIdx.x < 5)
    partial[((size_t)slot * gridDim.x + bx) * 5 + threadIdx.x] =
        ((sm[threadIdx.x][0] + sm[threadIdx.x][1]) + sm[threadIdx.x][2]) + sm[threadIdx.x][3];
}

// per slot: (sum Uhat^2, sum U, sum U^2, sum Uhat U, sum d2 (U2p - 2 Uhat + U2m), sum Uhat)
__global__ __launch_bounds__(256) void k_risk_reduce(const double* __restrict__ pa, uint32_t na,
                                                     const double* __restrict__ pb, uint32_t nb_,
                                                     double* __restrict__ sums, const double* __restrict__ pk,
                                                     double* __restrict__ kept) {
  __shared__ double sm[6][256];
  const uint32_t slot = blockIdx.x;
  if (pk) {  // kept ranks of the slot (profiling): the blocks' counts
    double c = 0.0;
    for (uint32_t k = threadIdx.x; k < na; k += blockDim.x) c += pk[(size_t)slot * na + k];
    sm[0][threadIdx.x] = c;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) sm[0][threadIdx.x] += sm[0][threadIdx.x + o];
      __syncthreads();
    }
    if (threadIdx.x == 0) kept[slot] = sm[0][0];
    __syncthreads();
  }
  double s[6] = {0, 0, 0, 0, 0, 0};
  for (uint32_t k = threadIdx.x; k < nb_; k += blockDim.x) {
    const double* P = pb + ((size_t)slot * nb_ + k) * 5;
    s[0] += P[0];  // sum Uhat^2
    s[5] += P[1];  // sum Uhat
    s[3] += P[2];  // sum Uhat U
    s[1] += P[3];  // sum U
    s[2] += P[4];  // sum U^2
  }
  for (uint32_t k = threadIdx.x; k < na; k += blockDim.x) s[4] += pa[(size_t)slot * na + k];
  for (int c = 0; c < 6; ++c) sm[c][threadIdx.x] = s[c];
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o)
      for (int c = 0; c < 6; ++c) sm[c][threadIdx.x] += sm[c][threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x < 6) sums[(size_t)slot * 6 + threadIdx.x] = sm[threadIdx.x][0];
}

// shape table of the instantiated SVD / evaluation kernels (patch size,
// window frames, lanes per patch in the evaluation, lanes per matrix in the
// SVD)
#define PG_SHAPES(X) \
  X(4, 3, 16, 4) X(4, 5, 16, 4) X(4, 7, 16, 4) X(4, 9, 16, 4) X(4, 11, 16, 4) X(4, 13, 16, 4) X(4, 15, 16, 4) \
  X(2, 3, 16, 4) X(3, 3, 16, 4) X(3, 5, 16, 4) X(3, 7, 16, 4) X(3, 9, 16, 4) \
  X(6, 7, 64, 16) X(6, 15, 64, 16) X(6, 31, 64, 32) X(8, 7, 64, 16) X(8, 15, 64, 16) X(8, 31, 64, 32)

// second lanes-per-matrix mapping of the SVD for each shape, timed against
// the first by the autotuner of the batched path
#define PG_SVD_ALT(X) \
  X(4, 3, 8) X(4, 5, 8) X(4, 7, 8) X(4, 9, 8) X(4, 11, 8) X(4, 13, 8) X(4, 15, 8) \
  X(2, 3, 2) X(3, 3, 8) X(3, 5, 8) X(3, 7, 8) X(3, 9, 8) \
  X(6, 7, 32) X(6, 15, 32) X(6, 31, 64) X(8, 7, 32) X(8, 15, 32) X(8, 31, 64)

// Every other shape, FP32 only: the window length is known at run time
// (the kernels' RT form) up to NM frames.  (patch size, NM, lanes per
// matrix in the SVD, lanes per patch in the divergence factors and, for
// patches above 4 x 4, the evaluation; patches up to 4 x 4 evaluate fused,
// one thread per patch)
#define PG_GENERIC(X) \
  X(2, 16, 4, 16) X(3, 16, 4, 16) X(4, 16, 4, 16) \
  X(5, 16, 16, 64) X(5, 32, 16, 64) X(6, 16, 16, 64) X(6, 32, 32, 64) X(7, 16, 16, 64) X(7, 32, 32, 64) \
  X(8, 16, 32, 64) X(8, 32, 64, 64) X(16, 16, 64, 64)

bool shape_exact(uint32_t bs, uint32_t nw) {
#define PG_CHK(B, N, L, LS) \
  if (bs == B && nw == N) return true;
  PG_SHAPES(PG_CHK)
#undef PG_CHK
  return false;
}

// the generic family that runs a shape: its NM, or 0
uint32_t shape_generic(uint32_t bs, uint32_t nw) {
  if (nw < 1) return 0;
#define PG_GCHK(B, NM, L, LE) \
  if (bs == B && nw <= NM) return NM;
  PG_GENERIC(PG_GCHK)
#undef PG_GCHK
  return 0;
}

uint32_t svd_lanes(uint32_t bs, uint32_t nw, int which) {
#define PG_L1(B, N, LE, L) \
  if (which == 0 && bs == B && nw == N) return L;
  PG_SHAPES(PG_L1)
#undef PG_L1
#define PG_L2(B, N, L) \
  if (which == 1 && bs == B && nw == N) return L;
  PG_SVD_ALT(PG_L2)
#undef PG_L2
  return 0;
}

bool svd_shape_supported(uint32_t bs, uint32_t nw, bool fp64) {
  return shape_exact(bs, nw) || (!fp64 && shape_generic(bs, nw) != 0);
}

hipError_t launch_svd(const void* uvar, const int32_t* pos, const uint32_t* sel, uint32_t nb, uint32_t nvar,
                      uint32_t nvar_stored, const Shape& sh, void* factors, int max_sweeps, double tol,
                      uint32_t* stats, hipStream_t s, int mapping) {
  if (nvar > nvar_stored) return hipErrorInvalidValue;
  const uint32_t lim = sh.nr - sh.bs;
  const uint32_t groups = nvar * sh.p_sel;
#define PG_SVD_F32(B, N, L)                                                                                      \
  {                                                                                                            \
    const dim3 grid((groups * L + 255) / 256, nb);                                                             \
    hipLaunchKernelGGL((k_svd<B, N, L, float>), grid, dim3(256), 0, s, (const float*)uvar, pos, sel, nvar,     \
                       nvar_stored, sh.p_sel, sh.p_all, sh.nr, sh.fs, sh.nw, sh.vol, lim, sh.fstride, max_sweeps,            \
                       (float)(tol * tol), (float*)factors, stats);                                                \
    return hipGetLastError();                                                                                  \
  }
#define PG_SVD(B, N, LE, L)                                                                                      \
  if (sh.bs == B && sh.nw == N && (sh.fp64 || mapping == 0)) {                                                 \
    if (sh.fp64) {                                                                                             \
      /* FP64: at least 16 lanes per matrix (fewer rows per lane fit the registers); the 64-row matrices     \
         of patch 8 take 32 (15 frames) and 64 (31 frames) */                                                  \
      constexpr int LD = (B == 8 && N == 31) ? 64 : (B == 8 && N == 15) ? 32 : (L < 16 ? 16 : L);            \
      const dim3 grid((groups * LD + 255) / 256, nb);                                                          \
      hipLaunchKernelGGL((k_svd<B, N, LD, double>), grid, dim3(256), 0, s, (const double*)uvar, pos, sel, nvar, \
                         nvar_stored, sh.p_sel, sh.p_all, sh.nr, sh.fs, sh.nw, sh.vol, lim, sh.fstride, max_sweeps, tol * tol, \
                         (double*)factors, stats);                                                               \
      return hipGetLastError();                                                                                \
    }                                                                                                          \
    PG_SVD_F32(B, N, L)                                                                                        \
  }
  PG_SHAPES(PG_SVD)
#undef PG_SVD
#define PG_SVD2(B, N, L)                                                                                         \
  if (sh.bs == B && sh.nw == N && mapping == 1) PG_SVD_F32(B, N, L)
  PG_SVD_ALT(PG_SVD2)
#undef PG_SVD2
  // an exact shape without an alternative mapping: the primary one
#define PG_SVD3(B, N, LE, L)                                                                                     \
  if (sh.bs == B && sh.nw == N && !sh.fp64) PG_SVD_F32(B, N, L)
  PG_SHAPES(PG_SVD3)
#undef PG_SVD3
#undef PG_SVD_F32
  if (sh.fp64) return hipErrorInvalidValue;
  const uint32_t nm = shape_generic(sh.bs, sh.nw);
#define PG_SVDG(B, NM, L, LE)                                                                                    \
  if (sh.bs == B && nm == NM) {                                                                                \
    const dim3 grid((groups * L + 255) / 256, nb);                                                             \
    hipLaunchKernelGGL((k_svd<B, NM, L, float, true>), grid, dim3(256), 0, s, (const float*)uvar, pos, sel,    \
                       nvar, nvar_stored, sh.p_sel, sh.p_all, sh.nr, sh.fs, sh.nw, sh.vol, lim, sh.fstride, max_sweeps,      \
                       (float)(tol * tol), (float*)factors, stats);                                                \
    return hipGetLastError();                                                                                  \
  }
  PG_GENERIC(PG_SVDG)
#undef PG_SVDG
  return hipErrorInvalidValue;
}

hipError_t launch_static_rows(const void* factors, const int32_t* place, const WinMeta* meta, uint32_t nb,
                              uint32_t nvar_stored, const Shape& sh, void* W0, void* S01, hipStream_t s, uint32_t nk) {
  if (!fused_eval(sh) || nvar_stored == 0) return hipErrorInvalidValue;
  if (nk >= sh.nw) nk = 0;
  const uint32_t nf = nk ? nk + 1 : sh.nw;
  const dim3 grid((uint32_t)(((size_t)sh.p_sel * nf * 4 + 255) / 256), nb);
#define PG_SR(B)                                                                                                 \
  case B:                                                                                                      \
    if (sh.fp64)                                                                                               \
      hipLaunchKernelGGL((k_static_rows<B, double>), grid, dim3(256), 0, s, (const double*)factors, place, meta, \
                         nvar_stored, sh.p_sel, sh.nw, nk, sh.fstride, (double*)W0, (double*)S01);              \
    else                                                                                                       \
      hipLaunchKernelGGL((k_static_rows<B, float>), grid, dim3(256), 0, s, (const float*)factors, place, meta,   \
                         nvar_stored, sh.p_sel, sh.nw, nk, sh.fstride, (float*)W0, (float*)S01);                \
    return hipGetLastError();
  switch (sh.bs) {
    PG_SR(2) PG_SR(3) PG_SR(4)
    default: return hipErrorInvalidValue;
  }
#undef PG_SR
}

bool fused_eval(const Shape& sh) { return sh.bs <= 4; }

uint32_t factor_stride(uint32_t nw, uint32_t m) {
  const uint32_t sr = (nw + 3) & ~3u;
  return ((nw * sr + nw * m + nw + 31) & ~31u) | 32u;
}

uint32_t eval_blocks(const Shape& sh, uint32_t nvar_run) {
  (void)nvar_run;  // one group (one thread with the fused gather) per patch either way
  if (fused_eval(sh)) return (sh.p_sel + 255) / 256;
  return (sh.p_sel + 3) / 4;  // one 64-lane group per patch
}

hipError_t launch_evaluation(const Shape& sh, const void* factors, const void* divq, const WinMeta* meta,
                             const uint32_t* wl, uint32_t na, uint32_t nvar_stored, uint32_t nvar_run, int exp_w,
                             const int32_t* inv, void* rows, double* partial, const double* slot_lam, double* kept,
                             hipStream_t s) {
  if (fused_eval(sh))
    return launch_round_prep(factors, nvar_run == 3 ? divq : nullptr, meta, wl, na, nvar_stored, sh, exp_w,
                             slot_lam, rows, nvar_run == 3 ? partial : nullptr, kept, s);
  return launch_eval(factors, meta, wl, na, nvar_stored, nvar_run, sh, exp_w, divq, inv, rows, partial, s, slot_lam,
                     kept);
}

size_t singular_pair_bytes(const Shape& sh) { return (size_t)sh.nw * sh.p_sel * 2 * (sh.fp64 ? 8 : 4); }

hipError_t launch_round_prep(const void* factors, const void* qdiv, const WinMeta* meta, const uint32_t* wl,
                             uint32_t na, uint32_t nvar_stored, const Shape& sh, int exp_w, const double* slot_lam,
                             void* dr, double* partial, double* kept, hipStream_t s) {
  const uint32_t nblk = eval_blocks(sh, 1);
#define PG_RP(B, N, L, LS)                                                                                       \
  if (sh.bs == B && sh.nw == N) {                                                                              \
    if constexpr (B <= 4) {                                                                                    \
      if (sh.fp64)                                                                                             \
        hipLaunchKernelGGL((k_round_prep<B, N, double>), dim3(nblk, na), dim3(256), 0, s,                      \
                           (const double*)factors, (const double*)qdiv, meta, wl, nvar_stored, sh.p_sel, sh.nw, \
                           sh.fstride, exp_w, slot_lam, (double*)dr, partial, kept);                            \
      else                                                                                                     \
        hipLaunchKernelGGL((k_round_prep<B, N, float>), dim3(nblk, na), dim3(256), 0, s, (const float*)factors, \
                           (const float*)qdiv, meta, wl, nvar_stored, sh.p_sel, sh.nw, sh.fstride, exp_w,       \
                           slot_lam, (float*)dr, partial, kept);                                                \
      return hipGetLastError();                                                                                \
    }                                                                                                          \
  }
  PG_SHAPES(PG_RP)
#undef PG_RP
  if (sh.fp64) return hipErrorInvalidValue;
  const uint32_t nm = shape_generic(sh.bs, sh.nw);
#define PG_RPG(B, NM, L, LE)                                                                                     \
  if constexpr (B <= 4) {                                                                                      \
    if (sh.bs == B && nm == NM) {                                                                              \
      hipLaunchKernelGGL((k_round_prep<B, NM, float, true>), dim3(nblk, na), dim3(256), 0, s,                  \
                         (const float*)factors, (const float*)qdiv, meta, wl, nvar_stored, sh.p_sel, sh.nw,     \
                         sh.fstride, exp_w, slot_lam, (float*)dr, partial, kept);                               \
      return hipGetLastError();                                                                                \
    }                                                                                                          \
  }
  PG_GENERIC(PG_RPG)
#undef PG_RPG
  return hipErrorInvalidValue;
}

hipError_t launch_eval(const void* factors, const WinMeta* meta, const uint32_t* wl, uint32_t na,
                       uint32_t nvar_stored, uint32_t nvar_run, const Shape& sh, int exp_w, const void* divq,
                       const int32_t* inv, void* B0, double* partial, hipStream_t s, const double* slot_lam,
                       double* kept) {
  if (nvar_run == 3 && !divq) return hipErrorInvalidValue;  // launch_divq after the SVD
  const uint32_t nblk = eval_blocks(sh, nvar_run);
#define PG_EV(B, N, L, LS)                                                                                       \
  if (sh.bs == B && sh.nw == N) {                                                                              \
    if constexpr (L == 64) {                                                                                   \
      if (sh.fp64)                                                                                             \
        hipLaunchKernelGGL((k_eval<B, N, L, double>), dim3(nblk, na), dim3(256), 0, s, (const double*)factors,  \
                           meta, wl, nvar_stored, nvar_run, sh.p_sel, sh.nw, sh.fstride, exp_w,                 \
                           (const double*)divq, inv, (double*)B0, partial, slot_lam, kept);                    \
      else                                                                                                     \
        hipLaunchKernelGGL((k_eval<B, N, L, float>), dim3(nblk, na), dim3(256), 0, s, (const float*)factors,    \
                           meta, wl, nvar_stored, nvar_run, sh.p_sel, sh.nw, sh.fstride, exp_w,                 \
                           (const float*)divq, inv, (float*)B0, partial, slot_lam, kept);                      \
      return hipGetLastError();                                                                                \
    }                                                                                                          \
  }
  PG_SHAPES(PG_EV)
#undef PG_EV
  if (sh.fp64) return hipErrorInvalidValue;
  const uint32_t nm = shape_generic(sh.bs, sh.nw);
#define PG_EVG(B, NM, L, LE)                                                                                     \
  if constexpr (LE == 64 && B > 4) {                                                                           \
    if (sh.bs == B && nm == NM) {                                                                              \
      hipLaunchKernelGGL((k_eval<B, NM, LE, float, true>), dim3(nblk, na), dim3(256), 0, s,                    \
                         (const float*)factors, meta, wl, nvar_stored, nvar_run, sh.p_sel, sh.nw, sh.fstride,   \
                         exp_w, (const float*)divq, inv, (float*)B0, partial, slot_lam, kept);                 \
      return hipGetLastError();                                                                                \
    }                                                                                                          \
  }
  PG_GENERIC(PG_EVG)
#undef PG_EVG
  return hipErrorInvalidValue;
}

hipError_t launch_divq(const void* factors, const int32_t* pos, const uint32_t* sel, const void* dw, uint32_t nb,
                       uint32_t nvar, const Shape& sh, void* divq, hipStream_t s) {
  const uint32_t lim = sh.nr - sh.bs;
#define PG_DQ(B, N, L, LS)                                                                                       \
  if (sh.bs == B && sh.nw == N) {                                                                              \
    const dim3 grid((nvar * sh.p_sel + 256 / L - 1) / (256 / L), nb);                                          \
    if (sh.fp64)                                                                                               \
      hipLaunchKernelGGL((k_divq<B, N, L, double>), grid, dim3(256), 0, s, (const double*)factors, pos, sel,    \
                         (const double*)dw, nvar, sh.p_sel, sh.p_all, sh.nr, sh.fs, sh.nw, sh.vol, lim,         \
                         sh.fstride, (double*)divq);                                                           \
    else                                                                                                       \
      hipLaunchKernelGGL((k_divq<B, N, L, float>), grid, dim3(256), 0, s, (const float*)factors, pos, sel,      \
                         (const float*)dw, nvar, sh.p_sel, sh.p_all, sh.nr, sh.fs, sh.nw, sh.vol, lim,          \
                         sh.fstride, (float*)divq);                                                            \
    return hipGetLastError();                                                                                  \
  }
  PG_SHAPES(PG_DQ)
#undef PG_DQ
  if (sh.fp64) return hipErrorInvalidValue;
  const uint32_t nm = shape_generic(sh.bs, sh.nw);
#define PG_DQG(B, NM, L, LE)                                                                                     \
  if (sh.bs == B && nm == NM) {                                                                                \
    const dim3 grid((nvar * sh.p_sel + 256 / LE - 1) / (256 / LE), nb);                                       \
    hipLaunchKernelGGL((k_divq<B, NM, LE, float, true>), grid, dim3(256), 0, s, (const float*)factors, pos,   \
                       sel, (const float*)dw, nvar, sh.p_sel, sh.p_all, sh.nr, sh.fs, sh.nw, sh.vol, lim,      \
                       sh.fstride, (float*)divq);                                                              \
    return hipGetLastError();                                                                                  \
  }
  PG_GENERIC(PG_DQG)
#undef PG_DQG
  return hipErrorInvalidValue;
}

// tiled gather for patches up to 4 x 4 (32 x 8 voxel tiles), else one
// 256-voxel run per block
uint32_t gather_frames(const Shape& sh, uint32_t nk) {
  if (sh.bs > 4 || nk == 0 || nk > sh.nw) return sh.nw;
  return nk;
}

uint32_t gather_blocks(const Shape& sh, int mode, uint32_t nk) {
  if (sh.bs <= 4)
    return ((sh.nr + kTileRows - 1) / kTileRows) * ((sh.nr + kTileCols - 1) / kTileCols) *
           (mode == 1 ? 1 : gather_frames(sh, nk));
  return ((mode == 1 ? sh.fs : sh.vol) + 255) / 256;
}

hipError_t launch_gather(const GatherSrc& src, const int32_t* off, const double* invw, const double* u64,
                         const WinMeta* meta, const uint32_t* wl, uint32_t na, const Shape& sh, int mode, double* out,
                         double* partial, hipStream_t s, uint32_t nk) {
  const uint32_t nblk = gather_blocks(sh, mode, nk);
  const uint32_t nkk = (mode == 1) ? 1 : gather_frames(sh, nk);
#define PG_GT(B)                                                                                                 \
  case B:                                                                                                      \
    if (sh.fp64)                                                                                               \
      hipLaunchKernelGGL((k_gather_tile<B, double>), dim3(nblk, na), dim3(256), 0, s, (const double*)src.factors, \
                         (const double*)src.dr, (const double*)src.S01, (const double*)src.W0,                 \
                         src.list, off, invw, u64, meta, wl, src.slot_lam, src.exp_w,                            \
                         sh.nr, sh.fs, sh.nw,                                                                  \
                         sh.vol, sh.n1, sh.p_sel, sh.fstride, src.nvar_stored, mode, nkk, out, partial);       \
    else                                                                                                       \
      hipLaunchKernelGGL((k_gather_tile<B, float>), dim3(nblk, na), dim3(256), 0, s, (const float*)src.factors,  \
                         (const float*)src.dr, (const float*)src.S01, (const float*)src.W0,                    \
                         src.list, off, invw, u64, meta, wl, src.slot_lam, src.exp_w,                             \
                         sh.nr, sh.fs, sh.nw,                                                                  \
                         sh.vol, sh.n1, sh.p_sel, sh.fstride, src.nvar_stored, mode, nkk, out, partial);       \
    break;
#define PG_GA(B)                                                                                                 \
  case B:                                                                                                      \
    if (sh.fp64)                                                                                               \
      hipLaunchKernelGGL((k_gather_direct<B, double>), dim3(nblk, na), dim3(256), 0, s, (const double*)src.B0,   \
                         off, invw, u64, meta, wl, sh.nr, sh.fs, sh.nw, sh.vol, sh.n1, sh.p_sel, mode, out,      \
                         partial);                                                                             \
    else                                                                                                       \
      hipLaunchKernelGGL((k_gather_direct<B, float>), dim3(nblk, na), dim3(256), 0, s, (const float*)src.B0,     \
                         off, invw, u64, meta, wl, sh.nr, sh.fs, sh.nw, sh.vol, sh.n1, sh.p_sel, mode, out,      \
                         partial);                                                                             \
    break;
  switch (sh.bs) {
    PG_GT(2) PG_GT(3) PG_GT(4)
    PG_GA(5) PG_GA(6) PG_GA(7) PG_GA(8) PG_GA(16)
    default: return hipErrorInvalidValue;
  }
#undef PG_GA
#undef PG_GT
  return hipGetLastError();
}

// ------------------------------------------------------------- lambda grid --
// Per window b: G thresholds log-spaced over [0.01, 10] times a scale from
// the median of the sorted singular values of `nsamp` sampled patches (1 /
// median^2 with exponential weighting, whose threshold acts on s^2, the
// median itself without).  One block per window; the values are sorted in
// LDS (bitonic, padded with +inf) and the median is the mean of the middle
// pair for an even count.
template <class T>
__global__ __launch_bounds__(256) void k_grid_lambdas(const T* __restrict__ factors, uint32_t nvar_stored,
                                                      uint32_t p_sel, uint32_t fstride, uint32_t s_at, uint32_t n,
                                                      const uint32_t* __restrict__ sample, uint32_t nsamp,
                                                      uint32_t G, int exp_w, double* __restrict__ lam) {
  constexpr int CAPV = 8192;
  __shared__ float v[CAPV];
  const uint32_t b = blockIdx.x, cnt = min(nsamp * n, (uint32_t)CAPV);
  uint32_t P2 = 1;
  while (P2 < cnt) P2 <<= 1;
  for (uint32_t e = threadIdx.x; e < P2; e += blockDim.x)
    v[e] = (e < cnt) ? (float)factors[((size_t)b * nvar_stored * p_sel + sample[e / n]) * fstride + s_at + e % n]
                     : INFINITY;
  __syncthreads();
  for (uint32_t k = 2; k <= P2; k <<= 1)
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      for (uint32_t e = threadIdx.x; e < P2; e += blockDim.x) {
        const uint32_t q = e ^ j;
        if (q > e) {
          const bool up = (e & k) == 0;
          const float a = v[e], c = v[q];
          if ((a > c) == up) {
            v[e] = c;
            v[q] = a;
          }
        }
      }
      __syncthreads();
    }
  const double med = (cnt % 2) ? (double)v[cnt / 2] : 0.5 * ((double)v[cnt / 2 - 1] + (double)v[cnt / 2]);
  const double m = fmax(med, 1e-12);
  const double scale = exp_w ? 1.0 / (m * m) : m;
  for (uint32_t j = threadIdx.x; j < G; j += blockDim.x)
    lam[(size_t)b * G + j] = scale * pow(10.0, -2.0 + 3.0 * (double)j / (double)(G > 1 ? G - 1 : 1));
}

// PGURE risk of every (window, candidate) from the GPU sums, as the driver
// evaluates it (src/pgure.hpp:136, mu and sigma swapped at the constructor
// call, the delta1 term zero), then the first candidate of least risk
// becomes the window's lambda.  One thread per window.
__global__ void k_grid_select(const double* __restrict__ sums, const double* __restrict__ noise, uint32_t nb,
                              uint32_t G, double OoN, double eps2, const double* __restrict__ lam,
                              WinMeta* __restrict__ meta, double* __restrict__ risk) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const double alpha = noise[3 * b], mup = noise[3 * b + 2], sigp = noise[3 * b + 1];
  const double sigmasq = sigp * sigp;
  uint32_t best = 0;
  double rbest = 0;
  for (uint32_t j = 0; j < G; ++j) {
    const double* S = sums + ((size_t)b * G + j) * 6;
    const double S1 = S[0] - 2.0 * S[3] + S[2];
    const double r =
        OoN * (S1 - (alpha + mup) * S[1] - (2 * sigmasq * alpha / (eps2 * eps2) * S[4]) + (2 * mup * S[5]) + mup) -
        sigmasq;
    risk[(size_t)b * G + j] = r;
    if (j == 0 || r < rbest) {
      rbest = r;
      best = j;
    }
  }
  meta[b].lambda = lam[(size_t)b * G + best];
}

hipError_t launch_grid_lambdas(const void* factors, uint32_t nvar_stored, const Shape& sh, uint32_t s_at,
                               const uint32_t* sample, uint32_t nsamp, uint32_t nb, uint32_t G, int exp_w,
                               double* lam, hipStream_t s) {
  if (sh.fp64)
    hipLaunchKernelGGL(k_grid_lambdas<double>, dim3(nb), dim3(256), 0, s, (const double*)factors, nvar_stored,
                       sh.p_sel, sh.fstride, s_at, sh.nw, sample, nsamp, G, exp_w, lam);
  else
    hipLaunchKernelGGL(k_grid_lambdas<float>, dim3(nb), dim3(256), 0, s, (const float*)factors, nvar_stored,
                       sh.p_sel, sh.fstride, s_at, sh.nw, sample, nsamp, G, exp_w, lam);
  return hipGetLastError();
}

hipError_t launch_grid_select(const double* sums, const double* noise, uint32_t nb, uint32_t G, double OoN,
                              double eps2, const double* lam, WinMeta* meta, double* risk, hipStream_t s) {
  hipLaunchKernelGGL(k_grid_select, dim3((nb + 63) / 64), dim3(64), 0, s, sums, noise, nb, G, OoN, eps2, lam, meta,
                     risk);
  return hipGetLastError();
}

hipError_t launch_risk_reduce(const double* pa, uint32_t na_blk, const double* pb, uint32_t nb_blk, uint32_t na,
                              double* sums, hipStream_t s, const double* pk, double* kept) {
  hipLaunchKernelGGL(k_risk_reduce, dim3(na), dim3(256), 0, s, pa, na_blk, pb, nb_blk, sums, pk, kept);
  return hipGetLastError();
}

}  // namespace pgsvt
