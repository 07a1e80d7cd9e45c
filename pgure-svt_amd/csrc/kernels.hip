// kernels.hip -- gfx950 kernels of the PGURE-SVT patchwise SVT path.
//
// Layouts (all column-major "cubes" as in the reference, element (r, c, f) at
// r + nr*(c + nc*f)):
//   X, Z     double [nf][nc][nr]            input / median-filtered input
//   u64,w64  double [nb][vol]               normalised window (risk, ARPS)
//   u32      float  [nb][nvar][vol]         variant windows (SVD, reconstruct)
//   pos      int32  [nb][nw][p_all]         patch position, row | col << 16
//   W        int32  [nb][vol]               overlap count per voxel
//   factors  float  [nb][nvar][p_sel][fstride]  V (n x n row-major), s (n)
//   acc      int64  [nb][nvar][vol]         fixed-point overlap-add sums
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>

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
__global__ void k_prepare(const double* __restrict__ X, const double* __restrict__ Z, const WinMeta* __restrict__ meta,
                          const uint8_t* __restrict__ d2, uint32_t fs, uint32_t vol, uint32_t nvar, double eps2,
                          double d2neg, double d2pos, double* __restrict__ u64, double* __restrict__ w64,
                          float* __restrict__ u32) {
  const uint32_t b = blockIdx.y;
  const WinMeta m = meta[b];
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < vol; v += gridDim.x * blockDim.x) {
    const size_t src = (size_t)m.s0 * fs + v;
    const double u = X[src] / m.uMax;
    u64[(size_t)b * vol + v] = u;
    if (Z) w64[(size_t)b * vol + v] = Z[src] / m.wMax;
    float* o = u32 + (size_t)b * nvar * vol + v;
    o[0] = (float)u;
    if (nvar == 3) {
      const double dl = d2[v] ? d2neg : d2pos;
      o[vol] = (float)(u + dl * eps2);
      o[2 * (size_t)vol] = (float)(u - dl * eps2);
    }
  }
}

hipError_t launch_prepare(const double* X, const double* Z, const WinMeta* meta, uint32_t nb, const uint8_t* d2,
                          const Shape& sh, uint32_t nvar, double eps2, double d2neg, double d2pos, double* u64,
                          double* w64, float* u32, hipStream_t s) {
  const dim3 grid(std::min<uint32_t>((sh.vol + 255) / 256, 1024), nb);
  hipLaunchKernelGGL(k_prepare, grid, dim3(256), 0, s, X, Z, meta, d2, sh.fs, sh.vol, nvar, eps2, d2neg, d2pos,
                     u64, w64, u32);
  return hipGetLastError();
}

// reference-frame positions: patches(0,i,ref) = i % (1+nyMbs),
// patches(1,i,ref) = i / (1+nxMbs) (src/arps.hpp:60-64, 85-89, 114-118)
__global__ void k_fill_ref(const WinMeta* __restrict__ meta, uint32_t p_all, uint32_t n1, int32_t* __restrict__ pos,
                           uint32_t nw) {
  const uint32_t b = blockIdx.y;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= p_all) return;
  pos[((size_t)b * nw + meta[b].ref) * p_all + i] = (int32_t)((i % n1) | ((i / n1) << 16));
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
constexpr int kArpsMaxMw = 15;
constexpr int kArpsWords = ((2 * kArpsMaxMw + 1) * (2 * kArpsMaxMw + 1) + 31) / 32;
constexpr int kArpsThreads = 128;

template <int BS>
__device__ __forceinline__ double arps_cost(const double (&ref)[BS * BS], const double* __restrict__ F, uint32_t nr,
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
  double ref[BS * BS];
  const double* S = win + (size_t)src * fs + i + (size_t)nr * j;
#pragma unroll
  for (int q = 0; q < BS * BS; ++q) ref[q] = S[(q % BS) + (size_t)nr * (q / BS)];
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
template <int BS>
__global__ __launch_bounds__(kArpsThreads) void k_arps(const double* __restrict__ w64, const int16_t* __restrict__ jobs,
                                                       uint32_t max_jobs, uint32_t nr, uint32_t nc, uint32_t fs,
                                                       uint32_t nw, uint32_t vol, uint32_t p_all, int n1, int mw,
                                                       int32_t* __restrict__ pos) {
  __shared__ uint32_t sv[kArpsWords * kArpsThreads];
  const uint32_t b = blockIdx.z, job = blockIdx.y;
  const int it = blockIdx.x * blockDim.x + threadIdx.x;
  const int16_t* J = jobs + ((size_t)b * max_jobs + job) * 4;
  const int sa = J[0], da = J[1], sb = J[2], db = J[3];
  if (sa < 0 || it >= (int)p_all) return;
  Visit vis{sv + threadIdx.x, kArpsThreads, 2 * mw + 1};
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

hipError_t launch_arps(const double* w64, const int16_t* jobs, uint32_t max_jobs, uint32_t nb, const Shape& sh,
                       int mw, int32_t* pos, hipStream_t s) {
  if (mw > kArpsMaxMw) return hipErrorInvalidValue;
  const dim3 grid((sh.p_all + kArpsThreads - 1) / kArpsThreads, max_jobs, nb);
  const dim3 blk(kArpsThreads);
#define PG_ARPS(B)                                                                                              \
  case B:                                                                                                       \
    hipLaunchKernelGGL(k_arps<B>, grid, blk, 0, s, w64, jobs, max_jobs, sh.nr, sh.nc, sh.fs, sh.nw, sh.vol,     \
                       sh.p_all, (int)sh.n1, mw, pos);                                                          \
    break;
  switch (sh.bs) {
    PG_ARPS(2) PG_ARPS(3) PG_ARPS(4) PG_ARPS(5) PG_ARPS(6) PG_ARPS(7) PG_ARPS(8)
    default: return hipErrorInvalidValue;
  }
#undef PG_ARPS
  return hipGetLastError();
}

// ------------------------------------------------------- overlap counts --
// weights of src/svt.hpp:157-159: one per patch covering a voxel
template <int BS>
__global__ void k_count(const int32_t* __restrict__ pos, const uint32_t* __restrict__ sel, uint32_t p_sel,
                        uint32_t p_all, uint32_t nr, uint32_t fs, uint32_t nw, uint32_t vol, uint32_t lim,
                        int32_t* __restrict__ W) {
  const uint32_t b = blockIdx.z, k = blockIdx.y;
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= p_sel) return;
  const int32_t pk = pos[((size_t)b * nw + k) * p_all + sel[p]];
  const uint32_t row = min((uint32_t)(pk & 0xffff), lim), col = min((uint32_t)(pk >> 16), lim);
  int32_t* Wk = W + (size_t)b * vol + (size_t)k * fs + row + (size_t)nr * col;
#pragma unroll
  for (int q = 0; q < BS * BS; ++q) atomicAdd(Wk + (q % BS) + (size_t)nr * (q / BS), 1);
}

hipError_t launch_count(const int32_t* pos, const uint32_t* sel, uint32_t nb, const Shape& sh, int32_t* W,
                        hipStream_t s) {
  const dim3 grid((sh.p_sel + 255) / 256, sh.nw, nb);
  const uint32_t lim = sh.nr - sh.bs;
#define PG_CNT(B)                                                                                                 \
  case B:                                                                                                         \
    hipLaunchKernelGGL(k_count<B>, grid, dim3(256), 0, s, pos, sel, sh.p_sel, sh.p_all, sh.nr, sh.fs, sh.nw,      \
                       sh.vol, lim, W);                                                                           \
    break;
  switch (sh.bs) {
    PG_CNT(2) PG_CNT(3) PG_CNT(4) PG_CNT(5) PG_CNT(6) PG_CNT(7) PG_CNT(8)
    default: return hipErrorInvalidValue;
  }
#undef PG_CNT
  return hipGetLastError();
}

// --------------------------------------------------- batched Jacobi SVD --
// One-sided (Hestenes) Jacobi on the m x n Casorati matrix of each patch
// (arma::svd_econ at src/svt.hpp:111).  A group of L lanes holds one matrix,
// lane q owning row q of X and row q of V; column dot products are
// butterflies across the group (DPP inside a 16-lane row, then lane swaps
// for 32/64-lane groups).  Every lane of a group ends the butterfly with the
// bit-identical sum, so all lanes apply the same rotation.  Column norms are
// carried through the sweep with the exact Jacobi update (a -= t c, b += t c)
// and refreshed at the start of every sweep, so a pair costs one butterfly.
template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xF, 0xF, false));
}

template <int L>
__device__ __forceinline__ float group_sum(float v) {
  v += dpp_f<0xB1>(v);   // quad_perm [1,0,3,2]
  v += dpp_f<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp_f<0x141>(v);  // row_half_mirror
  v += dpp_f<0x140>(v);  // row_mirror
  if (L >= 32) v += __shfl_xor(v, 16);
  if (L >= 64) v += __shfl_xor(v, 32);
  return v;
}

template <int BS, int N, int L>
__global__ __launch_bounds__(256) void k_svd(const float* __restrict__ u32, const int32_t* __restrict__ pos,
                                             const uint32_t* __restrict__ sel, uint32_t nvar, uint32_t p_sel,
                                             uint32_t p_all, uint32_t nr, uint32_t fs, uint32_t nw, uint32_t vol,
                                             uint32_t lim, uint32_t fstride, int max_sweeps, float tol2,
                                             float* __restrict__ factors, uint32_t* __restrict__ stats) {
  constexpr int M = BS * BS;
  static_assert(M <= L && N <= L, "group too small");
  const uint32_t b = blockIdx.y;
  const int lane = threadIdx.x % L;
  const uint32_t g = (blockIdx.x * blockDim.x + threadIdx.x) / L;
  const uint32_t var = g / p_sel, p = g % p_sel;
  const bool valid = var < nvar;  // uniform per group
  const uint32_t it = valid ? sel[p] : 0;

  float x[N], v[N];
  const int rr = lane % BS, cc = lane / BS;
  const float* U = u32 + ((size_t)b * nvar + (valid ? var : 0)) * vol;
#pragma unroll
  for (int k = 0; k < N; ++k) {
    float val = 0.f;
    if (valid && lane < M) {
      const int32_t pk = pos[((size_t)b * nw + k) * p_all + it];
      const uint32_t row = min((uint32_t)(pk & 0xffff), lim), col = min((uint32_t)(pk >> 16), lim);
      val = U[(size_t)k * fs + (row + rr) + (size_t)nr * (col + cc)];
    }
    x[k] = val;
    v[k] = (lane == k) ? 1.f : 0.f;
  }

  float a[N];
  int sweep = 0;
  for (; sweep < max_sweeps; ++sweep) {
#pragma unroll
    for (int k = 0; k < N; ++k) a[k] = group_sum<L>(x[k] * x[k]);
    // sum of t^2 over the sweep: zero exactly when no pair rotated.  A
    // floating-point chain (never reassociated) keeps this in one register.
    float tsum = 0.f;
#pragma unroll
    for (int i = 0; i < N - 1; ++i) {
#pragma unroll
      for (int j = i + 1; j < N; ++j) {
        // rotation zeroing the (i, j) inner product: zeta = (b - a) / 2c,
        // t = sign(zeta) / (|zeta| + sqrt(1 + zeta^2)), cs = 1/sqrt(1+t^2)
        const float c = group_sum<L>(x[i] * x[j]);
        const bool rot = c * c > tol2 * a[i] * a[j];
        const float zeta = (a[j] - a[i]) * __builtin_amdgcn_rcpf(c + c);
        const float den = fabsf(zeta) + __builtin_amdgcn_sqrtf(fmaf(zeta, zeta, 1.f));
        const float t = rot ? copysignf(__builtin_amdgcn_rcpf(den), zeta) : 0.f;
        const float cs = __builtin_amdgcn_rsqf(fmaf(t, t, 1.f));
        const float sn = cs * t;
        const float xi = x[i], xj = x[j];
        x[i] = cs * xi - sn * xj;
        x[j] = sn * xi + cs * xj;
        const float vi = v[i], vj = v[j];
        v[i] = cs * vi - sn * vj;
        v[j] = sn * vi + cs * vj;
        a[i] = fmaf(-t, c, a[i]);
        a[j] = fmaf(t, c, a[j]);
        tsum = fmaf(t, t, tsum);
        // Pin the V update next to its rotation.  V only feeds the next
        // sweep, so without this the compiler sinks every V update to the
        // end of the sweep and keeps (cs, sn) of all rotations live at once
        // (about 2 n(n-1)/2 extra registers).  The empty statement only
        // names the two registers; it has no instruction and no memory
        // access.
        asm volatile("" : "+v"(v[i]), "+v"(v[j]));
      }
    }
    if (!__any(tsum > 0.f)) break;
  }
  if (stats && lane == 0 && valid) atomicMax(stats, (uint32_t)sweep + 1);
  if (!valid) return;
  float* F = factors + (((size_t)b * nvar + var) * p_sel + p) * fstride;
  float sq = 0.f;
#pragma unroll
  for (int k = 0; k < N; ++k) {
    const float nk = group_sum<L>(x[k] * x[k]);
    if (lane == k) sq = sqrtf(nk);
  }
  if (lane < N) {
#pragma unroll
    for (int k = 0; k < N; ++k) F[lane * N + k] = v[k];
    F[N * N + lane] = sq;
  }
}

// ------------------------------------------ threshold + reconstruction --
// src/svt.hpp:121-167 with src/utils.hpp:96-106.  B = U diag(g(s)) V^T is
// formed as (X V) diag(g(s)/s) V^T from the stored V and s and the original
// Casorati matrix, so only n*n+n floats per patch are kept.  Each lane
// reconstructs its row of B and adds it into the fixed-point accumulators.
template <int BS, int N, int L>
__global__ __launch_bounds__(256) void k_recon(const float* __restrict__ factors, const float* __restrict__ u32,
                                               const int32_t* __restrict__ pos, const uint32_t* __restrict__ sel,
                                               const WinMeta* __restrict__ meta, uint32_t nvar_stored,
                                               uint32_t nvar_run, uint32_t p_sel, uint32_t p_all, uint32_t nr,
                                               uint32_t fs, uint32_t nw, uint32_t vol, uint32_t lim, uint32_t fstride,
                                               int exp_w, int only_ref, const uint32_t* __restrict__ wl,
                                               unsigned long long* __restrict__ acc) {
  constexpr int M = BS * BS, G = 256 / L, VS = N * N;
  __shared__ float sV[G * VS];
  __shared__ float sS[G * N];
  const uint32_t b = wl[blockIdx.y];
  const int grp = threadIdx.x / L, lane = threadIdx.x % L;
  const uint32_t g = blockIdx.x * G + grp;
  const uint32_t var = g / p_sel, p = g % p_sel;
  const bool valid = var < nvar_run;
  if (valid && lane < N) {
    const float* F = factors + (((size_t)b * nvar_stored + var) * p_sel + p) * fstride;
#pragma unroll
    for (int k = 0; k < N; ++k) sV[grp * VS + lane * N + k] = F[lane * N + k];
    sS[grp * N + lane] = F[N * N + lane];
  }
  __syncthreads();
  if (!valid || lane >= M) return;

  const double lam = meta[b].lambda;
  float s[N], d[N];
  float smax = 0.f;
#pragma unroll
  for (int l = 0; l < N; ++l) {
    s[l] = sS[grp * N + l];
    smax = fmaxf(smax, s[l]);
  }
  if (exp_w) {
    const float kk = (float)(-0.5 * lam);
#pragma unroll
    for (int l = 0; l < N; ++l) {
      const float w = fabsf(smax * expf(kk * (s[l] * s[l])));
      const float gl = fmaxf(s[l] - w, 0.f);
      d[l] = s[l] > 0.f ? gl / s[l] : 0.f;
    }
  } else {
    const float lf = (float)lam;
#pragma unroll
    for (int l = 0; l < N; ++l) {
      const float gl = fmaxf(s[l] - lf, 0.f);
      d[l] = s[l] > 0.f ? gl / s[l] : 0.f;
    }
  }

  const uint32_t it = sel[p];
  const int rr = lane % BS, cc = lane / BS;
  const float* U = u32 + ((size_t)b * nvar_stored + var) * vol;
  float x[N];
  uint32_t addr[N];
#pragma unroll
  for (int k = 0; k < N; ++k) {
    const int32_t pk = pos[((size_t)b * nw + k) * p_all + it];
    const uint32_t row = min((uint32_t)(pk & 0xffff), lim), col = min((uint32_t)(pk >> 16), lim);
    addr[k] = (uint32_t)k * fs + (row + rr) + nr * (col + cc);
    x[k] = U[addr[k]];
  }
  float y[N];
#pragma unroll
  for (int l = 0; l < N; ++l) {
    float t = 0.f;
#pragma unroll
    for (int j = 0; j < N; ++j) t = fmaf(x[j], sV[grp * VS + j * N + l], t);
    y[l] = t * d[l];
  }
  unsigned long long* A = acc + (size_t)b * vol;  // variant-0 accumulator
  const int ref = (int)meta[b].ref;
#pragma unroll
  for (int k = 0; k < N; ++k) {
    if (only_ref && k != ref) continue;
    float bq = 0.f;
#pragma unroll
    for (int l = 0; l < N; ++l) bq = fmaf(y[l], sV[grp * VS + k * N + l], bq);
    const long long q = __float2ll_rn(bq * kFixScale);
    atomicAdd(A + addr[k], (unsigned long long)q);
  }
}

// --------------------------------------- fused reconstruction + risk --
// One lambda evaluation of PGURE<T>::CalculatePGURE (src/pgure.hpp:128-136)
// for every patch of a window.  A group reconstructs its patch for the three
// decomposed variants (U, U + eps2 delta2, U - eps2 delta2).  Only the
// unperturbed reconstruction is needed as a volume (for sum Uhat^2); every
// other risk term is linear in the overlap-added values, so with
// Uhat_v = acc_v / W_v it is accumulated per contribution:
//   sum Uhat                  = sum_c B0_c / W_v
//   sum Uhat U                = sum_c B0_c U_v / W_v
//   sum d2 (U2p - 2Uhat + U2m) = sum_c d2_v (B1_c - 2 B0_c + B2_c) / W_v
// (c runs over patch contributions, v is the voxel c lands on).  Only B0 is
// scattered into the fixed-point volume; the sums are FP64 per thread, then
// reduced in a fixed order.
template <int BS, int N, int L>
__global__ __launch_bounds__(256) void k_recon_risk(
    const float* __restrict__ factors, const float* __restrict__ u32, const int32_t* __restrict__ pos,
    const uint32_t* __restrict__ sel, const WinMeta* __restrict__ meta, const uint32_t* __restrict__ wl,
    uint32_t p_sel, uint32_t p_all, uint32_t nr, uint32_t fs, uint32_t nw, uint32_t vol, uint32_t lim,
    uint32_t fstride, int exp_w, const double* __restrict__ invw, const double* __restrict__ u64,
    const uint8_t* __restrict__ d2, double d2neg, double d2pos, unsigned long long* __restrict__ acc,
    double* __restrict__ partial) {
  constexpr int M = BS * BS, G = 256 / L, VS = N * N;
  __shared__ float sV[G * VS];
  __shared__ float sS[G * N];
  __shared__ double red[3][256];
  const uint32_t slot = blockIdx.y, b = wl[slot];
  const int grp = threadIdx.x / L, lane = threadIdx.x % L;
  const uint32_t p = blockIdx.x * G + grp;
  const bool valid = p < p_sel;
  const bool row = valid && lane < M;
  const double lam = meta[b].lambda;
  const uint32_t it = valid ? sel[p] : 0;
  const int rr = lane % BS, cc = lane / BS;
  uint32_t addr[N];
#pragma unroll
  for (int k = 0; k < N; ++k) {
    uint32_t a = 0;
    if (row) {
      const int32_t pk = pos[((size_t)b * nw + k) * p_all + it];
      const uint32_t r0 = min((uint32_t)(pk & 0xffff), lim), c0 = min((uint32_t)(pk >> 16), lim);
      a = (uint32_t)k * fs + (r0 + rr) + nr * (c0 + cc);
    }
    addr[k] = a;
  }
  float B0[N];
  double Cd[N];
#pragma unroll
  for (int k = 0; k < N; ++k) {
    B0[k] = 0.f;
    Cd[k] = 0.0;
  }
#pragma unroll 1
  for (int var = 0; var < 3; ++var) {
    __syncthreads();  // sV / sS reuse across variants
    if (valid && lane < N) {
      const float* F = factors + (((size_t)b * 3 + var) * p_sel + p) * fstride;
#pragma unroll
      for (int k = 0; k < N; ++k) sV[grp * VS + lane * N + k] = F[lane * N + k];
      sS[grp * N + lane] = F[N * N + lane];
    }
    __syncthreads();
    if (row) {
      const float* S = sS + grp * N;
      float smax = 0.f;
#pragma unroll
      for (int l = 0; l < N; ++l) smax = fmaxf(smax, S[l]);
      const float kk = (float)(-0.5 * lam), lf = (float)lam;
      const float* U = u32 + ((size_t)b * 3 + var) * vol;
      float x[N];
#pragma unroll
      for (int k = 0; k < N; ++k) x[k] = U[addr[k]];
      float y[N];
#pragma unroll
      for (int l = 0; l < N; ++l) {
        float t = 0.f;
#pragma unroll
        for (int j = 0; j < N; ++j) t = fmaf(x[j], sV[grp * VS + j * N + l], t);
        const float sl = S[l];
        const float w = exp_w ? fabsf(smax * expf(kk * (sl * sl))) : lf;
        y[l] = sl > 0.f ? t * (fmaxf(sl - w, 0.f) / sl) : 0.f;
      }
      // Compiler-only barrier: re-read V from LDS for the second product
      // instead of keeping all n*n values of the first one in registers.
      asm volatile("" ::: "memory");
#pragma unroll
      for (int k = 0; k < N; ++k) {
        float bq = 0.f;
#pragma unroll
        for (int l = 0; l < N; ++l) bq = fmaf(y[l], sV[grp * VS + k * N + l], bq);
        if (var == 0)
          B0[k] = bq;
        else
          Cd[k] += (double)bq;
      }
    }
  }
  double s5 = 0, sx = 0, s4 = 0;
  if (row) {
    unsigned long long* A = acc + (size_t)b * vol;
    const double* IW = invw + (size_t)b * vol;
    const double* UU = u64 + (size_t)b * vol;
#pragma unroll
    for (int k = 0; k < N; ++k) {
      const uint32_t v = addr[k];
      atomicAdd(A + v, (unsigned long long)__float2ll_rn(B0[k] * kFixScale));
      const double iw = IW[v];
      const double b0 = (double)B0[k];
      s5 = fma(b0, iw, s5);
      sx = fma(b0 * UU[v], iw, sx);
      const double c2 = Cd[k] - 2.0 * b0;
      s4 = fma(c2 * (d2[v] ? d2neg : d2pos), iw, s4);
    }
  }
  red[0][threadIdx.x] = s5;
  red[1][threadIdx.x] = sx;
  red[2][threadIdx.x] = s4;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o)
      for (int c = 0; c < 3; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x < 3) partial[((size_t)slot * gridDim.x + blockIdx.x) * 3 + threadIdx.x] = red[threadIdx.x][0];
}

// sum Uhat^2, sum U, sum U^2 from the variant-0 volume (cleared as read)
__global__ __launch_bounds__(256) void k_risk_volume(unsigned long long* __restrict__ acc,
                                                     const double* __restrict__ invw,
                                                     const double* __restrict__ u64, const uint32_t* __restrict__ wl,
                                                     uint32_t vol, double* __restrict__ partial) {
  __shared__ double sm[3][256];
  const uint32_t slot = blockIdx.y, b = wl[slot];
  double q = 0, s2 = 0, su2 = 0;
  unsigned long long* A = acc + (size_t)b * vol;
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < vol; v += gridDim.x * blockDim.x) {
    const long long a = (long long)A[v];
    A[v] = 0ull;
    const double uh = ((double)a * kFixInv) * invw[(size_t)b * vol + v];
    const double U = u64[(size_t)b * vol + v];
    q = fma(uh, uh, q);
    s2 += U;
    su2 = fma(U, U, su2);
  }
  sm[0][threadIdx.x] = q;
  sm[1][threadIdx.x] = s2;
  sm[2][threadIdx.x] = su2;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o)
      for (int c = 0; c < 3; ++c) sm[c][threadIdx.x] += sm[c][threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x < 3) partial[((size_t)slot * gridDim.x + blockIdx.x) * 3 + threadIdx.x] = sm[threadIdx.x][0];
}

// per slot: (sum Uhat^2, sum U, sum U^2, sum Uhat U, sum d2 (..), sum Uhat)
__global__ __launch_bounds__(256) void k_risk_reduce2(const double* __restrict__ pa, uint32_t na,
                                                      const double* __restrict__ pb, uint32_t nb_,
                                                      double* __restrict__ sums) {
  __shared__ double sm[6][256];
  const uint32_t slot = blockIdx.x;
  double s[6] = {0, 0, 0, 0, 0, 0};
  for (uint32_t k = threadIdx.x; k < nb_; k += blockDim.x)
    for (int c = 0; c < 3; ++c) s[c] += pb[((size_t)slot * nb_ + k) * 3 + c];
  for (uint32_t k = threadIdx.x; k < na; k += blockDim.x) {
    s[5] += pa[((size_t)slot * na + k) * 3 + 0];  // sum Uhat
    s[3] += pa[((size_t)slot * na + k) * 3 + 1];  // sum Uhat U
    s[4] += pa[((size_t)slot * na + k) * 3 + 2];  // sum d2 (U2p - 2 Uhat + U2m)
  }
  for (int c = 0; c < 6; ++c) sm[c][threadIdx.x] = s[c];
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o)
      for (int c = 0; c < 6; ++c) sm[c][threadIdx.x] += sm[c][threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x < 6) sums[(size_t)slot * 6 + threadIdx.x] = sm[threadIdx.x][0];
}

__global__ void k_invw(const int32_t* __restrict__ W, size_t n, double* __restrict__ invw) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    invw[i] = W[i] > 0 ? 1.0 / (double)W[i] : 0.0;
}

hipError_t launch_invw(const int32_t* W, size_t n, double* invw, hipStream_t s) {
  const unsigned grid = (unsigned)std::min<size_t>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(k_invw, dim3(grid), dim3(256), 0, s, W, n, invw);
  return hipGetLastError();
}

// shape table of the instantiated SVD / reconstruction kernels
#define PG_SHAPES(X) \
  X(4, 3, 16) X(4, 5, 16) X(4, 7, 16) X(4, 9, 16) X(4, 11, 16) X(4, 13, 16) X(4, 15, 16) \
  X(2, 3, 16) X(3, 3, 16) X(3, 5, 16) X(3, 7, 16) X(3, 9, 16) \
  X(6, 7, 64) X(6, 15, 64) X(6, 31, 64) X(8, 7, 64) X(8, 15, 64) X(8, 31, 64)

bool svd_shape_supported(uint32_t bs, uint32_t nw) {
#define PG_CHK(B, N, L) \
  if (bs == B && nw == N) return true;
  PG_SHAPES(PG_CHK)
#undef PG_CHK
  return false;
}

hipError_t launch_svd(const float* u32, const int32_t* pos, const uint32_t* sel, uint32_t nb, uint32_t nvar,
                      const Shape& sh, float* factors, int max_sweeps, float tol, uint32_t* stats, hipStream_t s) {
  const uint32_t lim = sh.nr - sh.bs;
  const float tol2 = tol * tol;
#define PG_SVD(B, N, L)                                                                                          \
  if (sh.bs == B && sh.nw == N) {                                                                              \
    const uint32_t groups = nvar * sh.p_sel;                                                                   \
    const dim3 grid((groups * L + 255) / 256, nb);                                                             \
    hipLaunchKernelGGL((k_svd<B, N, L>), grid, dim3(256), 0, s, u32, pos, sel, nvar, sh.p_sel, sh.p_all,      \
                       sh.nr, sh.fs, sh.nw, sh.vol, lim, sh.fstride, max_sweeps, tol2, factors, stats);        \
    return hipGetLastError();                                                                                  \
  }
  PG_SHAPES(PG_SVD)
#undef PG_SVD
  return hipErrorInvalidValue;
}

hipError_t launch_recon(const float* factors, const float* u32, const int32_t* pos, const uint32_t* sel,
                        const WinMeta* meta, const uint32_t* wl, uint32_t nb, uint32_t nvar_stored,
                        uint32_t nvar_run, const Shape& sh, int exp_w, int only_ref, unsigned long long* acc,
                        hipStream_t s) {
  const uint32_t lim = sh.nr - sh.bs;
#define PG_REC(B, N, L)                                                                                          \
  if (sh.bs == B && sh.nw == N) {                                                                              \
    const uint32_t groups = nvar_run * sh.p_sel;                                                               \
    const dim3 grid((groups + (256 / L) - 1) / (256 / L), nb);                                                 \
    hipLaunchKernelGGL((k_recon<B, N, L>), grid, dim3(256), 0, s, factors, u32, pos, sel, meta, nvar_stored,  \
                       nvar_run, sh.p_sel, sh.p_all, sh.nr, sh.fs, sh.nw, sh.vol, lim, sh.fstride, exp_w,      \
                       only_ref, wl, acc);                                                                     \
    return hipGetLastError();                                                                                  \
  }
  PG_SHAPES(PG_REC)
#undef PG_REC
  return hipErrorInvalidValue;
}

uint32_t recon_risk_blocks(const Shape& sh) {
  const uint32_t G = (sh.m <= 16) ? 16 : (sh.m <= 32 ? 8 : 4);
  return (sh.p_sel + G - 1) / G;
}

hipError_t launch_recon_risk(const float* factors, const float* u32, const int32_t* pos, const uint32_t* sel,
                             const WinMeta* meta, const uint32_t* wl, uint32_t na, const Shape& sh, int exp_w,
                             const double* invw, const double* u64, const uint8_t* d2, double d2neg, double d2pos,
                             unsigned long long* acc, double* partial_a, double* partial_b, uint32_t nblk_b,
                             double* sums, hipStream_t s) {
  const uint32_t lim = sh.nr - sh.bs;
  const uint32_t nblk_a = recon_risk_blocks(sh);
  bool done = false;
#define PG_RR(B, N, L)                                                                                           \
  if (!done && sh.bs == B && sh.nw == N) {                                                                     \
    hipLaunchKernelGGL((k_recon_risk<B, N, L>), dim3(nblk_a, na), dim3(256), 0, s, factors, u32, pos, sel,     \
                       meta, wl, sh.p_sel, sh.p_all, sh.nr, sh.fs, sh.nw, sh.vol, lim, sh.fstride, exp_w, invw,  \
                       u64, d2, d2neg, d2pos, acc, partial_a);                                                 \
    done = true;                                                                                               \
  }
  PG_SHAPES(PG_RR)
#undef PG_RR
  if (!done) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_risk_volume, dim3(nblk_b, na), dim3(256), 0, s, acc, invw, u64, wl, sh.vol, partial_b);
  hipLaunchKernelGGL(k_risk_reduce2, dim3(na), dim3(256), 0, s, partial_a, nblk_a, partial_b, nblk_b, sums);
  return hipGetLastError();
}

// -------------------------------------------------------- PGURE risk --
__device__ __forceinline__ double fix_mean(unsigned long long& a, int32_t w) {
  const long long v = (long long)a;
  a = 0ull;
  return w > 0 ? ((double)v * kFixInv) / (double)w : 0.0;
}

// final frame: Y(:,:,t) = uMax * Recon(lambda)(:,:,ref)  (src/pguresvt.hpp:146-166)
__global__ void k_output(unsigned long long* __restrict__ acc, const int32_t* __restrict__ W,
                         const WinMeta* __restrict__ meta, uint32_t nvar_acc, uint32_t fs, uint32_t vol,
                         double* __restrict__ Y) {
  const uint32_t b = blockIdx.y;
  const uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x;
  if (pix >= fs) return;
  const WinMeta m = meta[b];
  const size_t v = (size_t)m.ref * fs + pix;
  (void)nvar_acc;
  const double val = fix_mean(acc[(size_t)b * vol + v], W[(size_t)b * vol + v]);
  Y[(size_t)m.t * fs + pix] = val * m.uMax;
}

hipError_t launch_output(unsigned long long* acc, const int32_t* W, const WinMeta* meta, uint32_t nb,
                         uint32_t nvar_acc, const Shape& sh, double* Y, hipStream_t s) {
  hipLaunchKernelGGL(k_output, dim3((sh.fs + 255) / 256, nb), dim3(256), 0, s, acc, W, meta, nvar_acc, sh.fs,
                     sh.vol, Y);
  return hipGetLastError();
}

// whole reconstructed window (variant 0 of batch slot 0), normalised units
__global__ void k_window_volume(unsigned long long* __restrict__ acc, const int32_t* __restrict__ W, uint32_t vol,
                                double* __restrict__ out) {
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= vol) return;
  out[v] = fix_mean(acc[v], W[v]);
}

hipError_t launch_window_volume(unsigned long long* acc, const int32_t* W, uint32_t nvar_acc, const Shape& sh,
                                double* out, hipStream_t s) {
  (void)nvar_acc;
  hipLaunchKernelGGL(k_window_volume, dim3((sh.vol + 255) / 256), dim3(256), 0, s, acc, W, sh.vol, out);
  return hipGetLastError();
}

}  // namespace pgsvt
