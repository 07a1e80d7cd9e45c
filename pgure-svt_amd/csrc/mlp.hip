// mlp.hip -- the learned-lambda proposal network on the GPU: BF16 matrix
// cores with FP32 accumulation for every product (forward and backward),
// FP32 master weights, Adam.  No reference counterpart: the reference
// always searches lambda (src/pgure.hpp:196-237).
//
// Also the feature gather: the log singular values of a fixed sample of
// patches of each window, read from the sorted SVD factors.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "kernels.hpp"

namespace pgsvt {
namespace {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// float -> bfloat16 bits, round to nearest even (NaN kept a NaN)
__device__ __forceinline__ short to_bf16(float x) {
  uint32_t u = __float_as_uint(x);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (short)((u >> 16) | 0x40);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}

__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.f + erff(x * 0.70710678118654752f)); }

__device__ __forceinline__ float gelu_grad(float x) {
  return 0.5f * (1.f + erff(x * 0.70710678118654752f)) + x * 0.39894228040143268f * expf(-0.5f * x * x);
}

// C(m, n) = act(sum_k A(m, k) B(k, n) + bias[n]) with A(m, k) = A[m*sam +
// k*sak], B(k, n) = B[k*sbk + n*sbn]; operands rounded to bf16, products
// summed in FP32 by v_mfma_f32_16x16x16_bf16.  One wave per 16 x 16 tile
// of C.  Lane l holds A rows l%16, k = 4(l/16)..+3 and B columns l%16,
// same k; it receives C rows 4(l/16)..+3 of column l%16.
__global__ __launch_bounds__(256) void k_gemm(const float* __restrict__ A, int64_t sam, int64_t sak,
                                              const float* __restrict__ B, int64_t sbk, int64_t sbn,
                                              float* __restrict__ C, int64_t ldc, int M, int N, int K,
                                              const float* __restrict__ bias, int act, float* __restrict__ pre) {
  const int wave = (int)((blockIdx.x * blockDim.x + threadIdx.x) / 64), lane = (int)(threadIdx.x % 64);
  const int tiles_n = (N + 15) / 16;
  const int tm = wave / tiles_n, tn = wave % tiles_n;
  if (tm * 16 >= M) return;  // uniform per wave: the matrix op runs with every lane
  const int i = tm * 16 + lane % 16, j = tn * 16 + lane % 16, kq = 4 * (lane / 16);
  v4f acc = {0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += 16) {
    v4s a, b;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int k = k0 + kq + v;
      a[v] = to_bf16((i < M && k < K) ? A[i * sam + k * sak] : 0.f);
      b[v] = to_bf16((j < N && k < K) ? B[k * sbk + j * sbn] : 0.f);
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
  }
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int r = tm * 16 + kq + v, c = tn * 16 + lane % 16;
    if (r < M && c < N) {
      const float x = acc[v] + (bias ? bias[c] : 0.f);
      if (pre) pre[r * ldc + c] = x;
      C[r * ldc + c] = act ? gelu(x) : x;
    }
  }
}

__global__ void k_gelu_bwd(const float* __restrict__ pre, const float* __restrict__ dy, float* __restrict__ dx,
                           int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dx[i] = dy[i] * gelu_grad(pre[i]);
}

// out[n] = sum over rows b of D[b*ld + n] (one thread per column, rows in order)
__global__ void k_colsum(const float* __restrict__ D, int rows, int cols, int64_t ld, float* __restrict__ out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cols) return;
  float s = 0.f;
  for (int b = 0; b < rows; ++b) s += D[b * ld + c];
  out[c] = s;
}

// mean squared error of y[b*ldy] against t[b]: dy = 2 (y - t) / n, loss
// in loss[0] (one block, FP64 sum)
__global__ __launch_bounds__(256) void k_mse(const float* __restrict__ y, int64_t ldy, const float* __restrict__ t,
                                             int n, float* __restrict__ dy, float* __restrict__ loss) {
  __shared__ double red[256];
  double s = 0.0;
  for (int b = threadIdx.x; b < n; b += blockDim.x) {
    const float e = y[b * ldy] - t[b];
    dy[b * ldy] = 2.f * e / (float)n;
    s += (double)e * e;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss[0] = (float)(red[0] / n);
}

__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int64_t n, float lr, float b1, float b2, float eps, float bc1,
                       float bc2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

// feat[b][i*n + r] = log(max(s_r, 1e-6)) of sampled patch i of window b
// (r-th largest singular value), then (x - mean) * inv_std when mean is set
__global__ void k_features(const float* __restrict__ factors, uint32_t nvar_stored, uint32_t p_sel,
                           uint32_t fstride, uint32_t s_at, uint32_t n, const uint32_t* __restrict__ sample,
                           uint32_t nsamp, uint32_t ld, const float* __restrict__ mean,
                           const float* __restrict__ inv_std, float* __restrict__ feat) {
  const uint32_t b = blockIdx.y;
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nsamp * n) return;
  const uint32_t i = e / n, r = e % n;
  const float s = factors[((size_t)b * nvar_stored * p_sel + sample[i]) * fstride + s_at + r];
  float x = logf(fmaxf(s, 1e-6f));
  if (mean) x = (x - mean[e]) * inv_std[e];
  feat[(size_t)b * ld + e] = x;
}

// out[r][c] = X[idx[r]][c]
__global__ void k_gather_rows(const float* __restrict__ X, const uint32_t* __restrict__ idx, uint32_t rows,
                              uint32_t cols, float* __restrict__ out) {
  const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (uint64_t)rows * cols) return;
  const uint32_t r = (uint32_t)(e / cols), c = (uint32_t)(e % cols);
  out[e] = X[(size_t)idx[r] * cols + c];
}

}  // namespace

hipError_t launch_gather_rows(const float* X, const uint32_t* idx, uint32_t rows, uint32_t cols, float* out,
                              hipStream_t s) {
  const uint64_t n = (uint64_t)rows * cols;
  hipLaunchKernelGGL(k_gather_rows, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, X, idx, rows, cols, out);
  return hipGetLastError();
}

hipError_t launch_gemm(const float* A, int64_t sam, int64_t sak, const float* B, int64_t sbk, int64_t sbn, float* C,
                       int64_t ldc, int M, int N, int K, const float* bias, int act, float* pre, hipStream_t s) {
  const int64_t waves = (int64_t)((M + 15) / 16) * ((N + 15) / 16);
  const int64_t blocks = (waves + 3) / 4;
  hipLaunchKernelGGL(k_gemm, dim3((uint32_t)blocks), dim3(256), 0, s, A, sam, sak, B, sbk, sbn, C, ldc, M, N, K,
                     bias, act, pre);
  return hipGetLastError();
}

hipError_t launch_gelu_bwd(const float* pre, const float* dy, float* dx, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_gelu_bwd, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, pre, dy, dx, n);
  return hipGetLastError();
}

hipError_t launch_colsum(const float* D, int rows, int cols, int64_t ld, float* out, hipStream_t s) {
  hipLaunchKernelGGL(k_colsum, dim3((cols + 255) / 256), dim3(256), 0, s, D, rows, cols, ld, out);
  return hipGetLastError();
}

hipError_t launch_mse(const float* y, int64_t ldy, const float* t, int n, float* dy, float* loss, hipStream_t s) {
  hipLaunchKernelGGL(k_mse, dim3(1), dim3(256), 0, s, y, ldy, t, n, dy, loss);
  return hipGetLastError();
}

hipError_t launch_adam(float* w, const float* g, float* m, float* v, int64_t n, float lr, float b1, float b2,
                       float eps, int step, hipStream_t s) {
  const float bc1 = 1.f - std::pow(b1, (float)step), bc2 = 1.f - std::pow(b2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, w, g, m, v, n, lr, b1, b2, eps, bc1,
                     bc2);
  return hipGetLastError();
}

hipError_t launch_features(const float* factors, uint32_t nvar_stored, uint32_t p_sel, uint32_t fstride,
                           uint32_t s_at, uint32_t n, const uint32_t* sample, uint32_t nsamp, uint32_t nb, uint32_t ld,
                           const float* mean, const float* inv_std, float* feat, hipStream_t s) {
  const dim3 grid((nsamp * n + 255) / 256, nb);
  hipLaunchKernelGGL(k_features, grid, dim3(256), 0, s, factors, nvar_stored, p_sel, fstride, s_at, n, sample, nsamp,
                     ld, mean, inv_std, feat);
  return hipGetLastError();
}

}  // namespace pgsvt
