// kernels.hpp -- launch wrappers of the gfx950 kernels (kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.hpp"

namespace pgsvt {

// Shapes shared by every launch of one denoise call.
struct Shape {
  uint32_t nr, nc;      // frame rows / cols (square in practice)
  uint32_t fs;          // nr * nc
  uint32_t nw;          // window frames (2*floor(Nt/2)+1)
  uint32_t vol;         // nw * fs
  uint32_t bs;          // patch size
  uint32_t m;           // bs*bs
  uint32_t n1;          // nr - bs + 1 (grid positions per axis)
  uint32_t p_all;       // n1 * n1
  uint32_t p_sel;       // selected patches (src/svt.hpp:61-92)
  uint32_t fstride;     // values per stored factor block (factor_stride)
  bool fp64 = false;    // SVD input, factors and reconstructions in FP64 (precision study)
};

// (patch size, window length) pairs the kernels run: the instantiated table
// in FP32 and FP64, any window up to 15 frames (patch 16) or 31 (patches
// 5 to 8) and every window of patches up to 4 x 4 in FP32
bool svd_shape_supported(uint32_t bs, uint32_t nw, bool fp64);

hipError_t launch_to_double(const void* raw, int dtype, double* out, size_t n, hipStream_t s);
hipError_t launch_frame_max(const double* X, uint32_t fs, uint32_t nf, double* out, hipStream_t s);
hipError_t launch_median(const double* X, double* Z, uint32_t nr, uint32_t nc, uint32_t nf, int r, hipStream_t s);
hipError_t launch_delta2(int64_t seed, uint64_t nvox, uint8_t* out, hipStream_t s);
// uvar: the SVD input variants, float, or double when sh.fp64
hipError_t launch_prepare(const double* X, const double* Z, const WinMeta* meta, uint32_t nb, const uint8_t* d2,
                          const Shape& sh, uint32_t nvar, double eps2, double d2neg, double d2pos, double* u64,
                          double* w64, void* uvar, hipStream_t s);
hipError_t launch_fill_ref(const WinMeta* meta, uint32_t nb, const Shape& sh, int32_t* pos, hipStream_t s);
hipError_t launch_arps(const double* w64, const int16_t* jobs, uint32_t max_jobs, uint32_t nb, const Shape& sh,
                       int mw, int32_t* pos, hipStream_t s);
// SVD of the first nvar of the nvar_stored variants of every window (uvar
// [nb][nvar_stored][vol] -> factors [nb][nvar_stored][p_sel][fstride]).
// mapping 0: the default lanes per matrix of the shape, 1: the alternative
// (FP32 only; see svd_lanes).
hipError_t launch_svd(const void* uvar, const int32_t* pos, const uint32_t* sel, uint32_t nb, uint32_t nvar,
                      uint32_t nvar_stored, const Shape& sh, void* factors, int max_sweeps, double tol,
                      uint32_t* stats, hipStream_t s, int mapping = 0);
// the static data of the fused evaluation (patches up to 4 x 4) from the U
// variant's factors: w_0 v_0[k] into W0 [nb][nw][p_sel][m] and the singular
// pairs (s_0, s_1) into S01 [nb][nw][p_sel][2], at the bucket places place
// [nb][nw][p_sel] (the buckets' inverse); for every frame (nk = 0), or for
// the frames that gathers with frame subset nk and the final frames
// (meta[b].ref) read
hipError_t launch_static_rows(const void* factors, const int32_t* place, const WinMeta* meta, uint32_t nb,
                              uint32_t nvar_stored, const Shape& sh, void* W0, void* S01, hipStream_t s,
                              uint32_t nk = 0);
uint32_t svd_lanes(uint32_t bs, uint32_t nw, int mapping);
// position buckets of every (window, frame): cnt and off are [nb][nw][n1*n1
// (+1)], list [nb][nw][p_sel]
// inv [nb][nw][p_sel]: bucket-sorted place of every patch
hipError_t launch_buckets(const int32_t* pos, const uint32_t* sel, uint32_t nb, const Shape& sh, int32_t* cnt,
                          int32_t* off, int32_t* list, int32_t* inv, hipStream_t s);
// 1/W and delta2/W per voxel, [nb][vol] (dw may be null)
hipError_t launch_weights(const int32_t* off, const uint8_t* d2, uint32_t nb, const Shape& sh, double d2neg,
                          double d2pos, double* invw, void* dw, hipStream_t s);
// one lambda evaluation of the listed slots; B0 [na][nw][p_sel][m] (rows in
// bucket order); partial [na][eval_blocks] (divergence sums, only with
// nvar_run == 3, which needs divq); kept (optional, nvar_run == 3):
// [na][eval_blocks] kept ranks of each block
uint32_t eval_blocks(const Shape& sh, uint32_t nvar_run);
hipError_t launch_eval(const void* factors, const WinMeta* meta, const uint32_t* wl, uint32_t na,
                       uint32_t nvar_stored, uint32_t nvar_run, const Shape& sh, int exp_w, const void* divq,
                       const int32_t* inv, void* B0, double* partial, hipStream_t s,
                       const double* slot_lam = nullptr, double* kept = nullptr);
// divergence factors w_r^T D v_r of every rank of every (patch, variant),
// D the patch's delta2/W along its trajectory (dw), after the SVD: divq
// [nb][p_sel][nvar][2][sr] (sr = nw rounded up to 4), per (patch, variant)
// its singular values (zero-padded) then its factors
hipError_t launch_divq(const void* factors, const int32_t* pos, const uint32_t* sel, const void* dw, uint32_t nb,
                       uint32_t nvar, const Shape& sh, void* divq, hipStream_t s);
// Patches up to 4 x 4 evaluate lambda in two launches: launch_round_prep
// (per patch that keeps rank 1: the factors d_r of its further kept ranks
// into dr [na][p_sel][nw], ended by a 0; with PGURE (qdiv = divq) the
// divergence sums into partial [na][eval_blocks]), then launch_gather,
// which forms d_0 and the patch rows from the window's static singular
// pairs and rows (written by launch_static_rows: S01 [nb][nw][p_sel][2], W0
// [nb][nw][p_sel][m] in bucket order; singular_pair_bytes per window), dr
// and the factors as it gathers.  Larger patches: launch_eval writes the
// rows to B0, launch_gather reads them.
bool fused_eval(const Shape& sh);
size_t singular_pair_bytes(const Shape& sh);
hipError_t launch_round_prep(const void* factors, const void* qdiv, const WinMeta* meta, const uint32_t* wl,
                             uint32_t na, uint32_t nvar_stored, const Shape& sh, int exp_w, const double* slot_lam,
                             void* dr, double* partial, double* kept, hipStream_t s);
// One lambda evaluation of the listed slots, either form: `rows` is dr
// (fused) or B0.
hipError_t launch_evaluation(const Shape& sh, const void* factors, const void* divq, const WinMeta* meta,
                             const uint32_t* wl, uint32_t na, uint32_t nvar_stored, uint32_t nvar_run, int exp_w,
                             const int32_t* inv, void* rows, double* partial, const double* slot_lam, double* kept,
                             hipStream_t s);
// what the gather reads: fused (bs <= 4) the factors, dr, the static
// singular pairs and rows with the bucket lists, and the lambdas (slot_lam
// per slot, or the windows' meta) with the weighting; otherwise the rows B0
struct GatherSrc {
  const void* B0 = nullptr;
  const void* factors = nullptr;
  const void* dr = nullptr;
  const void* S01 = nullptr;
  const void* W0 = nullptr;
  const int32_t* list = nullptr;
  uint32_t nvar_stored = 1;
  const double* slot_lam = nullptr;
  int exp_w = 0;
};
// values per stored factor block of one (patch, variant): V^T (nw rows of
// sr = nw rounded up to 4 values), W^T (nw rows of m values), the nw
// singular values, rounded up to an odd multiple of 32 values: every block
// starts a 128-byte line, so no line is written by two matrices (which run
// in different waves, often on different XCDs), and the stride is not a
// power of two (2048 bytes for 16 x 15 matrices slowed the SVD's stores by
// a tenth: they met in the same memory channels)
uint32_t factor_stride(uint32_t nw, uint32_t m);
// overlap-add gather: mode 0 risk sums into partial [na][gather_blocks][5]; mode 1
// final reference frames into out (the output cube); mode 2 the whole
// window of slot 0 into out.  nk (mode 0, patches up to 4 x 4): the sums
// cover nk frames spread over the window, k_i = i (nw - 1) / (nk - 1)
// (gather_frames: nw when nk is 0, above nw, or the patches are larger)
uint32_t gather_frames(const Shape& sh, uint32_t nk);
uint32_t gather_blocks(const Shape& sh, int mode, uint32_t nk = 0);
hipError_t launch_gather(const GatherSrc& src, const int32_t* off, const double* invw, const double* u64,
                         const WinMeta* meta, const uint32_t* wl, uint32_t na, const Shape& sh, int mode, double* out,
                         double* partial, hipStream_t s, uint32_t nk = 0);
// sums [na][6] = (sum Uhat^2, sum U, sum U^2, sum Uhat U,
// sum d2 (U2p - 2 Uhat + U2m), sum Uhat); with pk (the eval's kept-rank
// counts), kept [na] = their sum per window
hipError_t launch_risk_reduce(const double* pa, uint32_t na_blk, const double* pb, uint32_t nb_blk, uint32_t na,
                              double* sums, hipStream_t s, const double* pk = nullptr, double* kept = nullptr);

// hot-pixel filter of uint16 frames in place (one workgroup per frame);
// scratch: nf * nc uint32
hipError_t launch_hotpixel(uint16_t* seq, uint32_t nr, uint32_t nc, uint32_t nf, double threshold,
                           uint32_t* scratch, hipStream_t s);

// learned-lambda network (mlp.hip): C(m, n) = act(sum_k A(m, k) B(k, n) +
// bias[n]) with A(m, k) = A[m*sam + k*sak], B(k, n) = B[k*sbk + n*sbn], bf16
// operands, FP32 sums; act 1 = GELU (pre-activation kept in pre if set)
hipError_t launch_gemm(const float* A, int64_t sam, int64_t sak, const float* B, int64_t sbk, int64_t sbn, float* C,
                       int64_t ldc, int M, int N, int K, const float* bias, int act, float* pre, hipStream_t s);
hipError_t launch_gelu_bwd(const float* pre, const float* dy, float* dx, int64_t n, hipStream_t s);
hipError_t launch_colsum(const float* D, int rows, int cols, int64_t ld, float* out, hipStream_t s);
hipError_t launch_mse(const float* y, int64_t ldy, const float* t, int n, float* dy, float* loss, hipStream_t s);
hipError_t launch_adam(float* w, const float* g, float* m, float* v, int64_t n, float lr, float b1, float b2,
                       float eps, int step, hipStream_t s);
hipError_t launch_gather_rows(const float* X, const uint32_t* idx, uint32_t rows, uint32_t cols, float* out,
                              hipStream_t s);
// log singular values of sampled patches: feat[b*ld + i*n + r], standardised
// with mean / inv_std when given
hipError_t launch_features(const float* factors, uint32_t nvar_stored, uint32_t p_sel, uint32_t fstride,
                           uint32_t s_at, uint32_t n, const uint32_t* sample, uint32_t nsamp, uint32_t nb, uint32_t ld,
                           const float* mean, const float* inv_std, float* feat, hipStream_t s);

// fill n 32-bit words with v (graph-capturable, unlike a memset node)
hipError_t launch_fill32(void* p, uint64_t n, uint32_t v, hipStream_t s);
// rows of `width` 32-bit words, batch b: dst row r = src row r * step
// (batch strides src_batch and dst_batch in words)
hipError_t launch_copy_rows(const void* src, void* dst, uint32_t rows, uint32_t width, uint32_t step,
                            size_t src_batch, size_t dst_batch, uint32_t nb, hipStream_t s);

// Noise estimator's quadtree on the GPU (noise_gpu.hip).  Split tests of
// blocks [nblocks][3] (row, column, size) of frames frames [nf] of the FP64
// cube X (n x n frames) normalised by scale: stat [nf][nblocks] = the F
// statistic of SplitBlockQ.  Leaf statistics of leaves [nleaves][4] (frame,
// row, column, size): out [nleaves][2] = (robust mean, robust variance).
// Leaves of more than leaf_cache_elements() elements: those listed in big
// [nbig] have their values and Laplacians formed once into vals and laps
// from element boff[leaf] (boff [nleaves]); the others (boff ~0) form them
// again at every pass.
hipError_t launch_split_stats(const double* X, uint32_t n, double scale, const uint32_t* frames, uint32_t nf,
                              const uint32_t* blocks, uint32_t nblocks, double* stat, hipStream_t s);
uint32_t leaf_cache_elements();
hipError_t launch_leaf_stats(const double* X, uint32_t n, double scale, const uint32_t* leaves, uint32_t nleaves,
                             const uint32_t* big, uint32_t nbig, const uint64_t* boff, double* vals, double* laps,
                             double* out, hipStream_t s);
// lambda grid: G candidates per window from the sampled patches' median
// singular value; then risk of every candidate from the sums [nb*G][6] and
// the argmin into meta[b].lambda (risk [nb*G] kept)
hipError_t launch_grid_lambdas(const void* factors, uint32_t nvar_stored, const Shape& sh, uint32_t s_at,
                               const uint32_t* sample, uint32_t nsamp, uint32_t nb, uint32_t G, int exp_w,
                               double* lam, hipStream_t s);
hipError_t launch_grid_select(const double* sums, const double* noise, uint32_t nb, uint32_t G, double OoN,
                              double eps2, const double* lam, WinMeta* meta, double* risk, hipStream_t s);

}  // namespace pgsvt
