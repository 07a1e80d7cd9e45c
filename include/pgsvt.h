/*
 * pgsvt.h -- C ABI of the MI355X-native PGURE-SVT patchwise singular-value
 * thresholding path (library: pgure-svt_amd/pguresvt/libpgsvt.so).
 *
 * Plain pointers and sizes only.  All volumes are column-major "cubes" in the
 * reference's Armadillo layout: element (row r, col c, frame f) lives at
 * r + n_rows * (c + n_cols * f).
 *
 * Every entry point returns 0 on success or a negative PGSVT_E* code; the
 * message of the last failure on the calling thread is in pgsvt_last_error().
 */
#ifndef PGSVT_H
#define PGSVT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Input element types: the four instantiations PGURESVT<T1,double> of the
 * reference (pguresvt/_pguresvt.pyx:143-148, 181, 240, 299, 358). */
enum pgsvt_dtype { PGSVT_U8 = 0, PGSVT_U16 = 1, PGSVT_F32 = 2, PGSVT_F64 = 3 };

enum pgsvt_status {
  PGSVT_OK = 0,
  PGSVT_E_ARG = -1,      /* invalid argument / shape the reference cannot run */
  PGSVT_E_HIP = -2,      /* HIP runtime failure (allocation, launch, copy)   */
  PGSVT_E_NUMERIC = -3,  /* non-finite input or optimiser failure            */
  PGSVT_E_NOGPU = -4     /* no usable gfx950 device                          */
};

/* The 19 scalar parameters of the reference driver
 * PGURESVT(Y, estimates, X, trajLength, blockSize, blockOverlap, motionWindow,
 *          medianSize, noiseMethod, maxIter, nJobs, randomSeed, optimizePGURE,
 *          expWeighting, motionEstimation, lambdaEst, alphaEst, muEst,
 *          sigmaEst, tol)                       (src/pguresvt.hpp:17-38)
 * with the Python keyword names of pguresvt_{u8,u16,f,d}
 * (pguresvt/_pguresvt.pyx:151-168) and the same sentinels (-1 = estimate). */
typedef struct pgsvt_params {
  uint32_t trajectory_length;     /* trajLength, default 15                  */
  uint32_t patch_size;            /* blockSize, default 4                    */
  uint32_t patch_overlap;         /* blockOverlap (patch step), default 1    */
  uint32_t motion_window;         /* ARPS search half-width, default 7       */
  int64_t motion_filter;          /* medianSize; <=0 disables, default 5     */
  uint32_t noise_method;          /* 1..4, default 4                         */
  uint32_t max_iter;              /* optimiser evaluation cap, default 500   */
  int64_t n_jobs;                 /* host threads, -1 = all                  */
  int64_t random_seed;            /* pcg64 seed; <0 = nondeterministic       */
  int32_t optimize_pgure;         /* bool, default 1                         */
  int32_t exponential_weighting;  /* bool, default 1                         */
  int32_t motion_estimation;      /* bool, default 1                         */
  double lambda1;                 /* lambdaEst; start / fixed threshold      */
  double noise_alpha;             /* alphaEst, <0 = estimate                 */
  double noise_mu;                /* muEst,    <0 = estimate                 */
  double noise_sigma;             /* sigmaEst, <0 = estimate                 */
  double tol;                     /* relative f tolerance, default 1e-7      */
} pgsvt_params;

/* Fill *p with the defaults of pguresvt_{u8,u16,f,d}
 * (pguresvt/_pguresvt.pyx:151-168). */
void pgsvt_default_params(pgsvt_params* p);

/* Drop-in for PGURESVT<T1,double> (src/pguresvt.hpp:17-172) as bound by the
 * Cython entry points pguresvt_{u8,u16,f,d} (pguresvt/_pguresvt.pyx:151-384).
 *   X          input cube, n_rows x n_cols x n_frames, element type `dtype`
 *   Y          caller-allocated output cube of the same shape (float64)
 *   estimates  caller-allocated n_frames x 4 column-major matrix (float64):
 *              column 0 = lambda, 1 = alpha, 2 = mu, 3 = sigma per frame
 *              (src/pguresvt.hpp:150-153)
 * Runs on the current HIP device.  Deviation from the reference: inputs the
 * reference aborts on (frames < window, non-square frames, NaN) return
 * PGSVT_E_ARG / PGSVT_E_NUMERIC instead of terminating the process. */
int pgsvt_denoise(const void* X, int dtype, uint32_t n_rows, uint32_t n_cols,
                  uint32_t n_frames, const pgsvt_params* p, double* Y,
                  double* estimates);

/* Message of the last failure on this thread ("" if none). */
const char* pgsvt_last_error(void);

/* Library version string. */
const char* pgsvt_version(void);

/* pgsvt_denoise with X and Y in device memory of the current HIP device
 * (same shapes, layout and element types); `estimates` stays a host
 * n_frames x 4 column-major float64 matrix.  The reference has no device
 * variant: it is the PGURESVT<T1,double> call (src/pguresvt.hpp:17-172) for
 * callers whose frames are already resident in HBM, so no host<->device copy
 * of the cube happens.  Returns after Y is complete. */
int pgsvt_denoise_device(const void* X_dev, int dtype, uint32_t n_rows,
                         uint32_t n_cols, uint32_t n_frames,
                         const pgsvt_params* p, double* Y_dev,
                         double* estimates);

/* pgsvt_denoise_device that allocates no device memory.  No reference
 * counterpart: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Windows
 * are processed one at a time in a statically sized device array of 64 MiB
 * that each device the library runs on holds until the process ends; a
 * shape whose window needs more returns PGSVT_E_ARG with the size in the
 * message (64 x 64 frames fit with PGURE, 128 x 128 with a fixed lambda at
 * patch 4 and 15 frames).  Work goes to `stream` (a hipStream_t; NULL = the
 * null stream) and the call returns after Y is complete.  Calls of one
 * process are serialised (they share the array, which keeps the last
 * call's contents); calls that overlap on one device from several
 * processes each use their own copy.  FP32 only (pgsvt_set_precision(32)).
 * Results equal pgsvt_denoise_device's bit for bit. */
int pgsvt_denoise_device_noalloc(const void* X_dev, int dtype, uint32_t n_rows,
                                 uint32_t n_cols, uint32_t n_frames,
                                 const pgsvt_params* p, double* Y_dev,
                                 double* estimates, void* stream);

/* pgsvt_denoise restricted to output frames t_begin .. t_end-1 (the windows
 * of those frames read the input frames they need from the whole cube).  Y
 * holds the t_end - t_begin output frames in order, estimates is a
 * (t_end - t_begin) x 4 column-major matrix.  This is the per-GPU share when
 * the frames of one sequence are split across GPUs: the driver loop of
 * src/pguresvt.hpp:90-166 over a sub-range, with identical results. */
int pgsvt_denoise_range(const void* X, int dtype, uint32_t n_rows,
                        uint32_t n_cols, uint32_t n_frames,
                        const pgsvt_params* p, uint32_t t_begin,
                        uint32_t t_end, double* Y, double* estimates);

/* ------------------------------------------------------------------------
 * Lambda grid and batches (no reference counterpart: the reference searches
 * lambda with SBPLX, src/pgure.hpp:196-237).  Each window's PGURE risk is
 * evaluated at n_lambda candidates log-spaced over [0.01, 10] times a scale
 * from the median singular value of 256 fixed-sample patches (1 / median^2
 * with exponential weighting, the median without); the first candidate of
 * least risk is used.  All GPU work of a batch of windows (decomposition,
 * grid, risks, choice, final frames) runs as one captured HIP graph; the
 * SVD's lanes-per-matrix mapping is autotuned per shape on first use.
 * ------------------------------------------------------------------------ */
int pgsvt_denoise_grid(const void* X, int dtype, uint32_t n_rows,
                       uint32_t n_cols, uint32_t n_frames,
                       const pgsvt_params* p, uint32_t n_lambda, double* Y,
                       double* estimates);

/* One sequence of a batch: the arguments of pgsvt_denoise, and its status
 * on return. */
typedef struct pgsvt_job {
  const void* X;
  int dtype;
  uint32_t n_rows, n_cols, n_frames;
  pgsvt_params params;
  double* Y;
  double* estimates;
  int status;
} pgsvt_job;

/* Independent sequences of uneven shapes: `concurrency` at a time (largest
 * first), each on its own stream; n_lambda > 0 uses the lambda grid, 0 the
 * reference search.  Returns 0, or the first failing job's status with its
 * index in the message; every job's own status is set. */
int pgsvt_denoise_batch(pgsvt_job* jobs, uint32_t n_jobs, uint32_t n_lambda,
                        uint32_t concurrency);

/* ------------------------------------------------------------------------
 * Learned lambda (no reference counterpart: the reference always searches
 * lambda, src/pgure.hpp:196-237).  A network proposes each window's lambda
 * from n_in = 256 * Nw + 3 features: log(max(s, 1e-6)) of the sorted
 * singular values of 256 fixed-sample patches of the window (patch-major),
 * then the window's noise estimates (alpha, mu, sigma) from leaf statistics
 * computed once per frame at the sequence maximum and rescaled to the
 * window's maximum.  Network: n_in -> hidden -> hidden -> hidden -> 1 with
 * GELU, BF16 matrix cores, FP32 sums.
 * ------------------------------------------------------------------------ */
typedef struct pgsvt_lambda_model {
  uint32_t n_in;         /* feature count, 256 * Nw + 3 */
  uint32_t hidden;       /* hidden width */
  const float* mean;     /* [n_in] feature means */
  const float* inv_std;  /* [n_in] reciprocal feature standard deviations */
  const float* params;   /* W1 [n_in][hidden], b1, W2 [hidden][hidden], b2,
                            W3 [hidden][hidden], b3, W4 [hidden][1], b4 */
  float y_mean, y_std;   /* log(lambda) = y_mean + y_std * network output */
} pgsvt_lambda_model;

/* The raw (unstandardised) features of output frames t_begin .. t_end-1
 * into features [t_end - t_begin][256 * Nw + 3]; estimates as in
 * pgsvt_denoise_range with lambda = 0. */
int pgsvt_learned_features(const void* X, int dtype, uint32_t n_rows,
                           uint32_t n_cols, uint32_t n_frames,
                           const pgsvt_params* p, uint32_t t_begin,
                           uint32_t t_end, float* features,
                           double* estimates);

/* pgsvt_denoise with each window's lambda proposed by the model instead of
 * searched (optimize_pgure and lambda1 are ignored; the proposal is clamped
 * to the search's bounds [0, max(100, start)]).  Each proposal is probed
 * (pgsvt_set_learned_fallback): the window's PGURE risk at lambda / (1 +
 * delta), lambda and lambda (1 + delta), from the perturbed decompositions
 * of every 16th patch and the reconstruction sums of 5 of the window's
 * frames.  Where the parabola through the three predicts a relative risk
 * gap above the limit, lambda moves to the parabola's minimum and is
 * probed again, up to `refine` times; a window still above the limit is
 * searched (its full decompositions, the PGURE risk as pgsvt_denoise
 * evaluates it, SBPLX from the last estimate to a relative tolerance of
 * 1e-5) and its frame replaced.  estimates holds the final lambda and the noise
 * features' (alpha, mu, sigma). */
int pgsvt_denoise_learned(const void* X, int dtype, uint32_t n_rows,
                          uint32_t n_cols, uint32_t n_frames,
                          const pgsvt_params* p,
                          const pgsvt_lambda_model* model, double* Y,
                          double* estimates);
/* the same with X and Y in device memory */
int pgsvt_denoise_learned_device(const void* X_dev, int dtype, uint32_t n_rows,
                                 uint32_t n_cols, uint32_t n_frames,
                                 const pgsvt_params* p,
                                 const pgsvt_lambda_model* model,
                                 double* Y_dev, double* estimates);

/* Probe of pgsvt_denoise_learned: step delta in (0, 1) (default 0.1), the
 * predicted relative risk gap above which a window's lambda is refined or
 * searched (default 0.01; gap <= 0 turns the probe off), and the probes at
 * refined lambdas before the search (0-16, default 2).  Process-wide.  No
 * reference counterpart. */
int pgsvt_set_learned_fallback(double delta, double gap, uint32_t refine);
/* Windows that fell back to the search in the calling thread's last
 * pgsvt_denoise_learned(_device) call. */
int pgsvt_learned_fallbacks(void);

/* The network alone, for training and its tests: forward of n rows of X
 * ([n][n_in], host) into out [n]; gradient of mean((y - t)^2) into grad
 * (parameter layout) and the loss; Adam training for `epochs` passes of
 * shuffled mini-batches (params in: initial, out: trained; epoch_loss
 * [epochs] optional). */
int pgsvt_mlp_forward(const float* params, uint32_t n_in, uint32_t hidden,
                      const float* X, uint32_t n, float* out);
int pgsvt_mlp_gradient(const float* params, uint32_t n_in, uint32_t hidden,
                       const float* X, const float* t, uint32_t n,
                       float* grad, float* loss);
int pgsvt_mlp_train(const float* X, const float* t, uint32_t n,
                    uint32_t n_in, uint32_t hidden, uint32_t epochs,
                    uint32_t batch, float lr, uint64_t seed, float* params,
                    float* epoch_loss);

/* Arithmetic of the SVD input, the Jacobi SVD, the stored factors and the
 * reconstructions for calls started afterwards, process-wide: 32 (default,
 * FP32 with FP64 risk sums) or 64 (everything FP64: the GPU precision
 * reference).  The learned-lambda path runs in 32 only.  No reference
 * counterpart (the reference is FP64 throughout). */
int pgsvt_set_precision(int bits);

/* Lanes-per-matrix mapping of the batched Jacobi SVD for FP32 calls started
 * afterwards, process-wide: 0 = the primary mapping of each shape (default),
 * 1 = the alternative mapping (e.g. 8 instead of 4 lanes per 16 x 15
 * matrix), 2 = the lambda-grid path times both on the first batch of each
 * shape and keeps the faster (other paths use 0).  The mappings sum the
 * column dot products in different orders, so results differ in the last
 * bits between them; with 2 they are reproducible within a process only.
 * No reference counterpart (the reference calls LAPACK per patch,
 * src/svt.hpp:111). */
int pgsvt_set_svd_mapping(int mode);

/* Device memory is cached per device between calls (buffers of one call
 * are reused by the next call of the same shape), and pinned host staging
 * memory process-wide.  pgsvt_release_cache waits for the current device
 * and frees its cache and the unused pinned blocks.  No reference
 * counterpart (the reference allocates per window on the host). */
int pgsvt_release_cache(void);

/* Per-stage timing of later pgsvt_denoise* calls (no reference
 * counterpart; the reference prints wall times only, src/pguresvt.hpp:168).
 * pgsvt_profile_enable(1) clears and starts collection, (0) stops it;
 * (2) also counts the ranks each lambda evaluation keeps ("eval_kept_ranks",
 * at the price of one more reduction and copy per search round).
 * GPU stages are timed with HIP events on the library's stream.
 * pgsvt_profile_read writes a JSON object
 *   {"<stage>": {"launches": n, "ms": total, "units": work units}, ...}
 * into buf and returns 0, or returns the buffer size needed (bytes,
 * including the terminating NUL) when buf is NULL or len is too small. */
void pgsvt_profile_enable(int on);
int pgsvt_profile_read(char* buf, uint64_t len);

/* ------------------------------------------------------------------------
 * Window-level entry points.  They expose the stages of one time window of
 * the driver (src/pguresvt.hpp:90-166) so a caller can check each stage on
 * its own; `t` is the output-frame index whose window is processed.
 * Window size Nw = 2*floor(Nt/2)+1 with Nt the clamped trajectory length
 * (src/pguresvt.hpp:57-58); P_all = (n_rows-bs+1)*(n_cols-bs+1).
 * ------------------------------------------------------------------------ */

/* ARPS trajectories of window t: MotionEstimator<T>::Estimate
 * (src/arps.hpp:52-134).  patches: int32 [Nw][P_all][2] with [..][0] = row,
 * [..][1] = col, i.e. the reference icube(2, P_all, Nw) in its own layout. */
int pgsvt_window_trajectories(const void* X, int dtype, uint32_t n_rows,
                              uint32_t n_cols, uint32_t n_frames,
                              const pgsvt_params* p, uint32_t t,
                              int32_t* patches);

/* PGURE risk of window t at n_lambda thresholds: PGURE<T>::CalculatePGURE
 * (src/pgure.hpp:120-137) after the constructor (src/pgure.hpp:26-93).
 * noise[3] = (alpha, mu, sigma) exactly as the driver holds them before the
 * constructor call (src/pguresvt.hpp:133 passes them as alpha, sigma, mu). */
int pgsvt_window_risk(const void* X, int dtype, uint32_t n_rows,
                      uint32_t n_cols, uint32_t n_frames,
                      const pgsvt_params* p, uint32_t t, const double* noise,
                      const double* lambdas, uint32_t n_lambda, double* risk);

/* Thresholded reconstruction of the whole window t at lambda:
 * SVT<T>::Reconstruct (src/svt.hpp:121-167) of the unperturbed data, in
 * normalised units (before the `*= uMax` of src/pguresvt.hpp:147).
 * vol: float64 n_rows x n_cols x Nw, column-major. */
int pgsvt_window_reconstruct(const void* X, int dtype, uint32_t n_rows,
                             uint32_t n_cols, uint32_t n_frames,
                             const pgsvt_params* p, uint32_t t, double lambda,
                             double* vol);

/* Singular values of every selected patch of window t
 * (arma::svd_econ at src/svt.hpp:111): float64 [P][Nw], each row sorted
 * descending.  Returns the patch count P through *n_patches. */
int pgsvt_window_singular_values(const void* X, int dtype, uint32_t n_rows,
                                 uint32_t n_cols, uint32_t n_frames,
                                 const pgsvt_params* p, uint32_t t,
                                 double* s, uint32_t* n_patches);

/* Poisson-Gaussian noise estimate of a normalised window u (float64 cube
 * n_rows x n_cols x n_t): NoiseEstimator::Estimate (src/noise.hpp:35-153).
 * On entry alpha, mu and sigma hold user values (<0 = estimate). */
int pgsvt_noise_estimate(const double* u, uint32_t n_rows, uint32_t n_cols,
                         uint32_t n_t, uint32_t method, double* alpha,
                         double* mu, double* sigma);

/* Robust (mean, variance) of every quadtree leaf of one normalised n x n
 * frame, in the tree's order (src/noise.hpp:55-100): on the host
 * (on_gpu = 0, the estimator the driver uses) or on the GPU (on_gpu = 1,
 * the learned path's; frames of at least 16 x 16).  *count receives the
 * leaf count; mean and var hold up to cap values.  For the tests. */
int pgsvt_frame_leaves(const double* frame, uint32_t n, int on_gpu,
                       uint32_t cap, double* mean, double* var,
                       uint32_t* count);

/* The delta2 perturbation signs the GPU generates: PGURE<T>::
 * GenerateRandomPerturbations (src/pgure.hpp:167-186) with pcg64 seeded by
 * `seed` (src/pgure.hpp:52-59).  out[v] = 1 where delta2[v] is negative
 * (-sqrt(vQ/vP)), 0 where it is +sqrt(vP/vQ); n_vox values. */
int pgsvt_delta2_signs(int64_t seed, uint64_t n_vox, uint8_t* out);

/* ------------------------------------------------------------------------
 * Front-end support (the reference's command-line program, src/PGURE-SVT.cpp)
 * ------------------------------------------------------------------------ */

/* HotPixelFilter (src/hotpixel.hpp:19-64) on a host uint16 cube
 * n_rows x n_cols x n_frames, in place, on the GPU: pixels further than
 * threshold * MAD from the frame median are replaced by the median of their
 * 8 neighbours (border pixels by the frame median), in column-major order.
 * Deviation: |x - median| is the exact distance (the reference's uint16
 * arithmetic wraps below the median). */
int pgsvt_hotpixel_filter(uint16_t* seq, uint32_t n_rows, uint32_t n_cols,
                          uint32_t n_frames, double threshold);

/* Multi-page grayscale TIFF (uncompressed, 8/16-bit) as the front end
 * reads and writes it through libtiff (src/PGURE-SVT.cpp:118-174,
 * 204-240).  Cubes are height x width x pages, column-major, element
 * (r, c, k) = row r, column c of page k. */
int pgsvt_tiff_info(const char* path, uint32_t* width, uint32_t* height,
                    uint32_t* bits, uint32_t* pages);
int pgsvt_tiff_read_u16(const char* path, uint32_t first, uint32_t count,
                        uint16_t* out);
int pgsvt_tiff_write_u16(const char* path, const uint16_t* cube,
                         uint32_t width, uint32_t height, uint32_t pages);

#ifdef __cplusplus
}
#endif

#endif /* PGSVT_H */
