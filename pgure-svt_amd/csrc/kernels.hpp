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
  uint32_t fstride;     // floats per stored factor block (V n x n, s n)
};

bool svd_shape_supported(uint32_t bs, uint32_t nw);

hipError_t launch_to_double(const void* raw, int dtype, double* out, size_t n, hipStream_t s);
hipError_t launch_frame_max(const double* X, uint32_t fs, uint32_t nf, double* out, hipStream_t s);
hipError_t launch_median(const double* X, double* Z, uint32_t nr, uint32_t nc, uint32_t nf, int r, hipStream_t s);
hipError_t launch_delta2(int64_t seed, uint64_t nvox, uint8_t* out, hipStream_t s);
hipError_t launch_prepare(const double* X, const double* Z, const WinMeta* meta, uint32_t nb, const uint8_t* d2,
                          const Shape& sh, uint32_t nvar, double eps2, double d2neg, double d2pos, double* u64,
                          double* w64, float* u32, hipStream_t s);
hipError_t launch_fill_ref(const WinMeta* meta, uint32_t nb, const Shape& sh, int32_t* pos, hipStream_t s);
hipError_t launch_arps(const double* w64, const int16_t* jobs, uint32_t max_jobs, uint32_t nb, const Shape& sh,
                       int mw, int32_t* pos, hipStream_t s);
hipError_t launch_count(const int32_t* pos, const uint32_t* sel, uint32_t nb, const Shape& sh, int32_t* W,
                        hipStream_t s);
hipError_t launch_svd(const float* u32, const int32_t* pos, const uint32_t* sel, uint32_t nb, uint32_t nvar,
                      const Shape& sh, float* factors, int max_sweeps, float tol, uint32_t* stats, hipStream_t s);
// wl: device list of batch slots to process (nb entries)
hipError_t launch_recon(const float* factors, const float* u32, const int32_t* pos, const uint32_t* sel,
                        const WinMeta* meta, const uint32_t* wl, uint32_t nb, uint32_t nvar_stored,
                        uint32_t nvar_run, const Shape& sh, int exp_w, int only_ref, unsigned long long* acc,
                        hipStream_t s);
// fused lambda evaluation: threshold + reconstruct all three variants, the
// variant-0 volume into acc ([nb][vol]) and the linear risk sums; then the
// volume sums.  sums: double [na][6] = (sum Uhat^2, sum U, sum U^2,
// sum Uhat U, sum d2 (U2p - 2 Uhat + U2m), sum Uhat) per listed slot.
uint32_t recon_risk_blocks(const Shape& sh);
hipError_t launch_recon_risk(const float* factors, const float* u32, const int32_t* pos, const uint32_t* sel,
                             const WinMeta* meta, const uint32_t* wl, uint32_t na, const Shape& sh, int exp_w,
                             const double* invw, const double* u64, const uint8_t* d2, double d2neg, double d2pos,
                             unsigned long long* acc, double* partial_a, double* partial_b, uint32_t nblk_b,
                             double* sums, hipStream_t s);
hipError_t launch_invw(const int32_t* W, size_t n, double* invw, hipStream_t s);
hipError_t launch_output(unsigned long long* acc, const int32_t* W, const WinMeta* meta, uint32_t nb,
                         uint32_t nvar_acc, const Shape& sh, double* Y, hipStream_t s);
hipError_t launch_window_volume(unsigned long long* acc, const int32_t* W, uint32_t nvar_acc, const Shape& sh,
                                double* out, hipStream_t s);

}  // namespace pgsvt
