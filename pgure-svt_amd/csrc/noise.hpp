// noise.hpp -- host Poisson-Gaussian noise estimator of the PGURE-SVT path.
#pragma once

#include <array>
#include <cstdint>
#include <functional>
#include <memory>
#include <stdexcept>
#include <vector>

namespace pgsvt {

// An estimate the reference cannot form: noise method 1 reads the first
// round(0.05 * L) + 1 leaf statistics of the window (L = pixels per frame /
// leaves), through a bounds-checked copy that aborts the reference when
// fewer leaves exist (src/noise.hpp:115-121, 384-392).
struct NoiseError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// Robust (mean, variance) of every quadtree leaf of one normalised frame, in
// leaf order (src/noise.hpp:55-100).
struct FrameLeaves {
  std::vector<double> mean, var;
};

// NoiseEstimator::Estimate (src/noise.hpp:35-153) for a window of nt frames,
// each n x n and already normalised (frame k at u + k*n*n).  alpha/mu/sigma
// hold user values on entry (< 0 = estimate) and the estimates on return.
void estimate_noise(const double* u, uint32_t n, uint32_t nt, uint32_t method, double& alpha, double& mu,
                    double& sigma);

// Leaf statistics of one normalised n x n frame.
FrameLeaves frame_leaves(const double* f, uint32_t n);

// Split test of the quadtree: split(i, j, N) for the N x N block at row i,
// column j of the frame.
using SplitTest = std::function<bool(uint32_t, uint32_t, uint32_t)>;

// QuadTree (src/noise.hpp:55-100, 419-458): the leaves (row, column, size)
// in the reference's order, its repeated nodes included, for a split test.
std::vector<std::array<uint32_t, 3>> quadtree_leaves(uint32_t n, const SplitTest& split);

// F-test threshold of SplitBlockQ for N x N blocks (src/noise.hpp:173-221),
// or a negative value when such a block is never split (N <= 8, or N not in
// the table).
double split_threshold(uint32_t N);

// The window-level part of the estimate (robust line fit and the noise
// method) from the leaf statistics of its frames, in frame order.
void estimate_from_leaves(const std::vector<std::shared_ptr<const FrameLeaves>>& frames, uint32_t n,
                          uint32_t method, double& alpha, double& mu, double& sigma);

}  // namespace pgsvt
