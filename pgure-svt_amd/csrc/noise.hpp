// noise.hpp -- host Poisson-Gaussian noise estimator of the PGURE-SVT path.
#pragma once

#include <cstdint>
#include <functional>
#include <memory>
#include <mutex>
#include <unordered_map>
#include <vector>

namespace pgsvt {

// Robust (mean, variance) of every quadtree leaf of one normalised frame, in
// leaf order (src/noise.hpp:55-100).
struct FrameLeaves {
  std::vector<double> mean, var;
};

// Per-frame leaf statistics depend only on the frame and on the window
// maximum it is normalised by, so windows that share both share the result.
class LeafCache {
 public:
  // leaves of `frame` normalised by `scale`; `fill(buf)` writes the
  // normalised n x n frame into buf and is only called on a miss
  std::shared_ptr<const FrameLeaves> get(uint32_t frame, double scale, uint32_t n,
                                         const std::function<void(double*)>& fill);

 private:
  std::mutex mu_;
  std::unordered_map<uint64_t, std::shared_ptr<const FrameLeaves>> map_;
  std::unordered_map<uint64_t, double> scale_of_;
};

// NoiseEstimator::Estimate (src/noise.hpp:35-153) for a window of nt frames,
// each n x n and already normalised (frame k at u + k*n*n).  alpha/mu/sigma
// hold user values on entry (< 0 = estimate) and the estimates on return.
void estimate_noise(const double* u, uint32_t n, uint32_t nt, uint32_t method, double& alpha, double& mu,
                    double& sigma);

// Leaf statistics of one normalised n x n frame.
FrameLeaves frame_leaves(const double* f, uint32_t n);

// The window-level part of the estimate (robust line fit and the noise
// method) from the leaf statistics of its frames, in frame order.
void estimate_from_leaves(const std::vector<std::shared_ptr<const FrameLeaves>>& frames, uint32_t n,
                          uint32_t method, double& alpha, double& mu, double& sigma);

}  // namespace pgsvt
