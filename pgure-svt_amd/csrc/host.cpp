// host.cpp -- C ABI and host orchestration of the MI355X PGURE-SVT path.
//
// One call of pgsvt_denoise does what PGURESVT<T1,double> does
// (src/pguresvt.hpp:17-172), window by window, but windows are processed in
// batches that live in HBM together:
//   1. upload X, per-frame maxima, optional median pre-filter (GPU);
//   2. noise estimation per window on host threads (overlaps the GPU work);
//   3. per batch: normalised volumes, ARPS trajectories, overlap counts and
//      the batched Jacobi SVD of every patch of every window (GPU);
//   4. the lambda searches of all windows of the batch advance in lockstep:
//      each round evaluates one lambda per still-searching window in one
//      reconstruct + risk launch, then every search takes its next step;
//   5. the final reconstruction of each window's reference frame at its
//      lambda goes straight into the output cube.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <functional>
#include <fstream>
#include <future>
#include <map>
#include <memory>
#include <mutex>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include <sched.h>

#include "../../include/pgsvt.h"
#include "common.hpp"
#include "kernels.hpp"
#include "noise.hpp"
#include "sbplx.hpp"
#include "tiff.hpp"

namespace pgsvt {
namespace {

thread_local std::string g_err;

// arithmetic of the SVD / reconstruction stages for calls that start from
// now on: 32 (default) or 64 bits, see pgsvt_set_precision
std::atomic<int> g_precision{32};

// SVD lanes-per-matrix mapping (pgsvt_set_svd_mapping): 0 the primary
// mapping of each shape (default), 1 the alternative, 2 autotuned per shape
// by the grid path
std::atomic<int> g_svd_mapping{0};

// Learned-lambda fallback (pgsvt_set_learned_fallback): the probe's step
// delta (risks at lambda / (1 + delta), lambda, lambda (1 + delta)) and the
// predicted relative risk gap above which a window is searched; gap <= 0
// turns the probe off.  Windows searched by the last learned call of this
// thread: pgsvt_learned_fallbacks.
std::mutex g_fallback_mu;
double g_probe_delta = 0.1, g_probe_gap = 0.01;
uint32_t g_probe_refine = 2;
thread_local int g_last_fallbacks = 0;
// probe sample: every s-th selected patch in all three variants, s at most
// kProbeMaxStride and small enough to keep kProbeMinSample patches (the
// divergence sum is scaled up from the sample), the data terms over
// kProbeFrames frames spread over the window
constexpr uint32_t kProbeMaxStride = 64, kProbeMinSample = 1000, kProbeFrames = 5;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

struct HipError {
  hipError_t e;
  const char* what;
};

// Fault injection for tests of the error paths: PGSVT_FAULT="alloc:N" makes
// the N-th device buffer request of a call fail as out of memory (whether
// the cache or hipMalloc would serve it), "hip:N" the
// N-th checked HIP call (launch, copy, synchronisation) fail as a launch
// failure, after the real call succeeded.  Read at the start of every
// denoise call (fault_arm) and counted per calling thread; the failure
// unwinds like a real one.
struct FaultPlan {
  int kind = 0;  // 0 none, 1 allocation, 2 HIP call
  long left = 0;
};
thread_local FaultPlan g_fault;

void fault_arm() {
  g_fault = FaultPlan{};
  const char* e = std::getenv("PGSVT_FAULT");
  if (!e) return;
  const std::string v(e);
  const size_t c = v.find(':');
  if (c == std::string::npos) return;
  const std::string kind = v.substr(0, c);
  g_fault.left = std::atol(v.c_str() + c + 1);
  g_fault.kind = kind == "alloc" ? 1 : kind == "hip" ? 2 : 0;
}

bool fault_hit(int kind) {
  if (g_fault.kind != kind || g_fault.left <= 0) return false;
  if (--g_fault.left > 0) return false;
  g_fault.kind = 0;
  return true;
}

inline void hipchk(hipError_t e, const char* what) {
  if (e == hipSuccess && fault_hit(2)) throw HipError{hipErrorLaunchFailure, "injected failure after"};
  if (e != hipSuccess) throw HipError{e, what};
}

// ---- kernel timing (HIP events on the library stream), off by default ----
struct ProfEntry {
  std::string name;
  uint64_t launches = 0;
  double ms = 0, units = 0;
};
std::mutex g_prof_mu;
bool g_prof_on = false;
// also count the ranks each evaluation keeps (one more reduction and copy
// per search round)
bool g_prof_count = false;
std::vector<ProfEntry> g_prof;

struct ProfPending {
  int idx;
  hipEvent_t a, b;
  double units;
  bool own_a, own_b;  // events returned to the pool with this entry
};

class Profiler {
 public:
  explicit Profiler(hipStream_t s) : s_(s) {}
  ~Profiler() {
    flush();
    for (hipEvent_t e : pool_) (void)hipEventDestroy(e);
  }
  // time `launch` (which enqueues kernels on the stream) under `name`
  bool suspended = false;  // set while a graph is captured: launches only
  template <class F>
  void run(const char* name, double units, F launch) {
    if (!g_prof_on || suspended) {
      launch();
      return;
    }
    const hipEvent_t a = event(), b = event();
    hipchk(hipEventRecord(a, s_), "event record");
    launch();
    hipchk(hipEventRecord(b, s_), "event record");
    pend_.push_back({entry(name), a, b, units, true, true});
  }
  // two stages launched back to back, timed with three events: `name`
  // covers both, `n1` the first and `n2` the second (an event between
  // every pair of kernels delays the next dispatch; the shared middle event
  // saves three of six)
  template <class F1, class F2>
  void run2(const char* name, double units, const char* n1, double u1, F1 launch1, const char* n2, double u2,
            F2 launch2) {
    if (!g_prof_on || suspended) {
      launch1();
      launch2();
      return;
    }
    const hipEvent_t a = event(), m = event(), b = event();
    hipchk(hipEventRecord(a, s_), "event record");
    launch1();
    hipchk(hipEventRecord(m, s_), "event record");
    launch2();
    hipchk(hipEventRecord(b, s_), "event record");
    pend_.push_back({entry(name), a, b, units, true, true});
    pend_.push_back({entry(n1), a, m, u1, false, true});
    pend_.push_back({entry(n2), m, b, u2, false, false});
  }
  // read the events of everything already synchronised
  void flush() {
    if (pend_.empty()) return;
    (void)hipStreamSynchronize(s_);
    std::lock_guard<std::mutex> g(g_prof_mu);
    for (auto& p : pend_) {
      float ms = 0;
      if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess && p.idx < (int)g_prof.size()) {
        g_prof[p.idx].launches++;
        g_prof[p.idx].ms += ms;
        g_prof[p.idx].units += p.units;
      }
      if (p.own_a) pool_.push_back(p.a);
      if (p.own_b) pool_.push_back(p.b);
    }
    pend_.clear();
  }

 private:
  int entry(const char* name) {
    std::lock_guard<std::mutex> g(g_prof_mu);
    for (size_t k = 0; k < g_prof.size(); ++k)
      if (g_prof[k].name == name) return (int)k;
    g_prof.push_back(ProfEntry{name});
    return (int)g_prof.size() - 1;
  }
  hipStream_t s_;
  std::vector<ProfPending> pend_;
  std::vector<hipEvent_t> pool_;  // events of flushed launches, reused
  hipEvent_t event() {
    if (!pool_.empty()) {
      const hipEvent_t e = pool_.back();
      pool_.pop_back();
      return e;
    }
    hipEvent_t e;
    hipchk(hipEventCreate(&e), "event");
    return e;
  }
};

// wall-clock time of a host phase, recorded next to the kernel timings
// host-side entry: wall time `ms` and `units` of work (a pure count when
// ms is 0, such as the search's repeated lambdas)
void prof_host(const char* name, double ms, double units = 0) {
  std::lock_guard<std::mutex> g(g_prof_mu);
  if (!g_prof_on) return;
  for (auto& e : g_prof)
    if (e.name == name) {
      e.launches++;
      e.ms += ms;
      e.units += units;
      return;
    }
  g_prof.push_back(ProfEntry{name, 1, ms, units});
}

struct HostTimer {
  const char* name;
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  explicit HostTimer(const char* n) : name(n) {}
  ~HostTimer() {
    prof_host(name, std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
  }
};

// Device allocations are cached per device across calls: a released
// buffer goes back to a free list and a later request of at least its size
// (and at most 5/4 of it) takes it again, so repeated calls of one shape
// allocate and free nothing.  pgsvt_release_cache() frees the list; a
// failing hipMalloc empties it and retries once.
class DeviceCache {
 public:
  static DeviceCache& get() {
    static DeviceCache* c = new DeviceCache;  // never destroyed: outlives every caller
    return *c;
  }
  void* take(size_t bytes) {
    if (fault_hit(1)) throw HipError{hipErrorOutOfMemory, "hipMalloc (injected failure)"};
    int dev = 0;
    hipchk(hipGetDevice(&dev), "hipGetDevice");
    {
      std::lock_guard<std::mutex> g(mu_);
      auto& fl = free_[dev];
      auto it = fl.lower_bound(bytes);
      if (it != fl.end() && it->first <= bytes + bytes / 4) {
        void* p = it->second;
        held_[dev] -= it->first;
        size_of_[p] = it->first;
        fl.erase(it);
        return p;
      }
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      release(dev);
      hipchk(hipMalloc(&p, bytes), "hipMalloc");
    }
    std::lock_guard<std::mutex> g(mu_);
    size_of_[p] = bytes;
    return p;
  }
  void give(void* p) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return;
    std::lock_guard<std::mutex> g(mu_);
    auto it = size_of_.find(p);
    if (it == size_of_.end()) return;
    free_[dev].emplace(it->second, p);
    held_[dev] += it->second;
    size_of_.erase(it);
  }
  // free every cached block of device `dev`
  void release(int dev) {
    std::multimap<size_t, void*> fl;
    {
      std::lock_guard<std::mutex> g(mu_);
      fl.swap(free_[dev]);
      held_[dev] = 0;
    }
    for (auto& kv : fl) (void)hipFree(kv.second);
  }
  size_t held(int dev) {
    std::lock_guard<std::mutex> g(mu_);
    return held_[dev];
  }

 private:
  std::mutex mu_;
  std::map<int, std::multimap<size_t, void*>> free_;
  std::map<int, size_t> held_;
  std::map<void*, size_t> size_of_;
};

// device buffer owning its (cached) allocation
template <class T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  DBuf() {}
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() { release(); }
  void alloc(size_t count) {
    release();
    if (count == 0) return;
    p = (T*)DeviceCache::get().take(count * sizeof(T));
    n = count;
  }
  // Normal paths release a buffer only after the stream that used it was
  // synchronised; while an exception unwinds, kernels may still be queued,
  // so the device is drained before the block becomes reusable.
  void release() {
    if (p) {
      if (std::uncaught_exceptions() > 0) (void)hipDeviceSynchronize();
      DeviceCache::get().give(p);
    }
    p = nullptr;
    n = 0;
  }
  size_t bytes() const { return n * sizeof(T); }
};

// Pinned host staging for the small per-round copies (window metadata and
// list up, risk sums down), which then need no staging copy by the runtime.
// Pinned host blocks, kept for reuse across calls like the device blocks:
// allocating pinned memory costs far more than the copies it serves (the
// per-round staging of a call, the host copy of a device-resident input).
// A block of at least the asked size and at most twice it is reused; take()
// reports the block's size through `bytes`.
class PinnedCache {
 public:
  static PinnedCache& get() {
    static PinnedCache* c = new PinnedCache;  // never destroyed: outlives every caller
    return *c;
  }
  void* take(size_t& bytes) {
    {
      std::lock_guard<std::mutex> g(mu_);
      auto it = free_.lower_bound(bytes);
      if (it != free_.end() && it->first <= 2 * bytes) {
        void* p = it->second;
        bytes = it->first;
        free_.erase(it);
        return p;
      }
    }
    void* p = nullptr;
    hipchk(hipHostMalloc(&p, bytes, hipHostMallocDefault), "hipHostMalloc");
    return p;
  }
  void give(void* p, size_t bytes) {
    std::lock_guard<std::mutex> g(mu_);
    free_.emplace(bytes, p);
  }
  void release() {
    std::multimap<size_t, void*> fl;
    {
      std::lock_guard<std::mutex> g(mu_);
      fl.swap(free_);
    }
    for (auto& kv : fl) (void)hipHostFree(kv.second);
  }

 private:
  std::mutex mu_;
  std::multimap<size_t, void*> free_;
};

template <class T>
struct HBuf {
  T* p = nullptr;
  size_t n = 0;
  HBuf() {}
  HBuf(const HBuf&) = delete;
  HBuf& operator=(const HBuf&) = delete;
  ~HBuf() {
    if (p) PinnedCache::get().give(p, n * sizeof(T));
  }
  void alloc(size_t count) {
    if (count <= n) return;
    if (p) PinnedCache::get().give(p, n * sizeof(T));
    p = nullptr;
    n = 0;
    size_t bytes = count * sizeof(T);
    p = (T*)PinnedCache::get().take(bytes);
    n = bytes / sizeof(T);
  }
};

// one lock per device for the batch sizing of concurrent calls
std::mutex& batch_lock(int dev) {
  static std::mutex mu;
  static std::map<int, std::unique_ptr<std::mutex>>* locks = new std::map<int, std::unique_ptr<std::mutex>>;
  std::lock_guard<std::mutex> g(mu);
  auto& m = (*locks)[dev];
  if (!m) m.reset(new std::mutex);
  return *m;
}

struct Geometry {
  uint32_t Nt, fw, Nw;
};

Geometry geometry(const pgsvt_params& p) {
  Geometry g;
  g.Nt = (p.patch_size * p.patch_size < p.trajectory_length) ? p.patch_size * p.patch_size - 1 : p.trajectory_length;
  g.fw = (uint32_t)std::floor(g.Nt / 2);
  g.Nw = 2 * g.fw + 1;
  return g;
}

double elem(const void* X, int dtype, size_t k) {
  switch (dtype) {
    case PGSVT_U8: return (double)((const uint8_t*)X)[k];
    case PGSVT_U16: return (double)((const uint16_t*)X)[k];
    case PGSVT_F32: return (double)((const float*)X)[k];
    default: return ((const double*)X)[k];
  }
}

size_t dtype_size(int dtype) {
  switch (dtype) {
    case PGSVT_U8: return 1;
    case PGSVT_U16: return 2;
    case PGSVT_F32: return 4;
    default: return 8;
  }
}

// patch index set (src/svt.hpp:61-92), square frames
std::vector<uint32_t> patch_indices(uint32_t N, uint32_t bs, uint32_t ov) {
  const uint32_t n = N - bs, nd = n / ov;
  std::vector<uint32_t> j;
  j.reserve((size_t)(nd + 1) * (nd + 1) + 2 * (nd + 1));
  for (uint32_t i = 0; i < 1 + n; i += ov)
    for (uint32_t k = 0; k < 1 + n; k += ov) j.push_back(i * n + k);
  for (uint32_t i = 0; i < 1 + n; i += ov) j.push_back((n + 1) * n + i);  // right edge
  for (uint32_t i = 0; i < 1 + n; i += ov) j.push_back((n + 1) * i + n);  // bottom edge
  std::sort(j.begin(), j.end());
  j.erase(std::unique(j.begin(), j.end()), j.end());
  return j;
}

// Window placement (src/pguresvt.hpp:100-114, 155-166).
// Window of output frame t: first input frame s0 (src/pguresvt.hpp:100-
// 114), output frame ref (:155-166, with the driver's Nt) and ARPS reference
// frame aref (src/arps.hpp:55-118, with the window's frame count Nw).  Nt
// and Nw differ when the clamped trajectory length Nt is even (Nw = Nt + 1):
// then, near the end of the sequence, the output frame is the one after the
// ARPS reference frame.
void window_of(uint32_t t, uint32_t nf, const Geometry& g, uint32_t& s0, uint32_t& ref, uint32_t& aref) {
  if (t < g.fw) {
    s0 = 0;
    ref = aref = t;
  } else if (t >= nf - g.fw) {
    s0 = nf - 2 * g.fw - 1;
    ref = t - (nf - g.Nt);
    aref = t - (nf - g.Nw);
  } else {
    s0 = t - g.fw;
    ref = aref = g.fw;
  }
}

// ARPS schedule of one window (src/arps.hpp:52-134) as jobs of at most two
// passes; the second pass of a job reads the motion slot the first wrote.
struct Pass {
  int src, dst, slot;
};

bool arps_jobs(uint32_t t, uint32_t nf, const Geometry& g, std::vector<int16_t>& jobs) {
  // the motion estimator's Nt is its window's frame count (A.n_slices,
  // src/arps.hpp:35), not the driver's clamped trajectory length
  const uint32_t Nt = g.Nw, tw = g.fw;
  std::vector<Pass> ps;
  auto add = [&](int, int a, int b, int c) { ps.push_back({a, b, c}); };
  if (t < tw) {
    const uint32_t loopEnd = Nt - t - 1;
    for (uint32_t i = 0; i < loopEnd; ++i) add(i, t + i, t + i + 1, t + i);
    for (uint32_t i = 0; i < t; ++i) {
      const int n = -1 * (int)(i + 1);
      add(n, t + n + 1, t + n, t + n + 1);
    }
  } else if (t >= nf - tw) {
    const uint32_t endFrame = t - (nf - Nt);
    const uint32_t loopEnd = 2 * tw - endFrame;
    for (uint32_t i = 0; i < loopEnd; ++i) add(i, endFrame + i, endFrame + i + 1, endFrame + i);
    for (uint32_t i = 0; i < endFrame; ++i) {
      const int n = -1 * (int)(i + 1);
      if (2 * tw == endFrame)
        add(n, endFrame + n + 1, endFrame + n, endFrame + n);
      else
        add(n, endFrame + n + 1, endFrame + n, endFrame + n + 1);
    }
  } else {
    for (uint32_t i = 0; i < tw; ++i) add(i, tw + i, tw + i + 1, tw + i);
    for (uint32_t i = 0; i < tw; ++i) {
      const int n = -1 * (int)(i + 1);
      add(n, tw + n + 1, tw + n, tw + n + 1);
    }
  }
  // dependency: latest earlier pass writing the same slot
  std::vector<int> dep(ps.size(), -1), dependents(ps.size(), 0);
  for (size_t k = 0; k < ps.size(); ++k)
    for (size_t m = 0; m < k; ++m)
      if (ps[m].slot == ps[k].slot) dep[k] = (int)m;
  for (size_t k = 0; k < ps.size(); ++k)
    if (dep[k] >= 0) {
      if (dep[dep[k]] >= 0) return false;  // chains longer than two
      if (++dependents[dep[k]] > 1) return false;
    }
  jobs.clear();
  for (size_t k = 0; k < ps.size(); ++k) {
    if (dep[k] >= 0) continue;
    int16_t sb = -1, db = -1;
    for (size_t m = k + 1; m < ps.size(); ++m)
      if (dep[m] == (int)k) {
        sb = (int16_t)ps[m].src;
        db = (int16_t)ps[m].dst;
      }
    jobs.push_back((int16_t)ps[k].src);
    jobs.push_back((int16_t)ps[k].dst);
    jobs.push_back(sb);
    jobs.push_back(db);
  }
  return true;
}

int check_device() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return fail(PGSVT_E_NOGPU, "no HIP device");
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return fail(PGSVT_E_NOGPU, "no HIP device properties");
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(PGSVT_E_NOGPU, std::string("this library is built for gfx950, device is ") + prop.gcnArchName);
  return 0;
}

int validate(const void* X, int dtype, uint32_t nr, uint32_t nc, uint32_t nf, const pgsvt_params* p) {
  if (!X || !p) return fail(PGSVT_E_ARG, "null pointer");
  if (dtype < 0 || dtype > 3) return fail(PGSVT_E_ARG, "invalid dtype");
  if (nr != nc) return fail(PGSVT_E_ARG, "frames must be square: the reference patch tables assume n_rows == n_cols");
  if (nr > 65535) return fail(PGSVT_E_ARG, "frame side above 65535");
  if (p->patch_size < 2 || (p->patch_size > 8 && p->patch_size != 16) || p->patch_size > nr)
    return fail(PGSVT_E_ARG, "patch_size must be 2 to 8 or 16 and fit the frame");
  if (p->patch_overlap < 1 || p->patch_overlap > p->patch_size) return fail(PGSVT_E_ARG, "invalid patch_overlap");
  if (p->trajectory_length < 1 || p->trajectory_length % 2 == 0)
    return fail(PGSVT_E_ARG, "trajectory_length must be a positive odd integer");
  const Geometry g = geometry(*p);
  if (g.Nt < 1 || nf < g.Nw) return fail(PGSVT_E_ARG, "fewer frames than the temporal window");
  if (p->motion_estimation && (p->motion_window < 3 || p->motion_window % 2 == 0 || p->motion_window > 31))
    return fail(PGSVT_E_ARG, "motion_window must be odd, >= 3 and <= 31");
  if (!svd_shape_supported(p->patch_size, g.Nw, g_precision.load() == 64))
    return fail(PGSVT_E_ARG, "unsupported shape: patch " + std::to_string(p->patch_size) + " with a window of " +
                                 std::to_string(g.Nw) + " frames (patches 2-8 take windows up to 31 frames, patch 16 "
                                 "up to 15; the FP64 path only the instantiated table)");
  if (p->optimize_pgure && p->noise_method < 1) return fail(PGSVT_E_ARG, "invalid noise_method");
  return 0;
}

uint64_t seed_of(const pgsvt_params& p) {
  if (p.random_seed >= 0) return (uint64_t)p.random_seed;
  std::random_device rd;
  return ((uint64_t)rd() << 32) ^ rd();
}

// ------------------------------------------------------- host job pool --
// Noise estimation and lambda start points of every window, computed on
// host threads while the GPU works.  The work is one ordered job list: the
// leaf statistics of each distinct (frame, normalisation) a window needs,
// then the window's own fit, window after window.  Threads take jobs in
// list order, so a window's fit only waits for leaf jobs taken before it,
// and windows become ready roughly in order, as the search admits them.
struct WindowPrep {
  double alpha = -1, mu = -1, sigma = -1, start = 0;
  std::string err;       // non-empty: the estimate failed
  bool numeric = false;  // ... because the estimate is not defined (NoiseError)
  // rethrow a failure on the calling thread
  void check() const {
    if (err.empty()) return;
    if (numeric) throw NoiseError(err);
    throw std::runtime_error(err);
  }
};

class PrepPool {
 public:
  PrepPool(std::vector<std::function<void(PrepPool&)>> jobs, uint32_t windows, uint32_t threads)
      : jobs_(std::move(jobs)), res_(windows), ready_(windows, 0) {
    const uint32_t n = (uint32_t)jobs_.size();
    for (uint32_t k = 0; k < std::max(1u, std::min(threads, n)); ++k)
      th_.emplace_back([this, n] {
        for (;;) {
          const uint32_t j = next_.fetch_add(1);
          if (j >= n) return;
          jobs_[j](*this);
        }
      });
  }
  ~PrepPool() {
    for (auto& t : th_) t.join();
  }
  void publish(uint32_t t, WindowPrep r) {
    {
      std::lock_guard<std::mutex> g(mu_);
      res_[t] = std::move(r);
      ready_[t] = 1;
    }
    cv_.notify_all();
  }
  WindowPrep wait(uint32_t t) {
    std::unique_lock<std::mutex> g(mu_);
    cv_.wait(g, [&] { return ready_[t] != 0; });
    return res_[t];
  }
  bool ready(uint32_t t) {
    std::lock_guard<std::mutex> g(mu_);
    return ready_[t] != 0;
  }

 private:
  std::vector<std::function<void(PrepPool&)>> jobs_;
  std::vector<WindowPrep> res_;
  std::vector<char> ready_;
  std::vector<std::thread> th_;
  std::atomic<uint32_t> next_{0};
  std::mutex mu_;
  std::condition_variable cv_;
};

// ------------------------------------------------------------ pipeline --
struct Pipeline {
  const void* X = nullptr;
  int dtype = 0;
  uint32_t nr = 0, nc = 0, nf = 0;
  pgsvt_params p{};
  Geometry g{};
  Shape sh{};
  bool optimize = false, median = false;
  bool host_noise = false;  // the noise estimator runs on the host (search or learned lambda)
  uint32_t tb = 0;          // first output window of the call (output slot = t - tb)
  uint32_t nvar = 1;
  double eps1 = 1e-4, eps2 = 1e-2, d2neg = 0, d2pos = 0;
  uint64_t seed = 0;
  std::vector<uint32_t> sel;
  std::vector<double> fmaxX, fmaxZ;
  hipStream_t s = nullptr;

  DBuf<char> raw;
  DBuf<double> Xd, Zd, Yd;
  double* Yout = nullptr;           // final frames (Yd.p or a caller device buffer)
  HBuf<char> host_copy;             // host copy of a device-resident input (pinned)
  // the host copy is made on a stream of its own, beside the GPU work; the
  // host readers wait for `copied` (host_ready)
  hipStream_t cs = nullptr;
  hipEvent_t copied = nullptr;
  std::unique_ptr<Profiler> prof;
  DBuf<uint8_t> d2;
  DBuf<uint32_t> sel_d;
  // batch buffers
  uint32_t NB = 0;
  DBuf<WinMeta> meta_d;
  DBuf<uint32_t> wl_d;
  HBuf<WinMeta> meta_h;
  HBuf<uint32_t> wl_h;
  HBuf<double> sums_h, kept_h;
  DBuf<double> partial_k, kept_d;  // kept ranks of each eval block / window (profiling)
  DBuf<double> u64, w64, invw, partial, partial_b, sums;
  DBuf<char> dw;  // delta2 / W per voxel, in the reconstruction's type
  DBuf<char> divq;  // per (patch, variant) singular values and divergence factors (PGURE)
  // SVD input variants, factors and U reconstructions: FP32, or FP64 for
  // the precision study (sh.fp64); sized in bytes
  DBuf<char> uvar, factors, B0;  // B0: reconstructed rows (patches above 4 x 4) or the factors d_r (dr)
  DBuf<int32_t> pos, cnt, off, list, inv;
  // the fused evaluation's static data per window (patches up to 4 x 4):
  // rows w_0 v_0[k] and singular pairs (s_0, s_1) per (frame, bucket place)
  DBuf<char> w0b, s01b;
  uint32_t static_frames = 0;  // frame subset of the static data (0: every frame; see launch_static_rows)
  DBuf<int16_t> jobs_d;
  uint32_t max_jobs = 0, gather_blocks = 0;
  // learned-lambda probe: the patch subsample (sh_s: p_sel = its size) with
  // its three-variant factors and divergence factors, per-slot lambdas
  bool probe_on = false;
  uint32_t probe_stride = 1;
  Shape sh_s;
  DBuf<uint32_t> sel_s, wl_s;
  DBuf<char> factors_s, divq_s;
  DBuf<double> partial_s, lam_s;
  DBuf<uint32_t> stats;

  ~Pipeline() {
    // drain the streams before the member buffers go back to the caches
    if (s) (void)hipStreamSynchronize(s);
    if (cs) (void)hipStreamSynchronize(cs);
    prof.reset();
    if (copied) (void)hipEventDestroy(copied);
    if (cs) (void)hipStreamDestroy(cs);
    if (s) (void)hipStreamDestroy(s);
  }

  // the host copy of a device-resident input is complete (thread-safe)
  void host_ready() const {
    if (copied) hipchk(hipEventSynchronize(copied), "host copy of X");
  }

  // X_ is a host pointer, or a device pointer when on_device is set; the
  // host-side noise estimator then works on a host copy
  void setup(const void* X_, int dtype_, uint32_t nr_, uint32_t nc_, uint32_t nf_, const pgsvt_params& p_,
             bool on_device = false) {
    X = X_;
    dtype = dtype_;
    nr = nr_;
    nc = nc_;
    nf = nf_;
    p = p_;
    g = geometry(p);
    optimize = p.optimize_pgure != 0;
    median = p.motion_filter > 0;
    nvar = optimize ? 3 : 1;
    sh.nr = nr;
    sh.nc = nc;
    sh.fs = nr * nc;
    sh.nw = g.Nw;
    sh.vol = g.Nw * sh.fs;
    sh.bs = p.patch_size;
    sh.m = sh.bs * sh.bs;
    sh.n1 = nr - sh.bs + 1;
    sh.p_all = sh.n1 * sh.n1;
    sel = patch_indices(nr, sh.bs, p.patch_overlap);
    sh.p_sel = (uint32_t)sel.size();
    sh.fstride = factor_stride(g.Nw, sh.m);
    sh.fp64 = g_precision.load() == 64;
    const double vP = 0.5 + 0.5 * 1. / std::sqrt(1. * 1. + 4);
    const double vQ = 1 - vP;
    d2neg = -1 * std::sqrt(vQ / vP);
    d2pos = std::sqrt(vP / vQ);
    seed = seed_of(p);

    hipchk(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
    prof.reset(new Profiler(s));
    const size_t N = (size_t)sh.fs * nf;
    const void* src = X;
    if (on_device) {
      if (optimize || host_noise) {  // the noise estimator runs on the host
        host_copy.alloc(N * dtype_size(dtype));
        hipchk(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking), "hipStreamCreate");
        hipchk(hipEventCreateWithFlags(&copied, hipEventDisableTiming), "hipEventCreate");
        hipchk(hipMemcpyAsync(host_copy.p, X, N * dtype_size(dtype), hipMemcpyDeviceToHost, cs), "D2H X");
        hipchk(hipEventRecord(copied, cs), "event record");
        X = host_copy.p;
      }
    } else {
      raw.alloc(N * dtype_size(dtype));
      hipchk(hipMemcpyAsync(raw.p, X, raw.bytes(), hipMemcpyHostToDevice, s), "H2D X");
      src = raw.p;
    }
    Xd.alloc(N);
    hipchk(launch_to_double(src, dtype, Xd.p, N, s), "to_double");
    DBuf<double> fm;
    fm.alloc(nf);
    fmaxX.resize(nf);
    hipchk(launch_frame_max(Xd.p, sh.fs, nf, fm.p, s), "frame_max");
    hipchk(hipMemcpyAsync(fmaxX.data(), fm.p, nf * sizeof(double), hipMemcpyDeviceToHost, s), "D2H max");
    if (median) {
      Zd.alloc(N);
      prof->run("median_filter", (double)N, [&] {
        hipchk(launch_median(Xd.p, Zd.p, nr, nc, nf, (int)p.motion_filter, s), "median");
      });
      fmaxZ.resize(nf);
      hipchk(launch_frame_max(Zd.p, sh.fs, nf, fm.p, s), "frame_max Z");
      hipchk(hipMemcpyAsync(fmaxZ.data(), fm.p, nf * sizeof(double), hipMemcpyDeviceToHost, s), "D2H max");
    }
    if (optimize) {
      d2.alloc(sh.vol);
      hipchk(launch_delta2((int64_t)seed, sh.vol, d2.p, s), "delta2");
    }
    sel_d.alloc(sel.size());
    hipchk(hipMemcpyAsync(sel_d.p, sel.data(), sel.size() * sizeof(uint32_t), hipMemcpyHostToDevice, s), "H2D sel");
    hipchk(hipStreamSynchronize(s), "setup sync");
    raw.release();
  }

  WinMeta meta_of(uint32_t t) const {
    WinMeta m{};
    m.t = t;
    m.out = t - tb;
    window_of(t, nf, g, m.s0, m.ref, m.aref);
    double um = fmaxX[m.s0], wm = median ? fmaxZ[m.s0] : 0;
    for (uint32_t k = 1; k < g.Nw; ++k) {
      um = std::max(um, fmaxX[m.s0 + k]);
      if (median) wm = std::max(wm, fmaxZ[m.s0 + k]);
    }
    m.uMax = um;
    m.wMax = median ? wm : um;
    m.lambda = 0;
    return m;
  }

  size_t bytes_per_window() const {
    const size_t es = sh.fp64 ? 8 : 4;
    size_t b = 0;
    b += 8ull * sh.vol;                                  // u64
    if (median) b += 8ull * sh.vol;                      // w64
    b += es * sh.vol * nvar;                             // SVD input variants
    b += 4ull * sh.nw * sh.p_all;                        // pos
    b += 8ull * sh.nw * ((size_t)sh.n1 * sh.n1 + 1);     // bucket counts + offsets
    b += 8ull * sh.nw * sh.p_sel;                        // bucket lists + inverse
    b += 8ull * sh.vol * (optimize ? 2 : 1);             // 1/W, delta2/W
    b += es * sh.nw * (size_t)sh.p_sel * (fused_eval(sh) ? 1 : sh.m);  // B0 (or dr)
    if (fused_eval(sh)) {
      b += singular_pair_bytes(sh);                          // static singular pairs
      b += es * sh.nw * (size_t)sh.p_sel * sh.m;             // static rows
    }
    b += es * nvar * (size_t)sh.p_sel * sh.fstride;      // factors
    if (optimize) b += es * nvar * (size_t)sh.p_sel * 2 * ((sh.nw + 3) & ~3u);  // divq
    if (probe_on) b += es * 3 * (size_t)sh_s.p_sel * (sh.fstride + 2 * ((sh.nw + 3) & ~3u));  // probe factors
    return b;
  }

  void alloc_batch(uint32_t want) {
    int dev = 0;
    hipchk(hipGetDevice(&dev), "hipGetDevice");
    // Calls of this process size and allocate their batches one at a time
    // per device, so that a call sizing its batch sees the memory the calls
    // before it took (pgsvt_denoise_batch runs several calls at once).
    std::lock_guard<std::mutex> reserve(batch_lock(dev));
    size_t freeb = 0, total = 0;
    hipchk(hipMemGetInfo(&freeb, &total), "hipMemGetInfo");
    freeb += DeviceCache::get().held(dev);  // cached blocks are ours to reuse
    const size_t budget = std::min<size_t>((size_t)(freeb * 0.6), 64ull << 30);
    const size_t per = bytes_per_window();
    NB = (uint32_t)std::max<size_t>(1, std::min<size_t>(want, budget / per));
    if (const char* e = std::getenv("PGSVT_BATCH")) NB = std::max(1u, std::min<uint32_t>(NB, (uint32_t)atoi(e)));
    // Other processes can still take memory between the estimate and the
    // allocation: a batch that does not fit is halved (buffers of the failed
    // attempt go back to the cache, which a failing hipMalloc empties).
    for (;;) {
      try {
        alloc_batch_buffers();
        return;
      } catch (const HipError& e) {
        if (e.e != hipErrorOutOfMemory || NB == 1) throw;
        (void)hipGetLastError();
        NB = std::max(1u, NB / 2);
      }
    }
  }

  void alloc_batch_buffers() {
    meta_d.alloc(NB);
    wl_d.alloc(NB);
    meta_h.alloc(NB);
    wl_h.alloc(NB);
    sums_h.alloc((size_t)NB * 6);
    kept_h.alloc(NB);
    u64.alloc((size_t)NB * sh.vol);
    if (median) w64.alloc((size_t)NB * sh.vol);
    const size_t es = sh.fp64 ? 8 : 4;
    uvar.alloc((size_t)NB * sh.vol * nvar * es);
    pos.alloc((size_t)NB * sh.nw * sh.p_all);
    const size_t npos = (size_t)sh.n1 * sh.n1;
    cnt.alloc((size_t)NB * sh.nw * npos);
    off.alloc((size_t)NB * sh.nw * (npos + 1));
    list.alloc((size_t)NB * sh.nw * sh.p_sel);
    inv.alloc((size_t)NB * sh.nw * sh.p_sel);
    invw.alloc((size_t)NB * sh.vol);
    if (optimize) dw.alloc((size_t)NB * sh.vol * (sh.fp64 ? sizeof(double) : sizeof(float)));
    B0.alloc((size_t)NB * sh.nw * sh.p_sel * (fused_eval(sh) ? 1 : sh.m) * es);
    if (fused_eval(sh)) {
      w0b.alloc((size_t)NB * sh.nw * sh.p_sel * sh.m * es);
      s01b.alloc((size_t)NB * singular_pair_bytes(sh));
    }
    factors.alloc((size_t)NB * nvar * sh.p_sel * sh.fstride * es);
    if (optimize) divq.alloc((size_t)NB * nvar * sh.p_sel * 2 * ((sh.nw + 3) & ~3u) * es);
    gather_blocks = pgsvt::gather_blocks(sh, 0);
    partial.alloc((size_t)NB * eval_blocks(sh, 3));
    partial_b.alloc((size_t)NB * gather_blocks * 5);
    sums.alloc((size_t)NB * 6);
    partial_k.alloc((size_t)NB * eval_blocks(sh, 3));
    kept_d.alloc(NB);
    max_jobs = 2 * g.Nw;
    jobs_d.alloc((size_t)NB * max_jobs * 4);
    if (probe_on) {
      factors_s.alloc((size_t)NB * 3 * sh_s.p_sel * sh.fstride * es);
      divq_s.alloc((size_t)NB * 3 * sh_s.p_sel * 2 * ((sh.nw + 3) & ~3u) * es);
      partial_s.alloc((size_t)NB * eval_blocks(sh_s, 3));
      lam_s.alloc(NB);
      wl_s.alloc(NB);
    }
    stats.alloc(1);
    hipchk(launch_fill32(stats.p, stats.n, 0, s), "clear stats");
  }

  // volumes, trajectories, counts and factors of the windows in `ms`
  // metadata and ARPS schedules of the batch to the device (synchronous)
  void upload_batch(const std::vector<WinMeta>& ms) {
    const uint32_t nb = (uint32_t)ms.size();
    hipchk(hipMemcpyAsync(meta_d.p, ms.data(), nb * sizeof(WinMeta), hipMemcpyHostToDevice, s), "H2D meta");
    if (p.motion_estimation) {
      std::vector<int16_t> all((size_t)nb * max_jobs * 4, (int16_t)-1);
      for (uint32_t b = 0; b < nb; ++b) {
        std::vector<int16_t> j;
        if (!arps_jobs(ms[b].t, nf, g, j)) throw HipError{hipErrorInvalidValue, "unsupported ARPS schedule"};
        std::copy(j.begin(), j.end(), all.begin() + (size_t)b * max_jobs * 4);
      }
      hipchk(hipMemcpyAsync(jobs_d.p, all.data(), all.size() * sizeof(int16_t), hipMemcpyHostToDevice, s),
             "H2D jobs");
    }
    hipchk(hipStreamSynchronize(s), "upload sync");  // the host vectors are local
  }

  // the decomposition of the uploaded batch: kernels only (graph-capturable)
  void decompose_kernels(uint32_t nb, uint32_t nvar_run, int svd_mapping = 0) {
    hipchk(launch_prepare(Xd.p, median ? Zd.p : nullptr, meta_d.p, nb, d2.p, sh, nvar, eps2, d2neg, d2pos, u64.p,
                          median ? w64.p : nullptr, uvar.p, s),
           "prepare");
    hipchk(launch_fill32(pos.p, (size_t)nb * sh.nw * sh.p_all, 0, s), "clear pos");
    hipchk(launch_fill_ref(meta_d.p, nb, sh, pos.p, s), "fill_ref");
    if (p.motion_estimation) {
      // searches: one per (window pass, block)
      prof->run("arps", (double)nb * (g.Nw - 1) * sh.p_all, [&] {
        hipchk(launch_arps(median ? w64.p : u64.p, jobs_d.p, max_jobs, nb, sh, (int)p.motion_window, pos.p, s),
               "arps");
      });
    }
    prof->run("buckets_weights", (double)nb * sh.nw * sh.p_sel, [&] {
      hipchk(launch_buckets(pos.p, sel_d.p, nb, sh, cnt.p, off.p, list.p, inv.p, s), "buckets");
      hipchk(launch_weights(off.p, d2.p, nb, sh, d2neg, d2pos, invw.p, optimize ? dw.p : nullptr, s), "weights");
    });
    // units: matrices decomposed
    prof->run("svd_jacobi", (double)nb * nvar_run * sh.p_sel, [&] {
      // rotation threshold |c| > tol sqrt(ab): 1e-6 in FP32, 1e-12 in FP64;
      hipchk(launch_svd(uvar.p, pos.p, sel_d.p, nb, nvar_run, nvar, sh, factors.p, 30, sh.fp64 ? 1e-12 : 1e-6,
                        stats.p, s, svd_mapping),
             "svd");
    });
    if (fused_eval(sh)) {
      // units: bucket places (patch x frame) of the static rows and pairs
      const uint32_t nf = (static_frames && static_frames < sh.nw) ? static_frames + 1 : sh.nw;
      prof->run("static_rows", (double)nb * nf * sh.p_sel, [&] {
        hipchk(launch_static_rows(factors.p, inv.p, meta_d.p, nb, nvar, sh, w0b.p, s01b.p, s, static_frames),
               "static rows");
      });
    }
    if (optimize && nvar_run == 3) {
      // units: matrices
      prof->run("divergence_factors", (double)nb * nvar_run * sh.p_sel, [&] {
        hipchk(launch_divq(factors.p, pos.p, sel_d.p, dw.p, nb, nvar_run, sh, divq.p, s), "divergence factors");
      });
    }
  }

  // what the gather reads (see GatherSrc)
  GatherSrc gather_src() const {
    GatherSrc g;
    g.B0 = B0.p;
    g.factors = factors.p;
    g.dr = B0.p;
    g.S01 = s01b.p;
    g.W0 = w0b.p;
    g.list = list.p;
    g.nvar_stored = nvar;
    g.exp_w = p.exponential_weighting;
    return g;
  }

  // one lambda evaluation of the slots listed in wl (device), their lambdas
  // from meta or slot_lam
  void evaluate(const uint32_t* wl, uint32_t na, uint32_t nvar_run, const double* slot_lam, double* part,
                double* kept) {
    hipchk(launch_evaluation(sh, factors.p, divq.p, meta_d.p, wl, na, nvar, nvar_run, p.exponential_weighting, inv.p,
                             B0.p, part, slot_lam, kept, s),
           "eval");
  }

  // the learned-lambda probe's patch subsample (before alloc_batch)
  void setup_probe() {
    std::vector<uint32_t> sub;
    probe_stride = std::max(1u, std::min(kProbeMaxStride, sh.p_sel / kProbeMinSample));
    for (uint32_t q = 0; q < sh.p_sel; q += probe_stride) sub.push_back(sel[q]);
    sh_s = sh;
    sh_s.p_sel = (uint32_t)sub.size();
    sel_s.alloc(sub.size());
    hipchk(hipMemcpyAsync(sel_s.p, sub.data(), sub.size() * 4, hipMemcpyHostToDevice, s), "H2D probe sample");
    hipchk(hipStreamSynchronize(s), "probe sample sync");  // sub is local
    probe_on = true;
  }

  // the subsample's three variants and their divergence factors (after
  // decompose_kernels of the same batch): the U variant's factor blocks
  // copied from the main decomposition, the perturbed variants decomposed
  void decompose_probe(uint32_t nb) {
    prof->run("probe_decompose", (double)nb * 2 * sh_s.p_sel, [&] {
      const size_t fw = sh.fstride;  // 32-bit words per factor block (FP32)
      hipchk(launch_copy_rows(factors.p, factors_s.p, sh_s.p_sel, (uint32_t)fw, probe_stride,
                              (size_t)nvar * sh.p_sel * fw, (size_t)3 * sh_s.p_sel * fw, nb, s),
             "probe U factors");
      hipchk(launch_svd(uvar.p + sizeof(float) * sh.vol, pos.p, sel_s.p, nb, 2, 3, sh_s,
                        factors_s.p + sizeof(float) * sh_s.p_sel * fw, 30, 1e-6, stats.p, s,
                        g_svd_mapping.load() == 1 ? 1 : 0),
             "probe svd");
      hipchk(launch_divq(factors_s.p, pos.p, sel_s.p, dw.p, nb, 3, sh_s, divq_s.p, s), "probe divergence factors");
    });
  }

  // PGURE risks of the batch windows `wins` (slots of ms) at three lambdas
  // each, lam [wins][3], into risk [wins][3]: the data sums over kProbeFrames frames of the window
  // (all patches reconstructed) and the divergence sum over the patch
  // subsample, each scaled to the whole window.  The same sample for every
  // lambda of a window, so the three risks differ by the lambda only.
  void probe_risks(const std::vector<WinMeta>& ms, const std::vector<WindowPrep>& prep,
                   const std::vector<uint32_t>& wins, const std::vector<double>& lam, std::vector<double>& risk) {
    const uint32_t nb = (uint32_t)ms.size(), nslot = 3 * (uint32_t)wins.size();
    const uint32_t nk = gather_frames(sh, kProbeFrames);
    const uint32_t gblk = pgsvt::gather_blocks(sh, 0, nk), eblk = eval_blocks(sh_s, 3);
    const double fk = (double)sh.nw / nk, fp = (double)sh.p_sel / sh_s.p_sel;
    hipchk(hipMemcpyAsync(meta_d.p, ms.data(), nb * sizeof(WinMeta), hipMemcpyHostToDevice, s), "H2D meta");
    risk.assign(nslot, 0.0);
    std::vector<uint32_t> wl(NB);
    for (uint32_t c0 = 0; c0 < nslot; c0 += NB) {
      const uint32_t na = std::min(NB, nslot - c0);
      for (uint32_t i = 0; i < na; ++i) wl[i] = wins[(c0 + i) / 3];
      hipchk(hipMemcpyAsync(wl_s.p, wl.data(), na * 4, hipMemcpyHostToDevice, s), "H2D probe list");
      hipchk(hipMemcpyAsync(lam_s.p, lam.data() + c0, na * 8, hipMemcpyHostToDevice, s), "H2D probe lambdas");
      prof->run("probe_eval", (double)na * sh.vol, [&] {
        evaluate(wl_s.p, na, 1, lam_s.p, nullptr, nullptr);
        GatherSrc gs = gather_src();
        gs.slot_lam = lam_s.p;
        hipchk(launch_gather(gs, off.p, invw.p, u64.p, meta_d.p, wl_s.p, na, sh, 0, nullptr, partial_b.p, s, nk),
               "probe gather");
        hipchk(launch_round_prep(factors_s.p, divq_s.p, meta_d.p, wl_s.p, na, 3, sh_s, p.exponential_weighting,
                                 lam_s.p, nullptr, partial_s.p, nullptr, s),
               "probe divergence");
        hipchk(launch_risk_reduce(partial_s.p, eblk, partial_b.p, gblk, na, sums.p, s), "probe reduce");
      });
      hipchk(hipMemcpyAsync(sums_h.p, sums.p, (size_t)na * 6 * 8, hipMemcpyDeviceToHost, s), "D2H probe sums");
      hipchk(hipStreamSynchronize(s), "probe sync");  // wl is reused by the next chunk
      for (uint32_t i = 0; i < na; ++i) {
        double S[6];
        for (int q = 0; q < 6; ++q) S[q] = sums_h.p[(size_t)i * 6 + q] * (q == 4 ? fp : fk);
        risk[c0 + i] = this->risk(S, prep[wins[(c0 + i) / 3]]);
      }
    }
    prof->flush();
  }

  void decompose(std::vector<WinMeta>& ms, uint32_t nvar_run) {
    upload_batch(ms);
    decompose_kernels((uint32_t)ms.size(), nvar_run, (!sh.fp64 && g_svd_mapping.load() == 1) ? 1 : 0);
  }

  // PGURE risk (src/pgure.hpp:136) from the GPU sums; noise as the driver
  // holds it (alpha, mu, sigma) before the swapped constructor call.
  // S = (sum Uhat^2, sum U, sum U^2, sum Uhat U, sum d2 (U2p - 2 Uhat + U2m),
  // sum Uhat); sum (Uhat - U)^2 = sum Uhat^2 - 2 sum Uhat U + sum U^2.
  double risk(const double* S, const WindowPrep& w) const {
    const double alpha = w.alpha, mup = w.sigma, sigp = w.mu;  // constructor receives (alpha, sigma, mu)
    const double sigmasq = sigp * sigp;
    const double OoN = 1.0 / (sh.nr * sh.nc * sh.nw);
    const double S1 = S[0] - 2.0 * S[3] + S[2];
    const double t3 = 0.0;  // delta1 term: U1 == U (integer delta1 cube)
    return OoN * (S1 - (alpha + mup) * S[1] + (2 / eps1 * t3) - (2 * sigmasq * alpha / (eps2 * eps2) * S[4]) +
                  (2 * mup * S[5]) + mup) -
           sigmasq;
  }

  // evaluate the risk of the listed batch slots at their meta.lambda
  void eval_risk(std::vector<WinMeta>& ms, const std::vector<uint32_t>& list_, std::vector<double>& out) {
    const uint32_t na = (uint32_t)list_.size();
    // the staging buffers are free: the previous round was synchronised
    std::memcpy(meta_h.p, ms.data(), ms.size() * sizeof(WinMeta));
    std::memcpy(wl_h.p, list_.data(), na * sizeof(uint32_t));
    hipchk(hipMemcpyAsync(meta_d.p, meta_h.p, ms.size() * sizeof(WinMeta), hipMemcpyHostToDevice, s), "H2D meta");
    hipchk(hipMemcpyAsync(wl_d.p, wl_h.p, na * sizeof(uint32_t), hipMemcpyHostToDevice, s), "H2D list");
    // units: thresholded patch reconstructions (patch x variant)
    // profiling with rank counts: the evaluation also counts the ranks it
    // keeps (the factor rows it reads)
    const bool count = g_prof_count && !prof->suspended;
    // units of the whole evaluation: voxels of the evaluated windows
    // (gather_risk units: voxels)
    prof->run2(
        "search_eval", (double)na * sh.vol, "eval_reconstruct", (double)na * 3 * sh.p_sel,
        [&] { evaluate(wl_d.p, na, 3, nullptr, partial.p, count ? partial_k.p : nullptr); }, "gather_risk",
        (double)na * sh.vol, [&] {
          hipchk(launch_gather(gather_src(), off.p, invw.p, u64.p, meta_d.p, wl_d.p, na, sh, 0, nullptr, partial_b.p,
                               s),
                 "gather");
          hipchk(launch_risk_reduce(partial.p, eval_blocks(sh, 3), partial_b.p, gather_blocks, na, sums.p, s,
                                    count ? partial_k.p : nullptr, count ? kept_d.p : nullptr),
                 "reduce");
        });
    out.resize((size_t)na * 6);
    hipchk(hipMemcpyAsync(sums_h.p, sums.p, (size_t)na * 6 * sizeof(double), hipMemcpyDeviceToHost, s), "D2H sums");
    if (count)
      hipchk(hipMemcpyAsync(kept_h.p, kept_d.p, (size_t)na * sizeof(double), hipMemcpyDeviceToHost, s), "D2H kept");
    hipchk(hipStreamSynchronize(s), "risk sync");
    std::memcpy(out.data(), sums_h.p, (size_t)na * 6 * sizeof(double));
    if (count) {
      double k = 0;
      for (uint32_t i = 0; i < na; ++i) k += kept_h.p[i];
      prof_host("eval_kept_ranks", 0.0, k);
    }
    prof->flush();
  }

  // final frames of the batch into Yout
  void finish(std::vector<WinMeta>& ms) {
    const uint32_t nb = (uint32_t)ms.size();
    std::vector<uint32_t> lst(nb);
    for (uint32_t b = 0; b < nb; ++b) lst[b] = b;
    hipchk(hipMemcpyAsync(meta_d.p, ms.data(), nb * sizeof(WinMeta), hipMemcpyHostToDevice, s), "H2D meta");
    hipchk(hipMemcpyAsync(wl_d.p, lst.data(), nb * sizeof(uint32_t), hipMemcpyHostToDevice, s), "H2D list");
    prof->run("final_frame", (double)nb * sh.p_sel, [&] {
      evaluate(wl_d.p, nb, 1, nullptr, nullptr, nullptr);
      hipchk(launch_gather(gather_src(), off.p, invw.p, u64.p, meta_d.p, wl_d.p, nb, sh, 1, Yout, nullptr, s),
             "gather final");
    });
    hipchk(hipStreamSynchronize(s), "finish sync");
    prof->flush();
  }

  // accu(u) of the normalised window t in Armadillo's order (even / odd
  // partial sums), the lambda start point of src/pguresvt.hpp:139
  double host_window_accu(uint32_t t) const {
    host_ready();
    const WinMeta m = meta_of(t);
    const size_t base = (size_t)m.s0 * sh.fs;
    double a1 = 0, a2 = 0;
    size_t i = 0;
    for (; i + 1 < sh.vol; i += 2) {
      a1 += elem(X, dtype, base + i) / m.uMax;
      a2 += elem(X, dtype, base + i + 1) / m.uMax;
    }
    if (i < sh.vol) a1 += elem(X, dtype, base + i) / m.uMax;
    return a1 + a2;
  }

  // frame f normalised by scale (true division, as u /= uMax)
  void host_frame(uint32_t f, double scale, double* out) const {
    host_ready();
    const size_t base = (size_t)f * sh.fs;
    for (size_t k = 0; k < sh.fs; ++k) out[k] = elem(X, dtype, base + k) / scale;
  }
};

// A14: lockstep SBPLX over the batch slots of ms (slot b: ms[b]).  fill(b)
// makes prep[b] (noise estimates, start point) available, waiting if needed;
// ready(b) says whether it would not wait.  A window's search starts once
// its prep is there; every round evaluates the risk of the windows that
// asked for a lambda together.  On return ms[b].lambda is each window's
// last evaluated lambda (Q4).
// x0 (optional): per window a start point and initial step in place of the
// reference's (prep.start, its square root), with the relative tolerance
// tol_x0, for searches that start near the minimum.
struct SearchStart {
  std::vector<double> x, step;
  double tol = 0;
};

void lockstep_search(Pipeline& P, std::vector<WinMeta>& ms, std::vector<WindowPrep>& prep,
                     const std::function<bool(uint32_t)>& ready, const std::function<void(uint32_t)>& fill,
                     int& status, std::string& warn, const SearchStart* x0 = nullptr) {
  const pgsvt_params& p = P.p;
  const uint32_t nb = (uint32_t)ms.size();
  std::vector<std::unique_ptr<Sbplx1D>> opt(nb);
  std::vector<int> code(nb, Sbplx1D::kNeed);
  std::vector<char> started(nb, 0);
  uint32_t n_started = 0;
  auto admit = [&](uint32_t b) {
    fill(b);
    const double start = prep[b].start;
    const double ub = std::max(100.0, start);
    if (x0)
      opt[b].reset(new Sbplx1D(std::min(x0->x[b], ub), 0.0, ub, x0->step[b], x0->tol, 1e-12, (int)p.max_iter));
    else
      opt[b].reset(new Sbplx1D(start, 0.0, ub, std::sqrt(start), p.tol, 1e-12, (int)p.max_iter));
    started[b] = 1;
    n_started++;
  };
  HostTimer tr("host_search_rounds");
  std::vector<double> sums;
  // Risks already evaluated, per window.  The search asks for the very
  // same lambda again (a reflection clamped at the bound, a contraction
  // onto a vertex): about a quarter of its requests.  The risk of a
  // window is a deterministic function of lambda (fixed-order sums), so
  // a repeat is answered from here with the identical value.
  std::vector<std::vector<std::pair<double, double>>> seen(nb);
  auto recall = [&](uint32_t b, double l, double& r) {
    for (const auto& e : seen[b])
      if (e.first == l) {
        r = e.second;
        return true;
      }
    return false;
  };
  double repeats = 0;
  for (;;) {
    for (uint32_t b = 0; b < nb; ++b)
      if (!started[b] && ready(b)) admit(b);
    std::vector<uint32_t> list;
    for (uint32_t b = 0; b < nb; ++b) {
      if (!started[b] || code[b] != Sbplx1D::kNeed) continue;
      int rc;
      double r;
      while ((rc = opt[b]->run()) == Sbplx1D::kNeed && recall(b, opt[b]->pending(), r)) {
        opt[b]->supply(r);
        repeats += 1;
      }
      if (rc == Sbplx1D::kNeed) {
        ms[b].lambda = opt[b]->pending();
        list.push_back(b);
      } else {
        code[b] = rc;
      }
    }
    if (list.empty()) {
      if (n_started == nb) break;
      HostTimer tw("host_noise_wait");
      for (uint32_t b = 0; b < nb; ++b)
        if (!started[b]) {
          admit(b);
          break;
        }
      continue;
    }
    P.eval_risk(ms, list, sums);
    for (size_t k = 0; k < list.size(); ++k) {
      const uint32_t b = list[k];
      const double r = P.risk(&sums[k * 6], prep[b]);
      seen[b].emplace_back(ms[b].lambda, r);
      opt[b]->supply(r);
    }
  }
  prof_host("search_repeats", 0.0, repeats);
  for (uint32_t b = 0; b < nb; ++b) {
    ms[b].lambda = opt[b]->last_x();
    if (code[b] < 0) {
      status = PGSVT_E_NUMERIC;
      warn = "lambda optimisation failed (NLopt-equivalent error code " + std::to_string(code[b]) + ")";
    } else if (code[b] == Sbplx1D::kMaxEval) {
      std::fprintf(stderr,
                   "WARNING: optimization terminated after max_eval (%u) was reached.\n"
                   "Consider increasing the max_eval or increasing the convergence tolerance (tol=%g).\n",
                   p.max_iter, p.tol);
    }
  }
}

// Probe of three lambdas (ascending) and their risks: the predicted
// relative risk gap of the middle one, and where a search for the minimum
// should start.  The parabola through (log lambda, risk) has its minimum
// slope^2 / (4 curvature) below the middle risk, at its vertex (clamped to
// three probe steps from the middle).  Not convex, or points that coincide
// (lambda at a bound): the best observed improvement, or infinity when a
// neighbour is lower on a concave curve, and the lower neighbour as start.
struct ProbeFit {
  double gap, start;
};

ProbeFit probe_fit(const double* lam, const double* r) {
  const double lo = std::min(r[0], r[2]), scale = std::fabs(r[1]);
  const double best = (r[0] <= r[2]) ? lam[0] : lam[2];
  const double xm = std::log(lam[0]), x0 = std::log(lam[1]), xp = std::log(lam[2]);
  if (!std::isfinite(xm) || !std::isfinite(xp) || !(scale > 0)) return {INFINITY, lam[1]};
  if (!(xp > x0) || !(x0 > xm)) return {lo < r[1] ? (r[1] - lo) / scale : 0.0, lo < r[1] ? best : lam[1]};
  const double d1 = (r[1] - r[0]) / (x0 - xm), d2 = (r[2] - r[1]) / (xp - x0);
  const double a = (d2 - d1) / (xp - xm);  // q(x) = r0 + d1 (x - xm) + a (x - xm)(x - x0)
  const double b = d1 + a * (x0 - xm);     // q'(x0)
  if (!(a > 0)) return {lo < r[1] ? INFINITY : 0.0, lo < r[1] ? best : lam[1]};
  const double h = std::max(x0 - xm, xp - x0);
  const double xs = std::min(std::max(x0 - b / (2 * a), x0 - 3 * h), x0 + 3 * h);
  return {b * b / (4 * a) / scale, std::exp(xs)};
}

// Leaf statistics of the noise estimator (src/noise.hpp:55-100) for frames
// `frames` of the device cube Xd (n x n frames, FP64), normalised by scale,
// on the GPU: the split tests of every block of the complete quadtree whose
// size the F table holds (k_split_stats), the reference's recursion on the
// host with those decisions, then the statistics of each distinct leaf
// (k_leaf_stats), in the tree's order per frame.  Frames of at least 16 x 16
// (smaller ones never split and stay on the host).
std::vector<FrameLeaves> gpu_frame_leaves(const double* Xd, uint32_t n, const std::vector<uint32_t>& frames,
                                          double scale, hipStream_t s) {
  std::vector<uint32_t> cand;
  std::map<std::array<uint32_t, 3>, uint32_t> cand_at;
  std::function<void(uint32_t, uint32_t, uint32_t)> add = [&](uint32_t i, uint32_t j, uint32_t N) {
    if (split_threshold(N) < 0) return;
    cand_at[{i, j, N}] = (uint32_t)(cand.size() / 3);
    cand.insert(cand.end(), {i, j, N});
    const uint32_t h = N / 2;
    add(i, j, h);
    add(i + h, j, h);
    add(i, j + h, h);
    add(i + h, j + h, h);
  };
  add(0, 0, n);
  const uint32_t nf = (uint32_t)frames.size(), nc = (uint32_t)(cand.size() / 3);
  std::vector<double> stat((size_t)nf * nc);
  DBuf<uint32_t> frames_d, cand_d, leaves_d;
  DBuf<double> stat_d, out_d;
  if (nc > 0) {
    frames_d.alloc(nf);
    cand_d.alloc(cand.size());
    stat_d.alloc(stat.size());
    hipchk(hipMemcpyAsync(frames_d.p, frames.data(), nf * 4, hipMemcpyHostToDevice, s), "H2D leaf frames");
    hipchk(hipMemcpyAsync(cand_d.p, cand.data(), cand.size() * 4, hipMemcpyHostToDevice, s), "H2D split blocks");
    hipchk(launch_split_stats(Xd, n, scale, frames_d.p, nf, cand_d.p, nc, stat_d.p, s), "split stats");
    hipchk(hipMemcpyAsync(stat.data(), stat_d.p, stat.size() * 8, hipMemcpyDeviceToHost, s), "D2H split stats");
    hipchk(hipStreamSynchronize(s), "split stats sync");
  }
  // the reference's recursion per frame; each distinct leaf evaluated once
  std::vector<uint32_t> leaves;
  std::vector<std::vector<uint32_t>> order(nf);
  for (uint32_t fi = 0; fi < nf; ++fi) {
    const auto tree = quadtree_leaves(n, [&](uint32_t i, uint32_t j, uint32_t N) {
      const double thr = split_threshold(N);
      return thr >= 0 && stat[(size_t)fi * nc + cand_at.at({i, j, N})] > thr;
    });
    std::map<std::array<uint32_t, 3>, uint32_t> seen;
    for (const auto& nd : tree) {
      auto it = seen.find(nd);
      if (it == seen.end()) {
        it = seen.emplace(nd, (uint32_t)(leaves.size() / 4)).first;
        leaves.insert(leaves.end(), {frames[fi], nd[0], nd[1], nd[2]});
      }
      order[fi].push_back(it->second);
    }
  }
  const uint32_t nl = (uint32_t)(leaves.size() / 4);
  std::vector<double> stats((size_t)nl * 2);
  // leaves above the kernel's register cache: values and Laplacians formed
  // once, up to kLeafScratch elements in all (the rest form them per pass)
  constexpr uint64_t kLeafScratch = (uint64_t)1 << 25;
  std::vector<uint64_t> boff(nl, ~0ull);
  std::vector<uint32_t> big;
  uint64_t nbe = 0;
  for (uint32_t k = 0; k < nl; ++k) {
    const uint64_t N = (uint64_t)leaves[4 * k + 3] * leaves[4 * k + 3];
    if (N > leaf_cache_elements() && nbe + N <= kLeafScratch) {
      boff[k] = nbe;
      big.push_back(k);
      nbe += N;
    }
  }
  DBuf<uint32_t> big_d;
  DBuf<uint64_t> boff_d;
  DBuf<double> vals_d, laps_d;
  leaves_d.alloc(leaves.size());
  out_d.alloc(stats.size());
  boff_d.alloc(nl);
  hipchk(hipMemcpyAsync(leaves_d.p, leaves.data(), leaves.size() * 4, hipMemcpyHostToDevice, s), "H2D leaves");
  hipchk(hipMemcpyAsync(boff_d.p, boff.data(), (size_t)nl * 8, hipMemcpyHostToDevice, s), "H2D leaf offsets");
  if (!big.empty()) {
    big_d.alloc(big.size());
    vals_d.alloc(nbe);
    laps_d.alloc(nbe);
    hipchk(hipMemcpyAsync(big_d.p, big.data(), big.size() * 4, hipMemcpyHostToDevice, s), "H2D large leaves");
  }
  hipchk(launch_leaf_stats(Xd, n, scale, leaves_d.p, nl, big_d.p, (uint32_t)big.size(), boff_d.p, vals_d.p, laps_d.p,
                           out_d.p, s),
         "leaf stats");
  hipchk(hipMemcpyAsync(stats.data(), out_d.p, stats.size() * 8, hipMemcpyDeviceToHost, s), "D2H leaf stats");
  hipchk(hipStreamSynchronize(s), "leaf stats sync");
  std::vector<FrameLeaves> out(nf);
  for (uint32_t fi = 0; fi < nf; ++fi)
    for (uint32_t k : order[fi]) {
      out[fi].mean.push_back(stats[2 * (size_t)k]);
      out[fi].var.push_back(stats[2 * (size_t)k + 1]);
    }
  return out;
}

// Host threads for n_jobs (-1 = all): the CPUs this process may run on,
// capped by a cgroup CPU quota when one is set.
uint32_t thread_count(int64_t n_jobs) {
  if (n_jobs == 0 || n_jobs == 1) return 1;
  if (n_jobs > 0) return (uint32_t)n_jobs;
  uint32_t n = std::max(1u, std::thread::hardware_concurrency());
  cpu_set_t set;
  if (sched_getaffinity(0, sizeof set, &set) == 0) n = std::max(1, CPU_COUNT(&set));
  std::ifstream q("/sys/fs/cgroup/cpu.max");
  std::string quota;
  double period = 0;
  if (q >> quota >> period && quota != "max" && period > 0) {
    const double cpus = std::atof(quota.c_str()) / period;
    if (cpus >= 1) n = std::min(n, (uint32_t)std::ceil(cpus));
  }
  return n;
}

// ------------------------------------------------------- learned lambda --
// The proposal network n_in -> h -> h -> h -> 1 (GELU between layers).
// Parameters, flat FP32 in this order: W1 [n_in][h], b1 [h], W2 [h][h],
// b2 [h], W3 [h][h], b3 [h], W4 [h][1], b4 [1].  Every product runs on the
// BF16 matrix cores with FP32 sums (launch_gemm).
size_t mlp_param_count(uint32_t n_in, uint32_t h) {
  return (size_t)n_in * h + h + 2 * ((size_t)h * h + h) + h + 1;
}

class Mlp {
 public:
  Mlp(uint32_t n_in, uint32_t h, hipStream_t s) : n_in_(n_in), h_(h), s_(s) {
    P_.alloc(mlp_param_count(n_in, h));
    G_.alloc(P_.n);
  }
  size_t n_params() const { return P_.n; }
  float* params() { return P_.p; }
  float* grads() { return G_.p; }

  // forward of B rows of X ([B][n_in], device) into out ([B], device);
  // activations are kept for backward
  void forward(const float* X, uint32_t B, float* out) {
    ensure(B);
    const uint32_t h = h_;
    const float *W1 = P_.p, *b1 = W1 + (size_t)n_in_ * h, *W2 = b1 + h, *b2 = W2 + (size_t)h * h, *W3 = b2 + h,
                *b3 = W3 + (size_t)h * h, *W4 = b3 + h, *b4 = W4 + h;
    hipchk(launch_gemm(X, n_in_, 1, W1, h, 1, a1_.p, h, B, h, n_in_, b1, 1, p1_.p, s_), "mlp l1");
    hipchk(launch_gemm(a1_.p, h, 1, W2, h, 1, a2_.p, h, B, h, h, b2, 1, p2_.p, s_), "mlp l2");
    hipchk(launch_gemm(a2_.p, h, 1, W3, h, 1, a3_.p, h, B, h, h, b3, 1, p3_.p, s_), "mlp l3");
    hipchk(launch_gemm(a3_.p, h, 1, W4, 1, 1, out, 1, B, 1, h, b4, 0, nullptr, s_), "mlp l4");
  }

  // gradient of the mean squared error of the last forward's output y
  // against targets t (device, [B]) into grads(); loss into *loss_dev
  void backward(const float* X, const float* y, const float* t, uint32_t B, float* loss_dev) {
    const uint32_t h = h_;
    const float *W2 = P_.p + (size_t)n_in_ * h + h, *W3 = W2 + (size_t)h * h + h, *W4 = W3 + (size_t)h * h + h;
    float *gW1 = G_.p, *gb1 = gW1 + (size_t)n_in_ * h, *gW2 = gb1 + h, *gb2 = gW2 + (size_t)h * h, *gW3 = gb2 + h,
          *gb3 = gW3 + (size_t)h * h, *gW4 = gb3 + h, *gb4 = gW4 + h;
    hipchk(launch_mse(y, 1, t, (int)B, dy_.p, loss_dev, s_), "mse");
    // layer 4: y = a3 W4 + b4
    hipchk(launch_gemm(a3_.p, 1, h, dy_.p, 1, 1, gW4, 1, h, 1, B, nullptr, 0, nullptr, s_), "dW4");
    hipchk(launch_colsum(dy_.p, B, 1, 1, gb4, s_), "db4");
    hipchk(launch_gemm(dy_.p, 1, 1, W4, 1, 1, da_.p, h, B, h, 1, nullptr, 0, nullptr, s_), "da3");
    // layers 3, 2, 1: a = gelu(pre), pre = x W + b
    struct L {
      const float *x, *W, *pre;
      float *gW, *gb;
      uint32_t k;
    };
    const L layers[3] = {{a2_.p, W3, p3_.p, gW3, gb3, h}, {a1_.p, W2, p2_.p, gW2, gb2, h}, {X, nullptr, p1_.p, gW1, gb1, n_in_}};
    for (int l = 0; l < 3; ++l) {
      const L& c = layers[l];
      hipchk(launch_gelu_bwd(c.pre, da_.p, dp_.p, (int64_t)B * h, s_), "gelu bwd");
      hipchk(launch_gemm(c.x, 1, c.k, dp_.p, h, 1, c.gW, h, c.k, h, B, nullptr, 0, nullptr, s_), "dW");
      hipchk(launch_colsum(dp_.p, B, h, h, c.gb, s_), "db");
      if (c.W) hipchk(launch_gemm(dp_.p, h, 1, c.W, 1, h, da_.p, h, B, h, h, nullptr, 0, nullptr, s_), "dx");
    }
  }

 private:
  void ensure(uint32_t B) {
    if (B <= cap_) return;
    cap_ = B;
    for (DBuf<float>* b : {&a1_, &a2_, &a3_, &p1_, &p2_, &p3_, &da_, &dp_}) b->alloc((size_t)B * h_);
    dy_.alloc(B);
  }
  uint32_t n_in_, h_, cap_ = 0;
  hipStream_t s_;
  DBuf<float> P_, G_, a1_, a2_, a3_, p1_, p2_, p3_, da_, dp_, dy_;
};

struct StreamOwner {
  hipStream_t s = nullptr;
  StreamOwner() { hipchk(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate"); }
  ~StreamOwner() {
    (void)hipStreamSynchronize(s);
    (void)hipStreamDestroy(s);
  }
};

// fixed sample of patch indices for the learned-lambda features: nsamp
// distinct indices of [0, p_sel) (all of them, repeated, when p_sel is
// smaller), sorted; the same for every window and call
std::vector<uint32_t> feature_sample(uint32_t p_sel, uint32_t nsamp) {
  std::vector<uint32_t> idx(p_sel);
  for (uint32_t i = 0; i < p_sel; ++i) idx[i] = i;
  std::vector<uint32_t> out(nsamp);
  if (p_sel >= nsamp) {
    std::mt19937_64 rng(0x5EEDF00Dull);
    for (uint32_t i = 0; i < nsamp; ++i) {
      const uint32_t j = i + (uint32_t)(rng() % (p_sel - i));
      std::swap(idx[i], idx[j]);
      out[i] = idx[i];
    }
    std::sort(out.begin(), out.end());
  } else {
    for (uint32_t i = 0; i < nsamp; ++i) out[i] = i % p_sel;
  }
  return out;
}

constexpr uint32_t kFeatureSamples = 256;

// What one call computes.  mode 0: the reference driver (PGURE search or the
// fixed lambda); 1: lambda proposed by the learned network; 2: only the
// learned-lambda features of each window (no output frames); 3: PGURE risk
// over a grid of lambdas per window, least risk chosen, the whole GPU work of
// a batch captured in one graph.

// Lanes-per-matrix mapping of the SVD per (patch, window) shape.  The two
// mappings sum the column dot products in different orders, so their
// results differ in the last bits; a mapping that depends on timing would
// make a call's lambdas and frames depend on the process.  The default is
// therefore the fixed primary mapping of every shape; pgsvt_set_svd_mapping
// forces the alternative (1) or the autotuner (2): the first batch of a
// shape in the process times both mappings on its own matrices and the
// faster is kept for the shape (results then reproducible within the
// process only).
struct Autotune {
  std::mutex mu;
  std::map<std::pair<uint32_t, uint32_t>, int> svd;
  static Autotune& get() {
    static Autotune* a = new Autotune;
    return *a;
  }
};
struct RunSpec {
  uint32_t tb = 0, te = 0;  // output windows [tb, te)
  int mode = 0;             // 3: PGURE over a lambda grid, each batch one graph launch
  uint32_t grid = 64;       // mode 3: lambda candidates per window
  const pgsvt_lambda_model* model = nullptr;
  float* features = nullptr;  // mode 2: [te - tb][n_in]
};

// A1 driver (src/pguresvt.hpp:90-166) over output windows [rs.tb, rs.te):
// Y holds their frames in order (Y slot t - tb), est their (lambda, alpha,
// mu, sigma) as a (te - tb) x 4 column-major matrix.
int run_core(const void* X, int dtype, uint32_t nr, uint32_t nc, uint32_t nf, const pgsvt_params* pp, double* Y,
             double* est, bool on_device, const RunSpec& rs) {
  fault_arm();
  if (int rc = validate(X, dtype, nr, nc, nf, pp)) return rc;
  if (!(rs.tb < rs.te && rs.te <= nf)) return fail(PGSVT_E_ARG, "window range must satisfy t_begin < t_end <= n_frames");
  if (!est || (rs.mode != 2 && !Y) || (rs.mode == 2 && !rs.features)) return fail(PGSVT_E_ARG, "null output pointer");
  if (rs.mode == 1 && (!rs.model || !rs.model->params || !rs.model->mean || !rs.model->inv_std))
    return fail(PGSVT_E_ARG, "learned lambda needs a model");
  if (int rc = check_device()) return rc;
  HostTimer total_timer("host_total_call");
  const bool learned = rs.mode == 1 || rs.mode == 2;
  const uint32_t count = rs.te - rs.tb;
  pgsvt_params pl = *pp;
  // learned lambda: the U decompositions only, no search, unless the probe
  // is on (then every variant volume is prepared: the probe decomposes a
  // patch subsample in all three, and windows that fail it are searched)
  double probe_delta = 0, probe_gap_max = 0;
  uint32_t probe_refine = 0;
  {
    std::lock_guard<std::mutex> lk(g_fallback_mu);
    probe_delta = g_probe_delta;
    probe_gap_max = g_probe_gap;
    probe_refine = g_probe_refine;
  }
  const bool probe = rs.mode == 1 && probe_gap_max > 0;
  if (learned) pl.optimize_pgure = probe ? 1 : 0;
  g_last_fallbacks = 0;
  if (rs.mode == 3) pl.optimize_pgure = 1;  // the grid needs the risk's perturbed decompositions
  if (rs.mode == 3 && (rs.grid < 1 || rs.grid > 1024)) return fail(PGSVT_E_ARG, "lambda grid size must be 1..1024");
  Pipeline P;
  P.tb = rs.tb;
  P.host_noise = learned;
  {
    HostTimer t("host_setup");
    P.setup(X, dtype, nr, nc, nf, pl, on_device);
    if (probe) P.setup_probe();
  }
  const pgsvt_params& p = P.p;
  const Geometry& g = P.g;
  const uint32_t n_in = kFeatureSamples * g.Nw + 3;
  if (rs.mode == 1 && rs.model->n_in != n_in)
    return fail(PGSVT_E_ARG, "learned-lambda model expects " + std::to_string(rs.model->n_in) +
                                 " features, this shape gives " + std::to_string(n_in));
  if (learned && P.sh.fp64) return fail(PGSVT_E_ARG, "the learned-lambda path runs in FP32 only");
  const double lambda0 = (p.lambda1 >= 0.0) ? p.lambda1 : -1.0;
  const double alpha0 = (p.noise_alpha >= 0.0) ? p.noise_alpha : -1.0;
  const double mu0 = (p.noise_mu >= 0.0) ? p.noise_mu : -1.0;
  const double sigma0 = (p.noise_sigma >= 0.0) ? p.noise_sigma : -1.0;
  if (rs.mode != 2) {
    if (on_device) {
      P.Yout = Y;
    } else {
      P.Yd.alloc((size_t)P.sh.fs * count);
      P.Yout = P.Yd.p;
    }
  }
  // Buffers sized by the requested windows (small) come before the batch,
  // whose allocation halves the batch until it fits beside concurrent calls.
  // learned lambda: fixed patch sample, feature buffer, network
  std::unique_ptr<Mlp> net;
  DBuf<uint32_t> sample_d;
  DBuf<float> feat_d, nz_d, mean_d, istd_d, yout_d;
  const uint32_t nlog = kFeatureSamples * g.Nw;
  if (learned || rs.mode == 3) {
    const std::vector<uint32_t> smp = feature_sample(P.sh.p_sel, kFeatureSamples);
    sample_d.alloc(smp.size());
    hipchk(hipMemcpyAsync(sample_d.p, smp.data(), smp.size() * 4, hipMemcpyHostToDevice, P.s), "H2D sample");
    feat_d.alloc((size_t)count * n_in);
    nz_d.alloc((size_t)count * 3);
    yout_d.alloc(count);
    if (rs.mode == 1) {
      net.reset(new Mlp(n_in, rs.model->hidden, P.s));
      hipchk(hipMemcpyAsync(net->params(), rs.model->params, net->n_params() * 4, hipMemcpyHostToDevice, P.s),
             "H2D model");
      mean_d.alloc(n_in);
      istd_d.alloc(n_in);
      hipchk(hipMemcpyAsync(mean_d.p, rs.model->mean, n_in * 4, hipMemcpyHostToDevice, P.s), "H2D mean");
      hipchk(hipMemcpyAsync(istd_d.p, rs.model->inv_std, n_in * 4, hipMemcpyHostToDevice, P.s), "H2D inv_std");
    }
    hipchk(hipStreamSynchronize(P.s), "model sync");  // host vectors above are local
  }

  // grid mode: per-batch buffers and the captured graph
  const uint32_t G = rs.grid;
  DBuf<double> lam_g, sums_g, risk_g, noise_g;
  DBuf<uint32_t> wl_g;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  uint32_t graph_nb = 0;
  struct GraphOwner {
    hipGraph_t* g;
    hipGraphExec_t* e;
    ~GraphOwner() {
      if (*e) (void)hipGraphExecDestroy(*e);
      if (*g) (void)hipGraphDestroy(*g);
    }
  } graph_owner{&graph, &gexec};
  int svd_mapping = 0;
  if (rs.mode == 3) {
    lam_g.alloc((size_t)count * G);
    sums_g.alloc((size_t)count * G * 6);
    risk_g.alloc((size_t)count * G);
    wl_g.alloc((size_t)count * G);
    noise_g.alloc((size_t)count * 3);
  }
  {
    HostTimer t("host_alloc_batch");
    P.alloc_batch(count);
  }

  // Noise estimates on host threads.  Reference path: leaf statistics per
  // distinct (frame, window maximum), as the reference normalises each
  // window by its own maximum.  Learned path: leaf statistics once per frame,
  // normalised by the sequence maximum, then rescaled to each window's
  // maximum (means by the ratio, variances by its square) before the fit.
  struct LeafJob {
    uint32_t frame;
    double scale;
    std::promise<std::shared_ptr<const FrameLeaves>> pr;
    std::shared_future<std::shared_ptr<const FrameLeaves>> fu;
  };
  std::vector<std::unique_ptr<LeafJob>> leaf_jobs;
  std::unique_ptr<PrepPool> pool;
  // learned path: the GPU leaf statistics on a thread of their own (joined
  // before the pool and the jobs it serves go away)
  std::thread leaf_thread;
  struct Joiner {
    std::thread& t;
    ~Joiner() {
      if (t.joinable()) t.join();
    }
  } leaf_joiner{leaf_thread};
  const double gmax = *std::max_element(P.fmaxX.begin(), P.fmaxX.end());
  if (P.optimize || learned) {
    // learned lambda: the leaf statistics of every frame on the GPU (one
    // scale for all); the reference path keeps the host's bit-exact ones
    const bool leaves_on_gpu = learned && nr >= 16;
    std::map<std::pair<uint32_t, uint64_t>, uint32_t> key_of;
    std::vector<std::function<void(PrepPool&)>> jobs;
    for (uint32_t t = rs.tb; t < rs.te; ++t) {
      const WinMeta m = P.meta_of(t);
      const double scale = learned ? gmax : m.uMax;
      uint64_t sbits;
      std::memcpy(&sbits, &scale, 8);
      std::vector<uint32_t> need;
      for (uint32_t k = 0; k < g.Nw; ++k) {
        const auto key = std::make_pair(m.s0 + k, sbits);
        auto it = key_of.find(key);
        if (it == key_of.end()) {
          const uint32_t id = (uint32_t)leaf_jobs.size();
          leaf_jobs.emplace_back(new LeafJob{m.s0 + k, scale, {}, {}});
          LeafJob* J = leaf_jobs.back().get();
          J->fu = J->pr.get_future().share();
          if (!leaves_on_gpu)
            jobs.push_back([J, &P, nr](PrepPool&) {
              try {
                std::vector<double> buf((size_t)nr * nr);
                P.host_frame(J->frame, J->scale, buf.data());
                J->pr.set_value(std::make_shared<const FrameLeaves>(frame_leaves(buf.data(), nr)));
              } catch (...) {
                J->pr.set_exception(std::current_exception());
              }
            });
          it = key_of.emplace(key, id).first;
        }
        need.push_back(it->second);
      }
      const double ratio = gmax / m.uMax;
      jobs.push_back([t, need, ratio, learned, &leaf_jobs, &P, &p, &rs, g, nr, nc, lambda0, alpha0, mu0,
                      sigma0](PrepPool& pool) {
        WindowPrep w;
        w.alpha = alpha0;
        w.mu = mu0;
        w.sigma = sigma0;
        try {
          std::vector<std::shared_ptr<const FrameLeaves>> leaves;
          for (uint32_t id : need) {
            auto lv = leaf_jobs[id]->fu.get();
            if (learned && ratio != 1.0) {
              auto sc = std::make_shared<FrameLeaves>(*lv);
              for (double& v : sc->mean) v *= ratio;
              for (double& v : sc->var) v *= ratio * ratio;
              lv = sc;
            }
            leaves.push_back(lv);
          }
          estimate_from_leaves(leaves, nr, p.noise_method, w.alpha, w.mu, w.sigma);
          const double OoNxNyNt = 1.0 / (nc * nr * g.Nt);
          const double start = (lambda0 >= 0.0) ? lambda0 : P.host_window_accu(t) * OoNxNyNt;
          w.start = std::max(0.0, start);
        } catch (const NoiseError& e) {
          w.err = std::string("noise estimate: ") + e.what();
          w.numeric = true;
        } catch (const std::exception& e) {
          w.err = std::string("noise estimate: ") + e.what();
        } catch (...) {
          w.err = "noise estimate failed";
        }
        pool.publish(t - rs.tb, std::move(w));
      });
    }
    if (leaves_on_gpu) {
      // on a stream of their own, beside the decomposition; the window jobs
      // wait for the frames they need
      int dev = 0;
      hipchk(hipGetDevice(&dev), "hipGetDevice");
      std::vector<uint32_t> frames;
      for (const auto& J : leaf_jobs) frames.push_back(J->frame);
      const double* Xd = P.Xd.p;
      leaf_thread = std::thread([&leaf_jobs, frames, dev, nr, gmax, Xd]() {
        size_t done = 0;
        try {
          hipchk(hipSetDevice(dev), "hipSetDevice");
          hipStream_t s2 = nullptr;
          hipchk(hipStreamCreateWithFlags(&s2, hipStreamNonBlocking), "hipStreamCreate");
          struct StreamGuard {
            hipStream_t s;
            ~StreamGuard() {
              (void)hipStreamSynchronize(s);
              (void)hipStreamDestroy(s);
            }
          } guard{s2};
          HostTimer tl("noise_leaves_gpu");
          std::vector<FrameLeaves> lv = gpu_frame_leaves(Xd, nr, frames, gmax, s2);
          for (; done < leaf_jobs.size(); ++done)
            leaf_jobs[done]->pr.set_value(std::make_shared<const FrameLeaves>(std::move(lv[done])));
        } catch (...) {
          for (; done < leaf_jobs.size(); ++done) leaf_jobs[done]->pr.set_exception(std::current_exception());
        }
      });
    }
    pool.reset(new PrepPool(std::move(jobs), count, thread_count(p.n_jobs)));
  }


  int status = 0;
  std::string warn;
  for (uint32_t t0 = rs.tb; t0 < rs.te; t0 += P.NB) {
    const uint32_t nb = std::min(P.NB, rs.te - t0);
    std::vector<WinMeta> ms(nb);
    for (uint32_t b = 0; b < nb; ++b) ms[b] = P.meta_of(t0 + b);
    if (rs.mode == 3) {
      std::vector<WindowPrep> prep(nb);
      P.upload_batch(ms);
      std::vector<double> nz((size_t)nb * 3);
      {
        HostTimer tw("host_noise_wait");
        for (uint32_t b = 0; b < nb; ++b) {
          prep[b] = pool->wait(t0 + b - rs.tb);
          prep[b].check();
          nz[3 * b] = prep[b].alpha;
          nz[3 * b + 1] = prep[b].mu;
          nz[3 * b + 2] = prep[b].sigma;
        }
      }
      std::vector<uint32_t> wlg((size_t)nb * G), wlf(nb);
      for (uint32_t q = 0; q < nb * G; ++q) wlg[q] = q / G;
      for (uint32_t b = 0; b < nb; ++b) wlf[b] = b;
      hipchk(hipMemcpyAsync(noise_g.p, nz.data(), nz.size() * 8, hipMemcpyHostToDevice, P.s), "H2D noise");
      hipchk(hipMemcpyAsync(wl_g.p, wlg.data(), wlg.size() * 4, hipMemcpyHostToDevice, P.s), "H2D grid list");
      hipchk(hipMemcpyAsync(P.wl_d.p, wlf.data(), wlf.size() * 4, hipMemcpyHostToDevice, P.s), "H2D list");
      hipchk(hipStreamSynchronize(P.s), "grid upload sync");
      if (!gexec) {  // first batch of the call: the SVD mapping
        svd_mapping = P.sh.fp64 ? 0 : std::min(g_svd_mapping.load(), 1);
      }
      if (!gexec && g_svd_mapping.load() == 2) {  // autotuned once per shape
        auto& at = Autotune::get();
        const auto key = std::make_pair(P.sh.bs, P.sh.nw);
        bool known;
        {
          std::lock_guard<std::mutex> lk(at.mu);
          known = at.svd.count(key) != 0;
          if (known) svd_mapping = at.svd[key];
        }
        if (!known && !P.sh.fp64) {
          // time both mappings on this batch's own matrices
          P.decompose_kernels(nb, P.nvar, 0);
          float ms_map[2] = {0, 0};
          for (int m = 0; m < 2; ++m) {
            hipEvent_t e0, e1;
            hipchk(hipEventCreate(&e0), "event");
            hipchk(hipEventCreate(&e1), "event");
            hipchk(hipEventRecord(e0, P.s), "event record");
            hipchk(launch_svd(P.uvar.p, P.pos.p, P.sel_d.p, nb, P.nvar, P.nvar, P.sh, P.factors.p, 30, 1e-6, P.stats.p,
                              P.s,
                              m),
                   "svd autotune");
            hipchk(hipEventRecord(e1, P.s), "event record");
            hipchk(hipEventSynchronize(e1), "event sync");
            hipchk(hipEventElapsedTime(&ms_map[m], e0, e1), "event time");
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
          }
          svd_mapping = ms_map[1] < ms_map[0] ? 1 : 0;
          std::lock_guard<std::mutex> lk(at.mu);
          at.svd[key] = svd_mapping;
        }
      }
      if (!gexec || graph_nb != nb) {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        gexec = nullptr;
        graph = nullptr;
        // one graph: decomposition, lambda grid, risk of every candidate in
        // chunks of NB slots, argmin, final frames
        const uint32_t s_at = g.Nw * ((g.Nw + 3) & ~3u) + g.Nw * P.sh.m;
        const uint32_t slots = nb * G, C = P.NB;
        const double OoN = 1.0 / ((double)nr * nc * g.Nw);
        P.prof->suspended = true;
        hipchk(hipStreamBeginCapture(P.s, hipStreamCaptureModeThreadLocal), "begin capture");
        P.decompose_kernels(nb, P.nvar, svd_mapping);
        hipchk(launch_grid_lambdas(P.factors.p, P.nvar, P.sh, s_at, sample_d.p, kFeatureSamples, nb, G,
                                   p.exponential_weighting, lam_g.p, P.s),
               "grid lambdas");
        for (uint32_t c0 = 0; c0 < slots; c0 += C) {
          const uint32_t na = std::min(C, slots - c0);
          P.evaluate(wl_g.p + c0, na, 3, lam_g.p + c0, P.partial.p, nullptr);
          GatherSrc gs = P.gather_src();
          gs.slot_lam = lam_g.p + c0;
          hipchk(launch_gather(gs, P.off.p, P.invw.p, P.u64.p, P.meta_d.p, wl_g.p + c0, na, P.sh, 0, nullptr,
                               P.partial_b.p, P.s),
                 "grid gather");
          hipchk(launch_risk_reduce(P.partial.p, eval_blocks(P.sh, 3), P.partial_b.p, P.gather_blocks, na,
                                    sums_g.p + (size_t)c0 * 6, P.s),
                 "grid reduce");
        }
        hipchk(launch_grid_select(sums_g.p, noise_g.p, nb, G, OoN, P.eps2, lam_g.p, P.meta_d.p, risk_g.p, P.s),
               "grid select");
        P.evaluate(P.wl_d.p, nb, 1, nullptr, nullptr, nullptr);
        hipchk(launch_gather(P.gather_src(), P.off.p, P.invw.p, P.u64.p, P.meta_d.p, P.wl_d.p, nb, P.sh, 1, P.Yout,
                             nullptr, P.s),
               "gather final");
        const hipError_t ce = hipStreamEndCapture(P.s, &graph);
        P.prof->suspended = false;
        hipchk(ce, "end capture");
        hipchk(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0), "graph instantiate");
        graph_nb = nb;
      }
      P.prof->run("grid_graph", (double)nb * G, [&] { hipchk(hipGraphLaunch(gexec, P.s), "graph launch"); });
      hipchk(hipMemcpyAsync(ms.data(), P.meta_d.p, nb * sizeof(WinMeta), hipMemcpyDeviceToHost, P.s), "D2H meta");
      hipchk(hipStreamSynchronize(P.s), "graph sync");
      P.prof->flush();
      for (uint32_t b = 0; b < nb; ++b) {
        const uint32_t i = t0 + b - rs.tb;
        est[i + (size_t)count * 0] = ms[b].lambda;
        est[i + (size_t)count * 1] = prep[b].alpha;
        est[i + (size_t)count * 2] = prep[b].mu;
        est[i + (size_t)count * 3] = prep[b].sigma;
      }
      continue;
    }
    // the learned path gathers only the probe's frames and the output
    // frames: the static data of those frames suffice
    P.static_frames = learned ? (probe ? gather_frames(P.sh, kProbeFrames) : 1u) : 0u;
    P.decompose(ms, learned ? 1 : P.nvar);
    if (probe) P.decompose_probe(nb);
    std::vector<WindowPrep> prep(nb);
    std::vector<uint32_t> fallback;  // learned lambda: windows the probe sends to the search
    SearchStart restart;             // and where their searches start
    if (learned) {
      // log singular values of the sampled patches (standardised for the
      // network), then the noise features of each window as they arrive
      const uint32_t s_at = g.Nw * ((g.Nw + 3) & ~3u) + g.Nw * P.sh.m;
      P.prof->run("learned_features", (double)nb, [&] {
        hipchk(launch_features((const float*)P.factors.p, P.nvar, P.sh.p_sel, P.sh.fstride, s_at, g.Nw, sample_d.p,
                               kFeatureSamples, nb, n_in, rs.mode == 1 ? mean_d.p : nullptr,
                               rs.mode == 1 ? istd_d.p : nullptr, feat_d.p, P.s),
               "features");
      });
      std::vector<float> nz((size_t)nb * 3);
      {
        HostTimer tw("host_noise_wait");
        for (uint32_t b = 0; b < nb; ++b) {
          prep[b] = pool->wait(t0 + b - rs.tb);
          prep[b].check();
          const double raw[3] = {prep[b].alpha, prep[b].mu, prep[b].sigma};
          for (int q = 0; q < 3; ++q) {
            float v = (float)raw[q];
            if (rs.mode == 1) v = (v - rs.model->mean[nlog + q]) * rs.model->inv_std[nlog + q];
            nz[(size_t)b * 3 + q] = v;
          }
        }
      }
      // one contiguous upload, then into the feature rows on the device (a
      // strided upload from pageable memory is split into a copy per row)
      hipchk(hipMemcpyAsync(nz_d.p, nz.data(), nz.size() * 4, hipMemcpyHostToDevice, P.s), "H2D noise features");
      hipchk(launch_copy_rows(nz_d.p, feat_d.p + nlog, 1, 3, 1, 3, n_in, nb, P.s), "noise features");
      if (rs.mode == 2) {
        hipchk(hipMemcpyAsync(rs.features + (size_t)(t0 - rs.tb) * n_in, feat_d.p, (size_t)nb * n_in * 4,
                              hipMemcpyDeviceToHost, P.s),
               "D2H features");
        hipchk(hipStreamSynchronize(P.s), "features sync");
        for (uint32_t b = 0; b < nb; ++b) ms[b].lambda = 0.0;
      } else {
        std::vector<float> yh(nb);
        P.prof->run("learned_mlp", (double)nb, [&] { net->forward(feat_d.p, nb, yout_d.p); });
        hipchk(hipMemcpyAsync(yh.data(), yout_d.p, nb * 4, hipMemcpyDeviceToHost, P.s), "D2H lambda");
        hipchk(hipStreamSynchronize(P.s), "mlp sync");  // also orders the noise-feature upload
        for (uint32_t b = 0; b < nb; ++b) {
          const double l = std::exp((double)rs.model->y_mean + (double)rs.model->y_std * (double)yh[b]);
          ms[b].lambda = std::min(std::max(l, 0.0), std::max(100.0, prep[b].start));  // the search's bounds
        }
        P.prof->flush();
        if (probe) {
          // local optimality of each proposal: risks at lambda / f, lambda,
          // lambda f (clamped to the search's bounds).  A window whose
          // parabola predicts a gap above the limit moves to the parabola's
          // minimum and is probed again there, up to probe_refine times; a
          // window still above the limit then goes to the search, which
          // starts at the last estimate.
          const double f = 1.0 + probe_delta;
          std::vector<uint32_t> active(nb);
          for (uint32_t b = 0; b < nb; ++b) active[b] = b;
          double refined = 0;
          for (uint32_t round = 0; !active.empty(); ++round) {
            std::vector<double> lam(active.size() * 3), risk;
            for (size_t i = 0; i < active.size(); ++i) {
              const uint32_t b = active[i];
              const double l = ms[b].lambda, ub = std::max(100.0, prep[b].start);
              lam[3 * i] = l / f;
              lam[3 * i + 1] = l;
              lam[3 * i + 2] = std::min(l * f, ub);
            }
            P.probe_risks(ms, prep, active, lam, risk);
            std::vector<uint32_t> again;
            for (size_t i = 0; i < active.size(); ++i) {
              const uint32_t b = active[i];
              const ProbeFit pf = probe_fit(&lam[3 * i], &risk[3 * i]);
              if (pf.gap <= probe_gap_max) continue;  // accepted at its current lambda
              const double ub = std::max(100.0, prep[b].start);
              const double next = std::min(pf.start, ub);
              if (round < probe_refine) {
                ms[b].lambda = next;
                again.push_back(b);
                refined += 1;
              } else {
                fallback.push_back(b);
                restart.x.push_back(next);
                restart.step.push_back(std::max(std::fabs(next - ms[b].lambda), probe_delta * ms[b].lambda));
              }
            }
            active.swap(again);
          }
          prof_host("probe_refinements", 0.0, refined);
        }
      }
    } else if (P.optimize) {
      {
        HostTimer t("host_decompose_sync");
        hipchk(hipStreamSynchronize(P.s), "decompose sync");
      }
      // A window joins the lockstep search as soon as its noise estimate is
      // ready; the rounds run while the host threads finish the others.
      lockstep_search(
          P, ms, prep, [&](uint32_t b) { return pool->ready(t0 + b - rs.tb); },
          [&](uint32_t b) {
            prep[b] = pool->wait(t0 + b - rs.tb);
            prep[b].check();
          },
          status, warn);
    } else {
      for (uint32_t b = 0; b < nb; ++b) {
        ms[b].lambda = lambda0;
        prep[b].alpha = alpha0;
        prep[b].mu = mu0;
        prep[b].sigma = sigma0;
      }
    }
    if (rs.mode != 2) P.finish(ms);
    if (!fallback.empty()) {
      // the windows that failed the probe: full decompositions and the
      // reference search, then their frames again
      HostTimer tf("learned_fallback_search");
      std::vector<WinMeta> mf;
      std::vector<WindowPrep> pf;
      for (uint32_t b : fallback) {
        mf.push_back(ms[b]);
        pf.push_back(prep[b]);
      }
      P.static_frames = 0;  // the search gathers every frame
      P.decompose(mf, 3);
      hipchk(hipStreamSynchronize(P.s), "fallback decompose sync");
      // SBPLX from the probe's estimate of the minimum, steps of the size of
      // the correction, to a relative tolerance far below the risk gaps the
      // probe is there to remove
      restart.tol = std::max(p.tol, 1e-5);
      lockstep_search(P, mf, pf, [](uint32_t) { return true; }, [](uint32_t) {}, status, warn, &restart);
      P.finish(mf);
      for (size_t i = 0; i < fallback.size(); ++i) ms[fallback[i]].lambda = mf[i].lambda;
      g_last_fallbacks += (int)fallback.size();
      prof_host("learned_fallbacks", 0.0, (double)fallback.size());
    }
    for (uint32_t b = 0; b < nb; ++b) {
      const uint32_t i = t0 + b - rs.tb;
      est[i + (size_t)count * 0] = ms[b].lambda;
      est[i + (size_t)count * 1] = prep[b].alpha;
      est[i + (size_t)count * 2] = prep[b].mu;
      est[i + (size_t)count * 3] = prep[b].sigma;
    }
  }
  if (rs.mode != 2 && !on_device)
    hipchk(hipMemcpyAsync(Y, P.Yd.p, P.Yd.bytes(), hipMemcpyDeviceToHost, P.s), "D2H Y");
  hipchk(hipStreamSynchronize(P.s), "final sync");
  if (status) return fail(status, warn);
  return 0;
}

int run_denoise(const void* X, int dtype, uint32_t nr, uint32_t nc, uint32_t nf, const pgsvt_params* pp, double* Y,
                double* est, bool on_device) {
  RunSpec rs;
  rs.tb = 0;
  rs.te = nf;
  return run_core(X, dtype, nr, nc, nf, pp, Y, est, on_device, rs);
}

template <class F>
int guarded(F f) {
  try {
    return f();
  } catch (const NoiseError& e) {
    return fail(PGSVT_E_NUMERIC, e.what());
  } catch (const HipError& e) {
    return fail(PGSVT_E_HIP, std::string(e.what) + ": " + hipGetErrorString(e.e));
  } catch (const std::bad_alloc&) {
    return fail(PGSVT_E_HIP, "host allocation failed");
  } catch (const std::exception& e) {
    return fail(PGSVT_E_HIP, e.what());
  }
}

// --------------------------------------------- entry point without allocation --
// pgsvt_denoise_device_noalloc runs the reference driver window by window
// with every device buffer carved from one statically sized array of device
// memory, so that callers that cannot allocate (code under HIP graph capture,
// serving loops) can use it.  Each device the library runs on holds its own
// copy of the array from the first such call until the process ends; calls
// are serialised within the process (they share the array), and the array
// keeps the last call's contents.  The kernels, their launch shapes and the
// order of every sum are those of pgsvt_denoise_device, so the results are
// the same bit for bit.
constexpr size_t kNoAllocScratch = 64ull << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratch];

std::mutex g_scratch_mu;
std::map<int, unsigned char*> g_scratch_at;  // device -> address of its copy of `scratch`

// address of the current device's copy (g_scratch_mu held): the host-side
// name `scratch` is not a device address
unsigned char* scratch_address() {
  int dev = 0;
  hipchk(hipGetDevice(&dev), "hipGetDevice");
  auto it = g_scratch_at.find(dev);
  if (it != g_scratch_at.end()) return it->second;
  void* p = nullptr;
  hipchk(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
  g_scratch_at[dev] = (unsigned char*)p;
  return (unsigned char*)p;
}

// 256-byte aligned pieces of a region; with base == nullptr it only counts
struct Carve {
  unsigned char* base = nullptr;
  size_t used = 0;
  template <class T>
  T* take(size_t count) {
    used = (used + 255) & ~(size_t)255;
    T* p = base ? reinterpret_cast<T*>(base + used) : nullptr;
    used += count * sizeof(T);
    return p;
  }
};

// PGURE risk of src/pgure.hpp:136 from the six GPU sums (see Pipeline::risk)
double window_risk_of(const double* S, double alpha, double mu, double sigma, const Shape& sh, double eps2) {
  const double mup = sigma, sigp = mu;  // constructor receives (alpha, sigma, mu)
  const double sigmasq = sigp * sigp;
  const double OoN = 1.0 / (sh.nr * sh.nc * sh.nw);
  const double S1 = S[0] - 2.0 * S[3] + S[2];
  return OoN * (S1 - (alpha + mup) * S[1] + (2 / 1e-4 * 0.0) - (2 * sigmasq * alpha / (eps2 * eps2) * S[4]) +
                (2 * mup * S[5]) + mup) -
         sigmasq;
}

int run_noalloc(const void* X, int dtype, uint32_t nr, uint32_t nc, uint32_t nf, const pgsvt_params* pp, double* Y,
                double* est, hipStream_t s) {
  if (int rc = validate(X, dtype, nr, nc, nf, pp)) return rc;
  if (!Y || !est) return fail(PGSVT_E_ARG, "null output pointer");
  if (g_precision.load() != 32) return fail(PGSVT_E_ARG, "the entry point without allocation runs in FP32 only");
  if (int rc = check_device()) return rc;
  const pgsvt_params p = *pp;
  const Geometry g = geometry(p);
  const bool optimize = p.optimize_pgure != 0, median = p.motion_filter > 0;
  const uint32_t nvar = optimize ? 3 : 1;
  Shape sh;
  sh.nr = nr;
  sh.nc = nc;
  sh.fs = nr * nc;
  sh.nw = g.Nw;
  sh.vol = g.Nw * sh.fs;
  sh.bs = p.patch_size;
  sh.m = sh.bs * sh.bs;
  sh.n1 = nr - sh.bs + 1;
  sh.p_all = sh.n1 * sh.n1;
  const std::vector<uint32_t> sel = patch_indices(nr, sh.bs, p.patch_overlap);
  sh.p_sel = (uint32_t)sel.size();
  const uint32_t sr = (g.Nw + 3) & ~3u;
  sh.fstride = factor_stride(g.Nw, sh.m);
  const size_t npos = (size_t)sh.n1 * sh.n1, vol = sh.vol;
  const uint32_t max_jobs = 2 * g.Nw, eblk = eval_blocks(sh, 3), gblk = pgsvt::gather_blocks(sh, 0);
  const double vP = 0.5 + 0.5 * 1. / std::sqrt(1. * 1. + 4), vQ = 1 - vP;
  const double d2neg = -1 * std::sqrt(vQ / vP), d2pos = std::sqrt(vP / vQ), eps2 = 1e-2;
  // the layout: per call, then the window being processed
  struct Layout {
    uint8_t* d2;
    uint32_t *sel, *wl, *stats;
    double *fmax, *sums, *Xw, *Zw, *u64, *w64, *invw, *partial, *partial_b;
    WinMeta* meta;
    int16_t* jobs;
    float *uvar, *dw, *B0, *factors, *divq, *w0;
    int32_t *pos, *cnt, *off, *list, *inv;
    char* s01;
  } L{};
  auto carve = [&](Carve& c) {
    L.d2 = optimize ? c.take<uint8_t>(vol) : nullptr;
    L.sel = c.take<uint32_t>(sh.p_sel);
    L.wl = c.take<uint32_t>(1);
    L.stats = c.take<uint32_t>(1);
    L.fmax = c.take<double>(2 * g.Nw);
    L.sums = c.take<double>(6);
    L.meta = c.take<WinMeta>(1);
    L.jobs = c.take<int16_t>((size_t)max_jobs * 4);
    L.Xw = c.take<double>(vol);
    L.Zw = median ? c.take<double>(vol) : nullptr;
    L.u64 = c.take<double>(vol);
    L.w64 = median ? c.take<double>(vol) : nullptr;
    L.uvar = c.take<float>((size_t)nvar * vol);
    L.pos = c.take<int32_t>((size_t)g.Nw * sh.p_all);
    L.cnt = c.take<int32_t>((size_t)g.Nw * npos);
    L.off = c.take<int32_t>((size_t)g.Nw * (npos + 1));
    L.list = c.take<int32_t>((size_t)g.Nw * sh.p_sel);
    L.inv = c.take<int32_t>((size_t)g.Nw * sh.p_sel);
    L.invw = c.take<double>(vol);
    L.dw = optimize ? c.take<float>(vol) : nullptr;
    L.B0 = c.take<float>((size_t)g.Nw * sh.p_sel * (fused_eval(sh) ? 1 : sh.m));
    L.s01 = fused_eval(sh) ? c.take<char>(singular_pair_bytes(sh)) : nullptr;
    L.w0 = fused_eval(sh) ? c.take<float>((size_t)g.Nw * sh.p_sel * sh.m) : nullptr;
    L.factors = c.take<float>((size_t)nvar * sh.p_sel * sh.fstride);
    L.divq = optimize ? c.take<float>((size_t)nvar * sh.p_sel * 2 * sr) : nullptr;
    L.partial = c.take<double>(eblk);
    L.partial_b = c.take<double>((size_t)gblk * 5);
  };
  Carve count;
  carve(count);
  if (count.used > kNoAllocScratch)
    return fail(PGSVT_E_ARG, "this shape needs " + std::to_string(count.used) + " bytes of device scratch per window; "
                             "the entry point without allocation has " + std::to_string(kNoAllocScratch));
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  Carve at;
  at.base = scratch_address();
  carve(at);

  const size_t dsz = dtype_size(dtype);
  const uint64_t seed = seed_of(p);
  const double lambda0 = (p.lambda1 >= 0.0) ? p.lambda1 : -1.0;
  const double alpha0 = (p.noise_alpha >= 0.0) ? p.noise_alpha : -1.0;
  const double mu0 = (p.noise_mu >= 0.0) ? p.noise_mu : -1.0;
  const double sigma0 = (p.noise_sigma >= 0.0) ? p.noise_sigma : -1.0;
  const uint32_t zero = 0;
  hipchk(hipMemcpyAsync(L.sel, sel.data(), sel.size() * sizeof(uint32_t), hipMemcpyHostToDevice, s), "H2D sel");
  hipchk(hipMemcpyAsync(L.wl, &zero, sizeof zero, hipMemcpyHostToDevice, s), "H2D list");
  if (optimize) hipchk(launch_delta2((int64_t)seed, vol, L.d2, s), "delta2");
  hipchk(hipStreamSynchronize(s), "setup sync");  // the host vectors above are local
  GatherSrc gsrc;
  gsrc.B0 = L.B0;
  gsrc.factors = L.factors;
  gsrc.dr = L.B0;
  gsrc.S01 = L.s01;
  gsrc.exp_w = p.exponential_weighting;
  gsrc.W0 = L.w0;
  gsrc.list = L.list;
  gsrc.nvar_stored = nvar;
  std::vector<double> fmax(2 * g.Nw), sums(6), hw;
  int status = 0;
  std::string warn;
  for (uint32_t t = 0; t < nf; ++t) {
    WinMeta m{};
    m.t = t;
    m.out = t;
    uint32_t s0 = 0;
    window_of(t, nf, g, s0, m.ref, m.aref);
    m.s0 = 0;  // the window's frames are staged at the start of Xw
    hipchk(launch_to_double((const char*)X + (size_t)s0 * sh.fs * dsz, dtype, L.Xw, vol, s), "to_double");
    hipchk(launch_frame_max(L.Xw, sh.fs, g.Nw, L.fmax, s), "frame max");
    if (median) {
      hipchk(launch_median(L.Xw, L.Zw, nr, nc, g.Nw, (int)p.motion_filter, s), "median");
      hipchk(launch_frame_max(L.Zw, sh.fs, g.Nw, L.fmax + g.Nw, s), "frame max Z");
    }
    hipchk(hipMemcpyAsync(fmax.data(), L.fmax, fmax.size() * sizeof(double), hipMemcpyDeviceToHost, s), "D2H max");
    hipchk(hipStreamSynchronize(s), "max sync");
    m.uMax = *std::max_element(fmax.begin(), fmax.begin() + g.Nw);
    m.wMax = median ? *std::max_element(fmax.begin() + g.Nw, fmax.end()) : m.uMax;
    m.lambda = 0;
    std::vector<int16_t> jobs((size_t)max_jobs * 4, (int16_t)-1);
    if (p.motion_estimation) {
      std::vector<int16_t> j;
      if (!arps_jobs(t, nf, g, j)) throw HipError{hipErrorInvalidValue, "unsupported ARPS schedule"};
      std::copy(j.begin(), j.end(), jobs.begin());
    }
    hipchk(hipMemcpyAsync(L.meta, &m, sizeof m, hipMemcpyHostToDevice, s), "H2D meta");
    hipchk(hipMemcpyAsync(L.jobs, jobs.data(), jobs.size() * sizeof(int16_t), hipMemcpyHostToDevice, s), "H2D jobs");
    // decomposition (Pipeline::decompose_kernels for one window)
    hipchk(launch_prepare(L.Xw, L.Zw, L.meta, 1, L.d2, sh, nvar, eps2, d2neg, d2pos, L.u64, L.w64, L.uvar, s),
           "prepare");
    hipchk(launch_fill32(L.pos, (size_t)g.Nw * sh.p_all, 0, s), "clear pos");
    hipchk(launch_fill_ref(L.meta, 1, sh, L.pos, s), "fill_ref");
    if (p.motion_estimation)
      hipchk(launch_arps(median ? L.w64 : L.u64, L.jobs, max_jobs, 1, sh, (int)p.motion_window, L.pos, s), "arps");
    hipchk(launch_buckets(L.pos, L.sel, 1, sh, L.cnt, L.off, L.list, L.inv, s), "buckets");
    hipchk(launch_weights(L.off, L.d2, 1, sh, d2neg, d2pos, L.invw, L.dw, s), "weights");
    hipchk(launch_fill32(L.stats, 1, 0, s), "clear stats");
    hipchk(launch_svd(L.uvar, L.pos, L.sel, 1, nvar, nvar, sh, L.factors, 30, 1e-6, L.stats, s,
                      g_svd_mapping.load() == 1 ? 1 : 0),
           "svd");
    if (fused_eval(sh)) hipchk(launch_static_rows(L.factors, L.inv, L.meta, 1, nvar, sh, L.w0, L.s01, s), "static rows");
    if (optimize) hipchk(launch_divq(L.factors, L.pos, L.sel, L.dw, 1, nvar, sh, L.divq, s), "divergence factors");
    double alpha = alpha0, mu = mu0, sigma = sigma0, lambda = lambda0;
    if (optimize) {
      // noise estimate and start point from the normalised window on the host
      hw.resize(vol);
      hipchk(hipMemcpyAsync(hw.data(), L.Xw, vol * sizeof(double), hipMemcpyDeviceToHost, s), "D2H window");
      hipchk(hipStreamSynchronize(s), "window sync");
      for (double& v : hw) v /= m.uMax;
      estimate_noise(hw.data(), nr, g.Nw, p.noise_method, alpha, mu, sigma);
      double a1 = 0, a2 = 0;  // accu(u) in Armadillo's even / odd order
      size_t i = 0;
      for (; i + 1 < vol; i += 2) {
        a1 += hw[i];
        a2 += hw[i + 1];
      }
      if (i < vol) a1 += hw[i];
      const double start = std::max(0.0, (lambda0 >= 0.0) ? lambda0 : (a1 + a2) * (1.0 / (nc * nr * g.Nt)));
      Sbplx1D opt(start, 0.0, std::max(100.0, start), std::sqrt(start), p.tol, 1e-12, (int)p.max_iter);
      std::vector<std::pair<double, double>> seen;  // repeated lambdas answered as in the batched driver
      int rc;
      while ((rc = opt.run()) == Sbplx1D::kNeed) {
        const double l = opt.pending();
        double r = 0;
        bool known = false;
        for (const auto& e : seen)
          if (e.first == l) {
            r = e.second;
            known = true;
            break;
          }
        if (!known) {
          m.lambda = l;
          hipchk(hipMemcpyAsync(L.meta, &m, sizeof m, hipMemcpyHostToDevice, s), "H2D meta");
          hipchk(launch_evaluation(sh, L.factors, L.divq, L.meta, L.wl, 1, nvar, 3, p.exponential_weighting, L.inv,
                                   L.B0, L.partial, nullptr, nullptr, s),
                 "eval");
          hipchk(launch_gather(gsrc, L.off, L.invw, L.u64, L.meta, L.wl, 1, sh, 0, nullptr, L.partial_b, s), "gather");
          hipchk(launch_risk_reduce(L.partial, eblk, L.partial_b, gblk, 1, L.sums, s), "reduce");
          hipchk(hipMemcpyAsync(sums.data(), L.sums, 6 * sizeof(double), hipMemcpyDeviceToHost, s), "D2H sums");
          hipchk(hipStreamSynchronize(s), "risk sync");
          r = window_risk_of(sums.data(), alpha, mu, sigma, sh, eps2);
          seen.emplace_back(l, r);
        }
        opt.supply(r);
      }
      lambda = opt.last_x();
      if (rc < 0) {
        status = PGSVT_E_NUMERIC;
        warn = "lambda optimisation failed (NLopt-equivalent error code " + std::to_string(rc) + ")";
      } else if (rc == Sbplx1D::kMaxEval) {
        std::fprintf(stderr,
                     "WARNING: optimization terminated after max_eval (%u) was reached.\n"
                     "Consider increasing the max_eval or increasing the convergence tolerance (tol=%g).\n",
                     p.max_iter, p.tol);
      }
    }
    // the output frame at the window's lambda (Pipeline::finish)
    m.lambda = lambda;
    hipchk(hipMemcpyAsync(L.meta, &m, sizeof m, hipMemcpyHostToDevice, s), "H2D meta");
    hipchk(launch_evaluation(sh, L.factors, nullptr, L.meta, L.wl, 1, nvar, 1, p.exponential_weighting, L.inv, L.B0,
                             nullptr, nullptr, nullptr, s),
           "eval final");
    hipchk(launch_gather(gsrc, L.off, L.invw, L.u64, L.meta, L.wl, 1, sh, 1, Y, nullptr, s), "gather final");
    hipchk(hipStreamSynchronize(s), "window sync");  // m and jobs are reused by the next window
    est[t + (size_t)nf * 0] = lambda;
    est[t + (size_t)nf * 1] = alpha;
    est[t + (size_t)nf * 2] = mu;
    est[t + (size_t)nf * 3] = sigma;
  }
  if (status) return fail(status, warn);
  return 0;
}

// one window through the GPU stages; used by the window-level entry points
int one_window(const void* X, int dtype, uint32_t nr, uint32_t nc, uint32_t nf, const pgsvt_params* pp, uint32_t t,
               bool want_risk, Pipeline& P, std::vector<WinMeta>& ms) {
  if (int rc = validate(X, dtype, nr, nc, nf, pp)) return rc;
  if (t >= nf) return fail(PGSVT_E_ARG, "window index out of range");
  if (int rc = check_device()) return rc;
  pgsvt_params q = *pp;
  q.optimize_pgure = want_risk ? 1 : 0;
  P.setup(X, dtype, nr, nc, nf, q);
  P.alloc_batch(1);
  ms.assign(1, P.meta_of(t));
  P.decompose(ms, P.nvar);
  return 0;
}

}  // namespace
}  // namespace pgsvt

using namespace pgsvt;

extern "C" {

void pgsvt_default_params(pgsvt_params* p) {
  p->trajectory_length = 15;
  p->patch_size = 4;
  p->patch_overlap = 1;
  p->motion_window = 7;
  p->motion_filter = 5;
  p->noise_method = 4;
  p->max_iter = 500;
  p->n_jobs = -1;
  p->random_seed = -1;
  p->optimize_pgure = 1;
  p->exponential_weighting = 1;
  p->motion_estimation = 1;
  p->lambda1 = 0.0;
  p->noise_alpha = -1.0;
  p->noise_mu = -1.0;
  p->noise_sigma = -1.0;
  p->tol = 1e-7;
}

const char* pgsvt_last_error(void) { return g_err.c_str(); }

const char* pgsvt_version(void) { return "pgsvt 0.1.0 (gfx950)"; }

int pgsvt_set_precision(int bits) {
  if (bits != 32 && bits != 64) return fail(PGSVT_E_ARG, "precision must be 32 or 64 bits");
  g_precision.store(bits);
  return 0;
}

int pgsvt_set_learned_fallback(double delta, double gap, uint32_t refine) {
  if (!(delta > 0 && delta < 1)) return fail(PGSVT_E_ARG, "probe step delta must be in (0, 1)");
  if (!std::isfinite(gap)) return fail(PGSVT_E_ARG, "probe gap limit must be finite");
  if (refine > 16) return fail(PGSVT_E_ARG, "at most 16 probe refinements");
  std::lock_guard<std::mutex> lk(g_fallback_mu);
  g_probe_delta = delta;
  g_probe_gap = gap;
  g_probe_refine = refine;
  return 0;
}

int pgsvt_learned_fallbacks(void) { return g_last_fallbacks; }

int pgsvt_set_svd_mapping(int mode) {
  if (mode < 0 || mode > 2) return fail(PGSVT_E_ARG, "SVD mapping must be 0, 1 or 2");
  g_svd_mapping.store(mode);
  return 0;
}

int pgsvt_release_cache(void) {
  return guarded([&] {
    int dev = 0;
    hipchk(hipGetDevice(&dev), "hipGetDevice");
    hipchk(hipDeviceSynchronize(), "hipDeviceSynchronize");
    DeviceCache::get().release(dev);
    PinnedCache::get().release();
    return 0;
  });
}

int pgsvt_denoise(const void* X, int dtype, uint32_t n_rows, uint32_t n_cols, uint32_t n_frames,
                  const pgsvt_params* p, double* Y, double* estimates) {
  if (!Y || !estimates) return fail(PGSVT_E_ARG, "null output pointer");
  return guarded([&] { return run_denoise(X, dtype, n_rows, n_cols, n_frames, p, Y, estimates, false); });
}

int pgsvt_denoise_device(const void* X_dev, int dtype, uint32_t n_rows, uint32_t n_cols, uint32_t n_frames,
                         const pgsvt_params* p, double* Y_dev, double* estimates) {
  if (!Y_dev || !estimates) return fail(PGSVT_E_ARG, "null output pointer");
  return guarded([&] { return run_denoise(X_dev, dtype, n_rows, n_cols, n_frames, p, Y_dev, estimates, true); });
}

int pgsvt_denoise_device_noalloc(const void* X_dev, int dtype, uint32_t n_rows, uint32_t n_cols, uint32_t n_frames,
                                 const pgsvt_params* p, double* Y_dev, double* estimates, void* stream) {
  return guarded([&] {
    return run_noalloc(X_dev, dtype, n_rows, n_cols, n_frames, p, Y_dev, estimates, (hipStream_t)stream);
  });
}

int pgsvt_denoise_range(const void* X, int dtype, uint32_t n_rows, uint32_t n_cols, uint32_t n_frames,
                        const pgsvt_params* p, uint32_t t_begin, uint32_t t_end, double* Y, double* estimates) {
  return guarded([&] {
    RunSpec rs;
    rs.tb = t_begin;
    rs.te = t_end;
    return run_core(X, dtype, n_rows, n_cols, n_frames, p, Y, estimates, false, rs);
  });
}

int pgsvt_learned_features(const void* X, int dtype, uint32_t n_rows, uint32_t n_cols, uint32_t n_frames,
                           const pgsvt_params* p, uint32_t t_begin, uint32_t t_end, float* features,
                           double* estimates) {
  return guarded([&] {
    RunSpec rs;
    rs.tb = t_begin;
    rs.te = t_end;
    rs.mode = 2;
    rs.features = features;
    return run_core(X, dtype, n_rows, n_cols, n_frames, p, nullptr, estimates, false, rs);
  });
}

int pgsvt_denoise_learned(const void* X, int dtype, uint32_t n_rows, uint32_t n_cols, uint32_t n_frames,
                          const pgsvt_params* p, const pgsvt_lambda_model* model, double* Y, double* estimates) {
  return guarded([&] {
    RunSpec rs;
    rs.tb = 0;
    rs.te = n_frames;
    rs.mode = 1;
    rs.model = model;
    return run_core(X, dtype, n_rows, n_cols, n_frames, p, Y, estimates, false, rs);
  });
}

int pgsvt_denoise_learned_device(const void* X_dev, int dtype, uint32_t n_rows, uint32_t n_cols, uint32_t n_frames,
                                 const pgsvt_params* p, const pgsvt_lambda_model* model, double* Y_dev,
                                 double* estimates) {
  return guarded([&] {
    RunSpec rs;
    rs.tb = 0;
    rs.te = n_frames;
    rs.mode = 1;
    rs.model = model;
    return run_core(X_dev, dtype, n_rows, n_cols, n_frames, p, Y_dev, estimates, true, rs);
  });
}

int pgsvt_denoise_grid(const void* X, int dtype, uint32_t n_rows, uint32_t n_cols, uint32_t n_frames,
                       const pgsvt_params* p, uint32_t n_lambda, double* Y, double* estimates) {
  return guarded([&] {
    RunSpec rs;
    rs.tb = 0;
    rs.te = n_frames;
    rs.mode = 3;
    rs.grid = n_lambda;
    return run_core(X, dtype, n_rows, n_cols, n_frames, p, Y, estimates, false, rs);
  });
}

int pgsvt_denoise_batch(pgsvt_job* jobs, uint32_t n, uint32_t n_lambda, uint32_t concurrency) {
  if (!jobs && n) return fail(PGSVT_E_ARG, "null job list");
  // largest first (patches x window x frames), `concurrency` calls at once,
  // each on its own stream with its share of the host threads
  std::vector<uint32_t> order(n);
  for (uint32_t i = 0; i < n; ++i) order[i] = i;
  auto cost = [&](uint32_t i) {
    const pgsvt_job& j = jobs[i];
    const double pr = (double)j.n_rows - j.params.patch_size + 1;
    return pr * pr * j.params.patch_size * j.params.patch_size * j.params.trajectory_length * j.n_frames;
  };
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return cost(a) > cost(b); });
  const uint32_t K = std::max(1u, std::min(concurrency ? concurrency : 1u, std::max(1u, n)));
  const uint32_t share = std::max(1u, thread_count(-1) / K);
  std::atomic<uint32_t> next{0};
  std::vector<std::thread> th;
  std::mutex mu;
  int first_rc = 0;
  std::string first_msg;
  for (uint32_t k = 0; k < K; ++k)
    th.emplace_back([&] {
      for (;;) {
        const uint32_t q = next.fetch_add(1);
        if (q >= n) return;
        pgsvt_job& j = jobs[order[q]];
        pgsvt_params pp = j.params;
        if (pp.n_jobs < 0) pp.n_jobs = share;
        RunSpec rs;
        rs.tb = 0;
        rs.te = j.n_frames;
        rs.mode = n_lambda ? 3 : 0;
        rs.grid = n_lambda ? n_lambda : 64;
        j.status = guarded([&] { return run_core(j.X, j.dtype, j.n_rows, j.n_cols, j.n_frames, &pp, j.Y,
                                                 j.estimates, false, rs); });
        if (j.status) {
          std::lock_guard<std::mutex> lk(mu);
          if (!first_rc) {
            first_rc = j.status;
            first_msg = "job " + std::to_string(order[q]) + ": " + pgsvt_last_error();
          }
        }
      }
    });
  for (auto& t : th) t.join();
  if (first_rc) return fail(first_rc, first_msg);
  return 0;
}

void pgsvt_profile_enable(int on) {
  std::lock_guard<std::mutex> g(g_prof_mu);
  g_prof_on = on != 0;
  g_prof_count = on >= 2;
  g_prof.clear();
}

int pgsvt_profile_read(char* buf, uint64_t len) {
  std::string js = "{";
  {
    std::lock_guard<std::mutex> g(g_prof_mu);
    for (size_t k = 0; k < g_prof.size(); ++k) {
      char line[256];
      std::snprintf(line, sizeof line, "%s\"%s\": {\"launches\": %llu, \"ms\": %.6f, \"units\": %.1f}",
                    k ? ", " : "", g_prof[k].name.c_str(), (unsigned long long)g_prof[k].launches, g_prof[k].ms,
                    g_prof[k].units);
      js += line;
    }
  }
  js += "}";
  if (!buf || len < js.size() + 1) return (int)(js.size() + 1);
  std::memcpy(buf, js.c_str(), js.size() + 1);
  return 0;
}

int pgsvt_window_trajectories(const void* X, int dtype, uint32_t n_rows, uint32_t n_cols, uint32_t n_frames,
                              const pgsvt_params* p, uint32_t t, int32_t* patches) {
  return guarded([&] {
    Pipeline P;
    std::vector<WinMeta> ms;
    if (int rc = one_window(X, dtype, n_rows, n_cols, n_frames, p, t, false, P, ms)) return rc;
    std::vector<int32_t> packed((size_t)P.sh.nw * P.sh.p_all);
    hipchk(hipMemcpyAsync(packed.data(), P.pos.p, packed.size() * 4, hipMemcpyDeviceToHost, P.s), "D2H pos");
    hipchk(hipStreamSynchronize(P.s), "sync");
    for (size_t k = 0; k < packed.size(); ++k) {
      patches[2 * k] = packed[k] & 0xffff;
      patches[2 * k + 1] = (packed[k] >> 16) & 0xffff;
    }
    return 0;
  });
}

int pgsvt_window_risk(const void* X, int dtype, uint32_t n_rows, uint32_t n_cols, uint32_t n_frames,
                      const pgsvt_params* p, uint32_t t, const double* noise, const double* lambdas,
                      uint32_t n_lambda, double* risk) {
  return guarded([&] {
    Pipeline P;
    std::vector<WinMeta> ms;
    if (int rc = one_window(X, dtype, n_rows, n_cols, n_frames, p, t, true, P, ms)) return rc;
    WindowPrep w;
    w.alpha = noise[0];
    w.mu = noise[1];
    w.sigma = noise[2];
    std::vector<double> sums;
    for (uint32_t k = 0; k < n_lambda; ++k) {
      ms[0].lambda = lambdas[k];
      P.eval_risk(ms, {0u}, sums);
      risk[k] = P.risk(sums.data(), w);
    }
    return 0;
  });
}

int pgsvt_window_reconstruct(const void* X, int dtype, uint32_t n_rows, uint32_t n_cols, uint32_t n_frames,
                             const pgsvt_params* p, uint32_t t, double lambda, double* vol) {
  return guarded([&] {
    Pipeline P;
    std::vector<WinMeta> ms;
    if (int rc = one_window(X, dtype, n_rows, n_cols, n_frames, p, t, false, P, ms)) return rc;
    ms[0].lambda = lambda;
    const uint32_t zero = 0;
    hipchk(hipMemcpyAsync(P.meta_d.p, ms.data(), sizeof(WinMeta), hipMemcpyHostToDevice, P.s), "H2D meta");
    hipchk(hipMemcpyAsync(P.wl_d.p, &zero, sizeof(uint32_t), hipMemcpyHostToDevice, P.s), "H2D list");
    DBuf<double> out;
    out.alloc(P.sh.vol);
    P.evaluate(P.wl_d.p, 1, 1, nullptr, nullptr, nullptr);
    hipchk(launch_gather(P.gather_src(), P.off.p, P.invw.p, P.u64.p, P.meta_d.p, P.wl_d.p, 1, P.sh, 2, out.p,
                         nullptr, P.s),
           "gather");
    hipchk(hipMemcpyAsync(vol, out.p, out.bytes(), hipMemcpyDeviceToHost, P.s), "D2H vol");
    hipchk(hipStreamSynchronize(P.s), "sync");
    return 0;
  });
}

int pgsvt_window_singular_values(const void* X, int dtype, uint32_t n_rows, uint32_t n_cols, uint32_t n_frames,
                                 const pgsvt_params* p, uint32_t t, double* s, uint32_t* n_patches) {
  return guarded([&] {
    Pipeline P;
    std::vector<WinMeta> ms;
    if (int rc = one_window(X, dtype, n_rows, n_cols, n_frames, p, t, false, P, ms)) return rc;
    const size_t es = P.sh.fp64 ? 8 : 4, nval = (size_t)P.sh.p_sel * P.sh.fstride;
    std::vector<char> f(nval * es);
    hipchk(hipMemcpyAsync(f.data(), P.factors.p, f.size(), hipMemcpyDeviceToHost, P.s), "D2H factors");
    hipchk(hipStreamSynchronize(P.s), "sync");
    auto at = [&](size_t i) {
      return P.sh.fp64 ? reinterpret_cast<const double*>(f.data())[i] : (double)reinterpret_cast<const float*>(f.data())[i];
    };
    const uint32_t n = P.sh.nw;
    const uint32_t s_at = n * ((n + 3) & ~3u) + n * P.sh.m;  // after the factor lines
    for (uint32_t k = 0; k < P.sh.p_sel; ++k) {
      std::vector<double> v(n);
      for (uint32_t j = 0; j < n; ++j) v[j] = at((size_t)k * P.sh.fstride + s_at + j);
      std::sort(v.begin(), v.end(), [](double a, double b) { return a > b; });
      std::memcpy(s + (size_t)k * n, v.data(), n * sizeof(double));
    }
    *n_patches = P.sh.p_sel;
    return 0;
  });
}

int pgsvt_noise_estimate(const double* u, uint32_t n_rows, uint32_t n_cols, uint32_t n_t, uint32_t method,
                         double* alpha, double* mu, double* sigma) {
  if (!u || !alpha || !mu || !sigma) return fail(PGSVT_E_ARG, "null pointer");
  if (n_rows != n_cols) return fail(PGSVT_E_ARG, "quadtree noise estimation requires square frames");
  return guarded([&] {
    estimate_noise(u, n_rows, n_t, method, *alpha, *mu, *sigma);
    return 0;
  });
}

int pgsvt_frame_leaves(const double* frame, uint32_t n, int on_gpu, uint32_t cap, double* mean, double* var,
                       uint32_t* count) {
  if (!frame || !count || (cap > 0 && (!mean || !var))) return fail(PGSVT_E_ARG, "null pointer");
  if (n < 1) return fail(PGSVT_E_ARG, "empty frame");
  return guarded([&] {
    FrameLeaves fl;
    if (on_gpu) {
      if (n < 16) return fail(PGSVT_E_ARG, "leaf statistics on the GPU need frames of at least 16 x 16");
      if (int rc = check_device()) return rc;
      hipStream_t s = nullptr;
      hipchk(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
      struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() {
          (void)hipStreamSynchronize(s);
          (void)hipStreamDestroy(s);
        }
      } guard{s};
      DBuf<double> d;
      d.alloc((size_t)n * n);
      hipchk(hipMemcpyAsync(d.p, frame, (size_t)n * n * 8, hipMemcpyHostToDevice, s), "H2D frame");
      fl = std::move(gpu_frame_leaves(d.p, n, {0u}, 1.0, s)[0]);
    } else {
      fl = frame_leaves(frame, n);
    }
    *count = (uint32_t)fl.mean.size();
    if (fl.mean.size() > cap) return fail(PGSVT_E_ARG, "the frame has " + std::to_string(fl.mean.size()) +
                                                            " leaves, more than cap");
    std::copy(fl.mean.begin(), fl.mean.end(), mean);
    std::copy(fl.var.begin(), fl.var.end(), var);
    return 0;
  });
}

int pgsvt_hotpixel_filter(uint16_t* seq, uint32_t n_rows, uint32_t n_cols, uint32_t n_frames, double threshold) {
  return guarded([&] {
    if (!seq) return fail(PGSVT_E_ARG, "null sequence");
    if (n_rows < 1 || n_cols < 1 || n_frames < 1) return fail(PGSVT_E_ARG, "empty sequence");
    if (!(threshold >= 0.0)) return fail(PGSVT_E_ARG, "hot-pixel threshold must be >= 0");
    if (int rc = check_device()) return rc;
    hipStream_t s = nullptr;
    hipchk(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
    struct StreamGuard {
      hipStream_t s;
      ~StreamGuard() {
        (void)hipStreamSynchronize(s);
        (void)hipStreamDestroy(s);
      }
    } guard{s};
    const size_t n = (size_t)n_rows * n_cols * n_frames;
    DBuf<uint16_t> d;
    DBuf<uint32_t> cm;
    d.alloc(n);
    cm.alloc((size_t)n_frames * n_cols);
    hipchk(hipMemcpyAsync(d.p, seq, n * 2, hipMemcpyHostToDevice, s), "H2D seq");
    hipchk(launch_hotpixel(d.p, n_rows, n_cols, n_frames, threshold, cm.p, s), "hotpixel");
    hipchk(hipMemcpyAsync(seq, d.p, n * 2, hipMemcpyDeviceToHost, s), "D2H seq");
    hipchk(hipStreamSynchronize(s), "sync");
    return 0;
  });
}

int pgsvt_tiff_info(const char* path, uint32_t* width, uint32_t* height, uint32_t* bits, uint32_t* pages) {
  return guarded([&] {
    if (!path || !width || !height || !bits || !pages) return fail(PGSVT_E_ARG, "null argument");
    try {
      const TiffInfo i = tiff_info(path);
      *width = i.width;
      *height = i.height;
      *bits = i.bits;
      *pages = i.pages;
    } catch (const std::runtime_error& e) {  // unreadable or unsupported file
      return fail(PGSVT_E_ARG, e.what());
    }
    return 0;
  });
}

int pgsvt_tiff_read_u16(const char* path, uint32_t first, uint32_t count, uint16_t* out) {
  return guarded([&] {
    if (!path || !out) return fail(PGSVT_E_ARG, "null argument");
    try {
      tiff_read_u16(path, first, count, out);
    } catch (const std::runtime_error& e) {
      return fail(PGSVT_E_ARG, e.what());
    }
    return 0;
  });
}

int pgsvt_tiff_write_u16(const char* path, const uint16_t* cube, uint32_t width, uint32_t height, uint32_t pages) {
  return guarded([&] {
    if (!path || !cube) return fail(PGSVT_E_ARG, "null argument");
    try {
      tiff_write_u16(path, cube, width, height, pages);
    } catch (const std::runtime_error& e) {
      return fail(PGSVT_E_ARG, e.what());
    }
    return 0;
  });
}

int pgsvt_mlp_forward(const float* params, uint32_t n_in, uint32_t hidden, const float* X, uint32_t n, float* out) {
  return guarded([&] {
    if (!params || !X || !out || n_in == 0 || hidden == 0) return fail(PGSVT_E_ARG, "bad network arguments");
    if (int rc = check_device()) return rc;
    StreamOwner so;
    Mlp net(n_in, hidden, so.s);
    hipchk(hipMemcpyAsync(net.params(), params, net.n_params() * 4, hipMemcpyHostToDevice, so.s), "H2D params");
    DBuf<float> x, y;
    x.alloc((size_t)n * n_in);
    y.alloc(n);
    hipchk(hipMemcpyAsync(x.p, X, x.bytes(), hipMemcpyHostToDevice, so.s), "H2D X");
    net.forward(x.p, n, y.p);
    hipchk(hipMemcpyAsync(out, y.p, y.bytes(), hipMemcpyDeviceToHost, so.s), "D2H y");
    hipchk(hipStreamSynchronize(so.s), "sync");
    return 0;
  });
}

int pgsvt_mlp_gradient(const float* params, uint32_t n_in, uint32_t hidden, const float* X, const float* t,
                       uint32_t n, float* grad, float* loss) {
  return guarded([&] {
    if (!params || !X || !t || !grad || !loss || n_in == 0 || hidden == 0 || n == 0)
      return fail(PGSVT_E_ARG, "bad network arguments");
    if (int rc = check_device()) return rc;
    StreamOwner so;
    Mlp net(n_in, hidden, so.s);
    hipchk(hipMemcpyAsync(net.params(), params, net.n_params() * 4, hipMemcpyHostToDevice, so.s), "H2D params");
    DBuf<float> x, y, tt, l;
    x.alloc((size_t)n * n_in);
    y.alloc(n);
    tt.alloc(n);
    l.alloc(1);
    hipchk(hipMemcpyAsync(x.p, X, x.bytes(), hipMemcpyHostToDevice, so.s), "H2D X");
    hipchk(hipMemcpyAsync(tt.p, t, tt.bytes(), hipMemcpyHostToDevice, so.s), "H2D t");
    net.forward(x.p, n, y.p);
    net.backward(x.p, y.p, tt.p, n, l.p);
    hipchk(hipMemcpyAsync(grad, net.grads(), net.n_params() * 4, hipMemcpyDeviceToHost, so.s), "D2H grad");
    hipchk(hipMemcpyAsync(loss, l.p, 4, hipMemcpyDeviceToHost, so.s), "D2H loss");
    hipchk(hipStreamSynchronize(so.s), "sync");
    return 0;
  });
}

int pgsvt_mlp_train(const float* X, const float* t, uint32_t n, uint32_t n_in, uint32_t hidden, uint32_t epochs,
                    uint32_t batch, float lr, uint64_t seed, float* params, float* epoch_loss) {
  return guarded([&] {
    if (!X || !t || !params || n == 0 || n_in == 0 || hidden == 0 || batch == 0 || !(lr > 0.f))
      return fail(PGSVT_E_ARG, "bad training arguments");
    if (int rc = check_device()) return rc;
    StreamOwner so;
    hipStream_t s = so.s;
    Mlp net(n_in, hidden, s);
    const size_t np = net.n_params();
    hipchk(hipMemcpyAsync(net.params(), params, np * 4, hipMemcpyHostToDevice, s), "H2D params");
    DBuf<float> x, tt, xb, tb, y, m, v, l;
    DBuf<uint32_t> idx;
    x.alloc((size_t)n * n_in);
    tt.alloc(n);
    const uint32_t B = std::min(batch, n);
    xb.alloc((size_t)B * n_in);
    tb.alloc(B);
    y.alloc(B);
    idx.alloc(n);
    l.alloc(1);
    m.alloc(np);
    v.alloc(np);
    hipchk(hipMemcpyAsync(x.p, X, x.bytes(), hipMemcpyHostToDevice, s), "H2D X");
    hipchk(hipMemcpyAsync(tt.p, t, tt.bytes(), hipMemcpyHostToDevice, s), "H2D t");
    std::vector<float> zero(np, 0.f);
    hipchk(hipMemcpyAsync(m.p, zero.data(), np * 4, hipMemcpyHostToDevice, s), "H2D m");
    hipchk(hipMemcpyAsync(v.p, zero.data(), np * 4, hipMemcpyHostToDevice, s), "H2D v");
    std::mt19937_64 rng(seed);
    std::vector<uint32_t> order(n);
    for (uint32_t i = 0; i < n; ++i) order[i] = i;
    std::vector<float> losses;
    int step = 0;
    for (uint32_t ep = 0; ep < epochs; ++ep) {
      std::shuffle(order.begin(), order.end(), rng);
      hipchk(hipMemcpyAsync(idx.p, order.data(), n * 4, hipMemcpyHostToDevice, s), "H2D order");
      hipchk(hipStreamSynchronize(s), "sync order");  // `order` is reshuffled next epoch
      double sum = 0;
      uint32_t nbatch = 0;
      for (uint32_t b0 = 0; b0 + B <= n; b0 += B) {  // full batches
        hipchk(launch_gather_rows(x.p, idx.p + b0, B, n_in, xb.p, s), "gather X");
        hipchk(launch_gather_rows(tt.p, idx.p + b0, B, 1, tb.p, s), "gather t");
        net.forward(xb.p, B, y.p);
        net.backward(xb.p, y.p, tb.p, B, l.p);
        hipchk(launch_adam(net.params(), net.grads(), m.p, v.p, (int64_t)np, lr, 0.9f, 0.999f, 1e-8f, ++step, s),
               "adam");
        float lh = 0;
        hipchk(hipMemcpyAsync(&lh, l.p, 4, hipMemcpyDeviceToHost, s), "D2H loss");
        hipchk(hipStreamSynchronize(s), "sync loss");
        sum += lh;
        ++nbatch;
      }
      if (epoch_loss) epoch_loss[ep] = (float)(sum / std::max(1u, nbatch));
    }
    hipchk(hipMemcpyAsync(params, net.params(), np * 4, hipMemcpyDeviceToHost, s), "D2H params");
    hipchk(hipStreamSynchronize(s), "sync");
    return 0;
  });
}

int pgsvt_delta2_signs(int64_t seed, uint64_t n_vox, uint8_t* out) {
  return guarded([&] {
    if (int rc = check_device()) return rc;
    DBuf<uint8_t> d;
    d.alloc(n_vox);
    hipchk(launch_delta2(seed, n_vox, d.p, nullptr), "delta2");
    hipchk(hipMemcpy(out, d.p, n_vox, hipMemcpyDeviceToHost), "D2H");
    return 0;
  });
}

}  // extern "C"
