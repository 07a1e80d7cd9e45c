// common.hpp -- definitions shared by the host orchestrator and the gfx950
// kernels of the PGURE-SVT path.
#pragma once

#include <cstdint>

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define PG_HD __host__ __device__
#else
#define PG_HD
#endif

namespace pgsvt {

// Fixed-point scale of the overlap-add accumulators: contributions are
// rounded to multiples of 2^-32 and summed in int64, so the sum does not
// depend on the order the atomics land in (deterministic overlap-add).
constexpr float kFixScale = 4294967296.0f;   // 2^32
constexpr double kFixInv = 1.0 / 4294967296.0;

// Variants of the window decomposed for PGURE: the data U and the two
// perturbed copies U +/- eps2*delta2 (src/pgure.hpp:80-91).  The fourth copy
// U + eps1*delta1 of the reference equals U exactly (delta1 is an integer
// cube, so the scalar eps1 = 1e-4 is truncated to 0 before the product,
// src/pgure.hpp:80 with delta1 declared at :156) and is not decomposed.
constexpr int kVarU = 0, kVarP = 1, kVarM = 2, kMaxVariants = 3;

// Per-window description uploaded for every batch.
struct WinMeta {
  uint32_t t;      // output frame index
  uint32_t s0;     // first input frame of the window
  uint32_t ref;    // output frame inside the window (src/pguresvt.hpp:155-166)
  uint32_t aref;   // ARPS reference frame inside the window (src/arps.hpp:55-118)
  uint32_t out;    // slot of the output frame in the caller's Y
  double uMax;     // max of the raw window (src/pguresvt.hpp:116)
  double wMax;     // max of the motion-estimation window (:117)
  double lambda;   // threshold of the current evaluation
};

// ARPS pass: block matching from frame `src` to frame `dst` of the window,
// motion slot `slot` (src/arps.hpp:153).  Passes of one group run in order in
// one thread, so a pass can read the motion an earlier pass left in its slot.
struct ArpsPass {
  int16_t src, dst, slot, group;
};

// pcg64 = setseq_xsl_rr_128_64 with the default stream
// (src/pcg/pcg_random.hpp:166-169, 445-451, 501-503, 1085-1113).
struct U128 {
  uint64_t lo, hi;
};
PG_HD inline U128 u128_mul(U128 a, U128 b) {
  U128 r;
#ifdef __HIP_DEVICE_COMPILE__
  r.lo = a.lo * b.lo;
  r.hi = __umul64hi(a.lo, b.lo) + a.lo * b.hi + a.hi * b.lo;
#else
  unsigned __int128 p = (unsigned __int128)a.lo * b.lo;
  r.lo = (uint64_t)p;
  r.hi = (uint64_t)(p >> 64) + a.lo * b.hi + a.hi * b.lo;
#endif
  return r;
}
PG_HD inline U128 u128_add(U128 a, U128 b) {
  U128 r;
  r.lo = a.lo + b.lo;
  r.hi = a.hi + b.hi + (r.lo < a.lo ? 1u : 0u);
  return r;
}
PG_HD inline U128 pcg_mult() { return U128{4865540595714422341ULL, 2549297995355413924ULL}; }
PG_HD inline U128 pcg_inc() { return U128{1442695040888963407ULL, 6364136223846793005ULL}; }

// state after seed(s): bump(s + inc)  (pcg_random.hpp:501-503)
PG_HD inline U128 pcg_seed(uint64_t s) {
  return u128_add(u128_mul(u128_add(U128{s, 0}, pcg_inc()), pcg_mult()), pcg_inc());
}
// state advanced by `delta` steps (Brown's jump-ahead for an LCG)
PG_HD inline U128 pcg_advance(U128 state, uint64_t delta) {
  U128 acc_mult{1, 0}, acc_plus{0, 0}, cur_mult = pcg_mult(), cur_plus = pcg_inc();
  while (delta > 0) {
    if (delta & 1) {
      acc_mult = u128_mul(acc_mult, cur_mult);
      acc_plus = u128_add(u128_mul(acc_plus, cur_mult), cur_plus);
    }
    cur_plus = u128_mul(u128_add(cur_mult, U128{1, 0}), cur_plus);
    cur_mult = u128_mul(cur_mult, cur_mult);
    delta >>= 1;
  }
  return u128_add(u128_mul(acc_mult, state), acc_plus);
}
PG_HD inline U128 pcg_step(U128 s) { return u128_add(u128_mul(s, pcg_mult()), pcg_inc()); }
PG_HD inline uint64_t pcg_output(U128 s) {
  const unsigned rot = (unsigned)(s.hi >> 58) & 63u;
  const uint64_t x = s.hi ^ s.lo;
  return rot ? (x >> rot) | (x << (64 - rot)) : x;
}
// libstdc++ bernoulli_distribution(p) on one 64-bit draw: generate_canonical
// gives double(draw) * 2^-64, clamped below 1 (bits/random.tcc:3345-3380).
PG_HD inline bool bernoulli_draw(uint64_t draw, double p) {
  double u = (double)draw * 5.421010862427522170037264e-20;  // 2^-64
  if (u >= 1.0) u = 0.99999999999999988898;                  // nextafter(1, 0)
  return u < p;
}

}  // namespace pgsvt
