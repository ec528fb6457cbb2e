// Device-side building blocks shared by the libbz kernels (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/bz.h"

namespace bz {

constexpr int kWave = 64;

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC'11): 10 rounds of two 32x32->64 multiplies with the
// round keys bumped by the Weyl constants between rounds.
// ---------------------------------------------------------------------------
struct U4 {
  uint32_t a, b, c, d;
};

__host__ __device__ __forceinline__ uint32_t mulhi32(uint32_t a, uint32_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __umulhi(a, b);
#else
  return (uint32_t)(((uint64_t)a * (uint64_t)b) >> 32);
#endif
}

__host__ __device__ __forceinline__ U4 philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                              uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint32_t hi0 = mulhi32(0xD2511F53u, c0), lo0 = 0xD2511F53u * c0;
    const uint32_t hi1 = mulhi32(0xCD9E8D57u, c2), lo1 = 0xCD9E8D57u * c2;
    const uint32_t n0 = hi1 ^ c1 ^ k0;
    const uint32_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
  }
  return U4{c0, c1, c2, c3};
}

// Stream purposes (top byte of counter word 3).
enum : uint32_t { kPurposeCollide = 0, kPurposeGas = 1, kPurposeResample = 2, kPurposeInit = 3 };

// counter = {index, block, step_lo, (step_hi << 8) | purpose}, key = seed.
__host__ __device__ __forceinline__ U4 draw(uint64_t seed, uint32_t purpose, uint32_t index,
                                            uint32_t block, uint64_t step) {
  return philox(index, block, (uint32_t)step, ((uint32_t)(step >> 32) << 8) | purpose,
                (uint32_t)seed, (uint32_t)(seed >> 32));
}

// Exact conversions to [0,1): 53 bits from two words, 24 bits from one.
__host__ __device__ __forceinline__ double u53(uint32_t a, uint32_t b) {
  return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}
__host__ __device__ __forceinline__ float u24(uint32_t a) {
  return (float)(a >> 8) * (1.0f / 16777216.0f);
}

// Uniform k (0..4 -> reference R columns 0,1,2,3,5) of the collision stream.
template <typename T>
struct Uniforms;

template <>
struct Uniforms<double> {
  double u[5];
  __device__ __forceinline__ void first(uint64_t seed, uint32_t i, uint64_t step) {
    const U4 r = draw(seed, kPurposeCollide, i, 0, step);
    u[0] = u53(r.a, r.b);
    u[1] = u53(r.c, r.d);
  }
  __device__ __forceinline__ void second(uint64_t seed, uint32_t i, uint64_t step) {
    const U4 r = draw(seed, kPurposeCollide, i, 1, step);
    u[2] = u53(r.a, r.b);
    u[3] = u53(r.c, r.d);
  }
  __device__ __forceinline__ void third(uint64_t seed, uint32_t i, uint64_t step) {
    const U4 r = draw(seed, kPurposeCollide, i, 2, step);
    u[4] = u53(r.a, r.b);
  }
};

template <>
struct Uniforms<float> {
  float u[5];
  __device__ __forceinline__ void first(uint64_t seed, uint32_t i, uint64_t step) {
    const U4 r = draw(seed, kPurposeCollide, i, 0, step);
    u[0] = u24(r.a);
    u[1] = u24(r.b);
    u[2] = u24(r.c);
    u[3] = u24(r.d);
  }
  __device__ __forceinline__ void second(uint64_t, uint32_t, uint64_t) {}
  __device__ __forceinline__ void third(uint64_t seed, uint32_t i, uint64_t step) {
    const U4 r = draw(seed, kPurposeCollide, i, 1, step);
    u[4] = u24(r.a);
  }
};

// ---------------------------------------------------------------------------
// Wave / block helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ int lane_id() { return __lane_id(); }

template <typename V>
__device__ __forceinline__ V wave_sum(V v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

template <typename V>
__device__ __forceinline__ V wave_incl_scan(V v) {
  const int lane = lane_id();
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const V t = __shfl_up(v, off, kWave);
    if (lane >= off) v += t;
  }
  return v;
}

// Agent-scope relaxed atomics on 64-bit status words (tile look-back).
__device__ __forceinline__ uint64_t ld_acq(const uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_rel(uint64_t* p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Tile status word: [63:62] flag (1 aggregate, 2 inclusive prefix),
// [61:32] launch epoch, [31:0] value.
constexpr uint64_t kFlagAgg = 1ull << 62;
constexpr uint64_t kFlagPre = 2ull << 62;

// Decoupled look-back (single-pass chained scan).  Called by all 64 lanes of
// one wave; returns the exclusive prefix of tile t and publishes its
// inclusive prefix.  A tile's predecessors were claimed (by ticket) before
// it, by blocks that are running, and they publish their aggregates without
// waiting on anyone, so every spin below ends.
__device__ __forceinline__ uint32_t lookback(uint64_t* status, uint32_t t, uint32_t agg,
                                             uint32_t epoch) {
  const int lane = lane_id();
  const uint64_t ep = (uint64_t)(epoch & 0x3FFFFFFFu) << 32;
  if (t == 0) {
    if (lane == 0) st_rel(&status[0], kFlagPre | ep | agg);
    return 0;
  }
  if (lane == 0) st_rel(&status[t], kFlagAgg | ep | agg);
  uint32_t excl = 0;
  int64_t j = (int64_t)t - 1;
  while (true) {
    const int64_t idx = j - lane;
    uint64_t s = kFlagPre;  // virtual tile before 0: prefix 0
    if (idx >= 0) {
      do {
        s = ld_acq(&status[idx]);
      } while ((s >> 62) == 0 || (s & (0x3FFFFFFFull << 32)) != ep);
    }
    const uint64_t pre = __ballot((s >> 62) == 2);
    if (pre) {
      const int first = __ffsll((unsigned long long)pre) - 1;
      excl += wave_sum((uint32_t)(lane <= first ? (uint32_t)s : 0u));
      break;
    }
    excl += wave_sum((uint32_t)s);
    j -= kWave;
  }
  if (lane == 0) st_rel(&status[t], kFlagPre | ep | (uint32_t)(agg + excl));
  return excl;
}

// Device clock kept inside the context.
struct Clock {
  int64_t n;         // live particles (current state buffer)
  int64_t n_next;    // written by the producing kernel, committed by the next
  uint64_t step;     // keys the random stream
  double t;          // simulated time, s
  uint32_t epoch;    // look-back epoch, bumped per scan launch
  uint32_t ticket;   // dynamic tile counter
  uint32_t resample; // resampling pass requested (set by the check)
  uint32_t pad;
  unsigned long long cnt[BZ_NCOUNTERS];
};

}  // namespace bz
