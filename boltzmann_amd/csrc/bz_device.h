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

// Full 32x32 -> 64 product: one v_mad_u64_u32 on the device instead of a
// separate mul_hi / mul_lo pair.
__host__ __device__ __forceinline__ uint64_t mulwide32(uint32_t a, uint32_t b) {
  return (uint64_t)a * (uint64_t)b;
}

__host__ __device__ __forceinline__ U4 philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                              uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint64_t p0 = mulwide32(0xD2511F53u, c0);
    const uint64_t p1 = mulwide32(0xCD9E8D57u, c2);
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    c0 = n0;
    c1 = (uint32_t)p1;
    c2 = n2;
    c3 = (uint32_t)p0;
  }
  return U4{c0, c1, c2, c3};
}

// Stream purposes (top byte of counter word 3).
enum : uint32_t {
  kPurposeCollide = 0,
  kPurposeGas = 1,
  kPurposeResample = 2,
  kPurposeInit = 3,
  kPurposeFlight = 4,
  kPurposeDecide = 5
};

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

// The collision stream of particle i at a step, 5 uniforms (reference R
// columns 0,1,2,3,5).  u[0], the decision uniform compared with the
// collision probabilities, comes from a draw shared by the particles 1024
// slots apart: in FP32 the 4 particles i = 4096a + 1024b + c take word b of
// the draw of group 1024a + c (purpose kPurposeDecide), in FP64 the 2
// particles i = 2048a + 1024b + c take words 2b, 2b+1 of group 1024a + c.
// u[1..4], the scattering angles (the secondary's are u[3], u[4]), come from
// the particle's own draw, made only when it collides: FP32 one block, FP64
// a second block for the secondary's angles.
__host__ __device__ __forceinline__ uint32_t decision_group(uint32_t i, int per) {
  return per == 4 ? (((i >> 12) << 10) | (i & 1023u)) : (((i >> 11) << 10) | (i & 1023u));
}

// A kernel argument seed through an empty asm: the Philox round keys derived
// from it are then formed at each draw (scalar adds) instead of being hoisted
// out of a particle loop into scalar registers that spill to VGPR lanes.
__device__ __forceinline__ uint64_t opaque_seed(uint64_t s) {
  asm volatile("" : "+s"(s));
  return s;
}

template <typename T>
struct Decision;

// Keeps the last group's draw: a thread whose slots are 1024 apart draws
// once for 4 (FP32) or 2 (FP64) of them.
template <>
struct Decision<float> {
  uint32_t g = 0xFFFFFFFFu;
  U4 r;
  __device__ __forceinline__ float get(uint64_t seed, uint32_t i, uint64_t step) {
    const uint32_t gi = decision_group(i, 4);
    if (gi != g) {
      r = draw(opaque_seed(seed), kPurposeDecide, gi, 0, step);
      g = gi;
    }
    const uint32_t b = (i >> 10) & 3u;
    return u24(b == 0 ? r.a : (b == 1 ? r.b : (b == 2 ? r.c : r.d)));
  }
};
template <>
struct Decision<double> {
  uint32_t g = 0xFFFFFFFFu;
  U4 r;
  __device__ __forceinline__ double get(uint64_t seed, uint32_t i, uint64_t step) {
    const uint32_t gi = decision_group(i, 2);
    if (gi != g) {
      r = draw(opaque_seed(seed), kPurposeDecide, gi, 0, step);
      g = gi;
    }
    return ((i >> 10) & 1u) ? u53(r.c, r.d) : u53(r.a, r.b);
  }
};

template <typename T>
struct Uniforms;

template <>
struct Uniforms<double> {
  double u[5];
  __device__ __forceinline__ void scatter(uint64_t seed, uint32_t i, uint64_t step, bool ion) {
    const U4 r = draw(opaque_seed(seed), kPurposeCollide, i, 0, step);
    u[1] = u53(r.a, r.b);
    u[2] = u53(r.c, r.d);
    if (ion) {
      const U4 q = draw(opaque_seed(seed), kPurposeCollide, i, 1, step);
      u[3] = u53(q.a, q.b);
      u[4] = u53(q.c, q.d);
    }
  }
};

template <>
struct Uniforms<float> {
  float u[5];
  __device__ __forceinline__ void scatter(uint64_t seed, uint32_t i, uint64_t step, bool) {
    const U4 r = draw(opaque_seed(seed), kPurposeCollide, i, 0, step);
    u[1] = u24(r.a);
    u[2] = u24(r.b);
    u[3] = u24(r.c);
    u[4] = u24(r.d);
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

// Block-wide decoupled look-back (single-pass chained scan): every thread
// inspects LB predecessor status words per round (window BLOCK*LB), so a
// tile whose predecessors are all in flight at once sums their aggregates in
// one or two rounds instead of walking back one 64-tile window per memory
// round trip.  All threads of the block call it; returns the exclusive
// prefix and publishes the inclusive one.  s_min / s_red are LDS scratch
// (1 and BLOCK/64 words).  Tiles are processed by co-resident blocks in
// increasing order per block and publish their aggregates before looking
// back, so every wait below ends.
template <int BLOCK, int LB>
__device__ __forceinline__ uint32_t lookback_block(uint64_t* status, uint32_t t, uint32_t agg,
                                                   uint32_t epoch, uint32_t* s_min,
                                                   uint32_t* s_red) {
  constexpr int NW = BLOCK / kWave;
  constexpr int WIN = BLOCK * LB;
  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1), wid = tid / kWave;
  const uint64_t ep = (uint64_t)(epoch & 0x3FFFFFFFu) << 32;
  if (t == 0) {
    if (tid == 0) st_rel(&status[0], kFlagPre | ep | agg);
    return 0;
  }
  if (tid == 0) st_rel(&status[t], kFlagAgg | ep | agg);
  uint32_t excl = 0;
  int64_t j = (int64_t)t - 1;
  while (true) {
    uint64_t s[LB];
    uint32_t my_first = 0xFFFFFFFFu;
#pragma unroll
    for (int k = 0; k < LB; ++k) {
      const int off = tid + k * BLOCK;  // distance back from j
      const int64_t idx = j - off;
      s[k] = kFlagPre;  // virtual tile before 0: prefix 0
      if (idx >= 0) {
        s[k] = ld_acq(&status[idx]);
        while ((s[k] >> 62) == 0 || (s[k] & (0x3FFFFFFFull << 32)) != ep) {
          __builtin_amdgcn_s_sleep(2);  // back off instead of hammering the line
          s[k] = ld_acq(&status[idx]);
        }
      }
      if ((s[k] >> 62) == 2 && (uint32_t)off < my_first) my_first = (uint32_t)off;
    }
    if (tid == 0) *s_min = 0xFFFFFFFFu;
    __syncthreads();
    if (my_first != 0xFFFFFFFFu) atomicMin(s_min, my_first);
    __syncthreads();
    const uint32_t first = *s_min;  // nearest predecessor holding a prefix
    uint32_t part = 0;
#pragma unroll
    for (int k = 0; k < LB; ++k) {
      const uint32_t off = (uint32_t)(tid + k * BLOCK);
      if (off <= first) part += (uint32_t)s[k];
    }
    part = wave_sum(part);
    if (lane == 0) s_red[wid] = part;
    __syncthreads();
    uint32_t tot = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) tot += s_red[w];
    excl += tot;
    __syncthreads();
    if (first != 0xFFFFFFFFu) break;
    j -= WIN;
  }
  if (tid == 0) st_rel(&status[t], kFlagPre | ep | (uint32_t)(agg + excl));
  return excl;
}

// Binning key of one live particle: cell = int((x + dx/2)/dx) clamped to the
// nodes (1-D; 0 otherwise), energy bin = number of thresholds <= e, key =
// cell * nb + bin (the key of bz_bin_sort).
struct BinKeys {
  double dx;
  double thr[8];
  int nb;
  int n_nodes;
  int n_keys;  // the key of a dead slot
  int dim;
  int ew;      // 1: bin the energy after (e*w)/w (BZ_REF_ENERGY_ROUNDTRIP)
};

// A wave-uniform FP64 constant held in a VGPR pair.  A kernel whose uniform
// arguments outgrow the scalar register file has them spilled to VGPR lanes
// and read back (v_readlane plus hazard wait states) inside its loops; the
// constants the per-particle path reads most are kept in VGPRs instead.
__device__ __forceinline__ double in_vgpr(double v) {
  double r;
  asm volatile("v_mov_b64 %0, %1" : "=v"(r) : "s"(v));
  return r;
}

// int(x / dx) for x >= 0, equal to the FP64 quotient's truncation, without
// the divide: the product with the reciprocal is within a few ulps of the
// quotient, so away from a cell edge (more than 1e-12 relative) it truncates
// to the same integer; near an edge, and for x < 0, the quotient itself.
__device__ __forceinline__ int cell_of(double x, double dx, double inv_dx) {
  const double q = x * inv_dx;
  const double t = trunc(q);
  const double margin = 1e-12 * q;
  if (q - t > margin && t + 1.0 - q > margin) return (int)t;
  // rare: an empty asm keeps the compiler from evaluating the divide for
  // every particle and selecting
  asm volatile("" ::: "memory");
  return (int)(x / dx);
}

// The energy thresholds of a binning in the compute type: from the kernel
// arguments, or a copy staged in LDS (BinThrLds) by a kernel that bins
// every particle it touches.
template <typename T>
struct BinThrArgs {
  const BinKeys& k;
  __device__ __forceinline__ T operator[](int q) const { return (T)k.thr[q]; }
};
template <typename T>
struct BinThrLds {
  __attribute__((address_space(3))) const T* t;
  __device__ __forceinline__ T operator[](int q) const { return t[q]; }
};

template <typename T, typename TH>
__device__ __forceinline__ uint32_t bin_key(const BinKeys& k, const TH& thr, T x, T e, T w) {
  if (k.ew) e = (e * w) / w;
  int cell = 0;
  if (k.dim == 1) {
    cell = cell_of((double)x + 0.5 * k.dx, k.dx, 1.0 / k.dx);
    cell = cell < 0 ? 0 : (cell >= k.n_nodes ? k.n_nodes - 1 : cell);
  }
  // thresholds past the binning's nb - 1 are +inf (an infinite energy,
  // which passes them, is kept in the last bin)
  int b = 0;
#pragma unroll
  for (int q = 0; q < 7; ++q) b += (e >= thr[q]) ? 1 : 0;
  b = b < k.nb - 1 ? b : k.nb - 1;
  return (uint32_t)(cell * k.nb + b);
}

template <typename T>
__device__ __forceinline__ uint32_t bin_key(const BinKeys& k, T x, T e, T w) {
  return bin_key<T>(k, BinThrArgs<T>{k}, x, e, w);
}

// Stable compaction blocks: block b of the grid owns a contiguous run of
// whole tiles of kCompTile slots out of n.
constexpr int kCompBlock = 256;
constexpr int kCompItems = 4;
constexpr int kCompTile = kCompBlock * kCompItems;

__device__ __forceinline__ void comp_range(int64_t n, int64_t* lo, int64_t* hi) {
  const int64_t tiles = (n + kCompTile - 1) / kCompTile;
  const int64_t per = (tiles + gridDim.x - 1) / gridDim.x;
  int64_t a = (int64_t)blockIdx.x * per * kCompTile;
  int64_t b = a + per * kCompTile;
  *lo = a < n ? a : n;
  *hi = b < n ? b : n;
}

// Device clock kept inside the context.
struct Clock {
  int64_t n;         // live particles (current state buffer)
  int64_t n_next;    // written by the producing kernel, committed by the next
  int64_t n_prev;    // slot count an in-place push started from (its births go after)
  uint64_t step;     // keys the random stream
  double t;          // simulated time, s
  uint32_t epoch;    // look-back epoch, bumped per scan launch
  uint32_t ticket;   // dynamic tile counter
  uint32_t resample; // resampling pass requested (set by the check)
  uint32_t pad;
  unsigned long long cnt[BZ_NCOUNTERS];
  // bits of an upper bound of every weight (a positive double's bits order
  // like the double): set from the state after a clock set, raised by every
  // merge; the deposit's fixed-point scale follows from it
  unsigned long long w_max_bits;
};

// ---------------------------------------------------------------------------
// Deterministic node moments of the 1-D deposit.  Each contribution v >= 0
// (a weight, or a weight times an energy) is scaled by 2^S, S = 48 minus the
// binary exponent of the weight bound, truncated to an integer and added to
// a 64-bit word; a wrap of that word (the returned old value says so) adds a
// carry to a 32-bit carry word.  Integer sums do not depend on the order the
// adds arrive in, so the node sums are the same on every run.  The push
// keeps the weight and energy sums (moments 0 and 2) per block in LDS, 64-bit
// words [2][nodes] and one carry word per node holding both moments' carries
// in 16-bit fields; the ionization sum (moment 1, added only where a particle
// ionizes) goes to global 64-bit words and carry words shared by all blocks.
// Resolution 2^-48 of the weight bound per contribution; range 2^80 units
// (2^32 bound-weights or bound-weight-eV) per node and block in LDS, 2^96 in
// the global words.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double dep_scale_of(unsigned long long wmax_bits) {
  if (!wmax_bits) return 1.0;
  int e;
  frexp(__longlong_as_double((long long)wmax_bits), &e);
  return ldexp(1.0, 48 - e);
}

// v * scale truncated, as a 64-bit word and the carries above it.  Below
// 2^64 the two 32-bit halves are converted separately: h = floor(t / 2^32)
// and t - h * 2^32 are exact, and the low half truncates to the same word
// as t itself would.
__device__ __forceinline__ unsigned long long dep_fixed_split(double v, double scale,
                                                              uint32_t* qh) {
  const double t = v * scale;  // v >= 0; scale is a power of two
  if (t < 18446744073709551616.0) {
    *qh = 0;
    const double h = floor(t * 2.3283064365386962890625e-10);
    const double l = fma(-h, 4294967296.0, t);
    return ((unsigned long long)(uint32_t)h << 32) | (unsigned long long)(uint32_t)l;
  }
  // rare: an empty asm keeps the compiler from evaluating this branch for
  // every contribution and selecting
  asm volatile("" ::: "memory");
  const double h = floor(t * 5.421010862427522170e-20);  // beyond 64 bits
  *qh = (uint32_t)h;
  return (unsigned long long)(t - h * 18446744073709551616.0);
}

// LDS moments: slot 0 (weight) or 1 (energy) of [2][nn] words, 16-bit carry
// fields of the node's carry word
__device__ __forceinline__ void dep_fixed_add(unsigned long long* lo, uint32_t* hi, int slot,
                                              int nn, int node, double v, double scale) {
  uint32_t qh;
  const unsigned long long q = dep_fixed_split(v, scale, &qh);
  const unsigned long long old = atomicAdd(&lo[slot * nn + node], q);
  const uint32_t carry = qh + (old + q < old ? 1u : 0u);
  if (carry) atomicAdd(&hi[node], carry << (16 * slot));
}

__device__ __forceinline__ double dep_fixed_value(unsigned long long lo, uint32_t carries,
                                                  double inv_scale) {
  return ((double)carries * 18446744073709551616.0 + (double)lo) * inv_scale;
}

__device__ __forceinline__ uint32_t dep_carries(uint32_t hi, int slot) {
  return (hi >> (16 * slot)) & 0xFFFFu;
}

// a carry field past half its range: the block's sums are close to wrapping
__device__ __forceinline__ bool dep_near_range(uint32_t hi) {
  return ((hi >> 15) & 1u) || ((hi >> 31) & 1u);
}

// global moment 1: the node's 64-bit word and its own 32-bit carry word
__device__ __forceinline__ void dep_fixed_add_global(unsigned long long* lo, uint32_t* hi,
                                                     int node, double v, double scale) {
  uint32_t qh;
  const unsigned long long q = dep_fixed_split(v, scale, &qh);
  const unsigned long long old = atomicAdd(&lo[node], q);
  const uint32_t carry = qh + (old + q < old ? 1u : 0u);
  if (carry) atomicAdd(&hi[node], carry);
}

}  // namespace bz
