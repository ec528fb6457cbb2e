// Energy/cell binning, stable LSD radix sort by (cell, energy bin), and
// ES-WPNC merge / split resampling, for gfx950.
//
// Binning (bin_particles.py of the 1/3/5-bin strategies) builds per-cell,
// per-bin particle lists with atomic slots, so the list order depends on
// scheduling.  Here every particle gets key = cell*nb + bin and the state is
// ordered by a stable sort, so a segment lists its particles in ascending
// index order on every run.
//
// Resampling (boltzmann.py _do_resampling) selects particles by searching a
// cumulative weight sum.  The sum is formed in 64-bit fixed point with one
// power-of-two scale per segment, so the prefix sums are exact integers and
// the selection does not depend on the summation order.
#include "bz_launch.h"

namespace bz {

constexpr int kSortBlock = 256;
constexpr int kSortItems = 16;
constexpr int kSortTile = kSortBlock * kSortItems;  // 4096
constexpr int kScanChunk = 4096;

struct SortWs {
  uint32_t* keys_a;
  uint32_t* keys_b;
  uint32_t* vals_b;
  uint32_t* hist;       // [256][tiles]
  uint32_t* chunks;     // chunk sums of the hist scan
  int64_t* cdf;         // resample: fixed-point prefix per sorted position
  int32_t* copies;      // resample: split copies per sorted position
  int32_t* plan;        // resample: [kind, new_count, new_off, hw_count] per key
  uint32_t* sel;        // resample: selected sorted position per new particle
  // resample: [n_split, n_merge], then the split keys at [2, 2 + kstride)
  // and the merge keys at [2 + kstride, 2 + 2 kstride)
  int32_t* klist;
  int32_t kstride;
  int64_t max_tiles;
  int64_t max_chunks;
  // in-place resampling: members of the split / merge segments, grouped by a
  // sort of their own (keys and slots, two buffers each)
  uint32_t* mkeys;
  uint32_t* mslot;
  uint32_t* mkeys2;
  uint32_t* mslot2;
  int32_t* ip_counts;   // [max_keys+1] slots per key (key-order lists)
  int32_t* ip_hcnt;     // [max_keys+1] slots per key of an in-place pass; zero between passes
  int32_t* ip_moff;     // [max_keys+2] member offsets per key
  int32_t* ip_mcnt;     // [max_keys+1] members per key
  int32_t* ip_cand;     // [max_keys+1] 0 untouched, 1 split candidate, 2 merge; [max_keys]: fire
  uint32_t* ip_bcnt;    // members per extraction wave, then their exclusive prefixes
  Clock* ip_mclk;       // member count (the radix kernels read it as a clock)
  long long* ip_chunk;  // per 1024-key chunk: totals and prefixes of the plan scans
  int32_t* ip_tmp;      // [2][max_keys+1] chunk-local split / merge list positions
  uint32_t* ip_bits;    // [max_keys/32 + 1] candidate keys as a bit mask
};

// member extraction: waves of 4-wave blocks, each wave a contiguous slot range
constexpr int kIpBlocks = 2048;
constexpr int kIpWaves = kIpBlocks * 4;
// plan scans: 1024 keys per block (n_keys <= 65534: at most 64 chunks); per
// chunk 8 words: member total, appended, splits, merges and their prefixes;
// then the fire word and the slot count the pass's appends start from
constexpr int kKeyChunk = 1024;
constexpr int kMaxKeyChunks = 64;
constexpr int kIpChunkWords = 8 * kMaxKeyChunks + 8;

static int64_t tiles_for(int64_t cap) { return cap / kSortTile + 2; }

static SortWs carve(void* base, int64_t cap, int32_t max_keys) {
  SortWs w;
  char* p = (char*)base;
  auto take = [&](size_t bytes) {
    char* r = p;
    p += (bytes + 255) & ~(size_t)255;
    return r;
  };
  w.max_tiles = tiles_for(cap);
  w.max_chunks = (256 * w.max_tiles) / kScanChunk + 2;
  w.keys_a = (uint32_t*)take(4 * cap);
  w.keys_b = (uint32_t*)take(4 * cap);
  w.vals_b = (uint32_t*)take(4 * cap);
  w.hist = (uint32_t*)take(4 * 256 * w.max_tiles);
  w.chunks = (uint32_t*)take(4 * (w.max_chunks + 1));
  w.cdf = (int64_t*)take(8 * cap);
  w.copies = (int32_t*)take(4 * cap);
  w.plan = (int32_t*)take(16 * (size_t)(max_keys + 1));
  w.sel = (uint32_t*)take(4 * cap);
  w.kstride = max_keys + 1;
  w.klist = (int32_t*)take(4 * (2 + 2 * (size_t)w.kstride));
  w.mkeys = (uint32_t*)take(4 * cap);
  w.mslot = (uint32_t*)take(4 * cap);
  w.mkeys2 = (uint32_t*)take(4 * cap);
  w.mslot2 = (uint32_t*)take(4 * cap);
  w.ip_counts = (int32_t*)take(4 * ((size_t)max_keys + 1));
  w.ip_hcnt = (int32_t*)take(4 * ((size_t)max_keys + 1));
  w.ip_moff = (int32_t*)take(4 * ((size_t)max_keys + 2));
  w.ip_mcnt = (int32_t*)take(4 * ((size_t)max_keys + 1));
  w.ip_cand = (int32_t*)take(4 * ((size_t)max_keys + 1));
  w.ip_bcnt = (uint32_t*)take(4 * (size_t)kIpWaves);
  w.ip_mclk = (Clock*)take(sizeof(Clock));
  w.ip_chunk = (long long*)take(8 * (size_t)kIpChunkWords);
  w.ip_tmp = (int32_t*)take(8 * ((size_t)max_keys + 1));
  w.ip_bits = (uint32_t*)take(4 * ((size_t)max_keys / 32 + 2));
  return w;
}

size_t sort_workspace_bytes(int64_t capacity, int32_t max_keys) {
  const int64_t cap = capacity > 0 ? capacity : 1;
  const int64_t tiles = tiles_for(cap);
  const int64_t chunks = (256 * tiles) / kScanChunk + 2;
  size_t b = 0;
  auto add = [&](size_t bytes) { b += (bytes + 255) & ~(size_t)255; };
  add(4 * cap);
  add(4 * cap);
  add(4 * cap);
  add(4 * 256 * tiles);
  add(4 * (chunks + 1));
  add(8 * cap);
  add(4 * cap);
  add(16 * (size_t)(max_keys + 1));
  add(4 * cap);
  add(4 * (2 + 2 * ((size_t)max_keys + 1)));
  for (int k = 0; k < 4; ++k) add(4 * cap);
  add(4 * ((size_t)max_keys + 1));
  add(4 * ((size_t)max_keys + 1));
  add(4 * ((size_t)max_keys + 2));
  add(4 * ((size_t)max_keys + 1));
  add(4 * ((size_t)max_keys + 1));
  add(4 * (size_t)kIpWaves);
  add(sizeof(Clock));
  add(8 * (size_t)kIpChunkWords);
  add(8 * ((size_t)max_keys + 1));
  add(4 * ((size_t)max_keys / 32 + 2));
  return b + 1024;
}

// ---------------------------------------------------------------------------
// keys + per-key counts
// ---------------------------------------------------------------------------
struct KeyArgs {
  const void* x;
  const void* vx;
  const void* vy;
  const void* vz;
  const void* w;
  double kev;
  BinKeys bk;  // live keys 0..n_keys-1; dead particles get key n_keys
};

// Key per slot.  Each thread takes kKeyItems slots, all loads issued
// before use.  The per-key offsets and counts are read off the sorted keys
// afterwards (k_key_bounds), so this pass needs neither atomics nor an LDS
// histogram.
constexpr int kKeyItems = 4;

// key = cell * nb + energy bin of one slot; n_keys for a dead slot (w <= 0)
template <typename T>
__device__ __forceinline__ uint32_t slot_key(const KeyArgs& a, T x, T vx, T vy, T vz, T w) {
  if (!(w > (T)0)) return (uint32_t)a.bk.n_keys;
  const T e = (T)a.kev * (vx * vx + vy * vy + vz * vz);
  return bin_key<T>(a.bk, x, e, w);
}

template <typename T>
__global__ __launch_bounds__(kSortBlock) void k_keys(KeyArgs a, const Clock* clk, uint32_t* keys) {
  const int64_t n = clk->n;
  const int lane = threadIdx.x & 63;
  const int64_t base = (int64_t)blockIdx.x * kSortBlock * kKeyItems + (threadIdx.x - lane) * kKeyItems;
  if (base >= n) return;
  const T* x = (const T*)a.x;
  const T* vx = (const T*)a.vx;
  const T* vy = (const T*)a.vy;
  const T* vz = (const T*)a.vz;
  const T* w = (const T*)a.w;
  T qx[kKeyItems], qa[kKeyItems], qb[kKeyItems], qc[kKeyItems], qw[kKeyItems];
#pragma unroll
  for (int k = 0; k < kKeyItems; ++k) {
    const int64_t i = base + k * kWave + lane;
    const bool v = i < n;
    qw[k] = v ? w[i] : (T)0;
    qx[k] = v && a.bk.dim == 1 ? x[i] : (T)0;
    qa[k] = v ? vx[i] : (T)0;
    qb[k] = v ? vy[i] : (T)0;
    qc[k] = v ? vz[i] : (T)0;
  }
#pragma unroll
  for (int k = 0; k < kKeyItems; ++k) {
    const int64_t i = base + k * kWave + lane;
    if (i < n) keys[i] = slot_key<T>(a, qx[k], qa[k], qb[k], qc[k], qw[k]);
  }
}

// Counts per key without sorting (the input of the resample trigger): the
// keys of k_keys, counted in an LDS histogram per block (lds_keys counters;
// 0: straight into the global counts) and added to counts with integer
// atomics (exact, so the result does not depend on the order).
template <typename T>
__global__ __launch_bounds__(256) void k_key_hist(KeyArgs a, const Clock* clk, int32_t* counts,
                                                  int lds_keys) {
  extern __shared__ uint32_t s_c[];
  const int64_t n = clk->n;
  for (int j = threadIdx.x; j < lds_keys; j += 256) s_c[j] = 0u;
  __syncthreads();
  const T* x = (const T*)a.x;
  const T* vx = (const T*)a.vx;
  const T* vy = (const T*)a.vy;
  const T* vz = (const T*)a.vz;
  const T* w = (const T*)a.w;
  const int64_t stride = (int64_t)gridDim.x * 256 * kKeyItems;
  for (int64_t base = (int64_t)blockIdx.x * 256 * kKeyItems + threadIdx.x; base < n;
       base += stride) {
    T q[kKeyItems][5];
#pragma unroll
    for (int k = 0; k < kKeyItems; ++k) {
      const int64_t i = base + k * 256;
      const bool v = i < n;
      q[k][0] = v && a.bk.dim == 1 ? x[i] : (T)0;
      q[k][1] = v ? vx[i] : (T)0;
      q[k][2] = v ? vy[i] : (T)0;
      q[k][3] = v ? vz[i] : (T)0;
      q[k][4] = v ? w[i] : (T)0;
    }
#pragma unroll
    for (int k = 0; k < kKeyItems; ++k) {
      if (base + k * 256 >= n) continue;
      const uint32_t key = slot_key<T>(a, q[k][0], q[k][1], q[k][2], q[k][3], q[k][4]);
      if (lds_keys) atomicAdd(&s_c[key], 1u);
      else atomicAdd(&counts[key], 1);
    }
  }
  __syncthreads();
  for (int j = threadIdx.x; j < lds_keys; j += 256)
    if (s_c[j]) atomicAdd(&counts[j], (int32_t)s_c[j]);
}

// Offsets of every key from the sorted keys in one pass: position s sits
// between keys kp = skeys[s-1] (or -1) and kc = skeys[s] (or m past the
// end), so s is the first position of every key in (kp, kc]; keys that do
// not occur get the start of the next one.  offsets[k] = lower bound of k,
// k = 0..m, with offsets[m] = n.  A wave covers 256 consecutive positions
// (4 per lane, loads issued together); the left neighbour comes from the
// next lane down.
constexpr int kBoundsTile = 1024;  // positions per 256-thread block

__global__ __launch_bounds__(256) void k_key_bounds(const uint32_t* __restrict__ skeys,
                                                    const Clock* clk, int m,
                                                    int32_t* offsets) {
  const int64_t n = clk->n;
  const int lane = threadIdx.x & 63;
  // positions 0..n (n closes the last keys), tiles of kBoundsTile walked by
  // the grid: the work follows the count on the device, not the capacity
  for (int64_t tb = (int64_t)blockIdx.x * kBoundsTile; tb <= n;
       tb += (int64_t)gridDim.x * kBoundsTile) {
  const int64_t base = tb + (threadIdx.x - lane) * 4;
  if (base > n) continue;  // whole wave past the end (wave-uniform)
  int kc[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int64_t s = base + k * 64 + lane;
    kc[k] = s < n ? (int)skeys[s] : m;
  }
  const int before = (lane == 0 && base > 0) ? (int)skeys[base - 1] : -1;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int up = __shfl_up(kc[k], 1, 64);
    const int last = __shfl(k > 0 ? kc[k > 0 ? k - 1 : 0] : 0, 63, 64);
    const int kp = lane > 0 ? up : (k > 0 ? last : before);
    const int64_t s = base + k * 64 + lane;
    if (s <= n) {
      const int hi = kc[k] < m ? kc[k] : m;
      for (int q = kp + 1; q <= hi; ++q) offsets[q] = (int32_t)s;
    }
  }
  }
}

// counts[k] = offsets[k+1] - offsets[k], k = 0..m-1; the sort's size is
// added to the sorted-slots counter
__global__ __launch_bounds__(256) void k_key_counts(const int32_t* __restrict__ offsets, int m,
                                                    int32_t* counts, Clock* clk) {
  if (blockIdx.x == 0 && threadIdx.x == 0)
    clk->cnt[BZ_CNT_SORTED] += (unsigned long long)clk->n;
  for (int k = blockIdx.x * 256 + threadIdx.x; k < m; k += gridDim.x * 256)
    counts[k] = offsets[k + 1] - offsets[k];
}

__global__ void k_zero_i32(int32_t* p, int64_t m) {
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < m;
       j += (int64_t)gridDim.x * blockDim.x)
    p[j] = 0;
}

// ---------------------------------------------------------------------------
// radix pass: upsweep histogram, scan, stable downsweep scatter
// ---------------------------------------------------------------------------
// tiles of kSortTile keys in the count the clock holds; the histogram of a
// pass is laid out [digit][tile] with that many tiles per digit
__device__ __forceinline__ int64_t sort_tiles(const Clock* clk) {
  return (clk->n + kSortTile - 1) / kSortTile;
}

__global__ __launch_bounds__(kSortBlock) void k_upsweep(const uint32_t* keys, const Clock* clk,
                                                        int shift, uint32_t* hist) {
  __shared__ uint32_t s_h[256];
  const int64_t n = clk->n;
  const int64_t tiles = sort_tiles(clk);
  for (int64_t t = blockIdx.x; t < tiles; t += gridDim.x) {
    __syncthreads();  // the previous tile's counts are written
    s_h[threadIdx.x] = 0;
    __syncthreads();
    // thread t counts the kSortItems consecutive keys from base + t*kSortItems
    // (the tile's histogram does not depend on the split) and adds one run of
    // equal digits at a time: nearly sorted keys cut the LDS atomics on a few
    // hot words by the run length
    const int64_t i0 = t * kSortTile + (int64_t)threadIdx.x * kSortItems;
    uint32_t kv[kSortItems];
    if (i0 + kSortItems <= n) {
      const uint4* k4 = (const uint4*)(keys + i0);
#pragma unroll
      for (int q = 0; q < kSortItems / 4; ++q) {
        const uint4 v = k4[q];
        kv[4 * q] = v.x;
        kv[4 * q + 1] = v.y;
        kv[4 * q + 2] = v.z;
        kv[4 * q + 3] = v.w;
      }
    } else {
#pragma unroll
      for (int r = 0; r < kSortItems; ++r) kv[r] = i0 + r < n ? keys[i0 + r] : 0u;
    }
    uint32_t cur = 0u, run = 0u;
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
      if (i0 + r < n) {
        const uint32_t d = (kv[r] >> shift) & 255u;
        if (run && d == cur) {
          ++run;
        } else {
          if (run) atomicAdd(&s_h[cur], run);
          cur = d;
          run = 1u;
        }
      }
    }
    if (run) atomicAdd(&s_h[cur], run);
    __syncthreads();
    hist[(int64_t)threadIdx.x * tiles + t] = s_h[threadIdx.x];
  }
}

// generic 3-phase exclusive scan of uint32 data[0..m)
// (m = 256 digits x the pass's tiles, from the clock's count)
__global__ __launch_bounds__(1024) void k_chunk_reduce(const uint32_t* data, const Clock* clk,
                                                       uint32_t* chunks) {
  __shared__ uint32_t s_w[16];
  const int64_t m = 256 * sort_tiles(clk);
  const int64_t base = (int64_t)blockIdx.x * kScanChunk;
  if (base >= m) return;  // block-uniform
  uint32_t s = 0;
  for (int k = threadIdx.x; k < kScanChunk; k += 1024) {
    const int64_t i = base + k;
    if (i < m) s += data[i];
  }
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int w = 0; w < 16; ++w) t += s_w[w];
    chunks[blockIdx.x] = t;
  }
}

__global__ __launch_bounds__(1024) void k_chunk_scan(uint32_t* chunks, const Clock* clk) {
  // sequential carry over 1024-wide block scans
  __shared__ uint32_t s_w[16];
  __shared__ uint32_t s_carry;
  const int64_t nch = (256 * sort_tiles(clk) + kScanChunk - 1) / kScanChunk;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  if (tid == 0) s_carry = 0;
  __syncthreads();
  for (int64_t base = 0; base < nch; base += 1024) {
    const int64_t i = base + tid;
    const uint32_t v = i < nch ? chunks[i] : 0u;
    const uint32_t inc = wave_incl_scan(v);
    if (lane == 63) s_w[wid] = inc;
    __syncthreads();
    if (wid == 0) {
      uint32_t t = lane < 16 ? s_w[lane] : 0u;
      t = wave_incl_scan(t);
      if (lane < 16) s_w[lane] = t;
    }
    __syncthreads();
    const uint32_t excl = s_carry + (wid ? s_w[wid - 1] : 0u) + inc - v;
    if (i < nch) chunks[i] = excl;
    __syncthreads();
    if (tid == 0) s_carry += s_w[15];
    __syncthreads();
  }
}

__global__ __launch_bounds__(1024) void k_chunk_apply(uint32_t* data, const Clock* clk,
                                                      const uint32_t* chunks) {
  __shared__ uint32_t s_w[16];
  __shared__ uint32_t s_carry;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int64_t m = 256 * sort_tiles(clk);
  const int64_t base = (int64_t)blockIdx.x * kScanChunk;
  if (base >= m) return;  // block-uniform
  if (tid == 0) s_carry = chunks[blockIdx.x];
  __syncthreads();
  for (int k = 0; k < kScanChunk; k += 1024) {
    const int64_t i = base + k + tid;
    const uint32_t v = i < m ? data[i] : 0u;
    const uint32_t inc = wave_incl_scan(v);
    if (lane == 63) s_w[wid] = inc;
    __syncthreads();
    if (wid == 0) {
      uint32_t t = lane < 16 ? s_w[lane] : 0u;
      t = wave_incl_scan(t);
      if (lane < 16) s_w[lane] = t;
    }
    __syncthreads();
    const uint32_t excl = s_carry + (wid ? s_w[wid - 1] : 0u) + inc - v;
    if (i < m) data[i] = excl;
    __syncthreads();
    if (tid == 0) s_carry += s_w[15];
    __syncthreads();
  }
}

// Stable scatter of one tile.  Wave w owns the tile's w-th quarter
// (kSortItems rows of 64 consecutive items), so the stable order is wave,
// then row, then lane.  Pass 1, no barrier: each wave ranks its rows with a
// wave-level match (8 ballots) against its own running digit counts in LDS
// and keeps every item's rank within the wave.  One barrier, then each digit's
// counts are turned into exclusive offsets across the waves; one more, and
// every item goes to tile base + wave offset + rank.
__global__ __launch_bounds__(kSortBlock) void k_downsweep(const uint32_t* keys_in,
                                                          const uint32_t* vals_in,
                                                          uint32_t* keys_out, uint32_t* vals_out,
                                                          const Clock* clk, int shift,
                                                          const uint32_t* hist,
                                                          int identity_vals) {
  constexpr int NW = kSortBlock / kWave;
  static_assert(kSortBlock == 256, "one digit per thread in the offset pass");
  __shared__ uint32_t s_base[256];
  __shared__ uint32_t s_wc[NW][256];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int64_t n = clk->n;
  const int64_t tiles = sort_tiles(clk);
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (int64_t t = blockIdx.x; t < tiles; t += gridDim.x) {
    __syncthreads();  // the previous tile is done with the LDS
    s_base[tid] = hist[(int64_t)tid * tiles + t];
#pragma unroll
    for (int j = 0; j < 256 / kWave; ++j) s_wc[wid][j * kWave + lane] = 0u;
    const int64_t w0 = t * kSortTile + (int64_t)wid * (kSortItems * kWave);
    uint32_t key[kSortItems], val[kSortItems], pre[kSortItems];
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
      const int64_t i = w0 + r * kWave + lane;
      const bool valid = i < n;
      key[r] = valid ? keys_in[i] : 0u;
      val[r] = valid ? (identity_vals ? (uint32_t)i : vals_in[i]) : 0u;
    }
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
      const bool valid = w0 + r * kWave + lane < n;
      const uint32_t d = (key[r] >> shift) & 255u;
      uint64_t m = __ballot(valid);
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        const uint64_t bb = __ballot((d >> b) & 1u);
        m &= ((d >> b) & 1u) ? bb : ~bb;
      }
      const uint32_t rank = (uint32_t)__popcll(m & lt);
      const uint32_t before = s_wc[wid][d];  // this wave's items of digit d in earlier rows
      pre[r] = before + rank;
      if (valid && rank == 0) s_wc[wid][d] = before + (uint32_t)__popcll(m);
    }
    __syncthreads();
    {
      uint32_t acc = 0u;
#pragma unroll
      for (int w = 0; w < NW; ++w) {
        const uint32_t c = s_wc[w][tid];
        s_wc[w][tid] = acc;
        acc += c;
      }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
      if (w0 + r * kWave + lane < n) {
        const uint32_t d = (key[r] >> shift) & 255u;
        const uint32_t off = s_base[d] + s_wc[wid][d] + pre[r];
        keys_out[off] = key[r];
        vals_out[off] = val[r];
      }
    }
  }
}

// One radix pass over the count in `clk`: histogram, scan, stable scatter.
// Grids are sized to what is resident and walk the tiles the count needs.
static void radix_pass(const LaunchEnv& env, const SortWs& ws, const Clock* clk,
                       const uint32_t* kin, const uint32_t* vin, uint32_t* kout, uint32_t* vout,
                       int shift, int identity) {
  const int64_t tg = ws.max_tiles < (int64_t)env.num_cu * 8 ? ws.max_tiles
                                                            : (int64_t)env.num_cu * 8;
  hipLaunchKernelGGL(k_upsweep, dim3((unsigned)tg), dim3(kSortBlock), 0, env.stream, kin, clk,
                     shift, ws.hist);
  hipLaunchKernelGGL(k_chunk_reduce, dim3((unsigned)ws.max_chunks), dim3(1024), 0, env.stream,
                     (const uint32_t*)ws.hist, clk, ws.chunks);
  hipLaunchKernelGGL(k_chunk_scan, dim3(1), dim3(1024), 0, env.stream, ws.chunks, clk);
  hipLaunchKernelGGL(k_chunk_apply, dim3((unsigned)ws.max_chunks), dim3(1024), 0, env.stream,
                     ws.hist, clk, (const uint32_t*)ws.chunks);
  timing_begin(env, 4);
  hipLaunchKernelGGL(k_downsweep, dim3((unsigned)tg), dim3(kSortBlock), 0, env.stream, kin, vin,
                     kout, vout, clk, shift, (const uint32_t*)ws.hist, identity);
  timing_end(env, 4);
}

static void key_bounds(const LaunchEnv& env, const Clock* clk, const uint32_t* skeys, int mk,
                       int32_t* offsets) {
  const int64_t bmax = (env.capacity + kBoundsTile) / kBoundsTile;
  const int64_t bg = bmax < (int64_t)env.num_cu * 8 ? bmax : (int64_t)env.num_cu * 8;
  hipLaunchKernelGGL(k_key_bounds, dim3((unsigned)bg), dim3(256), 0, env.stream, skeys, clk, mk,
                     offsets);
}

// Stable sort of the keys in ws.keys_a (slots [0, n)) over key values
// 0..n_keys, then the offsets and counts of every key.  The permutation goes
// to perm; the sorted keys end in keys_b (one 8-bit pass) or keys_a (two),
// returned in *skeys.
static int sort_keys(const LaunchEnv& env, const SortWs& ws, int n_keys, int32_t* counts,
                     int32_t* offsets, uint32_t* perm, const uint32_t** skeys) {
  if (n_keys + 1 > 65535) return fail("sort: at most 65534 keys (two 8-bit passes)");
  const int passes = (n_keys + 1) <= 256 ? 1 : 2;
  // the first of two passes parks the permutation in vals_b, or in sel when
  // the caller's perm is vals_b itself (sel is free until the resampling)
  uint32_t* park = perm == ws.vals_b ? ws.sel : ws.vals_b;
  uint32_t* kin = ws.keys_a;
  const uint32_t* vin = nullptr;
  for (int ps = 0; ps < passes; ++ps) {
    const bool last = ps == passes - 1;
    uint32_t* kout = ps == 0 ? ws.keys_b : ws.keys_a;
    uint32_t* vout = last ? perm : park;
    radix_pass(env, ws, env.clock, kin, vin, kout, vout, 8 * ps, ps == 0 ? 1 : 0);
    const int rc = check_launch("radix pass");
    if (rc) return rc;
    kin = kout;
    vin = vout;
  }
  // offsets of every key from the sorted keys, then the counts
  const int mk = n_keys + 1;  // key values 0..n_keys (n_keys: dead slots)
  key_bounds(env, env.clock, kin, mk, offsets);
  hipLaunchKernelGGL(k_key_counts, dim3((mk + 255) / 256), dim3(256), 0, env.stream,
                     (const int32_t*)offsets, mk, counts, env.clock);
  if (skeys) *skeys = kin;
  return check_launch("key counts");
}

BinKeys bin_keys(const bz_params_t* p, const double* thr, int nb) {
  BinKeys bk;
  const int n_nodes = p->dim == 1 ? p->n_nodes : 1;
  bk.dx = p->dx;
  // unused thresholds are +inf: no energy reaches them
  for (int k = 0; k < 8; ++k) bk.thr[k] = (thr && k < nb - 1) ? thr[k] : HUGE_VAL;
  bk.nb = nb;
  bk.n_nodes = n_nodes;
  bk.n_keys = n_nodes * nb;
  bk.dim = p->dim;
  bk.ew = (p->ref_flags & BZ_REF_ENERGY_ROUNDTRIP) ? 1 : 0;
  return bk;
}

static KeyArgs key_args(const bz_state_t* s, const bz_params_t* p, const double* thr, int nb) {
  KeyArgs ka;
  ka.x = s->x;
  ka.vx = s->vx;
  ka.vy = s->vy;
  ka.vz = s->vz;
  ka.w = s->w;
  ka.kev = p->kev;
  ka.bk = bin_keys(p, thr, nb);
  return ka;
}

int launch_bin_count(const LaunchEnv& env, const bz_state_t* s, const bz_params_t* p,
                     const double* thr, int nb, int32_t* counts) {
  const KeyArgs ka = key_args(s, p, thr, nb);
  const int m = ka.bk.n_keys + 1;
  // an LDS histogram per block while it fits in 64 KiB, global atomics above
  const int lds_keys = m <= 16384 ? m : 0;
  hipLaunchKernelGGL(k_zero_i32, dim3(64), dim3(256), 0, env.stream, counts, (int64_t)m);
  if (env.precision == BZ_F64)
    hipLaunchKernelGGL(k_key_hist<double>, dim3(env.num_cu * 2), dim3(256),
                       (size_t)lds_keys * 4, env.stream, ka, (const Clock*)env.clock, counts,
                       lds_keys);
  else
    hipLaunchKernelGGL(k_key_hist<float>, dim3(env.num_cu * 2), dim3(256),
                       (size_t)lds_keys * 4, env.stream, ka, (const Clock*)env.clock, counts,
                       lds_keys);
  return check_launch("k_key_hist");
}

int launch_bin_sort(const LaunchEnv& env, const bz_state_t* s, const bz_params_t* p,
                    const double* thr, int nb, int32_t* counts, int32_t* offsets,
                    uint32_t* perm, bool keys_ready) {
  const int n_nodes = p->dim == 1 ? p->n_nodes : 1;
  const int n_keys = n_nodes * nb;
  if (n_keys + 1 > 65535) return fail("bin_sort: at most 65534 keys (two 8-bit passes)");
  const int64_t cap = env.capacity;
  SortWs ws = carve(env.sort_ws, cap, env.max_keys);
  const KeyArgs ka = key_args(s, p, thr, nb);
  timing_begin(env, 1);
  int rc = 0;
  if (!keys_ready) {
    // (otherwise the push that left this state wrote the keys already)
    auto kk = env.precision == BZ_F64 ? k_keys<double> : k_keys<float>;
    // one thread per kKeyItems slots of the capacity; blocks past the count exit
    const int64_t kgrid =
        (cap + (int64_t)kSortBlock * kKeyItems - 1) / ((int64_t)kSortBlock * kKeyItems);
    hipLaunchKernelGGL(kk, dim3((unsigned)(kgrid > 0 ? kgrid : 1)), dim3(kSortBlock), 0,
                       env.stream, ka, (const Clock*)env.clock, ws.keys_a);
    rc = check_launch("k_keys");
    if (rc) return rc;
  }
  rc = sort_keys(env, ws, n_keys, counts, offsets, perm, nullptr);
  timing_end(env, 1);
  return rc;
}

// ---------------------------------------------------------------------------
// The reference's bin_particles on its own arrays (1bin/kernels/
// bin_particles.py:7-22, 3bin :7-37, 5bin :7-54): row j of tosum (w in
// column 0, the energy in column 3) with w > 0 and 0 <= curr_xbins[j] <= N
// goes to the list of its cell and energy bin, the bin being the number of
// thresholds at or below the energy (the reference's chain of `<`
// compares: a NaN energy matches none of them and is listed nowhere when
// there are several bins).  Lists are filled in ascending row order (the
// reference's atomic slots give one order of many) up to each bin's
// max_ppc; write_inds get the full counts, as the reference's atomic
// counters do.
// ---------------------------------------------------------------------------
struct RefBins {
  double thr[8];
  int64_t* list[8];
  int64_t* inds[8];
  int64_t max_ppc[8];
};

__global__ __launch_bounds__(256) void k_ref_keys(const double* __restrict__ tosum, int64_t cols,
                                                  const int64_t* __restrict__ cx,
                                                  const Clock* clk, int n_nodes, int nb,
                                                  RefBins rb, uint32_t* keys) {
  const int64_t n = clk->n;
  const uint32_t dead = (uint32_t)(n_nodes * nb);
  for (int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x; j < n;
       j += (int64_t)gridDim.x * 256) {
    const double w = tosum[j * cols];
    const double e = tosum[j * cols + 3];
    const int64_t cc = cx[j];
    uint32_t key = dead;
    if (w > 0.0 && cc >= 0 && cc < n_nodes && (nb == 1 || e == e)) {
      int b = 0;
      for (int q = 0; q < nb - 1; ++q) b += e >= rb.thr[q] ? 1 : 0;
      key = (uint32_t)(cc * nb + b);
    }
    keys[j] = key;
  }
}

__global__ __launch_bounds__(256) void k_ref_lists(const uint32_t* __restrict__ skeys,
                                                   const uint32_t* __restrict__ perm,
                                                   const int32_t* __restrict__ offsets,
                                                   const int32_t* __restrict__ counts,
                                                   const Clock* clk, int n_nodes, int nb,
                                                   RefBins rb) {
  const int64_t n = clk->n;
  const int n_keys = n_nodes * nb;
  const int64_t stride = (int64_t)gridDim.x * 256;
  for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < n_keys; k += stride)
    rb.inds[k % nb][k / nb] = counts[k];
  for (int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x; p < n; p += stride) {
    const uint32_t key = skeys[p];
    if (key >= (uint32_t)n_keys) continue;
    const int b = (int)(key % (uint32_t)nb), cell = (int)(key / (uint32_t)nb);
    const int64_t j = p - offsets[key];
    if (j < rb.max_ppc[b]) rb.list[b][j * n_nodes + cell] = (int64_t)perm[p];
  }
}

__global__ void k_set_count(Clock* clk, int64_t n) { clk->n = n; }

int launch_bin_lists(const LaunchEnv& env, const double* tosum, int64_t cols,
                     const int64_t* curr_xbins, int64_t nc, int n_cells, int nb,
                     const double* thr, int64_t* const* lists, int64_t* const* write_inds,
                     const int64_t* max_ppc) {
  const int n_nodes = n_cells + 1;
  const int n_keys = n_nodes * nb;
  if (n_keys > env.max_keys) return fail("bin lists: %d keys exceed the context's max_keys", n_keys);
  if (nc > env.capacity) return fail("bin lists: %lld rows exceed the capacity", (long long)nc);
  SortWs ws = carve(env.sort_ws, env.capacity, env.max_keys);
  RefBins rb;
  for (int q = 0; q < 8; ++q) {
    rb.thr[q] = q < nb - 1 ? thr[q] : HUGE_VAL;
    rb.list[q] = q < nb ? lists[q] : nullptr;
    rb.inds[q] = q < nb ? write_inds[q] : nullptr;
    rb.max_ppc[q] = q < nb ? max_ppc[q] : 0;
  }
  hipLaunchKernelGGL(k_set_count, dim3(1), dim3(1), 0, env.stream, env.clock, nc);
  const int grid = env.num_cu * 4;
  hipLaunchKernelGGL(k_ref_keys, dim3(grid), dim3(256), 0, env.stream, tosum, cols, curr_xbins,
                     (const Clock*)env.clock, n_nodes, nb, rb, ws.keys_a);
  int rc = check_launch("k_ref_keys");
  if (rc) return rc;
  const uint32_t* skeys = nullptr;
  uint32_t* perm = (uint32_t*)ws.copies;
  rc = sort_keys(env, ws, n_keys, ws.ip_counts, ws.ip_moff, perm, &skeys);
  if (rc) return rc;
  hipLaunchKernelGGL(k_ref_lists, dim3(grid), dim3(256), 0, env.stream, skeys,
                     (const uint32_t*)perm, (const int32_t*)ws.ip_moff,
                     (const int32_t*)ws.ip_counts, (const Clock*)env.clock, n_nodes, nb, rb);
  return check_launch("k_ref_lists");
}

// Batched swarms: key = instance (held in x), n_inst for a dead slot.
template <typename T>
__global__ __launch_bounds__(256) void k_inst_keys(const T* __restrict__ x,
                                                   const T* __restrict__ w, const Clock* clk,
                                                   int n_inst, uint32_t* keys) {
  const int64_t n = clk->n;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * 256) {
    uint32_t key = (uint32_t)n_inst;
    if (w[i] > (T)0) {
      int k = (int)x[i];
      key = (uint32_t)(k < 0 ? 0 : (k >= n_inst ? n_inst - 1 : k));
    }
    keys[i] = key;
  }
}

int launch_group_sort(const LaunchEnv& env, const bz_state_t* s, int n_inst, int32_t* counts,
                      int32_t* offsets, uint32_t* perm, const uint32_t** skeys) {
  SortWs ws = carve(env.sort_ws, env.capacity, env.max_keys);
  const unsigned grid = (unsigned)(env.num_cu * 8);
  if (env.precision == BZ_F64)
    hipLaunchKernelGGL(k_inst_keys<double>, dim3(grid), dim3(256), 0, env.stream,
                       (const double*)s->x, (const double*)s->w, (const Clock*)env.clock, n_inst,
                       ws.keys_a);
  else
    hipLaunchKernelGGL(k_inst_keys<float>, dim3(grid), dim3(256), 0, env.stream,
                       (const float*)s->x, (const float*)s->w, (const Clock*)env.clock, n_inst,
                       ws.keys_a);
  const int rc = check_launch("k_inst_keys");
  if (rc) return rc;
  return sort_keys(env, ws, n_inst, counts, offsets, perm, skeys);
}

// ---------------------------------------------------------------------------
// ES-WPNC resampling
// ---------------------------------------------------------------------------
enum : int { kCopy = 0, kSplit = 1, kMerge = 2 };

struct ResArgs {
  const void* in[5];
  void* out[5];
  bz_resample_policy_t pol;
  double kev;
  uint64_t seed;
  int n_keys;
  int prec;
  const double* U;   // caller's uniforms U[key*u_stride + m], or NULL: Philox
  int64_t u_stride;
};

// sorted keys of the final pass (keys_a for 2 passes, keys_b for 1)
__device__ __forceinline__ int key_at(const uint32_t* skeys, int64_t s) { return (int)skeys[s]; }

__device__ __forceinline__ int ceil_log2_i64(int64_t c) {
  int b = 0;
  while (((int64_t)1 << b) < c) ++b;
  return b;
}

template <typename T>
__device__ __forceinline__ double wt(const ResArgs& a, uint32_t idx) {
  return (double)((const T*)a.in[4])[idx];
}

// Phase A: heavy-particle counts of split candidates (one block per key).
template <typename T>
__global__ __launch_bounds__(256) void k_res_hw(ResArgs a, const int32_t* counts,
                                                const int32_t* offsets, const uint32_t* perm,
                                                const int32_t* targets, int32_t* plan) {
  const int k = blockIdx.x;
  const int c = counts[k];
  const int t = targets[k];
  __shared__ int s_sum;
  if (threadIdx.x == 0) s_sum = 0;
  __syncthreads();
  if (c > 0 && t > 0 && (double)c < a.pol.split_below * (double)t) {
    int h = 0;
    for (int j = threadIdx.x; j < c; j += 256)
      h += wt<T>(a, perm[offsets[k] + j]) > a.pol.min_wt ? 1 : 0;
    atomicAdd(&s_sum, h);
  }
  __syncthreads();
  if (threadIdx.x == 0) plan[4 * k + 3] = s_sum;
}

// Phase B (single block): kind and new count per key, trigger, offsets.
__global__ __launch_bounds__(1024) void k_res_plan(ResArgs a, const int32_t* counts,
                                                   const int32_t* targets, int32_t* plan,
                                                   int nb, Clock* clk, int32_t* klist,
                                                   int kstride) {
  __shared__ int s_fire;
  __shared__ int s_nlist[2];  // split and merge keys listed so far
  __shared__ int s_w[16];
  __shared__ int s_carry;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const bz_resample_policy_t& P = a.pol;
  if (tid == 0) {
    s_fire = (!P.ref_trigger && P.trigger_at <= 0.0) ? 1 : 0;  // <= 0: every check fires
    s_carry = 0;
    s_nlist[0] = 0;
    s_nlist[1] = 0;
  }
  __syncthreads();
  for (int k = tid; k < a.n_keys; k += 1024) {
    const int c = counts[k], t = targets[k];
    if (t <= 0 || c <= 0) continue;
    if (P.ref_trigger) {
      // only the last bin's maximum decides (3bin/boltzmann.py:1110-1113)
      if ((k % nb) == nb - 1 && (double)c >= 3.0 * (double)t) s_fire = 1;
    } else if ((double)c >= P.trigger_at * (double)t) {
      s_fire = 1;
    }
  }
  __syncthreads();
  const int fire = s_fire;
  int n_split = 0, n_merge = 0;
  unsigned long long n_copied = 0;  // particles k_res_copy moves unchanged
  for (int base = 0; base < a.n_keys; base += 1024) {
    const int k = base + tid;
    int kind = kCopy, nc = 0;
    if (k < a.n_keys) {
      const int c = counts[k], t = targets[k];
      nc = c;
      if (fire && t > 0) {
        const int keep = (int)(P.merge_keep * (double)t);
        if (c > 0 && (double)c < P.split_below * (double)t) {
          const int draws = (int)(P.split_to * (double)t) - c;
          if (plan[4 * k + 3] > 0 && draws > 0) {
            kind = kSplit;
            nc = c + draws;
            ++n_split;
            klist[2 + atomicAdd(&s_nlist[0], 1)] = k;
          }
        } else if ((double)c >= P.merge_at * (double)t && keep >= 1) {
          const int ce = (P.max_ppc > 0 && c > P.max_ppc) ? P.max_ppc : c;
          if (keep < ce) {
            kind = kMerge;
            nc = keep + (c - ce);
            ++n_merge;
            klist[2 + kstride + atomicAdd(&s_nlist[1], 1)] = k;
          }
        }
      }
      plan[4 * k + 0] = kind;
      plan[4 * k + 1] = nc;
      if (kind == kCopy) n_copied += (unsigned)c;
      else if (kind == kMerge && P.max_ppc > 0 && c > P.max_ppc)
        n_copied += (unsigned)(c - P.max_ppc);
    }
    const int inc = wave_incl_scan(nc);
    if (lane == 63) s_w[wid] = inc;
    __syncthreads();
    if (wid == 0) {
      int v = lane < 16 ? s_w[lane] : 0;
      v = wave_incl_scan(v);
      if (lane < 16) s_w[lane] = v;
    }
    __syncthreads();
    if (k < a.n_keys) plan[4 * k + 2] = s_carry + (wid ? s_w[wid - 1] : 0) + inc - nc;
    __syncthreads();
    if (tid == 0) s_carry += s_w[15];
    __syncthreads();
  }
  n_split = wave_sum(n_split);
  n_merge = wave_sum(n_merge);
  n_copied = wave_sum(n_copied);
  if (lane == 0) {
    if (n_split) atomicAdd(&clk->cnt[BZ_CNT_SPLIT], (unsigned long long)n_split);
    if (n_merge) atomicAdd(&clk->cnt[BZ_CNT_MERGE], (unsigned long long)n_merge);
    if (n_copied) atomicAdd(&clk->cnt[BZ_CNT_COPIED], n_copied);
  }
  if (tid == 0) {
    clk->n_next = s_carry;
    if (fire) atomicAdd(&clk->cnt[BZ_CNT_RESAMPLE], 1ull);
    klist[0] = s_nlist[0];
    klist[1] = s_nlist[1];
  }
}

// Copy segments (and the unmerged tail of clipped merges), per sorted
// position.  Each thread takes kCopyItems consecutive positions, with every
// load of a level (keys and permutation, then the plan, then the particle
// words) issued before any of it is used, so the gather's dependent loads
// overlap instead of running one particle at a time.
constexpr int kCopyItems = 4;

template <typename T>
__global__ __launch_bounds__(256) void k_res_copy(ResArgs a, const uint32_t* skeys,
                                                  const int32_t* offsets, const uint32_t* perm,
                                                  const int32_t* plan, const int32_t* counts,
                                                  const Clock* clk, int64_t cap) {
  const int64_t n_live = offsets[a.n_keys];
  const int64_t s0 = (int64_t)blockIdx.x * 256 * kCopyItems + threadIdx.x;
  if (s0 >= n_live) return;
  int key[kCopyItems];
  uint32_t src[kCopyItems];
#pragma unroll
  for (int k = 0; k < kCopyItems; ++k) {
    const int64_t s = s0 + k * 256;
    key[k] = s < n_live ? key_at(skeys, s) : -1;
    src[k] = s < n_live ? perm[s] : 0u;
  }
  int64_t dst[kCopyItems];
#pragma unroll
  for (int k = 0; k < kCopyItems; ++k) {
    dst[k] = -1;
    if (key[k] < 0) continue;
    const int kk = key[k];
    const int kind = plan[4 * kk];
    const int64_t j = s0 + k * 256 - offsets[kk];
    if (kind == kCopy) {
      dst[k] = plan[4 * kk + 2] + j;
    } else if (kind == kMerge) {
      const int c = counts[kk];
      const int ce = (a.pol.max_ppc > 0 && c > a.pol.max_ppc) ? a.pol.max_ppc : c;
      const int keep = plan[4 * kk + 1] - (c - ce);
      if (j >= ce) dst[k] = plan[4 * kk + 2] + keep + (j - ce);
    }
    if (dst[k] >= cap) dst[k] = -1;
  }
  T v[kCopyItems][5];
#pragma unroll
  for (int k = 0; k < kCopyItems; ++k)
#pragma unroll
    for (int f = 0; f < 5; ++f) v[k][f] = dst[k] >= 0 ? ((const T*)a.in[f])[src[k]] : (T)0;
#pragma unroll
  for (int k = 0; k < kCopyItems; ++k)
    if (dst[k] >= 0)
#pragma unroll
      for (int f = 0; f < 5; ++f) ((T*)a.out[f])[dst[k]] = v[k][f];
}

// block-wide helpers for the per-segment kernels (kSegBlock threads)
constexpr int kSegBlock = 512;
constexpr int kSegWaves = kSegBlock / 64;

// N sums at once over the block (one set of barriers); s_red holds N * kSegWaves
template <int N>
__device__ __forceinline__ void block_sum_n(double (&v)[N], double* s_red) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < N; ++q) v[q] = wave_sum(v[q]);
  __syncthreads();
  if (lane == 0)
#pragma unroll
    for (int q = 0; q < N; ++q) s_red[q * kSegWaves + wid] = v[q];
  __syncthreads();
#pragma unroll
  for (int q = 0; q < N; ++q) {
    double r = s_red[q * kSegWaves];
    for (int w = 1; w < kSegWaves; ++w) r += s_red[q * kSegWaves + w];
    v[q] = r;
  }
  __syncthreads();
}
__device__ __forceinline__ double block_sum_d(double v, double* s_red) {
  double t[1] = {v};
  block_sum_n<1>(t, s_red);
  return t[0];
}
__device__ __forceinline__ double block_max_d(double v, double* s_red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) s_red[wid] = v;
  __syncthreads();
  double r = s_red[0];
  for (int w = 1; w < kSegWaves; ++w) r = fmax(r, s_red[w]);
  __syncthreads();
  return r;
}

// Inclusive fixed-point prefix of q over sorted positions [s0, s0+c) into cdf.
// Returns the total.  q(j) is supplied by the functor; s_w64 holds kSegWaves.
// With samp, the last entry of every group of S positions also goes to
// samp[group] (the LDS index of a CDF kept in global memory).
template <typename Q>
__device__ int64_t block_fixed_scan(int64_t s0, int c, int64_t* cdf, Q q, int64_t* s_w64,
                                   int64_t* samp = nullptr, int S = 1) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int64_t carry = 0;
  for (int base = 0; base < c; base += kSegBlock) {
    const int j = base + threadIdx.x;
    const int64_t v = j < c ? q(j) : 0;
    const int64_t inc = wave_incl_scan(v);
    __syncthreads();
    if (lane == 63) s_w64[wid] = inc;
    __syncthreads();
    int64_t before = carry;
    for (int w = 0; w < wid; ++w) before += s_w64[w];
    if (j < c) {
      cdf[s0 + j] = before + inc;
      if (samp && ((j + 1) % S == 0 || j == c - 1)) samp[j / S] = before + inc;
    }
    for (int w = 0; w < kSegWaves; ++w) carry += s_w64[w];
  }
  __syncthreads();
  return carry;
}

// first j in [0, c) with cdf[s0+j] >= thr (left) or > thr (right)
__device__ __forceinline__ int search_cdf(const int64_t* cdf, int64_t s0, int c, int64_t thr,
                                          bool right) {
  int lo = 0, hi = c;  // answer in [lo, hi]
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    const int64_t v = cdf[s0 + mid];
    const bool go_left = right ? (v > thr) : (v >= thr);
    if (go_left) hi = mid;
    else lo = mid + 1;
  }
  return lo < c ? lo : c - 1;
}

// The same search (left) over a CDF in global memory with the last entry of
// every group of S in LDS (samp): the group from LDS, then at most S - 1
// global reads instead of a binary search's log2(c) dependent ones.
__device__ __forceinline__ int search_cdf_sampled(const int64_t* cdf, const int64_t* samp, int c,
                                                  int S, int64_t thr) {
  const int ng = (c + S - 1) / S;
  int lo = 0, hi = ng;  // first group whose last entry is >= thr
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (samp[mid] >= thr) hi = mid;
    else lo = mid + 1;
  }
  if (lo >= ng) return c - 1;
  const int j0 = lo * S, j1 = min(j0 + S, c);
  for (int j = j0; j < j1 - 1; ++j)
    if (cdf[j] >= thr) return j;
  return j1 - 1;
}

__device__ __forceinline__ double res_uniform(uint64_t seed, uint64_t step, uint32_t key,
                                              uint32_t m) {
  const U4 r = draw(seed, kPurposeResample, m >> 1, key, step);
  return (m & 1) ? u53(r.c, r.d) : u53(r.a, r.b);
}

// uniform m of segment `key`: the caller's stream when given, else Philox
__device__ __forceinline__ double seg_uniform(const ResArgs& a, uint64_t step, uint32_t key,
                                              uint32_t m) {
  return a.U ? a.U[(int64_t)key * a.u_stride + m] : res_uniform(a.seed, step, key, m);
}

// Split: weighted draws with replacement among w > min_wt; each drawn
// particle's weight is divided by (copies+1) and its copies follow it.
// Segments of up to kSplitLds particles keep the CDF, copy counts and output
// offsets in LDS; larger ones use the workspace.  Clones are written by all
// threads: output slot o belongs to the particle whose offset range holds o.
constexpr int kSplitLds = 3968;  // 3968 * 16 B stays under the 64 KiB default LDS limit

template <typename T, bool INPLACE>
__device__ __forceinline__ void res_split_key(int k, const ResArgs& a, const int32_t* counts,
                                              const int32_t* offsets, const uint32_t* perm,
                                              const int32_t* plan, const SortWs& ws,
                                              const Clock* clk, int64_t cap, int lds_cap) {
  extern __shared__ int64_t s_split[];  // cdf[lds_cap], copies/offs as int32 after
  __shared__ double s_red[kSegWaves];
  __shared__ int64_t s_w64[kSegWaves];
  __shared__ int s_wi[kSegWaves];
  const int c = counts[k];
  const int64_t s0 = offsets[k];
  const uint64_t step = clk->step;
  const double minw = a.pol.min_wt;
  const bool lds = c <= lds_cap;
  int64_t* cdf = lds ? s_split : ws.cdf + s0;
  int32_t* copies = lds ? (int32_t*)(s_split + lds_cap) : ws.copies + s0;
  int32_t* offs = lds ? (int32_t*)(s_split + lds_cap) + lds_cap : (int32_t*)ws.sel + s0;
  const T* Wt = (const T*)a.in[4];
  // fixed-point scale from the largest heavy weight
  double wmax = 0.0;
  for (int j = threadIdx.x; j < c; j += kSegBlock) {
    const double w = (double)Wt[perm[s0 + j]];
    if (w > minw) wmax = fmax(wmax, w);
    copies[j] = 0;
  }
  wmax = block_max_d(wmax, s_red);
  int ew;
  frexp(wmax, &ew);
  const int S = 61 - ew - ceil_log2_i64(c);
  auto q = [&](int j) -> int64_t {
    const double w = (double)Wt[perm[s0 + j]];
    return w > minw ? (int64_t)ldexp(w, S) : 0;
  };
  const int64_t Wq = block_fixed_scan(0, c, cdf, q, s_w64);
  // key order: plan holds the segment's new count; in place: the clones
  const int draws = INPLACE ? plan[4 * k + 1] : plan[4 * k + 1] - c;
  for (int m = threadIdx.x; m < draws; m += kSegBlock) {
    const double u = seg_uniform(a, step, (uint32_t)k, (uint32_t)m);
    const int64_t thr = (int64_t)(u * (double)Wq);
    atomicAdd(&copies[search_cdf(cdf, 0, c, thr, true)], 1);
  }
  __syncthreads();
  // exclusive offsets of (1 + copies) in segment order (in place: of the
  // clones alone, the particles themselves keep their slots)
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int carry = 0;
  for (int base = 0; base < c; base += kSegBlock) {
    const int j = base + threadIdx.x;
    const int reps = j < c ? (INPLACE ? 0 : 1) + copies[j] : 0;
    const int inc = wave_incl_scan(reps);
    __syncthreads();
    if (lane == 63) s_wi[wid] = inc;
    __syncthreads();
    int before = carry;
    for (int w = 0; w < wid; ++w) before += s_wi[w];
    if (j < c) offs[j] = before + inc - reps;
    for (int w = 0; w < kSegWaves; ++w) carry += s_wi[w];
  }
  __syncthreads();
  const int total = carry;  // key order: c + draws; in place: draws
  const int64_t out0 = plan[4 * k + 2];
  for (int o = threadIdx.x; o < total; o += kSegBlock) {
    const int64_t dst = out0 + o;
    if (dst >= cap) break;
    // largest j with offs[j] <= o: the member whose output range holds o
    int lo = 0, hi = c - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (offs[mid] <= o) lo = mid;
      else hi = mid - 1;
    }
    const uint32_t src = perm[s0 + lo];
    const int reps = 1 + copies[lo];
    T w = Wt[src];
    if (reps > 1) w = w / (T)reps;
#pragma unroll
    for (int f = 0; f < 4; ++f) ((T*)a.out[f])[dst] = ((const T*)a.in[f])[src];
    ((T*)a.out[4])[dst] = w;
  }
  if (INPLACE) {
    // the drawn particles keep their slots with the divided weight, written
    // once every clone has read the original
    __syncthreads();
    for (int j = threadIdx.x; j < c; j += kSegBlock) {
      if (!copies[j]) continue;
      const uint32_t src = perm[s0 + j];
      ((T*)a.out[4])[src] = Wt[src] / (T)(1 + copies[j]);
    }
  }
}

// Merge: keep K = int(merge_keep*target) particles chosen by stratified
// sampling of the weight CDF, each with weight W/K; velocities rescaled about
// the mean so that weight, momentum and energy are conserved.
constexpr int kMergeLds = 6144;  // 48 KiB of CDF

template <typename T, bool INPLACE>
__device__ __forceinline__ void res_merge_key(int k, const ResArgs& a, const int32_t* counts,
                                              const int32_t* offsets, const uint32_t* perm,
                                              const int32_t* targets, const int32_t* plan,
                                              const SortWs& ws, Clock* clk, int64_t cap,
                                              int lds_cap) {
  __shared__ double s_red[5 * kSegWaves];
  __shared__ int64_t s_w64[kSegWaves];
  const int c = counts[k];
  const int ce = (a.pol.max_ppc > 0 && c > a.pol.max_ppc) ? a.pol.max_ppc : c;
  // key order: plan holds the new count (K and the clipped tail); in place: K
  const int K = INPLACE ? plan[4 * k + 1] : plan[4 * k + 1] - (c - ce);
  const int64_t s0 = offsets[k];
  const int64_t out0 = plan[4 * k + 2];
  const uint64_t step = clk->step;
  const T* X = (const T*)a.in[0];
  const T* VX = (const T*)a.in[1];
  const T* VY = (const T*)a.in[2];
  const T* VZ = (const T*)a.in[3];
  const T* Wt = (const T*)a.in[4];
  // FP64 moments of the old set
  double w0 = 0, px = 0, py = 0, pz = 0, en = 0, wmax = 0;
  for (int j = threadIdx.x; j < ce; j += kSegBlock) {
    const uint32_t i = perm[s0 + j];
    const double w = (double)Wt[i];
    const double vx = VX[i], vy = VY[i], vz = VZ[i];
    w0 += w;
    px += w * vx;
    py += w * vy;
    pz += w * vz;
    en += w * (vx * vx + vy * vy + vz * vz);
    wmax = fmax(wmax, w);
  }
  double m5[5] = {w0, px, py, pz, en};
  block_sum_n<5>(m5, s_red);
  const double W = m5[0], Px = m5[1], Py = m5[2], Pz = m5[3], En = m5[4];
  wmax = block_max_d(wmax, s_red);
  const double mx = Px / W, my = Py / W, mz = Pz / W;
  double c2 = 0;
  for (int j = threadIdx.x; j < ce; j += kSegBlock) {
    const uint32_t i = perm[s0 + j];
    const double w = (double)Wt[i];
    const double dx_ = VX[i] - mx, dy_ = VY[i] - my, dz_ = VZ[i] - mz;
    c2 += w * (dx_ * dx_ + dy_ * dy_ + dz_ * dz_);
  }
  const double C_old = block_sum_d(c2, s_red);
  // exact fixed-point CDF of the weights
  int ew;
  frexp(wmax, &ew);
  const int S = 61 - ew - ceil_log2_i64(ce);
  auto q = [&](int j) -> int64_t { return (int64_t)ldexp((double)Wt[perm[s0 + j]], S); };
  // the CDF in LDS when it fits (the selection's binary searches read it);
  // else in the workspace, with every S-th entry in LDS
  extern __shared__ int64_t s_mcdf[];
  const bool lds = ce <= lds_cap;
  int64_t* cdf = lds ? s_mcdf : ws.cdf + s0;
  const int G = lds ? 1 : (ce + lds_cap - 1) / lds_cap;  // entries per LDS sample
  const int64_t Wq = block_fixed_scan(0, ce, cdf, q, s_w64, lds ? nullptr : s_mcdf, G);
  const double wbar = W / (double)K;
  // the merged weight may exceed every weight so far: raise the bound the
  // deposit scales its fixed-point sums by
  if (threadIdx.x == 0)
    atomicMax(&clk->w_max_bits, (unsigned long long)__double_as_longlong((double)(T)wbar));
  // stratified selection: r_k = (1/K)(U_k + k), searchsorted left
  double sx = 0, sy = 0, sz = 0, se = 0;
  for (int m = threadIdx.x; m < K; m += kSegBlock) {
    const double u = seg_uniform(a, step, (uint32_t)k, (uint32_t)m);
    const double r = (1.0 / (double)K) * (u + (double)m);
    const int64_t thr = (int64_t)(r * (double)Wq);
    const int j = lds ? search_cdf(cdf, 0, ce, thr, false)
                      : search_cdf_sampled(cdf, s_mcdf, ce, G, thr);
    if (out0 + m >= cap) continue;  // overflow: counted at commit
    ws.sel[out0 + m] = (uint32_t)(s0 + j);
    const uint32_t i = perm[s0 + j];
    const double vx = VX[i], vy = VY[i], vz = VZ[i];
    sx += vx;
    sy += vy;
    sz += vz;
    se += vx * vx + vy * vy + vz * vz;
  }
  double s4[4] = {sx, sy, sz, se};
  block_sum_n<4>(s4, s_red);
  const double Sx = s4[0], Sy = s4[1], Sz = s4[2], Se = s4[3];
  double hx, hy, hz, f;
  if (a.pol.ref_rescale) {
    // boltzmann.py:835-869: hats normalised by the original weight, factor from total energy
    hx = wbar * Sx / W;
    hy = wbar * Sy / W;
    hz = wbar * Sz / W;
    const double T_cell = (2.0 / 3.0) * (a.kev * En) / W;
    const double T_hat = (2.0 / 3.0) * (wbar * (a.kev * Se)) / W;
    f = T_hat > 0 ? sqrt(T_cell / T_hat) : 1.0;
  } else {
    hx = Sx / (double)K;
    hy = Sy / (double)K;
    hz = Sz / (double)K;
    double cn = 0;
    for (int m = threadIdx.x; m < K; m += kSegBlock) {
      if (out0 + m >= cap) break;
      const uint32_t i = perm[ws.sel[out0 + m]];
      const double dx_ = VX[i] - hx, dy_ = VY[i] - hy, dz_ = VZ[i] - hz;
      cn += dx_ * dx_ + dy_ * dy_ + dz_ * dz_;
    }
    const double C_new = wbar * block_sum_d(cn, s_red);
    f = C_new > 0 ? sqrt(C_old / C_new) : 1.0;
  }
  for (int m = threadIdx.x; m < K; m += kSegBlock) {
    const int64_t dst = out0 + m;
    if (dst >= cap) break;
    const uint32_t i = perm[ws.sel[dst]];
    ((T*)a.out[0])[dst] = X[i];
    ((T*)a.out[1])[dst] = (T)(mx + ((double)VX[i] - hx) * f);
    ((T*)a.out[2])[dst] = (T)(my + ((double)VY[i] - hy) * f);
    ((T*)a.out[3])[dst] = (T)(mz + ((double)VZ[i] - hz) * f);
    ((T*)a.out[4])[dst] = (T)wbar;
  }
  if (INPLACE) {
    // the merged members become holes once every selection has been read
    __syncthreads();
    for (int j = threadIdx.x; j < ce; j += kSegBlock) ((T*)a.out[4])[perm[s0 + j]] = (T)0;
  }
}

// Split / merge over the keys k_res_plan listed: a grid sized to what is
// resident walks the list, so the many keys that only copy cost no block
// (the split's 62 KiB of LDS lets two blocks share a CU).
template <typename T, bool INPLACE>
__global__ __launch_bounds__(kSegBlock) void k_res_split(ResArgs a, const int32_t* counts,
                                                   const int32_t* offsets, const uint32_t* perm,
                                                   const int32_t* plan, SortWs ws,
                                                   const Clock* clk, int64_t cap) {
  const int n_list = ws.klist[0];
  for (int li = blockIdx.x; li < n_list; li += gridDim.x) {
    __syncthreads();  // the previous key is done with the LDS
    res_split_key<T, INPLACE>(ws.klist[2 + li], a, counts, offsets, perm, plan, ws, clk, cap,
                              kSplitLds);
  }
}

template <typename T, bool INPLACE>
__global__ __launch_bounds__(kSegBlock) void k_res_merge(ResArgs a, const int32_t* counts,
                                                   const int32_t* offsets, const uint32_t* perm,
                                                   const int32_t* targets, const int32_t* plan,
                                                   SortWs ws, Clock* clk, int64_t cap) {
  const int n_list = ws.klist[1];
  for (int li = blockIdx.x; li < n_list; li += gridDim.x) {
    __syncthreads();
    res_merge_key<T, INPLACE>(ws.klist[2 + ws.kstride + li], a, counts, offsets, perm, targets,
                              plan, ws, clk, cap, kMergeLds);
  }
}

// The in-place pass runs its merges and splits in one launch, so a pass
// takes about the longer of the two (the merges: gathers of their members)
// rather than their sum.  Blocks of 39 KiB of LDS (four would fit a CU;
// the merge's 117 VGPRs leave room for two 512-thread blocks, and forcing
// more spills registers); larger segments keep their CDF in the workspace.
// Merges first in the walk.
constexpr int kSegSplitLds = 2432;  // x 16 B
constexpr int kSegMergeLds = 4864;  // x 8 B
constexpr size_t kSegLdsBytes = (size_t)kSegSplitLds * 16;
static_assert((size_t)kSegMergeLds * 8 == kSegLdsBytes, "one dynamic LDS size for both");

template <typename T>
__global__ __launch_bounds__(kSegBlock) void k_res_segments(ResArgs a, const int32_t* counts,
                                                            const int32_t* offsets,
                                                            const uint32_t* perm,
                                                            const int32_t* targets,
                                                            const int32_t* plan, SortWs ws,
                                                            Clock* clk, int64_t cap) {
  const int n_split = ws.klist[0], n_merge = ws.klist[1];
  for (int li = blockIdx.x; li < n_merge + n_split; li += gridDim.x) {
    __syncthreads();  // the previous segment is done with the LDS
    if (li < n_merge)
      res_merge_key<T, true>(ws.klist[2 + ws.kstride + li], a, counts, offsets, perm, targets,
                             plan, ws, clk, cap, kSegMergeLds);
    else
      res_split_key<T, true>(ws.klist[2 + (li - n_merge)], a, counts, offsets, perm, plan, ws, clk,
                             cap, kSegSplitLds);
  }
}

__global__ void k_res_commit(Clock* clk, int64_t cap) {
  int64_t n = clk->n_next;
  if (n > cap) {
    atomicAdd(&clk->cnt[BZ_CNT_OVERFLOW], (unsigned long long)(n - cap));
    n = cap;
  }
  clk->n = n;
  clk->n_next = n;
}

// ---------------------------------------------------------------------------
// In-place resampling, the reference's layout (boltzmann.py:812-823,
// :850-874): copy segments stay where they are, a split particle keeps its
// slot with weight w/(c+1) and its clones are appended, a merged segment's
// members become holes (w = 0) and its K particles are appended.  Only the
// members of split / merge segments are grouped (extracted in slot order and
// sorted by key), so a pass costs a key histogram, the extraction and a sort
// of the members instead of a sort and a copy of the whole state.  Members
// of a segment are its slots in ascending order, as after the full sort, so
// the selections are those of the key-order resampler.
// ---------------------------------------------------------------------------

// Slots per key from the keys the push (or k_keys) wrote.  An LDS histogram
// per block (lds_keys counters; 0: global atomics), added with integer
// atomics: exact and order independent.  A thread reads 16-byte vectors of
// consecutive keys and adds each run of equal keys in a vector with one
// atomic: neighbouring slots often share a key (a segment's appended merges
// and clones, a cell's particles loaded in order), and lanes adding to one
// LDS word serialise.
constexpr int kIpHistBlock = 1024;  // one block per CU: the LDS histogram of every key
constexpr int kIpHistVecs = 4;      // 16-byte vectors per thread in flight

__device__ __forceinline__ void ip_hist_add(uint32_t key, uint32_t c, uint32_t* s_c,
                                            int32_t* counts, int lds_keys) {
  if (lds_keys) atomicAdd(&s_c[key], c);
  else atomicAdd(&counts[key], (int32_t)c);
}

template <typename K>
__global__ __launch_bounds__(kIpHistBlock) void k_ip_hist(const K* __restrict__ keys,
                                                          const Clock* clk, int32_t* counts,
                                                          int lds_keys) {
  static_assert(sizeof(K) == 2 || sizeof(K) == 4, "16- or 32-bit keys");
  constexpr int KPV = 16 / (int)sizeof(K);  // keys per vector
  extern __shared__ uint32_t s_c[];
  const int64_t n = clk->n;
  for (int j = threadIdx.x; j < lds_keys; j += kIpHistBlock) s_c[j] = 0u;
  __syncthreads();
  // whole vectors (the key buffer is 16-byte aligned), then the tail
  const int64_t nv = n / KPV;
  const uint4* kv = (const uint4*)keys;
  const int64_t stride = (int64_t)gridDim.x * kIpHistBlock * kIpHistVecs;
  for (int64_t base = (int64_t)blockIdx.x * kIpHistBlock * kIpHistVecs + threadIdx.x; base < nv;
       base += stride) {
    uint4 v[kIpHistVecs];
#pragma unroll
    for (int k = 0; k < kIpHistVecs; ++k) {
      const int64_t i = base + k * kIpHistBlock;
      v[k] = i < nv ? kv[i] : make_uint4(0u, 0u, 0u, 0u);
    }
#pragma unroll
    for (int k = 0; k < kIpHistVecs; ++k) {
      if (base + k * kIpHistBlock >= nv) continue;
      const uint32_t w[4] = {v[k].x, v[k].y, v[k].z, v[k].w};
      uint32_t run = KPV == 8 ? (w[0] & 0xFFFFu) : w[0], c = 0;
#pragma unroll
      for (int q = 0; q < KPV; ++q) {
        const uint32_t key = KPV == 8 ? ((q & 1) ? w[q >> 1] >> 16 : w[q >> 1] & 0xFFFFu) : w[q];
        if (key != run) {
          ip_hist_add(run, c, s_c, counts, lds_keys);
          run = key;
          c = 0;
        }
        ++c;
      }
      ip_hist_add(run, c, s_c, counts, lds_keys);
    }
  }
  if (blockIdx.x == gridDim.x - 1)
    for (int64_t i = nv * KPV + threadIdx.x; i < n; i += kIpHistBlock)
      ip_hist_add((uint32_t)keys[i], 1u, s_c, counts, lds_keys);
  __syncthreads();
  for (int j = threadIdx.x; j < lds_keys; j += kIpHistBlock)
    if (s_c[j]) atomicAdd(&counts[j], (int32_t)s_c[j]);
}

// Block-wide exclusive scan of NV values per thread (1024 threads); tot gets
// the block totals.  s holds NV x 16 words.
template <int NV>
__device__ __forceinline__ void block_scan1024(long long (&v)[NV], long long (&tot)[NV],
                                               long long (*s)[16]) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  long long inc[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) inc[q] = wave_incl_scan(v[q]);
  if (lane == 63)
#pragma unroll
    for (int q = 0; q < NV; ++q) s[q][wid] = inc[q];
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    long long before = 0, t = 0;
    for (int w = 0; w < 16; ++w) {
      if (w < wid) before += s[q][w];
      t += s[q][w];
    }
    tot[q] = t;
    v[q] = before + inc[q] - v[q];
  }
  __syncthreads();
}

// The candidates and the trigger, one key per thread: cand[k] = 1 for a
// split candidate (its heavy-particle count is not known yet), 2 for a
// merge, 0 otherwise, as if the pass fired; chunk-local member offsets of the
// candidates into moff; a key that trips the trigger sets the fire word.
__global__ __launch_bounds__(kKeyChunk) void k_ip_cand(ResArgs a, const int32_t* counts,
                                                       const int32_t* targets, int nb,
                                                       int32_t* cand, int32_t* moff,
                                                       long long* chunk, uint32_t* bits) {
  __shared__ long long s[1][16];
  const bz_resample_policy_t& P = a.pol;
  const int k = blockIdx.x * kKeyChunk + threadIdx.x;
  int kind = 0, c = 0;
  bool trig = false;
  if (k < a.n_keys) {
    c = counts[k];
    const int t = targets[k];
    if (t > 0 && c > 0) {
      if (P.ref_trigger)
        // only the last bin's maximum decides (3bin/boltzmann.py:1110-1113)
        trig = (k % nb) == nb - 1 && (double)c >= 3.0 * (double)t;
      else
        trig = (double)c >= P.trigger_at * (double)t;
    }
    if (t > 0) {
      const int keep = (int)(P.merge_keep * (double)t);
      if (c > 0 && (double)c < P.split_below * (double)t) {
        if ((int)(P.split_to * (double)t) - c > 0) kind = 1;
      } else if ((double)c >= P.merge_at * (double)t && keep >= 1) {
        const int ce = (P.max_ppc > 0 && c > P.max_ppc) ? P.max_ppc : c;
        if (keep < ce) kind = 2;
      }
    }
    cand[k] = kind;
  }
  // the candidate keys as bits, 64 per wave (two words)
  const uint64_t cm = __ballot(kind != 0);
  // (words past the last key are never read: ip_listed checks the key first)
  if ((threadIdx.x & 63) == 0 && k < a.n_keys) {
    bits[k >> 5] = (uint32_t)cm;
    if (k + 32 < a.n_keys) bits[(k >> 5) + 1] = (uint32_t)(cm >> 32);
  }
  if (__any(trig) && (threadIdx.x & 63) == 0) atomicOr((int*)&chunk[8 * kMaxKeyChunks], 1);
  long long v[1] = {kind ? (long long)c : 0ll}, tot[1];
  block_scan1024<1>(v, tot, s);
  if (k < a.n_keys) moff[k] = (int32_t)v[0];
  if (threadIdx.x == 0) chunk[8 * blockIdx.x + 0] = tot[0];
}

// One block: the trigger (every check fires when trigger_at <= 0 and the
// reference rule is off), chunk prefixes of the member offsets and the member
// count (nothing is listed when the pass does not fire).  The fire word is
// cleared for the next pass; cand[n_keys] keeps the decision.
__global__ __launch_bounds__(64) void k_ip_cand_carry(ResArgs a, int nch, int32_t* cand,
                                                      long long* chunk, Clock* mclk) {
  const bz_resample_policy_t& P = a.pol;
  if (threadIdx.x != 0) return;
  int* fw = (int*)&chunk[8 * kMaxKeyChunks];
  const int fire = ((!P.ref_trigger && P.trigger_at <= 0.0) || *fw) ? 1 : 0;
  *fw = 0;
  cand[a.n_keys] = fire;
  long long acc = 0;
  for (int b = 0; b < nch; ++b) {
    const long long t = chunk[8 * b + 0];
    chunk[8 * b + 4] = acc;
    acc += t;
  }
  mclk->n = fire ? acc : 0;
  mclk->n_next = mclk->n;
}

// Extraction of the members of candidate segments, stable: global wave g
// owns a contiguous run of slots (whole 512-slot rows of 8 per lane);
// k_ip_wcount counts its members, k_ip_wscan turns the counts into
// exclusive offsets, k_ip_wscatter writes (key, slot) of the wave's members
// in slot order.  No barriers: each wave works alone.
__device__ __forceinline__ bool ip_listed(const uint32_t* bits, uint32_t key, int n_keys) {
  return key < (uint32_t)n_keys && ((bits[key >> 5] >> (key & 31)) & 1u);
}

// the candidate bits into LDS (every block of the extraction): the per-slot
// lookups are then LDS reads instead of scattered global loads
__device__ __forceinline__ void ip_stage_bits(const uint32_t* __restrict__ bits, int n_keys,
                                              uint32_t* s_bits) {
  const int words = n_keys / 32 + 1;
  for (int j = threadIdx.x; j < words; j += 256) s_bits[j] = bits[j];
  __syncthreads();
}

constexpr int kIpRow = 8;  // keys per lane and row
__device__ __forceinline__ void ip_wave_range(int64_t n, int64_t* lo, int64_t* hi) {
  const int64_t g = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int64_t rows = (n + 64 * kIpRow - 1) / (64 * kIpRow);
  const int64_t per = (rows + kIpWaves - 1) / kIpWaves;
  const int64_t a = g * per * 64 * kIpRow, b = a + per * 64 * kIpRow;
  *lo = a < n ? a : n;
  *hi = b < n ? b : n;
}

template <typename K>
__global__ __launch_bounds__(256) void k_ip_wcount(const K* __restrict__ keys,
                                                   const Clock* clk,
                                                   const int32_t* __restrict__ cand, int n_keys,
                                                   const uint32_t* __restrict__ bits,
                                                   uint32_t* wcnt) {
  extern __shared__ uint32_t s_bits[];
  const int lane = threadIdx.x & 63;
  int64_t lo, hi;
  ip_wave_range(clk->n, &lo, &hi);
  uint32_t c = 0;
  const int fire = cand[n_keys];
  if (fire) ip_stage_bits(bits, n_keys, s_bits);  // block-uniform
  if (fire) {
    for (int64_t b = lo; b < hi; b += 64 * kIpRow) {
      uint32_t kv[kIpRow];
#pragma unroll
      for (int u = 0; u < kIpRow; ++u) {
        const int64_t i = b + u * 64 + lane;
        kv[u] = i < hi ? (uint32_t)keys[i] : 0xFFFFFFFFu;
      }
#pragma unroll
      for (int u = 0; u < kIpRow; ++u) c += ip_listed(s_bits, kv[u], n_keys) ? 1u : 0u;
    }
  }
  c = wave_sum(c);
  if (lane == 0) wcnt[(int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)] = c;
}

// exclusive scan of the kIpWaves wave counts in place (one block, eight
// consecutive counts per thread)
static_assert(kIpWaves == 8 * 1024, "k_ip_wscan: eight wave counts per thread");
__global__ __launch_bounds__(1024) void k_ip_wscan(uint32_t* wcnt) {
  __shared__ long long s[1][16];
  uint4* p = (uint4*)wcnt + 2 * threadIdx.x;
  const uint4 lo = p[0], hi = p[1];
  uint32_t v[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
  long long x[1] = {0}, tot[1];
#pragma unroll
  for (int q = 0; q < 8; ++q) x[0] += v[q];
  block_scan1024<1>(x, tot, s);
  uint32_t run = (uint32_t)x[0];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const uint32_t t = v[q];
    v[q] = run;
    run += t;
  }
  p[0] = make_uint4(v[0], v[1], v[2], v[3]);
  p[1] = make_uint4(v[4], v[5], v[6], v[7]);
}

template <typename K>
__global__ __launch_bounds__(256) void k_ip_wscatter(const K* __restrict__ keys,
                                                     const Clock* clk,
                                                     const int32_t* __restrict__ cand,
                                                     int n_keys, const uint32_t* __restrict__ bits,
                                                     const uint32_t* wpre,
                                                     uint32_t* mkeys, uint32_t* mslot) {
  extern __shared__ uint32_t s_bits[];
  const int lane = threadIdx.x & 63;
  if (!cand[n_keys]) return;  // block-uniform
  ip_stage_bits(bits, n_keys, s_bits);
  int64_t lo, hi;
  ip_wave_range(clk->n, &lo, &hi);
  int64_t pos = wpre[(int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)];
  const uint64_t lt = (1ull << lane) - 1ull;
  for (int64_t b = lo; b < hi; b += 64 * kIpRow) {
    uint32_t kv[kIpRow];
#pragma unroll
    for (int u = 0; u < kIpRow; ++u) {
      const int64_t i = b + u * 64 + lane;
      kv[u] = i < hi ? (uint32_t)keys[i] : 0xFFFFFFFFu;
    }
#pragma unroll
    for (int u = 0; u < kIpRow; ++u) {
      const bool lst = ip_listed(s_bits, kv[u], n_keys);
      const uint64_t m = __ballot(lst);
      if (lst) {
        const int64_t d = pos + __popcll(m & lt);
        mkeys[d] = kv[u];
        mslot[d] = (uint32_t)(b + u * 64 + lane);
      }
      pos += __popcll(m);
    }
  }
}

// Member offsets and counts of every key (zero for keys that are not
// candidates, or when the pass does not fire), then whether a split
// candidate has a heavy particle (w > min_wt) among its grouped members (1)
// or none (0): the plan only asks whether the split draws, so the walk over
// a segment stops at the first 64 members that hold one.  One wave per key.
template <typename T>
__global__ __launch_bounds__(256) void k_ip_hw(ResArgs a, const int32_t* counts,
                                               const int32_t* cand, const long long* chunk,
                                               int32_t* moff, int32_t* mcnt,
                                               const uint32_t* mslot, int32_t* plan) {
  const int k = blockIdx.x * 4 + (int)(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (k >= a.n_keys) return;  // wave-uniform
  const int kc = cand[k];
  const int c = cand[a.n_keys] && kc ? counts[k] : 0;
  const int64_t off = chunk[8 * (k / kKeyChunk) + 4] + moff[k];
  if (lane == 0) {
    moff[k] = (int32_t)off;
    mcnt[k] = c;
  }
  if (kc != 1) return;  // wave-uniform
  bool any = false;
  for (int base = 0; base < c && !any; base += 64) {  // wave-uniform
    const int j = base + lane;
    any = __any(j < c && wt<T>(a, mslot[off + j]) > a.pol.min_wt);
  }
  if (lane == 0) plan[4 * k + 3] = any ? 1 : 0;
}

// Final kinds and appended counts, one key per thread, with chunk-local
// prefixes of the appended particles and of the split / merge lists.
__global__ __launch_bounds__(kKeyChunk) void k_ip_plan_local(ResArgs a, const int32_t* counts,
                                                             const int32_t* targets,
                                                             const int32_t* cand, int32_t* plan,
                                                             int32_t* tmp, long long* chunk,
                                                             int kstride) {
  __shared__ long long s[3][16];
  const bz_resample_policy_t& P = a.pol;
  const int k = blockIdx.x * kKeyChunk + threadIdx.x;
  int kind = kCopy, napp = 0;
  if (k < a.n_keys && cand[a.n_keys]) {
    const int kc = cand[k];
    if (kc == 1) {
      const int draws = (int)(P.split_to * (double)targets[k]) - counts[k];
      if (plan[4 * k + 3] > 0 && draws > 0) {
        kind = kSplit;
        napp = draws;
      }
    } else if (kc == 2) {
      kind = kMerge;
      napp = (int)(P.merge_keep * (double)targets[k]);
    }
  }
  long long v[3] = {(long long)napp, kind == kSplit ? 1ll : 0ll, kind == kMerge ? 1ll : 0ll};
  long long tot[3];
  block_scan1024<3>(v, tot, s);
  if (k < a.n_keys) {
    plan[4 * k + 0] = kind;
    plan[4 * k + 1] = napp;
    plan[4 * k + 2] = (int32_t)v[0];
    tmp[k] = (int32_t)(kind == kSplit ? v[1] : (kind == kMerge ? v[2] : -1));
  }
  if (threadIdx.x == 0) {
    chunk[8 * blockIdx.x + 1] = tot[0];
    chunk[8 * blockIdx.x + 2] = tot[1];
    chunk[8 * blockIdx.x + 3] = tot[2];
  }
}

// One block: chunk prefixes of appended / split / merge, the list lengths and
// the counters; the new slot count is committed here (clipped to the
// capacity, the excess counted as overflow), the count the appends start
// from kept in the chunk words for k_ip_plan_apply.  No kernel of the pass
// after this one reads the slot count.
__global__ __launch_bounds__(64) void k_ip_plan_carry(ResArgs a, int nch, const int32_t* cand,
                                                      long long* chunk, Clock* clk,
                                                      int32_t* klist, int64_t cap) {
  if (threadIdx.x != 0) return;
  long long app = 0, ns = 0, nm = 0;
  for (int b = 0; b < nch; ++b) {
    const long long ta = chunk[8 * b + 1], ts = chunk[8 * b + 2], tm = chunk[8 * b + 3];
    chunk[8 * b + 5] = app;
    chunk[8 * b + 6] = ns;
    chunk[8 * b + 7] = nm;
    app += ta;
    ns += ts;
    nm += tm;
  }
  klist[0] = (int32_t)ns;
  klist[1] = (int32_t)nm;
  const int64_t n0 = clk->n;
  chunk[8 * kMaxKeyChunks + 1] = n0;
  int64_t n1 = n0 + app;
  if (n1 > cap) {
    clk->cnt[BZ_CNT_OVERFLOW] += (unsigned long long)(n1 - cap);
    n1 = cap;
  }
  clk->n = n1;
  clk->n_next = n1;
  if (ns) clk->cnt[BZ_CNT_SPLIT] += (unsigned long long)ns;
  if (nm) clk->cnt[BZ_CNT_MERGE] += (unsigned long long)nm;
  if (cand[a.n_keys]) clk->cnt[BZ_CNT_RESAMPLE] += 1ull;
}

// The first appended slot of every key (after the slots in use before the
// pass, in key order) and the key lists in key order.  The pass's key
// histogram is read for the last time before this kernel: it is cleared
// here for the next pass.
__global__ __launch_bounds__(kKeyChunk) void k_ip_plan_apply(ResArgs a, int32_t* plan,
                                                             const int32_t* tmp,
                                                             const long long* chunk,
                                                             int32_t* klist, int kstride,
                                                             int32_t* hcnt) {
  const int k = blockIdx.x * kKeyChunk + threadIdx.x;
  if (blockIdx.x == 0 && threadIdx.x == 0) hcnt[a.n_keys] = 0;  // the holes' key
  if (k >= a.n_keys) return;
  hcnt[k] = 0;
  const long long* c = chunk + 8 * blockIdx.x;
  plan[4 * k + 2] = (int32_t)(chunk[8 * kMaxKeyChunks + 1] + c[5] + plan[4 * k + 2]);
  const int kind = plan[4 * k + 0];
  if (kind == kSplit) klist[2 + c[6] + tmp[k]] = k;
  else if (kind == kMerge) klist[2 + kstride + c[7] + tmp[k]] = k;
}

template <typename T>
static int launch_resample_t(const LaunchEnv& env, const bz_state_t* in, const bz_state_t* out,
                             const bz_params_t* p, const int32_t* counts, const int32_t* offsets,
                             const uint32_t* perm, const int32_t* targets, int32_t n_keys,
                             const bz_resample_policy_t* pol, const uint32_t* skeys_in,
                             int32_t** plan_out) {
  SortWs ws = carve(env.sort_ws, env.capacity, env.max_keys);
  // keys per cell: nb bins per node (1-D), or one key per instance (grouped)
  const int nb = skeys_in ? 1 : n_keys / (p->dim == 1 ? p->n_nodes : 1);
  if (!skeys_in && (nb < 1 || nb * (p->dim == 1 ? p->n_nodes : 1) != n_keys))
    return fail("resample: n_keys must be n_nodes * nb");
  // the final pass wrote the sorted keys to keys_a (two passes) or keys_b (one)
  const uint32_t* skeys = skeys_in ? skeys_in : ((n_keys + 1) <= 256 ? ws.keys_b : ws.keys_a);
  if (plan_out) *plan_out = ws.plan;
  ResArgs a;
  const void* iv[5] = {in->x, in->vx, in->vy, in->vz, in->w};
  void* ov[5] = {out->x, out->vx, out->vy, out->vz, out->w};
  for (int f = 0; f < 5; ++f) {
    a.in[f] = iv[f];
    a.out[f] = ov[f];
  }
  a.pol = *pol;
  a.kev = p->kev;
  a.seed = p->seed;
  a.n_keys = n_keys;
  a.prec = env.precision;
  a.U = pol->uniforms;
  a.u_stride = pol->u_stride;
  const int64_t cap = env.capacity;
  timing_begin(env, 2);
  hipLaunchKernelGGL(k_res_hw<T>, dim3(n_keys), dim3(256), 0, env.stream, a, counts, offsets,
                     perm, targets, ws.plan);
  hipLaunchKernelGGL(k_res_plan, dim3(1), dim3(1024), 0, env.stream, a, counts, targets, ws.plan,
                     nb, env.clock, ws.klist, ws.kstride);
  const int64_t cgrid = (cap + 256 * kCopyItems - 1) / (256 * kCopyItems);
  timing_begin(env, 3);
  hipLaunchKernelGGL(k_res_copy<T>, dim3((unsigned)(cgrid > 0 ? cgrid : 1)), dim3(256), 0,
                     env.stream, a, skeys, offsets, perm, (const int32_t*)ws.plan, counts,
                     (const Clock*)env.clock, cap);
  timing_end(env, 3);
  // at most n_keys listed keys; blocks beyond the list exit at once
  const int sgrid = n_keys < 2 * env.num_cu ? n_keys : 2 * env.num_cu;
  const int mgrid = n_keys < 2 * env.num_cu ? n_keys : 2 * env.num_cu;
  hipLaunchKernelGGL((k_res_split<T, false>), dim3(sgrid), dim3(kSegBlock), (size_t)kSplitLds * 16,
                     env.stream, a, counts, offsets, perm, (const int32_t*)ws.plan, ws,
                     (const Clock*)env.clock, cap);
  hipLaunchKernelGGL((k_res_merge<T, false>), dim3(mgrid), dim3(kSegBlock), (size_t)kMergeLds * 8,
                     env.stream, a, counts, offsets, perm, targets, (const int32_t*)ws.plan, ws,
                     env.clock, cap);
  hipLaunchKernelGGL(k_res_commit, dim3(1), dim3(1), 0, env.stream, env.clock, cap);
  timing_end(env, 2);
  return check_launch("resample");
}

// Stable sort of the extracted members (ws.mkeys, ws.mslot; count in
// ws.ip_mclk) by key, then the member offsets and counts of every key.
static int sort_members(const LaunchEnv& env, const SortWs& ws, int n_keys) {
  const int passes = (n_keys + 1) <= 256 ? 1 : 2;
  uint32_t* kin = ws.mkeys;
  uint32_t* vin = ws.mslot;
  for (int ps = 0; ps < passes; ++ps) {
    uint32_t* kout = ps == 0 ? ws.mkeys2 : ws.mkeys;
    uint32_t* vout = ps == 0 ? ws.mslot2 : ws.mslot;
    radix_pass(env, ws, ws.ip_mclk, kin, vin, kout, vout, 8 * ps, 0);
    const int rc = check_launch("member radix pass");
    if (rc) return rc;
    kin = kout;
    vin = vout;
  }
  // one pass leaves the members in the second buffers
  if (passes == 1) {
    hipMemcpyAsync(ws.mkeys, ws.mkeys2, 4 * (size_t)env.capacity, hipMemcpyDeviceToDevice,
                   env.stream);
    hipMemcpyAsync(ws.mslot, ws.mslot2, 4 * (size_t)env.capacity, hipMemcpyDeviceToDevice,
                   env.stream);
  }
  return 0;
}

template <typename T, typename K>
static int launch_resample_inplace_t(const LaunchEnv& env, const bz_state_t* s,
                                     const bz_params_t* p, const K* keys, int nb,
                                     const int32_t* targets, int32_t n_keys,
                                     const bz_resample_policy_t* pol) {
  SortWs ws = carve(env.sort_ws, env.capacity, env.max_keys);
  const int64_t cap = env.capacity;
  if (cap > 0x7FFFFFFFll) return fail("resample in place: capacity above 2^31 slots");
  ResArgs a;
  const void* iv[5] = {s->x, s->vx, s->vy, s->vz, s->w};
  for (int f = 0; f < 5; ++f) {
    a.in[f] = iv[f];
    a.out[f] = (void*)iv[f];  // in place
  }
  a.pol = *pol;
  a.kev = p->kev;
  a.seed = p->seed;
  a.n_keys = n_keys;
  a.prec = env.precision;
  a.U = pol->uniforms;
  a.u_stride = pol->u_stride;
  timing_begin(env, 2);
  // slots per key
  const int m = n_keys + 1;
  const int lds_keys = m <= 40960 ? m : 0;
  if ((size_t)lds_keys * 4 > 64 * 1024 &&
      hipFuncSetAttribute((const void*)k_ip_hist<K>, hipFuncAttributeMaxDynamicSharedMemorySize,
                          lds_keys * 4) != hipSuccess)
    return fail("hipFuncSetAttribute(key histogram LDS) failed");
  const int nch = (n_keys + kKeyChunk - 1) / kKeyChunk;
  if (nch > kMaxKeyChunks) return fail("resample: at most %d keys", kMaxKeyChunks * kKeyChunk);
  // The histogram and the fire word the candidate kernel raises start at
  // zero: the workspace is cleared when the context is made, and every pass
  // clears them again (k_ip_plan_apply, k_ip_cand_carry), also one that
  // stops on a launch error (bail)
  auto bail = [&](int rc) {
    hipLaunchKernelGGL(k_zero_i32, dim3(64), dim3(256), 0, env.stream, ws.ip_hcnt, (int64_t)m);
    return rc;
  };
  timing_begin(env, 6);
  hipLaunchKernelGGL(k_ip_hist<K>, dim3(env.num_cu), dim3(kIpHistBlock), (size_t)lds_keys * 4,
                     env.stream, keys, (const Clock*)env.clock, ws.ip_hcnt, lds_keys);
  timing_end(env, 6);
  // candidates, trigger and member offsets (one key per thread, 1024-key chunks)
  hipLaunchKernelGGL(k_ip_cand, dim3(nch), dim3(kKeyChunk), 0, env.stream, a,
                     (const int32_t*)ws.ip_hcnt, targets, nb, ws.ip_cand, ws.ip_moff,
                     ws.ip_chunk, ws.ip_bits);
  hipLaunchKernelGGL(k_ip_cand_carry, dim3(1), dim3(64), 0, env.stream, a, nch, ws.ip_cand,
                     ws.ip_chunk, ws.ip_mclk);
  // members of the candidate segments, in slot order, then grouped by key
  const size_t bits_lds = 4 * ((size_t)n_keys / 32 + 1);
  hipLaunchKernelGGL(k_ip_wcount<K>, dim3(kIpBlocks), dim3(256), bits_lds, env.stream, keys,
                     (const Clock*)env.clock, (const int32_t*)ws.ip_cand, n_keys,
                     (const uint32_t*)ws.ip_bits, ws.ip_bcnt);
  hipLaunchKernelGGL(k_ip_wscan, dim3(1), dim3(1024), 0, env.stream, ws.ip_bcnt);
  hipLaunchKernelGGL(k_ip_wscatter<K>, dim3(kIpBlocks), dim3(256), bits_lds, env.stream, keys,
                     (const Clock*)env.clock, (const int32_t*)ws.ip_cand, n_keys,
                     (const uint32_t*)ws.ip_bits, (const uint32_t*)ws.ip_bcnt, ws.mkeys,
                     ws.mslot);
  int rc = check_launch("member extraction");
  if (rc) return bail(rc);
  rc = sort_members(env, ws, n_keys);
  if (rc) return bail(rc);
  hipLaunchKernelGGL(k_ip_hw<T>, dim3((n_keys + 3) / 4), dim3(256), 0, env.stream, a,
                     (const int32_t*)ws.ip_hcnt, (const int32_t*)ws.ip_cand,
                     (const long long*)ws.ip_chunk, ws.ip_moff, ws.ip_mcnt,
                     (const uint32_t*)ws.mslot, ws.plan);
  hipLaunchKernelGGL(k_ip_plan_local, dim3(nch), dim3(kKeyChunk), 0, env.stream, a,
                     (const int32_t*)ws.ip_hcnt, targets, (const int32_t*)ws.ip_cand, ws.plan,
                     ws.ip_tmp, ws.ip_chunk, ws.kstride);
  hipLaunchKernelGGL(k_ip_plan_carry, dim3(1), dim3(64), 0, env.stream, a, nch,
                     (const int32_t*)ws.ip_cand, ws.ip_chunk, env.clock, ws.klist, cap);
  hipLaunchKernelGGL(k_ip_plan_apply, dim3(nch), dim3(kKeyChunk), 0, env.stream, a, ws.plan,
                     (const int32_t*)ws.ip_tmp, (const long long*)ws.ip_chunk, ws.klist,
                     ws.kstride, ws.ip_hcnt);
  // merges and splits together: every segment of a pass resident at once
  // where the list allows (at most 6 blocks per CU; 2 are resident at once)
  const int ggrid = n_keys < 6 * env.num_cu ? n_keys : 6 * env.num_cu;
  timing_begin(env, 5);
  hipLaunchKernelGGL(k_res_segments<T>, dim3(ggrid), dim3(kSegBlock), kSegLdsBytes, env.stream, a,
                     (const int32_t*)ws.ip_mcnt, (const int32_t*)ws.ip_moff,
                     (const uint32_t*)ws.mslot, targets, (const int32_t*)ws.plan, ws, env.clock,
                     cap);
  timing_end(env, 5);
  timing_end(env, 2);
  return check_launch("resample in place");
}

// ---------------------------------------------------------------------------
// In-place compaction by hole filling.  The reference compacts the live
// particles after the push (post_processing_1D.py:26-38) with atomics, which
// leave their order open; here L = the live count, the k-th hole below L (in
// slot order) receives the k-th live particle at or above L (in slot order),
// and the count becomes L.  Live particles below L keep their slots, so the
// traffic is the weights read twice plus 40 B per moved particle, against a
// full copy of the state for a stable filter.  Blocks own the contiguous
// ranges of comp_range; block bb holds slot L.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kCompBlock) void k_fill_count(const T* __restrict__ w,
                                                           const Clock* clk, uint32_t* cnt) {
  __shared__ uint32_t s_red[kCompBlock / kWave];
  int64_t lo, hi;
  comp_range(clk->n, &lo, &hi);
  uint32_t c = 0;
  for (int64_t i = lo + threadIdx.x; i < hi; i += kCompBlock) c += w[i] > (T)0 ? 1u : 0u;
  c = wave_sum(c);
  if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) cnt[blockIdx.x] = s_red[0] + s_red[1] + s_red[2] + s_red[3];
}

// One block of 1024: L, the block bb holding slot L, the live slots of bb
// below L, then the hole offsets of blocks 0..bb and the mover offsets of
// blocks bb..grid-1.  off: [grid] hole offsets, [grid] mover offsets, H, bb,
// the holes of block bb.  Sets n_next = L.
template <typename T>
__global__ __launch_bounds__(1024) void k_fill_plan(const T* __restrict__ w, Clock* clk, int grid,
                                                    const uint32_t* cnt, int64_t* off) {
  __shared__ long long s_a[1024 / kWave], s_b[1024 / kWave];
  __shared__ long long s_tot;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int64_t n = clk->n;
  const int64_t tiles = (n + kCompTile - 1) / kCompTile;
  const int64_t per = (tiles + grid - 1) / grid;
  const int64_t span = per * kCompTile;  // slots per block range
  // L
  long long c = 0;
  for (int b = tid; b < grid; b += 1024) c += cnt[b];
  c = wave_sum(c);
  if (lane == 0) s_a[wid] = c;
  __syncthreads();
  if (tid == 0) {
    long long t = 0;
    for (int j = 0; j < 1024 / kWave; ++j) t += s_a[j];
    s_tot = t;
  }
  __syncthreads();
  const int64_t L = s_tot;
  const int bb = (span > 0 && L < n) ? (int)(L / span) : grid;  // grid: no movers
  // live slots of block bb below L
  long long lb = 0;
  if (bb < grid)
    for (int64_t i = (int64_t)bb * span + tid; i < L; i += 1024) lb += w[i] > (T)0 ? 1 : 0;
  lb = wave_sum(lb);
  __syncthreads();
  if (lane == 0) s_b[wid] = lb;
  __syncthreads();
  long long live_bb_below = 0;
  for (int j = 0; j < 1024 / kWave; ++j) live_bb_below += s_b[j];
  const long long holes_bb = bb < grid ? (L - (int64_t)bb * span) - live_bb_below : 0;
  // offsets: one thread per 2 blocks, then a block scan (grid <= 2048)
  long long h[2], m[2];
  for (int q = 0; q < 2; ++q) {
    const int b = 2 * tid + q;
    const int64_t lo = b < grid ? min((int64_t)b * span, n) : n;
    const int64_t hi = b < grid ? min(lo + span, n) : n;
    h[q] = (b < bb) ? (hi - lo) - (long long)cnt[b] : (b == bb ? holes_bb : 0);
    m[q] = (b > bb && b < grid) ? (long long)cnt[b]
                                : (b == bb ? (long long)cnt[b] - live_bb_below : 0);
  }
  long long hs = h[0] + h[1], ms = m[0] + m[1];
  const long long hi_incl = wave_incl_scan(hs), mi_incl = wave_incl_scan(ms);
  __syncthreads();
  if (lane == kWave - 1) {
    s_a[wid] = hi_incl;
    s_b[wid] = mi_incl;
  }
  __syncthreads();
  long long hb = 0, mb = 0, htot = 0;
  for (int j = 0; j < 1024 / kWave; ++j) {
    if (j < wid) {
      hb += s_a[j];
      mb += s_b[j];
    }
    htot += s_a[j];
  }
  const long long h0 = hb + hi_incl - hs, m0 = mb + mi_incl - ms;
  for (int q = 0; q < 2; ++q) {
    const int b = 2 * tid + q;
    if (b < grid) {
      off[b] = q == 0 ? h0 : h0 + h[0];
      off[grid + b] = q == 0 ? m0 : m0 + m[0];
    }
  }
  if (tid == 0) {
    off[2 * grid] = htot;
    off[2 * grid + 1] = bb;
    off[2 * grid + 2] = holes_bb;
    clk->n_next = L;
  }
}

// The holes below L and the movers at or above L of each block, ranked in
// slot order: hole_pos[k] / mover_pos[k] = the slot of the k-th one.
template <typename T>
__global__ __launch_bounds__(kCompBlock) void k_fill_rank(const T* __restrict__ w,
                                                          const Clock* clk, const int64_t* off,
                                                          uint32_t* hole_pos,
                                                          uint32_t* mover_pos) {
  constexpr int NW = kCompBlock / kWave;
  __shared__ uint64_t s_wsum[NW];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int grid = (int)gridDim.x;
  const int b = (int)blockIdx.x;
  const int bb = (int)off[2 * grid + 1];
  const int64_t L = clk->n_next;
  int64_t lo, hi;
  comp_range(clk->n, &lo, &hi);
  const int64_t hole0 = b <= bb ? off[b] : 0;
  const int64_t mover0 = b >= bb && b < grid ? off[grid + b] : 0;
  const int64_t holes_here = b == bb ? off[2 * grid + 2] : 0;
  int64_t seen = 0;  // flagged slots of this block before the sub-tile
  for (int64_t base = lo; base < hi; base += kCompTile) {
    uint64_t packed = 0;
#pragma unroll
    for (int k = 0; k < kCompItems; ++k) {
      const int64_t i = base + k * kCompBlock + tid;
      bool f = false;
      if (i < hi) f = i < L ? !(w[i] > (T)0) : (w[i] > (T)0);
      packed |= (uint64_t)(f ? 1u : 0u) << (16 * k);
    }
    const uint64_t incl = wave_incl_scan(packed);
    if (lane == kWave - 1) s_wsum[wid] = incl;
    __syncthreads();
    uint64_t before = 0, tot = 0;
#pragma unroll
    for (int q = 0; q < NW; ++q) {
      if (q < wid) before += s_wsum[q];
      tot += s_wsum[q];
    }
    const uint64_t excl = before + incl - packed;
    int64_t field_base = 0;
#pragma unroll
    for (int k = 0; k < kCompItems; ++k) {
      if ((packed >> (16 * k)) & 1u) {
        const int64_t i = base + k * kCompBlock + tid;
        const int64_t r = seen + field_base + (int64_t)((excl >> (16 * k)) & 0xFFFFu);
        if (i < L) hole_pos[hole0 + r] = (uint32_t)i;
        else mover_pos[mover0 + (r - holes_here)] = (uint32_t)i;
      }
      field_base += (int64_t)((tot >> (16 * k)) & 0xFFFFu);
    }
    seen += field_base;
    __syncthreads();  // s_wsum is rewritten by the next sub-tile
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_fill_move(bz_state_t s, const int64_t* off, int grid,
                                                   const uint32_t* hole_pos,
                                                   const uint32_t* mover_pos) {
  const int64_t H = off[2 * grid];
  T* f[5] = {(T*)s.x, (T*)s.vx, (T*)s.vy, (T*)s.vz, (T*)s.w};
  for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < H;
       k += (int64_t)gridDim.x * 256) {
    const uint32_t d = hole_pos[k], src = mover_pos[k];
#pragma unroll
    for (int q = 0; q < 5; ++q) f[q][d] = f[q][src];
  }
}

template <typename T>
static int launch_fill_holes_t(const LaunchEnv& env, const bz_state_t* s) {
  SortWs ws = carve(env.sort_ws, env.capacity, env.max_keys);
  // the block counts [grid] and offsets [2*grid + 3] in the tile histogram
  // buffer (4*256 B per tile of the capacity)
  const int64_t bytes = (int64_t)4 * 256 * ws.max_tiles;
  int64_t fit = (bytes - 64) / 20;
  int grid = env.num_cu * 8;
  if (grid > 2048) grid = 2048;
  if (grid > fit) grid = (int)fit;
  if (grid < 1) return fail("hole filling: workspace too small");
  uint32_t* cnt = ws.hist;
  int64_t* off = (int64_t*)(ws.hist + ((grid + 1) & ~1));
  hipLaunchKernelGGL(k_fill_count<T>, dim3(grid), dim3(kCompBlock), 0, env.stream,
                     (const T*)s->w, (const Clock*)env.clock, cnt);
  hipLaunchKernelGGL(k_fill_plan<T>, dim3(1), dim3(1024), 0, env.stream, (const T*)s->w,
                     env.clock, grid, (const uint32_t*)cnt, off);
  hipLaunchKernelGGL(k_fill_rank<T>, dim3(grid), dim3(kCompBlock), 0, env.stream,
                     (const T*)s->w, (const Clock*)env.clock, (const int64_t*)off, ws.mkeys,
                     ws.mslot);
  timing_begin(env, 7);
  hipLaunchKernelGGL(k_fill_move<T>, dim3(env.num_cu * 4), dim3(256), 0, env.stream, *s,
                     (const int64_t*)off, grid, (const uint32_t*)ws.mkeys,
                     (const uint32_t*)ws.mslot);
  timing_end(env, 7);
  hipLaunchKernelGGL(k_res_commit, dim3(1), dim3(1), 0, env.stream, env.clock, env.capacity);
  return check_launch("hole filling");
}

int launch_fill_holes(const LaunchEnv& env, const bz_state_t* s) {
  return env.precision == BZ_F64 ? launch_fill_holes_t<double>(env, s)
                                 : launch_fill_holes_t<float>(env, s);
}

int launch_resample_inplace(const LaunchEnv& env, const bz_state_t* s, const bz_params_t* p,
                            const void* keys, int keys16, int nb, const int32_t* targets,
                            int32_t n_keys, const bz_resample_policy_t* pol) {
  if (n_keys + 1 > 65535) return fail("resample: at most 65534 keys");
  const uint16_t* k16 = (const uint16_t*)keys;
  const uint32_t* k32 = (const uint32_t*)keys;
  if (env.precision == BZ_F64)
    return keys16 ? launch_resample_inplace_t<double>(env, s, p, k16, nb, targets, n_keys, pol)
                  : launch_resample_inplace_t<double>(env, s, p, k32, nb, targets, n_keys, pol);
  return keys16 ? launch_resample_inplace_t<float>(env, s, p, k16, nb, targets, n_keys, pol)
                : launch_resample_inplace_t<float>(env, s, p, k32, nb, targets, n_keys, pol);
}

// the slot keys of a state (k_keys) into the sort input, for a caller whose
// state has no keys from the push
int launch_slot_keys(const LaunchEnv& env, const bz_state_t* s, const bz_params_t* p,
                     const double* thr, int nb) {
  SortWs ws = carve(env.sort_ws, env.capacity, env.max_keys);
  const KeyArgs ka = key_args(s, p, thr, nb);
  auto kk = env.precision == BZ_F64 ? k_keys<double> : k_keys<float>;
  const int64_t kgrid = (env.capacity + (int64_t)kSortBlock * kKeyItems - 1) /
                        ((int64_t)kSortBlock * kKeyItems);
  hipLaunchKernelGGL(kk, dim3((unsigned)(kgrid > 0 ? kgrid : 1)), dim3(kSortBlock), 0,
                     env.stream, ka, (const Clock*)env.clock, ws.keys_a);
  return check_launch("k_keys");
}

int launch_resample(const LaunchEnv& env, const bz_state_t* in, const bz_state_t* out,
                    const bz_params_t* p, const int32_t* counts, const int32_t* offsets,
                    const uint32_t* perm, const int32_t* targets, int32_t n_keys,
                    const bz_resample_policy_t* pol) {
  if (env.precision == BZ_F64)
    return launch_resample_t<double>(env, in, out, p, counts, offsets, perm, targets, n_keys,
                                     pol, nullptr, nullptr);
  return launch_resample_t<float>(env, in, out, p, counts, offsets, perm, targets, n_keys, pol,
                                  nullptr, nullptr);
}

int launch_resample_grouped(const LaunchEnv& env, const bz_state_t* in, const bz_state_t* out,
                            const bz_params_t* p, const int32_t* counts, const int32_t* offsets,
                            const uint32_t* perm, const uint32_t* skeys, const int32_t* targets,
                            int32_t n_keys, const bz_resample_policy_t* pol, int32_t** plan_out) {
  if (env.precision == BZ_F64)
    return launch_resample_t<double>(env, in, out, p, counts, offsets, perm, targets, n_keys,
                                     pol, skeys, plan_out);
  return launch_resample_t<float>(env, in, out, p, counts, offsets, perm, targets, n_keys, pol,
                                  skeys, plan_out);
}

void* sort_input_keys(const LaunchEnv& env) {
  return carve(env.sort_ws, env.capacity, env.max_keys).keys_a;
}

void* sort_scratch_keys(const LaunchEnv& env) {
  return carve(env.sort_ws, env.capacity, env.max_keys).keys_b;
}

void* sort_scratch_perm(const LaunchEnv& env) {
  return carve(env.sort_ws, env.capacity, env.max_keys).vals_b;
}

}  // namespace bz
