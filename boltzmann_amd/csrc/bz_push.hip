// Fused electron step for gfx950: MCC + push + walls + node deposit, in one
// pass over the SoA state, with births and the removal of absorbed particles
// handled in one of two ways.
//
// 1-D discharges and batched 0-D swarms (in place; for the discharges the
// reference's electron_kernel layout):
//   each block owns a contiguous run of tiles of BLOCK*ITEMS particles
//   HBM read  x, vx, vy, vz, w            (5 words / particle)
//   registers collision, push, walls, secondary
//   LDS       per-block FP64 node histograms (ne, ni_src, nrg), or the
//             per-instance moments of the swarms (x holds the instance)
//   HBM write x, vx, vy, vz back to the same slot (swarms: vx, vy, vz);
//             w = 0 for an absorbed particle (w of a survivor is unchanged
//             and not rewritten)
//   births    compacted per block into a staging buffer, then appended
//             after the live range by k_push_post in parent order, in the
//             same launch that sums the partial rows and the counters
//   No block waits on another, so any grid size and residency works.
//   Absorbed slots stay as holes (w = 0) until a compaction or a resampling
//   pass rewrites the state.
// The single 0-D swarm (compacting):
//   scan      per-particle output count (alive + born) -> block scan ->
//             decoupled look-back across tiles -> output slot
//   HBM write survivors, each secondary right after its parent
// Histograms are flushed once per block to a partial buffer and summed over
// blocks in a fixed order (k_push_post / k_dep_reduce), so the moments do not
// depend on scheduling across blocks.
#include "bz_physics.h"

namespace bz {

struct PushArgs {
  bz_params_t p;
  bz_xsec_t xs;
  const void* in[5];
  void* out[5];
  const void* E;
  const void* R;
  const void* G;
  Clock* clk;
  uint64_t* status;
  double* partials;  // [gridDim][3][n_nodes] doubles, then [gridDim][8] u32 counters
  int64_t capacity;
  int sched;         // 0 static tile order (cooperative launch), 1 ticket counter
  // batched 0-D swarms (DIM 2): instance k owns particles [off[k], off[k+1])
  const int64_t* inst_off;
  const void* inst_E;  // [n_inst] field of each instance, V/cm
  int n_inst;
  uint32_t* block_births;  // in-place mode: births of each block
  int queue_mode;          // 1-D push candidate queue: 1 always, 0 where walks are long
  int e_lds_off;           // 1-D in-place push: LDS byte offset of the staged field (0: global)
  // 1-D push: binning key of every slot it leaves (NULL: none), the input
  // of the next bin sort
  uint32_t* keys;
  int keys16;  // keys written as 16-bit words
  BinKeys bk;
  // queued 1-D push: per wave a ring of kQCap candidates' states [5][kQCap]
  // (x, vx, vy, vz, w in the state type), so a batch reads its particles
  // from a small buffer the wave keeps hot instead of the state
  void* qbuf;
  // 1-D: the ionization moment's global fixed-point words (k_push_post
  // reads and clears them)
  unsigned long long* m1_lo;
  uint32_t* m1_hi;
  // 1-D collision counters (NULL: off): per node sum w of the elastic and of
  // the excitation collisions (fixed point, carry words) and the number of
  // ionizations; k_push_post folds them into coll_acc [3][n_nodes]
  unsigned long long* coll_lo;
  uint32_t* coll_hi;
  unsigned long long* coll_cnt;
  double* coll_acc;
};

// The binning key of slot i into the sort input (32- or 16-bit words)
__device__ __forceinline__ void put_key(const PushArgs& a, int64_t i, uint32_t k) {
  if (a.keys16) ((uint16_t*)a.keys)[i] = (uint16_t)k;
  else a.keys[i] = k;
}

// A collision of the 1-D push into the counters: at the primary's node after
// the push, node 0 when it left the gap (the reference's collct rows,
// electron_kernel_numba.py:199,232, are reduced at the post-push position,
// x = 1e-8 for an absorbed particle, post_processing_1D.py:56); elastic and
// excitation events add the weight each (free-flight mode: every event of
// the step), an ionization counts 1.
template <typename T>
__device__ __forceinline__ void coll_add(const PushArgs& a, int nn, const StepOut<T>& o,
                                         const bz_params_t& p, double w, double scale) {
  if (!(o.n_el || o.n_exc || o.chan == kChanIon)) return;
  const int nd = o.alive ? dep_node((double)o.p.x, p) : 0;
  if (o.n_el) dep_fixed_add_global(a.coll_lo, a.coll_hi, nd, w * (double)o.n_el, scale);
  if (o.n_exc)
    dep_fixed_add_global(a.coll_lo + nn, a.coll_hi + nn, nd, w * (double)o.n_exc, scale);
  if (o.chan == kChanIon) atomicAdd(&a.coll_cnt[nd], 1ull);
}

// per-block counter row of the push (u32): births, absorbed, elastic,
// excitation, ionization, null, dropped, nu_max exceeded, live inputs
constexpr int kCntRow = 10;
// 1-D push: per-block birth counts, then per-(tile, wave) counts, in the
// context's status words
constexpr int kMaxPushBlocks = 4096;

// per-instance moments of a batched step (FP64, per block then summed)
enum : int {
  kImW = 0,     // sum w
  kImWE = 1,    // sum w * energy
  kImWVX = 2,   // sum w * vx
  kImEl = 3,    // weight of elastic events
  kImExc = 4,   // weight of excitation events
  kImIon = 5,   // weight of ionization events
  kImBirth = 6, // number of births
  kImWE2 = 7,   // sum w * energy^2
  kImN = 8
};

// 0-D swarms carry no position: x is neither read nor written (WITHX false)
template <typename T, bool WITHX>
__device__ __forceinline__ void store5(void* const* out, int64_t pos, const P5<T>& q) {
  if (WITHX) ((T*)out[0])[pos] = q.x;
  ((T*)out[1])[pos] = q.vx;
  ((T*)out[2])[pos] = q.vy;
  ((T*)out[3])[pos] = q.vz;
  ((T*)out[4])[pos] = q.w;
}

template <typename T, bool WITHX>
__device__ __forceinline__ void load5(const T* const* in, int64_t i, int64_t n, P5<T>& q) {
  if (i < n) {
    q.x = WITHX ? in[0][i] : (T)0;
    q.vx = in[1][i];
    q.vy = in[2][i];
    q.vz = in[3][i];
    q.w = in[4][i];
  } else {
    q.x = q.vx = q.vy = q.vz = q.w = (T)0;
  }
}

// LDS bytes of the staged cross-section table (energy, 3 sigma, 3 slopes, hints)
__host__ __device__ inline size_t table_lds_bytes(int n, int n_hint, int word) {
  return ((size_t)7 * n * word + (size_t)n_hint * 4 + 15) & ~(size_t)15;
}

// 1-D push: the table follows the [3][nn] histograms at a 16-byte boundary
__host__ __device__ inline int table_lds_offset_doubles(int nn) { return (3 * nn + 1) & ~1; }

// Copy m words global -> LDS with 16-byte accesses where both ends allow
// (the device tables are 256-byte aligned allocations; dst is 16-byte
// aligned by layout), scalar for the rest.
template <typename T, int BLOCK>
__device__ __forceinline__ void stage_lds(T* dst, const T* __restrict__ src, int m) {
  constexpr int V = 16 / sizeof(T);
  const bool vec = ((((uintptr_t)dst) | ((uintptr_t)src)) & 15) == 0;
  int done = 0;
  if (vec) {
    const int nv = m / V;
    const int4* s4 = (const int4*)src;
    int4* d4 = (int4*)dst;
    for (int j = threadIdx.x; j < nv; j += BLOCK) d4[j] = s4[j];
    done = nv * V;
  }
  for (int j = done + threadIdx.x; j < m; j += BLOCK) dst[j] = src[j];
}

template <typename T, int XS, int DIM, int RNG, int BLOCK, int ITEMS>
__global__ __launch_bounds__(BLOCK) void k_push(PushArgs a) {
  static_assert(DIM == 0, "the compacting push serves the single 0-D swarm");
  static_assert(ITEMS >= 1 && ITEMS <= 4, "packed scan holds 4 fields of 16 bits");
  constexpr int NW = BLOCK / kWave;
  constexpr int TILE = BLOCK * ITEMS;
  constexpr int LB = 1024 / BLOCK > 0 ? 1024 / BLOCK : 1;  // look-back window 1024 tiles
  constexpr bool WITHX = false;
  // dynamic LDS: the three sums, then the table
  extern __shared__ double s_dep[];
  __shared__ uint64_t s_wsum[NW];
  __shared__ uint32_t s_lbmin;
  __shared__ uint32_t s_lbred[NW];
  __shared__ uint32_t s_next;
  __shared__ unsigned long long s_cnt[kCntRow];

  const bz_params_t& p = a.p;
  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1);
  const int wid = tid / kWave;
  const int dep_words = 3;
  if (tid < kCntRow) s_cnt[tid] = 0ull;
  // stage the cross-section table in LDS (read-only for the whole launch)
  bz_xsec_t xs = a.xs;
  if (XS != BZ_XS_FIT) {
    const int n = a.xs.n;
    T* s_tab = (T*)(s_dep + dep_words);
    int32_t* s_hint = (int32_t*)(s_tab + 7 * n);
    const T* ge = (const T*)a.xs.energy;
    const T* gs = (const T*)a.xs.sig;
    const T* gd = (const T*)a.xs.slope;
    for (int j = tid; j < n; j += BLOCK) s_tab[j] = ge[j];
    for (int j = tid; j < 3 * n; j += BLOCK) {
      s_tab[n + j] = gs[j];
      s_tab[4 * n + j] = gd[j];
    }
    for (int j = tid; j < a.xs.n_hint; j += BLOCK) s_hint[j] = a.xs.hint[j];
    xs.energy = s_tab;
    xs.sig = s_tab + n;
    xs.slope = s_tab + 4 * n;
    xs.hint = s_hint;
  }

  const int64_t n = a.clk->n;
  const uint64_t step = a.clk->step;
  const uint32_t epoch = a.clk->epoch;
  const uint32_t ntiles = (uint32_t)((n + TILE - 1) / TILE);
  if (blockIdx.x == 0 && tid == 0 && n == 0) a.clk->n_next = 0;

  const T* in[5] = {(const T*)a.in[0], (const T*)a.in[1], (const T*)a.in[2], (const T*)a.in[3],
                    (const T*)a.in[4]};
  const T* __restrict__ E = (const T*)a.E;
  const T* R = (const T*)a.R;
  const T* G = (const T*)a.G;
  const T e_uniform = (DIM == 0) ? (E ? E[0] : (T)p.e_field0) : (T)0;

  uint32_t c_el = 0, c_ex = 0, c_io = 0, c_abs = 0, c_null = 0, c_over = 0, c_drop = 0;
  uint32_t c_live = 0;
  double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0;  // DIM 0 deposit in registers

  // Tile schedule.  sched 0: static order b, b+G, ... (cooperative launch, so
  // every block is resident).  sched 1: tiles claimed from a ticket counter by
  // running blocks (any residency).  Either way a tile's look-back only waits
  // on tiles owned by blocks that are running, so every wait ends.
  if (tid == 0) s_next = a.sched ? atomicAdd(&a.clk->ticket, 1u) : blockIdx.x;
  __syncthreads();
  uint32_t t = s_next;
  __syncthreads();  // s_next is rewritten at the top of the loop
  P5<T> cur[ITEMS];
#pragma unroll
  for (int k = 0; k < ITEMS; ++k)
    load5<T, WITHX>(in, (int64_t)t * TILE + k * BLOCK + tid, n, cur[k]);

  while (t < ntiles) {
    const int64_t base = (int64_t)t * TILE;
    // claim the next tile now; it is read after the scan barriers below
    if (tid == 0) s_next = a.sched ? atomicAdd(&a.clk->ticket, 1u) : t + gridDim.x;

    StepOut<T> o[ITEMS];
    uint64_t packed = 0;
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
      const int64_t i = base + (int64_t)k * BLOCK + tid;
      const T e_local = e_uniform;
      if (XS == BZ_XS_FLIGHT)
        o[k] = flight_particle<T, DIM>(cur[k], (uint32_t)i, step, p, xs, E, e_local);
      else
        o[k] = step_particle<T, XS, DIM, RNG>(cur[k], (uint32_t)i, step, p, xs, E, R, G, e_local);
      if (i < n && cur[k].w > (T)0) {
        ++c_live;
        c_el += o[k].n_el;
        c_ex += o[k].n_exc;
        c_io += o[k].chan == kChanIon;
        c_null += o[k].n_null;
        c_abs += !o[k].alive;
        c_over += o[k].over;
      }
      const uint32_t cnt = (uint32_t)(o[k].alive + o[k].born);
      packed |= (uint64_t)cnt << (16 * k);
      // deposit: survivors at their new position, secondaries where born
      if (o[k].alive) {
        const double w = (double)o[k].p.w;
        acc0 += w;
        acc1 += (o[k].chan == kChanIon) ? w : 0.0;
        acc2 += w * (double)o[k].e_post;
      }
      if (o[k].born) {
        const double w = (double)o[k].s.w;
        acc0 += w;
        acc2 += w * (double)o[k].e_sec;
      }
    }

    // block exclusive scan of the packed counts, in index order k-major
    const uint64_t incl = wave_incl_scan(packed);
    if (lane == kWave - 1) s_wsum[wid] = incl;
    __syncthreads();
    if (wid == 0) {
      uint64_t v = lane < NW ? s_wsum[lane] : 0ull;
      v = wave_incl_scan(v);
      if (lane < NW) s_wsum[lane] = v;
    }
    __syncthreads();
    const uint64_t excl = (wid ? s_wsum[wid - 1] : 0ull) + incl - packed;
    const uint64_t tot = s_wsum[NW - 1];
    const uint32_t t_next = s_next;
    uint32_t field_base[ITEMS];
    uint32_t agg = 0;
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
      field_base[k] = agg;
      agg += (uint32_t)((tot >> (16 * k)) & 0xFFFFu);
    }
    // prefetch the next tile so its HBM latency overlaps the look-back
#pragma unroll
    for (int k = 0; k < ITEMS; ++k)
      load5<T, WITHX>(in, (int64_t)t_next * TILE + k * BLOCK + tid, n, cur[k]);
    const uint32_t pre = lookback_block<BLOCK, LB>(a.status, t, agg, epoch, &s_lbmin, s_lbred);
    if (tid == 0 && t == ntiles - 1) a.clk->n_next = (int64_t)(pre + agg);
    const int64_t tile_pre = pre;
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
      const int64_t pos =
          tile_pre + field_base[k] + (int64_t)((excl >> (16 * k)) & 0xFFFFu);
      if (o[k].alive) {
        if (pos < a.capacity) store5<T, WITHX>(a.out, pos, o[k].p);
        else ++c_drop;
      }
      if (o[k].born) {
        const int64_t ps = pos + o[k].alive;
        if (ps < a.capacity) store5<T, WITHX>(a.out, ps, o[k].s);
        else ++c_drop;
      }
    }
    t = t_next;
    // s_wsum, s_next and the look-back scratch are rewritten only after the
    // look-back's barriers, which every thread has passed
  }

  // counters: wave sums -> LDS -> one row per block, summed by k_dep_reduce
  {
    const uint32_t v[kCntRow] = {c_io, c_abs, c_el, c_ex, c_io, c_null, c_drop, c_over, c_live, 0};
#pragma unroll
    for (int q = 0; q < kCntRow; ++q) {
      const uint32_t s = wave_sum(v[q]);
      if (lane == 0 && s) atomicAdd(&s_cnt[q], (unsigned long long)s);
    }
  }
  {
    // per-wave sums combined in wave order (no LDS atomics: reproducible)
    __shared__ double s_w0[3][NW];
    acc0 = wave_sum(acc0);
    acc1 = wave_sum(acc1);
    acc2 = wave_sum(acc2);
    if (lane == 0) {
      s_w0[0][wid] = acc0;
      s_w0[1][wid] = acc1;
      s_w0[2][wid] = acc2;
    }
    __syncthreads();
    if (tid < 3) {
      double s = 0.0;
      for (int w = 0; w < NW; ++w) s += s_w0[tid][w];
      s_dep[tid] = s;
    }
  }
  __syncthreads();
  const int m = 3;
  double* part = a.partials + (size_t)blockIdx.x * m;
  for (int j = tid; j < m; j += BLOCK) part[j] = s_dep[j];
  uint32_t* bcnt = (uint32_t*)(a.partials + (size_t)gridDim.x * m);
  if (tid < kCntRow) bcnt[(size_t)blockIdx.x * kCntRow + tid] = (uint32_t)s_cnt[tid];
}

// Sum the per-block partial histograms, in two levels inside one launch.
// Block (cx, cy) owns outputs [64 cx, 64 cx + 64) and partial rows
// [cy rpc, (cy + 1) rpc): lane l of wave w sums rows w, w+16, ... of its
// chunk for output l (coalesced across the lanes), the 16 wave sums are added
// in wave order and stored as level-2 row cy.  The last block to finish a
// column cx (an acquire-release counter) adds the level-2 rows in the same
// fixed pattern, so the result does not depend on timing.
constexpr int kRedWaves = 16;
constexpr int kRedThreads = kRedWaves * 64;

// m1 (optional): the 1-D push's global fixed-point ionization words, added
// to outputs [nn, 2 nn) at scale m1_inv and cleared for the next push.
struct M1Fold {
  unsigned long long* lo;
  uint32_t* hi;
  int nn;
  double inv;
};

__device__ __forceinline__ double m1_take(const M1Fold& f, int j) {
  if (!f.lo || j < f.nn || j >= 2 * f.nn) return 0.0;
  const int k = j - f.nn;
  const double v = dep_fixed_value(f.lo[k], f.hi[k], f.inv);
  f.lo[k] = 0ull;
  f.hi[k] = 0u;
  return v;
}

__device__ void dep_reduce_part(const double* __restrict__ partials, int nblocks, int m, int rpc,
                                double* l2, uint32_t* col_done, double* __restrict__ out,
                                double* acc, int cx, int cy, int ny, double (*s_part)[64],
                                int* s_last, const M1Fold& m1 = M1Fold{nullptr, nullptr, 0, 0.0}) {
  const int lane = threadIdx.x & 63;
  const int wid = threadIdx.x >> 6;
  const int j = cx * 64 + lane;
  const int r0 = cy * rpc;
  const int r1 = r0 + rpc < nblocks ? r0 + rpc : nblocks;
  // up to 16 rows per thread with every load issued before the first add,
  // so a 256-row chunk costs one memory round trip
  double sum = 0.0;
  if (j < m) {
    for (int b0 = r0 + wid; b0 < r1; b0 += 16 * kRedWaves) {
      double v[16];
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        const int b = b0 + u * kRedWaves;
        v[u] = b < r1 ? partials[(size_t)b * m + j] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 16; ++u) sum += v[u];
    }
  }
  s_part[wid][lane] = sum;
  __syncthreads();
  if (wid == 0) {
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < kRedWaves; ++w) t += s_part[w][lane];
    if (ny == 1) {
      if (j < m) {
        t += m1_take(m1, j);
        out[j] = t;
        if (acc) acc[j] += t;
      }
    } else if (j < m) {
      l2[(size_t)cy * m + j] = t;
      __threadfence();  // the level-2 row is visible before the column count moves
    }
  }
  if (ny > 1) {
    __syncthreads();
    if (threadIdx.x == 0) {
      const uint32_t before = __hip_atomic_fetch_add(&col_done[cx], 1u, __ATOMIC_ACQ_REL,
                                                     __HIP_MEMORY_SCOPE_AGENT);
      *s_last = before == (uint32_t)(ny - 1);
    }
    __syncthreads();
    if (*s_last) {
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      double t = 0.0;
      if (j < m)
        for (int y = wid; y < ny; y += kRedWaves) t += __hip_atomic_load(
            &l2[(size_t)y * m + j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      s_part[wid][lane] = t;
      __syncthreads();
      if (wid == 0) {
        double u = 0.0;
#pragma unroll
        for (int w = 0; w < kRedWaves; ++w) u += s_part[w][lane];
        if (j < m) {
          u += m1_take(m1, j);
          out[j] = u;
          if (acc) acc[j] += u;
        }
        if (lane == 0) col_done[cx] = 0u;  // ready for the next launch
      }
    }
  }
}

// Counters of the push's per-block rows into the clock (wave q sums counter
// q, all at once) and the slot count commit.  Runs in a block of its own.
__device__ void dep_commit(const double* __restrict__ partials, int nblocks, int m, Clock* clk,
                           int64_t cap, bool set_n, double advance_dt = 0.0) {
  const int lane = threadIdx.x & 63;
  const int q = threadIdx.x >> 6;
  const uint32_t* bcnt = (const uint32_t*)(partials + (size_t)nblocks * m);
  const int slot[9] = {BZ_CNT_BIRTHS, BZ_CNT_ABSORBED, BZ_CNT_ELASTIC, BZ_CNT_EXCITATION,
                       BZ_CNT_IONIZATION, BZ_CNT_NULL, BZ_CNT_OVERFLOW, BZ_CNT_NUMAX,
                       BZ_CNT_PSTEPS};
  static_assert(kRedWaves >= 9, "one wave per counter");
  if (q < 9) {
    unsigned long long s = 0;
    for (int b = lane; b < nblocks; b += 64) s += bcnt[(size_t)b * kCntRow + q];
    s = wave_sum(s);
    if (lane == 0 && s) clk->cnt[slot[q]] += s;
  }
  if (threadIdx.x == 0) {
    if (set_n) {
      // outputs past the capacity were dropped (and counted) by the push
      const int64_t nx = __hip_atomic_load(&clk->n_next, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
      clk->n = nx < cap ? nx : cap;
    }
    clk->epoch = clk->epoch + 1;
    clk->ticket = 0;
    if (advance_dt > 0.0) {
      // a step with no field stage after it (the batched swarms)
      clk->step = clk->step + 1;
      clk->t = clk->t + advance_dt;
    }
  }
}

__global__ __launch_bounds__(kRedThreads) void k_dep_reduce(
    const double* __restrict__ partials, int nblocks, int m, int rpc, double* l2,
    uint32_t* col_done, double* __restrict__ out, double* acc, Clock* clk, int commit,
    int64_t cap) {
  __shared__ double s_part[kRedWaves][64];
  __shared__ int s_last;
  if (commit && blockIdx.x == gridDim.x - 1) {
    // the extra column of blocks: counters and the slot count
    if (blockIdx.y == 0) dep_commit(partials, nblocks, m, clk, cap, true);
    return;
  }
  dep_reduce_part(partials, nblocks, m, rpc, l2, col_done, out, acc, blockIdx.x, blockIdx.y,
                  gridDim.y, s_part, &s_last);
}

// ---------------------------------------------------------------------------
// The push in place, with no barrier in the tile loop: the 1-D discharges
// (DIM 1) and the batched 0-D swarms (DIM 2).
//
// Block b owns tiles [t_first, t_end) (contiguous); within a tile, wave w
// owns the WSPAN = 64*ITEMS consecutive slots starting at t*TILE + w*WSPAN,
// item k of lane l being slot t*TILE + w*WSPAN + k*64 + l.  Every wave walks
// its slots tile after tile on its own: the next tile's particles are loaded
// before this tile's collision physics, survivors are written back to their
// slots, and births are staged in the wave's own slot range of the staging
// buffer in slot order, with the count per (tile, wave) stored for
// k_push_post.  The sums live in LDS (FP64 atomics) and are flushed once
// per block after the loop: node histograms [3][n_nodes] (DIM 1) or
// per-instance moments [n_inst][kImN] (DIM 2).
//
// Swarms: a 0-D particle has no position, so x holds its instance number
// (an exact small integer in either precision); the field of the instance is
// staged in LDS, x and w are never rewritten (no walls in 0-D), and a
// secondary inherits its parent's x, i.e. its instance.
// ---------------------------------------------------------------------------
// Candidate queue of the 1-D push: slots per wave (a batch of 64 runs as soon
// as 64 are queued, so one tile item never overflows it).
constexpr int kQCap = 128;
// tile items per wave below which the push does not queue
constexpr uint32_t kQueueMinItems = 16;

// the in-place push queues its collision candidates (tabulated sigma,
// Philox stream; the swarms use the null-collision table)
template <int XS, int DIM, int RNG>
struct QueuedPush {
  static constexpr bool value =
      (XS == BZ_XS_TABLE || (DIM == 1 && XS == BZ_XS_DIRECT)) && RNG == BZ_RNG_PHILOX;
};

// LDS bytes after the staged table: per-bucket bounds [n_hint] and the
// per-wave candidate queues [NW][kQCap]: slots, then their decision uniforms
__host__ __device__ inline size_t queue_lds_bytes(int n_hint, int word, int nw) {
  return (((size_t)n_hint * word + 15) & ~(size_t)15) + (size_t)nw * kQCap * (4 + word);
}

// The kImN per-instance moments of one swarm particle's step (all zero for a
// lane that is not counted): survivor and secondary weight, energy, energy
// squared and x momentum, the weight of the event's channel, the birth.
template <typename T>
__device__ __forceinline__ void swarm_moments(const StepOut<T>& o, T w_in, bool counted,
                                              double mv[kImN]) {
  const double w = counted ? (double)o.p.w : 0.0;
  const double ep = (double)o.e_post;
  const double ws = counted && o.born ? (double)o.s.w : 0.0;
  const double es = (double)o.e_sec;
  const double wc = counted ? (double)w_in : 0.0;
  mv[kImW] = w + ws;
  mv[kImWE] = w * ep + ws * es;
  mv[kImWVX] = w * (double)o.p.vx + ws * (double)o.s.vx;
  mv[kImEl] = o.chan == kChanElastic ? wc : 0.0;
  mv[kImExc] = o.chan == kChanExc ? wc : 0.0;
  mv[kImIon] = o.chan == kChanIon ? wc : 0.0;
  mv[kImBirth] = counted ? (double)o.born : 0.0;
  mv[kImWE2] = w * ep * ep + ws * es * es;
}

// Adds one item's moments to the wave's per-lane running sums of instance
// m_inst (wave-uniform, -1: none).  When the counted lanes leave m_inst the
// sums are added to the LDS moments first; lanes of several instances (an
// instance boundary, the births) add theirs to LDS directly.
__device__ __forceinline__ void swarm_accumulate(const double mv[kImN], bool counted, int inst,
                                                 int& m_inst, double macc[kImN], double* s_im,
                                                 int lane) {
  const uint64_t lm = __ballot(counted);
  if (!lm) return;
  if (!__all(!counted || inst == m_inst)) {
    if (m_inst >= 0) {
#pragma unroll
      for (int q = 0; q < kImN; ++q) {
        const double sq = wave_sum(macc[q]);
        macc[q] = 0.0;
        if (lane == 0 && sq != 0.0) atomicAdd(&s_im[m_inst * kImN + q], sq);
      }
    }
    const int inst0 = __shfl(inst, (int)__builtin_ctzll(lm), kWave);
    m_inst = __all(!counted || inst == inst0) ? inst0 : -1;
  }
  if (m_inst >= 0) {
#pragma unroll
    for (int q = 0; q < kImN; ++q) macc[q] += mv[q];
  } else if (counted) {
#pragma unroll
    for (int q = 0; q < kImN; ++q)
      if (mv[q] != 0.0) atomicAdd(&s_im[inst * kImN + q], mv[q]);
  }
}

// LDS doubles before the staged table: DIM 1 fixed-point histograms (2 x nn
// 64-bit sums, nn 32-bit carry words), DIM 2 moments and fields
__host__ __device__ inline int inplace_lds_doubles(int dim, int nn, int n_inst, int word) {
  if (dim == 1) return (2 * nn + (nn + 1) / 2 + 1) & ~1;
  const int d = n_inst * kImN + (n_inst * word + 7) / 8;
  return (d + 1) & ~1;
}

template <typename T, int XS, int DIM, int RNG, int BLOCK, int ITEMS, bool QUEUE>
__global__ __launch_bounds__(BLOCK) void k_push_inplace(PushArgs a) {
  static_assert(DIM == 1 || DIM == 2, "in-place push: 1-D discharges and batched swarms");
  constexpr int NW = BLOCK / kWave;
  constexpr int WSPAN = kWave * ITEMS;
  constexpr int TILE = BLOCK * ITEMS;
  extern __shared__ double s_dep[];
  __shared__ unsigned long long s_cnt[kCntRow];
  __shared__ uint32_t s_births;
  __shared__ T s_thr[8];  // the binning's energy thresholds (keys written here)

  constexpr bool kQueue = QUEUE && QueuedPush<XS, DIM, RNG>::value;
  // the FP64 constants of every particle's kick, walls, deposit node and
  // key in VGPRs (in_vgpr); the rest of the parameters stay uniform
  bz_params_t p = a.p;
  if (DIM == 1) {
    p.dx = in_vgpr(p.dx);
    p.half_dx = in_vgpr(p.half_dx);
    p.L_dx = in_vgpr(p.L_dx);
    p.dep_M = in_vgpr(p.dep_M);
    p.L = in_vgpr(p.L);
    p.afac = in_vgpr(p.afac);
    p.xfac = in_vgpr(p.xfac);
    p.kev = in_vgpr(p.kev);
  }
  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1);
  const int wid = tid / kWave;
  const int nn = DIM == 1 ? p.n_nodes : 0;
  const int n_inst = DIM == 2 ? a.n_inst : 0;
  const int sums = DIM == 1 ? 3 * nn : n_inst * kImN;
  T* s_pup = nullptr;       // kQueue: collision-probability bound per hint bucket
  uint32_t* s_q = nullptr;  // kQueue: candidate slots, kQCap per wave
  T* s_qr = nullptr;        // kQueue: their decision uniforms
  // The slot count first, then the first tile's particles: their round
  // trips overlap the staging of the sums, the table and the field below.
  const int64_t n = a.clk->n;
  const uint64_t step = a.clk->step;
  const uint32_t ntiles = (uint32_t)((n + TILE - 1) / TILE);
  const uint32_t per = (ntiles + gridDim.x - 1) / gridDim.x;
  const uint32_t t_first = (uint32_t)min((uint64_t)blockIdx.x * per, (uint64_t)ntiles);
  const uint32_t t_end = t_first + per < ntiles ? t_first + per : ntiles;
  const T* in[5] = {(const T*)a.in[0], (const T*)a.in[1], (const T*)a.in[2], (const T*)a.in[3],
                    (const T*)a.in[4]};
  P5<T> cur[ITEMS];
  {
    const int64_t b0 = (int64_t)t_first * TILE + wid * WSPAN;
#pragma unroll
    for (int k = 0; k < ITEMS; ++k)
      load5<T, true>(in, b0 + k * kWave + lane, t_first < t_end ? n : 0, cur[k]);
  }
  // DIM 1: fixed-point node sums (deterministic), DIM 2: FP64 moments
  unsigned long long* s_lo = (unsigned long long*)s_dep;
  uint32_t* s_hi = (uint32_t*)(s_lo + 2 * nn);
  const double dscale = DIM == 1 ? dep_scale_of(a.clk->w_max_bits) : 1.0;
  if (DIM == 1) {
    for (int j = tid; j < 2 * nn; j += BLOCK) s_lo[j] = 0ull;
    for (int j = tid; j < nn; j += BLOCK) s_hi[j] = 0u;
  } else {
    for (int j = tid; j < sums; j += BLOCK) s_dep[j] = 0.0;
  }
  T* s_ie = (T*)(s_dep + sums);  // DIM 2: field of every instance
  if (DIM == 2)
    for (int j = tid; j < n_inst; j += BLOCK) s_ie[j] = ((const T*)a.inst_E)[j];
  if (tid < kCntRow) s_cnt[tid] = 0ull;
  if (tid == 0) s_births = 0;
  if (tid < 8) s_thr[tid] = (T)a.bk.thr[tid];
  const BinThrLds<T> thr{(__attribute__((address_space(3))) const T*)s_thr};
  bz_xsec_t xs = a.xs;
  if (XS != BZ_XS_FIT) {
    const int n = a.xs.n;
    T* s_tab = (T*)(s_dep + inplace_lds_doubles(DIM, nn, n_inst, (int)sizeof(T)));
    int32_t* s_hint = (int32_t*)(s_tab + 7 * n);
    stage_lds<T, BLOCK>(s_tab, (const T*)a.xs.energy, n);
    stage_lds<T, BLOCK>(s_tab + n, (const T*)a.xs.sig, 3 * n);
    stage_lds<T, BLOCK>(s_tab + 4 * n, (const T*)a.xs.slope, 3 * n);
    for (int j = tid; j < a.xs.n_hint; j += BLOCK) s_hint[j] = a.xs.hint[j];
    xs.energy = s_tab;
    xs.sig = s_tab + n;
    xs.slope = s_tab + 4 * n;
    xs.hint = s_hint;
    if (kQueue) {
      s_pup = (T*)((char*)s_tab + table_lds_bytes(n, a.xs.n_hint, (int)sizeof(T)));
      s_q = (uint32_t*)((char*)s_pup + (((size_t)a.xs.n_hint * sizeof(T) + 15) & ~(size_t)15));
      s_qr = (T*)(s_q + NW * kQCap);
    }
  }
  // 1-D: the field of every cell in LDS (where it fits), so the kick reads
  // no global memory
  FieldRef<T> E;
  E.g = (const T*)a.E;
  E.lds = DIM == 1 && a.e_lds_off != 0;
  E.s = (__attribute__((address_space(3))) const T*)((char*)s_dep +
                                                     (E.lds ? a.e_lds_off : 0));
  if (E.lds) stage_lds<T, BLOCK>((T*)((char*)s_dep + a.e_lds_off), (const T*)a.E, p.n_cells);

  // the births stage reads the count from n_prev: the slot count itself is
  // committed while it runs
  if (blockIdx.x == 0 && tid == 0) a.clk->n_prev = n;
  T* st[5] = {(T*)a.in[0], (T*)a.in[1], (T*)a.in[2], (T*)a.in[3], (T*)a.in[4]};
  const T* R = (const T*)a.R;
  const T* G = (const T*)a.G;
  uint32_t* wave_births = a.block_births + kMaxPushBlocks;

  uint32_t c_el = 0, c_ex = 0, c_io = 0, c_abs = 0, c_null = 0, c_over = 0, c_live = 0;
  uint32_t w_births = 0;  // this wave's births over its tiles
  // DIM 2: per-lane running moments of the wave's current instance m_inst
  // (wave-uniform, -1: none); summed over the wave and added to LDS only
  // when the wave moves on to another instance
  double macc[kImN];
#pragma unroll
  for (int q = 0; q < kImN; ++q) macc[q] = 0.0;
  int m_inst = -1;
  // the queued path (below) only where a wave walks enough tiles for its
  // batches to fill: the last, partial batch re-reads its particles, a
  // latency a short walk would pay
  const bool queued = kQueue && !(p.gas_sigma > 0.0) &&
                      (a.queue_mode > 0 || per * ITEMS >= kQueueMinItems);
  __syncthreads();  // sums cleared, fields and table staged
  if (kQueue && queued) {
    for (int b = tid; b < xs.n_hint; b += BLOCK) s_pup[b] = (T)bucket_bound<T, XS>(b, xs, p);
    __syncthreads();
  }

  // Queued 1-D path (cold gas).  A live particle whose uniform R0 is at or
  // above its energy bucket's bound cannot collide: it is pushed right away
  // (the null branch of step_particle).  The others are queued per wave in
  // slot order and run through step_particle 64 at a time, so a wave does
  // the sigma interpolation, the probability and the scatter on full lanes.
  // Births are staged as in the general path: in the (tile, wave) slot span
  // of their parent, in slot order.  The wave's birth count of every tile
  // goes to wave_births once, in tile order, from lane 0: t_open is the tile
  // the batches are in, c_open its births so far; every tile below t_open
  // has been written.
  constexpr int PD = DIM == 1 ? 1 : 0;  // physics dimension
  if constexpr (kQueue) {
    if (queued) {
      // the wave's candidate ring, entries [qh, qh + qn) mod kQCap: slots and
      // decision uniforms in LDS; 1-D: the particles' states in global memory
      // too (a batch fills some ten tiles after its first candidate, when
      // their lines have left L2; the swarms re-read theirs from the state,
      // which measured faster there)
      constexpr bool kRing = DIM == 1;
      uint32_t* q = s_q + wid * kQCap;
      T* qr = s_qr + wid * kQCap;
      T* gq = (T*)a.qbuf + (size_t)(blockIdx.x * NW + wid) * 5 * kQCap;
      int qh = 0;
      Decision<T> dec[ITEMS];  // the decision draw of the group each item is in
      const uint64_t lt = (1ull << lane) - 1ull;
      const T e_bot = ((const T*)xs.energy)[0], e_top = ((const T*)xs.energy)[xs.n - 1];
      int qn = 0;                 // queued slots (wave-uniform)
      uint32_t t_open = t_first;  // wave-uniform
      uint32_t c_open = 0;
      auto advance_to = [&](uint32_t t_new) {
        if (lane == 0) {
          wave_births[(size_t)t_open * NW + wid] = c_open;
          for (uint32_t tt = t_open + 1; tt < t_new; ++tt) wave_births[(size_t)tt * NW + wid] = 0u;
        }
        t_open = t_new;
        c_open = 0;
      };
      auto run_batch = [&](int nb) {
        const bool act = lane < nb;
        const int e = (qh + lane) & (kQCap - 1);
        const uint32_t i = act ? q[e] : 0u;
        const T r0 = act ? qr[e] : (T)1;
        P5<T> c;
        if (kRing) {
          // the ring's global stores of this wave (other lanes' entries) are
          // complete and visible before they are read
          __threadfence_block();
          if (act) {
            c.x = gq[e];
            c.vx = gq[kQCap + e];
            c.vy = gq[2 * kQCap + e];
            c.vz = gq[3 * kQCap + e];
            c.w = gq[4 * kQCap + e];
          } else {
            c.x = c.vx = c.vy = c.vz = c.w = (T)0;
          }
        } else {
          load5<T, true>(in, (int64_t)i, act ? n : 0, c);
        }
        int inst = 0;
        T e_local = (T)0;
        if (DIM == 2) {
          inst = (int)c.x;
          inst = inst < 0 ? 0 : (inst >= n_inst ? n_inst - 1 : inst);
          e_local = s_ie[inst];
        }
        const StepOut<T> o =
            step_particle<T, XS, PD, RNG>(c, i, step, p, xs, E, R, G, e_local, &r0);
        if (act) {
          ++c_live;
          c_el += o.n_el;
          c_ex += o.n_exc;
          c_io += o.chan == kChanIon;
          c_null += o.n_null;
          c_abs += !o.alive;
          c_over += o.over;
          if (DIM == 2) {
            // velocities only; this batch's instances are mixed: per-lane sums
            st[1][i] = o.p.vx;
            st[2][i] = o.p.vy;
            st[3][i] = o.p.vz;
            double mv[kImN];
            swarm_moments<T>(o, c.w, true, mv);
  #pragma unroll
            for (int qq = 0; qq < kImN; ++qq)
              if (mv[qq] != 0.0) atomicAdd(&s_dep[inst * kImN + qq], mv[qq]);
          } else if (o.alive) {
            st[0][i] = o.p.x;
            st[1][i] = o.p.vx;
            st[2][i] = o.p.vy;
            st[3][i] = o.p.vz;
            const double w = (double)o.p.w;
            const int nd = dep_node((double)o.p.x, p);
            dep_fixed_add(s_lo, s_hi, 0, nn, nd, w, dscale);
            if (o.chan == kChanIon) dep_fixed_add_global(a.m1_lo, a.m1_hi, nd, w, dscale);
            dep_fixed_add(s_lo, s_hi, 1, nn, nd, w * (double)o.e_post, dscale);
          } else {
            st[4][i] = (T)0;
            if (DIM == 1 && o.chan == kChanIon && (p.ref_flags & BZ_REF_DEAD_IONIZER))
              dep_fixed_add_global(a.m1_lo, a.m1_hi, 0, 1.0, dscale);  // the reference's flag
          }
          if (DIM == 1 && a.coll_lo) coll_add<T>(a, nn, o, p, (double)c.w, dscale);
          if (DIM == 1 && a.keys)
            put_key(a, i, o.alive ? bin_key<T>(a.bk, thr, o.p.x, o.e_post, o.p.w)
                                  : (uint32_t)a.bk.n_keys);
        }
        // births: the batch's slots rise, so equal tiles are runs of lanes
        // (segments), taken in order
        const bool born = act && o.born;
        const uint64_t m = __ballot(born);
        const uint32_t te = act ? i / (uint32_t)TILE : 0xFFFFFFFFu;
        const uint32_t tprev = __shfl_up(te, 1, kWave);
        uint64_t heads = __ballot(act && (lane == 0 || te != tprev));
        while (heads) {
          const int s0 = (int)__builtin_ctzll(heads);
          heads &= heads - 1ull;
          const int s1 = heads ? (int)__builtin_ctzll(heads) : 64;
          const uint64_t below_s1 = s1 >= 64 ? ~0ull : ((1ull << s1) - 1ull);
          const uint64_t seg = ~((1ull << s0) - 1ull) & below_s1;
          const uint32_t ts = __shfl(te, s0, kWave);
          if (ts != t_open) advance_to(ts);
          if (born && ((seg >> lane) & 1ull)) {
            const uint32_t r = c_open + (uint32_t)__popcll(m & lt & seg);
            store5<T, true>(a.out, (int64_t)ts * TILE + wid * WSPAN + r, o.s);
            if (DIM == 1) {
              const double w = (double)o.s.w;
              const int nd = dep_node((double)o.s.x, p);
              dep_fixed_add(s_lo, s_hi, 0, nn, nd, w, dscale);
              dep_fixed_add(s_lo, s_hi, 1, nn, nd, w * (double)o.e_sec, dscale);
            }
          }
          c_open += (uint32_t)__popcll(m & seg);
        }
        w_births += (uint32_t)__popcll(m);
      };
      for (uint32_t t = t_first; t < t_end; ++t) {
        const int64_t bw = (int64_t)t * TILE + wid * WSPAN;
        P5<T> nxt[ITEMS];
  #pragma unroll
        for (int k = 0; k < ITEMS; ++k)
          load5<T, true>(in, bw + TILE + k * kWave + lane, t + 1 < t_end ? n : 0, nxt[k]);
  #pragma unroll
        for (int k = 0; k < ITEMS; ++k) {
          const int64_t i = bw + k * kWave + lane;
          const bool live = i < n && cur[k].w > (T)0;
          // every lane draws (the same group across the wave): a Philox
          // block once per 4 tiles of 1024 slots (2 in FP64)
          const T r0 = dec[k].get(p.seed, (uint32_t)i, step);
          bool cand = false;
          if (live) {
            const T s2 = cur[k].vx * cur[k].vx + cur[k].vy * cur[k].vy + cur[k].vz * cur[k].vz;
            const T e = (T)p.kev * s2;
            if (!(e < e_top)) {
              cand = true;  // beyond the table: the speed is not bounded
            } else {
              const T ec = e < e_bot ? e_bot : e;
              int b = xs_bucket<T>(ec, xs.hint_bits) - xs.hint_base;
              b = b < 0 ? 0 : (b >= xs.n_hint ? xs.n_hint - 1 : b);
              cand = r0 < s_pup[b];
            }
          }
          const bool plain = live && !cand;
          int inst = 0;
          T e_local = (T)0;
          if (DIM == 2) {
            inst = (int)cur[k].x;
            inst = inst < 0 ? 0 : (inst >= n_inst ? n_inst - 1 : inst);
            e_local = s_ie[inst];
          }
          StepOut<T> o;
          o.born = 0;
          o.chan = kChanNone;
          o.e_sec = (T)0;
          o.s = cur[k];
          push_tail<T, PD>(cur[k], cur[k].vx, cur[k].vy, cur[k].vz, p, E, e_local, o);
          if (plain) {
            ++c_live;
            ++c_null;
            if (DIM == 2) {
              st[1][i] = o.p.vx;  // only vx moves without a collision
            } else if (o.alive) {
              st[0][i] = o.p.x;
              st[1][i] = o.p.vx;
              const double w = (double)o.p.w;
              const int nd = dep_node((double)o.p.x, p);
              dep_fixed_add(s_lo, s_hi, 0, nn, nd, w, dscale);
              dep_fixed_add(s_lo, s_hi, 1, nn, nd, w * (double)o.e_post, dscale);
            } else {
              ++c_abs;
              st[4][i] = (T)0;
            }
          }
          // keys of the slots settled here (candidates get theirs in their batch)
          if (DIM == 1 && a.keys && i < n && !(live && cand))
            put_key(a, i, plain && o.alive ? bin_key<T>(a.bk, thr, o.p.x, o.e_post, o.p.w)
                                           : (uint32_t)a.bk.n_keys);
          if (DIM == 2) {
            // running moments of the wave's instance (the queued candidates
            // add theirs in their batch)
            double mv[kImN];
            swarm_moments<T>(o, cur[k].w, plain, mv);
            swarm_accumulate(mv, plain, inst, m_inst, macc, s_dep, lane);
          }
          const uint64_t cm = __ballot(live && cand);
          if (live && cand) {
            const int at = (qh + qn + (int)__popcll(cm & lt)) & (kQCap - 1);
            q[at] = (uint32_t)i;
            qr[at] = r0;
            if (kRing) {
              gq[at] = cur[k].x;
              gq[kQCap + at] = cur[k].vx;
              gq[2 * kQCap + at] = cur[k].vy;
              gq[3 * kQCap + at] = cur[k].vz;
              gq[4 * kQCap + at] = cur[k].w;
            }
          }
          qn += (int)__popcll(cm);
          if (qn >= 64) {
            run_batch(64);
            qh = (qh + 64) & (kQCap - 1);
            qn -= 64;
          }
        }
  #pragma unroll
        for (int k = 0; k < ITEMS; ++k) cur[k] = nxt[k];
      }
      if (qn > 0) run_batch(qn);
      if (t_first < t_end) advance_to(t_end);  // the last tile's count, zeros after it
    }
  }
  Decision<T> dec_nq[ITEMS];  // the decision draw of the group each item is in
  if (!queued)
  for (uint32_t t = t_first; t < t_end; ++t) {
    const int64_t bw = (int64_t)t * TILE + wid * WSPAN;
    // next tile first, so its HBM latency hides behind this tile's physics
    P5<T> nxt[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; ++k)
      load5<T, true>(in, bw + TILE + k * kWave + lane, t + 1 < t_end ? n : 0, nxt[k]);
    uint32_t tb = 0;  // births of this wave in this tile
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
      const int64_t i = bw + k * kWave + lane;
      const bool live = i < n && cur[k].w > (T)0;
      int inst = 0;
      T e_local = (T)0;
      if (DIM == 2) {
        inst = (int)cur[k].x;
        inst = inst < 0 ? 0 : (inst >= n_inst ? n_inst - 1 : inst);
        e_local = s_ie[inst];
      }
      constexpr int PD = DIM == 1 ? 1 : 0;  // physics dimension
      StepOut<T> o;
      if (XS == BZ_XS_FLIGHT) {
        o = flight_particle<T, PD>(cur[k], (uint32_t)i, step, p, xs, E, e_local);
      } else if (RNG == BZ_RNG_STREAM) {
        o = step_particle<T, XS, PD, RNG>(cur[k], (uint32_t)i, step, p, xs, E, R, G, e_local);
      } else {
        const T r0 = dec_nq[k].get(p.seed, (uint32_t)i, step);
        o = step_particle<T, XS, PD, RNG>(cur[k], (uint32_t)i, step, p, xs, E, R, G, e_local,
                                          &r0);
      }
      if (live) {
        ++c_live;
        c_el += o.n_el;
        c_ex += o.n_exc;
        c_io += o.chan == kChanIon;
        c_null += o.n_null;
        c_abs += !o.alive;
        c_over += o.over;
        if (DIM == 2) {
          st[1][i] = o.p.vx;
          st[2][i] = o.p.vy;
          st[3][i] = o.p.vz;
        } else if (o.alive) {
          st[0][i] = o.p.x;
          st[1][i] = o.p.vx;
          st[2][i] = o.p.vy;
          st[3][i] = o.p.vz;
          const double w = (double)o.p.w;
          const int nd = dep_node((double)o.p.x, p);
          dep_fixed_add(s_lo, s_hi, 0, nn, nd, w, dscale);
          if (o.chan == kChanIon) dep_fixed_add_global(a.m1_lo, a.m1_hi, nd, w, dscale);
          dep_fixed_add(s_lo, s_hi, 1, nn, nd, w * (double)o.e_post, dscale);
        } else {
          st[4][i] = (T)0;
          if (DIM == 1 && o.chan == kChanIon && (p.ref_flags & BZ_REF_DEAD_IONIZER))
            dep_fixed_add_global(a.m1_lo, a.m1_hi, 0, 1.0, dscale);  // the reference's flag
        }
        if (DIM == 1 && a.coll_lo) coll_add<T>(a, nn, o, p, (double)cur[k].w, dscale);
      }
      if (DIM == 1 && a.keys && i < n)
        put_key(a, i, live && o.alive ? bin_key<T>(a.bk, thr, o.p.x, o.e_post, o.p.w)
                                     : (uint32_t)a.bk.n_keys);
      if (DIM == 2) {
        // per-instance moments of this step; a wave whose lanes share one
        // instance (the common case: instances are contiguous between
        // population controls) sums in registers first
        double mv[kImN];
        swarm_moments<T>(o, cur[k].w, live, mv);
        swarm_accumulate(mv, live, inst, m_inst, macc, s_dep, lane);
      }
      const bool born = live && o.born;
      const uint64_t m = __ballot(born);
      if (born) {
        const uint32_t r = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
        store5<T, true>(a.out, bw + tb + r, o.s);
        if (DIM == 1) {
          const double w = (double)o.s.w;
          const int nd = dep_node((double)o.s.x, p);
          dep_fixed_add(s_lo, s_hi, 0, nn, nd, w, dscale);
          dep_fixed_add(s_lo, s_hi, 1, nn, nd, w * (double)o.e_sec, dscale);
        }
      }
      tb += (uint32_t)__popcll(m);
    }
    if (lane == 0) wave_births[(size_t)t * NW + wid] = tb;
    w_births += tb;
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) cur[k] = nxt[k];
  }

  if (DIM == 2 && m_inst >= 0) {
#pragma unroll
    for (int q = 0; q < kImN; ++q) {
      const double sq = wave_sum(macc[q]);
      if (lane == 0 && sq != 0.0) atomicAdd(&s_dep[m_inst * kImN + q], sq);
    }
  }
  {
    const uint32_t v[kCntRow] = {c_io, c_abs, c_el, c_ex, c_io, c_null, 0u, c_over, c_live, 0u};
#pragma unroll
    for (int q = 0; q < kCntRow; ++q) {
      const uint32_t s = wave_sum(v[q]);
      if (lane == 0 && s) atomicAdd(&s_cnt[q], (unsigned long long)s);
    }
    if (lane == 0 && w_births) atomicAdd(&s_births, w_births);
  }
  __syncthreads();
  double* part = a.partials + (size_t)blockIdx.x * sums;
  if (DIM == 1) {
    // rows 0 and 2 from LDS; row 1 is summed in the global words
    const double inv = 1.0 / dscale;
    for (int j = tid; j < sums; j += BLOCK) {
      const int m = j / nn, k = j - m * nn;
      part[j] = m == 1 ? 0.0
                       : dep_fixed_value(s_lo[(m >> 1) * nn + k], dep_carries(s_hi[k], m >> 1),
                                         inv);
    }
    bool near = false;
    for (int k = tid; k < nn; k += BLOCK) near |= dep_near_range(s_hi[k]);
    if (__any(near) && lane == 0) atomicAdd(&a.clk->cnt[BZ_CNT_DEP_RANGE], 1ull);
  } else {
    for (int j = tid; j < sums; j += BLOCK) part[j] = s_dep[j];
  }
  uint32_t* bcnt = (uint32_t*)(a.partials + (size_t)gridDim.x * sums);
  if (tid < kCntRow) bcnt[(size_t)blockIdx.x * kCntRow + tid] = (uint32_t)s_cnt[tid];
  if (tid == 0) a.block_births[blockIdx.x] = s_births;
}

// After k_push_inplace: the births of push block b go after the slots in use,
// behind the births of blocks 0..b-1 (summed here by this block itself), in
// the order (tile, wave, slot), which is parent order.  The block handling
// the last push block sets the slot count of the new state; births past the
// capacity are dropped and counted.  NT threads.
template <typename T, int NT>
__device__ void births_part(const PushArgs& a, int b, int nblk, int tile, int nw,
                            uint32_t* s_red) {
  constexpr int NWV = NT / 64;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int64_t n = a.clk->n_prev;
  const uint32_t ntiles = (uint32_t)((n + tile - 1) / tile);
  const uint32_t per = (ntiles + nblk - 1) / nblk;
  const uint32_t t_first = (uint32_t)min((uint64_t)b * per, (uint64_t)ntiles);
  const uint32_t t_end = t_first + per < ntiles ? t_first + per : ntiles;
  uint32_t s = 0;
  for (int j = tid; j < b; j += NT) s += a.block_births[j];
  s = wave_sum(s);
  if (lane == 0) s_red[wid] = s;
  __syncthreads();
  int64_t pre = 0;
#pragma unroll
  for (int w = 0; w < NWV; ++w) pre += s_red[w];
  const uint32_t mine = a.block_births[b];
  T* st[5] = {(T*)a.in[0], (T*)a.in[1], (T*)a.in[2], (T*)a.in[3], (T*)a.in[4]};
  const T* sg[5] = {(const T*)a.out[0], (const T*)a.out[1], (const T*)a.out[2],
                    (const T*)a.out[3], (const T*)a.out[4]};
  const uint32_t* wave_births = a.block_births + kMaxPushBlocks;
  const int wspan = tile / nw;
  const int64_t entries = (int64_t)(t_end - t_first) * nw;
  int64_t run = 0;  // births of this block's earlier entries
  if (mine) {
    for (int64_t e0 = 0; e0 < entries; e0 += NT) {
      const int64_t e = e0 + tid;
      const uint32_t c = e < entries ? wave_births[(size_t)t_first * nw + e] : 0u;
      // block exclusive scan of c
      uint32_t incl = c;
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const uint32_t v = __shfl_up(incl, off, 64);
        if (lane >= off) incl += v;
      }
      __syncthreads();
      if (lane == 63) s_red[wid] = incl;
      __syncthreads();
      uint32_t before = 0, tot = 0;
#pragma unroll
      for (int w = 0; w < NWV; ++w) {
        if (w < wid) before += s_red[w];
        tot += s_red[w];
      }
      const int64_t dst0 = n + pre + run + before + incl - c;
      const int64_t src0 = (int64_t)(t_first + e / nw) * tile + (e % nw) * wspan;
      for (uint32_t j = 0; j < c; ++j) {
        const int64_t dst = dst0 + j;
        if (dst >= a.capacity) break;
        T v[5];
#pragma unroll
        for (int f = 0; f < 5; ++f) {
          v[f] = sg[f][src0 + j];
          st[f][dst] = v[f];
        }
        if (a.keys) {
          const T e = (T)a.p.kev * (v[1] * v[1] + v[2] * v[2] + v[3] * v[3]);
          put_key(a, dst, bin_key<T>(a.bk, v[0], e, v[4]));
        }
      }
      run += tot;
    }
  }
  if (b == nblk - 1 && tid == 0) {
    const int64_t want = n + pre + mine;
    a.clk->n = want < a.capacity ? want : a.capacity;
    if (want > a.capacity)
      atomicAdd(&a.clk->cnt[BZ_CNT_OVERFLOW], (unsigned long long)(want - a.capacity));
  }
}

// After the in-place push, one launch: blocks [0, nbirth) append the births
// of push block b (births_part), the next nx*ny blocks sum the partial rows
// (dep_reduce_part), and the last block adds the counters to the clock.
// The slot count is committed by the births of the last push block.
template <typename T>
__global__ __launch_bounds__(kRedThreads) void k_push_post(PushArgs a, int tile, int nw,
                                                           int nbirth, int m, int rpc, int nx,
                                                           int ny, double* l2,
                                                           uint32_t* col_done,
                                                           double* __restrict__ out,
                                                           double* acc, double advance_dt) {
  __shared__ double s_part[kRedWaves][64];
  __shared__ int s_last;
  __shared__ uint32_t s_red[kRedWaves];
  const int b = blockIdx.x;
  if (b < nbirth) {
    births_part<T, kRedThreads>(a, b, nbirth, tile, nw, s_red);
  } else if (b < nbirth + nx * ny) {
    const int r = b - nbirth;
    M1Fold f{a.m1_lo, a.m1_hi, a.p.n_nodes, 1.0 / dep_scale_of(a.clk->w_max_bits)};
    if (!a.m1_lo) f.lo = nullptr;
    dep_reduce_part(a.partials, nbirth, m, rpc, l2, col_done, out, acc, r % nx, r / nx, ny,
                    s_part, &s_last, f);
  } else if (b == nbirth + nx * ny) {
    dep_commit(a.partials, nbirth, m, a.clk, a.capacity, false, advance_dt);
  } else if (a.coll_lo) {
    // the collision counters of this push into the accumulator, words cleared
    const int nn = a.p.n_nodes;
    const int k = (b - nbirth - nx * ny - 1) * kRedThreads + threadIdx.x;
    if (k < nn) {
      const double inv = 1.0 / dep_scale_of(a.clk->w_max_bits);
      for (int r = 0; r < 2; ++r) {
        const int j = r * nn + k;
        a.coll_acc[j] += dep_fixed_value(a.coll_lo[j], a.coll_hi[j], inv);
        a.coll_lo[j] = 0ull;
        a.coll_hi[j] = 0u;
      }
      a.coll_acc[2 * nn + k] += (double)a.coll_cnt[k];
      a.coll_cnt[k] = 0ull;
    }
  }
}

// The collision counters accumulated since the last reset to out[3][nn]
// (device memory); reset clears the accumulator.
__global__ void k_coll_read(double* acc, double* out, int nn, int reset) {
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < 3 * nn; j += gridDim.x * blockDim.x) {
    out[j] = acc[j];
    if (reset) acc[j] = 0.0;
  }
}


// ---------------------------------------------------------------------------
// Stable compaction of the live slots (w > 0).  Block b owns a contiguous run
// of slots; k_compact_count counts its live ones, k_compact_scatter sums the
// counts of blocks 0..b-1 itself and writes its live particles in slot order.
// ---------------------------------------------------------------------------

template <typename T>
__global__ __launch_bounds__(kCompBlock) void k_compact_count(const T* __restrict__ w,
                                                              const Clock* clk,
                                                              uint32_t* block_cnt) {
  __shared__ uint32_t s_red[kCompBlock / kWave];
  int64_t lo, hi;
  comp_range(clk->n, &lo, &hi);
  uint32_t c = 0;
  for (int64_t i = lo + threadIdx.x; i < hi; i += kCompBlock) c += w[i] > (T)0 ? 1u : 0u;
  c = wave_sum(c);
  if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) block_cnt[blockIdx.x] = s_red[0] + s_red[1] + s_red[2] + s_red[3];
}

template <typename T>
__global__ __launch_bounds__(kCompBlock) void k_compact_scatter(bz_state_t in, bz_state_t out,
                                                                Clock* clk,
                                                                const uint32_t* block_cnt) {
  constexpr int NW = kCompBlock / kWave;
  __shared__ uint32_t s_red[NW];
  __shared__ uint64_t s_wsum[NW];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int64_t n = clk->n;
  int64_t lo, hi;
  comp_range(n, &lo, &hi);
  uint32_t s = 0;
  for (int j = tid; j < (int)blockIdx.x; j += kCompBlock) s += block_cnt[j];
  s = wave_sum(s);
  if (lane == 0) s_red[wid] = s;
  __syncthreads();
  int64_t pos0 = (int64_t)s_red[0] + s_red[1] + s_red[2] + s_red[3];
  const T* src[5] = {(const T*)in.x, (const T*)in.vx, (const T*)in.vy, (const T*)in.vz,
                     (const T*)in.w};
  T* dst[5] = {(T*)out.x, (T*)out.vx, (T*)out.vy, (T*)out.vz, (T*)out.w};
  for (int64_t base = lo; base < hi; base += kCompTile) {
    uint64_t packed = 0;
#pragma unroll
    for (int k = 0; k < kCompItems; ++k) {
      const int64_t i = base + k * kCompBlock + tid;
      const bool live = i < hi && src[4][i] > (T)0;
      packed |= (uint64_t)(live ? 1u : 0u) << (16 * k);
    }
    const uint64_t incl = wave_incl_scan(packed);
    if (lane == kWave - 1) s_wsum[wid] = incl;
    __syncthreads();
    uint64_t before = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      if (w < wid) before += s_wsum[w];
      tot += s_wsum[w];
    }
    const uint64_t excl = before + incl - packed;
    uint32_t field_base = 0;
#pragma unroll
    for (int k = 0; k < kCompItems; ++k) {
      if ((packed >> (16 * k)) & 1u) {
        const int64_t i = base + k * kCompBlock + tid;
        const int64_t d = pos0 + field_base + (int64_t)((excl >> (16 * k)) & 0xFFFFu);
#pragma unroll
        for (int f = 0; f < 5; ++f) dst[f][d] = src[f][i];
      }
      field_base += (uint32_t)((tot >> (16 * k)) & 0xFFFFu);
    }
    pos0 += field_base;
    __syncthreads();  // s_wsum is rewritten by the next sub-tile
  }
  if (blockIdx.x == gridDim.x - 1 && tid == 0) clk->n_next = pos0;
}

__global__ void k_clock_commit_next(Clock* clk) { clk->n = clk->n_next; }

// Deposit of a state without pushing it (initial moments).  1-D: the
// fixed-point node sums of the push (deterministic); 0-D: per-wave sums
// added in wave order.
template <typename T, int DIM, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_deposit(bz_params_t p, const T* __restrict__ x,
                                                  const T* __restrict__ vx,
                                                  const T* __restrict__ vy,
                                                  const T* __restrict__ vz,
                                                  const T* __restrict__ w, Clock* clk,
                                                  double* partials) {
  extern __shared__ double s_dep[];
  __shared__ double s_w0[BLOCK / kWave][2];
  const int nn = DIM == 1 ? p.n_nodes : 1;
  unsigned long long* s_lo = (unsigned long long*)s_dep;
  uint32_t* s_hi = (uint32_t*)(s_lo + 2 * nn);
  const double dscale = dep_scale_of(clk->w_max_bits);
  if (DIM == 1) {
    for (int j = threadIdx.x; j < 2 * nn; j += BLOCK) s_lo[j] = 0ull;
    for (int j = threadIdx.x; j < nn; j += BLOCK) s_hi[j] = 0u;
  }
  __syncthreads();
  const int64_t n = clk->n;
  double a0 = 0.0, a2 = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * BLOCK) {
    const T wi = w[i];
    if (!(wi > (T)0)) continue;
    const T e = (T)p.kev * (vx[i] * vx[i] + vy[i] * vy[i] + vz[i] * vz[i]);
    const double wd = (double)wi;
    if (DIM == 1) {
      const int nd = dep_node((double)x[i], p);
      dep_fixed_add(s_lo, s_hi, 0, nn, nd, wd, dscale);
      dep_fixed_add(s_lo, s_hi, 1, nn, nd, wd * (double)e, dscale);
    } else {
      a0 += wd;
      a2 += wd * (double)e;
    }
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (DIM == 0) {
    a0 = wave_sum(a0);
    a2 = wave_sum(a2);
    if (lane == 0) {
      s_w0[wid][0] = a0;
      s_w0[wid][1] = a2;
    }
  }
  __syncthreads();
  double* part = partials + (size_t)blockIdx.x * 3 * nn;
  if (DIM == 1) {
    const double inv = 1.0 / dscale;
    for (int j = threadIdx.x; j < 3 * nn; j += BLOCK) {
      const int m = j / nn, k = j - m * nn;  // no ionization without a push
      part[j] = m == 1 ? 0.0
                       : dep_fixed_value(s_lo[(m >> 1) * nn + k], dep_carries(s_hi[k], m >> 1),
                                         inv);
    }
  } else if (threadIdx.x == 0) {
    double t0 = 0.0, t2 = 0.0;
    for (int q = 0; q < BLOCK / kWave; ++q) {
      t0 += s_w0[q][0];
      t2 += s_w0[q][1];
    }
    part[0] = t0;
    part[1] = 0.0;
    part[2] = t2;
  }
}

// Upper bound of the weights of a state: the largest weight's bits, by
// integer max (a positive double's bits order like the double), into the
// clock's bound; the deposit's fixed-point scale follows from it.
template <typename T>
__global__ __launch_bounds__(256) void k_dep_wmax(const T* __restrict__ w, Clock* clk) {
  const int64_t n = clk->n;
  double m = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const double v = (double)w[i];
    m = v > m ? v : m;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
  if ((threadIdx.x & 63) == 0 && m > 0.0)
    atomicMax(&clk->w_max_bits, (unsigned long long)__double_as_longlong(m));
}

}  // namespace bz

// ---------------------------------------------------------------------------
// Launch helpers (called from bz_api.hip)
// ---------------------------------------------------------------------------
#include "bz_launch.h"

namespace bz {

static_assert(5 * kQCap == kPushQueueWords && (kQCap & (kQCap - 1)) == 0,
              "the candidate ring: a power of two, sized as the context's queue buffer");

// Launch of k_dep_reduce: rows per chunk so that the level-2 rows fit.
// Geometry of the two-level reduction of nblocks rows of m sums: 64 outputs
// per block column (nx), row chunks of rpc rows (ny).
static int red_geometry(const LaunchEnv& env, int nblocks, int m, int* rpc_out, int* nx_out,
                        int* ny_out) {
  // one level up to 256 rows (at most 16 rows per thread), two above
  int64_t rpc = 256;
  const int64_t need = (int64_t)nblocks * m;
  if (need > rpc * env.dep_l2_doubles) rpc = (need + env.dep_l2_doubles - 1) / env.dep_l2_doubles;
  if (rpc > 65535) return fail("deposit reduction: too many partial rows");
  const int ny = (int)((nblocks + rpc - 1) / rpc);
  if ((int64_t)ny * m > env.dep_l2_doubles && ny > 1)
    return fail("deposit reduction: level-2 buffer");
  const int nx = (m + 63) / 64;
  if (nx > env.dep_cols) return fail("deposit reduction: %d output columns > %d", nx, env.dep_cols);
  *rpc_out = (int)rpc;
  *nx_out = nx;
  *ny_out = ny > 0 ? ny : 1;
  return 0;
}

int launch_dep_reduce(const LaunchEnv& env, int nblocks, int m, double* out, int commit,
                      int64_t cap, double* acc) {
  if (m < 1) return 0;
  int rpc = 0, nx = 0, ny = 0;
  const int rc = red_geometry(env, nblocks, m, &rpc, &nx, &ny);
  if (rc) return rc;
  hipLaunchKernelGGL(k_dep_reduce, dim3(nx + (commit ? 1 : 0), ny),
                     dim3(kRedWaves * 64), 0, env.stream,
                     (const double*)env.partials, nblocks, m, rpc, env.dep_l2, env.dep_cols_done,
                     out, acc, env.clock, commit, cap);
  return check_launch("k_dep_reduce");
}

int64_t dep_max_row(int32_t max_nodes, int32_t max_keys) {
  int64_t m = 4ll * max_nodes;  // the reference reduction interface: r = 4 columns
  if ((int64_t)kImN * max_keys > m) m = (int64_t)kImN * max_keys;
  return m;
}

int64_t push_partials_doubles(int num_cu, int32_t max_nodes, int32_t max_keys) {
  // up to 8 resident blocks per CU, each with its row of sums and 8 counters
  // (4 doubles); the field step also keeps two flux arrays here
  auto rows = [num_cu](int64_t row_doubles) {
    // a block's LDS holds its row, so at most 160 KiB / row blocks fit a CU
    int64_t per_cu = (160 * 1024) / (row_doubles * 8);
    per_cu = per_cu < 1 ? 1 : (per_cu > 8 ? 8 : per_cu);
    return (int64_t)num_cu * per_cu * row_doubles;
  };
  const int64_t hist = rows(3ll * max_nodes + kCntRow / 2);
  const int64_t inst = rows((int64_t)kImN * max_keys + kCntRow / 2);
  int64_t d = hist > inst ? hist : inst;
  if (d < 2ll * (max_nodes + 1)) d = 2ll * (max_nodes + 1);
  return d;
}

// What the push launch leaves for the stage after it.
struct PushLaunch {
  int grid;     // push blocks (partial rows)
  int tile;     // in-place push: slots per tile
  int nw;       // in-place push: waves per block
  bool inplace;
};

// 1-D discharges and batched swarms: k_push_inplace (k_push_post follows)
template <typename T, int XS, int DIM, int RNG, int BLOCK, int ITEMS, bool QUEUE>
static int launch_push_inplace_q(const LaunchEnv& env, PushArgs& a, PushLaunch* pl) {
  auto kern = k_push_inplace<T, XS, DIM, RNG, BLOCK, ITEMS, QUEUE>;
  const int nn = DIM == 1 ? a.p.n_nodes : 0;
  const int part_doubles = DIM == 1 ? 3 * nn : a.n_inst * kImN;
  size_t smem = (size_t)inplace_lds_doubles(DIM, nn, a.n_inst, (int)sizeof(T)) * sizeof(double);
  if (XS != BZ_XS_FIT) smem += table_lds_bytes(a.xs.n, a.xs.n_hint, (int)sizeof(T));
  if (QUEUE && QueuedPush<XS, DIM, RNG>::value)
    smem += queue_lds_bytes(a.xs.n_hint, (int)sizeof(T), BLOCK / kWave);
  if (smem > 160 * 1024) return fail("push: sums + table need %zu B of LDS (> 160 KiB)", smem);
  // 1-D: the field staged after the rest, where it fits without costing blocks
  const size_t smem_e = smem + (DIM == 1 ? (((size_t)a.p.n_cells * sizeof(T) + 15) & ~(size_t)15)
                                         : (size_t)0);
  const size_t smem_max = DIM == 1 && smem_e <= 160 * 1024 ? smem_e : smem;
  if (smem_max > 64 * 1024) {
    if (hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)smem_max) != hipSuccess)
      return fail("hipFuncSetAttribute(dynamic LDS) failed");
  }
  int occ = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, BLOCK, smem) != hipSuccess ||
      occ < 1)
    return fail("push kernel does not fit on a CU (LDS %zu B)", smem);
  a.e_lds_off = 0;
  if (smem_max > smem) {
    int occ_e = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ_e, kern, BLOCK, smem_max) ==
            hipSuccess &&
        (occ_e > 8 ? 8 : occ_e) >= (occ > 8 ? 8 : occ)) {
      a.e_lds_off = (int)smem;  // 16-byte aligned: every part above is
      smem = smem_max;
      occ = occ_e;
    }
  }
  int bpc = occ > 8 ? 8 : occ;
  if (env.push_bpc > 0 && env.push_bpc < bpc) bpc = env.push_bpc;
  int grid = env.num_cu * bpc;
  const int64_t cap_parts = env.partials_doubles / (part_doubles + kCntRow / 2);
  if (grid > cap_parts) grid = (int)cap_parts;
  if (grid > kMaxPushBlocks) grid = kMaxPushBlocks;
  if (QUEUE && QueuedPush<XS, DIM, RNG>::value) {
    // one candidate ring per wave
    const int64_t gq = env.push_q_waves / (BLOCK / kWave);
    if (grid > gq) grid = (int)gq;
  }
  if (grid < 1) return fail("partial sum buffer too small");
  a.qbuf = env.push_q;
  // per-(tile, wave) birth counts follow the per-block ones in the status words
  const int64_t max_tiles = env.capacity / (BLOCK * ITEMS) + 2;
  if (kMaxPushBlocks + max_tiles * (BLOCK / kWave) > 2 * env.status_words)
    return fail("status words too few for the birth counts");
  a.block_births = (uint32_t*)env.status;
  timing_begin(env, 0);
  hipLaunchKernelGGL(kern, dim3(grid), dim3(BLOCK), smem, env.stream, a);
  timing_end(env, 0);
  pl->grid = grid;
  pl->tile = BLOCK * ITEMS;
  pl->nw = BLOCK / kWave;
  pl->inplace = true;
  return check_launch("k_push_inplace");
}

// The queued variant where it can pay: forced, or when a full state would
// give each wave at least kQueueMinItems tile items (the kernel still
// checks the walk of the actual count).
template <typename T, int XS, int DIM, int RNG, int BLOCK, int ITEMS>
static int launch_push_inplace_t(const LaunchEnv& env, PushArgs& a, PushLaunch* pl) {
  if constexpr (QueuedPush<XS, DIM, RNG>::value) {
    // tile items per wave of a full state at one block per CU
    const int64_t walk = env.capacity / ((int64_t)BLOCK * env.num_cu);
    const bool want = env.push_queue > 0 || (env.push_queue == 0 && walk >= kQueueMinItems);
    if (want && !(a.p.gas_sigma > 0.0)) {
      a.queue_mode = env.push_queue > 0 ? 1 : 0;
      return launch_push_inplace_q<T, XS, DIM, RNG, BLOCK, ITEMS, true>(env, a, pl);
    }
  }
  a.queue_mode = 0;
  return launch_push_inplace_q<T, XS, DIM, RNG, BLOCK, ITEMS, false>(env, a, pl);
}

template <typename T, int XS, int DIM, int RNG, int BLOCK, int ITEMS>
static int launch_push_t(const LaunchEnv& env, PushArgs& a, PushLaunch* pl) {
  // the 1-D discharges and the batched swarms update in place; a single
  // 0-D swarm compacts
  if constexpr (DIM != 0) {
    return launch_push_inplace_t<T, XS, DIM, RNG, BLOCK, ITEMS>(env, a, pl);
  } else {
    auto kern = k_push<T, XS, DIM, RNG, BLOCK, ITEMS>;
    // the kernel's three sums, then the table
    const int part_doubles = 3;
    size_t smem = (size_t)4 * sizeof(double);
    if (XS != BZ_XS_FIT) smem += table_lds_bytes(a.xs.n, a.xs.n_hint, (int)sizeof(T));
    if (smem > 160 * 1024) return fail("push: deposit + table need %zu B of LDS (> 160 KiB)", smem);
    if (smem > 64 * 1024) {
      if (hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)smem) != hipSuccess)
        return fail("hipFuncSetAttribute(dynamic LDS) failed");
    }
    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, BLOCK, smem) != hipSuccess ||
        occ < 1)
      return fail("push kernel does not fit on a CU (LDS %zu B)", smem);
    // every block resident: occupancy-sized grid
    int grid = env.num_cu * (occ > 8 ? 8 : occ);
    // partial histograms [grid][part] doubles followed by [grid][kCntRow] u32 counters
    const int64_t cap_parts = env.partials_doubles / (part_doubles + kCntRow / 2);
    if (grid > cap_parts) grid = (int)cap_parts;
    if (grid < 1) return fail("deposit partial buffer too small");
    a.sched = env.push_sched;
    void* args[] = {&a};
    timing_begin(env, 0);
    hipError_t e = hipSuccess;
    if (a.sched == 0)
      // look-back waits on other blocks' tiles: every block must be resident
      e = hipLaunchCooperativeKernel((const void*)kern, dim3(grid), dim3(BLOCK), args,
                                     (unsigned)smem, env.stream);
    else
      hipLaunchKernelGGL(kern, dim3(grid), dim3(BLOCK), smem, env.stream, a);
    timing_end(env, 0);
    if (e != hipSuccess) return fail("k_push cooperative launch: %s", hipGetErrorString(e));
    pl->grid = grid;
    pl->tile = 0;
    pl->nw = 0;
    pl->inplace = false;
    return check_launch("k_push");
  }
}

template <typename T, int XS, int DIM, int RNG>
static int launch_push_shape(const LaunchEnv& env, PushArgs& a, int shape, PushLaunch* grid) {
  switch (shape) {
    case 0: return launch_push_t<T, XS, DIM, RNG, 256, 4>(env, a, grid);
    case 1: return launch_push_t<T, XS, DIM, RNG, 512, 2>(env, a, grid);
    case 2: return launch_push_t<T, XS, DIM, RNG, 256, 2>(env, a, grid);
    case 3: return launch_push_t<T, XS, DIM, RNG, 1024, 1>(env, a, grid);
    case 4: return launch_push_t<T, XS, DIM, RNG, 1024, 2>(env, a, grid);
    default: return fail("unknown push shape %d", shape);
  }
}

template <typename T>
static int launch_push_prec(const LaunchEnv& env, PushArgs& a, int shape, PushLaunch* grid) {
  const bz_params_t& p = a.p;
  const bool stream = p.rng_mode == BZ_RNG_STREAM;
  if (a.n_inst > 0) {
    if (p.xs_mode != BZ_XS_TABLE || stream)
      return fail("batched swarms use the tabulated model and the Philox stream");
    return launch_push_shape<T, BZ_XS_TABLE, 2, BZ_RNG_PHILOX>(env, a, shape, grid);
  }
  if (p.xs_mode == BZ_XS_FLIGHT) {
    if (stream) return fail("free-flight collisions draw from the Philox stream only");
    return p.dim == 1 ? launch_push_shape<T, BZ_XS_FLIGHT, 1, BZ_RNG_PHILOX>(env, a, shape, grid)
                      : launch_push_shape<T, BZ_XS_FLIGHT, 0, BZ_RNG_PHILOX>(env, a, shape, grid);
  }
  if (p.xs_mode == BZ_XS_DIRECT) {
    if (p.dim == 1)
      return stream ? launch_push_shape<T, BZ_XS_DIRECT, 1, BZ_RNG_STREAM>(env, a, 0, grid)
                    : launch_push_shape<T, BZ_XS_DIRECT, 1, BZ_RNG_PHILOX>(env, a, shape, grid);
    return stream ? launch_push_shape<T, BZ_XS_DIRECT, 0, BZ_RNG_STREAM>(env, a, 0, grid)
                  : launch_push_shape<T, BZ_XS_DIRECT, 0, BZ_RNG_PHILOX>(env, a, shape, grid);
  }
  if (p.xs_mode == BZ_XS_TABLE) {
    if (p.dim == 1)
      return stream ? launch_push_shape<T, BZ_XS_TABLE, 1, BZ_RNG_STREAM>(env, a, 0, grid)
                    : launch_push_shape<T, BZ_XS_TABLE, 1, BZ_RNG_PHILOX>(env, a, shape, grid);
    return stream ? launch_push_shape<T, BZ_XS_TABLE, 0, BZ_RNG_STREAM>(env, a, 0, grid)
                  : launch_push_shape<T, BZ_XS_TABLE, 0, BZ_RNG_PHILOX>(env, a, shape, grid);
  }
  if (p.dim == 1)
    return stream ? launch_push_shape<T, BZ_XS_FIT, 1, BZ_RNG_STREAM>(env, a, 0, grid)
                  : launch_push_shape<T, BZ_XS_FIT, 1, BZ_RNG_PHILOX>(env, a, 0, grid);
  return stream ? launch_push_shape<T, BZ_XS_FIT, 0, BZ_RNG_STREAM>(env, a, 0, grid)
                : launch_push_shape<T, BZ_XS_FIT, 0, BZ_RNG_PHILOX>(env, a, 0, grid);
}

// The stage after the push: births + reduction + counters in one launch
// (in place), or the reduction with the slot count commit (compacting push).
static int launch_push_post(const LaunchEnv& env, const PushArgs& a, const PushLaunch& pl, int m,
                            double* out, double* acc, double advance_dt = 0.0) {
  if (!pl.inplace) return launch_dep_reduce(env, pl.grid, m, out, 1, (int64_t)env.capacity, acc);
  int rpc = 0, nx = 0, ny = 0;
  const int rc = red_geometry(env, pl.grid, m, &rpc, &nx, &ny);
  if (rc) return rc;
  const unsigned coll_blocks =
      a.coll_lo ? (unsigned)((a.p.n_nodes + kRedThreads - 1) / kRedThreads) : 0u;
  const unsigned blocks = (unsigned)(pl.grid + nx * ny + 1) + coll_blocks;
  if (env.precision == BZ_F64)
    hipLaunchKernelGGL(k_push_post<double>, dim3(blocks), dim3(kRedThreads), 0, env.stream, a,
                       pl.tile, pl.nw, pl.grid, m, rpc, nx, ny, env.dep_l2, env.dep_cols_done,
                       out, acc, advance_dt);
  else
    hipLaunchKernelGGL(k_push_post<float>, dim3(blocks), dim3(kRedThreads), 0, env.stream, a,
                       pl.tile, pl.nw, pl.grid, m, rpc, nx, ny, env.dep_l2, env.dep_cols_done,
                       out, acc, advance_dt);
  return check_launch("k_push_post");
}

int launch_push(const LaunchEnv& env, const bz_state_t* in, const bz_state_t* out,
                const void* E, const bz_params_t* p, const bz_xsec_t* xs, const void* R,
                const void* G, double* moments, int shape) {
  PushArgs a;
  a.p = *p;
  if (xs) a.xs = *xs;
  else memset(&a.xs, 0, sizeof(a.xs));
  const void* iv[5] = {in->x, in->vx, in->vy, in->vz, in->w};
  void* ov[5] = {out->x, out->vx, out->vy, out->vz, out->w};
  for (int k = 0; k < 5; ++k) {
    a.in[k] = iv[k];
    a.out[k] = ov[k];
  }
  a.E = E;
  a.R = R;
  a.G = G;
  a.clk = env.clock;
  a.status = env.status;
  a.partials = env.partials;
  a.capacity = env.capacity;
  a.inst_off = nullptr;
  a.inst_E = nullptr;
  a.n_inst = 0;
  a.keys = nullptr;
  a.keys16 = 0;
  a.e_lds_off = 0;
  a.m1_lo = p->dim == 1 ? env.dep_m1_lo : nullptr;
  a.m1_hi = p->dim == 1 ? env.dep_m1_hi : nullptr;
  const bool coll = p->dim == 1 && env.coll_on;
  a.coll_lo = coll ? env.coll_lo : nullptr;
  a.coll_hi = coll ? env.coll_hi : nullptr;
  a.coll_cnt = coll ? env.coll_cnt : nullptr;
  a.coll_acc = coll ? env.coll_acc : nullptr;
  if (p->dim == 1 && env.push_keys_nb > 0) {
    a.keys = (uint32_t*)sort_input_keys(env);
    a.keys16 = env.push_keys16;
    a.bk = bin_keys(p, env.push_keys_thr, env.push_keys_nb);
  }
  PushLaunch pl;
  const int rc = env.precision == BZ_F64 ? launch_push_prec<double>(env, a, shape, &pl)
                                         : launch_push_prec<float>(env, a, shape, &pl);
  if (rc) return rc;
  return launch_push_post(env, a, pl, 3 * (p->dim == 1 ? p->n_nodes : 1), moments, nullptr);
}

template <typename T, int DIM>
static int launch_deposit_t(const LaunchEnv& env, const bz_state_t* s, const bz_params_t* p,
                            double* moments) {
  constexpr int BLOCK = 256;
  auto kern = k_deposit<T, DIM, BLOCK>;
  const int nn = DIM == 1 ? p->n_nodes : 1;
  const size_t smem = (size_t)inplace_lds_doubles(1, nn, 0, (int)sizeof(T)) * sizeof(double);
  if (smem > 64 * 1024) {
    if (hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)smem) != hipSuccess)
      return fail("hipFuncSetAttribute(dynamic LDS) failed");
  }
  int grid = env.num_cu * 2;
  const int64_t cap_parts = env.partials_doubles / (3 * nn);
  if (grid > cap_parts) grid = (int)cap_parts;
  hipLaunchKernelGGL(kern, dim3(grid), dim3(BLOCK), smem, env.stream, *p, (const T*)s->x,
                     (const T*)s->vx, (const T*)s->vy, (const T*)s->vz, (const T*)s->w,
                     env.clock, env.partials);
  int rc = check_launch("k_deposit");
  if (rc) return rc;
  return launch_dep_reduce(env, grid, 3 * nn, moments, 0, (int64_t)env.capacity);
}

template <typename T>
static int launch_compact_t(const LaunchEnv& env, const bz_state_t* in, const bz_state_t* out) {
  int grid = env.num_cu * 8;
  if (grid > 2 * env.status_words) grid = (int)(2 * env.status_words);
  uint32_t* cnt = (uint32_t*)env.status;
  hipLaunchKernelGGL(k_compact_count<T>, dim3(grid), dim3(kCompBlock), 0, env.stream,
                     (const T*)in->w, (const Clock*)env.clock, cnt);
  int rc = check_launch("k_compact_count");
  if (rc) return rc;
  hipLaunchKernelGGL(k_compact_scatter<T>, dim3(grid), dim3(kCompBlock), 0, env.stream, *in,
                     *out, env.clock, (const uint32_t*)cnt);
  rc = check_launch("k_compact_scatter");
  if (rc) return rc;
  hipLaunchKernelGGL(k_clock_commit_next, dim3(1), dim3(1), 0, env.stream, env.clock);
  return check_launch("k_clock_commit_next");
}

int launch_collision_read(const LaunchEnv& env, double* out, int nn, int reset) {
  hipLaunchKernelGGL(k_coll_read, dim3((3 * nn + 255) / 256), dim3(256), 0, env.stream,
                     env.coll_acc, out, nn, reset);
  return check_launch("k_coll_read");
}

int launch_compact(const LaunchEnv& env, const bz_state_t* in, const bz_state_t* out) {
  return env.precision == BZ_F64 ? launch_compact_t<double>(env, in, out)
                                 : launch_compact_t<float>(env, in, out);
}

int launch_dep_wmax(const LaunchEnv& env, const bz_state_t* s) {
  if (env.precision == BZ_F64)
    hipLaunchKernelGGL(k_dep_wmax<double>, dim3(env.num_cu * 2), dim3(256), 0, env.stream,
                       (const double*)s->w, env.clock);
  else
    hipLaunchKernelGGL(k_dep_wmax<float>, dim3(env.num_cu * 2), dim3(256), 0, env.stream,
                       (const float*)s->w, env.clock);
  return check_launch("k_dep_wmax");
}

int launch_deposit(const LaunchEnv& env, const bz_state_t* s, const bz_params_t* p,
                   double* moments) {
  if (env.precision == BZ_F64)
    return p->dim == 1 ? launch_deposit_t<double, 1>(env, s, p, moments)
                       : launch_deposit_t<double, 0>(env, s, p, moments);
  return p->dim == 1 ? launch_deposit_t<float, 1>(env, s, p, moments)
                     : launch_deposit_t<float, 0>(env, s, p, moments);
}

}  // namespace bz

// Collision uniforms of the Philox stream written out as R[c*stride + i]
// (columns 0,1,2,3,5; 4 and 6 zero): the numbers the in-kernel stream draws.
namespace bz {

template <typename T>
__global__ void k_fill_uniforms(T* R, int64_t stride, int64_t n, uint64_t seed, uint64_t step) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    Uniforms<T> U;
    U.u[0] = Decision<T>{}.get(seed, (uint32_t)i, step);
    U.scatter(seed, (uint32_t)i, step, true);
    R[i] = U.u[0];
    R[stride + i] = U.u[1];
    R[2 * stride + i] = U.u[2];
    R[3 * stride + i] = U.u[3];
    R[4 * stride + i] = (T)0;
    R[5 * stride + i] = U.u[4];
    R[6 * stride + i] = (T)0;
  }
}

int launch_fill_uniforms(const LaunchEnv& env, void* R, int64_t stride, int64_t n, uint64_t seed,
                         uint64_t step) {
  if (env.precision == BZ_F64)
    hipLaunchKernelGGL(k_fill_uniforms<double>, dim3(env.num_cu * 4), dim3(256), 0, env.stream,
                       (double*)R, stride, n, seed, step);
  else
    hipLaunchKernelGGL(k_fill_uniforms<float>, dim3(env.num_cu * 4), dim3(256), 0, env.stream,
                       (float*)R, stride, n, seed, step);
  return check_launch("k_fill_uniforms");
}

}  // namespace bz

// ---------------------------------------------------------------------------
// Batched 0-D swarms: the fused step over all instances in one launch, then
// the per-instance moments summed over blocks (k_dep_reduce commits the clock).
// ---------------------------------------------------------------------------
namespace bz {

int launch_push_swarm(const LaunchEnv& env, const bz_state_t* in, const bz_state_t* out,
                      const bz_params_t* p, const bz_xsec_t* xs, int n_inst, const void* inst_E,
                      double* inst_moments, double* acc, int shape) {
  PushArgs a;
  a.p = *p;
  a.xs = *xs;
  const void* iv[5] = {in->x, in->vx, in->vy, in->vz, in->w};
  void* ov[5] = {out->x, out->vx, out->vy, out->vz, out->w};
  for (int k = 0; k < 5; ++k) {
    a.in[k] = iv[k];
    a.out[k] = ov[k];
  }
  a.E = nullptr;
  a.R = nullptr;
  a.G = nullptr;
  a.clk = env.clock;
  a.status = env.status;
  a.partials = env.partials;
  a.capacity = env.capacity;
  a.inst_off = nullptr;
  a.inst_E = inst_E;
  a.n_inst = n_inst;
  a.keys = nullptr;
  a.keys16 = 0;
  a.e_lds_off = 0;
  a.m1_lo = nullptr;
  a.m1_hi = nullptr;
  a.coll_lo = nullptr;
  a.coll_hi = nullptr;
  a.coll_cnt = nullptr;
  a.coll_acc = nullptr;
  PushLaunch pl;
  const int rc = env.precision == BZ_F64 ? launch_push_prec<double>(env, a, shape, &pl)
                                         : launch_push_prec<float>(env, a, shape, &pl);
  if (rc) return rc;
  // the step ends here (no field stage): the clock advances by dt
  return launch_push_post(env, a, pl, n_inst * kImN, inst_moments, acc, p->dt);
}

}  // namespace bz
