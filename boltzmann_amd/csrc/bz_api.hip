// C ABI of libbz.so: context, workspace, clock and the stage entry points.
#include <stdio.h>
#include <stdlib.h>

#include <new>

#include "bz_launch.h"

namespace bz {

static thread_local char g_err[512] = "";

int fail(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return -1;
}

int check_launch(const char* what) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail("%s: %s", what, hipGetErrorString(e));
  return 0;
}

void timing_begin(const LaunchEnv& env, int kind) {
  Timing* t = env.timing;
  if (!t || !t->on || t->used[kind] >= t->cap) return;
  (void)hipEventRecord(t->start[kind][t->used[kind]], env.stream);
}

void timing_end(const LaunchEnv& env, int kind) {
  Timing* t = env.timing;
  if (!t || !t->on || t->used[kind] >= t->cap) return;
  (void)hipEventRecord(t->stop[kind][t->used[kind]], env.stream);
  t->used[kind] += 1;
  t->launches[kind] += 1;
}

}  // namespace bz

using namespace bz;

struct bz_ctx {
  int device;
  LaunchEnv env;
  Clock* host_clock;  // pinned mirror
  void* graphs;       // captured graphs (GraphSlots), created on first capture
  // state (its x array) whose binning keys the last 1-D push wrote into the
  // sort input, with the binning and energy factor they were made with; any
  // other call that changes a state or the sort workspace clears it
  const void* keys_state;
  BinKeys keys_bk;
  double keys_kev;
  int keys16;  // those keys are 16-bit words
  // the clock's weight bound was taken from the state loaded by the last
  // clock set (merges raise it on the device afterwards)
  bool wmax_valid;
};

static bool same_bins(const BinKeys& a, const BinKeys& b) {
  if (a.dx != b.dx || a.nb != b.nb || a.n_nodes != b.n_nodes || a.n_keys != b.n_keys ||
      a.dim != b.dim || a.ew != b.ew)
    return false;
  for (int k = 0; k < 8; ++k)
    if (a.thr[k] != b.thr[k]) return false;
  return true;
}

static void graph_release(bz_ctx* c);

#define BZ_HIP(call)                                                      \
  do {                                                                    \
    const hipError_t e_ = (call);                                         \
    if (e_ != hipSuccess) return fail("%s: %s", #call, hipGetErrorString(e_)); \
  } while (0)

#define BZ_CTX(ctx)                                        \
  do {                                                     \
    if (!(ctx)) return fail("null context");               \
    BZ_HIP(hipSetDevice((ctx)->device));                   \
  } while (0)

extern "C" {

const char* bz_last_error(void) { return g_err; }
int bz_version(void) { return 1; }

void bz_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  const U4 r = philox(ctr[0], ctr[1], ctr[2], ctr[3], key[0], key[1]);
  out[0] = r.a;
  out[1] = r.b;
  out[2] = r.c;
  out[3] = r.d;
}

int bz_ctx_create(bz_ctx_t* out, int device, int precision, int64_t capacity, int32_t max_nodes,
                  int32_t max_keys) {
  if (!out) return fail("null output pointer");
  *out = nullptr;
  if (precision != BZ_F32 && precision != BZ_F64) return fail("precision must be 4 or 8");
  if (capacity < 1 || capacity > 0x7FFFFFFFll)
    return fail("capacity must be in [1, 2^31-1] (32-bit scan values)");
  if (max_nodes < 1) return fail("max_nodes must be >= 1");
  if (max_keys < 1) max_keys = 1;
  BZ_HIP(hipSetDevice(device));
  hipDeviceProp_t prop;
  BZ_HIP(hipGetDeviceProperties(&prop, device));
  bz_ctx* c = new (std::nothrow) bz_ctx();
  if (!c) return fail("out of host memory");
  memset(c, 0, sizeof(*c));
  c->device = device;
  LaunchEnv& e = c->env;
  e.stream = nullptr;
  e.precision = precision;
  e.num_cu = prop.multiProcessorCount;
  e.capacity = capacity;
  e.max_nodes = max_nodes;
  e.max_keys = max_keys;
  e.push_shape = 0;
  e.push_bpc = 0;
  e.push_queue = 0;
  e.push_keys_nb = 0;
  e.push_keys16 = 0;
  // smallest push tile is 256 particles
  // look-back tile words, or the birth counts of the in-place 1-D push (u32:
  // 4096 per block, then one per tile and wave: at most capacity/64 + 64)
  e.status_words = capacity / 128 + 4096;
  e.partials_doubles = push_partials_doubles(e.num_cu, max_nodes, max_keys);
  const int64_t dep_m = dep_max_row(max_nodes, max_keys);
  e.dep_cols = (int)((dep_m + 63) / 64);
  e.dep_l2_doubles = 16 * dep_m;
  e.sort_ws_bytes = (int64_t)sort_workspace_bytes(capacity, max_keys);
  hipError_t st = hipMalloc((void**)&e.clock, sizeof(Clock));
  if (st == hipSuccess) st = hipMalloc((void**)&e.status, e.status_words * sizeof(uint64_t));
  if (st == hipSuccess) st = hipMalloc((void**)&e.partials, e.partials_doubles * sizeof(double));
  if (st == hipSuccess) st = hipMalloc(&e.sort_ws, e.sort_ws_bytes);
  // the in-place resampling's fire word starts at zero (its candidate kernels
  // clear it after every pass)
  if (st == hipSuccess) st = hipMemset(e.sort_ws, 0, e.sort_ws_bytes);
  if (st == hipSuccess) st = hipMalloc((void**)&e.dep_l2, e.dep_l2_doubles * sizeof(double));
  if (st == hipSuccess) st = hipMalloc((void**)&e.dep_cols_done, e.dep_cols * sizeof(uint32_t));
  if (st == hipSuccess) st = hipMemset(e.dep_cols_done, 0, e.dep_cols * sizeof(uint32_t));
  if (st == hipSuccess)
    st = hipMalloc((void**)&e.dep_m1_lo, (size_t)max_nodes * sizeof(unsigned long long));
  if (st == hipSuccess) st = hipMalloc((void**)&e.dep_m1_hi, (size_t)max_nodes * sizeof(uint32_t));
  if (st == hipSuccess)
    st = hipMemset(e.dep_m1_lo, 0, (size_t)max_nodes * sizeof(unsigned long long));
  if (st == hipSuccess) st = hipMemset(e.dep_m1_hi, 0, (size_t)max_nodes * sizeof(uint32_t));
  // collision counters: 2 fixed-point rows, their carries, the counts, the accumulator
  e.coll_on = 0;
  const size_t coll_bytes = (size_t)max_nodes * (16 + 8 + 8 + 24);
  if (st == hipSuccess) st = hipMalloc((void**)&e.coll_lo, coll_bytes);
  if (st == hipSuccess) st = hipMemset(e.coll_lo, 0, coll_bytes);
  if (st == hipSuccess) {
    e.coll_cnt = e.coll_lo + 2 * (size_t)max_nodes;
    e.coll_acc = (double*)(e.coll_cnt + max_nodes);
    e.coll_hi = (uint32_t*)(e.coll_acc + 3 * (size_t)max_nodes);
  }
  // candidate rings of the queued push (per wave; read only after written)
  e.push_q_waves = (int64_t)e.num_cu * kPushQueueWavesPerCu;
  if (st == hipSuccess)
    st = hipMalloc(&e.push_q, (size_t)e.push_q_waves * kPushQueueWords * (size_t)precision);
  if (st == hipSuccess) st = hipHostMalloc((void**)&c->host_clock, sizeof(Clock));
  if (st == hipSuccess) st = hipMemset(e.clock, 0, sizeof(Clock));
  if (st == hipSuccess) st = hipMemset(e.status, 0, e.status_words * sizeof(uint64_t));
  if (st == hipSuccess) st = hipDeviceSynchronize();
  if (st != hipSuccess) {
    const int rc = fail("bz_ctx_create: %s", hipGetErrorString(st));
    bz_ctx_destroy(c);
    return rc;
  }
  memset(c->host_clock, 0, sizeof(Clock));
  c->host_clock->epoch = 1;
  if (hipMemcpy(e.clock, c->host_clock, sizeof(Clock), hipMemcpyHostToDevice) != hipSuccess) {
    const int rc = fail("bz_ctx_create: clock upload failed");
    bz_ctx_destroy(c);
    return rc;
  }
  *out = c;
  return 0;
}

int bz_ctx_destroy(bz_ctx_t c) {
  if (!c) return 0;
  (void)hipSetDevice(c->device);
  (void)hipDeviceSynchronize();
  if (c->env.clock) (void)hipFree(c->env.clock);
  if (c->env.status) (void)hipFree(c->env.status);
  if (c->env.partials) (void)hipFree(c->env.partials);
  if (c->env.sort_ws) (void)hipFree(c->env.sort_ws);
  if (c->env.dep_l2) (void)hipFree(c->env.dep_l2);
  if (c->env.dep_cols_done) (void)hipFree(c->env.dep_cols_done);
  if (c->env.dep_m1_lo) (void)hipFree(c->env.dep_m1_lo);
  if (c->env.dep_m1_hi) (void)hipFree(c->env.dep_m1_hi);
  if (c->env.coll_lo) (void)hipFree(c->env.coll_lo);
  if (c->env.push_q) (void)hipFree(c->env.push_q);
  if (c->host_clock) (void)hipHostFree(c->host_clock);
  graph_release(c);
  if (Timing* t = c->env.timing) {
    for (int k = 0; k < kTimedKinds; ++k) {
      for (int i = 0; i < t->cap; ++i) {
        hipEventDestroy(t->start[k][i]);
        hipEventDestroy(t->stop[k][i]);
      }
      delete[] t->start[k];
      delete[] t->stop[k];
    }
    delete t;
  }
  delete c;
  return 0;
}

int bz_set_stream(bz_ctx_t c, void* stream) {
  if (!c) return fail("null context");
  c->env.stream = (hipStream_t)stream;
  return 0;
}

int bz_set_push_shape(bz_ctx_t c, int shape) {
  if (!c) return fail("null context");
  if (shape < 0 || shape > 4) return fail("push shape must be 0..4");
  c->env.push_shape = shape;
  return 0;
}

int bz_set_push_grid(bz_ctx_t c, int blocks_per_cu) {
  if (!c) return fail("null context");
  if (blocks_per_cu < 0 || blocks_per_cu > 8) return fail("blocks per CU must be 0..8");
  c->env.push_bpc = blocks_per_cu;
  return 0;
}

int bz_set_push_keys(bz_ctx_t c, int32_t nb, const double* thresholds_host) {
  if (!c) return fail("null context");
  if (nb < 0 || nb > 8) return fail("push keys: nb must be in [0,8]");
  if (nb > 1 && !thresholds_host) return fail("push keys: nb > 1 needs thresholds");
  c->env.push_keys_nb = nb;
  for (int k = 0; k < 8; ++k)
    c->env.push_keys_thr[k] = (thresholds_host && k < nb - 1) ? thresholds_host[k] : 0.0;
  c->keys_state = nullptr;
  return 0;
}

int bz_set_push_keys16(bz_ctx_t c, int32_t on) {
  if (!c) return fail("null context");
  c->env.push_keys16 = on ? 1 : 0;
  c->keys_state = nullptr;
  return 0;
}

int bz_set_push_queue(bz_ctx_t c, int mode) {
  if (!c) return fail("null context");
  if (mode < -1 || mode > 1) return fail("push queue mode must be -1, 0 or 1");
  c->env.push_queue = mode;
  return 0;
}

int bz_set_push_schedule(bz_ctx_t c, int sched) {
  if (!c) return fail("null context");
  if (sched != 0 && sched != 1) return fail("push schedule must be 0 or 1");
  c->env.push_sched = sched;
  return 0;
}

int bz_timing_enable(bz_ctx_t c, int on) {
  BZ_CTX(c);
  if (!c->env.timing) {
    Timing* t = new (std::nothrow) Timing();
    if (!t) return fail("out of host memory");
    memset(t, 0, sizeof(*t));
    t->cap = 2048;
    for (int k = 0; k < kTimedKinds; ++k) {
      t->start[k] = new hipEvent_t[t->cap];
      t->stop[k] = new hipEvent_t[t->cap];
      for (int i = 0; i < t->cap; ++i) {
        BZ_HIP(hipEventCreate(&t->start[k][i]));
        BZ_HIP(hipEventCreate(&t->stop[k][i]));
      }
    }
    c->env.timing = t;
  }
  BZ_HIP(hipStreamSynchronize(c->env.stream));
  Timing* t = c->env.timing;
  if (on) {
    for (int k = 0; k < kTimedKinds; ++k) {
      t->used[k] = 0;
      t->launches[k] = 0;
    }
  }
  t->on = on ? 1 : 0;
  return 0;
}

int bz_timing_get(bz_ctx_t c, double* ms3, int64_t* launches3) {
  BZ_CTX(c);
  BZ_HIP(hipStreamSynchronize(c->env.stream));
  Timing* t = c->env.timing;
  for (int k = 0; k < 3; ++k) {  // the three stage kinds the caller's arrays hold
    double s = 0.0;
    if (t) {
      for (int i = 0; i < t->used[k]; ++i) {
        float ms = 0.f;
        BZ_HIP(hipEventElapsedTime(&ms, t->start[k][i], t->stop[k][i]));
        s += ms;
      }
    }
    if (ms3) ms3[k] = s;
    if (launches3) launches3[k] = t ? t->used[k] : 0;
  }
  return 0;
}

int bz_timing_get_n(bz_ctx_t c, int n, double* ms, int64_t* launches) {
  BZ_CTX(c);
  if (n < 0 || n > kTimedKinds) return fail("timing: at most %d kinds", kTimedKinds);
  BZ_HIP(hipStreamSynchronize(c->env.stream));
  Timing* t = c->env.timing;
  for (int k = 0; k < n; ++k) {
    double s = 0.0;
    if (t) {
      for (int i = 0; i < t->used[k]; ++i) {
        float e = 0.f;
        BZ_HIP(hipEventElapsedTime(&e, t->start[k][i], t->stop[k][i]));
        s += e;
      }
    }
    if (ms) ms[k] = s;
    if (launches) launches[k] = t ? t->used[k] : 0;
  }
  return 0;
}

int bz_synchronize(bz_ctx_t c) {
  BZ_CTX(c);
  BZ_HIP(hipStreamSynchronize(c->env.stream));
  return 0;
}

int bz_clock_set(bz_ctx_t c, int64_t n, uint64_t step, double t) {
  BZ_CTX(c);
  if (n < 0 || n > c->env.capacity) return fail("count %lld outside [0, capacity]", (long long)n);
  c->keys_state = nullptr;
  BZ_HIP(hipStreamSynchronize(c->env.stream));
  BZ_HIP(hipMemcpy(c->host_clock, c->env.clock, sizeof(Clock), hipMemcpyDeviceToHost));
  c->host_clock->n = n;
  c->host_clock->n_next = n;
  c->host_clock->step = step;
  c->host_clock->t = t;
  c->host_clock->ticket = 0;
  c->host_clock->w_max_bits = 0ull;
  BZ_HIP(hipMemcpy(c->env.clock, c->host_clock, sizeof(Clock), hipMemcpyHostToDevice));
  c->wmax_valid = false;
  return 0;
}

// the weight bound of a freshly set state, before its first deposit
static int ensure_wmax(bz_ctx_t c, const bz_state_t* s) {
  if (c->wmax_valid) return 0;
  const int rc = launch_dep_wmax(c->env, s);
  if (!rc) c->wmax_valid = true;
  return rc;
}

int bz_clock_get(bz_ctx_t c, int64_t* n, uint64_t* step, double* t, uint64_t* counters) {
  BZ_CTX(c);
  BZ_HIP(hipMemcpyAsync(c->host_clock, c->env.clock, sizeof(Clock), hipMemcpyDeviceToHost,
                        c->env.stream));
  BZ_HIP(hipStreamSynchronize(c->env.stream));
  if (n) *n = c->host_clock->n;
  if (step) *step = c->host_clock->step;
  if (t) *t = c->host_clock->t;
  if (counters)
    for (int k = 0; k < BZ_NCOUNTERS; ++k) counters[k] = c->host_clock->cnt[k];
  return 0;
}

int bz_counters_reset(bz_ctx_t c) {
  BZ_CTX(c);
  BZ_HIP(hipStreamSynchronize(c->env.stream));
  BZ_HIP(hipMemcpy(c->host_clock, c->env.clock, sizeof(Clock), hipMemcpyDeviceToHost));
  memset(c->host_clock->cnt, 0, sizeof(c->host_clock->cnt));
  BZ_HIP(hipMemcpy(c->env.clock, c->host_clock, sizeof(Clock), hipMemcpyHostToDevice));
  return 0;
}

}  // extern "C"

__global__ void k_clock_advance(Clock* clk, double dt) {
  clk->step += 1;
  clk->t += dt;
}

extern "C" {

int bz_clock_advance(bz_ctx_t c, double dt) {
  BZ_CTX(c);
  hipLaunchKernelGGL(k_clock_advance, dim3(1), dim3(1), 0, c->env.stream, c->env.clock, dt);
  return check_launch("k_clock_advance");
}

static int check_params(const bz_ctx* c, const bz_params_t* p) {
  if (!p) return fail("null params");
  if (p->dim != 0 && p->dim != 1) return fail("dim must be 0 or 1");
  if (p->dim == 1) {
    if (p->n_cells < 1 || p->n_nodes < 1) return fail("n_cells / n_nodes must be >= 1");
    if (p->n_nodes > c->env.max_nodes)
      return fail("n_nodes %d exceeds context max_nodes %d", p->n_nodes, c->env.max_nodes);
    if (p->bnf < 1) return fail("bnf must be >= 1");
    if (!(p->dx > 0.0) || !(p->L > 0.0)) return fail("dx and L must be > 0");
  }
  if (p->xs_mode != BZ_XS_FIT && p->xs_mode != BZ_XS_TABLE && p->xs_mode != BZ_XS_FLIGHT &&
      p->xs_mode != BZ_XS_DIRECT)
    return fail("bad xs_mode");
  if (p->xs_mode == BZ_XS_FLIGHT && !(p->nu_max > 0.0))
    return fail("free-flight collisions need nu_max > 0");
  if (p->rng_mode != BZ_RNG_PHILOX && p->rng_mode != BZ_RNG_STREAM) return fail("bad rng_mode");
  if (p->push_order != BZ_PUSH_REF && p->push_order != BZ_PUSH_LEAPFROG)
    return fail("bad push_order");
  return 0;
}

int bz_push_mcc(bz_ctx_t c, const bz_state_t* in, const bz_state_t* out, const void* E,
                const bz_params_t* p, const bz_xsec_t* xs, const void* R, const void* gas_v,
                double* moments) {
  BZ_CTX(c);
  int rc = check_params(c, p);
  if (rc) return rc;
  if (!in || !out || !moments) return fail("null state or moments pointer");
  if (in->x == out->x || in->w == out->w) return fail("in and out state must not alias");
  if (p->dim == 1 && !E) return fail("1-D push needs the field array E");
  if (p->xs_mode != BZ_XS_FIT) {
    if (!xs || !xs->energy || !xs->sig || !xs->slope || !xs->hint || xs->n < 2 ||
        xs->n_hint < 1)
      return fail("table mode needs a cross-section table");
    if (xs->hint_bits < 1 || xs->hint_bits > 20) return fail("hint_bits must be in [1,20]");
  }
  if (p->rng_mode == BZ_RNG_STREAM) {
    if (!R) return fail("stream mode needs the uniform array R");
    if (p->r_stride < 1) return fail("stream mode needs r_stride >= 1");
  }
  c->keys_state = nullptr;
  if (p->dim == 1 && c->env.push_keys_nb > 0 && c->env.push_keys16 &&
      (int64_t)p->n_nodes * c->env.push_keys_nb >= 65535)
    return fail("16-bit push keys: n_nodes * nb must be below 65535");
  rc = ensure_wmax(c, in);
  if (rc) return rc;
  rc = launch_push(c->env, in, out, E, p, xs, R, gas_v, moments, c->env.push_shape);
  if (!rc && p->dim == 1 && c->env.push_keys_nb > 0) {
    c->keys_state = in->x;
    c->keys_bk = bin_keys(p, c->env.push_keys_thr, c->env.push_keys_nb);
    c->keys_kev = p->kev;
    c->keys16 = c->env.push_keys16;
  }
  return rc;
}

int bz_deposit(bz_ctx_t c, const bz_state_t* s, const bz_params_t* p, double* moments) {
  BZ_CTX(c);
  int rc = check_params(c, p);
  if (rc) return rc;
  if (!s || !moments) return fail("null state or moments pointer");
  rc = ensure_wmax(c, s);
  if (rc) return rc;
  return launch_deposit(c->env, s, p, moments);
}

int bz_compact(bz_ctx_t c, const bz_state_t* in, const bz_state_t* out) {
  BZ_CTX(c);
  if (!in || !out) return fail("null state");
  if (in->w == out->w) return fail("bz_compact needs two distinct state buffers");
  c->keys_state = nullptr;
  return launch_compact(c->env, in, out);
}

int bz_bin_lists(bz_ctx_t c, const double* tosum, int64_t tosum_cols, const int64_t* curr_xbins,
                 int64_t nc, int32_t n_cells, int32_t nb, const double* thresholds_host,
                 int64_t* const* lists, int64_t* const* write_inds, const int64_t* max_ppc) {
  BZ_CTX(c);
  if (nb < 1 || nb > 8) return fail("nb must be in [1,8]");
  if (nc < 0) return fail("negative row count");
  if (n_cells < 1) return fail("n_cells must be >= 1");
  if (tosum_cols < 4) return fail("tosum needs at least 4 columns (w ... energy)");
  if (nc > 0 && (!tosum || !curr_xbins)) return fail("null tosum / curr_xbins");
  if (!lists || !write_inds || !max_ppc || (nb > 1 && !thresholds_host))
    return fail("null lists, write_inds, max_ppc or thresholds");
  for (int b = 0; b < nb; ++b) {
    if (!write_inds[b]) return fail("null write_inds[%d]", b);
    if (max_ppc[b] < 0) return fail("negative max_ppc[%d]", b);
    if (max_ppc[b] > 0 && !lists[b]) return fail("null list[%d]", b);
  }
  c->keys_state = nullptr;
  return launch_bin_lists(c->env, tosum, tosum_cols, curr_xbins, nc, n_cells, nb,
                          thresholds_host, lists, write_inds, max_ppc);
}

int bz_set_collision_counts(bz_ctx_t c, int32_t on) {
  BZ_CTX(c);
  c->env.coll_on = on ? 1 : 0;
  return 0;
}

int bz_collision_counts(bz_ctx_t c, double* out, int32_t n_nodes, int32_t reset) {
  BZ_CTX(c);
  if (!out) return fail("null output");
  if (n_nodes < 1 || n_nodes > c->env.max_nodes) return fail("n_nodes outside [1, max_nodes]");
  return launch_collision_read(c->env, out, n_nodes, reset);
}

int bz_compact_inplace(bz_ctx_t c, const bz_state_t* s) {
  BZ_CTX(c);
  if (!s || !s->w) return fail("null state");
  c->keys_state = nullptr;
  return launch_fill_holes(c->env, s);
}

int bz_field_step(bz_ctx_t c, const bz_field_params_t* fp, const double* moments, double* ni,
                  double* ns, double* Te, double* g2, double* V, double* E, void* E_t,
                  const double* chol, int32_t advance_time) {
  BZ_CTX(c);
  if (!fp) return fail("null field params");
  if (fp->n < 3) return fail("field needs N >= 3 cells");
  if (fp->n + 1 > c->env.max_nodes) return fail("N+1 exceeds context max_nodes");
  if (!moments || !ni || !ns || !Te || !g2 || !V || !E || !E_t || !chol)
    return fail("null field array");
  return launch_field(c->env, fp, moments, ni, ns, Te, g2, V, E, E_t, chol, advance_time);
}

int bz_field_substep(bz_ctx_t c, const bz_field_params_t* fp, double* moments,
                     const double* src, double* ni, double* ns, double* Te, double* g2,
                     double* V, double* E, void* E_t, const double* chol) {
  BZ_CTX(c);
  if (!fp) return fail("null field params");
  if (fp->n < 3) return fail("field needs N >= 3 cells");
  if (fp->n + 1 > c->env.max_nodes) return fail("N+1 exceeds context max_nodes");
  if (!moments || !src || !ni || !ns || !Te || !g2 || !V || !E || !E_t || !chol)
    return fail("null field array");
  bz_field_params_t q = *fp;
  q.skip_moments = 1;
  return launch_field(c->env, &q, moments, ni, ns, Te, g2, V, E, E_t, chol, 1, src);
}

int bz_closure_source(bz_ctx_t c, const bz_mlp_t* net, int32_t n_cells, const double* E,
                      const double* moments, double gas_density_m3, double dt, double* src,
                      double* acc) {
  BZ_CTX(c);
  if (!net || !net->w) return fail("closure source: null network");
  if (n_cells < 1) return fail("closure source: need cells");
  if (!E || !moments || !src) return fail("closure source: null pointer");
  if (!(gas_density_m3 > 0.0) || !(dt > 0.0)) return fail("closure source: need N_gas, dt > 0");
  return launch_closure_source(c->env, net, n_cells, E, moments, gas_density_m3, dt, src, acc);
}

int bz_closure_commit(bz_ctx_t c, int32_t n_nodes, const double* moments, double* acc,
                      double* ni, double* g2, double* Te, double* tally) {
  BZ_CTX(c);
  if (n_nodes < 1) return fail("closure commit: need nodes");
  if (!moments || !acc || !ni || !g2 || !Te || !tally) return fail("closure commit: null pointer");
  return launch_closure_commit(c->env, n_nodes, moments, acc, ni, g2, Te, tally);
}

int bz_bin_count(bz_ctx_t c, const bz_state_t* s, const bz_params_t* p,
                 const double* thresholds_host, int32_t nb, int32_t* counts) {
  BZ_CTX(c);
  int rc = check_params(c, p);
  if (rc) return rc;
  if (nb < 1 || nb > 8) return fail("nb must be in [1,8]");
  if (nb > 1 && !thresholds_host) return fail("nb > 1 needs thresholds");
  const int64_t n_keys = (int64_t)(p->dim == 1 ? p->n_nodes : 1) * nb;
  if (n_keys > c->env.max_keys) return fail("n_keys %lld exceeds context max_keys", (long long)n_keys);
  if (!s || !counts) return fail("null bin/count pointer");
  return launch_bin_count(c->env, s, p, thresholds_host, nb, counts);
}

int bz_bin_sort(bz_ctx_t c, const bz_state_t* s, const bz_params_t* p,
                const double* thresholds_host, int32_t nb, int32_t* counts, int32_t* offsets,
                uint32_t* perm) {
  BZ_CTX(c);
  int rc = check_params(c, p);
  if (rc) return rc;
  if (nb < 1 || nb > 8) return fail("nb must be in [1,8]");
  if (nb > 1 && !thresholds_host) return fail("nb > 1 needs thresholds");
  const int64_t n_keys = (int64_t)(p->dim == 1 ? p->n_nodes : 1) * nb;
  if (n_keys > c->env.max_keys) return fail("n_keys %lld exceeds context max_keys", (long long)n_keys);
  if (n_keys > 65534) return fail("n_keys must be <= 65534 (two 8-bit radix passes)");
  if (!s || !counts || !offsets || !perm) return fail("null bin/sort pointer");
  // the keys the last push wrote serve when they are this state's, made
  // with this binning
  const bool ready = c->keys_state && c->keys_state == s->x && c->keys_kev == p->kev &&
                     !c->keys16 && same_bins(c->keys_bk, bin_keys(p, thresholds_host, nb));
  c->keys_state = nullptr;  // the sort overwrites its input keys
  return launch_bin_sort(c->env, s, p, thresholds_host, nb, counts, offsets, perm, ready);
}

int bz_resample(bz_ctx_t c, const bz_state_t* in, const bz_state_t* out, const bz_params_t* p,
                const int32_t* counts, const int32_t* offsets, const uint32_t* perm,
                const int32_t* targets, int32_t n_keys, const bz_resample_policy_t* pol) {
  BZ_CTX(c);
  int rc = check_params(c, p);
  if (rc) return rc;
  if (!in || !out || !counts || !offsets || !perm || !targets || !pol)
    return fail("null resample pointer");
  if (in->x == out->x) return fail("in and out state must not alias");
  if (n_keys < 1 || n_keys > c->env.max_keys) return fail("bad n_keys");
  c->keys_state = nullptr;
  return launch_resample(c->env, in, out, p, counts, offsets, perm, targets, n_keys, pol);
}

int bz_resample_inplace(bz_ctx_t c, const bz_state_t* s, const bz_params_t* p,
                        const double* thresholds_host, int32_t nb, const int32_t* targets,
                        int32_t n_keys, const bz_resample_policy_t* pol) {
  BZ_CTX(c);
  int rc = check_params(c, p);
  if (rc) return rc;
  if (!s || !targets || !pol) return fail("null resample pointer");
  if (nb < 1 || nb > 8) return fail("nb must be in [1,8]");
  if (nb > 1 && !thresholds_host) return fail("nb > 1 needs thresholds");
  const int64_t nk = (int64_t)(p->dim == 1 ? p->n_nodes : 1) * nb;
  if (n_keys != nk) return fail("n_keys must be n_nodes * nb = %lld", (long long)nk);
  if (n_keys < 1 || n_keys > c->env.max_keys) return fail("bad n_keys");
  if (pol->uniforms && pol->u_stride < 1) return fail("resample uniforms need u_stride >= 1");
  const bool ready = c->keys_state && c->keys_state == s->x && c->keys_kev == p->kev &&
                     same_bins(c->keys_bk, bin_keys(p, thresholds_host, nb));
  const int keys16 = ready ? c->keys16 : 0;  // launch_slot_keys writes 32-bit keys
  c->keys_state = nullptr;  // the state changes
  if (!ready) {
    rc = launch_slot_keys(c->env, s, p, thresholds_host, nb);
    if (rc) return rc;
  }
  return launch_resample_inplace(c->env, s, p, sort_input_keys(c->env), keys16, nb, targets,
                                 n_keys, pol);
}

}  // extern "C"

extern "C" int bz_reduction(bz_ctx_t c, int32_t r, int32_t M, int32_t RM, double* dw,
                            const double* dv, const double* dx, int64_t ptcl) {
  BZ_CTX(c);
  if (r < 1 || M < 1 || RM < 1 || RM > M) return fail("reduction: need r, M >= 1, 1 <= RM <= M");
  if (ptcl < 0) return fail("reduction: negative particle count");
  if (!dw || (ptcl > 0 && (!dv || !dx))) return fail("reduction: null pointer");
  return launch_reduction(c->env, r, M, RM, dw, dv, dx, ptcl);
}

namespace bz {
int launch_fill_uniforms(const LaunchEnv& env, void* R, int64_t stride, int64_t n, uint64_t seed,
                         uint64_t step);
}

extern "C" int bz_fill_uniforms(bz_ctx_t c, void* R, int64_t stride, int64_t n, uint64_t seed,
                                uint64_t step) {
  BZ_CTX(c);
  if (n < 0 || stride < n) return fail("fill_uniforms: need 0 <= n <= stride");
  if (n > 0 && !R) return fail("fill_uniforms: null array");
  return launch_fill_uniforms(c->env, R, stride, n, seed, step);
}

// ---------------------------------------------------------------------------
// Batched swarms
// ---------------------------------------------------------------------------
namespace bz {
int launch_swarm_step(const LaunchEnv& env, const bz_state_t* in, const bz_state_t* out,
                      const bz_params_t* p, const bz_xsec_t* xs, int n_inst, int64_t* off,
                      const void* inst_E, double* mom, double* acc, int accumulate);
int launch_swarm_control(const LaunchEnv& env, const bz_state_t* in, const bz_state_t* out,
                         const bz_params_t* p, int n_inst, int64_t* off, const int32_t* targets,
                         int32_t* counts, int32_t* offsets, const bz_resample_policy_t* pol);
int launch_swarm_eedf(const LaunchEnv& env, const bz_state_t* s, int n_inst, const int64_t* off,
                      int n_bins, double e_lo, double e_hi, double kev, double* out);
}  // namespace bz

extern "C" {

int bz_swarm_step(bz_ctx_t c, const bz_state_t* in, const bz_state_t* out, const bz_params_t* p,
                  const bz_xsec_t* xs, int32_t n_inst, int64_t* inst_off, const void* inst_E,
                  double* inst_moments, double* inst_acc, int32_t accumulate) {
  BZ_CTX(c);
  if (!p || p->dim != 0) return fail("swarm step: params must be 0-D");
  if (p->xs_mode != BZ_XS_TABLE || p->rng_mode != BZ_RNG_PHILOX)
    return fail("swarm step: tabulated model and Philox stream only");
  if (n_inst < 1 || n_inst > 1024) return fail("swarm step: 1 <= n_inst <= 1024");
  if (!in || !out || !xs || !inst_E || !inst_moments) return fail("swarm step: null pointer");
  if (accumulate && !inst_acc) return fail("swarm step: accumulate needs inst_acc");
  if (in->vx == out->vx) return fail("in and out state must not alias");
  c->keys_state = nullptr;
  return launch_swarm_step(c->env, in, out, p, xs, n_inst, inst_off, inst_E, inst_moments,
                           inst_acc, accumulate);
}

int bz_swarm_control(bz_ctx_t c, const bz_state_t* in, const bz_state_t* out,
                     const bz_params_t* p, int32_t n_inst, int64_t* inst_off,
                     const int32_t* targets, int32_t* counts_ws, int32_t* offsets_ws,
                     const bz_resample_policy_t* pol) {
  BZ_CTX(c);
  if (n_inst < 1 || n_inst > c->env.max_keys) return fail("swarm control: bad n_inst");
  if (!in || !out || !p || !inst_off || !targets || !counts_ws || !offsets_ws || !pol)
    return fail("swarm control: null pointer");
  if (in->vx == out->vx) return fail("in and out state must not alias");
  c->keys_state = nullptr;
  return launch_swarm_control(c->env, in, out, p, n_inst, inst_off, targets, counts_ws,
                              offsets_ws, pol);
}

int bz_mlp_forward(bz_ctx_t c, const bz_mlp_t* net, const float* x, float* y, int64_t n) {
  BZ_CTX(c);
  if (!net || !net->w) return fail("mlp forward: null network");
  if (n < 0) return fail("mlp forward: negative size");
  if (n > 0 && (!x || !y)) return fail("mlp forward: null pointer");
  for (int i = 0; i < 3; ++i)
    if (!(net->out_scale[i] != 0.0)) return fail("mlp: out_scale must be non-zero");
  return launch_mlp_forward(c->env, net, x, y, n);
}

int64_t bz_mlp_workspace_floats(int64_t n) { return n > 0 ? mlp_workspace_floats(n) : 0; }

int bz_mlp_grad(bz_ctx_t c, const bz_mlp_t* net, const float* x, const float* t, int64_t n,
                float* ws, float* grad, float* loss) {
  BZ_CTX(c);
  if (!net || !net->w) return fail("mlp grad: null network");
  if (n < 1) return fail("mlp grad: empty batch");
  if (!x || !t || !ws || !grad) return fail("mlp grad: null pointer");
  for (int i = 0; i < 3; ++i)
    if (!(net->out_scale[i] != 0.0)) return fail("mlp: out_scale must be non-zero");
  return launch_mlp_grad(c->env, net, x, t, n, ws, grad, loss);
}

int bz_mlp_adam(bz_ctx_t c, const bz_adam_t* opt, float* master, float* m, float* v,
                const float* grad, uint16_t* w) {
  BZ_CTX(c);
  if (!opt || !master || !m || !v || !grad || !w) return fail("mlp adam: null pointer");
  if (!(opt->bias1 > 0.f) || !(opt->bias2 > 0.f)) return fail("mlp adam: bias terms must be > 0");
  return launch_mlp_adam(c->env, opt, master, m, v, grad, w);
}

int bz_mlp_round(bz_ctx_t c, const float* master, uint16_t* w) {
  BZ_CTX(c);
  if (!master || !w) return fail("mlp round: null pointer");
  return launch_mlp_round(c->env, master, w);
}

int bz_swarm_eedf(bz_ctx_t c, const bz_state_t* s, int32_t n_inst, const int64_t* inst_off,
                  int32_t n_bins, double e_lo, double e_hi, double kev, double* out) {
  BZ_CTX(c);
  if (n_inst < 1 || n_bins < 1 || n_bins > 8192) return fail("swarm eedf: bad sizes");
  if (!(e_lo > 0.0) || !(e_hi > e_lo)) return fail("swarm eedf: need 0 < e_lo < e_hi");
  if (!s || !inst_off || !out) return fail("swarm eedf: null pointer");
  return launch_swarm_eedf(c->env, s, n_inst, inst_off, n_bins, e_lo, e_hi, kev, out);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// hipGraph capture of enqueued stages (device clock makes replays valid)
// ---------------------------------------------------------------------------
namespace {
constexpr int kMaxGraphs = 64;
struct GraphSlots {
  hipGraphExec_t exec[kMaxGraphs];
  int used[kMaxGraphs];
  int saved_sched;
  int saved_timing;
  int capturing;
};
GraphSlots* slots_of(bz_ctx* c) {
  if (!c->graphs) {
    GraphSlots* s = new (std::nothrow) GraphSlots();
    if (!s) return nullptr;
    memset(s, 0, sizeof(GraphSlots));
    c->graphs = s;
  }
  return (GraphSlots*)c->graphs;
}
}  // namespace

static void graph_release(bz_ctx* c) {
  GraphSlots* g = (GraphSlots*)c->graphs;
  if (!g) return;
  for (int k = 0; k < kMaxGraphs; ++k)
    if (g->used[k]) hipGraphExecDestroy(g->exec[k]);
  delete g;
  c->graphs = nullptr;
}

extern "C" {

int bz_graph_begin(bz_ctx_t c) {
  BZ_CTX(c);
  if (!c->env.stream) return fail("graph capture needs a non-default stream (bz_set_stream)");
  GraphSlots* g = slots_of(c);
  if (!g) return fail("too many contexts with graphs on this thread");
  if (g->capturing) return fail("capture already in progress");
  c->keys_state = nullptr;  // keys are reused within one capture only
  g->saved_sched = c->env.push_sched;
  c->env.push_sched = 1;  // cooperative launches are not captured: ticket schedule
  g->saved_timing = c->env.timing ? c->env.timing->on : 0;
  if (c->env.timing) c->env.timing->on = 0;
  BZ_HIP(hipStreamBeginCapture(c->env.stream, hipStreamCaptureModeRelaxed));
  g->capturing = 1;
  return 0;
}

int bz_graph_end(bz_ctx_t c, int32_t* id) {
  BZ_CTX(c);
  GraphSlots* g = slots_of(c);
  if (!g || !g->capturing) return fail("no capture in progress");
  hipGraph_t graph = nullptr;
  const hipError_t e = hipStreamEndCapture(c->env.stream, &graph);
  g->capturing = 0;
  c->keys_state = nullptr;
  c->env.push_sched = g->saved_sched;
  if (c->env.timing) c->env.timing->on = g->saved_timing;
  if (e != hipSuccess) return fail("hipStreamEndCapture: %s", hipGetErrorString(e));
  int slot = -1;
  for (int k = 0; k < kMaxGraphs; ++k)
    if (!g->used[k]) {
      slot = k;
      break;
    }
  if (slot < 0) {
    hipGraphDestroy(graph);
    return fail("no free graph slot");
  }
  const hipError_t e2 = hipGraphInstantiate(&g->exec[slot], graph, nullptr, nullptr, 0);
  hipGraphDestroy(graph);
  if (e2 != hipSuccess) return fail("hipGraphInstantiate: %s", hipGetErrorString(e2));
  g->used[slot] = 1;
  if (id) *id = slot;
  return 0;
}

int bz_graph_launch(bz_ctx_t c, int32_t id) {
  BZ_CTX(c);
  GraphSlots* g = slots_of(c);
  if (!g || id < 0 || id >= kMaxGraphs || !g->used[id]) return fail("bad graph id");
  c->keys_state = nullptr;  // a replay may change states and the sort input
  BZ_HIP(hipGraphLaunch(g->exec[id], c->env.stream));
  return 0;
}

int bz_graph_destroy(bz_ctx_t c, int32_t id) {
  BZ_CTX(c);
  GraphSlots* g = slots_of(c);
  if (!g || id < 0 || id >= kMaxGraphs || !g->used[id]) return fail("bad graph id");
  hipGraphExecDestroy(g->exec[id]);
  g->used[id] = 0;
  return 0;
}

}  // extern "C"
