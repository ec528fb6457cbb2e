// Batched 0-D swarms (many independent electron swarms at different E/N in
// one launch), their population control and diagnostics.
//
// The state of all instances is one SoA array; instance k owns particles
// [off[k], off[k+1]).  The fused push (DIM 2) looks each particle's instance
// up in LDS, applies that instance's field and sums per-instance moments;
// births stay beside their parents, so instances remain contiguous and the
// new offsets are the old ones plus the per-instance birth counts.
#include "bz_launch.h"

namespace bz {

constexpr int kImNs = 8;  // moments per instance (layout of bz_push.hip)
constexpr int kImBirthS = 6;

// new offsets from the birth counts; running sums of the moments
__global__ __launch_bounds__(1024) void k_swarm_offsets(int n_inst, int64_t* off,
                                                        const double* __restrict__ mom,
                                                        double* acc, int accumulate,
                                                        int64_t cap) {
  __shared__ int64_t s_w[16];
  __shared__ int64_t s_carry;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  if (tid == 0) s_carry = 0;
  if (accumulate && acc)
    for (int j = tid; j < n_inst * kImNs; j += 1024) acc[j] += mom[j];
  __syncthreads();
  for (int base = 0; base < n_inst; base += 1024) {
    const int k = base + tid;
    int64_t c = 0;
    if (k < n_inst) c = off[k + 1] - off[k] + (int64_t)mom[k * kImNs + kImBirthS];
    const int64_t inc = wave_incl_scan(c);
    if (lane == 63) s_w[wid] = inc;
    __syncthreads();
    if (wid == 0) {
      int64_t v = lane < 16 ? s_w[lane] : 0;
      v = wave_incl_scan(v);
      if (lane < 16) s_w[lane] = v;
    }
    __syncthreads();
    const int64_t excl = s_carry + (wid ? s_w[wid - 1] : 0) + inc - c;
    __syncthreads();  // every old offset of this chunk has been read
    if (k < n_inst) off[k] = excl < cap ? excl : cap;
    if (tid == 0) s_carry += s_w[15];
    __syncthreads();
  }
  if (tid == 0) off[n_inst] = s_carry < cap ? s_carry : cap;
}

// counts / offsets / keys-per-position / identity permutation of the grouped state
__global__ __launch_bounds__(256) void k_swarm_prep(int n_inst, const int64_t* off,
                                                    int32_t* counts, int32_t* offsets,
                                                    uint32_t* skeys, uint32_t* perm) {
  const int k = blockIdx.x;
  const int64_t a = off[k], b = off[k + 1];
  if (threadIdx.x == 0) {
    counts[k] = (int32_t)(b - a);
    offsets[k] = (int32_t)a;
    if (k == n_inst - 1) {
      counts[n_inst] = 0;
      offsets[n_inst] = (int32_t)b;
      offsets[n_inst + 1] = (int32_t)b;
    }
  }
  for (int64_t s = a + threadIdx.x; s < b; s += 256) {
    skeys[s] = (uint32_t)k;
    perm[s] = (uint32_t)s;
  }
}

__global__ void k_swarm_offsets_from_plan(int n_inst, const int32_t* plan, int64_t* off,
                                          const Clock* clk) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k <= n_inst; k += gridDim.x * blockDim.x)
    off[k] = k < n_inst ? (int64_t)plan[4 * k + 2] : clk->n;
}

// EEDF: per instance, weight per log-spaced energy bin (one block per instance)
template <typename T>
__global__ __launch_bounds__(256) void k_swarm_eedf(const T* __restrict__ vx,
                                                    const T* __restrict__ vy,
                                                    const T* __restrict__ vz,
                                                    const T* __restrict__ w, const int64_t* off,
                                                    int n_bins, double lg_lo, double inv_dlg,
                                                    double kev, double* out) {
  extern __shared__ double s_h[];
  const int k = blockIdx.x;
  for (int j = threadIdx.x; j < n_bins; j += 256) s_h[j] = 0.0;
  __syncthreads();
  const int64_t a = off[k], b = off[k + 1];
  for (int64_t i = a + threadIdx.x; i < b; i += 256) {
    const double e = kev * ((double)vx[i] * vx[i] + (double)vy[i] * vy[i] + (double)vz[i] * vz[i]);
    int bin = e > 0.0 ? (int)floor((log10(e) - lg_lo) * inv_dlg) : 0;
    bin = bin < 0 ? 0 : (bin >= n_bins ? n_bins - 1 : bin);
    atomicAdd(&s_h[bin], (double)w[i]);
  }
  __syncthreads();
  for (int j = threadIdx.x; j < n_bins; j += 256) out[(size_t)k * n_bins + j] = s_h[j];
}

int launch_swarm_step(const LaunchEnv& env, const bz_state_t* in, const bz_state_t* out,
                      const bz_params_t* p, const bz_xsec_t* xs, int n_inst, int64_t* off,
                      const void* inst_E, double* mom, double* acc, int accumulate) {
  int rc = launch_push_swarm(env, in, out, p, xs, n_inst, off, inst_E, mom, env.push_shape);
  if (rc) return rc;
  hipLaunchKernelGGL(k_swarm_offsets, dim3(1), dim3(1024), 0, env.stream, n_inst, off,
                     (const double*)mom, acc, accumulate, (int64_t)env.capacity);
  return check_launch("k_swarm_offsets");
}

int launch_swarm_control(const LaunchEnv& env, const bz_state_t* in, const bz_state_t* out,
                         const bz_params_t* p, int n_inst, int64_t* off, const int32_t* targets,
                         int32_t* counts, int32_t* offsets, const bz_resample_policy_t* pol) {
  uint32_t* skeys = (uint32_t*)sort_scratch_keys(env);
  uint32_t* perm = (uint32_t*)sort_scratch_perm(env);
  hipLaunchKernelGGL(k_swarm_prep, dim3(n_inst), dim3(256), 0, env.stream, n_inst,
                     (const int64_t*)off, counts, offsets, skeys, perm);
  int rc = check_launch("k_swarm_prep");
  if (rc) return rc;
  int32_t* plan = nullptr;
  rc = launch_resample_grouped(env, in, out, p, counts, offsets, perm, skeys, targets, n_inst,
                               pol, &plan);
  if (rc) return rc;
  hipLaunchKernelGGL(k_swarm_offsets_from_plan, dim3((n_inst + 256) / 256), dim3(256), 0,
                     env.stream, n_inst, (const int32_t*)plan, off, (const Clock*)env.clock);
  return check_launch("k_swarm_offsets_from_plan");
}

int launch_swarm_eedf(const LaunchEnv& env, const bz_state_t* s, int n_inst, const int64_t* off,
                      int n_bins, double e_lo, double e_hi, double kev, double* out) {
  const double lg_lo = log10(e_lo), lg_hi = log10(e_hi);
  const double inv = (double)n_bins / (lg_hi - lg_lo);
  const size_t smem = (size_t)n_bins * sizeof(double);
  if (env.precision == BZ_F64)
    hipLaunchKernelGGL(k_swarm_eedf<double>, dim3(n_inst), dim3(256), smem, env.stream,
                       (const double*)s->vx, (const double*)s->vy, (const double*)s->vz,
                       (const double*)s->w, off, n_bins, lg_lo, inv, kev, out);
  else
    hipLaunchKernelGGL(k_swarm_eedf<float>, dim3(n_inst), dim3(256), smem, env.stream,
                       (const float*)s->vx, (const float*)s->vy, (const float*)s->vz,
                       (const float*)s->w, off, n_bins, lg_lo, inv, kev, out);
  return check_launch("k_swarm_eedf");
}

}  // namespace bz
