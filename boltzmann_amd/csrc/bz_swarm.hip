// Batched 0-D swarms (many independent electron swarms at different E/N in
// one launch), their population control and diagnostics.
//
// The state of all instances is one SoA array.  A 0-D particle has no
// position, so x holds its instance number.  The fused push (DIM 2 of
// k_push_inplace) reads the instance from x, applies that instance's field,
// updates the velocities in place and sums per-instance moments; births are
// appended after the slots in use (in parent order) and inherit their
// parent's instance.  Population control sorts the slots by instance
// (stable), merges or splits each instance towards its target, and leaves
// instance k in [off[k], off[k+1]), the layout the EEDF diagnostic reads.
#include "bz_launch.h"

namespace bz {

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
  (void)off;  // instances are identified by x; offsets come from the population control
  return launch_push_swarm(env, in, out, p, xs, n_inst, inst_E, mom, accumulate ? acc : nullptr,
                           env.push_shape);
}

int launch_swarm_control(const LaunchEnv& env, const bz_state_t* in, const bz_state_t* out,
                         const bz_params_t* p, int n_inst, int64_t* off, const int32_t* targets,
                         int32_t* counts, int32_t* offsets, const bz_resample_policy_t* pol) {
  uint32_t* perm = (uint32_t*)sort_scratch_perm(env);
  const uint32_t* skeys = nullptr;
  int rc = launch_group_sort(env, in, n_inst, counts, offsets, perm, &skeys);
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
