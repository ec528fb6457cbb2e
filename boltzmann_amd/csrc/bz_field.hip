// FP64 grid step on device: host moment update, explicit drift-diffusion of
// the heavies and the tridiagonal Poisson solve, in one single-block kernel
// so a whole kinetic step needs no host round trip.
//
// Reference: boltzmann.py:643-647 (ni += ni_src, Te, g2), heavies_kernel_1D.py:5-64,
// EF_kernel_1D.py:8-44 with the banded Cholesky factor of boltzmann.py:744-755.
#include "bz_launch.h"

namespace bz {

constexpr int kFieldBlock = 1024;

__global__ __launch_bounds__(kFieldBlock) void k_field(bz_field_params_t fp,
                                                       const double* __restrict__ mom,
                                                       double* ni, double* ns, double* Te,
                                                       double* g2, double* V, double* E,
                                                       void* E_t, int prec,
                                                       const double* __restrict__ chol,
                                                       double* Ji, double* Js, Clock* clk,
                                                       int advance) {
  extern __shared__ double s_f[];  // rhs[N-1], tempy[N-1]
  const int N = fp.n;
  const int nn = N + 1;
  const int tid = threadIdx.x;
  const double* ne = mom;
  const double* nsrc = mom + nn;
  const double* nrg = mom + 2 * nn;
  double* rhs = s_f;
  double* ty = s_f + (N - 1);
  const double* diag = chol;
  const double* low = chol + (N - 1);

  // host moment update (boltzmann.py:643-647)
  for (int i = tid; i < nn; i += kFieldBlock) {
    ni[i] = ni[i] + nsrc[i];
    double te = ((2.0 / 3.0) * nrg[i]) / ne[i];
    if (te != te) te = 0.0;  // nan_to_num
    Te[i] = te;
    g2[i] = g2[i] + nsrc[i];
  }
  __syncthreads();
  double t = clk->t;
  if (advance) t = t + fp.dt;
  __syncthreads();

  if (fp.fluid_on) {
    const double dx = fp.dx, dt = fp.dt, mu = fp.mu_i, Di = fp.D_i, Ds = fp.D_s;
    const double e0 = E[0] > 0.0 ? E[0] : 0.0;
    const double eN = E[N - 1] < 0.0 ? E[N - 1] : 0.0;
    // boundary values (heavies_kernel_1D.py:26-35)
    if (tid == 0) {
      ni[0] = ni[1] * ((Di / dx - 0.5 * mu * e0) / (Di / dx + 0.5 * mu * e0));
      ni[N] = ni[N - 1] * ((Di / dx + 0.5 * mu * eN) / (Di / dx - 0.5 * mu * eN));
      ns[0] = 0.0;
      ns[N] = 0.0;
    }
    __syncthreads();
    for (int i = tid; i < N; i += kFieldBlock) {
      Ji[i] = mu * (0.5 * (ni[i] + ni[i + 1])) * E[i] - Di * (ni[i + 1] - ni[i]) / dx;
      Js[i] = -Ds * (ns[i + 1] - ns[i]) / dx;
    }
    __syncthreads();
    for (int i = 1 + tid; i < N; i += kFieldBlock) {
      ni[i] = ni[i] - dt * (Ji[i] - Ji[i - 1]) / dx;
      ns[i] = ns[i] - dt * (Js[i] - Js[i - 1]) / dx;
    }
    __syncthreads();
    if (tid == 0) {
      ni[0] = ni[1] * ((Di / dx - 0.5 * mu * e0) / (Di / dx + 0.5 * mu * e0));
      ni[N] = ni[N - 1] * ((Di / dx + 0.5 * mu * eN) / (Di / dx - 0.5 * mu * eN));
      ns[0] = 0.0;
      ns[N] = 0.0;
    }
    __syncthreads();
  }

  if (fp.poisson_on) {
    const double dx = fp.dx;
    const double vl = fp.v0 * sin(2.0 * 3.141592653589793 * fp.freq * t);
    for (int i = tid; i < N - 1; i += kFieldBlock) rhs[i] = (1.0 / fp.eps) * (ni[i + 1] - ne[i + 1]);
    __syncthreads();
    if (tid == 0) {
      rhs[0] = rhs[0] + (1.0 / (dx * dx)) * fp.v0 * sin(2.0 * 3.141592653589793 * fp.freq * t);
      ty[0] = rhs[0] / diag[0];
      for (int i = 1; i < N - 1; ++i) ty[i] = (1.0 / diag[i]) * (rhs[i] - low[i] * ty[i - 1]);
      V[N - 2] = ty[N - 2] / diag[N - 2];
      for (int i = 1; i < N - 1; ++i)
        V[N - 2 - i] = (1.0 / diag[N - 2 - i]) * (ty[N - 2 - i] - low[N - 1 - i] * V[N - 1 - i]);
    }
    __syncthreads();
    for (int i = tid; i < N; i += kFieldBlock) {
      double e;
      if (i == 0) e = -(V[0] - vl) / dx;
      else if (i == N - 1) e = -(0.0 - V[N - 2]) / dx;
      else e = -(V[i] - V[i - 1]) / dx;
      E[i] = e;
    }
    __syncthreads();
  }
  for (int i = tid; i < N; i += kFieldBlock) {
    if (prec == BZ_F64) ((double*)E_t)[i] = E[i];
    else ((float*)E_t)[i] = (float)E[i];
  }
  if (tid == 0 && advance) {
    clk->t = t;
    clk->step = clk->step + 1;
  }
}

int launch_field(const LaunchEnv& env, const bz_field_params_t* fp, const double* moments,
                 double* ni, double* ns, double* Te, double* g2, double* V, double* E,
                 void* E_t, const double* chol, int advance) {
  const int N = fp->n;
  const size_t smem = (size_t)2 * (N - 1) * sizeof(double);
  if (smem > 160 * 1024) return fail("field solve limited to N <= 10240 cells");
  if (smem > 64 * 1024) {
    if (hipFuncSetAttribute((const void*)k_field, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)smem) != hipSuccess)
      return fail("hipFuncSetAttribute(field LDS) failed");
  }
  // fluxes live in the partial buffer (free between push launches)
  if (env.partials_doubles < 2ll * (N + 1)) return fail("workspace too small for fluxes");
  double* Ji = env.partials;
  double* Js = env.partials + (N + 1);
  hipLaunchKernelGGL(k_field, dim3(1), dim3(kFieldBlock), smem, env.stream, *fp, moments, ni, ns,
                     Te, g2, V, E, E_t, env.precision, chol, Ji, Js, env.clock, advance);
  return check_launch("k_field");
}

}  // namespace bz
