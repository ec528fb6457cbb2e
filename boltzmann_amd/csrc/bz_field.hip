// FP64 grid step on device: host moment update, explicit drift-diffusion of
// the heavies and the tridiagonal Poisson solve, in one single-block kernel
// so a whole kinetic step needs no host round trip.
//
// Reference: boltzmann.py:643-647 (ni += ni_src, Te, g2), heavies_kernel_1D.py:5-64,
// EF_kernel_1D.py:8-44 with the banded Cholesky factor of boltzmann.py:744-755.
#include "bz_launch.h"

namespace bz {


// Bernoulli function B(x) = x / (e^x - 1), B(0) = 1.
__device__ __forceinline__ double bernoulli(double x) {
  if (fabs(x) < 1e-4) return 1.0 - 0.5 * x + x * x * (1.0 / 12.0);
  return x / expm1(x);
}

// Ion / metastable transport with Scharfetter-Gummel fluxes
//   G_{i+1/2} = a_i n_i - b_i n_{i+1},  a_i = (D/dx) B(-Pe_i),  b_i = (D/dx) B(Pe_i),
//   Pe_i = mu E_i dx / D,
// explicit in time, sub-cycled so that every sub-step keeps the update a
// convex combination (h (a_i + b_{i-1}) / dx <= 1): positive and stable for
// any field.  Wall nodes follow the reference's rule (heavies_kernel_1D.py:26-35)
// restated for these fluxes: the wall face carries no inflow, i.e. its flux
// is zero when E pushes ions into the domain (n_0 = n_1 exp(-Pe_0)) and pure
// drift outflow otherwise (n_0 = n_1).  Metastables diffuse with the
// reference's central flux and zero wall density, on the same sub-steps.
// Uses LDS: a, b, Gi, Gs [N each] (the Poisson scratch, not yet in use).
template <int BLOCK>
__device__ void sg_heavies(const bz_field_params_t& fp, int N, const double* E, double* ni,
                           double* ns, double* lds) {
  __shared__ double s_rate;
  __shared__ int s_m;
  const int tid = threadIdx.x;
  const double dx = fp.dx, mu = fp.mu_i, Di = fp.D_i, Ds = fp.D_s;
  double* a = lds;
  double* b = lds + N;
  double* Gi = lds + 2 * N;
  double* Gs = lds + 3 * N;
  const double dd = Di / dx;
  for (int i = tid; i < N; i += BLOCK) {
    const double pe = mu * E[i] * dx / Di;
    a[i] = dd * bernoulli(-pe);
    b[i] = dd * bernoulli(pe);
  }
  if (tid == 0) s_rate = 0.0;
  __syncthreads();
  // largest diagonal rate over interior nodes (max is order independent)
  double r = 0.0;
  for (int i = 1 + tid; i < N; i += BLOCK) r = fmax(r, (a[i] + b[i - 1]) / dx);
  r = fmax(r, 2.0 * Ds / (dx * dx));
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) r = fmax(r, __shfl_xor(r, off, 64));
  if ((tid & 63) == 0) atomicMax((unsigned long long*)&s_rate, (unsigned long long)__double_as_longlong(r));
  __syncthreads();
  if (tid == 0) {
    // non-negative doubles order like their bit patterns
    const double rate = s_rate;
    int m = (int)ceil(fp.dt * rate);
    if (m < 1) m = 1;
    if (fp.max_substeps > 0 && m > fp.max_substeps) m = fp.max_substeps;
    s_m = m;
  }
  __syncthreads();
  const int m = s_m;
  const double h = fp.dt / m;
  const double pe0 = mu * fmax(E[0], 0.0) * dx / Di;
  const double peN = mu * fmin(E[N - 1], 0.0) * dx / Di;
  const double f0 = exp(-pe0), fN = exp(peN);
  for (int sub = 0; sub < m; ++sub) {
    if (tid == 0) {
      ni[0] = ni[1] * f0;
      ni[N] = ni[N - 1] * fN;
      ns[0] = 0.0;
      ns[N] = 0.0;
    }
    __syncthreads();
    for (int i = tid; i < N; i += BLOCK) {
      Gi[i] = a[i] * ni[i] - b[i] * ni[i + 1];
      Gs[i] = -Ds * (ns[i + 1] - ns[i]) / dx;
    }
    __syncthreads();
    for (int i = 1 + tid; i < N; i += BLOCK) {
      ni[i] = ni[i] - h * (Gi[i] - Gi[i - 1]) / dx;
      ns[i] = ns[i] - h * (Gs[i] - Gs[i - 1]) / dx;
    }
    __syncthreads();
  }
  if (tid == 0) {
    ni[0] = ni[1] * f0;
    ni[N] = ni[N - 1] * fN;
    ns[0] = 0.0;
    ns[N] = 0.0;
  }
  __syncthreads();
}

// Exclusive scan of affine maps y -> A y + B over the block (thread order):
// returns the value entering thread tid's chunk, i.e. the composition of the
// maps of threads 0..tid-1 applied to 0 (0 for thread 0).  Waves scan with
// shuffles; the wave totals are composed in wave order.
template <int BLOCK>
__device__ double affine_scan_in(double A, double B, double* s_wa, double* s_wb) {
  constexpr int NW = BLOCK / 64;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const double pa = __shfl_up(A, off, 64), pb = __shfl_up(B, off, 64);
    if (lane >= off) {
      B = A * pb + B;
      A = A * pa;
    }
  }
  if (lane == 63) {
    s_wa[wid] = A;
    s_wb[wid] = B;
  }
  __syncthreads();
  double y = 0.0;  // value entering this wave
  for (int w = 0; w < wid && w < NW; ++w) y = s_wa[w] * y + s_wb[w];
  const double ia = __shfl_up(A, 1, 64), ib = __shfl_up(B, 1, 64);
  if (lane > 0) y = ia * y + ib;
  __syncthreads();  // s_wa, s_wb are reused by the next scan
  return y;
}

// mom is not __restrict__: a closure sub-step updates its ne through ne_rw.
// STAGE: the grid is small enough for ni and ns to live in LDS through the
// heavy-species sub-steps and the Poisson right-hand side (one global read
// and one write instead of round trips on every sub-step).
template <int BLOCK, bool STAGE>
__global__ __launch_bounds__(BLOCK) void k_field(bz_field_params_t fp,
                                                       const double* mom,
                                                       double* ni, double* ns, double* Te,
                                                       double* g2, double* V, double* E,
                                                       void* E_t, int prec,
                                                       const double* __restrict__ chol,
                                                       double* Ji, double* Js, Clock* clk,
                                                       int advance, const double* src,
                                                       double* ne_rw) {
  // a, b, Gi, Gs [N each] of the heavy species, then rhs, tempy, 1/diag,
  // lower [N-1 each] of the Poisson solve (STAGE: ni, ns [N+1 each] after)
  extern __shared__ double s_f[];
  __shared__ double s_wa[BLOCK / 64], s_wb[BLOCK / 64];
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
  if (!fp.skip_moments) {
    for (int i = tid; i < nn; i += BLOCK) {
      ni[i] = ni[i] + nsrc[i];
      double te = ((2.0 / 3.0) * nrg[i]) / ne[i];
      if (te != te) te = 0.0;  // nan_to_num
      Te[i] = te;
      g2[i] = g2[i] + nsrc[i];
    }
  }
  if (src) {
    // closure sub-step: ionization from the learned rate makes ion-electron pairs
    for (int i = tid; i < nn; i += BLOCK) {
      const double q = src[i];
      ni[i] = ni[i] + q;
      g2[i] = g2[i] + q;
      ne_rw[i] = ne_rw[i] + q;
    }
  }
  __syncthreads();
  double t = clk->t;
  if (advance) t = t + fp.dt;
  __syncthreads();

  // ion density read by the Poisson right-hand side
  const double* ni_p = ni;
  if (fp.fluid_on && fp.ion_scheme == 1) {
    if (STAGE) {
      // after the heavy-species scratch a, b, Gi, Gs [N each]
      double* s_ni = s_f + 4 * N;
      double* s_ns = s_ni + nn;
      for (int i = tid; i < nn; i += BLOCK) {
        s_ni[i] = ni[i];
        s_ns[i] = ns[i];
      }
      __syncthreads();
      sg_heavies<BLOCK>(fp, N, E, s_ni, s_ns, s_f);
      for (int i = tid; i < nn; i += BLOCK) {
        ni[i] = s_ni[i];
        ns[i] = s_ns[i];
      }
      ni_p = s_ni;
    } else {
      sg_heavies<BLOCK>(fp, N, E, ni, ns, s_f);
    }
  } else if (fp.fluid_on) {
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
    for (int i = tid; i < N; i += BLOCK) {
      Ji[i] = mu * (0.5 * (ni[i] + ni[i + 1])) * E[i] - Di * (ni[i + 1] - ni[i]) / dx;
      Js[i] = -Ds * (ns[i + 1] - ns[i]) / dx;
    }
    __syncthreads();
    for (int i = 1 + tid; i < N; i += BLOCK) {
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
    const int M = N - 1;
    double* inv = s_f + 2 * M;  // 1/diag
    double* lw = s_f + 3 * M;   // lower band
    for (int i = tid; i < M; i += BLOCK) {
      rhs[i] = (1.0 / fp.eps) * (ni_p[i + 1] - ne[i + 1]);
      inv[i] = 1.0 / diag[i];
      lw[i] = low[i];
    }
    __syncthreads();
    if (tid == 0)
      rhs[0] = rhs[0] + (1.0 / (dx * dx)) * fp.v0 * sin(2.0 * 3.141592653589793 * fp.freq * t);
    __syncthreads();
    // Both triangular solves are first-order linear recurrences
    // y_i = A_i y_{i-1} + B_i; each thread composes the maps of its chunk,
    // a block scan composes across chunks, then each chunk is re-run from
    // its incoming value with the reference's per-row expression.
    const int chunk = (M + BLOCK - 1) / BLOCK;
    const int lo = tid * chunk;
    const int hi = lo + chunk < M ? lo + chunk : M;
    // forward: U^T y = rhs (EF_kernel_1D.py:31-33)
    {
      double A = 1.0, B = 0.0;
      for (int i = lo; i < hi; ++i) {
        double ai, bi;
        if (i == 0) {
          ai = 0.0;
          bi = rhs[0] / diag[0];
        } else {
          ai = -(inv[i] * lw[i]);
          bi = inv[i] * rhs[i];
        }
        B = ai * B + bi;
        A = ai * A;
      }
      double y = affine_scan_in<BLOCK>(A, B, s_wa, s_wb);
      for (int i = lo; i < hi; ++i) {
        y = (i == 0) ? rhs[0] / diag[0] : inv[i] * (rhs[i] - lw[i] * y);
        ty[i] = y;
      }
      __syncthreads();
    }
    // backward: U V = y (EF_kernel_1D.py:35-37), index r = M-1-i
    {
      double A = 1.0, B = 0.0;
      for (int r = lo; r < hi; ++r) {
        const int i = M - 1 - r;
        double ai, bi;
        if (r == 0) {
          ai = 0.0;
          bi = ty[i] / diag[i];
        } else {
          ai = -(inv[i] * lw[i + 1]);
          bi = inv[i] * ty[i];
        }
        B = ai * B + bi;
        A = ai * A;
      }
      double v = affine_scan_in<BLOCK>(A, B, s_wa, s_wb);
      for (int r = lo; r < hi; ++r) {
        const int i = M - 1 - r;
        v = (r == 0) ? ty[i] / diag[i] : inv[i] * (ty[i] - lw[i + 1] * v);
        V[i] = v;
      }
      __syncthreads();
    }
    for (int i = tid; i < N; i += BLOCK) {
      double e;
      if (i == 0) e = -(V[0] - vl) / dx;
      else if (i == N - 1) e = -(0.0 - V[N - 2]) / dx;
      else e = -(V[i] - V[i - 1]) / dx;
      E[i] = e;
    }
    __syncthreads();
  }
  for (int i = tid; i < N; i += BLOCK) {
    if (prec == BZ_F64) ((double*)E_t)[i] = E[i];
    else ((float*)E_t)[i] = (float)E[i];
  }
  if (tid == 0 && advance) {
    clk->t = t;
    clk->step = clk->step + 1;
  }
}

template <int BLOCK, bool STAGE>
static int launch_field_t(const LaunchEnv& env, const bz_field_params_t* fp,
                          const double* moments, double* ni, double* ns, double* Te, double* g2,
                          double* V, double* E, void* E_t, const double* chol, int advance,
                          const double* src) {
  const int N = fp->n;
  // heavy-species scratch a, b, Gi, Gs (N each), reused by the Poisson
  // solve for rhs, y, 1/diag, lower band (N-1 each) [+ staged ni, ns]
  size_t smem = (size_t)4 * N * sizeof(double);
  if (STAGE) smem += (size_t)2 * (N + 1) * sizeof(double);
  // 1 KiB of the 160 KiB stays for the kernel's own LDS variables
  if (smem > 159 * 1024) return fail("field solve limited to N <= 5088 cells (LDS)");
  if (smem > 64 * 1024) {
    if (hipFuncSetAttribute((const void*)k_field<BLOCK, STAGE>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem) != hipSuccess)
      return fail("hipFuncSetAttribute(field LDS) failed");
  }
  // fluxes live in the partial buffer (free between push launches)
  if (env.partials_doubles < 2ll * (N + 1)) return fail("workspace too small for fluxes");
  double* Ji = env.partials;
  double* Js = env.partials + (N + 1);
  hipLaunchKernelGGL((k_field<BLOCK, STAGE>), dim3(1), dim3(BLOCK), smem, env.stream, *fp,
                     moments, ni, ns, Te, g2, V, E, E_t, env.precision, chol, Ji, Js, env.clock,
                     advance, src, const_cast<double*>(moments));
  return check_launch("k_field");
}

// One block: 256 threads while the grid is small (fewer threads in every
// barrier and a shorter scan; ni and ns staged in LDS), 1024 for the large
// grids.
int launch_field(const LaunchEnv& env, const bz_field_params_t* fp, const double* moments,
                 double* ni, double* ns, double* Te, double* g2, double* V, double* E,
                 void* E_t, const double* chol, int advance, const double* src) {
  if (fp->n <= 1024)
    return launch_field_t<256, true>(env, fp, moments, ni, ns, Te, g2, V, E, E_t, chol, advance,
                                     src);
  return launch_field_t<1024, false>(env, fp, moments, ni, ns, Te, g2, V, E, E_t, chol, advance,
                                     src);
}

}  // namespace bz

// ---------------------------------------------------------------------------
// Kinetic update of a closure-driven run: the sub-steps since the last update
// added the learned ionization acc to ni; the kinetic step counted nsrc over
// the same interval.  Replace one by the other (ni += nsrc - acc, g2 alike),
// refresh Te = 2/3 nrg/ne, and add the tally's sums:
//   tally[0] += sum nsrc, tally[1] += sum acc, tally[2] += |sum nsrc - sum acc|,
//   tally[3] += 1, tally[4] += sum ne, tally[5] += sum nrg.
// Sums run in a fixed order (deterministic).  acc is cleared.
namespace bz {

__global__ __launch_bounds__(1024) void k_closure_commit(int nn, const double* __restrict__ mom,
                                                         double* __restrict__ acc,
                                                         double* __restrict__ ni,
                                                         double* __restrict__ g2,
                                                         double* __restrict__ Te,
                                                         double* __restrict__ tally) {
  __shared__ double s_red[4][16];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const double* ne = mom;
  const double* nsrc = mom + nn;
  const double* nrg = mom + 2 * nn;
  double s_k = 0.0, s_a = 0.0, s_n = 0.0, s_e = 0.0;
  for (int i = tid; i < nn; i += 1024) {
    const double d = nsrc[i] - acc[i];
    ni[i] = ni[i] + d;
    g2[i] = g2[i] + d;
    double te = ((2.0 / 3.0) * nrg[i]) / ne[i];
    if (te != te) te = 0.0;
    Te[i] = te;
    s_k += nsrc[i];
    s_a += acc[i];
    s_n += ne[i];
    s_e += nrg[i];
    acc[i] = 0.0;
  }
  double v[4] = {s_k, s_a, s_n, s_e};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const double r = wave_sum(v[q]);
    if (lane == 0) s_red[q][wid] = r;
  }
  __syncthreads();
  if (tid == 0) {
    double t[4] = {0.0, 0.0, 0.0, 0.0};
    for (int q = 0; q < 4; ++q)
      for (int w = 0; w < 16; ++w) t[q] += s_red[q][w];
    tally[0] += t[0];
    tally[1] += t[1];
    tally[2] += fabs(t[0] - t[1]);
    tally[3] += 1.0;
    tally[4] += t[2];
    tally[5] += t[3];
  }
}

int launch_closure_commit(const LaunchEnv& env, int nn, const double* mom, double* acc, double* ni,
                          double* g2, double* Te, double* tally) {
  hipLaunchKernelGGL(k_closure_commit, dim3(1), dim3(1024), 0, env.stream, nn, mom, acc, ni, g2,
                     Te, tally);
  return check_launch("k_closure_commit");
}

}  // namespace bz
