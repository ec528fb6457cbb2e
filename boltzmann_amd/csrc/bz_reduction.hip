// The reference's native reduction interface, deterministic, for gfx950.
//
// reduction(r, p, M, RM, dw, dv, dx, ptcl) (kernels/reduction/cuda_reduction/
// __init__.py:11-15 -> cuda_red.calls, kernels-new.cu:76-132): for particle
// i < ptcl the fine bin b = int(dx[i] * M), the coarse row k = b / bnf with
// bnf = ceil(M/RM) (the redcheck block, kernels-new.cu:86-90), and
// dw[r*k + j] += dv[r*i + j] for the r value columns; rows RM..M-1 of dw are
// zeroed.  The reference adds with FP64 atomics (reda, :56-73) and folds the
// fine rows in place with a cross-block race (redcheck, :8-50), so its sums
// change from run to run.  Here every column is summed in signed fixed point:
// a pre-pass takes the largest |value| of each column, each value is scaled
// by 2^(48 - its binary exponent) and rounded to a 64-bit integer, and a
// block adds them into 64-bit LDS words with 32-bit carry words (a 96-bit
// two's-complement sum per row); the blocks' rows are then added as 128-bit
// integers.  Integer sums do not depend on the order or on how the particles
// were split over blocks, so the result is the same on every run and for any
// grid.  Resolution 2^-48 of the column's largest |value|; range 2^79 of
// those units per row and block.
#include "bz_launch.h"

namespace bz {

constexpr int kRedBlock = 256;

// Per column j < r: the bits of max |dv[r*i + j]| into colmax[j] (the bits of
// a non-negative double order like the double).
__global__ __launch_bounds__(kRedBlock) void k_red_absmax(int r, const double* __restrict__ dv,
                                                          int64_t ptcl,
                                                          unsigned long long* colmax) {
  for (int j = 0; j < r; ++j) {
    double m = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kRedBlock + threadIdx.x; i < ptcl;
         i += (int64_t)gridDim.x * kRedBlock)
      m = fmax(m, fabs(dv[(int64_t)r * i + j]));
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
    if ((threadIdx.x & 63) == 0 && m > 0.0)
      atomicMax(&colmax[j], (unsigned long long)__double_as_longlong(m));
  }
}

__device__ __forceinline__ double red_scale(unsigned long long bits) {
  if (!bits) return 1.0;
  int e;
  frexp(__longlong_as_double((long long)bits), &e);
  return ldexp(1.0, 48 - e);
}

// 128-bit two's complement (signed hi, unsigned lo) to double
__device__ __forceinline__ double red_value(unsigned long long lo, long long hi, double inv) {
  if (hi >= 0) return ((double)hi * 18446744073709551616.0 + (double)lo) * inv;
  // magnitude: (-hi - 1) * 2^64 + (2^64 - lo), or -hi * 2^64 when lo = 0
  const unsigned long long mlo = ~lo + 1ull;
  const double mhi = (double)(-hi - (lo ? 1 : 0));
  return -(mhi * 18446744073709551616.0 + (double)mlo) * inv;
}

// Block rows of one column at a time: LDS words lo[RM] (u64) and carries
// hi[RM] (i32), stored as the integer pair (lo, hi) to
// rows[block][r*k + j][2] (so the sum over blocks is exact).
__global__ __launch_bounds__(kRedBlock) void k_red_fixed(int r, int M, int RM, int bnf,
                                                         const double* __restrict__ dv,
                                                         const double* __restrict__ dx,
                                                         int64_t ptcl,
                                                         const unsigned long long* colmax,
                                                         unsigned long long* rows) {
  extern __shared__ unsigned long long s_lo[];  // [RM], then int32 hi[RM]
  int32_t* s_hi = (int32_t*)(s_lo + RM);
  unsigned long long* part = rows + (size_t)blockIdx.x * RM * r * 2;
  for (int j = 0; j < r; ++j) {
    for (int k = threadIdx.x; k < RM; k += kRedBlock) {
      s_lo[k] = 0ull;
      s_hi[k] = 0;
    }
    __syncthreads();
    const double scale = red_scale(colmax[j]);
    for (int64_t i = (int64_t)blockIdx.x * kRedBlock + threadIdx.x; i < ptcl;
         i += (int64_t)gridDim.x * kRedBlock) {
      const double v = dv[(int64_t)r * i + j];
      if (v == 0.0) continue;
      const int b = (int)(dx[i] * (double)M);
      int k = b / bnf;
      k = k < 0 ? 0 : (k >= RM ? RM - 1 : k);
      const long long q = (long long)rint(v * scale);  // |q| <= 2^48
      const unsigned long long uq = (unsigned long long)q;
      const unsigned long long old = atomicAdd(&s_lo[k], uq);
      const int32_t dh = (old + uq < old ? 1 : 0) - (q < 0 ? 1 : 0);
      if (dh) atomicAdd(&s_hi[k], dh);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < RM; k += kRedBlock) {
      part[2 * ((size_t)r * k + j)] = s_lo[k];
      part[2 * ((size_t)r * k + j) + 1] = (unsigned long long)(long long)s_hi[k];
    }
    __syncthreads();  // the LDS is cleared for the next column
  }
}

// dw[q] = the exact integer sum over blocks of rows[b][q], scaled back, for
// q < RM*r (column q % r); rows RM..M-1 zeroed.
__global__ __launch_bounds__(kRedBlock) void k_red_sum(const unsigned long long* __restrict__ rows,
                                                       int nblocks, int r, int64_t m,
                                                       int64_t mtot,
                                                       const unsigned long long* colmax,
                                                       double* dw) {
  for (int64_t q = (int64_t)blockIdx.x * kRedBlock + threadIdx.x; q < mtot;
       q += (int64_t)gridDim.x * kRedBlock) {
    double s = 0.0;
    if (q < m) {
      unsigned long long lo = 0ull;
      long long hi = 0;
      for (int b = 0; b < nblocks; ++b) {
        const unsigned long long bl = rows[2 * ((size_t)b * m + q)];
        const long long bh = (long long)rows[2 * ((size_t)b * m + q) + 1];
        const unsigned long long nl = lo + bl;
        hi += bh + (nl < lo ? 1 : 0);
        lo = nl;
      }
      s = red_value(lo, hi, 1.0 / red_scale(colmax[q % r]));
    }
    dw[q] = s;
  }
}

__global__ void k_red_clear(unsigned long long* p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0ull;
}

// The reduction on caller scratch of `scratch_doubles` doubles: the column
// bounds (r words) then one row of RM*r doubles per block.
int launch_reduction_scratch(hipStream_t stream, int num_cu, int r, int M, int RM, double* dw,
                             const double* dv, const double* dx, int64_t ptcl, double* scratch,
                             int64_t scratch_doubles) {
  if (r < 1 || r > 64 || M < 1 || RM < 1 || RM > M)
    return fail("reduction: need 1 <= r <= 64, M >= 1, 1 <= RM <= M");
  const int bnf = (M + RM - 1) / RM;  // ceil(M/RM), the redcheck block size
  const int64_t m = (int64_t)RM * r;
  const size_t smem = (size_t)RM * 12;
  if (smem > 160 * 1024) return fail("reduction: %d rows exceed the LDS of a block", RM);
  if (smem > 64 * 1024 &&
      hipFuncSetAttribute((const void*)k_red_fixed, hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)smem) != hipSuccess)
    return fail("hipFuncSetAttribute(reduction LDS) failed");
  unsigned long long* colmax = (unsigned long long*)scratch;
  unsigned long long* rows = (unsigned long long*)scratch + 64;
  int64_t grid = (int64_t)num_cu * 2;
  const int64_t fit = (scratch_doubles - 64) / (2 * m);
  if (grid > fit) grid = fit;
  if (grid < 1)
    return fail("reduction: %lld rows x %d columns need %lld doubles of scratch, %lld available",
                (long long)RM, r, (long long)(2 * m + 64), (long long)scratch_doubles);
  // the column bounds start at zero (a fill kernel: no memset node when captured)
  hipLaunchKernelGGL(k_red_clear, dim3(1), dim3(64), 0, stream, colmax, r);
  if (ptcl > 0)
    hipLaunchKernelGGL(k_red_absmax, dim3((unsigned)grid), dim3(kRedBlock), 0, stream, r, dv, ptcl,
                       colmax);
  hipLaunchKernelGGL(k_red_fixed, dim3((unsigned)grid), dim3(kRedBlock), smem, stream, r, M, RM,
                     bnf, dv, dx, ptcl, (const unsigned long long*)colmax, rows);
  const int64_t mtot = (int64_t)M * r;
  const int64_t sg = (mtot + kRedBlock - 1) / kRedBlock;
  hipLaunchKernelGGL(k_red_sum, dim3((unsigned)(sg < 4096 ? sg : 4096)), dim3(kRedBlock), 0,
                     stream, (const unsigned long long*)rows, (int)grid, r, m, mtot,
                     (const unsigned long long*)colmax, dw);
  return check_launch("reduction");
}

int launch_reduction(const LaunchEnv& env, int r, int M, int RM, double* dw, const double* dv,
                     const double* dx, int64_t ptcl) {
  return launch_reduction_scratch(env.stream, env.num_cu, r, M, RM, dw, dv, dx, ptcl,
                                  env.partials, env.partials_doubles);
}

}  // namespace bz

// ---------------------------------------------------------------------------
// The reduction for callers that cannot allocate (code under graph capture,
// serving loops, harnesses that forbid hipMalloc): no context, and its
// scratch (the column bounds and the blocks' integer rows) in a file-scope
// device array of kRedStaticDoubles doubles, 8 MiB that every GPU the library
// runs on holds until the process ends.  Its address on each device is taken
// once, with that device current, on the first call there.  Every call on a
// GPU shares that GPU's array, so calls that overlap on one GPU race; the
// column bounds are cleared by a kernel at each call.
// ---------------------------------------------------------------------------
#include <mutex>

namespace bz {

constexpr int64_t kRedStaticDoubles = (int64_t)1 << 20;
static __device__ __align__(256) double scratch[kRedStaticDoubles];

constexpr int kStaticMaxDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[kStaticMaxDevices];
static int g_scratch_cu[kStaticMaxDevices];

}  // namespace bz

extern "C" int bz_reduction_static(int32_t r, int32_t M, int32_t RM, double* dw,
                                   const double* dv, const double* dx, int64_t ptcl,
                                   void* hip_stream) {
  using namespace bz;
  if (r < 1 || r > 64 || M < 1 || RM < 1 || RM > M)
    return fail("bz_reduction_static: need 1 <= r <= 64, M >= 1, 1 <= RM <= M");
  if (ptcl < 0) return fail("bz_reduction_static: negative particle count");
  if (!dw || (ptcl > 0 && (!dv || !dx))) return fail("bz_reduction_static: null pointer");
  const int64_t need = 64 + 2 * (int64_t)RM * r;
  if (need > kRedStaticDoubles)
    return fail("bz_reduction_static: %d rows x %d columns need %lld doubles of scratch, "
                "more than the %lld its static array holds",
                RM, r, (long long)need, (long long)kRedStaticDoubles);
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return fail("bz_reduction_static: no current device");
  if (dev < 0 || dev >= kStaticMaxDevices)
    return fail("bz_reduction_static: device %d beyond %d", dev, kStaticMaxDevices);
  double* ptr = nullptr;
  int ncu = 0;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
      // the array's address on this device (the host-side name is only a
      // stand-in, never passed to a kernel)
      void* p = nullptr;
      const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !p)
        return fail("bz_reduction_static: hipGetSymbolAddress: %s", hipGetErrorString(e));
      int n = 0;
      const hipError_t e2 = hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev);
      if (e2 != hipSuccess || n < 1)
        return fail("bz_reduction_static: compute unit count: %s", hipGetErrorString(e2));
      g_scratch_ptr[dev] = (double*)p;
      g_scratch_cu[dev] = n;
    }
    ptr = g_scratch_ptr[dev];
    ncu = g_scratch_cu[dev];
  }
  return launch_reduction_scratch((hipStream_t)hip_stream, ncu, r, M, RM, dw, dv, dx, ptcl, ptr,
                                  kRedStaticDoubles);
}
