// Learned closure: a 3 -> 64 -> 64 -> 64 -> 3 SiLU perceptron with BF16
// weights and FP32 accumulation.  Inference, the gradient of the batch loss
// and the Adam update are hand-written kernels.
//
// Parameter layout (flat, BZ_MLP_NPARAM entries, row-major W[out][in]):
//   W1[64][3] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[3][64] b4[3]
// The FP32 master copy is what Adam updates; every forward and backward pass
// reads the BF16 copy (rounded to nearest even), with activations and all
// sums in FP32.  The network is far too small for the matrix cores to pay:
// a batch of 64 samples is 0.5 MFLOP per layer, so each wave keeps its
// samples' activations in registers and LDS and streams the weights from
// LDS as broadcast reads.
#include "bz_launch.h"

namespace bz {
namespace mlp {

constexpr int kIn = BZ_MLP_IN;
constexpr int kH = BZ_MLP_HIDDEN;
constexpr int kOut = BZ_MLP_OUT;
constexpr int oW1 = 0;
constexpr int oB1 = oW1 + kH * kIn;
constexpr int oW2 = oB1 + kH;
constexpr int oB2 = oW2 + kH * kH;
constexpr int oW3 = oB2 + kH;
constexpr int oB3 = oW3 + kH * kH;
constexpr int oW4 = oB3 + kH;
constexpr int oB4 = oW4 + kOut * kH;
constexpr int kN = oB4 + kOut;
static_assert(kN == BZ_MLP_NPARAM, "parameter layout");
constexpr int kWlds = (kN + 3) & ~3;   // LDS floats holding the weights
constexpr int kStride = kWlds + 4;      // partial-gradient row: params, loss, pad
constexpr int kTile = 64;               // samples per gradient block (one wave)
constexpr int kRow = kH + 1;            // padded LDS activation row

// x_n = (x - in_shift) * in_scale;  y = y_n * out_scale + out_shift;
// targets are normalised with out_inv = 1 / out_scale
struct Norm {
  float in_shift[kIn], in_scale[kIn], out_shift[kOut], out_scale[kOut], out_inv[kOut];
  double in_lo[kIn], in_hi[kIn];
};

// a feature clamped to the trained domain (none where lo >= hi)
__device__ __forceinline__ double clamp_feature(const Norm& nm, int k, double v) {
  if (nm.in_lo[k] < nm.in_hi[k]) v = fmin(fmax(v, nm.in_lo[k]), nm.in_hi[k]);
  return v;
}

__device__ __forceinline__ float bf16_f32(uint16_t h) {
  return __uint_as_float(((uint32_t)h) << 16);
}

__device__ __forceinline__ uint16_t f32_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu)) return 0x7fc0;  // NaN
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

// Stage the BF16 parameters into LDS as FP32 (block-cooperative).
__device__ __forceinline__ void stage_weights(const uint16_t* __restrict__ w, float* Wl) {
  for (int i = threadIdx.x; i < kWlds; i += blockDim.x) Wl[i] = i < kN ? bf16_f32(w[i]) : 0.f;
  __syncthreads();
}

__device__ __forceinline__ void silu(float z, float& a, float& d) {
  const float s = 1.f / (1.f + expf(-z));
  a = z * s;
  d = s * (1.f + z * (1.f - s));  // d silu / dz
}

// One hidden layer: 64 (or 3) inputs held in registers, outputs written to
// the thread's own LDS row (activation, and silu' when training).  Only one
// 64-wide vector is live in registers at a time.
template <int K, bool SAVE_D>
__device__ __forceinline__ void layer(const float* __restrict__ Wl, int oW, int oB,
                                      const float (&in)[K], float* a_row, float* d_row) {
#pragma unroll 2
  for (int j = 0; j < kH; ++j) {
    float acc = Wl[oB + j];
#pragma unroll
    for (int k = 0; k < K; ++k) acc = fmaf(Wl[oW + j * K + k], in[k], acc);
    float a, d;
    silu(acc, a, d);
    a_row[j] = a;
    if (SAVE_D) d_row[j] = d;
  }
}

__device__ __forceinline__ void load_row(const float* row, float (&h)[kH]) {
#pragma unroll
  for (int k = 0; k < kH; ++k) h[k] = row[k];
}

__device__ __forceinline__ void output(const float* __restrict__ Wl, const float (&h)[kH],
                                       float (&y)[kOut]) {
#pragma unroll
  for (int o = 0; o < kOut; ++o) {
    float acc = Wl[oB4 + o];
#pragma unroll
    for (int k = 0; k < kH; ++k) acc = fmaf(Wl[oW4 + o * kH + k], h[k], acc);
    y[o] = acc;
  }
}

// Forward pass of one normalised sample through the thread's LDS rows.
template <bool SAVE_D>
__device__ __forceinline__ void forward(const float* __restrict__ Wl, const float (&x)[kIn],
                                        float* A1, float* A2, float* A3, float* P1, float* P2,
                                        float* P3, float (&y)[kOut]) {
  layer<kIn, SAVE_D>(Wl, oW1, oB1, x, A1, P1);
  float h[kH];
  load_row(A1, h);
  layer<kH, SAVE_D>(Wl, oW2, oB2, h, A2, P2);
  load_row(A2, h);
  layer<kH, SAVE_D>(Wl, oW3, oB3, h, A3, P3);
  load_row(A3, h);
  output(Wl, h, y);
}

// ---------------------------------------------------------------------------
constexpr int kFwdBlock = 128;

__global__ void __launch_bounds__(kFwdBlock) k_mlp_forward(Norm nm,
                                                           const uint16_t* __restrict__ w,
                                                           const float* __restrict__ x,
                                                           float* __restrict__ y, int64_t n) {
  __shared__ float Wl[kWlds];
  __shared__ float rows[kFwdBlock * kRow];
  stage_weights(w, Wl);
  float* row = rows + threadIdx.x * kRow;
  for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < n;
       s += (int64_t)gridDim.x * blockDim.x) {
    float in[kIn], out[kOut];
#pragma unroll
    for (int i = 0; i < kIn; ++i) in[i] = (x[s * kIn + i] - nm.in_shift[i]) * nm.in_scale[i];
    // inference keeps one row per thread: each layer reads it into
    // registers before overwriting it
    layer<kIn, false>(Wl, oW1, oB1, in, row, nullptr);
    float h[kH];
    load_row(row, h);
    layer<kH, false>(Wl, oW2, oB2, h, row, nullptr);
    load_row(row, h);
    layer<kH, false>(Wl, oW3, oB3, h, row, nullptr);
    load_row(row, h);
    output(Wl, h, out);
#pragma unroll
    for (int o = 0; o < kOut; ++o) y[s * kOut + o] = fmaf(out[o], nm.out_scale[o], nm.out_shift[o]);
  }
}

// delta_prev[k] = (sum_j W[j][k] delta[j]) * silu'_prev[k], written over the
// silu' row; delta is read from the thread's own row into registers first.
__device__ __forceinline__ void back_hidden(const float* __restrict__ Wl, int oW,
                                            const float* d_row, float* p_row) {
  float d[kH];
  load_row(d_row, d);
#pragma unroll 2
  for (int k = 0; k < kH; ++k) {
    float acc = 0.f;
#pragma unroll
    for (int j = 0; j < kH; ++j) acc = fmaf(Wl[oW + j * kH + k], d[j], acc);
    p_row[k] = acc * p_row[k];
  }
}

// dW[j][c] = sum_q delta[q][j] a_prev[q][c];  db[c] = sum_q delta[q][c]
__device__ __forceinline__ void weight_grad(const float* Dl, const float* Ap, int c, float* out,
                                            int oW, int oB) {
  float acc[kH];
#pragma unroll
  for (int j = 0; j < kH; ++j) acc[j] = 0.f;
  float bsum = 0.f;
  for (int q = 0; q < kTile; ++q) {
    const float a = Ap[q * kRow + c];
    bsum += Dl[q * kRow + c];
#pragma unroll
    for (int j = 0; j < kH; ++j) acc[j] = fmaf(Dl[q * kRow + j], a, acc[j]);
  }
#pragma unroll
  for (int j = 0; j < kH; ++j) out[oW + j * kH + c] = acc[j];
  out[oB + c] = bsum;
}

// Per-block gradient of L = (1/2n) sum_s |y_s - t_s|^2 (normalised space)
// over kTile samples, one thread per sample for the forward pass and the
// delta recursion, one thread per weight column for the weight gradients.
// Writes a row of kStride floats: the parameter gradient, then the loss.
__global__ void __launch_bounds__(kTile) k_mlp_grad(Norm nm, const uint16_t* __restrict__ w,
                                                    const float* __restrict__ x,
                                                    const float* __restrict__ t, int64_t n,
                                                    float inv_n, float* __restrict__ part) {
  extern __shared__ float lds[];
  float* Wl = lds;                    // kWlds
  float* X = Wl + kWlds;              // [kTile][4] normalised inputs
  float* D4 = X + kTile * 4;          // [kTile][4] output deltas
  float* A1 = D4 + kTile * 4;         // [kTile][kRow] activations
  float* A2 = A1 + kTile * kRow;
  float* A3 = A2 + kTile * kRow;
  float* P1 = A3 + kTile * kRow;      // silu'(z), then the layer's delta
  float* P2 = P1 + kTile * kRow;
  float* P3 = P2 + kTile * kRow;
  stage_weights(w, Wl);

  const int s = threadIdx.x;
  const int64_t g = (int64_t)blockIdx.x * kTile + s;
  const bool live = g < n;
  float loss = 0.f;
  {
    float xin[kIn], tt[kOut], y[kOut];
#pragma unroll
    for (int i = 0; i < kIn; ++i) {
      xin[i] = live ? (x[g * kIn + i] - nm.in_shift[i]) * nm.in_scale[i] : 0.f;
      X[s * 4 + i] = xin[i];
    }
#pragma unroll
    for (int o = 0; o < kOut; ++o)
      tt[o] = live ? (t[g * kOut + o] - nm.out_shift[o]) * nm.out_inv[o] : 0.f;
    const int r = s * kRow;
    forward<true>(Wl, xin, A1 + r, A2 + r, A3 + r, P1 + r, P2 + r, P3 + r, y);
    float d4[kOut];
#pragma unroll
    for (int o = 0; o < kOut; ++o) {
      const float e = live ? y[o] - tt[o] : 0.f;
      d4[o] = e * inv_n;
      loss = fmaf(0.5f * e, e * inv_n, loss);
      D4[s * 4 + o] = d4[o];
    }
    // delta recursion on the thread's own rows (no other thread reads them yet)
#pragma unroll 4
    for (int k = 0; k < kH; ++k) {
      float acc = 0.f;
#pragma unroll
      for (int o = 0; o < kOut; ++o) acc = fmaf(Wl[oW4 + o * kH + k], d4[o], acc);
      P3[r + k] = acc * P3[r + k];
    }
    back_hidden(Wl, oW3, P3 + r, P2 + r);
    back_hidden(Wl, oW2, P2 + r, P1 + r);
  }
  __syncthreads();

  float* out = part + (int64_t)blockIdx.x * kStride;
  const int c = threadIdx.x;  // weight column (input index) owned by this thread
  weight_grad(P3, A2, c, out, oW3, oB3);
  weight_grad(P2, A1, c, out, oW2, oB2);
  // layer 1: dW1[c][i] = sum_s delta1[s][c] * x[s][i]; db1[c]
  {
    float acc[kIn] = {0.f, 0.f, 0.f};
    float bsum = 0.f;
    for (int q = 0; q < kTile; ++q) {
      const float d = P1[q * kRow + c];
      bsum += d;
#pragma unroll
      for (int i = 0; i < kIn; ++i) acc[i] = fmaf(d, X[q * 4 + i], acc[i]);
    }
#pragma unroll
    for (int i = 0; i < kIn; ++i) out[oW1 + c * kIn + i] = acc[i];
    out[oB1 + c] = bsum;
  }
  // output layer: dW4[o][c] = sum_s d4[s][o] * a3[s][c]; db4[o]
  {
    float acc[kOut] = {0.f, 0.f, 0.f};
    for (int q = 0; q < kTile; ++q) {
      const float a = A3[q * kRow + c];
#pragma unroll
      for (int o = 0; o < kOut; ++o) acc[o] = fmaf(D4[q * 4 + o], a, acc[o]);
    }
#pragma unroll
    for (int o = 0; o < kOut; ++o) out[oW4 + o * kH + c] = acc[o];
    if (c < kOut) {
      float b = 0.f;
      for (int q = 0; q < kTile; ++q) b += D4[q * 4 + c];
      out[oB4 + c] = b;
    }
  }
  // block loss (wave reduction; the block is one wave)
  loss = wave_sum(loss);
  if (c == 0) out[kWlds] = loss;
}

// Deterministic sum of the per-block rows (fixed order, FP64).
__global__ void __launch_bounds__(256) k_mlp_reduce(const float* __restrict__ part, int nblocks,
                                                    float* __restrict__ grad,
                                                    float* __restrict__ loss) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p > kN) return;
  const int col = p < kN ? p : kWlds;
  double acc = 0.0;
  for (int b = 0; b < nblocks; ++b) acc += (double)part[(int64_t)b * kStride + col];
  if (p < kN)
    grad[p] = (float)acc;
  else if (loss)
    *loss = (float)acc;
}

__global__ void __launch_bounds__(256) k_mlp_adam(bz_adam_t opt, float* __restrict__ master,
                                                  float* __restrict__ m, float* __restrict__ v,
                                                  const float* __restrict__ grad,
                                                  uint16_t* __restrict__ wbf) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kN) return;
  const float g = grad[p];
  const float mp = opt.beta1 * m[p] + (1.f - opt.beta1) * g;
  const float vp = opt.beta2 * v[p] + (1.f - opt.beta2) * g * g;
  m[p] = mp;
  v[p] = vp;
  const float mh = mp / opt.bias1;
  const float vh = vp / opt.bias2;
  const float w = master[p] - opt.lr * mh / (sqrtf(vh) + opt.eps);
  master[p] = w;
  wbf[p] = f32_bf16(w);
}

__global__ void __launch_bounds__(256) k_mlp_round(const float* __restrict__ master,
                                                   uint16_t* __restrict__ wbf) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < kN) wbf[p] = f32_bf16(master[p]);
}

Norm norm_of(const bz_mlp_t* net) {
  Norm nm;
  for (int i = 0; i < kIn; ++i) {
    nm.in_shift[i] = (float)net->in_shift[i];
    nm.in_scale[i] = (float)net->in_scale[i];
    nm.in_lo[i] = net->in_lo[i];
    nm.in_hi[i] = net->in_hi[i];
  }
  for (int o = 0; o < kOut; ++o) {
    nm.out_shift[o] = (float)net->out_shift[o];
    nm.out_scale[o] = (float)net->out_scale[o];
    nm.out_inv[o] = (float)(1.0 / net->out_scale[o]);
  }
  return nm;
}

size_t grad_lds_bytes() { return sizeof(float) * (kWlds + 8 * kTile + 6 * kTile * kRow); }

}  // namespace mlp

int64_t mlp_workspace_floats(int64_t n) {
  return ((n + mlp::kTile - 1) / mlp::kTile) * (int64_t)mlp::kStride;
}

int launch_mlp_forward(const LaunchEnv& env, const bz_mlp_t* net, const float* x, float* y,
                       int64_t n) {
  if (n <= 0) return 0;
  int64_t blocks = (n + mlp::kFwdBlock - 1) / mlp::kFwdBlock;
  const int64_t cap = (int64_t)env.num_cu * 4;
  if (blocks > cap) blocks = cap;
  hipLaunchKernelGGL(mlp::k_mlp_forward, dim3((unsigned)blocks), dim3(mlp::kFwdBlock), 0,
                     env.stream,
                     mlp::norm_of(net), net->w, x, y, n);
  return check_launch("k_mlp_forward");
}

int launch_mlp_grad(const LaunchEnv& env, const bz_mlp_t* net, const float* x, const float* t,
                    int64_t n, float* ws, float* grad, float* loss) {
  const int64_t blocks = (n + mlp::kTile - 1) / mlp::kTile;
  if (blocks > (1 << 20)) return fail("mlp grad: batch too large (%lld)", (long long)n);
  const size_t lds = mlp::grad_lds_bytes();
  static bool attr = false;
  if (!attr) {
    const hipError_t e = hipFuncSetAttribute((const void*)mlp::k_mlp_grad,
                                             hipFuncAttributeMaxDynamicSharedMemorySize,
                                             (int)lds);
    if (e != hipSuccess) return fail("mlp grad: LDS attribute: %s", hipGetErrorString(e));
    attr = true;
  }
  hipLaunchKernelGGL(mlp::k_mlp_grad, dim3((unsigned)blocks), dim3(mlp::kTile), lds, env.stream,
                     mlp::norm_of(net), net->w, x, t, n, (float)(1.0 / (double)n), ws);
  int rc = check_launch("k_mlp_grad");
  if (rc) return rc;
  hipLaunchKernelGGL(mlp::k_mlp_reduce, dim3((mlp::kN + 1 + 255) / 256), dim3(256), 0,
                     env.stream, ws, (int)blocks, grad, loss);
  return check_launch("k_mlp_reduce");
}

int launch_mlp_adam(const LaunchEnv& env, const bz_adam_t* opt, float* master, float* m, float* v,
                    const float* grad, uint16_t* wbf) {
  hipLaunchKernelGGL(mlp::k_mlp_adam, dim3((mlp::kN + 255) / 256), dim3(256), 0, env.stream,
                     *opt, master, m, v, grad, wbf);
  return check_launch("k_mlp_adam");
}

int launch_mlp_round(const LaunchEnv& env, const float* master, uint16_t* wbf) {
  hipLaunchKernelGGL(mlp::k_mlp_round, dim3((mlp::kN + 255) / 256), dim3(256), 0, env.stream,
                     master, wbf);
  return check_launch("k_mlp_round");
}

}  // namespace bz

// ---------------------------------------------------------------------------
// Closure source on the grid: one wave per node, one lane per hidden unit
// (each lane keeps its rows of W1..W3 and its column of W4 in registers), so
// a node's three layers take ~200 lane-parallel FMAs instead of 8.8k serial
// ones: the launch latency, not the arithmetic, bounds a sub-step.
namespace bz {
namespace mlp {

__global__ void __launch_bounds__(256) k_closure_source(Norm nm, const uint16_t* __restrict__ w,
                                                        int N, const double* __restrict__ E,
                                                        const double* __restrict__ mom,
                                                        double gas_m3, double dt,
                                                        double* __restrict__ src,
                                                        double* __restrict__ acc) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int nn = N + 1;
  float w1[kIn], w2[kH], w3[kH], w4[kOut];
#pragma unroll
  for (int i = 0; i < kIn; ++i) w1[i] = bf16_f32(w[oW1 + lane * kIn + i]);
#pragma unroll
  for (int k = 0; k < kH; ++k) {
    w2[k] = bf16_f32(w[oW2 + lane * kH + k]);
    w3[k] = bf16_f32(w[oW3 + lane * kH + k]);
  }
#pragma unroll
  for (int o = 0; o < kOut; ++o) w4[o] = bf16_f32(w[oW4 + o * kH + lane]);
  const float b1 = bf16_f32(w[oB1 + lane]), b2 = bf16_f32(w[oB2 + lane]),
              b3 = bf16_f32(w[oB3 + lane]);
  const float b40 = bf16_f32(w[oB4]);
  for (int node = blockIdx.x * 4 + wave; node < nn; node += gridDim.x * 4) {
    // local reduced field (node value: mean |E| of the adjacent cells)
    double en;
    if (node == 0) en = fabs(E[0]);
    else if (node == N) en = fabs(E[N - 1]);
    else en = 0.5 * (fabs(E[node - 1]) + fabs(E[node]));
    const double en_td = fmax(en * 100.0 / gas_m3 * 1e21, 1e-3);
    const double ne = mom[node];
    const double me = ne > 0.0 ? mom[2 * nn + node] / ne : 0.0;
    const double lne = log10(fmax(ne * 1e6, 1.0));
    float x[kIn];
    x[0] = ((float)clamp_feature(nm, 0, log10(en_td)) - nm.in_shift[0]) * nm.in_scale[0];
    x[1] = ((float)clamp_feature(nm, 1, me) - nm.in_shift[1]) * nm.in_scale[1];
    x[2] = ((float)clamp_feature(nm, 2, lne) - nm.in_shift[2]) * nm.in_scale[2];
    float z = b1, a, d;
#pragma unroll
    for (int i = 0; i < kIn; ++i) z = fmaf(w1[i], x[i], z);
    silu(z, a, d);
    float h = a;
    z = b2;
#pragma unroll
    for (int k = 0; k < kH; ++k) z = fmaf(w2[k], __shfl(h, k, 64), z);
    silu(z, a, d);
    h = a;
    z = b3;
#pragma unroll
    for (int k = 0; k < kH; ++k) z = fmaf(w3[k], __shfl(h, k, 64), z);
    silu(z, a, d);
    h = a;
    const float y0 = wave_sum(w4[0] * h) + b40;  // log10 k_ion, normalised
    if (lane == 0) {
      const double lk = (double)fmaf(y0, nm.out_scale[0], nm.out_shift[0]);
      const double s = pow(10.0, lk) * gas_m3 * ne * dt;
      src[node] = s;
      if (acc) acc[node] += s;
    }
  }
}

}  // namespace mlp

int launch_closure_source(const LaunchEnv& env, const bz_mlp_t* net, int N, const double* E,
                          const double* mom, double gas_m3, double dt, double* src, double* acc) {
  const int nn = N + 1;
  int blocks = (nn + 3) / 4;
  if (blocks > env.num_cu) blocks = env.num_cu;
  hipLaunchKernelGGL(mlp::k_closure_source, dim3(blocks), dim3(256), 0, env.stream,
                     mlp::norm_of(net), net->w, N, E, mom, gas_m3, dt, src, acc);
  return check_launch("k_closure_source");
}

}  // namespace bz
