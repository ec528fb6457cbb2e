// Per-particle physics of the electron step (device side).
//
// Mirrors kernels/electron_kernel_numba.py:39-287 of the reference for the
// fit model and adds the tabulated three-channel null-collision model.  All
// arithmetic on the decision path is + - * / sqrt and compares, evaluated in
// the order the reference writes it, compiled with -ffp-contract=off.
#pragma once

#include "bz_device.h"

namespace bz {

template <typename T>
struct P5 {
  T x, vx, vy, vz, w;
};

// Argon fits of the reference (electron_kernel_numba.py:67-171):
// log10(sigma) = (a*L + b)*L + c with L = log10(E/eV), piecewise in E.
__constant__ double kElBreak[11] = {0.000780396, 0.006098667, 0.0476269, 0.1015574,
                                    0.1943527,   0.2995781,   1.097156,  2.339524,
                                    4.988691,    10.63765,    22.68322};
__constant__ double kElCoef[12][3] = {
    {-0.02704763, -0.23720051, -19.67900951}, {-0.08847237, -0.6084786, -20.24111992},
    {-0.37608274, -1.8167778, -21.51414308},  {-1.4874467, -4.82420619, -23.55745478},
    {-0.9870356, -4.2206026, -23.44715988},   {14.28063581, 18.92275458, -14.70556113},
    {-2.12069169, 0.7555105, -19.71943671},   {-0.3585636, 0.79246666, -19.71260274},
    {1.25262128, -0.40002029, -19.48909003},  {-2.28332905, 4.89566076, -21.46789173},
    {-1.47508661, 2.82263476, -20.1928456},   {-0.11090525, -0.77506081, -17.81752307}};
__constant__ double kIonBreak[7] = {15.77742, 15.81964, 15.96196, 16.44162,
                                    18.05837, 23.50771, 41.87498};
__constant__ double kIonCoef[8][3] = {
    {-2149239.99337822, 5150451.29014456, -3085665.64874862},
    {-227229.83440255, 545171.14839945, -327016.69508374},
    {-20173.68563579, 48599.13254342, -29290.96100603},
    {-1838.90888364, 4491.58838596, -2763.82866252},
    {-186.98923927, 477.04784782, -324.75732578},
    {-24.48843901, 69.43965607, -69.145292411},
    {-4.33101714, 14.47124519, -31.66224458},
    {-0.68345667, 2.6053539, -21.99823511}};

__device__ __forceinline__ double dlog10(double v) { return log10(v); }
__device__ __forceinline__ float dlog10(float v) { return log10f(v); }
__device__ __forceinline__ double dpow10(double v) { return pow(10.0, v); }
__device__ __forceinline__ float dpow10(float v) { return powf(10.0f, v); }
// e^v - 1 and e^v in plain IEEE operations, written the same way in the
// oracle (orc_expm1), so the collision probabilities of the direct and the
// fit modes agree bit for bit between the device and the CPU check:
// v = k ln2 + r with k = floor(v / ln2 + 1/2) and r = (v - k ln2_hi) -
// k ln2_lo (Cody-Waite split; k ln2_hi is exact), e^r - 1 = r + r^2 P(r)
// with P the Taylor polynomial in Horner form (e^r - 1 to the r^15 term in
// FP64, r^8 in FP32; |r| <= 0.35), then e^v - 1 = 2^k (e^r - 1) + (2^k - 1).
// v is clamped to [-80, 80] (e^-80 is a normal FP32 number, so no
// denormal is formed; below it e^v - 1 rounds to -1 in either precision).
template <typename T>
struct ExpSplit;
template <>
struct ExpSplit<double> {
  static constexpr double inv_ln2 = 1.4426950408889634;
  static constexpr double ln2_hi = 6.93147180369123816490e-01;
  static constexpr double ln2_lo = 1.90821492927058770002e-10;
  __device__ __forceinline__ static double poly(double r) {
    double q = 7.647163731819816e-13;
    q = q * r + 1.1470745597729725e-11;
    q = q * r + 1.6059043836821613e-10;
    q = q * r + 2.08767569878681e-09;
    q = q * r + 2.505210838544172e-08;
    q = q * r + 2.755731922398589e-07;
    q = q * r + 2.755731922398589e-06;
    q = q * r + 2.48015873015873e-05;
    q = q * r + 0.0001984126984126984;
    q = q * r + 0.001388888888888889;
    q = q * r + 0.008333333333333333;
    q = q * r + 0.041666666666666664;
    q = q * r + 0.16666666666666666;
    q = q * r + 0.5;
    return q;
  }
};
template <>
struct ExpSplit<float> {
  static constexpr float inv_ln2 = 1.44269504f;
  static constexpr float ln2_hi = 0.693359375f;
  static constexpr float ln2_lo = -2.12194440e-4f;
  __device__ __forceinline__ static float poly(float r) {
    float q = 2.48015873e-05f;
    q = q * r + 1.98412698e-04f;
    q = q * r + 1.38888889e-03f;
    q = q * r + 8.33333333e-03f;
    q = q * r + 4.16666667e-02f;
    q = q * r + 1.66666667e-01f;
    q = q * r + 0.5f;
    return q;
  }
};

// e^r - 1 of the reduced argument and k
template <typename T>
__device__ __forceinline__ T expm1_reduced(T v, int* k) {
  using C = ExpSplit<T>;
  v = v < (T)-80 ? (T)-80 : (v > (T)80 ? (T)80 : v);
  const T kf = floor(v * C::inv_ln2 + (T)0.5);
  const T r = (v - kf * C::ln2_hi) - kf * C::ln2_lo;
  *k = (int)kf;
  return r + (r * r) * C::poly(r);
}

template <typename T>
__device__ __forceinline__ T dexpm1(T v) {
  int k;
  const T e = expm1_reduced<T>(v, &k);
  if (k == 0) return e;
  const T s = ldexp((T)1, k);
  return s * e + (s - (T)1);
}

template <typename T>
__device__ __forceinline__ T dexp(T v) {
  int k;
  const T e = expm1_reduced<T>(v, &k);
  return ldexp((T)1 + e, k);
}
__device__ __forceinline__ double dlog(double v) { return log(v); }
__device__ __forceinline__ float dlog(float v) { return logf(v); }
__device__ __forceinline__ double dsqrt(double v) { return sqrt(v); }
__device__ __forceinline__ float dsqrt(float v) { return sqrtf(v); }
__device__ __forceinline__ double dlog1p(double v) { return log1p(v); }
__device__ __forceinline__ float dlog1p(float v) { return log1pf(v); }
// sin(pi v), cos(pi v): the argument reduction is exact, so an angle 2*pi*u
// with u in [0,1) costs no large-argument reduction.
__device__ __forceinline__ void dsincospi(double v, double* s, double* c) { sincospi(v, s, c); }
__device__ __forceinline__ void dsincospi(float v, float* s, float* c) { sincospif(v, s, c); }

template <typename T>
__device__ __forceinline__ T fit_poly(const double (*coef)[3], int seg, T L) {
  const T a = (T)coef[seg][0], b = (T)coef[seg][1], c = (T)coef[seg][2];
  return (a * L + b) * L + c;
}

// Fit model collision thresholds (electron_kernel_numba.py:115-195).
template <typename T>
__device__ __forceinline__ void fit_thresholds(T e, T vmag, const bz_params_t& p, T& r0, T& r1,
                                               T& r2) {
  const T L = dlog10(e);
  int seg = 0;
#pragma unroll
  for (int k = 0; k < 11; ++k) seg += (e >= (T)kElBreak[k]) ? 1 : 0;
  T sig0 = dpow10(fit_poly<T>(kElCoef, seg, L));
  const T eion = (T)p.e_ion;
  T sig2;
  if (e < eion) {
    sig2 = (T)0;
  } else if (e == eion) {
    sig2 = (T)1;  // no interval matches: 10**0 (electron_kernel_numba.py:164-174)
  } else {
    int s2 = 0;
#pragma unroll
    for (int k = 0; k < 7; ++k) s2 += (e >= (T)kIonBreak[k]) ? 1 : 0;
    sig2 = dpow10(fit_poly<T>(kIonCoef, s2, L));
  }
  const T nn = (T)p.nn;
  sig0 *= nn;
  sig2 *= nn;
  const T sig_tot = sig0 + sig2;
  const T P = (T)1 - dexp(-(T)p.dt * vmag * sig_tot);
  const T pcst = P * ((T)1 / sig_tot);
  r0 = sig0 * pcst;
  r1 = r0;  // no excitation channel in the fit model
  r2 = r0 + sig2 * pcst;
}

// Tabulated model: sigma*N piecewise linear in E, clamped to the table,
// zero below each threshold; null-collision thresholds nu_k * P_null/nu_max.
template <typename T>
__device__ __forceinline__ int xs_bucket(T e, int bits);
template <>
__device__ __forceinline__ int xs_bucket<double>(double e, int bits) {
  return (int)((uint64_t)__double_as_longlong(e) >> (52 - bits));
}
template <>
__device__ __forceinline__ int xs_bucket<float>(float e, int bits) {
  return (int)(__float_as_uint(e) >> (23 - bits));
}

// Cumulative collision probabilities r_j = sum_{c<=j} (sigma_c N v) * scale of
// the tabulated model, sigma linear in E between table nodes.
template <typename T>
__device__ __forceinline__ void table_rates(T e, T vmag, const bz_params_t& p,
                                            const bz_xsec_t& xs, T scale, T& r0, T& r1, T& r2,
                                            bool& over);

template <typename T>
__device__ __forceinline__ void table_thresholds(T e, T vmag, const bz_params_t& p,
                                                 const bz_xsec_t& xs, T& r0, T& r1, T& r2,
                                                 bool& over) {
  table_rates<T>(e, vmag, p, xs, (T)p.pc, r0, r1, r2, over);
  over = r2 > (T)p.p_null;
}

// Direct model: cumulative frequencies nu_j of the tabulated sigma, then
// P = 1 - exp(-nu_t dt) and r_j = nu_j * P / nu_t (the reference's
// P * sigma_j / sigma_t, electron_kernel_numba.py:186-212).
template <typename T>
__device__ __forceinline__ void direct_thresholds(T e, T vmag, const bz_params_t& p,
                                                  const bz_xsec_t& xs, T& r0, T& r1, T& r2) {
  bool unused;
  table_rates<T>(e, vmag, p, xs, (T)1, r0, r1, r2, unused);
  if (r2 > (T)0) {
    const T P = -dexpm1(-(T)p.dt * r2);
    const T pcst = P * ((T)1 / r2);
    r0 = r0 * pcst;
    r1 = r1 * pcst;
    r2 = r2 * pcst;
  } else {
    r0 = r1 = r2 = (T)0;
  }
}

template <typename T>
__device__ __forceinline__ void table_rates(T e, T vmag, const bz_params_t& p,
                                            const bz_xsec_t& xs, T scale, T& r0, T& r1, T& r2,
                                            bool& over) {
  const T* E = (const T*)xs.energy;
  const T* S = (const T*)xs.sig;
  const T* D = (const T*)xs.slope;
  const int n = xs.n;
  const T e0 = E[0], e1 = E[n - 1];
  T ec = e < e0 ? e0 : e;
  ec = ec > e1 ? e1 : ec;
  int b = xs_bucket<T>(ec, xs.hint_bits) - xs.hint_base;
  b = b < 0 ? 0 : (b >= xs.n_hint ? xs.n_hint - 1 : b);
  int j = xs.hint[b];
  // a hint bucket spans one or two table intervals at the tables' spacing:
  // two branch-free steps, then a loop the host bounds to 7 steps in all
#pragma unroll
  for (int it = 0; it < 2; ++it) {
    const int q = j + 1 <= n - 2 ? j + 1 : n - 2;  // in range, so read unconditionally
    j += (q == j + 1 && E[q] <= ec) ? 1 : 0;
  }
  for (int it = 2; it < 8; ++it) {
    if (j + 1 <= n - 2 && E[j + 1] <= ec) ++j;
    else break;
  }
  const T d = ec - E[j];
  T s0 = S[j] + D[j] * d;
  T s1 = S[n + j] + D[n + j] * d;
  T s2 = S[2 * n + j] + D[2 * n + j] * d;
  if (e < (T)p.e_exc) s1 = (T)0;
  if (e < (T)p.e_ion) s2 = (T)0;
  r0 = (s0 * vmag) * scale;
  r1 = r0 + (s1 * vmag) * scale;
  r2 = r1 + (s2 * vmag) * scale;
  over = r2 > (T)1;
}

// Upper bound of the collision probability a particle can have in energy
// bucket b of the table's hint (the bucket holds energies with the same top
// bits, below bucket_upper): the largest sigma of every channel over the
// table nodes that bracket the bucket, times the speed at the bucket's upper
// energy, gives nu_up >= nu_t(e) for every e in it; the bound is P =
// 1 - exp(-nu_up dt) (direct) or nu_up pc (null collision), raised by 1e-4
// relative to cover the rounding of the per-particle evaluation.  A uniform
// R0 at or above it means no collision, exactly as the full evaluation
// decides.  A bucket whose null-collision bound passes P_null returns 2 (every
// particle of it takes the full path, which counts nu_max overruns).
template <typename T>
__device__ __forceinline__ double bucket_upper(int b_abs, int bits);
template <>
__device__ __forceinline__ double bucket_upper<double>(int b_abs, int bits) {
  return __longlong_as_double((long long)((uint64_t)(b_abs + 1) << (52 - bits)));
}
template <>
__device__ __forceinline__ double bucket_upper<float>(int b_abs, int bits) {
  return (double)__uint_as_float((uint32_t)(b_abs + 1) << (23 - bits));
}

template <typename T, int XS>
__device__ double bucket_bound(int b, const bz_xsec_t& xs, const bz_params_t& p) {
  const T* E = (const T*)xs.energy;
  const T* S = (const T*)xs.sig;
  const int n = xs.n;
  const int j0 = xs.hint[b];
  int j1 = b + 1 < xs.n_hint ? xs.hint[b + 1] + 1 : n - 1;
  j1 = j1 < n - 1 ? j1 : n - 1;
  double sig_up = 0.0;
  for (int c = 0; c < 3; ++c) {
    double m = 0.0;
    for (int j = j0; j <= j1; ++j) m = fmax(m, (double)S[c * n + j]);
    sig_up += m;
  }
  double e_hi = bucket_upper<T>(b + xs.hint_base, xs.hint_bits);
  e_hi = fmin(e_hi, (double)E[n - 1]);
  const double nu_up = sig_up * sqrt(e_hi / p.kev);
  const double P = XS == BZ_XS_DIRECT ? -expm1(-p.dt * nu_up) : nu_up * p.pc;
  const double Pm = P * (1.0 + 1e-4) + 1e-30;
  if (XS != BZ_XS_DIRECT && Pm > p.p_null) return 2.0;
  return Pm;
}

template <typename T>
__device__ __forceinline__ void iso_dir(T vnew, T ra, T rb, T& vx, T& vy, T& vz) {
  const T cchi = (T)2 * ra - (T)1;
  T s2 = (T)1 - cchi * cchi;
  const T schi = dsqrt(s2 > (T)0 ? s2 : (T)0);
  T sphi, cphi;
  dsincospi((T)2 * rb, &sphi, &cphi);  // phi = 2 pi rb
  vx = vnew * schi * cphi;
  vy = vnew * schi * sphi;
  vz = vnew * cchi;
}

template <typename T>
__device__ __forceinline__ void gas_velocity(uint64_t seed, uint32_t i, uint64_t step, T sigma,
                                             T& gx, T& gy, T& gz);
template <>
__device__ __forceinline__ void gas_velocity<double>(uint64_t seed, uint32_t i, uint64_t step,
                                                     double sigma, double& gx, double& gy,
                                                     double& gz) {
  const U4 a = draw(seed, kPurposeGas, i, 0, step);
  const U4 b = draw(seed, kPurposeGas, i, 1, step);
  const double g0 = u53(a.a, a.b), g1 = u53(a.c, a.d), g2 = u53(b.a, b.b), g3 = u53(b.c, b.d);
  const double r1 = sqrt(-2.0 * log(1.0 - g0)), r2 = sqrt(-2.0 * log(1.0 - g2));
  double s, c;
  sincospi(2.0 * g1, &s, &c);
  gx = sigma * (r1 * c);
  gy = sigma * (r1 * s);
  gz = sigma * (r2 * cospi(2.0 * g3));
}
template <>
__device__ __forceinline__ void gas_velocity<float>(uint64_t seed, uint32_t i, uint64_t step,
                                                    float sigma, float& gx, float& gy,
                                                    float& gz) {
  const U4 a = draw(seed, kPurposeGas, i, 0, step);
  const float g0 = u24(a.a), g1 = u24(a.b), g2 = u24(a.c), g3 = u24(a.d);
  const float r1 = sqrtf(-2.0f * logf(1.0f - g0)), r2 = sqrtf(-2.0f * logf(1.0f - g2));
  float s, c;
  sincospif(2.0f * g1, &s, &c);
  gx = sigma * (r1 * c);
  gy = sigma * (r1 * s);
  gz = sigma * (r2 * cospif(2.0f * g3));
}

enum : int { kChanNone = 0, kChanElastic = 1, kChanExc = 2, kChanIon = 3 };

template <typename T>
struct StepOut {
  P5<T> p;      // primary after collision, push and walls
  P5<T> s;      // secondary (valid when born)
  T e_post;     // primary energy after push, eV
  T e_sec;      // secondary energy, eV
  int alive;
  int born;
  int chan;     // kChanIon when a secondary was born, else the last event's channel
  int over;
  int n_el, n_exc, n_null;  // events of the step (several in free-flight mode)
};

// The field of the cells for the 1-D kick: a plain pointer, or FieldRef,
// which reads the copy staged in LDS through an LDS-typed pointer (a ds_read
// that waits for nothing else in flight) and global memory otherwise.
template <typename T>
struct FieldRef {
  const T* g;
  __attribute__((address_space(3))) const T* s;
  bool lds;
  __device__ __forceinline__ T operator[](int c) const { return lds ? s[c] : g[c]; }
};

// Push and walls after the collision (electron_kernel_numba.py:257-287):
// velocity (vx, vy, vz) after the collision, field of the particle's cell
// (1-D) or of its swarm, the configured update order; fills o.p, o.alive
// and o.e_post.
template <typename T, int DIM, typename EF>
__device__ __forceinline__ void push_tail(const P5<T>& in, T vx, T vy, T vz,
                                          const bz_params_t& p, const EF& E,
                                          T e_local, StepOut<T>& o) {
  T x = in.x;
  T ef;
  if (DIM == 1) {
    int c = cell_of((double)in.x, p.dx, 1.0 / p.dx);
    c = c < 0 ? 0 : (c >= p.n_cells ? p.n_cells - 1 : c);
    ef = E[c];
  } else {
    ef = e_local;  // 0-D: uniform field of the swarm (or of its instance)
  }
  if (p.push_order == BZ_PUSH_LEAPFROG) {
    // kick with E at the step's starting cell, then drift with the new v
    vx = vx - (T)p.afac * ((T)100 * ef);
    if (DIM == 1) x = x + (T)p.xfac * vx;
  } else {
    // reference order: drift with the post-collision v, then kick
    if (DIM == 1) x = x + (T)p.xfac * vx;
    vx = vx - (T)p.afac * ((T)100 * ef);
  }
  if (DIM == 1) o.alive = ((T)0 < x && x < (T)p.L) ? 1 : 0;
  else o.alive = 1;
  o.p.x = x;
  o.p.vx = vx;
  o.p.vy = vy;
  o.p.vz = vz;
  o.p.w = in.w;
  o.e_post = (T)p.kev * (vx * vx + vy * vy + vz * vz);
}

// One particle through collision + push + walls
// (electron_kernel_numba.py:39-287; dead particles are dropped, not pushed).
// r0p: the particle's decision uniform when the caller has drawn it (the
// Philox stream; null: drawn here).
template <typename T, int XS, int DIM, int RNG, typename EF>
__device__ __forceinline__ StepOut<T> step_particle(const P5<T> in, uint32_t i, uint64_t step,
                                                    const bz_params_t& p, const bz_xsec_t& xs,
                                                    const EF& E, const T* R,
                                                    const T* G, T e_local,
                                                    const T* r0p = nullptr) {
  StepOut<T> o;
  o.alive = 0;
  o.born = 0;
  o.chan = kChanNone;
  o.over = 0;
  o.n_el = o.n_exc = o.n_null = 0;
  o.e_post = (T)0;
  o.e_sec = (T)0;
  o.s = in;
  if (!(in.w > (T)0)) {
    o.p = in;
    return o;
  }
  T gx = (T)0, gy = (T)0, gz = (T)0;
  if (RNG == BZ_RNG_STREAM) {
    if (G) {
      const int64_t st = p.r_stride;
      gx = G[i];
      gy = G[st + i];
      gz = G[2 * st + i];
    }
  } else if (p.gas_sigma > 0.0) {
    // a uniform branch; the empty asm keeps the compiler from evaluating the
    // draw for every particle and selecting zero when the gas is cold
    asm volatile("" ::: "memory");
    gas_velocity<T>(p.seed, i, step, (T)p.gas_sigma, gx, gy, gz);
  }
  T vx = in.vx, vy = in.vy, vz = in.vz;
  const T ux = vx - gx, uy = vy - gy, uz = vz - gz;
  const T s = ux * ux + uy * uy + uz * uz;
  const T vmag = dsqrt(s);
  const T e = (T)p.kev * s;

  Uniforms<T> U;
  if (RNG == BZ_RNG_STREAM) {
    const int64_t st = p.r_stride;
    U.u[0] = R[i];
    U.u[1] = R[st + i];
    U.u[2] = R[2 * st + i];
    U.u[3] = R[3 * st + i];
    U.u[4] = R[5 * st + i];
  } else {
    U.u[0] = r0p ? *r0p : Decision<T>{}.get(p.seed, i, step);
  }

  T r0, r1, r2;
  bool over = false;
  if (XS == BZ_XS_FIT) fit_thresholds<T>(e, vmag, p, r0, r1, r2);
  else if (XS == BZ_XS_DIRECT) direct_thresholds<T>(e, vmag, p, xs, r0, r1, r2);
  else table_thresholds<T>(e, vmag, p, xs, r0, r1, r2, over);
  o.over = over ? 1 : 0;

  const T R0 = U.u[0];
  T tx = (T)0, ty = (T)0, tz = (T)0;
  if (R0 < r0) o.chan = kChanElastic;
  else if (XS != BZ_XS_FIT && R0 < r1) o.chan = kChanExc;
  else if ((XS != BZ_XS_FIT && R0 < r2) || (XS == BZ_XS_FIT && r0 < R0 && R0 < r2))
    o.chan = kChanIon;
  if (o.chan != kChanNone) {
    // one scatter for every channel (a wave with mixed channels runs it
    // once); an ionization also takes the secondary's angles
    if (RNG != BZ_RNG_STREAM) U.scatter(p.seed, i, step, o.chan == kChanIon);
    T vnew;
    if (o.chan == kChanElastic) {
      vnew = vmag * (T)p.fel;
    } else {
      const T enew = o.chan == kChanExc ? e - (T)p.e_exc : fabs((T)0.5 * (e - (T)p.e_ion));
      vnew = dsqrt(((T)2 * enew) / (T)p.jm);
    }
    iso_dir<T>(vnew, U.u[1], U.u[2], vx, vy, vz);
    if (o.chan == kChanIon) {
      iso_dir<T>(vnew, U.u[3], U.u[4], tx, ty, tz);
      o.born = 1;
      o.s.x = in.x;
      o.s.vx = tx;
      o.s.vy = ty;
      o.s.vz = tz;
      o.s.w = in.w;
      o.e_sec = (T)p.kev * (tx * tx + ty * ty + tz * tz);
    }
  }

  o.n_el = o.chan == kChanElastic;
  o.n_exc = o.chan == kChanExc;
  o.n_null = o.chan == kChanNone;
  push_tail<T, DIM>(in, vx, vy, vz, p, E, e_local, o);
  return o;
}

// Uniforms of free-flight event k of particle i: flight, channel and the two
// scattering angles; the secondary's two angles come from a further block.
template <typename T>
struct FlightDraws;
template <>
struct FlightDraws<double> {
  __device__ __forceinline__ static void event(uint64_t seed, uint32_t i, uint64_t step,
                                               uint32_t k, double u[4]) {
    const U4 r = draw(seed, kPurposeFlight, i, 3 * k, step);
    const U4 q = draw(seed, kPurposeFlight, i, 3 * k + 1, step);
    u[0] = u53(r.a, r.b);
    u[1] = u53(r.c, r.d);
    u[2] = u53(q.a, q.b);
    u[3] = u53(q.c, q.d);
  }
  __device__ __forceinline__ static void secondary(uint64_t seed, uint32_t i, uint64_t step,
                                                   uint32_t k, double u[2]) {
    const U4 r = draw(seed, kPurposeFlight, i, 3 * k + 2, step);
    u[0] = u53(r.a, r.b);
    u[1] = u53(r.c, r.d);
  }
};
template <>
struct FlightDraws<float> {
  __device__ __forceinline__ static void event(uint64_t seed, uint32_t i, uint64_t step,
                                               uint32_t k, float u[4]) {
    const U4 r = draw(seed, kPurposeFlight, i, 2 * k, step);
    u[0] = u24(r.a);
    u[1] = u24(r.b);
    u[2] = u24(r.c);
    u[3] = u24(r.d);
  }
  __device__ __forceinline__ static void secondary(uint64_t seed, uint32_t i, uint64_t step,
                                                   uint32_t k, float u[2]) {
    const U4 r = draw(seed, kPurposeFlight, i, 2 * k + 1, step);
    u[0] = u24(r.a);
    u[1] = u24(r.b);
  }
};

constexpr int kMaxFlights = 256;

// One particle through a step of exact free-flight null-collision MC: flight
// times are drawn from exp(nu_max); at the end of each flight a real
// collision of channel j happens with probability nu_j(e) / nu_max, so any
// number of collisions fits in one step and the step length is not limited
// by the collision frequency.  The field is the cell-constant E of the
// particle's starting cell for the whole step (as in the single-collision
// step).  Each flight moves x with the velocity at its start and then
// applies the acceleration; a particle leaving the gap at a flight's end is
// absorbed.  One secondary at most per step: a further ionizing event in
// the same step is treated as a null event.  The secondary starts at the
// collision point and is not moved in its birth step.  Tabulated sigma and
// the Philox stream only.
template <typename T, int DIM, typename EF>
__device__ __forceinline__ StepOut<T> flight_particle(const P5<T> in, uint32_t i, uint64_t step,
                                                      const bz_params_t& p, const bz_xsec_t& xs,
                                                      const EF& E, T e_local) {
  StepOut<T> o;
  o.alive = 0;
  o.born = 0;
  o.chan = kChanNone;
  o.over = 0;
  o.n_el = o.n_exc = o.n_null = 0;
  o.e_post = (T)0;
  o.e_sec = (T)0;
  o.s = in;
  if (!(in.w > (T)0)) {
    o.p = in;
    return o;
  }
  T ef;
  if (DIM == 1) {
    int c = cell_of((double)in.x, p.dx, 1.0 / p.dx);
    c = c < 0 ? 0 : (c >= p.n_cells ? p.n_cells - 1 : c);
    ef = E[c];
  } else {
    ef = e_local;
  }
  const T acc = (T)(p.afac / p.dt) * ((T)100 * ef);  // dv_x/dt, m/s^2
  const T inv_nu = (T)(1.0 / p.nu_max);
  T x = in.x, vx = in.vx, vy = in.vy, vz = in.vz;
  T tau = (T)p.dt;
  int alive = 1;
  for (int k = 0; k < kMaxFlights; ++k) {
    T u[4];
    FlightDraws<T>::event(p.seed, i, step, (uint32_t)k, u);
    const T tf = -dlog1p(-u[0]) * inv_nu;
    const bool hit = tf < tau;
    const T h = hit ? tf : tau;
    if (p.push_order == BZ_PUSH_LEAPFROG) {
      vx = vx - acc * h;
      if (DIM == 1) x = x + ((T)100 * h) * vx;
    } else {
      if (DIM == 1) x = x + ((T)100 * h) * vx;
      vx = vx - acc * h;
    }
    if (DIM == 1 && !((T)0 < x && x < (T)p.L)) {
      alive = 0;
      break;
    }
    if (!hit) {
      tau = (T)0;
      break;
    }
    tau = tau - tf;
    // collision at the end of the flight
    const T s = vx * vx + vy * vy + vz * vz;
    const T vmag = dsqrt(s);
    const T e = (T)p.kev * s;
    T r0, r1, r2;
    bool over = false;
    table_rates<T>(e, vmag, p, xs, inv_nu, r0, r1, r2, over);
    o.over |= over ? 1 : 0;
    const T R = u[1];
    if (R < r0) {
      ++o.n_el;
      o.chan = kChanElastic;
      iso_dir<T>(vmag * (T)p.fel, u[2], u[3], vx, vy, vz);
    } else if (R < r1) {
      ++o.n_exc;
      o.chan = kChanExc;
      const T enew = e - (T)p.e_exc;
      iso_dir<T>(dsqrt(((T)2 * enew) / (T)p.jm), u[2], u[3], vx, vy, vz);
    } else if (R < r2 && !o.born) {
      const T eej = fabs((T)0.5 * (e - (T)p.e_ion));
      const T vnew = dsqrt(((T)2 * eej) / (T)p.jm);
      iso_dir<T>(vnew, u[2], u[3], vx, vy, vz);
      T q[2], tx, ty, tz;
      FlightDraws<T>::secondary(p.seed, i, step, (uint32_t)k, q);
      iso_dir<T>(vnew, q[0], q[1], tx, ty, tz);
      o.born = 1;
      o.s.x = x;
      o.s.vx = tx;
      o.s.vy = ty;
      o.s.vz = tz;
      o.s.w = in.w;
      o.e_sec = (T)p.kev * (tx * tx + ty * ty + tz * tz);
    } else {
      ++o.n_null;
    }
  }
  if (alive && tau > (T)0) {
    // flight budget exhausted (nu_max dt far above kMaxFlights): rest ballistic
    if (p.push_order == BZ_PUSH_LEAPFROG) {
      vx = vx - acc * tau;
      if (DIM == 1) x = x + ((T)100 * tau) * vx;
    } else {
      if (DIM == 1) x = x + ((T)100 * tau) * vx;
      vx = vx - acc * tau;
    }
    if (DIM == 1 && !((T)0 < x && x < (T)p.L)) alive = 0;
    o.over = 1;
  }
  if (o.born) o.chan = kChanIon;
  o.alive = alive;
  o.p.x = x;
  o.p.vx = vx;
  o.p.vy = vy;
  o.p.vz = vz;
  o.p.w = in.w;
  o.e_post = (T)p.kev * (vx * vx + vy * vy + vz * vz);
  return o;
}

// Node of the deposit: fine bin int((x+dx/2)/(L+dx) * M), then fine -> node
// by integer division by bnf (kernels-new.cu:63 and :8-50).  The quotient
// is formed with the reciprocal of L+dx: it and the correctly rounded
// quotient differ by a few ulp, so away from an integer (more than 1e-12
// relative) both truncate to the same bin; next to one the division runs.
__device__ __forceinline__ int dep_node(double x, const bz_params_t& p) {
  const double xh = x + p.half_dx;
  const double yf = (xh * (1.0 / p.L_dx)) * p.dep_M;
  const double fy = floor(yf);
  const double margin = 1e-12 * yf;
  int b;
  if (yf >= 0.0 && yf - fy > margin && fy + 1.0 - yf > margin) b = (int)fy;
  else b = (int)((xh / p.L_dx) * p.dep_M);
  // k = b / bnf without the integer divide: the FP32 quotient is within one
  // of it for |b| < 2^24 and one step corrects it; for b < 0 either rounding
  // ends at node 0 after the clamp
  const int bnf = p.bnf;
  int k;
  if ((bnf & (bnf - 1)) == 0) {
    k = b >> __builtin_ctz((unsigned)bnf);  // a power of two: floor(b / bnf)
  } else {
    k = (int)((float)b * (1.0f / (float)bnf));
    if (k * bnf > b) --k;
    else if ((k + 1) * bnf <= b) ++k;
  }
  return k < 0 ? 0 : (k >= p.n_nodes ? p.n_nodes - 1 : k);
}

}  // namespace bz
