// CPU oracle of the electron kinetic step -- TEST INFRASTRUCTURE ONLY.
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
// load this library, and only as the checker (or as the timed CPU baseline).
// It is a restatement, in plain sequential C++, of the reference algorithm:
//
//   collision + push + walls  1bin/kernels/electron_kernel_numba.py:39-287
//   deposit                   1bin/kernels/post_processing_1D.py:44-56,
//                             1bin/kernels/reduction/cuda_reduction/kernels-new.cu:56-132
//   compaction                1bin/kernels/post_processing_1D.py:26-38
//   host moment update        1bin/boltzmann.py:643-647
//   heavies                   1bin/kernels/heavies_kernel_1D.py:5-64
//   Poisson                   1bin/kernels/EF_kernel_1D.py:8-44
//   binning                   1bin/kernels/bin_particles.py:7-31 (3bin :39-78, 5bin :86-140)
//   resampling                1bin/boltzmann.py:799-916 (3bin :813-1113, 5bin :843-1334)
//
// plus the definitions this library adds on top of the reference: the Philox
// stream, the tabulated three-channel null-collision model, the canonical
// particle order (secondary right after its parent; segments in key order)
// and the order-independent fixed-point selection of the resampler.  It is
// written independently of the device code and compiled with
// -ffp-contract=off so its arithmetic is plain IEEE in source order.
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

extern "C" {

// ---------------------------------------------------------------------------
// Philox4x32-10, Salmon et al. SC'11: round function
//   (hi0,lo0) = M0*c0, (hi1,lo1) = M1*c2,
//   c' = (hi1^c1^k0, lo1, hi0^c3^k1, lo0); keys bumped by the Weyl constants.
// ---------------------------------------------------------------------------
void orc_philox(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
  uint32_t c[4] = {ctr_in[0], ctr_in[1], ctr_in[2], ctr_in[3]};
  uint32_t k[2] = {key_in[0], key_in[1]};
  for (int round = 0; round < 10; ++round) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t n[4] = {(uint32_t)(p1 >> 32) ^ c[1] ^ k[0], (uint32_t)p1,
                           (uint32_t)(p0 >> 32) ^ c[3] ^ k[1], (uint32_t)p0};
    memcpy(c, n, sizeof(c));
    k[0] += 0x9E3779B9u;
    k[1] += 0xBB67AE85u;
  }
  memcpy(out, c, sizeof(c));
}

static void orc_draw(uint64_t seed, uint32_t purpose, uint32_t index, uint32_t block,
                     uint64_t step, uint32_t out[4]) {
  const uint32_t ctr[4] = {index, block, (uint32_t)step,
                           ((uint32_t)(step >> 32) << 8) | purpose};
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  orc_philox(ctr, key, out);
}

static double orc_u53(uint32_t a, uint32_t b) {
  return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) / 9007199254740992.0;
}
static float orc_u24(uint32_t a) { return (float)(a >> 8) / 16777216.0f; }

// The collision uniforms of step `step` for particles 0..n-1, laid out as the
// reference's R matrix transposed: R[c*stride + i], columns 0,1,2,3,5 used.
// Column 0 is the decision uniform (compared with the collision
// probabilities): one Philox draw of purpose 5 serves the particles 1024
// slots apart of a group, FP32: i = 4096a + 1024b + c takes word b of the
// draw of group 1024a + c; FP64: i = 2048a + 1024b + c takes words 2b, 2b+1
// of group 1024a + c.  Columns 1,2,3,5 are the scattering angles from the
// particle's own draws (purpose 0): FP32 the four words of block 0, FP64
// block 0 for columns 1, 2 and block 1 for columns 3, 5.  The device draws
// the same numbers (the stream is this build's; the reference draws cuRAND).
static void orc_uniforms_at(uint64_t seed, uint64_t step, int64_t n, int prec, void* R,
                            int64_t stride, uint32_t first) {
  for (int64_t j = 0; j < n; ++j) {
    const uint32_t i = first + (uint32_t)j;
    uint32_t d[4], b0[4], b1[4];
    orc_draw(seed, 0, i, 0, step, b0);
    if (prec == 8) {
      orc_draw(seed, 5, ((i >> 11) << 10) | (i & 1023u), 0, step, d);
      orc_draw(seed, 0, i, 1, step, b1);
      const int h = (int)((i >> 10) & 1u);
      double* r = (double*)R;
      r[0 * stride + j] = orc_u53(d[2 * h], d[2 * h + 1]);
      r[1 * stride + j] = orc_u53(b0[0], b0[1]);
      r[2 * stride + j] = orc_u53(b0[2], b0[3]);
      r[3 * stride + j] = orc_u53(b1[0], b1[1]);
      r[4 * stride + j] = 0.0;
      r[5 * stride + j] = orc_u53(b1[2], b1[3]);
      r[6 * stride + j] = 0.0;
    } else {
      orc_draw(seed, 5, ((i >> 12) << 10) | (i & 1023u), 0, step, d);
      float* r = (float*)R;
      r[0 * stride + j] = orc_u24(d[(i >> 10) & 3u]);
      r[1 * stride + j] = orc_u24(b0[0]);
      r[2 * stride + j] = orc_u24(b0[1]);
      r[3 * stride + j] = orc_u24(b0[2]);
      r[4 * stride + j] = 0.0f;
      r[5 * stride + j] = orc_u24(b0[3]);
      r[6 * stride + j] = 0.0f;
    }
  }
}

void orc_collision_uniforms(uint64_t seed, uint64_t step, int64_t n, int prec, void* R,
                            int64_t stride) {
  orc_uniforms_at(seed, step, n, prec, R, stride, 0u);
}

// Gas velocity of particles 0..n-1 at `step` from the Philox stream (purpose
// 1), G[c*stride + i], c = x, y, z: Box-Muller with std `sigma` per
// component, the draw the device makes when the gas is warm.  The reference
// draws it with cp.random.multivariate_normal(0, Eg*I) (1bin/boltzmann.py:589);
// the normal deviates here are r1 cos(2 pi g1), r1 sin(2 pi g1) and
// r2 cos(2 pi g3) with r = sqrt(-2 ln(1 - g)).  sin/cos of pi*v use the
// exactly reduced m_sincospi below (defined after this block).
void orc_gas_velocities(uint64_t seed, uint64_t step, int64_t n, int prec, double sigma, void* G,
                        int64_t stride);

// Resampling draws: uniform m of segment `key` at step `step` (FP64).
double orc_resample_uniform(uint64_t seed, uint64_t step, uint32_t key, uint32_t m) {
  uint32_t b[4];
  orc_draw(seed, 2, m >> 1, key, step, b);
  return (m & 1) ? orc_u53(b[2], b[3]) : orc_u53(b[0], b[1]);
}

// ---------------------------------------------------------------------------
// Parameters (filled by oracle/oracle.py from the physical inputs)
// ---------------------------------------------------------------------------
typedef struct {
  double dt, xfac, afac, kev, jm, fel, L, dx, half_dx, L_dx, dep_M, nn, e_ion, e_exc, pc, e_field0,
      nu_max;
  int32_t dim, xs_mode, n_cells, n_nodes, bnf, nthreads;
  uint64_t seed;  // Philox key of the free-flight draws (xs_mode 2)
  uint64_t step;  // step number keying those draws
  int32_t push_order;   // 0: x then v (reference), 1: v then x (leapfrog)
  int32_t births_last;  // 1: secondaries after all survivors (reference layout),
                        // 0: each secondary right after its parent
  int32_t ref_flags;    // 1: dead ionizer's flag at node 0, 2: (e*w)/w binning energy
  int32_t pad;
} orc_params;

typedef struct {
  const void* energy;
  const void* sig;
  const void* slope;
  int32_t n;
  int32_t pad;
} orc_table;

}  // extern "C"

namespace {

// Reference argon fits (electron_kernel_numba.py:67-171).
const double el_i[11] = {0.000780396, 0.006098667, 0.0476269, 0.1015574, 0.1943527, 0.2995781,
                         1.097156,    2.339524,    4.988691,  10.63765,  22.68322};
const double el_a[12] = {-0.02704763, -0.08847237, -0.37608274, -1.4874467,
                         -0.9870356,  14.28063581, -2.12069169, -0.3585636,
                         1.25262128,  -2.28332905, -1.47508661, -0.11090525};
const double el_b[12] = {-0.23720051, -0.6084786,  -1.8167778,  -4.82420619,
                         -4.2206026,  18.92275458, 0.7555105,   0.79246666,
                         -0.40002029, 4.89566076,  2.82263476,  -0.77506081};
const double el_c[12] = {-19.67900951, -20.24111992, -21.51414308, -23.55745478,
                         -23.44715988, -14.70556113, -19.71943671, -19.71260274,
                         -19.48909003, -21.46789173, -20.1928456,  -17.81752307};
const double io_i[7] = {15.77742, 15.81964, 15.96196, 16.44162, 18.05837, 23.50771, 41.87498};
const double io_a[8] = {-2149239.99337822, -227229.83440255, -20173.68563579, -1838.90888364,
                        -186.98923927,     -24.48843901,     -4.33101714,     -0.68345667};
const double io_b[8] = {5150451.29014456, 545171.14839945, 48599.13254342, 4491.58838596,
                        477.04784782,     69.43965607,     14.47124519,    2.6053539};
const double io_c[8] = {-3085665.64874862, -327016.69508374, -29290.96100603, -2763.82866252,
                        -324.75732578,     -69.145292411,    -31.66224458,    -21.99823511};

inline double m_log10(double v) { return log10(v); }
inline float m_log10(float v) { return log10f(v); }
inline double m_p10(double v) { return pow(10.0, v); }
inline float m_p10(float v) { return powf(10.0f, v); }
// e^v - 1 and e^v from plain IEEE operations (no libm), so that the
// collision probabilities of the direct and fit modes are the same numbers
// the device computes: v = k ln2 + r, k = floor(v / ln2 + 1/2), r = (v -
// k ln2_hi) - k ln2_lo (Cody-Waite, k ln2_hi exact); e^r - 1 = r + r^2 P(r),
// P the Taylor polynomial 1/2! + r/3! + ... in Horner form (to 1/15! in
// FP64, 1/8! in FP32, |r| <= ln2/2); e^v - 1 = 2^k (e^r - 1) + (2^k - 1);
// v clamped to [-80, 80].
// Pinned against libm's expm1 / exp (within 2 ulp) in the oracle tests.
static double orc_expm1_red(double v, int* k) {
  v = v < -80.0 ? -80.0 : (v > 80.0 ? 80.0 : v);
  const double kf = floor(v * 1.4426950408889634 + 0.5);
  const double r = (v - kf * 6.93147180369123816490e-01) - kf * 1.90821492927058770002e-10;
  static const double c[13] = {1.1470745597729725e-11, 1.6059043836821613e-10,
                               2.08767569878681e-09,   2.505210838544172e-08,
                               2.755731922398589e-07,  2.755731922398589e-06,
                               2.48015873015873e-05,   0.0001984126984126984,
                               0.001388888888888889,   0.008333333333333333,
                               0.041666666666666664,   0.16666666666666666,
                               0.5};
  double q = 7.647163731819816e-13;
  for (int j = 0; j < 13; ++j) q = q * r + c[j];
  *k = (int)kf;
  return r + (r * r) * q;
}
static float orc_expm1_red(float v, int* k) {
  v = v < -80.0f ? -80.0f : (v > 80.0f ? 80.0f : v);
  const float kf = floorf(v * 1.44269504f + 0.5f);
  const float r = (v - kf * 0.693359375f) - kf * -2.12194440e-4f;
  static const float c[6] = {1.98412698e-04f, 1.38888889e-03f, 8.33333333e-03f,
                             4.16666667e-02f, 1.66666667e-01f, 0.5f};
  float q = 2.48015873e-05f;
  for (int j = 0; j < 6; ++j) q = q * r + c[j];
  *k = (int)kf;
  return r + (r * r) * q;
}
template <typename T>
static T orc_expm1(T v) {
  int k;
  const T e = orc_expm1_red(v, &k);
  if (k == 0) return e;
  const T s = std::ldexp((T)1, k);
  return s * e + (s - (T)1);
}
template <typename T>
static T orc_exp(T v) {
  int k;
  const T e = orc_expm1_red(v, &k);
  return std::ldexp((T)1 + e, k);
}
inline double m_exp(double v) { return orc_exp(v); }
inline float m_exp(float v) { return orc_exp(v); }
inline double m_expm1(double v) { return orc_expm1(v); }
inline float m_expm1(float v) { return orc_expm1(v); }
inline double m_sqrt(double v) { return sqrt(v); }
inline float m_sqrt(float v) { return sqrtf(v); }
// sin(pi v), cos(pi v) with an exact argument reduction: n = rint(2v) and
// r = v - n/2 in [-1/4, 1/4] are exact, pi*r is one rounding, and the
// quadrant n mod 4 maps sin/cos(pi r) to the result.  FP32 callers round the
// FP64 value once.
inline void m_sincospi(double v, double* s, double* c) {
  const double n = nearbyint(2.0 * v);
  const double r = v - 0.5 * n;
  const double sr = sin(3.141592653589793 * r), cr = cos(3.141592653589793 * r);
  switch (((long long)n) & 3) {
    case 0: *s = sr; *c = cr; break;
    case 1: *s = cr; *c = -sr; break;
    case 2: *s = -sr; *c = -cr; break;
    default: *s = -cr; *c = sr; break;
  }
}
inline void m_sincospi(float v, float* s, float* c) {
  double sd, cd;
  m_sincospi((double)v, &sd, &cd);
  *s = (float)sd;
  *c = (float)cd;
}

}  // namespace

extern "C" void orc_gas_velocities(uint64_t seed, uint64_t step, int64_t n, int prec,
                                   double sigma, void* G, int64_t stride) {
  for (int64_t j = 0; j < n; ++j) {
    const uint32_t i = (uint32_t)j;
    uint32_t a[4], b[4];
    orc_draw(seed, 1, i, 0, step, a);
    if (prec == 8) {
      orc_draw(seed, 1, i, 1, step, b);
      const double g0 = orc_u53(a[0], a[1]), g1 = orc_u53(a[2], a[3]);
      const double g2 = orc_u53(b[0], b[1]), g3 = orc_u53(b[2], b[3]);
      const double r1 = sqrt(-2.0 * log(1.0 - g0)), r2 = sqrt(-2.0 * log(1.0 - g2));
      double s1, c1, s3, c3;
      m_sincospi(2.0 * g1, &s1, &c1);
      m_sincospi(2.0 * g3, &s3, &c3);
      double* g = (double*)G;
      g[j] = sigma * (r1 * c1);
      g[stride + j] = sigma * (r1 * s1);
      g[2 * stride + j] = sigma * (r2 * c3);
    } else {
      const float g0 = orc_u24(a[0]), g1 = orc_u24(a[1]), g2 = orc_u24(a[2]), g3 = orc_u24(a[3]);
      const float r1 = sqrtf(-2.0f * logf(1.0f - g0)), r2 = sqrtf(-2.0f * logf(1.0f - g2));
      float s1, c1, s3, c3;
      m_sincospi(2.0f * g1, &s1, &c1);
      m_sincospi(2.0f * g3, &s3, &c3);
      const float sg = (float)sigma;
      float* g = (float*)G;
      g[j] = sg * (r1 * c1);
      g[stride + j] = sg * (r1 * s1);
      g[2 * stride + j] = sg * (r2 * c3);
    }
  }
}

namespace {

template <typename T>
T poly(double a, double b, double c, T L) {
  return ((T)a * L + (T)b) * L + (T)c;
}

// The reference sums indicator * polynomial over the intervals.
template <typename T>
void fit_cross_sections(T e, const orc_params& P, T* s_el, T* s_ion) {
  const T L = m_log10(e);
  T lg0 = (T)0;
  for (int k = 0; k < 12; ++k) {
    const bool lo = (k == 0) || ((T)el_i[k - 1] <= e);
    const bool hi = (k == 11) || (e < (T)el_i[k]);
    if (lo && hi) lg0 = lg0 + poly<T>(el_a[k], el_b[k], el_c[k], L);
  }
  T lg2 = (T)0;
  for (int k = 0; k < 8; ++k) {
    const bool lo = (k == 0) ? ((T)P.e_ion < e) : ((T)io_i[k - 1] <= e);
    const bool hi = (k == 7) || (e < (T)io_i[k]);
    if (lo && hi) lg2 = lg2 + poly<T>(io_a[k], io_b[k], io_c[k], L);
  }
  T s0 = m_p10(lg0);
  T s2 = m_p10(lg2);
  if (e < (T)P.e_ion) s2 = (T)0;
  *s_el = s0 * (T)P.nn;
  *s_ion = s2 * (T)P.nn;
}

}  // namespace

// the oracle's e^v - 1 and e^v (the collision probabilities' exponential)
extern "C" void orc_expm1_f64(int64_t n, const double* v, double* em1, double* ex) {
  for (int64_t i = 0; i < n; ++i) {
    em1[i] = orc_expm1(v[i]);
    ex[i] = orc_exp(v[i]);
  }
}
extern "C" void orc_expm1_f32(int64_t n, const float* v, float* em1, float* ex) {
  for (int64_t i = 0; i < n; ++i) {
    em1[i] = orc_expm1(v[i]);
    ex[i] = orc_exp(v[i]);
  }
}

// sigma_el * nn and sigma_ion * nn of the reference fits at energies e[0..n)
extern "C" void orc_fit_sigma_f64(const orc_params* P, int64_t n, const double* e, double* s_el,
                                  double* s_ion) {
  for (int64_t i = 0; i < n; ++i) fit_cross_sections<double>(e[i], *P, &s_el[i], &s_ion[i]);
}
extern "C" void orc_fit_sigma_f32(const orc_params* P, int64_t n, const float* e, float* s_el,
                                  float* s_ion) {
  for (int64_t i = 0; i < n; ++i) fit_cross_sections<float>(e[i], *P, &s_el[i], &s_ion[i]);
}

namespace {

template <typename T>
void table_cross_sections(T e, const orc_params& P, const orc_table& tab, T s[3]) {
  const T* E = (const T*)tab.energy;
  const T* S = (const T*)tab.sig;
  const T* D = (const T*)tab.slope;
  const int n = tab.n;
  T ec = e;
  if (ec < E[0]) ec = E[0];
  if (ec > E[n - 1]) ec = E[n - 1];
  // largest j in [0, n-2] with E[j] <= ec
  int lo = 0, hi = n - 2;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (E[mid] <= ec) lo = mid;
    else hi = mid - 1;
  }
  const T d = ec - E[lo];
  for (int c = 0; c < 3; ++c) s[c] = S[c * n + lo] + D[c * n + lo] * d;
  if (e < (T)P.e_exc) s[1] = (T)0;
  if (e < (T)P.e_ion) s[2] = (T)0;
}

template <typename T>
void scatter(T speed, T ra, T rb, T* vx, T* vy, T* vz) {
  const T cos_chi = (T)2 * ra - (T)1;
  T q = (T)1 - cos_chi * cos_chi;
  if (q < (T)0) q = (T)0;
  const T sin_chi = m_sqrt(q);
  T sphi, cphi;
  m_sincospi((T)2 * rb, &sphi, &cphi);  // phi = 2 pi rb
  *vx = speed * sin_chi * cphi;
  *vy = speed * sin_chi * sphi;
  *vz = speed * cos_chi;
}

template <typename T>
struct Rec {
  T x, vx, vy, vz, w;
};

int node_of(double x, const orc_params& P) {
  if (P.dim == 0) return 0;
  const double tx = (x + P.half_dx) / P.L_dx;
  const int fine = (int)(tx * P.dep_M);
  int k = fine / P.bnf;
  if (k < 0) k = 0;
  if (k >= P.n_nodes) k = P.n_nodes - 1;
  return k;
}

// One particle; returns number of output records (0, 1 or 2) in out[].
template <typename T>
int one_particle(int64_t i, const orc_params& P, const T* x, const T* vx, const T* vy,
                 const T* vz, const T* w, const T* E, const T* R, int64_t rs, const T* G,
                 const orc_table* tab, Rec<T> out[2], T e_out[2], int* chan, int* ionized) {
  *chan = 0;
  *ionized = 0;
  const T wi = w[i];
  if (!(wi > (T)0)) return 0;
  T gx = 0, gy = 0, gz = 0;
  if (G) {
    gx = G[i];
    gy = G[rs + i];
    gz = G[2 * rs + i];
  }
  T ux = vx[i] - gx, uy = vy[i] - gy, uz = vz[i] - gz;
  T e = ux * ux + uy * uy + uz * uz;
  const T v_mag = m_sqrt(e);
  e = (T)P.kev * e;
  T R0, R1, R2, R3, R5;
  if (R) {
    R0 = R[i];
    R1 = R[rs + i];
    R2 = R[2 * rs + i];
    R3 = R[3 * rs + i];
    R5 = R[5 * rs + i];
  } else {
    // no stream given: the Philox uniforms of (seed, step, i), drawn here
    T u[7];
    orc_uniforms_at(P.seed, P.step, 1, sizeof(T) == 8 ? 8 : 4, u, 1, (uint32_t)i);
    R0 = u[0];
    R1 = u[1];
    R2 = u[2];
    R3 = u[3];
    R5 = u[5];
  }
  int c = 0;
  if (P.xs_mode == 0) {
    T s0, s2;
    fit_cross_sections<T>(e, P, &s0, &s2);
    const T tot = s0 + s2;
    const T Pc = (T)1 - m_exp(-(T)P.dt * v_mag * tot);
    const T pcst = Pc * ((T)1 / tot);
    const T g0 = s0 * pcst;
    const T g2 = g0 + s2 * pcst;
    if (R0 < g0) c = 1;
    else if (g0 < R0 && R0 < g2) c = 3;
  } else if (P.xs_mode == 3) {
    // direct: P = 1 - exp(-nu_t dt), thresholds nu_j * P / nu_t
    // (electron_kernel_numba.py:186-212 with the tabulated sigma)
    T s[3];
    table_cross_sections<T>(e, P, *tab, s);
    T g0 = s[0] * v_mag;
    T g1 = g0 + s[1] * v_mag;
    T g2 = g1 + s[2] * v_mag;
    if (g2 > (T)0) {
      const T Pc = -m_expm1(-(T)P.dt * g2);
      const T pcst = Pc * ((T)1 / g2);
      g0 = g0 * pcst;
      g1 = g1 * pcst;
      g2 = g2 * pcst;
    } else {
      g0 = g1 = g2 = (T)0;
    }
    if (R0 < g0) c = 1;
    else if (R0 < g1) c = 2;
    else if (R0 < g2) c = 3;
  } else {
    T s[3];
    table_cross_sections<T>(e, P, *tab, s);
    const T pc = (T)P.pc;
    const T g0 = (s[0] * v_mag) * pc;
    const T g1 = g0 + (s[1] * v_mag) * pc;
    const T g2 = g1 + (s[2] * v_mag) * pc;
    if (R0 < g0) c = 1;
    else if (R0 < g1) c = 2;
    else if (R0 < g2) c = 3;
  }
  *chan = c;
  T nvx = vx[i], nvy = vy[i], nvz = vz[i];
  int born = 0;
  Rec<T> sec = {0, 0, 0, 0, 0};
  T e_sec = 0;
  if (c == 1) {
    scatter<T>(v_mag * (T)P.fel, R1, R2, &nvx, &nvy, &nvz);
  } else if (c == 2) {
    const T e_new = e - (T)P.e_exc;
    scatter<T>(m_sqrt(((T)2 * e_new) / (T)P.jm), R1, R2, &nvx, &nvy, &nvz);
  } else if (c == 3) {
    T e_ej = (T)0.5 * (e - (T)P.e_ion);
    if (e_ej < (T)0) e_ej = -e_ej;
    const T sp = m_sqrt(((T)2 * e_ej) / (T)P.jm);
    scatter<T>(sp, R1, R2, &nvx, &nvy, &nvz);
    T tx, ty, tz;
    scatter<T>(sp, R3, R5, &tx, &ty, &tz);
    born = 1;
    sec.x = x[i];
    sec.vx = tx;
    sec.vy = ty;
    sec.vz = tz;
    sec.w = wi;
    e_sec = (T)P.kev * (tx * tx + ty * ty + tz * tz);
  }
  T nx = x[i];
  T ef;
  if (P.dim == 1) {
    int cell = (int)((double)x[i] / P.dx);
    if (cell < 0) cell = 0;
    if (cell >= P.n_cells) cell = P.n_cells - 1;
    ef = E[cell];
  } else {
    ef = E ? E[0] : (T)P.e_field0;
  }
  if (P.push_order == 1) {
    nvx = nvx - (T)P.afac * ((T)100 * ef);
    if (P.dim == 1) nx = nx + (T)P.xfac * nvx;
  } else {
    // electron_kernel_numba.py:257-264: X-ADV with the post-collision v, then V-ADV
    if (P.dim == 1) nx = nx + (T)P.xfac * nvx;
    nvx = nvx - (T)P.afac * ((T)100 * ef);
  }
  const bool alive = (P.dim == 0) || ((T)0 < nx && nx < (T)P.L);
  int k = 0;
  if (alive) {
    out[k] = Rec<T>{nx, nvx, nvy, nvz, wi};
    e_out[k] = (T)P.kev * (nvx * nvx + nvy * nvy + nvz * nvz);
    ++k;
    *ionized = (c == 3);
  }
  if (born) {
    out[k] = sec;
    e_out[k] = e_sec;
    ++k;
  }
  return k;
}

inline float m_log1p(float v) { return log1pf(v); }
inline double m_log1p(double v) { return std::log1p(v); }

// Free-flight draws: event k of particle i uses Philox blocks 3k, 3k+1 (FP64)
// or 2k (FP32) of purpose 4; the secondary's angles come from block 3k+2 /
// 2k+1.
void flight_event(uint64_t seed, uint32_t i, uint64_t step, uint32_t k, double u[4]) {
  uint32_t r[4], q[4];
  orc_draw(seed, 4, i, 3 * k, step, r);
  orc_draw(seed, 4, i, 3 * k + 1, step, q);
  u[0] = orc_u53(r[0], r[1]);
  u[1] = orc_u53(r[2], r[3]);
  u[2] = orc_u53(q[0], q[1]);
  u[3] = orc_u53(q[2], q[3]);
}
void flight_event(uint64_t seed, uint32_t i, uint64_t step, uint32_t k, float u[4]) {
  uint32_t r[4];
  orc_draw(seed, 4, i, 2 * k, step, r);
  for (int j = 0; j < 4; ++j) u[j] = orc_u24(r[j]);
}
void flight_secondary(uint64_t seed, uint32_t i, uint64_t step, uint32_t k, double u[2]) {
  uint32_t r[4];
  orc_draw(seed, 4, i, 3 * k + 2, step, r);
  u[0] = orc_u53(r[0], r[1]);
  u[1] = orc_u53(r[2], r[3]);
}
void flight_secondary(uint64_t seed, uint32_t i, uint64_t step, uint32_t k, float u[2]) {
  uint32_t r[4];
  orc_draw(seed, 4, i, 2 * k + 1, step, r);
  u[0] = orc_u24(r[0]);
  u[1] = orc_u24(r[1]);
}

struct Events {
  int el, exc, ion, null_, alive;
};

// Exact free-flight null collision over one step (any number of events):
// flights ~ exp(nu_max); at a flight's end channel j with probability
// nu_j(e)/nu_max; E of the starting cell for the whole step; x moves with
// the flight's initial velocity, then the acceleration applies; absorbed
// when a flight ends outside the gap; one secondary per step at most
// (further ionizing events are null), born at the collision point and not
// moved in its birth step.
template <typename T>
int one_particle_flight(int64_t i, uint64_t step, const orc_params& P, const T* x, const T* vx,
                        const T* vy, const T* vz, const T* w, const T* E, const orc_table* tab,
                        Rec<T> out[2], T e_out[2], Events* ev) {
  *ev = Events{0, 0, 0, 0, 0};
  const T wi = w[i];
  if (!(wi > (T)0)) return 0;
  T ef;
  if (P.dim == 1) {
    int cell = (int)((double)x[i] / P.dx);
    if (cell < 0) cell = 0;
    if (cell >= P.n_cells) cell = P.n_cells - 1;
    ef = E[cell];
  } else {
    ef = E ? E[0] : (T)P.e_field0;
  }
  const T acc = (T)(P.afac / P.dt) * ((T)100 * ef);
  const T inv_nu = (T)(1.0 / P.nu_max);
  T px = x[i], pvx = vx[i], pvy = vy[i], pvz = vz[i];
  T tau = (T)P.dt;
  int alive = 1, born = 0;
  Rec<T> sec = {0, 0, 0, 0, 0};
  T e_sec = 0;
  for (int k = 0; k < 256; ++k) {
    T u[4];
    flight_event(P.seed, (uint32_t)i, step, (uint32_t)k, u);
    const T tf = -m_log1p(-u[0]) * inv_nu;
    const bool hit = tf < tau;
    const T h = hit ? tf : tau;
    if (P.push_order == 1) {
      pvx = pvx - acc * h;
      if (P.dim == 1) px = px + ((T)100 * h) * pvx;
    } else {
      if (P.dim == 1) px = px + ((T)100 * h) * pvx;
      pvx = pvx - acc * h;
    }
    if (P.dim == 1 && !((T)0 < px && px < (T)P.L)) {
      alive = 0;
      break;
    }
    if (!hit) {
      tau = (T)0;
      break;
    }
    tau = tau - tf;
    T s2 = pvx * pvx + pvy * pvy + pvz * pvz;
    const T v_mag = m_sqrt(s2);
    const T e = (T)P.kev * s2;
    T sg[3];
    table_cross_sections<T>(e, P, *tab, sg);
    const T r0 = (sg[0] * v_mag) * inv_nu;
    const T r1 = r0 + (sg[1] * v_mag) * inv_nu;
    const T r2 = r1 + (sg[2] * v_mag) * inv_nu;
    const T R = u[1];
    if (R < r0) {
      ++ev->el;
      scatter<T>(v_mag * (T)P.fel, u[2], u[3], &pvx, &pvy, &pvz);
    } else if (R < r1) {
      ++ev->exc;
      const T e_new = e - (T)P.e_exc;
      scatter<T>(m_sqrt(((T)2 * e_new) / (T)P.jm), u[2], u[3], &pvx, &pvy, &pvz);
    } else if (R < r2 && !born) {
      T e_ej = (T)0.5 * (e - (T)P.e_ion);
      if (e_ej < (T)0) e_ej = -e_ej;
      const T sp = m_sqrt(((T)2 * e_ej) / (T)P.jm);
      scatter<T>(sp, u[2], u[3], &pvx, &pvy, &pvz);
      T q[2], tx, ty, tz;
      flight_secondary(P.seed, (uint32_t)i, step, (uint32_t)k, q);
      scatter<T>(sp, q[0], q[1], &tx, &ty, &tz);
      born = 1;
      ev->ion = 1;
      sec = Rec<T>{px, tx, ty, tz, wi};
      e_sec = (T)P.kev * (tx * tx + ty * ty + tz * tz);
    } else {
      ++ev->null_;
    }
  }
  if (alive && tau > (T)0) {
    if (P.push_order == 1) {
      pvx = pvx - acc * tau;
      if (P.dim == 1) px = px + ((T)100 * tau) * pvx;
    } else {
      if (P.dim == 1) px = px + ((T)100 * tau) * pvx;
      pvx = pvx - acc * tau;
    }
    if (P.dim == 1 && !((T)0 < px && px < (T)P.L)) alive = 0;
  }
  ev->alive = alive;
  int k = 0;
  if (alive) {
    out[k] = Rec<T>{px, pvx, pvy, pvz, wi};
    e_out[k] = (T)P.kev * (pvx * pvx + pvy * pvy + pvz * pvz);
    ++k;
  }
  if (born) {
    out[k] = sec;
    e_out[k] = e_sec;
    ++k;
  }
  return k;
}

template <typename T>
int64_t push_step(const orc_params* P, int64_t n, const T* x, const T* vx, const T* vy,
                  const T* vz, const T* w, const T* E, const T* R, int64_t rs, const T* G,
                  const orc_table* tab, T* ox, T* ovx, T* ovy, T* ovz, T* ow, int64_t cap,
                  double* mom, uint64_t* cnt, double* coll) {
  const int nn = P->dim == 1 ? P->n_nodes : 1;
  int nt = P->nthreads > 0 ? P->nthreads : 1;
#ifndef _OPENMP
  nt = 1;
#endif
  // fixed chunking: chunk k handles [k*n/nt, (k+1)*n/nt)
  std::vector<int64_t> chunk_cnt(nt + 1, 0);
  std::vector<std::vector<double>> part(nt, std::vector<double>(3 * nn, 0.0));
  // collision counters (coll != NULL): per node, sum w of the elastic and of
  // the excitation collisions and the number of ionizations, at the
  // primary's position after the push, node 0 when it left the gap (the
  // reference's collct rows, electron_kernel_numba.py:199,232, deposited at
  // x = 1e-8 for an absorbed particle, post_processing_1D.py:56)
  std::vector<std::vector<double>> cpart(coll ? nt : 0, std::vector<double>(3 * nn, 0.0));
  std::vector<std::vector<uint64_t>> ccnt(nt, std::vector<uint64_t>(8, 0));
  std::vector<std::vector<Rec<T>>> buf(nt), sec(nt);
  const bool births_last = P->births_last != 0;
#pragma omp parallel for num_threads(nt) schedule(static, 1)
  for (int k = 0; k < nt; ++k) {
    const int64_t lo = n * k / nt, hi = n * (k + 1) / nt;
    std::vector<Rec<T>>& b = buf[k];
    b.reserve((size_t)(hi - lo) + 16);
    double* pm = part[k].data();
    uint64_t* cc = ccnt[k].data();
    for (int64_t i = lo; i < hi; ++i) {
      Rec<T> o[2];
      T eo[2];
      int chan, ion, m, born;
      if (P->xs_mode == 2) {
        Events ev;
        m = one_particle_flight<T>(i, P->step, *P, x, vx, vy, vz, w, E, tab, o, eo, &ev);
        ion = ev.alive && ev.ion;
        born = ev.ion;
        if (w[i] > (T)0) {
          cc[2] += ev.el;
          cc[3] += ev.exc;
          cc[4] += ev.ion;
          cc[0] += ev.ion;
          cc[5] += ev.null_;
          if (!ev.alive) cc[1] += 1;
          if (coll && (ev.el || ev.exc || ev.ion) && P->dim == 1) {
            const int nd = ev.alive ? node_of((double)o[0].x, *P) : 0;
            double* cp = cpart[k].data();
            cp[nd] += (double)w[i] * (double)ev.el;
            cp[nn + nd] += (double)w[i] * (double)ev.exc;
            cp[2 * nn + nd] += ev.ion ? 1.0 : 0.0;
          }
        }
      } else {
        m = one_particle<T>(i, *P, x, vx, vy, vz, w, E, R, rs, G, tab, o, eo, &chan, &ion);
        born = chan == 3;
        if (w[i] > (T)0) {
          cc[chan == 0 ? 5 : (chan == 1 ? 2 : (chan == 2 ? 3 : 4))] += 1;
          if (chan == 3) cc[0] += 1;
          const bool alive_primary = (m == 2) || (m == 1 && chan != 3);
          if (!alive_primary) cc[1] += 1;
          // the reference reduces the rows of absorbed particles too: an
          // ionizer that left keeps its unscaled flag 1 at x = 1e-8, node 0
          // (electron_kernel_numba.py:231,267-270; post_processing_1D.py:48-55)
          if (!alive_primary && chan == 3 && (P->ref_flags & 1) && P->dim == 1) pm[nn] += 1.0;
          if (coll && chan != 0 && P->dim == 1) {
            const int nd = alive_primary ? node_of((double)o[0].x, *P) : 0;
            double* cp = cpart[k].data();
            if (chan == 1) cp[nd] += (double)w[i];
            else if (chan == 2) cp[nn + nd] += (double)w[i];
            else cp[2 * nn + nd] += 1.0;
          }
        }
      }
      // the secondary, when one was born, is the last record
      for (int q = 0; q < m; ++q) {
        const int nd = node_of((double)o[q].x, *P);
        const double wd = (double)o[q].w;
        pm[nd] += wd;
        if (q == 0 && ion) pm[nn + nd] += wd;
        pm[2 * nn + nd] += wd * (double)eo[q];
        const bool is_sec = births_last && born && q == m - 1;
        (is_sec ? sec[k] : b).push_back(o[q]);
      }
    }
    chunk_cnt[k + 1] = (int64_t)b.size();
  }
  for (int k = 0; k < nt; ++k) chunk_cnt[k + 1] += chunk_cnt[k];
  const int64_t n_surv = chunk_cnt[nt];
  std::vector<int64_t> sec_cnt(nt + 1, 0);
  for (int k = 0; k < nt; ++k) sec_cnt[k + 1] = sec_cnt[k] + (int64_t)sec[k].size();
  const int64_t total = n_surv + sec_cnt[nt];
  uint64_t dropped = 0;
  auto put = [&](int64_t pos, const Rec<T>& r) {
    if (pos < cap) {
      ox[pos] = r.x;
      ovx[pos] = r.vx;
      ovy[pos] = r.vy;
      ovz[pos] = r.vz;
      ow[pos] = r.w;
      return 0;
    }
    return 1;
  };
#pragma omp parallel for num_threads(nt) schedule(static, 1) reduction(+ : dropped)
  for (int k = 0; k < nt; ++k) {
    int64_t pos = chunk_cnt[k];
    for (const Rec<T>& r : buf[k]) dropped += put(pos++, r);
    pos = n_surv + sec_cnt[k];
    for (const Rec<T>& r : sec[k]) dropped += put(pos++, r);
  }
  for (int j = 0; j < 3 * nn; ++j) {
    double s = 0.0;
    for (int k = 0; k < nt; ++k) s += part[k][j];
    mom[j] = s;
  }
  if (coll)
    for (int j = 0; j < 3 * nn; ++j) {
      double s = 0.0;
      for (int k = 0; k < nt; ++k) s += cpart[k][j];
      coll[j] = s;
    }
  for (int q = 0; q < 8; ++q) {
    uint64_t s = 0;
    for (int k = 0; k < nt; ++k) s += ccnt[k][q];
    cnt[q] = s;
  }
  cnt[6] = dropped;
  return total < cap ? total : cap;
}

}  // namespace

extern "C" {

int64_t orc_push_step_f64(const orc_params* P, int64_t n, const double* x, const double* vx,
                          const double* vy, const double* vz, const double* w, const double* E,
                          const double* R, int64_t rs, const double* G, const orc_table* tab,
                          double* ox, double* ovx, double* ovy, double* ovz, double* ow,
                          int64_t cap, double* mom, uint64_t* cnt, double* coll) {
  return push_step<double>(P, n, x, vx, vy, vz, w, E, R, rs, G, tab, ox, ovx, ovy, ovz, ow, cap,
                           mom, cnt, coll);
}

int64_t orc_push_step_f32(const orc_params* P, int64_t n, const float* x, const float* vx,
                          const float* vy, const float* vz, const float* w, const float* E,
                          const float* R, int64_t rs, const float* G, const orc_table* tab,
                          float* ox, float* ovx, float* ovy, float* ovz, float* ow, int64_t cap,
                          double* mom, uint64_t* cnt, double* coll) {
  return push_step<float>(P, n, x, vx, vy, vz, w, E, R, rs, G, tab, ox, ovx, ovy, ovz, ow, cap,
                          mom, cnt, coll);
}

// Deposit of a state (no push): ne, ni_src (=0), nrg per node.
void orc_deposit_f64(const orc_params* P, int64_t n, const double* x, const double* vx,
                     const double* vy, const double* vz, const double* w, double* mom) {
  const int nn = P->dim == 1 ? P->n_nodes : 1;
  for (int j = 0; j < 3 * nn; ++j) mom[j] = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    if (!(w[i] > 0.0)) continue;
    const double e = P->kev * (vx[i] * vx[i] + vy[i] * vy[i] + vz[i] * vz[i]);
    const int nd = node_of(x[i], *P);
    mom[nd] += w[i];
    mom[2 * nn + nd] += w[i] * e;
  }
}

void orc_deposit_f32(const orc_params* P, int64_t n, const float* x, const float* vx,
                     const float* vy, const float* vz, const float* w, double* mom) {
  const int nn = P->dim == 1 ? P->n_nodes : 1;
  for (int j = 0; j < 3 * nn; ++j) mom[j] = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    if (!(w[i] > 0.0f)) continue;
    const float e = (float)P->kev * (vx[i] * vx[i] + vy[i] * vy[i] + vz[i] * vz[i]);
    const int nd = node_of((double)x[i], *P);
    mom[nd] += (double)w[i];
    mom[2 * nn + nd] += (double)w[i] * (double)e;
  }
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Binning and ES-WPNC resampling
// ---------------------------------------------------------------------------
extern "C" {
typedef struct {
  double split_below, merge_at, merge_keep, trigger_at, min_wt, split_to;
  int32_t ref_trigger, max_ppc, ref_rescale, pad;
} orc_policy;
}

namespace {

template <typename T>
int bin_key(const orc_params& P, int nb, const double* thr, T x, T vx, T vy, T vz, T w,
            int n_keys) {
  if (!(w > (T)0)) return n_keys;
  int cell = 0;
  if (P.dim == 1) {
    // curr_xbins = int((x + dx/2)/dx)  (boltzmann.py:658)
    cell = (int)(((double)x + 0.5 * P.dx) / P.dx);
    if (cell < 0) cell = 0;
    if (cell >= P.n_nodes) cell = P.n_nodes - 1;
  }
  T e = (T)P.kev * (vx * vx + vy * vy + vz * vz);
  // the reference bins tosum[:,3] after it was multiplied and divided by w
  // (post_processing_1D.py:53,62)
  if (P.ref_flags & 2) e = (e * w) / w;
  // bin_particles.py: LE < thr[0] <= ... ; bins ascend with energy
  int b = 0;
  for (int k = 0; k < nb - 1; ++k)
    if (e >= (T)thr[k]) b = k + 1;
  return cell * nb + b;
}

int ceil_log2(int64_t c) {
  int b = 0;
  while (((int64_t)1 << b) < c) ++b;
  return b;
}

// first j with cdf[j] >= thr (left) or > thr (right), clamped to c-1
int search(const std::vector<int64_t>& cdf, int64_t thr, bool right) {
  auto it = right ? std::upper_bound(cdf.begin(), cdf.end(), thr)
                  : std::lower_bound(cdf.begin(), cdf.end(), thr);
  int j = (int)(it - cdf.begin());
  return j < (int)cdf.size() ? j : (int)cdf.size() - 1;
}

template <typename T>
void bin_sort(const orc_params& P, int nb, const double* thr, int64_t n, const T* x, const T* vx,
              const T* vy, const T* vz, const T* w, int32_t* counts, int32_t* offsets,
              uint32_t* perm) {
  const int nn = P.dim == 1 ? P.n_nodes : 1;
  const int n_keys = nn * nb;
  std::vector<int> key(n);
  for (int k = 0; k <= n_keys; ++k) counts[k] = 0;
  for (int64_t i = 0; i < n; ++i) {
    key[i] = bin_key<T>(P, nb, thr, x[i], vx[i], vy[i], vz[i], w[i], n_keys);
    counts[key[i]] += 1;
  }
  offsets[0] = 0;
  for (int k = 0; k <= n_keys; ++k) offsets[k + 1] = offsets[k] + counts[k];
  std::vector<int32_t> fill(offsets, offsets + n_keys + 1);
  for (int64_t i = 0; i < n; ++i) perm[fill[key[i]]++] = (uint32_t)i;  // stable
}

// Resampling uniform m of segment k: the caller's stream U[k*u_stride + m]
// when given, the Philox stream otherwise.
inline double res_u(const double* U, int64_t u_stride, uint64_t seed, uint64_t step, int k,
                    int m) {
  return U ? U[(int64_t)k * u_stride + m]
           : orc_resample_uniform(seed, step, (uint32_t)k, (uint32_t)m);
}

// What one resampling pass does to one (cell, bin) segment.
struct SegAction {
  int kind = 0;                    // 0 untouched, 1 split, 2 merge
  int ce = 0;                      // merge: members taking part (max_ppc clip)
  std::vector<int> copies;         // split: clones of each member
  std::vector<uint32_t> sel;       // merge: the selected slots, in draw order
  std::vector<double> nv[3];       // merge: velocities of the merged particles
  double wbar = 0;                 // merge: their weight
};

// Binning, trigger and the split / merge decisions and draws of every
// segment (boltzmann.py:799-916; 3bin:813-1113; 5bin:843-1334).  Members of
// a segment are its slots in ascending order.  Fills counts/offsets/perm.
template <typename T>
bool resample_plan(const orc_params& P, const orc_policy& pol, int nb, const double* thr,
                   const int32_t* targets, uint64_t seed, uint64_t step, int64_t n, const T* x,
                   const T* vx, const T* vy, const T* vz, const T* w, const double* U,
                   int64_t u_stride, std::vector<int32_t>& counts, std::vector<int32_t>& offsets,
                   std::vector<uint32_t>& perm, std::vector<SegAction>& act, uint64_t* cnt) {
  const int nn = P.dim == 1 ? P.n_nodes : 1;
  const int n_keys = nn * nb;
  counts.assign(n_keys + 1, 0);
  offsets.assign(n_keys + 2, 0);
  perm.assign(n > 0 ? n : 1, 0);
  act.assign(n_keys, SegAction());
  bin_sort<T>(P, nb, thr, n, x, vx, vy, vz, w, counts.data(), offsets.data(), perm.data());
  bool fire = !pol.ref_trigger && pol.trigger_at <= 0.0;  // <= 0: every check fires
  for (int k = 0; k < n_keys; ++k) {
    const int c = counts[k], t = targets[k];
    if (t <= 0 || c <= 0) continue;
    if (pol.ref_trigger) {
      if (k % nb == nb - 1 && (double)c >= 3.0 * t) fire = true;
    } else if ((double)c >= pol.trigger_at * t) {
      fire = true;
    }
  }
  if (!fire) return false;
  cnt[10] += 1;
  for (int k = 0; k < n_keys; ++k) {
    const int c = counts[k], t = targets[k];
    const uint32_t* ids = perm.data() + offsets[k];
    SegAction& A = act[k];
    if (t <= 0) continue;
    const int keep = (int)(pol.merge_keep * (double)t);
    if (c > 0 && (double)c < pol.split_below * t) {
      const int draws = (int)(pol.split_to * (double)t) - c;
      int hw = 0;
      for (int j = 0; j < c; ++j) hw += (double)w[ids[j]] > pol.min_wt;
      if (hw == 0 || draws <= 0) continue;
      // split (boltzmann.py:806-823): draws with replacement, p ~ w over w > min_wt
      A.kind = 1;
      cnt[8] += 1;
      double wmax = 0;
      for (int j = 0; j < c; ++j)
        if ((double)w[ids[j]] > pol.min_wt) wmax = std::max(wmax, (double)w[ids[j]]);
      int ew;
      frexp(wmax, &ew);
      const int S = 61 - ew - ceil_log2(c);
      std::vector<int64_t> cdf(c);
      int64_t acc = 0;
      for (int j = 0; j < c; ++j) {
        const double wj = (double)w[ids[j]];
        acc += wj > pol.min_wt ? (int64_t)ldexp(wj, S) : 0;
        cdf[j] = acc;
      }
      A.copies.assign(c, 0);
      for (int m = 0; m < draws; ++m) {
        const double u = res_u(U, u_stride, seed, step, k, m);
        const int64_t th = (int64_t)(u * (double)acc);
        A.copies[search(cdf, th, true)] += 1;
      }
    } else if ((double)c >= pol.merge_at * t && keep >= 1) {
      const int ce = (pol.max_ppc > 0 && c > pol.max_ppc) ? pol.max_ppc : c;
      if (keep >= ce) continue;
      // merge (boltzmann.py:825-877)
      A.kind = 2;
      A.ce = ce;
      cnt[9] += 1;
      double W = 0, Px = 0, Py = 0, Pz = 0, En = 0, wmax = 0;
      for (int j = 0; j < ce; ++j) {
        const uint32_t i = ids[j];
        const double wi = w[i], a = vx[i], b = vy[i], cc = vz[i];
        W += wi;
        Px += wi * a;
        Py += wi * b;
        Pz += wi * cc;
        En += wi * (a * a + b * b + cc * cc);
        wmax = std::max(wmax, wi);
      }
      const double mx = Px / W, my = Py / W, mz = Pz / W;
      double Cold = 0;
      for (int j = 0; j < ce; ++j) {
        const uint32_t i = ids[j];
        const double a = vx[i] - mx, b = vy[i] - my, cc = vz[i] - mz;
        Cold += (double)w[i] * (a * a + b * b + cc * cc);
      }
      int ew;
      frexp(wmax, &ew);
      const int S = 61 - ew - ceil_log2(ce);
      std::vector<int64_t> cdf(ce);
      int64_t acc = 0;
      for (int j = 0; j < ce; ++j) {
        acc += (int64_t)ldexp((double)w[ids[j]], S);
        cdf[j] = acc;
      }
      const int K = keep;
      A.wbar = W / (double)K;
      A.sel.resize(K);
      double Sx = 0, Sy = 0, Sz = 0, Se = 0;
      for (int m = 0; m < K; ++m) {
        const double u = res_u(U, u_stride, seed, step, k, m);
        const double r = (1.0 / (double)K) * (u + (double)m);
        const int64_t th = (int64_t)(r * (double)acc);
        A.sel[m] = ids[search(cdf, th, false)];
        const double a = vx[A.sel[m]], b = vy[A.sel[m]], cc = vz[A.sel[m]];
        Sx += a;
        Sy += b;
        Sz += cc;
        Se += a * a + b * b + cc * cc;
      }
      double hx, hy, hz, f;
      if (pol.ref_rescale) {
        hx = A.wbar * Sx / W;
        hy = A.wbar * Sy / W;
        hz = A.wbar * Sz / W;
        const double Tc = (2.0 / 3.0) * (P.kev * En) / W;
        const double Th = (2.0 / 3.0) * (A.wbar * (P.kev * Se)) / W;
        f = Th > 0 ? sqrt(Tc / Th) : 1.0;
      } else {
        hx = Sx / K;
        hy = Sy / K;
        hz = Sz / K;
        double Cn = 0;
        for (int m = 0; m < K; ++m) {
          const double a = vx[A.sel[m]] - hx, b = vy[A.sel[m]] - hy, cc = vz[A.sel[m]] - hz;
          Cn += a * a + b * b + cc * cc;
        }
        Cn *= A.wbar;
        f = Cn > 0 ? sqrt(Cold / Cn) : 1.0;
      }
      for (int q = 0; q < 3; ++q) A.nv[q].resize(K);
      for (int m = 0; m < K; ++m) {
        const uint32_t i = A.sel[m];
        A.nv[0][m] = mx + ((double)vx[i] - hx) * f;
        A.nv[1][m] = my + ((double)vy[i] - hy) * f;
        A.nv[2][m] = mz + ((double)vz[i] - hz) * f;
      }
    }
  }
  return true;
}

// Key-order layout: every segment written to out in key order; a split
// particle is followed by its clones, a merged segment is its K particles
// followed by the members past the max_ppc clip.
template <typename T>
int64_t resample(const orc_params& P, const orc_policy& pol, int nb, const double* thr,
                 const int32_t* targets, uint64_t seed, uint64_t step, int64_t n, const T* x,
                 const T* vx, const T* vy, const T* vz, const T* w, T* ox, T* ovx, T* ovy,
                 T* ovz, T* ow, int64_t cap, uint64_t* cnt, const double* U, int64_t u_stride) {
  std::vector<int32_t> counts, offsets;
  std::vector<uint32_t> perm;
  std::vector<SegAction> act;
  resample_plan<T>(P, pol, nb, thr, targets, seed, step, n, x, vx, vy, vz, w, U, u_stride,
                   counts, offsets, perm, act, cnt);
  const int n_keys = (int)act.size();
  int64_t o = 0;
  auto emit = [&](T a, T b, T c_, T d, T e) {
    if (o < cap) {
      ox[o] = a;
      ovx[o] = b;
      ovy[o] = c_;
      ovz[o] = d;
      ow[o] = e;
    } else {
      cnt[6] += 1;
    }
    ++o;
  };
  for (int k = 0; k < n_keys; ++k) {
    const int c = counts[k];
    const uint32_t* ids = perm.data() + offsets[k];
    const SegAction& A = act[k];
    if (A.kind == 0) {
      for (int j = 0; j < c; ++j) {
        const uint32_t i = ids[j];
        emit(x[i], vx[i], vy[i], vz[i], w[i]);
      }
    } else if (A.kind == 1) {
      for (int j = 0; j < c; ++j) {
        const uint32_t i = ids[j];
        const int reps = 1 + A.copies[j];
        T wn = w[i];
        if (reps > 1) wn = wn / (T)reps;
        for (int r = 0; r < reps; ++r) emit(x[i], vx[i], vy[i], vz[i], wn);
      }
    } else {
      for (size_t m = 0; m < A.sel.size(); ++m)
        emit(x[A.sel[m]], (T)A.nv[0][m], (T)A.nv[1][m], (T)A.nv[2][m], (T)A.wbar);
      for (int j = A.ce; j < c; ++j) {
        const uint32_t i = ids[j];
        emit(x[i], vx[i], vy[i], vz[i], w[i]);
      }
    }
  }
  return o < cap ? o : cap;
}

// In-place layout, the reference's own (boltzmann.py:812-823, :850-874): a
// split particle keeps its slot with weight w/(c+1) and its c clones are
// appended; a merged segment's members 0..ce-1 become holes (w = 0) and its
// K particles are appended; every other slot is left as it is.  The appended
// particles follow the n slots in key order, a split particle's clones in
// member order.  State arrays hold `cap` slots; returns the new slot count.
template <typename T>
int64_t resample_inplace(const orc_params& P, const orc_policy& pol, int nb, const double* thr,
                         const int32_t* targets, uint64_t seed, uint64_t step, int64_t n, T* x,
                         T* vx, T* vy, T* vz, T* w, int64_t cap, uint64_t* cnt, const double* U,
                         int64_t u_stride) {
  std::vector<int32_t> counts, offsets;
  std::vector<uint32_t> perm;
  std::vector<SegAction> act;
  if (!resample_plan<T>(P, pol, nb, thr, targets, seed, step, n, x, vx, vy, vz, w, U, u_stride,
                        counts, offsets, perm, act, cnt))
    return n;
  const int n_keys = (int)act.size();
  int64_t o = n;
  auto emit = [&](T a, T b, T c_, T d, T e) {
    if (o < cap) {
      x[o] = a;
      vx[o] = b;
      vy[o] = c_;
      vz[o] = d;
      w[o] = e;
    } else {
      cnt[6] += 1;
    }
    ++o;
  };
  for (int k = 0; k < n_keys; ++k) {
    const uint32_t* ids = perm.data() + offsets[k];
    const SegAction& A = act[k];
    if (A.kind == 1) {
      for (int j = 0; j < counts[k]; ++j) {
        if (!A.copies[j]) continue;
        const uint32_t i = ids[j];
        const T wn = w[i] / (T)(1 + A.copies[j]);
        for (int r = 0; r < A.copies[j]; ++r) emit(x[i], vx[i], vy[i], vz[i], wn);
        w[i] = wn;
      }
    } else if (A.kind == 2) {
      for (size_t m = 0; m < A.sel.size(); ++m)
        emit(x[A.sel[m]], (T)A.nv[0][m], (T)A.nv[1][m], (T)A.nv[2][m], (T)A.wbar);
      for (int j = 0; j < A.ce; ++j) w[ids[j]] = (T)0;
    }
  }
  return o < cap ? o : cap;
}

}  // namespace

extern "C" {

void orc_bin_sort_f64(const orc_params* P, int nb, const double* thr, int64_t n, const double* x,
                      const double* vx, const double* vy, const double* vz, const double* w,
                      int32_t* counts, int32_t* offsets, uint32_t* perm) {
  bin_sort<double>(*P, nb, thr, n, x, vx, vy, vz, w, counts, offsets, perm);
}
void orc_bin_sort_f32(const orc_params* P, int nb, const double* thr, int64_t n, const float* x,
                      const float* vx, const float* vy, const float* vz, const float* w,
                      int32_t* counts, int32_t* offsets, uint32_t* perm) {
  bin_sort<float>(*P, nb, thr, n, x, vx, vy, vz, w, counts, offsets, perm);
}

// U (optional): the caller's resampling uniforms, U[key*u_stride + m]
int64_t orc_resample_f64(const orc_params* P, const orc_policy* pol, int nb, const double* thr,
                         const int32_t* targets, uint64_t seed, uint64_t step, int64_t n,
                         const double* x, const double* vx, const double* vy, const double* vz,
                         const double* w, double* ox, double* ovx, double* ovy, double* ovz,
                         double* ow, int64_t cap, uint64_t* cnt, const double* U,
                         int64_t u_stride) {
  return resample<double>(*P, *pol, nb, thr, targets, seed, step, n, x, vx, vy, vz, w, ox, ovx,
                          ovy, ovz, ow, cap, cnt, U, u_stride);
}
int64_t orc_resample_f32(const orc_params* P, const orc_policy* pol, int nb, const double* thr,
                         const int32_t* targets, uint64_t seed, uint64_t step, int64_t n,
                         const float* x, const float* vx, const float* vy, const float* vz,
                         const float* w, float* ox, float* ovx, float* ovy, float* ovz, float* ow,
                         int64_t cap, uint64_t* cnt, const double* U, int64_t u_stride) {
  return resample<float>(*P, *pol, nb, thr, targets, seed, step, n, x, vx, vy, vz, w, ox, ovx,
                         ovy, ovz, ow, cap, cnt, U, u_stride);
}
int64_t orc_resample_inplace_f64(const orc_params* P, const orc_policy* pol, int nb,
                                 const double* thr, const int32_t* targets, uint64_t seed,
                                 uint64_t step, int64_t n, double* x, double* vx, double* vy,
                                 double* vz, double* w, int64_t cap, uint64_t* cnt,
                                 const double* U, int64_t u_stride) {
  return resample_inplace<double>(*P, *pol, nb, thr, targets, seed, step, n, x, vx, vy, vz, w,
                                  cap, cnt, U, u_stride);
}
int64_t orc_resample_inplace_f32(const orc_params* P, const orc_policy* pol, int nb,
                                 const double* thr, const int32_t* targets, uint64_t seed,
                                 uint64_t step, int64_t n, float* x, float* vx, float* vy,
                                 float* vz, float* w, int64_t cap, uint64_t* cnt, const double* U,
                                 int64_t u_stride) {
  return resample_inplace<float>(*P, *pol, nb, thr, targets, seed, step, n, x, vx, vy, vz, w,
                                 cap, cnt, U, u_stride);
}

// Grid step: moment update, heavies, Poisson (FP64).
//   chol = [diag(N-1), lower(N-1)] of the banded Cholesky factor (upper form,
//   as returned by scipy.linalg.cholesky_banded for boltzmann.py:744-755).
// Bernoulli function B(x) = x / (e^x - 1) of the Scharfetter-Gummel flux.
static double orc_bernoulli(double x) {
  if (std::fabs(x) < 1e-4) return 1.0 - 0.5 * x + x * x * (1.0 / 12.0);
  return x / std::expm1(x);
}

// Heavies with Scharfetter-Gummel fluxes, sub-cycled to the positivity bound
// h (a_i + b_{i-1}) / dx <= 1; wall nodes carry no inflow (the reference's
// wall rule of heavies_kernel_1D.py:26-35 restated for these fluxes).
static void orc_sg_heavies(int N, double dx, double dt, double mu, double Di, double Ds,
                           int max_sub, const double* E, double* ni, double* ns) {
  std::vector<double> a(N), b(N), Gi(N), Gs(N);
  const double dd = Di / dx;
  for (int i = 0; i < N; ++i) {
    const double pe = mu * E[i] * dx / Di;
    a[i] = dd * orc_bernoulli(-pe);
    b[i] = dd * orc_bernoulli(pe);
  }
  double r = 2.0 * Ds / (dx * dx);
  for (int i = 1; i < N; ++i) r = std::max(r, (a[i] + b[i - 1]) / dx);
  int m = (int)std::ceil(dt * r);
  if (m < 1) m = 1;
  if (max_sub > 0 && m > max_sub) m = max_sub;
  const double h = dt / m;
  const double f0 = std::exp(-(mu * std::max(E[0], 0.0) * dx / Di));
  const double fN = std::exp(mu * std::min(E[N - 1], 0.0) * dx / Di);
  auto bc = [&]() {
    ni[0] = ni[1] * f0;
    ni[N] = ni[N - 1] * fN;
    ns[0] = 0.0;
    ns[N] = 0.0;
  };
  for (int s = 0; s < m; ++s) {
    bc();
    for (int i = 0; i < N; ++i) {
      Gi[i] = a[i] * ni[i] - b[i] * ni[i + 1];
      Gs[i] = -Ds * (ns[i + 1] - ns[i]) / dx;
    }
    for (int i = 1; i < N; ++i) {
      ni[i] = ni[i] - h * (Gi[i] - Gi[i - 1]) / dx;
      ns[i] = ns[i] - h * (Gs[i] - Gs[i - 1]) / dx;
    }
  }
  bc();
}

void orc_field_step(int N, double dx, double dt, double eps, double v0, double freq, double mu_i,
                    double D_i, double D_s, int fluid_on, int poisson_on, int ion_scheme,
                    int max_substeps, double t, const double* mom, double* ni, double* ns,
                    double* Te, double* g2, double* V, double* E, const double* chol) {
  const int nn = N + 1;
  const double* ne = mom;
  const double* nsrc = mom + nn;
  const double* nrg = mom + 2 * nn;
  for (int i = 0; i < nn; ++i) {
    ni[i] = ni[i] + nsrc[i];
    double te = (2.0 / 3.0) * nrg[i] / ne[i];
    Te[i] = (te != te) ? 0.0 : te;
    g2[i] = g2[i] + nsrc[i];
  }
  if (fluid_on && ion_scheme == 1) {
    orc_sg_heavies(N, dx, dt, mu_i, D_i, D_s, max_substeps, E, ni, ns);
  } else if (fluid_on) {
    std::vector<double> Ji(N), Js(N);
    auto bc = [&]() {
      ni[0] = ni[1] * ((D_i / dx - 0.5 * mu_i * std::max(E[0], 0.0)) /
                       (D_i / dx + 0.5 * mu_i * std::max(E[0], 0.0)));
      ni[N] = ni[N - 1] * ((D_i / dx + 0.5 * mu_i * std::min(E[N - 1], 0.0)) /
                           (D_i / dx - 0.5 * mu_i * std::min(E[N - 1], 0.0)));
      ns[0] = 0;
      ns[N] = 0;
    };
    bc();
    for (int i = 0; i < N; ++i) {
      Ji[i] = mu_i * (0.5 * (ni[i] + ni[i + 1])) * E[i] - D_i * (ni[i + 1] - ni[i]) / dx;
      Js[i] = -D_s * (ns[i + 1] - ns[i]) / dx;
    }
    for (int i = 1; i < N; ++i) {
      ni[i] = ni[i] - dt * (Ji[i] - Ji[i - 1]) / dx;
      ns[i] = ns[i] - dt * (Js[i] - Js[i - 1]) / dx;
    }
    bc();
  }
  if (poisson_on) {
    const double* diag = chol;
    const double* low = chol + (N - 1);
    std::vector<double> rhs(N - 1), ty(N - 1);
    for (int i = 0; i < N - 1; ++i) rhs[i] = (1.0 / eps) * (ni[i + 1] - ne[i + 1]);
    rhs[0] = rhs[0] + (1.0 / (dx * dx)) * v0 * sin(2 * 3.141592653589793 * freq * t);
    ty[0] = rhs[0] / diag[0];
    for (int i = 1; i < N - 1; ++i) ty[i] = (1.0 / diag[i]) * (rhs[i] - low[i] * ty[i - 1]);
    V[N - 2] = ty[N - 2] / diag[N - 2];
    for (int i = 1; i < N - 1; ++i)
      V[N - 2 - i] = (1.0 / diag[N - 2 - i]) * (ty[N - 2 - i] - low[N - 2 - i + 1] * V[N - 2 - i + 1]);
    for (int i = 0; i < N - 2; ++i) E[i + 1] = -(V[i + 1] - V[i]) / dx;
    E[0] = -(V[0] - v0 * sin(2 * 3.141592653589793 * freq * t)) / dx;
    E[N - 1] = -(0 - V[N - 2]) / dx;
  }
}

}  // extern "C"
