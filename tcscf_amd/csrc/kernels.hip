// Device kernels of the grid-integral hot path (gfx950):
//   - STO basis tabulation (values + Cartesian gradients, real solid harmonics)
//   - pair-density packing P[n2, p(j,i)] = w2 A2[n2,j] A2[n2,i]
//   - fused kernel generation + contraction F = Kw . P on MFMA, in FP64,
//     FP32 and BF16 hi/lo split; Kw(r1, r2) is generated per tile in
//     registers (FP64/FP32) or LDS (BF16) and never stored.
#include <hip/hip_runtime.h>

#include <type_traits>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <vector>

#include "common.hpp"
#include "device_util.hpp"

namespace tc {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef unsigned short us8 __attribute__((ext_vector_type(8)));

constexpr double kFar = 1.0e6;  // padding coordinate: far from every grid point
constexpr int kVFold = 4;       // FP64 fold period (32-point steps) of the BF16 V tiles
// V launches above this many (r1 point, r3 point, pair) units fold every step:
// on a long launch the four-step fold loses its speed (measured: 65,536 x
// 65,536 points, 820 pairs, 3.5e12 units: 98-105 ms against 87-88 ms folding
// every step; 30,200 x 9,700 points, 465 pairs: 2.6 ms against 2.9 ms)
constexpr double kVLongWork = 2e12;
constexpr int kOchiMaxTerms = TCSCF_OCHI_MAX_TERMS;

// Output layouts of the 16x16 MFMA accumulators: element i of lane l = 16 g + r
// holds D[row][r] with row = d_row_*(g, i).  (Pinned by tcscf_selftest_mfma.)
__host__ __device__ constexpr int d_row_f64(int g, int i) { return g + 4 * i; }
__host__ __device__ constexpr int d_row_f32(int g, int i) { return 4 * g + i; }
__host__ __device__ constexpr int d_row_bf16(int g, int i) { return 4 * g + i; }

/* ------------------------------------------------------------------------ */
/* scratch buffers                                                           */
/* ------------------------------------------------------------------------ */
namespace {
struct ScratchBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
};
std::mutex g_scratch_mu;
ScratchBuf g_scratch[64][kNumSlots];
}  // namespace

void* scratch(int slot, size_t bytes, int* rc) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) {
    set_error("scratch: hipGetDevice failed");
    *rc = TCSCF_EHIP;
    return nullptr;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  ScratchBuf& b = g_scratch[dev][slot];
  if (b.bytes < bytes) {
    if (b.ptr) {
      // Older work queued on any stream may still read the old buffer.
      if (hipDeviceSynchronize() != hipSuccess || hipFree(b.ptr) != hipSuccess) {
        set_error("scratch: release failed");
        *rc = TCSCF_EHIP;
        return nullptr;
      }
      b.ptr = nullptr;
      b.bytes = 0;
    }
    size_t want = std::max(bytes, (size_t)1 << 20);
    if (hipMalloc(&b.ptr, want) != hipSuccess) {
      b.ptr = nullptr;
      set_error("scratch: hipMalloc of " + std::to_string(want) + " bytes failed");
      *rc = TCSCF_ENOMEM;
      return nullptr;
    }
    b.bytes = want;
  }
  *rc = 0;
  return b.ptr;
}

/* ------------------------------------------------------------------------ */
/* STO basis tabulation                                                      */
/* ------------------------------------------------------------------------ */
constexpr int kMaxFnsPerLaunch = 96;

struct StoTable {
  int nb;
  int code[kMaxFnsPerLaunch];      // n | l << 8 | (m + 8) << 16 | kind << 24
  double zeta[kMaxFnsPerLaunch];   // exponent of e^{-zeta r} (alpha, or b = Z / n)
  double norm[kMaxFnsPerLaunch];
};

// Associated Laguerre polynomial L_n^(k)(x) by the three-term recurrence
// (m + 1) L_{m+1} = (2m + 1 + k - x) L_m - (m + k) L_{m-1}; 0 for n < 0.
__device__ inline double laguerre(int n, int k, double x) {
  if (n < 0) return 0.0;
  double lm1 = 1.0, l = 1.0 + k - x;
  if (n == 0) return lm1;
  for (int m = 1; m < n; ++m) {
    const double lp = ((2 * m + 1 + k - x) * l - (m + k) * lm1) / (m + 1);
    lm1 = l;
    l = lp;
  }
  return l;
}

// Real solid harmonic S_lm(x, y, z) (homogeneous of degree l, unit-normalised
// on the sphere) and its gradient.  Standard real harmonics without the
// Condon-Shortley phase: p_{+1} ~ x, p_{-1} ~ y, p_0 ~ z.
__device__ inline void solid_harmonic(int l, int m, double x, double y, double z, double* S, double* g) {
  g[0] = g[1] = g[2] = 0;
  switch (l) {
    case 0:
      *S = 0.28209479177387814;
      return;
    case 1: {
      const double c = 0.48860251190291992;
      if (m == 1) { *S = c * x; g[0] = c; }
      else if (m == -1) { *S = c * y; g[1] = c; }
      else { *S = c * z; g[2] = c; }
      return;
    }
    case 2: {
      const double c = 1.0925484305920792, c0 = 0.31539156525252005, c2 = 0.54627421529603959;
      switch (m) {
        case -2: *S = c * x * y; g[0] = c * y; g[1] = c * x; return;
        case -1: *S = c * y * z; g[1] = c * z; g[2] = c * y; return;
        case 0: *S = c0 * (2 * z * z - x * x - y * y); g[0] = -2 * c0 * x; g[1] = -2 * c0 * y; g[2] = 4 * c0 * z; return;
        case 1: *S = c * x * z; g[0] = c * z; g[2] = c * x; return;
        default: *S = c2 * (x * x - y * y); g[0] = 2 * c2 * x; g[1] = -2 * c2 * y; return;
      }
    }
    default: {
      const double c3 = 0.59004358992664352, c2 = 2.8906114426405538, c1 = 0.45704579946446572,
                   c0 = 0.37317633259011539, c2b = 1.4453057213202769;
      switch (m) {
        case -3: *S = c3 * (3 * x * x * y - y * y * y); g[0] = 6 * c3 * x * y; g[1] = 3 * c3 * (x * x - y * y); return;
        case -2: *S = c2 * x * y * z; g[0] = c2 * y * z; g[1] = c2 * x * z; g[2] = c2 * x * y; return;
        case -1:
          *S = c1 * y * (4 * z * z - x * x - y * y);
          g[0] = -2 * c1 * x * y; g[1] = c1 * (4 * z * z - x * x - 3 * y * y); g[2] = 8 * c1 * y * z;
          return;
        case 0:
          *S = c0 * z * (2 * z * z - 3 * x * x - 3 * y * y);
          g[0] = -6 * c0 * x * z; g[1] = -6 * c0 * y * z; g[2] = c0 * (6 * z * z - 3 * x * x - 3 * y * y);
          return;
        case 1:
          *S = c1 * x * (4 * z * z - x * x - y * y);
          g[0] = c1 * (4 * z * z - 3 * x * x - y * y); g[1] = -2 * c1 * x * y; g[2] = 8 * c1 * x * z;
          return;
        case 2: *S = c2b * z * (x * x - y * y); g[0] = 2 * c2b * x * z; g[1] = -2 * c2b * y * z; g[2] = c2b * (x * x - y * y); return;
        default: *S = c3 * (x * x * x - 3 * x * y * y); g[0] = 3 * c3 * (x * x - y * y); g[1] = -6 * c3 * x * y; return;
      }
    }
  }
}

// phi = g(r) S_lm(x,y,z), grad phi = g'(r) (x,y,z)/r S + g(r) grad S, with
//   STO:       g = N r^k e^{-zeta r}, k = n-1-l
//              (grad written as N e^{-zeta r} [ (k r^{k-2} - zeta r^{k-1}) S (x,y,z) + r^k grad S ]);
//   Ochi:      g = N L_n^(2l+2)(2 zeta r) e^{-zeta r};
//   hydrogen:  g = N L_(n-l-1)^(2l+1)(2 zeta r) e^{-zeta r}, zeta = Z / n;
// and d/dx L_q^(k)(x) = -L_(q-1)^(k+1)(x).
__global__ __launch_bounds__(256) void basis_eval_kernel(const double* __restrict__ pts, int64_t n, StoTable t,
                                                         int nb_total, int b0, double* __restrict__ values,
                                                         double* __restrict__ grads) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n * t.nb) return;
  const int64_t k = idx / t.nb;
  const int b = (int)(idx - k * t.nb);
  const double x = pts[k], y = pts[n + k], z = pts[2 * n + k];
  const double r = sqrt(x * x + y * y + z * z);
  const int code = t.code[b];
  const int nq = code & 0xff, l = (code >> 8) & 0xff, m = ((code >> 16) & 0xff) - 8, kind = code >> 24;
  const double zeta = t.zeta[b];
  double S, gS[3];
  solid_harmonic(l, m, x, y, z, &S, gS);
  const double e = t.norm[b] * exp(-zeta * r);
  const int64_t o = k * nb_total + b0 + b;
  if (kind == TCSCF_BASIS_STO) {
    const int kp = nq - 1 - l;
    const double rk = kp == 0 ? 1.0 : pow(r, (double)kp);
    values[o] = e * rk * S;
    if (grads) {
      const double rkm1 = kp == 0 ? 1.0 / r : pow(r, (double)(kp - 1));
      const double rkm2 = kp == 0 ? 0.0 : pow(r, (double)(kp - 2));
      const double radial = (kp * rkm2 - zeta * rkm1) * S;
      grads[3 * o + 0] = e * (radial * x + rk * gS[0]);
      grads[3 * o + 1] = e * (radial * y + rk * gS[1]);
      grads[3 * o + 2] = e * (radial * z + rk * gS[2]);
    }
    return;
  }
  // Laguerre degree q and order kl of the radial polynomial
  const int q = kind == TCSCF_BASIS_OCHI ? nq : nq - l - 1;
  const int kl = kind == TCSCF_BASIS_OCHI ? 2 * l + 2 : 2 * l + 1;
  const double xl = 2 * zeta * r;
  const double L = laguerre(q, kl, xl);
  values[o] = e * L * S;
  if (grads) {
    const double dg = e * (-2 * zeta * laguerre(q - 1, kl + 1, xl) - zeta * L);  // g'(r)
    const double radial = dg / r * S;
    grads[3 * o + 0] = radial * x + e * L * gS[0];
    grads[3 * o + 1] = radial * y + e * L * gS[1];
    grads[3 * o + 2] = radial * z + e * L * gS[2];
  }
}

/* ------------------------------------------------------------------------ */
/* Pair densities and padded coordinates                                     */
/* ------------------------------------------------------------------------ */
__device__ inline void pair_of(int64_t p, int* j, int* i) {
  int ii = (int)((sqrt(8.0 * (double)p + 1.0) - 1.0) * 0.5);
  while ((int64_t)(ii + 1) * (ii + 2) / 2 <= p) ++ii;
  while ((int64_t)ii * (ii + 1) / 2 > p) --ii;
  *i = ii;
  *j = (int)(p - (int64_t)ii * (ii + 1) / 2);
}

// Row-major P[k][p] = P(ord[k], p) (ld = ldp), rows k < n2p, zero outside (n2, npair).
template <class T>
__global__ __launch_bounds__(256) void pack_pairs_rowmajor(const double* __restrict__ A, const double* __restrict__ w,
                                                           const int* __restrict__ ord, int64_t n2, int nb,
                                                           int64_t npair, int64_t n2p, int64_t ldp,
                                                           T* __restrict__ P) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n2p * ldp) return;
  const int64_t k = idx / ldp, p = idx - k * ldp;
  double v = 0;
  if (k < n2 && p < npair) {
    const int64_t ks = ord ? ord[k] : k;
    int j, i;
    pair_of(p, &j, &i);
    v = w[ks] * A[ks * nb + j] * A[ks * nb + i];
  }
  P[idx] = (T)v;
}

__device__ inline unsigned short bf16_bits_rne(float f) {
  unsigned int u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ inline float bf16_to_float(unsigned short h) { return __uint_as_float((unsigned int)h << 16); }

// BF16 split of P in MFMA B-fragment order for v_mfma_f32_16x16x32_bf16:
// fragment (kb, cb) = 64 lanes x 8 bf16, lane l = 16 g + r holds
// P[ord[32 kb + 8 g + e]][16 cb + r], e = 0..7, where ord is the summation order
// of the r2 points (NULL = grid order).  NSPLIT terms: P ~ hi + mid (+ lo), each
// plane nkb*ncb*512 bf16, planes back to back.
template <int NSPLIT>
__global__ __launch_bounds__(256) void pack_pairs_bf16(const double* __restrict__ A, const double* __restrict__ w,
                                                       const int* __restrict__ ord, int64_t n2, int nb, int64_t npair,
                                                       int64_t nkb, int64_t ncb, unsigned short* __restrict__ planes) {
  // one fragment lane: its 8 bf16 of every plane (points 32 kb + 8 g .. + 7,
  // pair 16 cb + r) as one 16-byte store per plane
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t lanes = nkb * ncb * 64, total = lanes * 8;
  if (idx >= lanes) return;
  const int64_t frag = idx >> 6;
  const int lane = (int)(idx & 63);
  const int64_t kb = frag / ncb, cb = frag - kb * ncb;
  const int g = lane >> 4, r = lane & 15;
  const int64_t k0 = 32 * kb + 8 * g, p = 16 * cb + r;
  int j = 0, i = 0;
  if (p < npair) pair_of(p, &j, &i);
  unsigned short h[NSPLIT][8];
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const int64_t k = k0 + e;
    double v = 0;
    if (k < n2 && p < npair) {
      const int64_t ks = ord ? ord[k] : k;
      v = w[ks] * A[ks * nb + j] * A[ks * nb + i];
    }
#pragma unroll
    for (int t = 0; t < NSPLIT; ++t) {
      h[t][e] = bf16_bits_rne((float)v);
      v -= (double)bf16_to_float(h[t][e]);
    }
  }
#pragma unroll
  for (int t = 0; t < NSPLIT; ++t) {
    uint4 q;
    q.x = h[t][0] | ((unsigned)h[t][1] << 16);
    q.y = h[t][2] | ((unsigned)h[t][3] << 16);
    q.z = h[t][4] | ((unsigned)h[t][5] << 16);
    q.w = h[t][6] | ((unsigned)h[t][7] << 16);
    *(uint4*)(planes + t * total + idx * 8) = q;
  }
}

// Padded SoA coordinates (3 x n2p) in FP64 and as float4 {x, y, z, 0}, in the
// summation order ord (NULL = grid order).
__global__ void pad_coords(const double* __restrict__ pts, const int* __restrict__ ord, int64_t n, int64_t np,
                           double* __restrict__ out, float4* __restrict__ out4) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= np) return;
  double x = kFar, y = kFar, z = kFar;
  if (k < n) {
    const int64_t ks = ord ? ord[k] : k;
    x = pts[ks];
    y = pts[n + ks];
    z = pts[2 * n + ks];
  }
  out[k] = x;
  out[np + k] = y;
  out[2 * np + k] = z;
  out4[k] = make_float4((float)x, (float)y, (float)z, 0.f);
}

/* ------------------------------------------------------------------------ */
/* Kernel functions K(d), d = r1 - r2; coincident points give 0             */
/* ------------------------------------------------------------------------ */
__device__ inline double rsqrt_f64(double s) {
  double y = __builtin_amdgcn_rsq(s);
  double e = fma(-s * y, y, 1.0);
  y = fma(0.5 * y, e, y);
  e = fma(-s * y, y, 1.0);
  y = fma(0.5 * y, e, y);
  return y;
}
__device__ inline double rsqrt_any(double s) { return rsqrt_f64(s); }
__device__ inline float rsqrt_any(float s) { return __builtin_amdgcn_rsqf(s); }

// The kernel formulas, each with one fixed rounding sequence (contraction off),
// so every kernel functor -- including the batched one -- produces the same bits.
template <class T>
__device__ inline T dist2(T dx, T dy, T dz) {
#pragma clang fp contract(off)
  return dx * dx + dy * dy + dz * dz;
}
// a / b: exact division in FP64; in FP32 (the BF16 and FP32 contractions,
// whose K is rounded to bf16 pieces or accumulated in FP32 anyway) the
// hardware reciprocal (1 ulp) in place of the IEEE division sequence.
__device__ inline double quot(double a, double b) { return a / b; }
__device__ inline float quot(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

template <class T>
__device__ inline T k_tc_scalar(T s, T y, double a12, double c2a2) {
#pragma clang fp contract(off)
  const T t = s * y + T(a12);
  const T t2 = t * t;
  return T(2) * y - quot(T(c2a2), t2 * t2);
}
template <class T>
__device__ inline T k_tc_smooth(T s, T y, double a12) {
#pragma clang fp contract(off)
  const T t = s * y + T(a12);
  const T t2 = t * t;
  return quot(T(1), t2 * t2);
}
template <class T>
__device__ inline T k_grad_scale(T s, T y, double a12) {
#pragma clang fp contract(off)
  const T t = s * y + T(a12);
  return quot(T(a12) * y, t * t);
}

struct KernCoulomb {
  static constexpr int ncomp = 1;
  static constexpr bool absolute = false;
  template <class T>
  __device__ inline void operator()(T dx, T dy, T dz, T* out) const {
    const T s = dist2(dx, dy, dz);
    out[0] = s > T(0) ? rsqrt_any(s) : T(0);
  }
};

// 2/r12 - 2 c^2 a12^2 / (r12 + a12)^4  (r12-only Ochi term, c = c_same or c_oppo)
struct KernTCScalar {
  static constexpr int ncomp = 1;
  static constexpr bool absolute = false;
  double a12, c2a2;  // c2a2 = 2 c^2 a12^2
  template <class T>
  __device__ inline void operator()(T dx, T dy, T dz, T* out) const {
    const T s = dist2(dx, dy, dz);
    const T y = rsqrt_any(s);
    out[0] = s > T(0) ? k_tc_scalar(s, y, a12, c2a2) : T(0);
  }
};

// 1 / (r12 + a12)^4: the smooth part of the TC scalar kernel, which is
// 2/r12 - 2 c^2 a12^2 / (r12 + a12)^4 (coincident pairs excluded).
struct KernTCSmooth {
  static constexpr int ncomp = 1;
  static constexpr bool absolute = false;
  double a12;
  template <class T>
  __device__ inline void operator()(T dx, T dy, T dz, T* out) const {
    const T s = dist2(dx, dy, dz);
    const T y = rsqrt_any(s);
    out[0] = s > T(0) ? k_tc_smooth(s, y, a12) : T(0);
  }
};

// The reference's general Ochi Jastrow (src/jastrowFunctions.hpp:9-115),
//   u(r1, r2) = sum_t c_t (r12/(r12+a12))^i_t (r1/(r1+a))^j_t (r2/(r2+a))^k_t,
// with the coefficients c_t of one spin case already selected.
struct OchiTerms {
  double a, a12;
  int nt;
  int i[kOchiMaxTerms], j[kOchiMaxTerms], k[kOchiMaxTerms];
  double c[kOchiMaxTerms];
};

template <class T>
__device__ inline T sqrt_any(T s);
template <>
__device__ inline double sqrt_any(double s) { return __builtin_sqrt(s); }
template <>
__device__ inline float sqrt_any(float s) { return __builtin_sqrtf(s); }

template <class T>
__device__ inline T ipow(T x, int n) {
  T r = T(1);
  for (int e = 0; e < n; ++e) r = r * x;
  return r;
}

// grad_p u(p, q) = r12L (p - q) + rpL p (Ochi::gradient(p, q), jastrowFunctions.hpp:72-108),
// given d = p - q, s = |d|^2 > 0 and r12 = |d|.  At p = 0 the r_p terms have no
// defined direction; they contribute 0 there.
template <class T>
__device__ inline void ochi_gradient(const OchiTerms& u, T px, T py, T pz, T qx, T qy, T qz, T dx, T dy, T dz, T s,
                                     T r12, T* g) {
#pragma clang fp contract(off)
  const T sp = dist2(px, py, pz), rp = sqrt_any(sp), rq = sqrt_any(dist2(qx, qy, qz));
  const T pT = rp / (rp + T(u.a)), qT = rq / (rq + T(u.a)), tT = r12 / (r12 + T(u.a12));
  const T pS = sp > T(0) ? T(u.a) / (sp * (rp + T(u.a))) : T(0);
  const T tS = T(u.a12) / (s * (r12 + T(u.a12)));
  T pL = T(0), tL = T(0);
  for (int t = 0; t < u.nt; ++t) {
    const T f = T(u.c[t]) * ipow(tT, u.i[t]) * ipow(pT, u.j[t]) * ipow(qT, u.k[t]);
    pL = pL + f * T(u.j[t]) * pS;
    tL = tL + f * T(u.i[t]) * tS;
  }
  g[0] = tL * dx + pL * px;
  g[1] = tL * dy + pL * py;
  g[2] = tL * dz + pL * pz;
}

// GRAD = false: 2/r12 - |grad_1 u(r1,r2)|^2 - |grad_1 u(r2,r1)|^2 (the TC scalar
// integrand of src/unitTests/HartreeFockTest.cpp:276-296); GRAD = true:
// grad_1 u(r1, r2) (:300-316).  Needs r1 and r2 themselves, not only r1 - r2.
template <bool GRAD>
struct KernOchi {
  static constexpr int ncomp = GRAD ? 3 : 1;
  static constexpr bool absolute = true;
  OchiTerms u;
  template <class T>
  __device__ inline void eval(T ax, T ay, T az, T bx, T by, T bz, T* out) const {
#pragma clang fp contract(off)
    const T dx = ax - bx, dy = ay - by, dz = az - bz;
    const T s = dist2(dx, dy, dz);
    if (!(s > T(0))) {
      for (int c = 0; c < ncomp; ++c) out[c] = T(0);
      return;
    }
    const T r12 = sqrt_any(s);
    T g[3];
    ochi_gradient(u, ax, ay, az, bx, by, bz, dx, dy, dz, s, r12, g);
    if constexpr (GRAD) {
      out[0] = g[0];
      out[1] = g[1];
      out[2] = g[2];
    } else {
      T h[3];
      ochi_gradient(u, bx, by, bz, ax, ay, az, T(0) - dx, T(0) - dy, T(0) - dz, s, r12, h);
      out[0] = T(2) / r12 - (g[0] * g[0] + g[1] * g[1] + g[2] * g[2]) - (h[0] * h[0] + h[1] * h[1] + h[2] * h[2]);
    }
  }
};

// The TC scalar kernel of both spins in one contraction: component 0 is
// 1/r12 (5 products), component 1 the smooth 1/(r12 + a12)^4 (3 products),
// and the epilogue writes F_spin = 2 F_C - c2a2_spin G (c2a2 = 2 c^2 a12^2)
// for both spins from the FP64 sums.
struct KernTCPair {
  static constexpr int ncomp = 2;
  static constexpr int ncomp5 = 1;  // components that take the five-product split
  static constexpr int nout = 2;
  static constexpr bool absolute = false;
  double a12, c2a2_same, c2a2_oppo;
  template <class T>
  __device__ inline void operator()(T dx, T dy, T dz, T* out) const {
    const T s = dist2(dx, dy, dz);
    const T y = rsqrt_any(s);
    out[0] = s > T(0) ? y : T(0);
    out[1] = s > T(0) ? k_tc_smooth(s, y, a12) : T(0);
  }
  __device__ inline void finish(const double* v, double* o) const {
#pragma clang fp contract(off)
    o[0] = 2 * v[0] - c2a2_same * v[1];
    o[1] = 2 * v[0] - c2a2_oppo * v[1];
  }
};

// Components that take the five-product split when NPROD = 5 (default: all);
// functors with `nout` combine their components in the epilogue.
template <class K, class = void>
struct Comp5 : std::integral_constant<int, K::ncomp> {};
template <class K>
struct Comp5<K, std::void_t<decltype(K::ncomp5)>> : std::integral_constant<int, K::ncomp5> {};
template <class K, class = void>
struct Combines : std::false_type {};
template <class K>
struct Combines<K, std::void_t<decltype(K::nout)>> : std::true_type {};

// K at r1 = a, r2 = b for any kernel functor: functors marked `absolute` see
// both points, the others only d = a - b.
template <class Kern, class T>
__device__ inline void kernel_eval(const Kern& kern, T ax, T ay, T az, T bx, T by, T bz, T* out) {
  if constexpr (Kern::absolute) kern.eval(ax, ay, az, bx, by, bz, out);
  else kern(ax - bx, ay - by, az - bz, out);
}

// grad_1 u(r1, r2) = c a12 / (r12 (r12 + a12)^2) (r1 - r2), with c = 1.
struct KernJastrowGrad {
  static constexpr int ncomp = 3;
  static constexpr bool absolute = false;
  double a12;
  template <class T>
  __device__ inline void operator()(T dx, T dy, T dz, T* out) const {
#pragma clang fp contract(off)
    const T s = dist2(dx, dy, dz);
    const T y = rsqrt_any(s);
    const T gg = s > T(0) ? k_grad_scale(s, y, a12) : T(0);
    out[0] = gg * dx;
    out[1] = gg * dy;
    out[2] = gg * dz;
  }
};

/* ------------------------------------------------------------------------ */
/* FP64 contraction: v_mfma_f64_16x16x4_f64                                  */
/* A (16x4): lane l holds A[l%16][l/16]; B (4x16): B[l/16][l%16];            */
/* D (16x16): d4 element i of lane l is D[l/16 + 4i][l%16].                  */
/* Workgroup = 4 waves as 2 (rows) x 2 (cols); wave tile 16*MB x 64.         */
/* ------------------------------------------------------------------------ */
template <int MB, class Kern>
__global__ __launch_bounds__(256) void fgemm_f64_kernel(const double* __restrict__ p1, int64_t n1,
                                                        const double* __restrict__ x2, int64_t n2p,
                                                        const double* __restrict__ P, int64_t ldp,
                                                        void* __restrict__ Fout, int out_f32, int64_t npair,
                                                        double scale, Kern kern) {
  constexpr int NC = Kern::ncomp;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = lane >> 4, r = lane & 15;
  const int64_t rowW = (int64_t)blockIdx.y * (32 * MB) + (wave >> 1) * (16 * MB);
  const int64_t colW = (int64_t)blockIdx.x * 128 + (wave & 1) * 64;
  double ax[MB], ay[MB], az[MB];
#pragma unroll
  for (int m = 0; m < MB; ++m) {
    const int64_t row = rowW + 16 * m + r;
    const bool ok = row < n1;
    ax[m] = ok ? p1[row] : -kFar;
    ay[m] = ok ? p1[n1 + row] : -kFar;
    az[m] = ok ? p1[2 * n1 + row] : -kFar;
  }
  d4 acc[NC][MB][4];
#pragma unroll
  for (int c = 0; c < NC; ++c)
#pragma unroll
    for (int m = 0; m < MB; ++m)
#pragma unroll
      for (int n = 0; n < 4; ++n) acc[c][m][n] = d4{0, 0, 0, 0};
  const double* __restrict__ y2 = x2 + n2p;
  const double* __restrict__ z2 = x2 + 2 * n2p;
  const double* __restrict__ Pb = P + colW + r;
#pragma unroll 2
  for (int64_t k0 = 0; k0 < n2p; k0 += 4) {
    const int64_t k = k0 + g;
    const double bx = x2[k], by = y2[k], bz = z2[k];
    double b[4];
#pragma unroll
    for (int n = 0; n < 4; ++n) b[n] = Pb[k * ldp + 16 * n];
    double a[MB][NC];
#pragma unroll
    for (int m = 0; m < MB; ++m) kernel_eval(kern, ax[m], ay[m], az[m], bx, by, bz, a[m]);
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
      for (int m = 0; m < MB; ++m)
#pragma unroll
        for (int n = 0; n < 4; ++n)
          acc[c][m][n] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[m][c], b[n], acc[c][m][n], 0, 0, 0);
  }
#pragma unroll
  for (int c = 0; c < NC; ++c)
#pragma unroll
    for (int m = 0; m < MB; ++m)
#pragma unroll
      for (int n = 0; n < 4; ++n)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int64_t row = rowW + 16 * m + d_row_f64(g, i), col = colW + 16 * n + r;
          if (row < n1 && col < npair) {
            const int64_t o = (row * NC + c) * npair + col;
            const double v = scale * acc[c][m][n][i];
            if (out_f32) ((float*)Fout)[o] = (float)v;
            else ((double*)Fout)[o] = v;
          }
        }
}

/* ------------------------------------------------------------------------ */
/* FP32 contraction: v_mfma_f32_16x16x4_f32 (same fragment layout as FP64),  */
/* FP32 accumulation inside 64-point chunks, FP64 sums across chunks.        */
/* ------------------------------------------------------------------------ */
template <int MB, class Kern>
__global__ __launch_bounds__(256) void fgemm_f32_kernel(const double* __restrict__ p1, int64_t n1,
                                                        const float4* __restrict__ x2, int64_t n2p,
                                                        const float* __restrict__ P, int64_t ldp,
                                                        void* __restrict__ Fout, int out_f32, int64_t npair,
                                                        double scale, Kern kern) {
  constexpr int NC = Kern::ncomp;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = lane >> 4, r = lane & 15;
  const int64_t rowW = (int64_t)blockIdx.y * (32 * MB) + (wave >> 1) * (16 * MB);
  const int64_t colW = (int64_t)blockIdx.x * 128 + (wave & 1) * 64;
  float ax[MB], ay[MB], az[MB];
#pragma unroll
  for (int m = 0; m < MB; ++m) {
    const int64_t row = rowW + 16 * m + r;
    const bool ok = row < n1;
    ax[m] = ok ? (float)p1[row] : (float)-kFar;
    ay[m] = ok ? (float)p1[n1 + row] : (float)-kFar;
    az[m] = ok ? (float)p1[2 * n1 + row] : (float)-kFar;
  }
  f4 acc[NC][MB][4];
  double sum[NC][MB][4][4];
#pragma unroll
  for (int c = 0; c < NC; ++c)
#pragma unroll
    for (int m = 0; m < MB; ++m)
#pragma unroll
      for (int n = 0; n < 4; ++n) {
        acc[c][m][n] = f4{0, 0, 0, 0};
#pragma unroll
        for (int i = 0; i < 4; ++i) sum[c][m][n][i] = 0;
      }
  const float* __restrict__ Pb = P + colW + r;
  for (int64_t kc = 0; kc < n2p; kc += 64) {
    const int64_t kend = kc + 64 < n2p ? kc + 64 : n2p;
#pragma unroll 2
    for (int64_t k0 = kc; k0 < kend; k0 += 4) {
      const int64_t k = k0 + g;
      const float4 q = x2[k];
      float b[4];
#pragma unroll
      for (int n = 0; n < 4; ++n) b[n] = Pb[k * ldp + 16 * n];
      float a[MB][NC];
#pragma unroll
      for (int m = 0; m < MB; ++m) kernel_eval(kern, ax[m], ay[m], az[m], q.x, q.y, q.z, a[m]);
#pragma unroll
      for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int m = 0; m < MB; ++m)
#pragma unroll
          for (int n = 0; n < 4; ++n)
            acc[c][m][n] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[m][c], b[n], acc[c][m][n], 0, 0, 0);
    }
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
      for (int m = 0; m < MB; ++m)
#pragma unroll
        for (int n = 0; n < 4; ++n) {
#pragma unroll
          for (int i = 0; i < 4; ++i) sum[c][m][n][i] += (double)acc[c][m][n][i];
          acc[c][m][n] = f4{0, 0, 0, 0};
        }
  }
#pragma unroll
  for (int c = 0; c < NC; ++c)
#pragma unroll
    for (int m = 0; m < MB; ++m)
#pragma unroll
      for (int n = 0; n < 4; ++n)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int64_t row = rowW + 16 * m + d_row_f32(g, i), col = colW + 16 * n + r;
          if (row < n1 && col < npair) {
            const int64_t o = (row * NC + c) * npair + col;
            const double v = scale * sum[c][m][n][i];
            if (out_f32) ((float*)Fout)[o] = (float)v;
            else ((double*)Fout)[o] = v;
          }
        }
}

/* ------------------------------------------------------------------------ */
/* BF16 split contraction: v_mfma_f32_16x16x32_bf16                          */
/* A (16x32): lane l = 16g + r holds A[r][8g .. 8g+7]; B (32x16): lane l     */
/* holds B[8g .. 8g+7][r]; D rows d_row_bf16(g, i).                          */
/* K ~ Kh + Kl (bf16), P ~ Ph + Pm (+ Pl) (bf16, pre-split in fragment      */
/* order).  NPROD = 3: Kh Ph + Kh Pm + Kl Ph (P split in two);               */
/* NPROD = 5: + Kh Pl + Kl Pm (P split in three); NPROD = 4: + Kh Pl only    */
/* (Kl Pm is a K rounding residual times a P one: independent per pair, it   */
/* averages out, where the P residuals, shared by every r1 row, do not).     */
/* Products are exact in FP32; the FP32 accumulator covers FOLD 32-point     */
/* k-steps and is then folded into FP64 sums (long FP32 runs lose the small  */
/* terms of the large-dynamic-range pair densities).  NPROD = 4 takes FOLD = */
/* 2 and issues each step's products smallest first (Kh Pl, Kl Ph, Kh Pm,   */
/* Kh Ph), so the corrections add at their own magnitude.  The three-       */
/* component V tiles (Jastrow gradient) take FOLD = 4 at every NPROD (FOLD  */
/* = 1 on launches above kVLongWork units): their fold is a third of their  */
/* VALU work at FOLD = 1, and V's error against FP64 does not move with it. */
/* Workgroup = 4 waves sharing 16*MB rows; wave w owns columns               */
/* [64 w, 64 w + 64) of the workgroup's 256.  Each k-step the workgroup      */
/* generates its K tile once, split hi/lo, into LDS in A-fragment order      */
/* (double buffered); B fragments and the next step's r2 coordinates are     */
/* prefetched into registers one step ahead.                                 */
/* ------------------------------------------------------------------------ */
// One workgroup tile of the BF16 contraction: rows row0 .. row0 + 16 MB of r1,
// 16-column blocks cbw .. cbw + 15 of the pairs, all of r2.  Output element
// (row, component c, pair col) goes to Fout[row * ld_row + c * npair + col].
// ktile: the workgroup's K tile (LDS, [2][NC][MB][2][64] fragments); crd: a
// two-slot LDS ring of r2 coordinates (32 points per step), filled two steps
// ahead by direct-to-LDS loads of wave 0 so no register waits on them.
template <int MB, int NPROD, int FOLD, class Kern, int NB = 4>
__device__ __forceinline__ void bf16_tile(us8 (*ktile)[Kern::ncomp][MB][2][64], float4 (*crd)[32],
                                          const double* __restrict__ p1,
                                          int64_t n1, const float4* __restrict__ x2, int64_t n2p,
                                          const us8* __restrict__ P0, int64_t plane, int64_t ncb,
                                          void* __restrict__ Fout, void* __restrict__ Fout2, int out_f32,
                                          int64_t npair, int64_t ld_row, double scale, Kern kern, int64_t row0,
                                          int64_t cbw, int64_t sb = 0, int64_t ns = -1,
                                          double* __restrict__ part = nullptr) {
  constexpr int NC = Kern::ncomp;
  static_assert(NB == 2 || NB == 4, "16-column blocks per wave: 2 or 4");
  constexpr int NSPLIT = NPROD == 3 ? 2 : 3;
  constexpr int VALS = MB * 64 * 8;       // K values per component per k-step
  constexpr int EPT = VALS / 256;         // per thread
  static_assert(EPT >= 2 && EPT <= 8 && VALS % 256 == 0 && EPT % 2 == 0, "MB must be 1, 2 or 4");
  static_assert(FOLD == 1 || FOLD == 2 || FOLD == 4, "FP64 fold every 1, 2 or 4 steps");
  static_assert(NPROD >= 3 && NPROD <= 5, "3, 4 or 5 products");
  constexpr bool REV = NPROD == 4;
  typedef __bf16 bf2 __attribute__((ext_vector_type(2)));
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int g = lane >> 4, r = lane & 15;
  const int64_t cb0 = cbw + wave * NB;

  // this thread generates elements ge .. ge+EPT-1 of fragment lane gl of m-block gm
  const int gi = tid * EPT;
  const int gm = gi / 512;
  const int gl = (gi / 8) & 63;
  const int ge = gi & 7;
  const int gk = 8 * (gl >> 4) + ge;
  float ax, ay, az;
  {
    const int64_t grow = row0 + 16 * gm + (gl & 15);
    const bool ok = grow < n1;
    ax = ok ? (float)p1[grow] : (float)-kFar;
    ay = ok ? (float)p1[n1 + grow] : (float)-kFar;
    az = ok ? (float)p1[2 * n1 + grow] : (float)-kFar;
  }

  // Coordinates of step t: lanes 0..31 of wave 0 load them into a register
  // (fetch_coords) two steps ahead and store them into crd[t & 1] before the
  // barrier of the step that issued the load (store_coords).  A plain load, so
  // the only wait is for that load: the B fragments loaded after it stay in
  // flight across the barrier.  (A direct-to-LDS load here made every later
  // LDS access wait for all outstanding loads, B included.)
  float4 cnext = make_float4(0.f, 0.f, 0.f, 0.f);
  auto fetch_coords = [&](int64_t t) {
    if (wave == 0 && lane < 32) cnext = x2[(sb + t) * 32 + lane];
  };
  auto store_coords = [&](int64_t t) {
    if (wave == 0 && lane < 32) crd[t & 1][lane] = cnext;
  };
  // K = hi + lo, both bf16 (round to nearest even, v_cvt_pk_bf16_f32), for the
  // step whose coordinates sit in crd[buf]
  auto generate = [&](int buf) {
    float4 qn[EPT];
#pragma unroll
    for (int e = 0; e < EPT; ++e) qn[e] = crd[buf][gk + e];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      unsigned hw[EPT / 2], lw[EPT / 2];
#pragma unroll
      for (int e = 0; e < EPT; e += 2) {
#pragma clang fp contract(off)
        float v0[NC], v1[NC];
        kernel_eval(kern, ax, ay, az, qn[e].x, qn[e].y, qn[e].z, v0);
        kernel_eval(kern, ax, ay, az, qn[e + 1].x, qn[e + 1].y, qn[e + 1].z, v1);
        bf2 h;
        h[0] = (__bf16)v0[c];
        h[1] = (__bf16)v1[c];
        const unsigned hb = __builtin_bit_cast(unsigned, h);
        bf2 l;
        l[0] = (__bf16)(v0[c] - __uint_as_float(hb << 16));
        l[1] = (__bf16)(v1[c] - __uint_as_float(hb & 0xffff0000u));
        hw[e / 2] = hb;
        lw[e / 2] = __builtin_bit_cast(unsigned, l);
      }
      unsigned* dh = (unsigned*)((unsigned short*)&ktile[buf][c][gm][0][gl] + ge);
      unsigned* dl = (unsigned*)((unsigned short*)&ktile[buf][c][gm][1][gl] + ge);
#pragma unroll
      for (int e = 0; e < EPT / 2; ++e) {
        dh[e] = hw[e];
        dl[e] = lw[e];
      }
    }
  };

  auto load_b = [&](us8 (&dst)[NSPLIT][NB], int64_t s) {
#pragma unroll
    for (int t = 0; t < NSPLIT; ++t)
#pragma unroll
      for (int n = 0; n < NB; ++n) dst[t][n] = P0[t * plane + ((sb + s) * ncb + cb0 + n) * 64 + lane];
  };

  double sum[NC][MB][NB][4];
#pragma unroll
  for (int c = 0; c < NC; ++c)
#pragma unroll
    for (int m = 0; m < MB; ++m)
#pragma unroll
      for (int n = 0; n < NB; ++n)
#pragma unroll
        for (int i = 0; i < 4; ++i) sum[c][m][n][i] = 0;
  f4 acc[NC][MB][NB];
  auto fold_all = [&] {
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
      for (int m = 0; m < MB; ++m)
#pragma unroll
        for (int n = 0; n < NB; ++n)
#pragma unroll
          for (int i = 0; i < 4; ++i) sum[c][m][n][i] += (double)acc[c][m][n][i];
  };

  const int64_t nsteps = ns < 0 ? n2p / 32 : ns;  // steps sb .. sb + nsteps - 1 of r2
  us8 bqA[NSPLIT][NB], bqB[NSPLIT][NB];  // B fragments, ping-pong between steps
  fetch_coords(0);
  store_coords(0);
  if (nsteps > 1) {
    fetch_coords(1);
    store_coords(1);
  }
  __syncthreads();
  generate(0);
  load_b(bqA, 0);
  __syncthreads();

  // One 32-point step: MFMAs on ktile[s & 1] and cur, B of step s + 1 into nxt,
  // K of step s + 1 into the other tile.  FRESH: the step starts a new FP32
  // partial (zero accumulator operand); the fold to FP64 follows every FOLD
  // steps.  The step between the coordinate fetch and the barrier is one
  // basic block: the loads past the last step re-read the last step and the K
  // generated after the last step is not used.  With no branch around the
  // loads, the wait before each MFMA counts only loads older than its operand,
  // so the B fragments of step s + 1 stay in flight through step s (behind a
  // branch, the join took the count of the path without them and every MFMA
  // waited for the loads just issued).  With several kernel components the
  // scheduler interleaves each wave's MFMAs with its generation and fold;
  // with one component (RUN) scheduling fences keep the MFMAs one run while
  // the other wave of the SIMD fills the VALU (measured faster for one
  // component, slower for three).
  constexpr bool RUN = NC == 1;
  auto fence = [] {
    if constexpr (RUN) __builtin_amdgcn_sched_barrier(0);
  };
  auto step = [&](int64_t s, us8 (&cur)[NSPLIT][NB], us8 (&nxt)[NSPLIT][NB], auto fresh, auto foldend) {
    constexpr bool FRESH = decltype(fresh)::value;
    constexpr bool FOLDEND = decltype(foldend)::value;
    const int buf = (int)(s & 1);
    fetch_coords(s + 2 < nsteps ? s + 2 : nsteps - 1);  // for crd[s & 1], read last by generate(s) one step ago
    load_b(nxt, s + 1 < nsteps ? s + 1 : s);
    fence();
    const f4 zero = {0, 0, 0, 0};
    // Product-major: every accumulator of a component takes one product before
    // any takes the next, so consecutive MFMAs are independent and each
    // accumulator's next product issues MB * NB MFMAs later (a chain per
    // accumulator stalls on its own result).  The per-element order of the
    // products (hh, hm, lh, hl, lm) is unchanged.
    auto mfma_all = [&](int c, const bf8 (&a)[MB], int sp, bool first) {
#pragma unroll
      for (int m = 0; m < MB; ++m)
#pragma unroll
        for (int n = 0; n < NB; ++n)
          acc[c][m][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[m], __builtin_bit_cast(bf8, cur[sp][n]),
                                                                 first ? zero : acc[c][m][n], 0, 0, 0);
    };
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      bf8 ah[MB], al[MB];
#pragma unroll
      for (int m = 0; m < MB; ++m) {
        ah[m] = __builtin_bit_cast(bf8, ktile[buf][c][m][0][lane]);
        al[m] = __builtin_bit_cast(bf8, ktile[buf][c][m][1][lane]);
      }
      // products b = 0 hh, 1 hm, 2 lh, 3 hl, 4 lm; np of them (folded at
      // compile time: c and j are unrolled), smallest first for NPROD = 4
      const int np = (NPROD >= 4 && c < Comp5<Kern>::value) ? NPROD : 3;
#pragma unroll
      for (int j = 0; j < 5; ++j)
        if (j < np) {
          const int b = REV ? np - 1 - j : j;
          const int sp = (b == 0 || b == 2) ? 0 : (b == 3 ? NSPLIT - 1 : 1);
          if (b == 2 || b == 4) mfma_all(c, al, sp, FRESH && j == 0);
          else mfma_all(c, ah, sp, FRESH && j == 0);
        }
    }
    fence();
    generate(buf ^ 1);
    store_coords(s + 2);
    fence();
    if constexpr (FOLDEND) fold_all();
    __syncthreads();
  };
  using T = std::integral_constant<bool, true>;
  using Fb = std::integral_constant<bool, false>;
  using Fz = std::integral_constant<bool, FOLD == 1>;
  if constexpr (FOLD == 4) {
    for (int64_t s = 0; s < nsteps; s += 4) {
      step(s, bqA, bqB, T{}, Fb{});
      if (s + 1 < nsteps) step(s + 1, bqB, bqA, Fb{}, Fb{});
      if (s + 2 < nsteps) step(s + 2, bqA, bqB, Fb{}, Fb{});
      if (s + 3 < nsteps) step(s + 3, bqB, bqA, Fb{}, T{});
    }
    if (nsteps & 3) fold_all();  // the last, incomplete group of four
  } else {
    for (int64_t s = 0; s < nsteps; s += 2) {
      step(s, bqA, bqB, T{}, Fz{});
      if (s + 1 < nsteps) step(s + 1, bqB, bqA, Fz{}, T{});
    }
    if (FOLD == 2 && (nsteps & 1)) fold_all();  // the last, unpaired step
  }
  if (part) {
    // a piece of a split tile: the raw FP64 sums, [c][16 MB rows][64 NB cols]
    constexpr int TR = 16 * MB, TC = 64 * NB;
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
      for (int m = 0; m < MB; ++m)
#pragma unroll
        for (int n = 0; n < NB; ++n)
#pragma unroll
          for (int i = 0; i < 4; ++i)
            part[((int64_t)c * TR + 16 * m + d_row_bf16(g, i)) * TC + 16 * (wave * NB + n) + r] = sum[c][m][n][i];
    return;
  }
  if constexpr (Combines<Kern>::value) {
    // outputs o = 0, 1 of kern.finish to Fout / Fout2, row-major n1 x npair each
    static_assert(Kern::nout == 2, "two outputs");
#pragma unroll
    for (int m = 0; m < MB; ++m)
#pragma unroll
      for (int n = 0; n < NB; ++n)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int64_t row = row0 + 16 * m + d_row_bf16(g, i), col = 16 * (cb0 + n) + r;
          if (row < n1 && col < npair) {
            double v[NC], o[2];
#pragma unroll
            for (int c = 0; c < NC; ++c) v[c] = scale * sum[c][m][n][i];
            kern.finish(v, o);
            const int64_t e = row * ld_row + col;
            if (out_f32) {
              ((float*)Fout)[e] = (float)o[0];
              ((float*)Fout2)[e] = (float)o[1];
            } else {
              ((double*)Fout)[e] = o[0];
              ((double*)Fout2)[e] = o[1];
            }
          }
        }
  } else {
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
      for (int m = 0; m < MB; ++m)
#pragma unroll
        for (int n = 0; n < NB; ++n)
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int64_t row = row0 + 16 * m + d_row_bf16(g, i), col = 16 * (cb0 + n) + r;
            if (row < n1 && col < npair) {
              const int64_t o = row * ld_row + c * npair + col;
              const double v = scale * sum[c][m][n][i];
              if (out_f32) ((float*)Fout)[o] = (float)v;
              else ((double*)Fout)[o] = v;
            }
          }
  }
}

/* ------------------------------------------------------------------------ */
/* Software-pipelined BF16 tile (FP64 fold after every step)                 */
/* ------------------------------------------------------------------------ */
// The tile, products and rounding of bf16_tile, with each 32-point step laid
// out so that its VALU work issues between MFMAs instead of after them:
//   - step s folds step s - 1's FP32 partials into the FP64 sums accumulator
//     by accumulator, each fold just before that accumulator's fresh product
//     (a register dependency on the fold's result keeps that order);
//   - the K tiles of step s + 1 are generated between step s's other
//     products;
//   - B fragments come through a raw buffer view whose scalar offset carries
//     the step and the split plane, so no lane computes an address;
//   - every wave loads and stores the 32 r2 coordinates of a step (the same
//     values), so a step has no branch and is one scheduling region.
// Workgroup: NW waves over 16 MB rows; wave w owns the 16-column blocks
// cbw + w NB .. + NB - 1 of the pairs.  Per accumulator and element the
// product order (hh, hm, lh, then hl, lm for the components that take five
// products) and the fold after every step are those of bf16_tile: the results
// are bitwise equal.
template <int MB, int NPROD, class Kern, int NB, int NW, int FOLD = 1, bool REV = false>
__device__ __forceinline__ void bf16_tile_pipe(us8 (*ktile)[Kern::ncomp][MB][2][64], float4 (*crd)[32],
                                               const double* __restrict__ p1, int64_t n1,
                                               const float4* __restrict__ x2, int64_t n2p,
                                               const us8* __restrict__ P0, int64_t plane, int64_t ncb,
                                               void* __restrict__ Fout, void* __restrict__ Fout2, int out_f32,
                                               int64_t npair, int64_t ld_row, double scale, Kern kern, int64_t row0,
                                               int64_t cbw, int64_t sb, int64_t ns, double* __restrict__ part) {
  constexpr int NC = Kern::ncomp;
  constexpr int C5 = NPROD >= 4 ? Comp5<Kern>::value : 0;  // components with NPROD (4 or 5) products
  constexpr int NSPLIT = NPROD == 3 ? 2 : 3;
  constexpr int VALS = MB * 64 * 8;     // K values per component per k-step
  constexpr int EPT = VALS / (64 * NW);  // per thread
  static_assert(EPT >= 2 && EPT <= 8 && VALS % (64 * NW) == 0 && EPT % 2 == 0, "K values per thread: 2 .. 8");
  constexpr int NACC = NC * MB * NB;
  constexpr int NROUND = NC * 2 + C5 * (NPROD - 3);  // product rounds after the fresh one
  typedef __bf16 bf2 __attribute__((ext_vector_type(2)));
  typedef unsigned u4 __attribute__((ext_vector_type(4)));
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int g = lane >> 4, r = lane & 15;
  const int64_t cb0 = cbw + wave * NB;

  // this thread generates elements ge .. ge+EPT-1 of fragment lane gl of m-block gm
  const int gi = tid * EPT;
  const int gm = gi / 512;
  const int gl = (gi / 8) & 63;
  const int ge = gi & 7;
  const int gk = 8 * (gl >> 4) + ge;
  float ax, ay, az;
  {
    const int64_t grow = row0 + 16 * gm + (gl & 15);
    const bool ok = grow < n1;
    ax = ok ? (float)p1[grow] : (float)-kFar;
    ay = ok ? (float)p1[n1 + grow] : (float)-kFar;
    az = ok ? (float)p1[2 * n1 + grow] : (float)-kFar;
  }

  // B planes: lane and the wave's first column block in the vector offset,
  // block n in the instruction offset (n * 1 KiB), step and plane in the
  // scalar offset.  Every offset lies inside the planes by construction.
  const __amdgpu_buffer_rsrc_t pv = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<us8*>(P0), (short)0, (int)((int64_t)NSPLIT * plane * 16), kBufferWord3);
  const int voff = (int)(cb0 * 1024) + lane * 16;
  const int plane_bytes = (int)(plane * 16);
  auto load_b = [&](us8 (&dst)[NSPLIT][NB], int64_t s) {
    const int so = __builtin_amdgcn_readfirstlane((int)((sb + s) * ncb * 1024));
#pragma unroll
    for (int t = 0; t < NSPLIT; ++t)
#pragma unroll
      for (int n = 0; n < NB; ++n)
        dst[t][n] = __builtin_bit_cast(us8, __builtin_amdgcn_raw_buffer_load_b128(pv, voff + n * 1024,
                                                                                   so + t * plane_bytes, 0));
  };
  // r2 coordinates: lanes l and l + 32 of every wave load point l of a step
  const __amdgpu_buffer_rsrc_t cv =
      __builtin_amdgcn_make_buffer_rsrc(const_cast<float4*>(x2), (short)0, (int)(n2p * 16), kBufferWord3);
  const int coff = (lane & 31) * 16;
  u4 cnext;
  auto fetch_coords = [&](int64_t t) {
    cnext = __builtin_amdgcn_raw_buffer_load_b128(cv, coff, __builtin_amdgcn_readfirstlane((int)((sb + t) * 512)), 0);
  };
  auto store_coords = [&](int64_t t) { crd[t & 1][lane & 31] = __builtin_bit_cast(float4, cnext); };
  // K = hi + lo (bf16, round to nearest even) of every component at points
  // qn[e], qn[e + 1]: the hi and lo words of the pair
  auto gen_pair = [&](const float4 (&qn)[EPT], int e, unsigned (&hw)[NC], unsigned (&lw)[NC]) {
#pragma clang fp contract(off)
    float v0[NC], v1[NC];
    kernel_eval(kern, ax, ay, az, qn[e].x, qn[e].y, qn[e].z, v0);
    kernel_eval(kern, ax, ay, az, qn[e + 1].x, qn[e + 1].y, qn[e + 1].z, v1);
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      bf2 h;
      h[0] = (__bf16)v0[c];
      h[1] = (__bf16)v1[c];
      const unsigned hb = __builtin_bit_cast(unsigned, h);
      bf2 l;
      l[0] = (__bf16)(v0[c] - __uint_as_float(hb << 16));
      l[1] = (__bf16)(v1[c] - __uint_as_float(hb & 0xffff0000u));
      hw[c] = hb;
      lw[c] = __builtin_bit_cast(unsigned, l);
    }
  };
  auto store_k = [&](const unsigned (&hw)[EPT / 2][NC], const unsigned (&lw)[EPT / 2][NC], int buf) {
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      unsigned* dh = (unsigned*)((unsigned short*)&ktile[buf][c][gm][0][gl] + ge);
      unsigned* dl = (unsigned*)((unsigned short*)&ktile[buf][c][gm][1][gl] + ge);
#pragma unroll
      for (int e = 0; e < EPT / 2; ++e) {
        dh[e] = hw[e][c];
        dl[e] = lw[e][c];
      }
    }
  };

  double sum[NC][MB][NB][4];
#pragma unroll
  for (int c = 0; c < NC; ++c)
#pragma unroll
    for (int m = 0; m < MB; ++m)
#pragma unroll
      for (int n = 0; n < NB; ++n)
#pragma unroll
        for (int i = 0; i < 4; ++i) sum[c][m][n][i] = 0;
  f4 acc[NC][MB][NB];
  const int64_t nsteps = ns, last = nsteps - 1;  // steps sb .. sb + ns - 1 of r2
  us8 bqA[NSPLIT][NB], bqB[NSPLIT][NB];  // B fragments, ping-pong between steps
  // prologue: coordinates of steps 0 and 1 in the ring, step 2's in cnext,
  // K(0) in ktile[0], B(0) in bqA
  fetch_coords(0);
  store_coords(0);
  fetch_coords(last < 1 ? last : 1);
  store_coords(1);
  fetch_coords(last < 2 ? last : 2);
  __syncthreads();
  {
    float4 q0[EPT];
#pragma unroll
    for (int e = 0; e < EPT; ++e) q0[e] = crd[0][gk + e];
    unsigned hw[EPT / 2][NC], lw[EPT / 2][NC];
#pragma unroll
    for (int e = 0; e < EPT; e += 2) gen_pair(q0, e, hw[e / 2], lw[e / 2]);
    store_k(hw, lw, 0);
  }
  load_b(bqA, 0);
  __syncthreads();

  // Product j of a component with np products: the K part (0 hi, 1 lo) and
  // the P plane.  The base order is hh, hm, lh, hl, lm; REV issues it
  // backwards, the small corrections first into a fresh partial and hh last,
  // so the corrections round at their own magnitude.
  auto kpart = [](int np, int j) {
    const int b = REV ? np - 1 - j : j;
    return (b == 2 || b == 4) ? 1 : 0;
  };
  auto plane_of = [](int np, int j) {
    const int b = REV ? np - 1 - j : j;
    return b == 0 || b == 2 ? 0 : (b == 3 ? NSPLIT - 1 : 1);
  };
  // zero accumulator operand of the fresh products after the first step: an
  // empty asm statement redefines it after each accumulator's fold (it stays
  // zero), a register dependency that keeps each fold before the product
  // that overwrites its accumulator
  f4 zc = {0, 0, 0, 0};
  // FRESH: the step starts a new FP32 partial (every FOLD steps); FOLDP: the
  // previous partial is folded first (a fresh step other than the first)
  auto step = [&](int64_t s, us8 (&cur)[NSPLIT][NB], us8 (&nxt)[NSPLIT][NB], auto fresh, auto foldp) {
    constexpr bool FRESH = decltype(fresh)::value;
    constexpr bool FOLDP = decltype(foldp)::value;
    const int buf = (int)(s & 1);
    // --- region 1: the ring store, the loads (coordinates first, so waiting
    // for them one step later leaves the B loads in flight), the LDS reads
    store_coords(s + 2);  // ring slot s & 1 was last read for K(s), before the previous barrier
    fetch_coords(s + 3 <= last ? s + 3 : last);
    __builtin_amdgcn_sched_barrier(0);
    load_b(nxt, s + 1 <= last ? s + 1 : last);
    bf8 ak[NC][2][MB];  // K hi / lo fragments
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
      for (int m = 0; m < MB; ++m) {
        ak[c][0][m] = __builtin_bit_cast(bf8, ktile[buf][c][m][0][lane]);
        ak[c][1][m] = __builtin_bit_cast(bf8, ktile[buf][c][m][1][lane]);
      }
    float4 qn[EPT];  // coordinates of step s + 1
#pragma unroll
    for (int e = 0; e < EPT; ++e) qn[e] = crd[buf ^ 1][gk + e];
    __builtin_amdgcn_sched_barrier(0);
    // --- region 2: per accumulator the first product of the step, after the
    // fold of the previous partial (8 VALU) in a step that starts a new one
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const int np = c < C5 ? NPROD : (NPROD >= 4 ? 3 : NPROD);
      const int kp = kpart(np, 0), sp = plane_of(np, 0);
#pragma unroll
      for (int m = 0; m < MB; ++m)
#pragma unroll
        for (int n = 0; n < NB; ++n) {
          const bf8 b = __builtin_bit_cast(bf8, cur[sp][n]);
          if constexpr (!FRESH) {
            acc[c][m][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ak[c][kp][m], b, acc[c][m][n], 0, 0, 0);
          } else if constexpr (!FOLDP) {
            const f4 zero = {0, 0, 0, 0};
            acc[c][m][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ak[c][kp][m], b, zero, 0, 0, 0);
          } else {
#pragma unroll
            for (int i = 0; i < 4; ++i) sum[c][m][n][i] += (double)acc[c][m][n][i];
            asm("; fold %1" : "+v"(zc) : "v"(sum[c][m][n][3]));
            acc[c][m][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ak[c][kp][m], b, zc, 0, 0, 0);
          }
        }
    }
#pragma unroll
    for (int k = 0; k < NACC; ++k) {
      if constexpr (FOLDP) __builtin_amdgcn_sched_group_barrier(0x002, 8, 0);
      __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
    }
    __builtin_amdgcn_sched_barrier(0);
    // --- region 3: the other products, component by component, with the K
    // pairs of step s + 1 spread through them in source order (which the
    // group solver follows) two VALU per MFMA, the K tile stores last
    unsigned hw[EPT / 2][NC], lw[EPT / 2][NC];
    auto gen_after = [&](int rd) {
#pragma unroll
      for (int e = 0; e < EPT; e += 2)
        if ((e / 2) * NROUND / (EPT / 2) == rd) gen_pair(qn, e, hw[e / 2], lw[e / 2]);
    };
    int rd = 0;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const int np = c < C5 ? NPROD : (NPROD >= 4 ? 3 : NPROD);
#pragma unroll
      for (int j = 1; j < 5; ++j)
        if (j < np) {  // folded at compile time (c and j are unrolled)
          const int kp = kpart(np, j), sp = plane_of(np, j);
#pragma unroll
          for (int m = 0; m < MB; ++m)
#pragma unroll
            for (int n = 0; n < NB; ++n)
              acc[c][m][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                  ak[c][kp][m], __builtin_bit_cast(bf8, cur[sp][n]), acc[c][m][n], 0, 0, 0);
          gen_after(rd++);
        }
    }
    store_k(hw, lw, buf ^ 1);
#pragma unroll
    for (int k = 0; k < NROUND * MB * NB; ++k) {
      __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
      __builtin_amdgcn_sched_group_barrier(0x002, 2, 0);
    }
    __builtin_amdgcn_sched_barrier(0);
    __syncthreads();
  };
  using Tt = std::integral_constant<bool, true>;
  using Ff = std::integral_constant<bool, false>;
  using Odd = std::integral_constant<bool, FOLD == 1>;  // odd steps start a partial only at FOLD 1
  step(0, bqA, bqB, Tt{}, Ff{});
  int64_t s = 1;
  for (; s + 1 < nsteps; s += 2) {
    step(s, bqB, bqA, Odd{}, Odd{});
    step(s + 1, bqA, bqB, Tt{}, Tt{});
  }
  if (s < nsteps) step(s, bqB, bqA, Odd{}, Odd{});
#pragma unroll
  for (int c = 0; c < NC; ++c)
#pragma unroll
    for (int m = 0; m < MB; ++m)
#pragma unroll
      for (int n = 0; n < NB; ++n)
#pragma unroll
        for (int i = 0; i < 4; ++i) sum[c][m][n][i] += (double)acc[c][m][n][i];
  if (part) {
    // a piece of a split tile: the raw FP64 sums, [c][16 MB rows][16 NW NB cols]
    constexpr int TR = 16 * MB, TC = 16 * NW * NB;
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
      for (int m = 0; m < MB; ++m)
#pragma unroll
        for (int n = 0; n < NB; ++n)
#pragma unroll
          for (int i = 0; i < 4; ++i)
            part[((int64_t)c * TR + 16 * m + d_row_bf16(g, i)) * TC + 16 * (wave * NB + n) + r] = sum[c][m][n][i];
    return;
  }
#pragma unroll
  for (int m = 0; m < MB; ++m)
#pragma unroll
    for (int n = 0; n < NB; ++n)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int64_t row = row0 + 16 * m + d_row_bf16(g, i), col = 16 * (cb0 + n) + r;
        if (row < n1 && col < npair) {
          if constexpr (Combines<Kern>::value) {
            // outputs o = 0, 1 of kern.finish to Fout / Fout2, row-major n1 x npair each
            static_assert(Kern::nout == 2, "two outputs");
            double v[NC], o[2];
#pragma unroll
            for (int c = 0; c < NC; ++c) v[c] = scale * sum[c][m][n][i];
            kern.finish(v, o);
            const int64_t e = row * ld_row + col;
            if (out_f32) {
              ((float*)Fout)[e] = (float)o[0];
              ((float*)Fout2)[e] = (float)o[1];
            } else {
              ((double*)Fout)[e] = o[0];
              ((double*)Fout2)[e] = o[1];
            }
          } else {
#pragma unroll
            for (int c = 0; c < NC; ++c) {
              const int64_t o = row * ld_row + c * npair + col;
              const double v = scale * sum[c][m][n][i];
              if (out_f32) ((float*)Fout)[o] = (float)v;
              else ((double*)Fout)[o] = v;
            }
          }
        }
      }
}

// NW waves per workgroup (256 threads: two workgroups per CU; 512: one)
// Workgroup t = blockIdx.x is tile t: column group t % gx, row block t / gx.
template <int MB, int NPROD, class Kern, int NB, int NW, int FOLD = 1, bool REV = false>
__global__ __launch_bounds__(64 * NW, 8 / NW) void fgemm_bf16_pipe_kernel(
    const double* __restrict__ p1, int64_t n1, const float4* __restrict__ x2, int64_t n2p,
    const us8* __restrict__ P0, int64_t plane, int64_t ncb, void* __restrict__ Fout, void* __restrict__ Fout2,
    int out_f32, int64_t npair, double scale, Kern kern, int64_t gx) {
  __shared__ us8 ktile[2][Kern::ncomp][MB][2][64];
  __shared__ float4 crd[2][32];
  const int64_t t = blockIdx.x;
  bf16_tile_pipe<MB, NPROD, Kern, NB, NW, FOLD, REV>(ktile, crd, p1, n1, x2, n2p, P0, plane, ncb, Fout, Fout2, out_f32,
                                                     npair, Combines<Kern>::value ? npair : Kern::ncomp * npair,
                                                     scale, kern, (t / gx) * (16 * MB), (t % gx) * (NW * NB), 0,
                                                     n2p / 32, nullptr);
}

// Split tail: the tiles t0 .. t0 + W / nst - 1 (the last, partly filled round of
// a launch) cut into gridDim.x runs of r2 steps, one per workgroup, over the
// tail's steps in tile order.  Run w covers tail steps [b(w), b(w + 1)), b(w)
// = w W / gridDim.x moved down to a multiple of four steps of its tile, so
// each piece's FP32 partials (two or four steps each) group the steps as the
// unsplit tile does.  A run is
// at most one tile's worth, so at most two pieces (the end of one tile, the
// start of the next); piece k of run w leaves its raw FP64 sums in part slot
// 2 w + k for bf16_pipe_fixup.
__host__ __device__ inline int64_t split_bound(int64_t w, int64_t W, int64_t S, int64_t nst) {
  const int64_t b = w * W / S, t0 = b / nst * nst;
  return t0 + (b - t0) / 4 * 4;
}

template <int MB, int NPROD, class Kern, int NB, int NW, int FOLD = 1, bool REV = false>
__global__ __launch_bounds__(64 * NW, 8 / NW) void fgemm_bf16_pipe_tail_kernel(
    const double* __restrict__ p1, int64_t n1, const float4* __restrict__ x2, int64_t n2p,
    const us8* __restrict__ P0, int64_t plane, int64_t ncb, int64_t npair, Kern kern, int64_t gx, int64_t t0,
    int64_t W, double* __restrict__ part) {
  __shared__ us8 ktile[2][Kern::ncomp][MB][2][64];
  __shared__ float4 crd[2][32];
  constexpr int64_t TE = (int64_t)Kern::ncomp * 16 * MB * 16 * NW * NB;  // doubles per piece
  const int64_t nst = n2p / 32, S = gridDim.x, w = blockIdx.x;
  const int64_t g1 = split_bound(w + 1, W, S, nst);
  int k = 0;
  for (int64_t g = split_bound(w, W, S, nst); g < g1 && k < 2; ++k) {
    const int64_t tr = g / nst, s0 = g - tr * nst, s1 = g1 - tr * nst < nst ? g1 - tr * nst : nst;
    const int64_t t = t0 + tr;
    if (k) __syncthreads();  // the previous piece's LDS reads are done
    bf16_tile_pipe<MB, NPROD, Kern, NB, NW, FOLD, REV>(ktile, crd, p1, n1, x2, n2p, P0, plane, ncb, nullptr, nullptr,
                                                       0, npair, 0, 1.0, kern, (t / gx) * (16 * MB),
                                                       (t % gx) * (NW * NB), s0, s1 - s0, part + (2 * w + k) * TE);
    g = tr * nst + s1;
  }
}

// The split tiles' outputs: per element the pieces' FP64 sums added in step
// order, then the epilogue of bf16_tile_pipe (scale, combine, storage type).
// Grid: (element chunks of 256, tail tiles).
template <class Kern, int TR, int TC>
__global__ __launch_bounds__(256) void bf16_pipe_fixup(const double* __restrict__ part, int64_t n1, int64_t npair,
                                                       int64_t gx, int64_t t0, int64_t nst, int64_t W, int64_t S,
                                                       void* __restrict__ Fout, void* __restrict__ Fout2, int out_f32,
                                                       double scale, Kern kern) {
  constexpr int NC = Kern::ncomp;
  constexpr int64_t TE = (int64_t)NC * TR * TC;
  const int64_t tr = blockIdx.y, t = t0 + tr;
  const int e = blockIdx.x * 256 + threadIdx.x;  // element (row_local, col_local)
  if (e >= TR * TC) return;
  const int rl = e / TC, cl = e - rl * TC;
  const int64_t row = (t / gx) * TR + rl, col = (t % gx) * TC + cl;
  if (row >= n1 || col >= npair) return;
  // the runs that cover tile tr: from the first w with b(w + 1) > tr nst
  const int64_t lo = tr * nst, hi = lo + nst;
  int64_t w = lo * S / W;
  while (w > 0 && split_bound(w, W, S, nst) > lo) --w;
  while (split_bound(w + 1, W, S, nst) <= lo) ++w;
  double v[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) v[c] = 0;
  for (; w < S && split_bound(w, W, S, nst) < hi; ++w) {
    const int64_t k = tr - split_bound(w, W, S, nst) / nst;  // piece of run w that lies in tile tr
    const double* pc = part + (2 * w + k) * TE;
#pragma unroll
    for (int c = 0; c < NC; ++c) v[c] += pc[((int64_t)c * TR + rl) * TC + cl];
  }
  if constexpr (Combines<Kern>::value) {
    double x[NC], o[2];
#pragma unroll
    for (int c = 0; c < NC; ++c) x[c] = scale * v[c];
    kern.finish(x, o);
    const int64_t e2 = row * npair + col;
    if (out_f32) {
      ((float*)Fout)[e2] = (float)o[0];
      ((float*)Fout2)[e2] = (float)o[1];
    } else {
      ((double*)Fout)[e2] = o[0];
      ((double*)Fout2)[e2] = o[1];
    }
  } else {
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const int64_t o = row * (NC * npair) + c * npair + col;
      if (out_f32) ((float*)Fout)[o] = (float)(scale * v[c]);
      else ((double*)Fout)[o] = scale * v[c];
    }
  }
}

// Two workgroups per CU (a 256-VGPR budget) where the kernel fits it without
// spills: one component, the three-product split, or a mixed split.
template <class Kern, int NPROD>
constexpr int bf16_min_blocks() {
  return (Kern::ncomp == 1 || NPROD == 3 || Comp5<Kern>::value < Kern::ncomp) ? 2 : 1;
}

// Workgroup t = blockIdx.x is tile t: column group t % gx, row block t / gx.
template <int MB, int NPROD, int FOLD, class Kern, int NB>
__global__ __launch_bounds__(256, (bf16_min_blocks<Kern, NPROD>())) void fgemm_bf16_kernel(
    const double* __restrict__ p1, int64_t n1, const float4* __restrict__ x2, int64_t n2p,
    const us8* __restrict__ P0, int64_t plane, int64_t ncb, void* __restrict__ Fout, void* __restrict__ Fout2,
    int out_f32, int64_t npair, double scale, Kern kern, int64_t gx) {
  __shared__ us8 ktile[2][Kern::ncomp][MB][2][64];
  __shared__ float4 crd[2][32];
  const int64_t t = blockIdx.x;
  bf16_tile<MB, NPROD, FOLD, Kern, NB>(ktile, crd, p1, n1, x2, n2p, P0, plane, ncb, Fout, Fout2, out_f32, npair,
                                       Combines<Kern>::value ? npair : Kern::ncomp * npair, scale, kern,
                                       (t / gx) * (16 * MB), (t % gx) * (4 * NB));
}

// The split tail of fgemm_bf16_kernel (runs and pieces as for the pipelined
// kernel's fgemm_bf16_pipe_tail_kernel).
template <int MB, int NPROD, int FOLD, class Kern, int NB>
__global__ __launch_bounds__(256, (bf16_min_blocks<Kern, NPROD>())) void fgemm_bf16_tail_kernel(
    const double* __restrict__ p1, int64_t n1, const float4* __restrict__ x2, int64_t n2p,
    const us8* __restrict__ P0, int64_t plane, int64_t ncb, int64_t npair, Kern kern, int64_t gx, int64_t t0,
    int64_t W, double* __restrict__ part) {
  __shared__ us8 ktile[2][Kern::ncomp][MB][2][64];
  __shared__ float4 crd[2][32];
  constexpr int64_t TE = (int64_t)Kern::ncomp * 16 * MB * 64 * NB;  // doubles per piece
  const int64_t nst = n2p / 32, S = gridDim.x, w = blockIdx.x;
  const int64_t g1 = split_bound(w + 1, W, S, nst);
  int k = 0;
  for (int64_t g = split_bound(w, W, S, nst); g < g1 && k < 2; ++k) {
    const int64_t tr = g / nst, s0 = g - tr * nst, s1 = g1 - tr * nst < nst ? g1 - tr * nst : nst;
    const int64_t t = t0 + tr;
    if (k) __syncthreads();  // the previous piece's LDS reads are done
    bf16_tile<MB, NPROD, FOLD, Kern, NB>(ktile, crd, p1, n1, x2, n2p, P0, plane, ncb, nullptr, nullptr, 0, npair, 0,
                                         1.0, kern, (t / gx) * (16 * MB), (t % gx) * (4 * NB), s0, s1 - s0,
                                         part + (2 * w + k) * TE);
    g = tr * nst + s1;
  }
}

/* ------------------------------------------------------------------------ */
/* MFMA fragment self-test: D = A B for one fragment of each instruction.    */
/* A, B, D given row-major as FP64 on the host side of the test.             */
/* ------------------------------------------------------------------------ */
__global__ void mfma_selftest_kernel(int which, const double* A, const double* B, double* D) {
  const int l = threadIdx.x, g = l >> 4, r = l & 15;
  if (which == 0) {
    d4 c = {0, 0, 0, 0};
    c = __builtin_amdgcn_mfma_f64_16x16x4f64(A[r * 4 + g], B[g * 16 + r], c, 0, 0, 0);
    for (int i = 0; i < 4; ++i) D[l * 4 + i] = c[i];
  } else if (which == 1) {
    f4 c = {0, 0, 0, 0};
    c = __builtin_amdgcn_mfma_f32_16x16x4f32((float)A[r * 4 + g], (float)B[g * 16 + r], c, 0, 0, 0);
    for (int i = 0; i < 4; ++i) D[l * 4 + i] = c[i];
  } else {
    us8 a, b;
    for (int e = 0; e < 8; ++e) {
      a[e] = bf16_bits_rne((float)A[r * 32 + 8 * g + e]);
      b[e] = bf16_bits_rne((float)B[(8 * g + e) * 16 + r]);
    }
    f4 c = {0, 0, 0, 0};
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8, a), __builtin_bit_cast(bf8, b), c, 0, 0, 0);
    for (int i = 0; i < 4; ++i) D[l * 4 + i] = c[i];
  }
}

}  // namespace tc

/* ======================================================================== */
/* C ABI                                                                      */
/* ======================================================================== */
using namespace tc;

extern "C" size_t tcscf_F_workspace_bytes(int64_t n2, int nb, int precision);

namespace {

template <int MB, class Kern>
int launch_f64(const double* p1, int64_t n1, const double* x2, int64_t n2p, const double* P, int64_t ldp, void* F,
               int out_f32, int64_t npair, double scale, Kern kern, hipStream_t st) {
  dim3 grid((unsigned)(ldp / 128), (unsigned)((n1 + 32 * MB - 1) / (32 * MB)));
  hipLaunchKernelGGL((fgemm_f64_kernel<MB, Kern>), grid, dim3(256), 0, st, p1, n1, x2, n2p, P, ldp, F, out_f32, npair,
                     scale, kern);
  TC_HIP(hipGetLastError());
  return 0;
}

template <int MB, class Kern>
int launch_f32(const double* p1, int64_t n1, const float4* x2, int64_t n2p, const float* P, int64_t ldp, void* F,
               int out_f32, int64_t npair, double scale, Kern kern, hipStream_t st) {
  dim3 grid((unsigned)(ldp / 128), (unsigned)((n1 + 32 * MB - 1) / (32 * MB)));
  hipLaunchKernelGGL((fgemm_f32_kernel<MB, Kern>), grid, dim3(256), 0, st, p1, n1, x2, n2p, P, ldp, F, out_f32, npair,
                     scale, kern);
  TC_HIP(hipGetLastError());
  return 0;
}

// Layout of the software-pipelined kernel per kernel functor (TCSCF_BF16_PIPE
// selects another for A/B timing; 0 = bf16_tile).  One component (the
// Coulomb kernel): 64 rows x 256 columns per 8-wave workgroup (mode 2, each B
// fragment feeds 64 rows: half the vector-memory traffic of 32-row tiles at
// the K values per thread of the 4-wave tile; Ne 5.07 -> 4.53 ms); 1 = 32 x
// 256 per 4 waves, 3 = 64 x 128 per 4 waves.  Two components (the F pair):
// 32 x 256 per 8 waves (mode 1; TC-Ne 9.45 -> 7.80 ms), 2 = 64 x 128 per 8
// waves.  Three (V): bf16_tile, which measured faster (3.27 ms) than the one
// spill-free pipelined layout, 32 x 128 per 8 waves (3.95 ms, mode 1).
// Kernels whose functor evaluates behind a branch keep bf16_tile.
int bf16_pipe_mode(int ncomp) {
  static const int env = [] {
    const char* e = std::getenv("TCSCF_BF16_PIPE");
    return e ? std::atoi(e) : -1;
  }();
  if (env >= 0) return env;
  return ncomp == 1 ? 2 : (ncomp == 2 ? 1 : 0);
}

// Set while a caller-provided workspace is in use (the no-allocation entry
// point): the split tail's partial sums would need library scratch.
thread_local bool t_fixed_ws = false;

// Resident workgroups of kernel fn on the current device (CUs x workgroups per
// CU), kept per (device, kernel).
int resident_slots(const void* fn, int block) {
  static std::mutex mu;
  static std::map<std::pair<int, const void*>, int> slots;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  std::lock_guard<std::mutex> lock(mu);
  auto f = slots.find({dev, fn});
  if (f != slots.end()) return f->second;
  int ncu = 0, per = 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn, block, 0) != hipSuccess)
    return 0;
  return slots[{dev, fn}] = ncu * per;
}

// The resident slots to split the last round of a launch of kernel fn over,
// or 0 for no split: the split is off (TCSCF_BF16_SPLIT=0, a caller-provided
// workspace), the launch fits one round, or its last round is over 90% full
// or has under four r2 steps per slot.
int64_t split_slots(const void* fn, int block, int64_t tiles, int64_t nst) {
  const char* env = std::getenv("TCSCF_BF16_SPLIT");
  if ((env && std::atoi(env) == 0) || t_fixed_ws) return 0;
  const int64_t S = resident_slots(fn, block);
  const int64_t r = S ? tiles % S : 0;
  return S && tiles > S && r > 0 && r * 10 <= S * 9 && r * nst >= 4 * S ? S : 0;
}

// The launch's tiles run in rounds of the resident slots; a last round of r
// tiles (r < slots) leaves slots - r of them idle for a whole tile.  Unless
// TCSCF_BF16_SPLIT=0, that round is split: the full rounds run as usual, the
// last r tiles are cut along r2 into one equal run per slot (at most two
// pieces each, partial sums in library scratch) and a fix-up pass adds each
// tile's pieces in step order.  The split tiles' FP64 sums then associate
// differently from an unsplit tile's (last-bit differences); the tiles of the
// full rounds are unchanged.
template <int MB, int NB, int NW, int NPROD, class Kern, int FOLD = 1, bool REV = false>
int launch_bf16_pipe(const double* p1, int64_t n1, const float4* x2, int64_t n2p, const us8* P0, int64_t plane,
                     int64_t ncb, void* F, int out_f32, int64_t npair, double scale, Kern kern, hipStream_t st,
                     void* F2) {
  constexpr int NC = Kern::ncomp, TR = 16 * MB, TC = 16 * NW * NB;
  const int64_t gx = ncb / (NW * NB), gy = (n1 + TR - 1) / TR, tiles = gx * gy, nst = n2p / 32;
  const int64_t S = split_slots((const void*)fgemm_bf16_pipe_kernel<MB, NPROD, Kern, NB, NW, FOLD, REV>, 64 * NW,
                                tiles, nst);
  const int64_t r = S ? tiles % S : 0, full = tiles - r;
  const bool split = S > 0;
  double* part = nullptr;
  if (split) {
    int rc = 0;
    part = (double*)scratch(kSlotSplit, (size_t)(2 * S) * NC * TR * TC * sizeof(double), &rc);
    if (rc) return rc;
  }
  const int64_t main_tiles = split ? full : tiles;
  if (main_tiles > 0) {
    hipLaunchKernelGGL((fgemm_bf16_pipe_kernel<MB, NPROD, Kern, NB, NW, FOLD, REV>), dim3((unsigned)main_tiles),
                       dim3(64 * NW), 0, st, p1, n1, x2, n2p, P0, plane, ncb, F, F2, out_f32, npair, scale, kern, gx);
    TC_HIP(hipGetLastError());
  }
  if (split) {
    const int64_t W = r * nst;
    hipLaunchKernelGGL((fgemm_bf16_pipe_tail_kernel<MB, NPROD, Kern, NB, NW, FOLD, REV>), dim3((unsigned)S),
                       dim3(64 * NW), 0, st, p1, n1, x2, n2p, P0, plane, ncb, npair, kern, gx, full, W, part);
    TC_HIP(hipGetLastError());
    hipLaunchKernelGGL((bf16_pipe_fixup<Kern, TR, TC>), dim3((unsigned)((TR * TC + 255) / 256), (unsigned)r),
                       dim3(256), 0, st, part, n1, npair, gx, full, nst, W, S, F, F2, out_f32, scale, kern);
    TC_HIP(hipGetLastError());
  }
  return 0;
}
// Accumulation scheme of the Coulomb kernel (A/B): 0 = FP64 fold every step,
// products hh first; 1 = every step, corrections first; 2 = every second step,
// corrections first; 3 = every second step, hh first.
int bf16_scheme() {
  static const int v = [] {
    const char* e = std::getenv("TCSCF_BF16_SCHEME");
    return e ? std::atoi(e) : 0;
  }();
  return v;
}

template <class Kern>
struct PipeServed : std::integral_constant<bool, std::is_same<Kern, KernCoulomb>::value ||
                                                     std::is_same<Kern, KernTCPair>::value ||
                                                     std::is_same<Kern, KernJastrowGrad>::value> {};

template <int MB, int NPROD, int FOLD, class Kern, int NB = 4>
int launch_bf16(const double* p1, int64_t n1, const float4* x2, int64_t n2p, const us8* P0, int64_t plane, int64_t ncb,
                void* F, int out_f32, int64_t npair, double scale, Kern kern, hipStream_t st, void* F2) {
  dim3 grid((unsigned)(ncb / (4 * NB)), (unsigned)((n1 + 16 * MB - 1) / (16 * MB)));
  constexpr int NSPLIT = NPROD == 3 ? 2 : 3;
  constexpr int NC = Kern::ncomp;
  if constexpr (PipeServed<Kern>::value && FOLD == (NPROD == 4 ? 2 : 1) && MB == (NC == 1 ? 2 : 1) && NB == 4) {
    // the pipelined kernel's buffer offsets are 32-bit byte offsets
    const int mode = bf16_pipe_mode(NC);
    if (mode && (int64_t)NSPLIT * plane * 16 < ((int64_t)1 << 31) && n2p * 16 < ((int64_t)1 << 31)) {
#define TC_PIPE(M, B, W) \
  return launch_bf16_pipe<M, B, W, NPROD, Kern, FOLD, NPROD == 4>(p1, n1, x2, n2p, P0, plane, ncb, F, out_f32, npair, \
                                                                  scale, kern, st, F2)
      if constexpr (NC == 1) {
        if (mode == 1) TC_PIPE(2, 4, 4);
        if (mode == 3) TC_PIPE(4, 2, 4);
        if constexpr (NPROD == 5) {
          const int sch = bf16_scheme();
          if (sch == 1)
            return launch_bf16_pipe<4, 2, 8, NPROD, Kern, 1, true>(p1, n1, x2, n2p, P0, plane, ncb, F, out_f32, npair,
                                                                   scale, kern, st, F2);
          if (sch == 2)
            return launch_bf16_pipe<4, 2, 8, NPROD, Kern, 2, true>(p1, n1, x2, n2p, P0, plane, ncb, F, out_f32, npair,
                                                                   scale, kern, st, F2);
          if (sch == 3)
            return launch_bf16_pipe<4, 2, 8, NPROD, Kern, 2, false>(p1, n1, x2, n2p, P0, plane, ncb, F, out_f32, npair,
                                                                    scale, kern, st, F2);
        }
        TC_PIPE(4, 2, 8);
      } else if constexpr (NC == 2) {
        if (mode == 2) TC_PIPE(4, 1, 8);
        TC_PIPE(2, 2, 8);
      } else {
        TC_PIPE(2, 1, 8);
      }
#undef TC_PIPE
    }
  }
  // the plain tile, with the last round split as for the pipelined kernel
  constexpr int TR = 16 * MB, TC = 64 * NB;
  const int64_t gx = grid.x, tiles = gx * grid.y, nst = n2p / 32;
  const int64_t S = split_slots((const void*)fgemm_bf16_kernel<MB, NPROD, FOLD, Kern, NB>, 256, tiles, nst);
  const int64_t r = S ? tiles % S : 0, full = tiles - r;
  double* part = nullptr;
  if (S) {
    int rc = 0;
    part = (double*)scratch(kSlotSplit, (size_t)(2 * S) * NC * TR * TC * sizeof(double), &rc);
    if (rc) return rc;
  }
  if (full > 0) {
    hipLaunchKernelGGL((fgemm_bf16_kernel<MB, NPROD, FOLD, Kern, NB>), dim3((unsigned)full), dim3(256), 0, st, p1, n1,
                       x2, n2p, P0, plane, ncb, F, F2, out_f32, npair, scale, kern, gx);
    TC_HIP(hipGetLastError());
  }
  if (S) {
    const int64_t W = r * nst;
    hipLaunchKernelGGL((fgemm_bf16_tail_kernel<MB, NPROD, FOLD, Kern, NB>), dim3((unsigned)S), dim3(256), 0, st, p1,
                       n1, x2, n2p, P0, plane, ncb, npair, kern, gx, full, W, part);
    TC_HIP(hipGetLastError());
    hipLaunchKernelGGL((bf16_pipe_fixup<Kern, TR, TC>), dim3((unsigned)((TR * TC + 255) / 256), (unsigned)r),
                       dim3(256), 0, st, part, n1, npair, gx, full, nst, W, S, F, F2, out_f32, scale, kern);
    TC_HIP(hipGetLastError());
  }
  return 0;
}

// BF16 kernel variant: rows per workgroup 16*MB, FP64 fold every FOLD steps,
// NB 16-column blocks per wave.  TCSCF_BF16_CFG="MB,FOLD[,NB]" selects one for
// tuning (its FOLD applies to the one- and two-component kernels at three or
// five products; V always folds every kVFold steps).
struct Bf16Cfg {
  int mb, fold, nb;
};
Bf16Cfg bf16_cfg(int ncomp) {
  static Bf16Cfg env = [] {
    Bf16Cfg c{0, 0, 4};
    if (const char* e = std::getenv("TCSCF_BF16_CFG")) std::sscanf(e, "%d,%d,%d", &c.mb, &c.fold, &c.nb);
    return c;
  }();
  Bf16Cfg c = env.mb ? env : Bf16Cfg{2, 1, 4};
  if (ncomp != 1) c = Bf16Cfg{1, c.fold, 4};
  return c;
}

template <int NPROD, int MB, class Kern>
int launch_bf16_fold(Bf16Cfg c, const double* p1, int64_t n1, const float4* x2, int64_t n2p, const us8* P0,
                     int64_t plane, int64_t ncb, void* F, int out_f32, int64_t npair, double scale, Kern kern,
                     hipStream_t st, void* F2) {
  if constexpr (Kern::ncomp == 3) {  // V (three components): FP64 fold every kVFold steps
    if ((double)n1 * (double)n2p * (double)npair > kVLongWork)
      return launch_bf16<MB, NPROD, 1>(p1, n1, x2, n2p, P0, plane, ncb, F, out_f32, npair, scale, kern, st, F2);
    return launch_bf16<MB, NPROD, kVFold>(p1, n1, x2, n2p, P0, plane, ncb, F, out_f32, npair, scale, kern, st, F2);
  } else if constexpr (NPROD == 4) {  // four products: FP64 fold every second step
    if (MB == 4 && c.nb == 2)
      return launch_bf16<MB, NPROD, 2, Kern, 2>(p1, n1, x2, n2p, P0, plane, ncb, F, out_f32, npair, scale, kern, st, F2);
    return launch_bf16<MB, NPROD, 2>(p1, n1, x2, n2p, P0, plane, ncb, F, out_f32, npair, scale, kern, st, F2);
  } else {
    if (MB == 4 && c.nb == 2) {
      if (c.fold == 2)
        return launch_bf16<MB, NPROD, 2, Kern, 2>(p1, n1, x2, n2p, P0, plane, ncb, F, out_f32, npair, scale, kern, st,
                                                  F2);
      return launch_bf16<MB, NPROD, 1, Kern, 2>(p1, n1, x2, n2p, P0, plane, ncb, F, out_f32, npair, scale, kern, st, F2);
    }
    if (c.fold == 2)
      return launch_bf16<MB, NPROD, 2>(p1, n1, x2, n2p, P0, plane, ncb, F, out_f32, npair, scale, kern, st, F2);
    return launch_bf16<MB, NPROD, 1>(p1, n1, x2, n2p, P0, plane, ncb, F, out_f32, npair, scale, kern, st, F2);
  }
}

template <int NPROD, class Kern>
int launch_bf16_mb(const double* p1, int64_t n1, const float4* x2, int64_t n2p, const us8* P0, int64_t plane,
                   int64_t ncb, void* F, int out_f32, int64_t npair, double scale, Kern kern, hipStream_t st,
                   void* F2 = nullptr) {
  const Bf16Cfg c = bf16_cfg(Kern::ncomp);
  if constexpr (Kern::ncomp == 1) {
    switch (c.mb) {
      case 4:
        return launch_bf16_fold<NPROD, 4>(c, p1, n1, x2, n2p, P0, plane, ncb, F, out_f32, npair, scale, kern, st, F2);
      default:
        return launch_bf16_fold<NPROD, 2>(c, p1, n1, x2, n2p, P0, plane, ncb, F, out_f32, npair, scale, kern, st, F2);
    }
  } else {
    return launch_bf16_fold<NPROD, 1>(c, p1, n1, x2, n2p, P0, plane, ncb, F, out_f32, npair, scale, kern, st, F2);
  }
}

// (Fs, Fo) <- (2 Fs - cs Fo, 2 Fs - co Fo) elementwise.
template <class T>
__global__ __launch_bounds__(256) void tc_pair_combine(T* __restrict__ Fs, T* __restrict__ Fo, int64_t len, double cs,
                                                       double co) {
#pragma clang fp contract(off)
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < len; e += (int64_t)gridDim.x * 256) {
    const double fc = Fs[e], g = Fo[e];
    Fs[e] = (T)(2 * fc - cs * g);
    Fo[e] = (T)(2 * fc - co * g);
  }
}

// ws_fixed: the workspace (tcscf_F_workspace_bytes of it) given by the caller;
// null takes it from the library's scratch slot.
template <class Kern>
int dispatch(int precision, const double* pts1, int64_t n1, const double* pts2, const double* w2, const int* ord2,
             int64_t n2, const double* A2, int nb, double scale, void* F, int out_f32, Kern kern, hipStream_t st,
             void* F2 = nullptr, char* ws_fixed = nullptr) {
  const int64_t npair = npair_of(nb);
  int rc = 0;
  struct FixedWs {  // a caller-provided workspace: no other scratch either
    bool prev;
    explicit FixedWs(bool on) : prev(t_fixed_ws) { t_fixed_ws = on; }
    ~FixedWs() { t_fixed_ws = prev; }
  } fixed_guard(ws_fixed != nullptr);
  if (is_bf16(precision)) {
    const int nsplit = precision == TCSCF_BF16X3 ? 2 : 3;
    const int64_t n2p = round_up(n2, 32), ncb = round_up(npair, 256) / 16, nkb = n2p / 32;
    char* ws = ws_fixed ? ws_fixed : (char*)scratch(kSlotPairs, tcscf_F_workspace_bytes(n2, nb, precision), &rc);
    if (rc) return rc;
    unsigned short* planes = (unsigned short*)ws;
    const int64_t plane = nkb * ncb * 512;  // bf16 per plane
    double* xd = (double*)(planes + nsplit * plane);
    float4* x4 = (float4*)(xd + 3 * n2p);
    const unsigned gp = (unsigned)((plane / 8 + 255) / 256);  // one thread per fragment lane
    if (nsplit == 2)
      hipLaunchKernelGGL(pack_pairs_bf16<2>, dim3(gp), dim3(256), 0, st, A2, w2, ord2, n2, nb, npair, nkb, ncb, planes);
    else
      hipLaunchKernelGGL(pack_pairs_bf16<3>, dim3(gp), dim3(256), 0, st, A2, w2, ord2, n2, nb, npair, nkb, ncb, planes);
    TC_HIP(hipGetLastError());
    hipLaunchKernelGGL(pad_coords, dim3((unsigned)((n2p + 255) / 256)), dim3(256), 0, st, pts2, ord2, n2, n2p, xd, x4);
    TC_HIP(hipGetLastError());
    // planes are addressed in 16-byte fragment-lane units
    const us8* P0 = (const us8*)planes;
    const int64_t plane8 = plane / 8;
    if (precision == TCSCF_BF16X3)
      return launch_bf16_mb<3>(pts1, n1, x4, n2p, P0, plane8, ncb, F, out_f32, npair, scale, kern, st, F2);
    if (precision == TCSCF_BF16X4)
      return launch_bf16_mb<4>(pts1, n1, x4, n2p, P0, plane8, ncb, F, out_f32, npair, scale, kern, st, F2);
    return launch_bf16_mb<5>(pts1, n1, x4, n2p, P0, plane8, ncb, F, out_f32, npair, scale, kern, st, F2);
  }
  if constexpr (Combines<Kern>::value) {
    TC_CHECK(false, "combined-output kernels run on the BF16 path only");
  } else {
  const int64_t n2p = round_up(n2, 16), ldp = round_up(npair, 128);
  char* ws = ws_fixed ? ws_fixed : (char*)scratch(kSlotPairs, tcscf_F_workspace_bytes(n2, nb, precision), &rc);
  if (rc) return rc;
  const size_t el = precision == TCSCF_FP64 ? 8 : 4;
  double* xd = (double*)(ws + (size_t)n2p * ldp * el);
  float4* x4 = (float4*)(xd + 3 * n2p);
  const int64_t tot = n2p * ldp;
  hipLaunchKernelGGL(pad_coords, dim3((unsigned)((n2p + 255) / 256)), dim3(256), 0, st, pts2, ord2, n2, n2p, xd, x4);
  TC_HIP(hipGetLastError());
  if (precision == TCSCF_FP64) {
    hipLaunchKernelGGL(pack_pairs_rowmajor<double>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, A2, w2, ord2,
                       n2, nb, npair, n2p, ldp, (double*)ws);
    TC_HIP(hipGetLastError());
    if constexpr (Kern::ncomp == 1) return launch_f64<4>(pts1, n1, xd, n2p, (const double*)ws, ldp, F, out_f32, npair, scale, kern, st);
    else return launch_f64<2>(pts1, n1, xd, n2p, (const double*)ws, ldp, F, out_f32, npair, scale, kern, st);
  }
  hipLaunchKernelGGL(pack_pairs_rowmajor<float>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, A2, w2, ord2, n2,
                     nb, npair, n2p, ldp, (float*)ws);
  TC_HIP(hipGetLastError());
  if constexpr (Kern::ncomp == 1) return launch_f32<4>(pts1, n1, x4, n2p, (const float*)ws, ldp, F, out_f32, npair, scale, kern, st);
  else return launch_f32<1>(pts1, n1, x4, n2p, (const float*)ws, ldp, F, out_f32, npair, scale, kern, st);
  }
  return TCSCF_EINVAL;
}

}  // namespace

namespace tc {

int compute_F_into(const double* pts1, int64_t n1, const double* pts2, const double* w2, const int32_t* ord2,
                   int64_t n2, const double* A2, int nb, int kernel, const tcscf_jastrow* jp, int same_spin,
                   int precision, double scale, void* F, int out_dtype, void* stream, char* ws_fixed) {
  TC_CHECK(pts1 && pts2 && w2 && A2 && F && n1 > 0 && n2 > 0 && nb > 0, "tcscf_compute_F: bad arguments");
  TC_CHECK(precision >= 0 && precision <= 4, "tcscf_compute_F: unknown precision");
  TC_CHECK(out_dtype == TCSCF_F64 || out_dtype == TCSCF_F32, "tcscf_compute_F: unknown out_dtype");
  TC_CHECK(n1 < ((int64_t)1 << 31) && n2 < ((int64_t)1 << 31), "tcscf_compute_F: grid too large");
  hipStream_t st = (hipStream_t)stream;
  const int of = out_dtype == TCSCF_F32;
  switch (kernel) {
    case TCSCF_K_COULOMB:
      return dispatch(precision, pts1, n1, pts2, w2, ord2, n2, A2, nb, scale, F, of, KernCoulomb{}, st, nullptr,
                      ws_fixed);
    case TCSCF_K_TC_SCALAR: {
      TC_CHECK(jp && jp->a12 > 0, "tcscf_compute_F: TC kernel needs jastrow parameters with a12 > 0");
      const double c = same_spin ? jp->c_same : jp->c_oppo;
      return dispatch(precision, pts1, n1, pts2, w2, ord2, n2, A2, nb, scale, F, of,
                      KernTCScalar{jp->a12, 2 * c * c * jp->a12 * jp->a12}, st, nullptr, ws_fixed);
    }
    case TCSCF_K_JASTROW_GRAD:
      TC_CHECK(jp && jp->a12 > 0, "tcscf_compute_F: gradient kernel needs jastrow parameters with a12 > 0");
      return dispatch(precision, pts1, n1, pts2, w2, ord2, n2, A2, nb, scale, F, of, KernJastrowGrad{jp->a12}, st,
                      nullptr, ws_fixed);
    case TCSCF_K_TC_SMOOTH:
      TC_CHECK(jp && jp->a12 > 0, "tcscf_compute_F: TC kernel needs jastrow parameters with a12 > 0");
      return dispatch(precision, pts1, n1, pts2, w2, ord2, n2, A2, nb, scale, F, of, KernTCSmooth{jp->a12}, st,
                      nullptr, ws_fixed);
    default:
      TC_CHECK(false, "tcscf_compute_F: unknown kernel");
  }
  return TCSCF_EINVAL;
}

}  // namespace tc

extern "C" {

int tcscf_basis_eval_fns(const double* pts, int64_t n, const tcscf_basis_fn* fns, int nb, double* values,
                         double* grads, void* stream) {
  TC_CHECK(pts && fns && values && n > 0 && nb > 0, "tcscf_basis_eval: bad arguments");
  for (int b0 = 0; b0 < nb; b0 += kMaxFnsPerLaunch) {
    StoTable t;
    t.nb = std::min(kMaxFnsPerLaunch, nb - b0);
    for (int b = 0; b < t.nb; ++b) {
      const tcscf_basis_fn& f = fns[b0 + b];
      TC_CHECK(f.l >= std::abs(f.m) && f.l <= 3 && f.alpha > 0, "tcscf_basis_eval: need |m| <= l <= 3, alpha > 0");
      double zeta = f.alpha, norm = 0;
      if (f.kind == TCSCF_BASIS_STO) {
        TC_CHECK(f.n >= 1 && f.n >= f.l, "tcscf_basis_eval: STO needs n >= 1, n >= l");
        norm = std::pow(2 * f.alpha, f.n) * std::sqrt(2 * f.alpha / std::tgamma(2.0 * f.n + 1));
      } else if (f.kind == TCSCF_BASIS_OCHI) {
        TC_CHECK(f.n >= 0, "tcscf_basis_eval: Ochi function needs n >= 0");
        // (2 alpha)^(l + 3/2) / sqrt((n + 2l + 2)! / n!)
        norm = std::pow(2 * f.alpha, f.l + 1.5) / std::sqrt(std::exp(std::lgamma(f.n + 2.0 * f.l + 3) - std::lgamma(f.n + 1.0)));
      } else if (f.kind == TCSCF_BASIS_HYDROGEN) {
        TC_CHECK(f.n > f.l, "tcscf_basis_eval: hydrogen-like function needs n > l");
        zeta = f.alpha / f.n;  // b = Z / n
        // -(2b)^(l + 3/2) / sqrt(2 n (n + l)! / (n - l - 1)!)
        norm = -std::pow(2 * zeta, f.l + 1.5) /
               std::sqrt(2.0 * f.n * std::exp(std::lgamma(f.n + f.l + 1.0) - std::lgamma(f.n - f.l + 0.0)));
      } else {
        TC_CHECK(false, "tcscf_basis_eval: unknown basis kind");
      }
      t.code[b] = f.n | (f.l << 8) | ((f.m + 8) << 16) | (f.kind << 24);
      t.zeta[b] = zeta;
      t.norm[b] = norm;
    }
    const int64_t total = n * t.nb;
    hipLaunchKernelGGL(basis_eval_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       pts, n, t, nb, b0, values, grads);
    TC_HIP(hipGetLastError());
  }
  return 0;
}

int tcscf_basis_eval(const double* pts, int64_t n, const tcscf_sto* fns, int nb, double* values, double* grads,
                     void* stream) {
  TC_CHECK(fns && nb > 0, "tcscf_basis_eval: bad arguments");
  std::vector<tcscf_basis_fn> f(nb);
  for (int b = 0; b < nb; ++b) f[b] = tcscf_basis_fn{TCSCF_BASIS_STO, fns[b].n, fns[b].l, fns[b].m, fns[b].zeta};
  return tcscf_basis_eval_fns(pts, n, f.data(), nb, values, grads, stream);
}

size_t tcscf_F_workspace_bytes(int64_t n2, int nb, int precision) {
  const int64_t npair = npair_of(nb);
  if (tc::is_bf16(precision)) {
    const int64_t n2p = round_up(n2, 32), ncb = round_up(npair, 256) / 16;
    const size_t nsplit = precision == TCSCF_BF16X3 ? 2 : 3;
    return (size_t)(n2p / 32) * ncb * 512 * 2 * nsplit + (size_t)n2p * (3 * 8 + 16) + 4096;
  }
  const int64_t n2p = round_up(n2, 16), ldp = round_up(npair, 128);
  const size_t el = precision == TCSCF_FP64 ? 8 : 4;
  return (size_t)n2p * ldp * el + (size_t)n2p * (3 * 8 + 16) + 4096;
}


int tcscf_compute_F(const double* pts1, int64_t n1, const double* pts2, const double* w2, const int32_t* ord2,
                    int64_t n2, const double* A2, int nb, int kernel, const tcscf_jastrow* jp, int same_spin,
                    int precision, double scale, void* F, int out_dtype, void* stream) {
  return tc::compute_F_into(pts1, n1, pts2, w2, ord2, n2, A2, nb, kernel, jp, same_spin, precision, scale, F,
                            out_dtype, stream, nullptr);
}

int tcscf_compute_F_ochi(const double* pts1, int64_t n1, const double* pts2, const double* w2, const int32_t* ord2,
                         int64_t n2, const double* A2, int nb, int kernel, const tcscf_ochi_jastrow* u, int same_spin,
                         int precision, double scale, void* F, int out_dtype, void* stream) {
  TC_CHECK(pts1 && pts2 && w2 && A2 && F && n1 > 0 && n2 > 0 && nb > 0 && u, "tcscf_compute_F_ochi: bad arguments");
  TC_CHECK(precision >= 0 && precision <= 4, "tcscf_compute_F_ochi: unknown precision");
  TC_CHECK(out_dtype == TCSCF_F64 || out_dtype == TCSCF_F32, "tcscf_compute_F_ochi: unknown out_dtype");
  TC_CHECK(n1 < ((int64_t)1 << 31) && n2 < ((int64_t)1 << 31), "tcscf_compute_F_ochi: grid too large");
  TC_CHECK(u->a > 0 && u->a12 > 0, "tcscf_compute_F_ochi: need a > 0 and a12 > 0");
  TC_CHECK(u->nterms >= 0 && u->nterms <= TCSCF_OCHI_MAX_TERMS, "tcscf_compute_F_ochi: 0 .. 8 terms");
  OchiTerms t{};
  t.a = u->a;
  t.a12 = u->a12;
  t.nt = u->nterms;
  for (int q = 0; q < u->nterms; ++q) {
    TC_CHECK(u->i[q] >= 0 && u->i[q] <= 16 && u->j[q] >= 0 && u->j[q] <= 16 && u->k[q] >= 0 && u->k[q] <= 16,
             "tcscf_compute_F_ochi: exponents must lie in 0 .. 16");
    t.i[q] = u->i[q];
    t.j[q] = u->j[q];
    t.k[q] = u->k[q];
    t.c[q] = same_spin ? u->c_same[q] : u->c_oppo[q];
  }
  hipStream_t st = (hipStream_t)stream;
  const int of = out_dtype == TCSCF_F32;
  if (kernel == TCSCF_K_TC_SCALAR)
    return dispatch(precision, pts1, n1, pts2, w2, ord2, n2, A2, nb, scale, F, of, KernOchi<false>{t}, st);
  TC_CHECK(kernel == TCSCF_K_JASTROW_GRAD, "tcscf_compute_F_ochi: kernel must be TCSCF_K_TC_SCALAR or TCSCF_K_JASTROW_GRAD");
  return dispatch(precision, pts1, n1, pts2, w2, ord2, n2, A2, nb, scale, F, of, KernOchi<true>{t}, st);
}

// F_same and F_oppo of the TC scalar kernel from the Coulomb contraction F_C
// (in Fs, at `precision`) and the contraction G of its smooth part (in Fo, at
// `smooth_precision`): F_spin = 2 F_C - 2 c_spin^2 a12^2 G, in place.
int tcscf_compute_F_tc_pair(const double* pts1, int64_t n1, const double* pts2, const double* w2,
                            const int32_t* ord2, int64_t n2, const double* A2, int nb, const tcscf_jastrow* jp,
                            int precision, int smooth_precision, double scale, void* Fs, void* Fo, int out_dtype,
                            void* stream) {
  TC_CHECK(jp && jp->a12 > 0, "tcscf_compute_F_tc_pair: jastrow parameters with a12 > 0 needed");
  TC_CHECK(Fs && Fo && Fs != Fo, "tcscf_compute_F_tc_pair: two distinct outputs needed");
  const double a2 = jp->a12 * jp->a12;
  const double cs = 2 * jp->c_same * jp->c_same * a2, co = 2 * jp->c_oppo * jp->c_oppo * a2;
  if ((precision == TCSCF_BF16X5 || precision == TCSCF_BF16X4) && smooth_precision == TCSCF_BF16X3) {
    // one launch: 1/r12 at five (four) products and the smooth part at three
    // share the three-way split pair densities and r12; the epilogue combines
    TC_CHECK(pts1 && pts2 && w2 && A2 && n1 > 0 && n2 > 0 && nb > 0, "tcscf_compute_F_tc_pair: bad arguments");
    TC_CHECK(out_dtype == TCSCF_F64 || out_dtype == TCSCF_F32, "tcscf_compute_F_tc_pair: unknown out_dtype");
    TC_CHECK(n1 < ((int64_t)1 << 31) && n2 < ((int64_t)1 << 31), "tcscf_compute_F_tc_pair: grid too large");
    return dispatch(precision, pts1, n1, pts2, w2, ord2, n2, A2, nb, scale, Fs, out_dtype == TCSCF_F32,
                    KernTCPair{jp->a12, cs, co}, (hipStream_t)stream, Fo);
  }
  int rc = tcscf_compute_F(pts1, n1, pts2, w2, ord2, n2, A2, nb, TCSCF_K_COULOMB, nullptr, 1, precision, scale, Fs,
                           out_dtype, stream);
  if (rc) return rc;
  if ((rc = tcscf_compute_F(pts1, n1, pts2, w2, ord2, n2, A2, nb, TCSCF_K_TC_SMOOTH, jp, 1, smooth_precision, scale,
                            Fo, out_dtype, stream)))
    return rc;
  const int64_t len = n1 * npair_of(nb);
  const unsigned grid = (unsigned)std::min<int64_t>((len + 255) / 256, 65536);
  hipStream_t st = (hipStream_t)stream;
  if (out_dtype == TCSCF_F32)
    hipLaunchKernelGGL(tc_pair_combine<float>, dim3(grid), dim3(256), 0, st, (float*)Fs, (float*)Fo, len, cs, co);
  else
    hipLaunchKernelGGL(tc_pair_combine<double>, dim3(grid), dim3(256), 0, st, (double*)Fs, (double*)Fo, len, cs, co);
  TC_HIP(hipGetLastError());
  return 0;
}

// Runs one fragment of MFMA `which` (0 f64 16x16x4, 1 f32 16x16x4, 2 bf16
// 16x16x32) with lane 16 g + r supplying A[r][k] and B[k][r] for its k slice;
// D receives the raw accumulator registers, D[4 l + i] (64 x 4, FP64).
int tcscf_selftest_mfma(int which, const double* A_dev, const double* B_dev, double* D_dev, void* stream) {
  TC_CHECK(which >= 0 && which <= 2, "tcscf_selftest_mfma: which");
  hipLaunchKernelGGL(mfma_selftest_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream, which, A_dev, B_dev, D_dev);
  TC_HIP(hipGetLastError());
  return 0;
}

}  // extern "C"

/* ======================================================================== */
/* Grid ERI: integrateAllR1R2 (src/integration/integrateAll.hpp:221-317)     */
/*   E[p(b1,b3)][col] = sum_n1 w1 A1[n1,b1] A1[n1,b3] T[n1, col]              */
/* as an FP64 MFMA GEMM E = P1w^T T (M = npair, N = ncol, K = n1), split over */
/* K.  T = F (packed pairs, USE_GRADIENT = false) or the gradient inner       */
/* integral (nb^2 ordered pairs (b2, b4), USE_GRADIENT = true).              */
/* ======================================================================== */
namespace tc {

// Workgroup = 4 waves (2 x 2) computing a 128 x 128 tile of E over the K range
// [kb, kb + kchunk) of blockIdx.z; partial tiles go to part[z][M][ncol].
__global__ __launch_bounds__(256) void eri_gemm_kernel(const double* __restrict__ P, int64_t ldp, int64_t M,
                                                       const double* __restrict__ T, int64_t ncol, int64_t n1,
                                                       int64_t kchunk, double* __restrict__ part) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = lane >> 4, r = lane & 15;
  const int64_t m0 = (int64_t)blockIdx.y * 128 + (wave >> 1) * 64;
  const int64_t c0 = (int64_t)blockIdx.x * 128 + (wave & 1) * 64;
  const int64_t kb = (int64_t)blockIdx.z * kchunk;
  const int64_t ke = kb + kchunk < n1 ? kb + kchunk : n1;
  d4 acc[4][4];
#pragma unroll
  for (int m = 0; m < 4; ++m)
#pragma unroll
    for (int n = 0; n < 4; ++n) acc[m][n] = d4{0, 0, 0, 0};
  for (int64_t k0 = kb; k0 < ke; k0 += 4) {
    const int64_t k = k0 + g;
    const bool kin = k < ke;
    double a[4], b[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) a[m] = kin ? P[k * ldp + m0 + 16 * m + r] : 0.0;  // zero beyond M
#pragma unroll
    for (int n = 0; n < 4; ++n) {
      const int64_t col = c0 + 16 * n + r;
      b[n] = (kin && col < ncol) ? T[k * ncol + col] : 0.0;
    }
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
      for (int n = 0; n < 4; ++n) acc[m][n] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[m], b[n], acc[m][n], 0, 0, 0);
  }
  double* out = part + (int64_t)blockIdx.z * M * ncol;
#pragma unroll
  for (int m = 0; m < 4; ++m)
#pragma unroll
    for (int n = 0; n < 4; ++n)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int64_t row = m0 + 16 * m + d_row_f64(g, i), col = c0 + 16 * n + r;
        if (row < M && col < ncol) out[row * ncol + col] = acc[m][n][i];
      }
}

__global__ void eri_reduce_kernel(const double* __restrict__ part, int nz, int64_t len, double* __restrict__ out) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= len) return;
  double s = 0;
  for (int z = 0; z < nz; ++z) s += part[(int64_t)z * len + e];
  out[e] = s;
}

// Ordered gradient pairs, 3 rows per r2 point:
// Q[3 k + c][b2 nb + b4] = w2[k] A2[k, b2] dA2[k, b4, c]; zero for k >= n2 or col >= nb^2.
__global__ __launch_bounds__(256) void pack_grad_pairs(const double* __restrict__ A2, const double* __restrict__ dA2,
                                                       const double* __restrict__ w2, int64_t n2, int nb, int64_t n2p,
                                                       int64_t ldq, double* __restrict__ Q) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= 3 * n2p * ldq) return;
  const int64_t row = idx / ldq, col = idx - row * ldq;
  const int64_t k = row / 3;
  const int c = (int)(row - 3 * k);
  double v = 0;
  if (k < n2 && col < (int64_t)nb * nb) {
    const int b2 = (int)(col / nb), b4 = (int)(col - (int64_t)b2 * nb);
    v = w2[k] * A2[k * nb + b2] * dA2[(k * nb + b4) * 3 + c];
  }
  Q[idx] = v;
}

// Gradient inner integral T[n1][col] = scale sum_{k, c} f_c(r1 - r2_k) Q[3 k + c][col]
// (evaluateR2GradientIntegral, src/integration/integrateAll.hpp:184-215) as one
// FP64 MFMA GEMM with K = 3 n2p (point-major, component-minor).  Same tiling as
// fgemm_f64_kernel: 4 waves as 2 x 2, wave tile 16 MB x 64.
template <int MB, class Kern>
__global__ __launch_bounds__(256) void fgemm_f64_vdot_kernel(const double* __restrict__ p1, int64_t n1,
                                                             const double* __restrict__ x2, int64_t n2p,
                                                             const double* __restrict__ Q, int64_t ldq,
                                                             double* __restrict__ T, int64_t ncol, double scale,
                                                             Kern kern) {
  static_assert(Kern::ncomp == 3, "vector kernel expected");
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = lane >> 4, r = lane & 15;
  const int64_t rowW = (int64_t)blockIdx.y * (32 * MB) + (wave >> 1) * (16 * MB);
  const int64_t colW = (int64_t)blockIdx.x * 128 + (wave & 1) * 64;
  double ax[MB], ay[MB], az[MB];
#pragma unroll
  for (int m = 0; m < MB; ++m) {
    const int64_t row = rowW + 16 * m + r;
    const bool ok = row < n1;
    ax[m] = ok ? p1[row] : -kFar;
    ay[m] = ok ? p1[n1 + row] : -kFar;
    az[m] = ok ? p1[2 * n1 + row] : -kFar;
  }
  d4 acc[MB][4];
#pragma unroll
  for (int m = 0; m < MB; ++m)
#pragma unroll
    for (int n = 0; n < 4; ++n) acc[m][n] = d4{0, 0, 0, 0};
  const double* __restrict__ Qb = Q + colW + r;
  for (int64_t k0 = 0; k0 < 3 * n2p; k0 += 4) {
    const int64_t k = k0 + g, pt = k / 3;
    const int c = (int)(k - 3 * pt);
    const double bx = x2[pt], by = x2[n2p + pt], bz = x2[2 * n2p + pt];
    double b[4];
#pragma unroll
    for (int n = 0; n < 4; ++n) b[n] = Qb[k * ldq + 16 * n];
    double a[MB];
#pragma unroll
    for (int m = 0; m < MB; ++m) {
      double v[3];
      kernel_eval(kern, ax[m], ay[m], az[m], bx, by, bz, v);
      a[m] = c == 0 ? v[0] : (c == 1 ? v[1] : v[2]);
    }
#pragma unroll
    for (int m = 0; m < MB; ++m)
#pragma unroll
      for (int n = 0; n < 4; ++n) acc[m][n] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[m], b[n], acc[m][n], 0, 0, 0);
  }
#pragma unroll
  for (int m = 0; m < MB; ++m)
#pragma unroll
    for (int n = 0; n < 4; ++n)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int64_t row = rowW + 16 * m + d_row_f64(g, i), col = colW + 16 * n + r;
        if (row < n1 && col < ncol) T[row * ncol + col] = scale * acc[m][n][i];
      }
}

// E (npair x ncol) = P1w^T T for T n1 x ncol (device); ws holds P1 and the partials.
int eri_contract(const double* A1, const double* w1, int64_t n1, int nb, const double* T, int64_t ncol, double* E,
                 double* ws, hipStream_t st) {
  const int64_t npair = npair_of(nb), ldp = round_up(npair, 128), n1p = round_up(n1, 16);
  const int nz = (int)std::min<int64_t>(64, std::max<int64_t>(1, n1 / 512));
  const int64_t kchunk = round_up((n1 + nz - 1) / nz, 4);
  double* P = ws;
  double* part = ws + n1p * ldp;
  const int64_t tot = n1p * ldp;
  hipLaunchKernelGGL(pack_pairs_rowmajor<double>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, A1, w1,
                     (const int*)nullptr, n1, nb, npair, n1p, ldp, P);
  TC_HIP(hipGetLastError());
  dim3 grid((unsigned)(round_up(ncol, 128) / 128), (unsigned)(ldp / 128), (unsigned)nz);
  hipLaunchKernelGGL(eri_gemm_kernel, grid, dim3(256), 0, st, P, ldp, npair, T, ncol, n1, kchunk, part);
  TC_HIP(hipGetLastError());
  const int64_t len = npair * ncol;
  hipLaunchKernelGGL(eri_reduce_kernel, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, st, part, nz, len, E);
  TC_HIP(hipGetLastError());
  return 0;
}

size_t eri_contract_doubles(int64_t n1, int nb, int64_t ncol) {
  const int64_t npair = npair_of(nb), ldp = round_up(npair, 128), n1p = round_up(n1, 16);
  const int64_t nz = std::min<int64_t>(64, std::max<int64_t>(1, n1 / 512));
  return (size_t)(n1p * ldp + nz * npair * ncol);
}

}  // namespace tc

extern "C" int tcscf_eri(const double* A1, const double* w1, int64_t n1, int nb, const void* F, double* eri,
                         void* stream) {
  using namespace tc;
  TC_CHECK(A1 && w1 && F && eri && n1 > 0 && nb > 0, "tcscf_eri: bad arguments");
  int rc = 0;
  double* ws = (double*)scratch(kSlotSmall, sizeof(double) * eri_contract_doubles(n1, nb, npair_of(nb)), &rc);
  if (rc) return rc;
  return eri_contract(A1, w1, n1, nb, (const double*)F, npair_of(nb), eri, ws, (hipStream_t)stream);
}

// F[k][p(j, i)] = A2[k][j] A2[k][i] / |r1_k - r2_k| (0 for coincident points).
__global__ __launch_bounds__(256) void qmc_pair_F(const double* __restrict__ p1, const double* __restrict__ p2,
                                                  const double* __restrict__ A2, int64_t n, int nb, int64_t npair,
                                                  double* __restrict__ F) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= n * npair) return;
  const int64_t k = e / npair, p = e - k * npair;
  int j, i;
  pair_of(p, &j, &i);
  const double dx = p1[k] - p2[k], dy = p1[n + k] - p2[n + k], dz = p1[2 * n + k] - p2[2 * n + k];
  const double s = dx * dx + dy * dy + dz * dz;
  F[e] = s > 0 ? A2[k * nb + j] * A2[k * nb + i] / __builtin_sqrt(s) : 0.0;
}

extern "C" int tcscf_eri_qmc6d(const double* pts1, const double* pts2, const double* w, int64_t n, const double* A1,
                               const double* A2, int nb, double* eri, void* stream) {
  using namespace tc;
  TC_CHECK(pts1 && pts2 && w && A1 && A2 && eri && n > 0 && nb > 0, "tcscf_eri_qmc6d: bad arguments");
  const int64_t npair = npair_of(nb);
  int rc = 0;
  double* ws = (double*)scratch(kSlotQmc, sizeof(double) * ((size_t)n * npair + eri_contract_doubles(n, nb, npair)), &rc);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(qmc_pair_F, dim3((unsigned)((n * npair + 255) / 256)), dim3(256), 0, st, pts1, pts2, A2, n, nb, npair,
                     ws);
  TC_HIP(hipGetLastError());
  return eri_contract(A1, w, n, nb, ws, npair, eri, ws + (size_t)n * npair, st);
}

namespace tc {
struct ComplexCols {
  int m[64];        // m of each function
  int partner[64];  // column of the real function with -m (itself for m = 0)
};

// The reference's complex functions g = R sph_legendre(l, |m|, theta) e^{i m phi}
// from the real table S (no Condon-Shortley phase in S):
//   g_{+u} = (-1)^u (S_{+u} + i S_{-u}) / sqrt2,  g_{-u} = (-1)^u (S_{+u} - i S_{-u}) / sqrt2.
// G1 (n x 2nb): [Re g | Im g] at the r1 points; G2 (n x nb): Re g at the r2
// points, which lie on the phi = 0 half-plane where g is real.
__global__ __launch_bounds__(256) void complex_columns(const double* __restrict__ A1, const double* __restrict__ A2,
                                                       int64_t n, int nb, ComplexCols cc, double* __restrict__ G1,
                                                       double* __restrict__ G2) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= n * nb) return;
  const int64_t k = e / nb;
  const int a = (int)(e - k * nb), m = cc.m[a], q = cc.partner[a];
  double re1 = A1[e], im1 = 0.0, re2 = A2[e];
  if (m != 0) {
    const double s = ((m < 0 ? -m : m) & 1) ? -M_SQRT1_2 : M_SQRT1_2;
    const double plus1 = m > 0 ? A1[e] : A1[k * nb + q], minus1 = m > 0 ? A1[k * nb + q] : A1[e];
    re1 = s * plus1;
    im1 = m > 0 ? s * minus1 : -s * minus1;
    re2 = s * (m > 0 ? A2[e] : A2[k * nb + q]);
  }
  G1[k * 2 * nb + a] = re1;
  G1[k * 2 * nb + nb + a] = im1;
  G2[e] = re2;
}

// eri[2 e], eri[2 e + 1] (e = ((a nb + b) nb + c) nb + d): the complex integral
// from Eh = [Re g | Im g]-pair rows x (b, d) columns.
__global__ __launch_bounds__(256) void complex_eri_assemble(const double* __restrict__ Eh, int nb, int64_t ncol,
                                                            double* __restrict__ eri) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t nb4 = (int64_t)nb * nb * nb * nb;
  if (e >= nb4) return;
  const int d = (int)(e % nb), c = (int)((e / nb) % nb), b = (int)((e / nb / nb) % nb), a = (int)(e / nb / nb / nb);
  auto p = [](int j, int i) { return j <= i ? (int64_t)i * (i + 1) / 2 + j : (int64_t)j * (j + 1) / 2 + i; };
  const int64_t col = p(b, d);
  const double re = Eh[p(a, c) * ncol + col] + Eh[p(nb + a, nb + c) * ncol + col];
  const double im = Eh[p(a, nb + c) * ncol + col] - Eh[p(c, nb + a) * ncol + col];
  eri[2 * e] = re;
  eri[2 * e + 1] = im;
}
}  // namespace tc

extern "C" int tcscf_eri_qmc5d_atomic(const double* pts1, const double* pts2, const double* w, int64_t n,
                                      const double* A1, const double* A2, int nb, const int* m_host,
                                      const int* partner_host, double* eri, void* stream) {
  using namespace tc;
  TC_CHECK(pts1 && pts2 && w && A1 && A2 && eri && m_host && partner_host && n > 0 && nb > 0,
           "tcscf_eri_qmc5d_atomic: bad arguments");
  TC_CHECK(nb <= 32, "tcscf_eri_qmc5d_atomic: at most 32 functions");
  ComplexCols cc{};
  for (int a = 0; a < nb; ++a) {
    const int q = partner_host[a];
    TC_CHECK(q >= 0 && q < nb && m_host[q] == -m_host[a] && (m_host[a] != 0 || q == a),
             "tcscf_eri_qmc5d_atomic: function " + std::to_string(a) + " needs its -m partner in the basis");
    cc.m[a] = m_host[a];
    cc.partner[a] = q;
  }
  const int64_t npair = npair_of(nb), nrow = npair_of(2 * nb);
  const size_t nG = (size_t)n * 3 * nb, nF = (size_t)n * npair, nE = (size_t)nrow * npair;
  int rc = 0;
  double* ws = (double*)scratch(kSlotQmc, sizeof(double) * (nG + nF + nE + eri_contract_doubles(n, 2 * nb, npair)), &rc);
  if (rc) return rc;
  double *G1 = ws, *G2 = ws + (size_t)n * 2 * nb, *F = ws + nG, *Eh = F + nF, *cws = Eh + nE;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(complex_columns, dim3((unsigned)((n * nb + 255) / 256)), dim3(256), 0, st, A1, A2, n, nb, cc, G1, G2);
  TC_HIP(hipGetLastError());
  hipLaunchKernelGGL(qmc_pair_F, dim3((unsigned)((n * npair + 255) / 256)), dim3(256), 0, st, pts1, pts2, G2, n, nb,
                     npair, F);
  TC_HIP(hipGetLastError());
  if ((rc = eri_contract(G1, w, n, 2 * nb, F, npair, Eh, cws, st))) return rc;
  const int64_t nb4 = (int64_t)nb * nb * nb * nb;
  hipLaunchKernelGGL(complex_eri_assemble, dim3((unsigned)((nb4 + 255) / 256)), dim3(256), 0, st, Eh, nb, npair, eri);
  TC_HIP(hipGetLastError());
  return 0;
}

extern "C" int tcscf_eri_gradient(const double* pts1, const double* A1, const double* w1, int64_t n1,
                                  const double* pts2, const double* w2, int64_t n2, const double* A2,
                                  const double* dA2, int nb, int kernel, const tcscf_jastrow* jastrow, double scale,
                                  double* eri, void* stream) {
  using namespace tc;
  TC_CHECK(pts1 && A1 && w1 && pts2 && w2 && A2 && dA2 && eri && n1 > 0 && n2 > 0 && nb > 0,
           "tcscf_eri_gradient: bad arguments");
  TC_CHECK(kernel == TCSCF_K_JASTROW_GRAD && jastrow && jastrow->a12 > 0,
           "tcscf_eri_gradient: needs the vector kernel TCSCF_K_JASTROW_GRAD and a12 > 0");
  hipStream_t st = (hipStream_t)stream;
  const int64_t ncol = (int64_t)nb * nb, ldq = round_up(ncol, 128), n2p = round_up(n2, 16);
  const size_t nq = (size_t)3 * n2p * ldq, nx = (size_t)3 * n2p, nt = (size_t)n1 * ncol;
  int rc = 0;
  char* ws = (char*)scratch(kSlotSmall, sizeof(float4) * n2p +
                                            sizeof(double) * (nq + nx + nt + eri_contract_doubles(n1, nb, ncol)), &rc);
  if (rc) return rc;
  float4* x4 = (float4*)ws;  // the scratch base is 256-byte aligned
  double* Q = (double*)(ws + sizeof(float4) * n2p);
  double* xd = Q + nq;
  double* T = xd + nx;
  double* rest = T + nt;
  hipLaunchKernelGGL(pad_coords, dim3((unsigned)((n2p + 255) / 256)), dim3(256), 0, st, pts2, (const int*)nullptr, n2,
                     n2p, xd, x4);
  TC_HIP(hipGetLastError());
  hipLaunchKernelGGL(pack_grad_pairs, dim3((unsigned)((nq + 255) / 256)), dim3(256), 0, st, A2, dA2, w2, n2, nb, n2p,
                     ldq, Q);
  TC_HIP(hipGetLastError());
  constexpr int MB = 2;
  const double c = jastrow->c_same;  // f = c_same grad_1 u(r1, r2)
  KernJastrowGrad kern{jastrow->a12};
  dim3 grid((unsigned)(ldq / 128), (unsigned)((n1 + 32 * MB - 1) / (32 * MB)));
  hipLaunchKernelGGL((fgemm_f64_vdot_kernel<MB, KernJastrowGrad>), grid, dim3(256), 0, st, pts1, n1, xd, n2p, Q, ldq,
                     T, ncol, scale * c, kern);
  TC_HIP(hipGetLastError());
  return eri_contract(A1, w1, n1, nb, T, ncol, eri, rest, st);
}
/* ======================================================================== */
/* threeElectronIntegrals (src/integration/integrateAll.hpp:369-442):        */
/*   Q(p,q,r,s,t,u) = sum_r1 w1 A_p A_s V(q,t).V(r,u)                          */
/* from the packed c = 1 Jastrow-gradient tensor V (n1 x 3 x npair), scaled   */
/* by c^2.  Validation-sized (nb <= 16): one thread per output, r1 points     */
/* staged through LDS in tiles.                                              */
/* ======================================================================== */
namespace tc {

constexpr int kQMaxNb = 16;
constexpr int kQTile = 16;

__global__ __launch_bounds__(256) void three_electron_kernel(const double* __restrict__ A1, const double* __restrict__ w1,
                                                             int64_t n1, const double* __restrict__ V, int nb,
                                                             double c2, double* __restrict__ Q) {
  __shared__ double sA[kQTile][kQMaxNb];
  __shared__ double sV[kQTile][3][kQMaxNb * (kQMaxNb + 1) / 2];
  const int64_t npair = (int64_t)nb * (nb + 1) / 2;
  const int64_t nq = (int64_t)nb * nb * nb * nb * nb * nb;
  const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int u = 0, t = 0, s = 0, r = 0, q = 0, p = 0;
  if (o < nq) {
    int64_t x = o;
    u = (int)(x % nb); x /= nb;
    t = (int)(x % nb); x /= nb;
    s = (int)(x % nb); x /= nb;
    r = (int)(x % nb); x /= nb;
    q = (int)(x % nb); x /= nb;
    p = (int)x;
  }
  const int pqt = q <= t ? t * (t + 1) / 2 + q : q * (q + 1) / 2 + t;
  const int pru = r <= u ? u * (u + 1) / 2 + r : r * (r + 1) / 2 + u;
  double acc = 0;
  for (int64_t k0 = 0; k0 < n1; k0 += kQTile) {
    const int kn = (int)(n1 - k0 < kQTile ? n1 - k0 : kQTile);
    __syncthreads();
    for (int e = threadIdx.x; e < kn * nb; e += blockDim.x) {
      const int kk = e / nb, b = e - kk * nb;
      sA[kk][b] = A1[(k0 + kk) * nb + b];
    }
    for (int e = threadIdx.x; e < kn * 3 * (int)npair; e += blockDim.x) {
      const int kk = e / (3 * (int)npair), rem = e - kk * 3 * (int)npair;
      const int c = rem / (int)npair, pp = rem - c * (int)npair;
      sV[kk][c][pp] = V[((k0 + kk) * 3 + c) * npair + pp];
    }
    __syncthreads();
    if (o < nq)
      for (int kk = 0; kk < kn; ++kk) {
        const double d = sV[kk][0][pqt] * sV[kk][0][pru] + sV[kk][1][pqt] * sV[kk][1][pru] +
                         sV[kk][2][pqt] * sV[kk][2][pru];
        acc += w1[k0 + kk] * sA[kk][p] * sA[kk][s] * d;
      }
  }
  if (o < nq) Q[o] = c2 * acc;
}

}  // namespace tc

extern "C" int tcscf_three_electron(const double* A1, const double* w1, int64_t n1, const double* V, int nb, double c,
                                    double* Q, void* stream) {
  using namespace tc;
  TC_CHECK(A1 && w1 && V && Q && n1 > 0 && nb > 0, "tcscf_three_electron: bad arguments");
  TC_CHECK(nb <= kQMaxNb, "tcscf_three_electron: nb > 16 (the nb^6 tensor is a validation tool)");
  const int64_t nq = (int64_t)nb * nb * nb * nb * nb * nb;
  hipLaunchKernelGGL(three_electron_kernel, dim3((unsigned)((nq + 255) / 256)), dim3(256), 0, (hipStream_t)stream, A1,
                     w1, n1, V, nb, c * c, Q);
  TC_HIP(hipGetLastError());
  return 0;
}

/* ======================================================================== */
/* Batched contraction: many (r1 grid, r2 grid, basis, kernel) instances of   */
/* uneven size in one persistent launch.  The work queue is the list of       */
/* (instance, 16 MB-row block, 256-column block, component) tiles, instances  */
/* ordered by descending r2 length; workgroups take tiles from an atomic      */
/* counter until the queue is empty.  Each tile is the same bf16_tile as the  */
/* single-instance kernel, so results match it bit for bit.                   */
/* ======================================================================== */
namespace tc {

// Any of the three kernels, one scalar component at a time (kind uniform per tile).
struct KernAny {
  static constexpr int ncomp = 1;
  static constexpr bool absolute = false;
  int kind, comp;
  double a12, c2a2;
  template <class T>
  __device__ inline void operator()(T dx, T dy, T dz, T* out) const {
#pragma clang fp contract(off)
    const T s = dist2(dx, dy, dz);
    const T y = rsqrt_any(s);
    if (kind == TCSCF_K_COULOMB) out[0] = s > T(0) ? y : T(0);
    else if (kind == TCSCF_K_TC_SCALAR) out[0] = s > T(0) ? k_tc_scalar(s, y, a12, c2a2) : T(0);
    else out[0] = (s > T(0) ? k_grad_scale(s, y, a12) : T(0)) * (comp == 0 ? dx : (comp == 1 ? dy : dz));
  }
};

struct BatchDesc {
  const double* p1;
  const float4* x2;
  const us8* P0;
  void* F;  // FP64, or FP32 when out_f32
  int64_t n1, n2p, plane, ncb, npair;
  int64_t rblocks, cblocks, tiles;  // row blocks (16 MB rows), 256-column blocks, all tiles
  double scale, a12, c2a2;
  int kind, ncomp, out_f32;
};

template <int MB, int NPROD>
__global__ __launch_bounds__(256, 2) void bf16_batch_kernel(const BatchDesc* __restrict__ desc,
                                                            const int64_t* __restrict__ start, int ndesc,
                                                            int64_t t_begin, int64_t t_end,
                                                            unsigned long long* __restrict__ counter) {
  __shared__ us8 ktile[2][1][MB][2][64];
  __shared__ float4 crd[2][32];
  __shared__ int64_t cur;
  for (;;) {
    if (threadIdx.x == 0) cur = t_begin + (int64_t)atomicAdd(counter, 1ull);
    __syncthreads();
    const int64_t t = cur;
    __syncthreads();  // every thread has read cur before it is replaced
    if (t >= t_end) break;
    int lo = 0, hi = ndesc - 1;  // last instance with start <= t
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (start[mid] <= t) lo = mid;
      else hi = mid - 1;
    }
    const BatchDesc& d = desc[lo];
    int64_t local = t - start[lo];
    const int comp = (int)(local % d.ncomp);
    local /= d.ncomp;
    const int64_t cb = local % d.cblocks, rb = local / d.cblocks;
    const KernAny kern{d.kind, comp, d.a12, d.c2a2};
    void* Fc = d.out_f32 ? (void*)((float*)d.F + comp * d.npair) : (void*)((double*)d.F + comp * d.npair);
    bf16_tile<MB, NPROD, NPROD == 4 ? 2 : 1, KernAny>(ktile, crd, d.p1, d.n1, d.x2, d.n2p, d.P0, d.plane, d.ncb, Fc,
                                                      nullptr, d.out_f32, d.npair, d.ncomp * d.npair, d.scale, kern,
                                                      rb * (16 * MB), cb * 16);
  }
}

// The V tensors of a batch: one tile is (instance, 16-row block, 256-column
// block) with all three components of the Jastrow gradient, which share r12,
// its rsqrt, the gradient scale and the B fragments -- the tile of the
// single-instance V kernel, so results match it bit for bit (below
// kVLongWork units per instance, where the single call folds every kVFold
// steps as this kernel does).
template <int NPROD>
__global__ __launch_bounds__(256, (bf16_min_blocks<KernJastrowGrad, NPROD>())) void bf16_batch_v_kernel(
    const BatchDesc* __restrict__ desc, const int64_t* __restrict__ start, int ndesc, int64_t t_end,
    unsigned long long* __restrict__ counter) {
  __shared__ us8 ktile[2][3][1][2][64];
  __shared__ float4 crd[2][32];
  __shared__ int64_t cur;
  for (;;) {
    if (threadIdx.x == 0) cur = (int64_t)atomicAdd(counter, 1ull);
    __syncthreads();
    const int64_t t = cur;
    __syncthreads();  // every thread has read cur before it is replaced
    if (t >= t_end) break;
    int lo = 0, hi = ndesc - 1;  // last instance with start <= t
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (start[mid] <= t) lo = mid;
      else hi = mid - 1;
    }
    const BatchDesc& d = desc[lo];
    const int64_t local = t - start[lo];
    const int64_t cb = local % d.cblocks, rb = local / d.cblocks;
    KernJastrowGrad kern;
    kern.a12 = d.a12;
    bf16_tile<1, NPROD, kVFold, KernJastrowGrad, 4>(ktile, crd, d.p1, d.n1, d.x2, d.n2p, d.P0, d.plane,
                                                                d.ncb, d.F, nullptr, d.out_f32, d.npair, 3 * d.npair,
                                                                d.scale, kern, rb * 16, cb * 16);
  }
}

__global__ void zero_counter(unsigned long long* c) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *c = 0ull;
}

// Autotuned row-block height per product count (3, 4 or 5), process-wide.
std::mutex g_tune_mu;
std::map<int, int> g_tuned_mb;

}  // namespace tc

extern "C" int tcscf_batch_compute_F(const tcscf_batch_item* items, int count, int precision, void* stream) {
  using namespace tc;
  TC_CHECK(items && count > 0, "tcscf_batch_compute_F: bad arguments");
  for (int k = 0; k < count; ++k)
    TC_CHECK(items[k].out_dtype == TCSCF_F64 || items[k].out_dtype == TCSCF_F32,
             "tcscf_batch_compute_F: item " + std::to_string(k) + ": unknown out_dtype");
  hipStream_t st = (hipStream_t)stream;
  if (!is_bf16(precision)) {  // FP64 / FP32: one launch per instance
    for (int k = 0; k < count; ++k) {
      const tcscf_batch_item& it = items[k];
      const int rc = tcscf_compute_F(it.pts1_dev, it.n1, it.pts2_dev, it.w2_dev, it.ord2_dev, it.n2, it.A2_dev, it.nb,
                                     it.kernel, &it.jastrow, it.same_spin, precision, it.scale, it.F_dev,
                                     it.out_dtype, stream);
      if (rc) return rc;
    }
    return 0;
  }
  const int nsplit = precision == TCSCF_BF16X3 ? 2 : 3;
  const int nprod = precision == TCSCF_BF16X3 ? 3 : (precision == TCSCF_BF16X4 ? 4 : 5);
  // workspace: per instance pair-density planes + padded coordinates, then the counter
  std::vector<size_t> off(count + 1, 0);
  for (int k = 0; k < count; ++k) {
    const tcscf_batch_item& it = items[k];
    TC_CHECK(it.pts1_dev && it.pts2_dev && it.w2_dev && it.A2_dev && it.F_dev && it.n1 > 0 && it.n2 > 0 && it.nb > 0,
             "tcscf_batch_compute_F: bad item");
    TC_CHECK(it.kernel >= TCSCF_K_COULOMB && it.kernel <= TCSCF_K_JASTROW_GRAD, "tcscf_batch_compute_F: unknown kernel");
    TC_CHECK(it.kernel == TCSCF_K_COULOMB || it.jastrow.a12 > 0, "tcscf_batch_compute_F: TC kernels need a12 > 0");
    off[k + 1] = off[k] + round_up(tcscf_F_workspace_bytes(it.n2, it.nb, precision), 256);
  }
  int rc = 0;
  char* ws = (char*)scratch(kSlotPairs, off[count] + 256, &rc);
  if (rc) return rc;
  unsigned long long* counter = (unsigned long long*)(ws + off[count]);  // [0] F queue, [1] V queue
  std::vector<BatchDesc> desc(count);
  std::vector<int> order(count);
  std::iota(order.begin(), order.end(), 0);
  for (int k = 0; k < count; ++k) {
    const tcscf_batch_item& it = items[k];
    const int64_t npair = npair_of(it.nb), n2p = round_up(it.n2, 32), ncb = round_up(npair, 256) / 16,
                  nkb = n2p / 32, plane = nkb * ncb * 512;
    unsigned short* planes = (unsigned short*)(ws + off[k]);
    double* xd = (double*)(planes + nsplit * plane);
    float4* x4 = (float4*)(xd + 3 * n2p);
    const unsigned gp = (unsigned)((plane / 8 + 255) / 256);  // one thread per fragment lane
    if (nsplit == 2)
      hipLaunchKernelGGL(pack_pairs_bf16<2>, dim3(gp), dim3(256), 0, st, it.A2_dev, it.w2_dev, it.ord2_dev, it.n2,
                         it.nb, npair, nkb, ncb, planes);
    else
      hipLaunchKernelGGL(pack_pairs_bf16<3>, dim3(gp), dim3(256), 0, st, it.A2_dev, it.w2_dev, it.ord2_dev, it.n2,
                         it.nb, npair, nkb, ncb, planes);
    TC_HIP(hipGetLastError());
    hipLaunchKernelGGL(pad_coords, dim3((unsigned)((n2p + 255) / 256)), dim3(256), 0, st, it.pts2_dev, it.ord2_dev,
                       it.n2, n2p, xd, x4);
    TC_HIP(hipGetLastError());
    BatchDesc& d = desc[k];
    d.p1 = it.pts1_dev;
    d.x2 = x4;
    d.P0 = (const us8*)planes;
    d.F = it.F_dev;
    d.out_f32 = it.out_dtype == TCSCF_F32;
    d.n1 = it.n1;
    d.n2p = n2p;
    d.plane = plane / 8;
    d.ncb = ncb;
    d.npair = npair;
    d.cblocks = ncb / 16;
    d.scale = it.scale;
    d.kind = it.kernel;
    d.ncomp = it.kernel == TCSCF_K_JASTROW_GRAD ? 3 : 1;
    d.a12 = it.jastrow.a12;
    const double c = it.same_spin ? it.jastrow.c_same : it.jastrow.c_oppo;
    d.c2a2 = 2 * c * c * it.jastrow.a12 * it.jastrow.a12;
  }
  // largest r2 first, so the tail of a queue holds the short tiles; the V
  // tensors (three components per tile) form a queue of their own
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return desc[a].n2p > desc[b].n2p; });
  std::vector<BatchDesc> sorted, vdesc;
  for (int k = 0; k < count; ++k) (desc[order[k]].ncomp == 3 ? vdesc : sorted).push_back(desc[order[k]]);
  const int nf = (int)sorted.size(), nv = (int)vdesc.size();

  int dev = 0, ncu = 256;
  TC_HIP(hipGetDevice(&dev));
  TC_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  // descriptors of a queue for one row-block height (16 MB rows; per tile ncomp
  // components, or all of them in one tile for the V queue), in a scratch slot
  auto upload_list = [&](std::vector<BatchDesc>& list, int mb, bool all_comp, int slot, int64_t* total,
                         const BatchDesc** d_desc, const int64_t** d_start) -> int {
    const int n = (int)list.size();
    std::vector<int64_t> start(n);
    int64_t t = 0;
    for (int k = 0; k < n; ++k) {
      list[k].rblocks = (list[k].n1 + 16 * mb - 1) / (16 * mb);
      list[k].tiles = list[k].rblocks * list[k].cblocks * (all_comp ? 1 : list[k].ncomp);
      start[k] = t;
      t += list[k].tiles;
    }
    *total = t;
    int r = 0;
    char* meta = (char*)scratch(slot, sizeof(BatchDesc) * n + sizeof(int64_t) * n + 256, &r);
    if (r) return r;
    TC_HIP(hipMemcpyAsync(meta, list.data(), sizeof(BatchDesc) * n, hipMemcpyHostToDevice, st));
    TC_HIP(hipMemcpyAsync(meta + round_up(sizeof(BatchDesc) * n, 256), start.data(), sizeof(int64_t) * n,
                          hipMemcpyHostToDevice, st));
    TC_HIP(hipStreamSynchronize(st));  // host vectors go out of scope
    *d_desc = (const BatchDesc*)meta;
    *d_start = (const int64_t*)(meta + round_up(sizeof(BatchDesc) * n, 256));
    return 0;
  };
  auto upload = [&](int mb, int64_t* total, const BatchDesc** d_desc, const int64_t** d_start) -> int {
    return upload_list(sorted, mb, false, kSlotBatchMeta + (mb - 1), total, d_desc, d_start);
  };
  auto launch = [&](int mb, const BatchDesc* dd, const int64_t* ds, int64_t t0, int64_t t1) -> int {
    hipLaunchKernelGGL(zero_counter, dim3(1), dim3(64), 0, st, counter);
    TC_HIP(hipGetLastError());
    const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(t1 - t0, (int64_t)ncu * 2));
#define TC_BATCH(M, P) \
  hipLaunchKernelGGL((bf16_batch_kernel<M, P>), dim3(grid), dim3(256), 0, st, dd, ds, nf, t0, t1, counter)
    if (nprod == 5) {
      if (mb == 1) TC_BATCH(1, 5);
      else TC_BATCH(2, 5);
    } else if (nprod == 4) {
      if (mb == 1) TC_BATCH(1, 4);
      else TC_BATCH(2, 4);
    } else {
      if (mb == 1) TC_BATCH(1, 3);
      else TC_BATCH(2, 3);
    }
#undef TC_BATCH
    TC_HIP(hipGetLastError());
    return 0;
  };
  // the V queue: one persistent launch over its three-component tiles
  if (nv) {
    int64_t vtotal;
    const BatchDesc* vd;
    const int64_t* vs;
    if ((rc = upload_list(vdesc, 1, true, kSlotBatchMetaV, &vtotal, &vd, &vs))) return rc;
    hipLaunchKernelGGL(zero_counter, dim3(1), dim3(64), 0, st, counter + 1);
    TC_HIP(hipGetLastError());
    const int per_cu = nprod == 5   ? bf16_min_blocks<KernJastrowGrad, 5>()
                       : nprod == 4 ? bf16_min_blocks<KernJastrowGrad, 4>()
                                    : bf16_min_blocks<KernJastrowGrad, 3>();
    const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(vtotal, (int64_t)ncu * per_cu));
    if (nprod == 5)
      hipLaunchKernelGGL((bf16_batch_v_kernel<5>), dim3(grid), dim3(256), 0, st, vd, vs, nv, vtotal, counter + 1);
    else if (nprod == 4)
      hipLaunchKernelGGL((bf16_batch_v_kernel<4>), dim3(grid), dim3(256), 0, st, vd, vs, nv, vtotal, counter + 1);
    else
      hipLaunchKernelGGL((bf16_batch_v_kernel<3>), dim3(grid), dim3(256), 0, st, vd, vs, nv, vtotal, counter + 1);
    TC_HIP(hipGetLastError());
  }
  if (!nf) {
    TC_HIP(hipStreamSynchronize(st));
    return 0;
  }

  // Autotuner: the row-block height (16 or 32 rows per tile) is timed on the
  // head of the queue (the largest tiles) the first time a precision is seen
  // in this process, and the faster one is kept.
  const int key = nprod;
  int mb_best = 0;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto f = g_tuned_mb.find(key);
    if (f != g_tuned_mb.end()) mb_best = f->second;
  }
  const char* env = std::getenv("TCSCF_BATCH_MB");
  if (env) mb_best = std::atoi(env) == 1 ? 1 : 2;
  if (!mb_best) {
    float best = 1e30f;
    hipEvent_t e0, e1;
    TC_HIP(hipEventCreate(&e0));
    TC_HIP(hipEventCreate(&e1));
    for (int mb : {1, 2}) {
      int64_t total;
      const BatchDesc* dd;
      const int64_t* ds;
      if ((rc = upload(mb, &total, &dd, &ds))) return rc;
      const int64_t probe = std::min<int64_t>(total, std::max<int64_t>(4 * ncu, total / 16));
      if ((rc = launch(mb, dd, ds, 0, probe))) return rc;  // warm-up
      TC_HIP(hipEventRecord(e0, st));
      if ((rc = launch(mb, dd, ds, 0, probe))) return rc;
      TC_HIP(hipEventRecord(e1, st));
      TC_HIP(hipEventSynchronize(e1));
      float ms = 0;
      TC_HIP(hipEventElapsedTime(&ms, e0, e1));
      // normalise to rows covered: the probe of mb = 1 covers half the rows per tile
      const float per_row = ms / (float)(probe * mb);
      if (per_row < best) {
        best = per_row;
        mb_best = mb;
      }
    }
    TC_HIP(hipEventDestroy(e0));
    TC_HIP(hipEventDestroy(e1));
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tuned_mb[key] = mb_best;
  }
  int64_t total;
  const BatchDesc* dd;
  const int64_t* ds;
  if ((rc = upload(mb_best, &total, &dd, &ds))) return rc;
  if ((rc = launch(mb_best, dd, ds, 0, total))) return rc;
  TC_HIP(hipStreamSynchronize(st));
  return 0;
}

// Rows per tile the autotuner chose for this precision (16 or 32), 0 before the
// first batch at that precision.
extern "C" int tcscf_batch_tuned_rows(int precision) {
  using namespace tc;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto f = g_tuned_mb.find(precision == TCSCF_BF16X3 ? 3 : (precision == TCSCF_BF16X4 ? 4 : 5));
  return f == g_tuned_mb.end() ? 0 : 16 * f->second;
}
