// Transcorrelated Fock build and SCF driver (TCHartreeFock::compute,
// src/HartreeFock.cpp:1007-1362), real basis.
//
// Inputs per r1 point n: packed F_same[n][p], F_oppo[n][p] (scalar TC kernel)
// and the Jastrow-gradient tensors, either as the reference has them
// (V_same[n][xyz][p] and V_oppo[n][xyz][p], independent) or, for a Jastrow
// whose opposite-spin coefficients are one multiple of the same-spin ones, as
// one c = 1 tensor V with V_same = c_same V, V_oppo = c_oppo V; basis values
// A[n][k] and gradients dA[n][k][xyz].
//
// Every term of the reference's G2 / G3 assembly (:1157-1224) is one of
//   (a) a rank-one product L_j R_i of per-point vectors      -> "vectors"
//   (b) a per-point 3-vector h dotted into V(n, j, i)        -> "V-linear"
// Pass 1 (tc_point_kernel, one wave per point) forms every per-point
// intermediate of the reference (:1118-1148) and writes the compact per-point
// vectors and the h vectors; pass 2 (tc_rank_kernel, FP64 MFMA) sums the
// rank-one terms over points per target matrix; pass 3 (tc_vlin_kernel) sums
// h . V over points.  Partials per workgroup are reduced in a fixed order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstring>
#include <limits>
#include <map>
#include <mutex>
#include <numeric>
#include <vector>

#include "common.hpp"
#include "dense.hpp"
#include "device_util.hpp"
#include "scf_steps.hpp"

namespace tc {
namespace tcs {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int kBlocks = 512;

constexpr int kWMain = 5;   // occupied orbitals of W = V C per pass over j (the first in the main pass)
constexpr int kWChunk = 2;  // orbitals of W per LDS round of |U|^2

// Compact per-point output, per spin s, vectors of nb doubles (lane i = basis
// index i):
//   0        u0 = w (faa A - FaiS)                      G2 : A (x) u0
//   1        u1 = w (VT . dA - Vai)                     VEC: A (x) u1
//   2        u2 = w (|VT|^2 - sNS - sNO_other) A + Z    G3 : A (x) u2
//   3        u3 = Z = 2 w (SVbiVT - VTSI . VT)          G3 : u3 (x) A
//   4..6     t_c = VTSI_c                               VEC: (-w t_c) (x) dA_c
//                                                       G3 : (2 w t_c) (x) t_c
//   7+3(t nmax+b)+c  W^t_bc = sum_j C_s[j][b] V^t_c(j, .)  G3 : (alpha_t W^t_bc) (x) W^t_bc
// with t = 0 for V_same and t = 1 for V_oppo (independent tensors only),
// alpha_0 = -2 w occ_s, alpha_1 = -2 w occ_o per (point, spin); with one
// proportional tensor t = 0 only and alpha_0 = -2 w (occ_s c_s^2 + occ_o c_o^2).
// x (x) y is the rank-one matrix x_j y_i.  (Vja == VTSI for symmetric V.)
// The reference's V(j,i) . scaledGradients terms (:1172-1174) are symmetric in
// (j, i) and cancel in the antisymmetrization G2 = scalar/2 + vec/2 - vec^T/2,
// so they are not formed.
constexpr int kVecFixed = 7;

// Fs, Fo and V hold the kernel's storage type ST (double, or float for the
// mixed-precision build); everything else, and all arithmetic, is FP64.
struct PointArgs {
  const void* Fs;     // [n1][npair]
  const void* Fo;     // [n1][npair]
  const void* V;      // [n1][3][npair]: V_same, or the c = 1 tensor
  const void* Vo;     // [n1][3][npair]: V_oppo (independent tensors), else unused
  const double* A;    // [n1][nb]
  const double* dA;   // [n1][nb][3]
  const double* w;    // [n1]
  const double* D;    // [2][nb][nb]
  const double* C;    // [2][nb][nmax]
  double c_same, c_oppo;  // 1, 1 for independent tensors
  int64_t n1, npair;
  int nb, nmax, nocc0, nocc1, nvec;
  double* vecs;       // [n1][2][nvec][nb]
  double* alpha;      // [n1][2 spins][2 tensors]
  double* hvec;       // [n1][2 spins][2 tensors][3]: the G3 V-linear vectors
};

// Cross-lane sums in registers: DPP moves within a row of 16 lanes (pairs at
// lane distance 1 and 2, then the half-row and row mirrors), the gfx950
// permlane swaps across rows (16 and 32 lanes apart).  Each step adds a lane's
// value to its partner's; fp addition is commutative, so both partners hold
// the same bits and every lane of the group ends with the identical sum.
template <int CTRL>
__device__ inline double dpp_move(double v) {
  const unsigned long long b = __double_as_longlong(v);
  const unsigned lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, 0xF, 0xF, false);
  const unsigned hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// v + v of the lane 16 (ROWS = 1) or 32 (ROWS = 2) lanes away.
template <int ROWS>
__device__ inline double swap_sum(double v) {
  const unsigned long long b = __double_as_longlong(v);
  const unsigned lo = (unsigned)b, hi = (unsigned)(b >> 32);
  unsigned __attribute__((ext_vector_type(2))) l, h;
  if (ROWS == 1) {
    l = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
    h = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
  } else {
    l = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
    h = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
  }
  const double x = __longlong_as_double((long long)(((unsigned long long)h[0] << 32) | l[0]));
  const double y = __longlong_as_double((long long)(((unsigned long long)h[1] << 32) | l[1]));
  return x + y;
}

__device__ inline double wave_sum(double v) {
  v += dpp_move<0xB1>(v);   // quad_perm [1,0,3,2]
  v += dpp_move<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp_move<0x141>(v);  // row_half_mirror
  v += dpp_move<0x140>(v);  // row_mirror
  return swap_sum<2>(swap_sum<1>(v));
}

// A wave-uniform value (e.g. a wave_sum result) moved to scalar registers.
__device__ inline double uniform(double v) {
  const unsigned long long b = __double_as_longlong(v);
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)b);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(b >> 32));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

__device__ inline void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Per-wave LDS doubles of the point vectors for NS spins: A, dA, then X, VTSI,
// kS and E per spin.
__host__ __device__ inline int64_t point_wave_elems(int nb, int ns) { return (4 + 10 * (int64_t)ns) * nb; }

// Sum of v over the lanes that share a basis index i (lane = jp * nbp + i).
__device__ inline double part_sum(double v, int nbp) {
  if (nbp == 16) v = swap_sum<1>(v);
  if (nbp <= 32) v = swap_sum<2>(v);
  return v;
}

// Doubles of LDS that the staged V rows ([nt][3][npair] of element type ST,
// nt tensors) of a wave take, 0 when not staged.
template <class ST>
__host__ __device__ inline int64_t staged_doubles(bool stage, int nb, int nt) {
  return stage ? (3 * nt * ((int64_t)nb * (nb + 1) / 2) * (int64_t)sizeof(ST) + 7) / 8 : 0;
}

// Pass 1: one wave per r1 point.  Lane = jp * nbp + i: i is the basis index
// (nbp = 16, 32 or 64 >= nb) and the 64 / nbp lane groups jp split every loop
// over a second basis index, so each lane walks nb * nbp / 64 terms.  The
// point's packed F_same / F_oppo / V rows are read at triangular (j, i)
// addresses, a gather that touches a separate cache line for almost every
// lane.  One pass over j forms every F- and V-linear intermediate (F / V
// against D, X and kS, and W = V C for the first kWMain orbitals), a second
// pass V . E once E is known.  STAGE copies the V rows (read in several
// passes) into the wave's LDS area by direct-to-LDS loads first, so those
// gathers hit LDS; it pays when 8 waves of a workgroup fit in LDS and hide the
// latency of the dependent passes.  D and C are shared in LDS, the per-point
// vectors live in a small per-wave LDS area.  NS = 1 (closed shell with
// identical spin densities and orbitals) forms spin 0 only, the spin-1 terms
// being the same numbers.  GV: V_same and V_oppo are independent tensors (the
// reference's VjiSame / VjiOppo); otherwise one c = 1 tensor scaled by c_same
// and c_oppo.
template <class ST, bool STAGE, int NS, bool GV>
__global__ __launch_bounds__((NS == 1 && sizeof(ST) == 4 && !GV) ? 768 : 512, 1) void tc_point_kernel(PointArgs a) {
  constexpr int NT = GV ? 2 : 1;  // V tensors
  extern __shared__ double smem[];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nwaves = blockDim.x >> 6;
  const int nb = a.nb, nmax = a.nmax;
  const int nbp = nb <= 16 ? 16 : (nb <= 32 ? 32 : 64), parts = 64 / nbp;
  const int64_t np = a.npair, nb2 = (int64_t)nb * nb;
  double* Ds = smem;                       // [NS][nb][nb]
  double* Cs = Ds + NS * nb2;              // [NS][nb][nmax]
  double* base = Cs + NS * (int64_t)nb * nmax + wave * (point_wave_elems(nb, NS) + staged_doubles<ST>(STAGE, nb, NT));
  double* As = base;                       // [nb]
  double* dAs = As + nb;                   // [nb][3]
  double* Xs = dAs + 3 * nb;               // [NS][nb]
  double* VT = Xs + NS * nb;               // VTSI [NS][nb][3]
  double* kS = VT + 3 * NS * nb;           // [NS][nb][3]
  double* Es = kS + 3 * NS * nb;           // [NS][nb][3]
  double* Wc = kS;                         // [NS][nb][kWChunk][3], once kS and Es are consumed
  ST* Sv = reinterpret_cast<ST*>(Es + 3 * NS * nb);  // STAGE: the point's V rows [NT][3][npair]
  for (int64_t e = tid; e < NS * nb2; e += blockDim.x) Ds[e] = a.D[e];
  for (int64_t e = tid; e < NS * (int64_t)nb * nmax; e += blockDim.x) Cs[e] = a.C[e];
  __syncthreads();

  const int i = lane & (nbp - 1), jp = lane / nbp;
  const bool act = i < nb;
  const bool own = act && jp == 0;  // the lane that stores / counts basis index i
  const int tri_i = i * (i + 1) / 2;
  const double cs = a.c_same, co = a.c_oppo;
  const int nocc[2] = {a.nocc0, a.nocc1};
  const int64_t nvw = (int64_t)gridDim.x * nwaves;
  auto pidx = [&](int j) -> int64_t { return j <= i ? tri_i + j : j * (j + 1) / 2 + i; };
  // body(j, v0, v1, v2) over the lane's j for one V tensor
  auto for_j = [&](const ST* vv, auto&& body) {
    if (act)
      for (int j = jp; j < nb; j += parts) {
        const int64_t pj = pidx(j);
        body(j, double(vv[pj]), double(vv[np + pj]), double(vv[2 * np + pj]));
      }
  };

  // wreg[s][b][c] += C_s[j][b0 + b] V_c(j, i) for the occupied orbitals b0 + b
  auto w_accumulate = [&](double (&wreg)[2][kWMain][3], int b0, int j, double v0, double v1, double v2) {
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const double* crow = Cs + ((int64_t)s * nb + j) * nmax + b0;
#pragma unroll
      for (int b = 0; b < kWMain; ++b)
        if (b0 + b < nocc[s]) {
          const double cb = crow[b];
          wreg[s][b][0] += cb * v0;
          wreg[s][b][1] += cb * v1;
          wreg[s][b][2] += cb * v2;
        }
    }
  };

  for (int64_t n = (int64_t)blockIdx.x * nwaves + wave; n < a.n1; n += nvw) {
    const ST* __restrict__ fs = (const ST*)a.Fs + n * np;
    const ST* __restrict__ fo = (const ST*)a.Fo + n * np;
    const ST* __restrict__ vv = (const ST*)a.V + n * 3 * np;
    const ST* __restrict__ vo = GV ? (const ST*)a.Vo + n * 3 * np : vv;
    if (STAGE) {
      // the V rows copied by direct-to-LDS loads (lane l of a load writes
      // dword k + l of the run), all loads in flight together, one wait
      const int64_t ndw = 3 * np * (int64_t)(sizeof(ST) / 4);
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const unsigned* g = reinterpret_cast<const unsigned*>(t ? vo : vv);
        unsigned* l = reinterpret_cast<unsigned*>(Sv) + t * ndw;
        for (int64_t k = 0; k < ndw; k += 64)
          if (k + lane < ndw)
            __builtin_amdgcn_global_load_lds(g + k + lane, (__attribute__((address_space(3))) void*)(l + k), 4, 0, 0);
      }
      asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
      vv = Sv;
      if (GV) vo = Sv + 3 * np;
    }
    if (own) {
      As[i] = a.A[n * nb + i];
      for (int c = 0; c < 3; ++c) dAs[i * 3 + c] = a.dA[(n * nb + i) * 3 + c];
    }
    wave_sync_lds();
    // X_s[i] = sum_k D_s[k][i] A[k];  kS_s[i] = sum_k D_s[k][i] dA[k]
    for (int s = 0; s < NS; ++s) {
      double x = 0, k0 = 0, k1 = 0, k2 = 0;
      if (act)
        for (int k = jp; k < nb; k += parts) {
          const double d = Ds[s * nb2 + k * nb + i];
          x += d * As[k];
          k0 += d * dAs[k * 3];
          k1 += d * dAs[k * 3 + 1];
          k2 += d * dAs[k * 3 + 2];
        }
      x = part_sum(x, nbp);
      k0 = part_sum(k0, nbp);
      k1 = part_sum(k1, nbp);
      k2 = part_sum(k2, nbp);
      if (own) {
        Xs[s * nb + i] = x;
        kS[(s * nb + i) * 3] = k0;
        kS[(s * nb + i) * 3 + 1] = k1;
        kS[(s * nb + i) * 3 + 2] = k2;
      }
    }
    wave_sync_lds();
    const double Ai = act ? As[i] : 0.0;
    double dAi[3] = {0, 0, 0};
    if (act)
      for (int c = 0; c < 3; ++c) dAi[c] = dAs[i * 3 + c];
    double occ[2];
    for (int s = 0; s < NS; ++s) {
      const double x = own ? Xs[s * nb + i] : 0.0;
      occ[s] = uniform(wave_sum(x * Ai));  // calculateOccupiedElectronDensity
    }
    // one pass over j: Faa, FaiS, V . D (both tensors), Vai, Q (= VTSI / c_s)
    // and W_s[i][b][c] = sum_j C_s[j][b] V_c(j, i) for the first kWMain orbitals
    double faa_p[2] = {0, 0}, fai[2] = {0, 0}, vd[2][3] = {{0, 0, 0}, {0, 0, 0}}, vai1[2] = {0, 0};
    double vdo[2][3] = {{0, 0, 0}, {0, 0, 0}};  // GV: V_oppo . D_s
    double q[2][3] = {{0, 0, 0}, {0, 0, 0}};
    double wreg[2][kWMain][3];
#pragma unroll
    for (int s = 0; s < NS; ++s)
#pragma unroll
      for (int b = 0; b < kWMain; ++b) wreg[s][b][0] = wreg[s][b][1] = wreg[s][b][2] = 0;
    if (act)
      for (int j = jp; j < nb; j += parts) {
        const int64_t pj = pidx(j);
        const double f_s = double(fs[pj]), f_o = double(fo[pj]);
        const double v0 = double(vv[pj]), v1 = double(vv[np + pj]), v2 = double(vv[2 * np + pj]);
        double o0 = 0, o1 = 0, o2 = 0;
        if (GV) {
          o0 = double(vo[pj]);
          o1 = double(vo[np + pj]);
          o2 = double(vo[2 * np + pj]);
        }
        w_accumulate(wreg, 0, j, v0, v1, v2);
#pragma unroll
        for (int s = 0; s < NS; ++s) {
          const double dsj = Ds[s * nb2 + j * nb + i], dbj = Ds[(NS == 1 ? s : 1 - s) * nb2 + j * nb + i];
          faa_p[s] += f_s * dsj + f_o * dbj;
          const double xj = Xs[s * nb + j];
          fai[s] += xj * f_s;
          vd[s][0] += v0 * dsj;
          vd[s][1] += v1 * dsj;
          vd[s][2] += v2 * dsj;
          if (GV) {
            vdo[s][0] += o0 * dsj;
            vdo[s][1] += o1 * dsj;
            vdo[s][2] += o2 * dsj;
          }
          const double* ks = kS + (s * nb + j) * 3;
          vai1[s] += v0 * ks[0] + v1 * ks[1] + v2 * ks[2];
          q[s][0] += xj * v0;
          q[s][1] += xj * v1;
          q[s][2] += xj * v2;
        }
      }
    // W blocks of kWMain orbitals per tensor t: (t = 0, b0 = 0) from the main
    // pass, the rest from passes of their own; each block goes to vout and,
    // kWChunk orbitals at a time through LDS (over kS and Es, consumed / not yet
    // written), adds its part of sum_ab |U^t_s[a][b]|^2,
    // U^t_s[a][b][c] = sum_l C_s[l][a] W^t_s[l][b][c] (calculateVab)
    double* vout = a.vecs + (int64_t)n * 2 * a.nvec * nb;
    double nacc[NT][2];
#pragma unroll
    for (int t = 0; t < NT; ++t) nacc[t][0] = nacc[t][1] = 0;
#pragma unroll
    for (int t = 0; t < NT; ++t)
      for (int b0 = 0; b0 < nmax; b0 += kWMain) {
        if (t > 0 || b0 > 0) {
#pragma unroll
          for (int s = 0; s < NS; ++s)
#pragma unroll
            for (int b = 0; b < kWMain; ++b) wreg[s][b][0] = wreg[s][b][1] = wreg[s][b][2] = 0;
          for_j(t ? vo : vv, [&](int j, double v0, double v1, double v2) { w_accumulate(wreg, b0, j, v0, v1, v2); });
        }
#pragma unroll
        for (int c0 = 0; c0 < kWMain; c0 += kWChunk) {
          if (b0 + c0 >= nmax) break;
#pragma unroll
          for (int s = 0; s < NS; ++s)
#pragma unroll
            for (int b = c0; b < c0 + kWChunk && b < kWMain; ++b)
#pragma unroll
              for (int c = 0; c < 3; ++c) {
                const double wv = part_sum(b0 + b < nocc[s] ? wreg[s][b][c] : 0.0, nbp);
                if (own) {
                  Wc[((s * nb + i) * kWChunk + b - c0) * 3 + c] = wv;
                  if (b0 + b < nmax)
                    vout[((int64_t)s * a.nvec + kVecFixed + 3 * (t * nmax + b0 + b) + c) * nb + i] = wv;
                }
              }
          wave_sync_lds();
          const int lim = std::min(nmax - b0 - c0, kWMain - c0), nbc = lim < kWChunk ? lim : kWChunk;
#pragma unroll
          for (int s = 0; s < NS; ++s) {
            const int nu = nocc[s] * nbc * 3;  // (a, b in the chunk, c)
            for (int u = lane; u < nu; u += 64) {
              const int aa = u / (nbc * 3), rem = u - aa * nbc * 3, bb = rem / 3, c = rem - bb * 3;
              double uv = 0;
              for (int l = 0; l < nb; ++l)
                uv += Cs[((int64_t)s * nb + l) * nmax + aa] * Wc[((s * nb + l) * kWChunk + bb) * 3 + c];
              nacc[t][s] += uv * uv;
            }
          }
          wave_sync_lds();
        }
      }
    double faa[2], VTv[2][3];
    double vdsum[2][3], vdosum[2][3];
    for (int s = 0; s < NS; ++s) {
      faa[s] = uniform(wave_sum(faa_p[s]));  // FaaSamePlusOppo
      for (int c = 0; c < 3; ++c) vdsum[s][c] = uniform(wave_sum(vd[s][c]));
      if (GV)
        for (int c = 0; c < 3; ++c) vdosum[s][c] = uniform(wave_sum(vdo[s][c]));
      fai[s] = part_sum(fai[s], nbp);
      vai1[s] = part_sum(vai1[s], nbp);
      for (int c = 0; c < 3; ++c) q[s][c] = part_sum(q[s][c], nbp);
    }
    // VaaSamePlusOppo == VTilde: V_same . D_s + V_oppo . D_s'
    for (int s = 0; s < NS; ++s)
      for (int c = 0; c < 3; ++c)
        VTv[s][c] = GV ? vdsum[s][c] + vdosum[NS == 1 ? s : 1 - s][c] : cs * vdsum[s][c] + co * vdsum[NS == 1 ? s : 1 - s][c];
    if (own)
      for (int s = 0; s < NS; ++s)
        for (int c = 0; c < 3; ++c) VT[(s * nb + i) * 3 + c] = cs * q[s][c];  // VTildeSubI (== Vja)
    wave_sync_lds();
    // E_s[k] = sum_l D_s[l][k] VTSI_s[l]
    for (int s = 0; s < NS; ++s) {
      double e0 = 0, e1 = 0, e2 = 0;
      if (act)
        for (int l = jp; l < nb; l += parts) {
          const double d = Ds[s * nb2 + l * nb + i];
          const double* t = VT + (s * nb + l) * 3;
          e0 += d * t[0];
          e1 += d * t[1];
          e2 += d * t[2];
        }
      e0 = part_sum(e0, nbp);
      e1 = part_sum(e1, nbp);
      e2 = part_sum(e2, nbp);
      if (own) {
        Es[(s * nb + i) * 3] = e0;
        Es[(s * nb + i) * 3 + 1] = e1;
        Es[(s * nb + i) * 3 + 2] = e2;
      }
    }
    wave_sync_lds();
    // SumOverb_Vbi_dot_VTildeb_s[i] = c_s sum_k V(k,i) . E_s[k]
    double sv1[2] = {0, 0};
    for_j(vv, [&](int k, double v0, double v1, double v2) {
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        const double* e = Es + (s * nb + k) * 3;
        sv1[s] += v0 * e[0] + v1 * e[1] + v2 * e[2];
      }
    });
    sv1[0] = part_sum(sv1[0], nbp);
    sv1[1] = part_sum(sv1[1], nbp);
    double scvt[2][3], nrmU[NT][2];
    for (int s = 0; s < NS; ++s) {
      const double x = own ? Xs[s * nb + i] : 0.0;
      for (int c = 0; c < 3; ++c) scvt[s][c] = uniform(wave_sum(own ? x * VT[(s * nb + i) * 3 + c] : 0.0));  // SumOverb_conjVTildeb_occb
#pragma unroll
      for (int t = 0; t < NT; ++t) nrmU[t][s] = uniform(wave_sum(nacc[t][s]));
    }
    const double wn = a.w[n];
    for (int s = 0; s < NS; ++s) {
      const int sb = NS == 1 ? s : 1 - s;
      double* vs = vout + (int64_t)s * a.nvec * nb;
      const double* vts = act ? VT + (s * nb + i) * 3 : nullptr;
      const double vtsi0 = act ? vts[0] : 0, vtsi1 = act ? vts[1] : 0, vtsi2 = act ? vts[2] : 0;
      const double vtdot = VTv[s][0] * vtsi0 + VTv[s][1] * vtsi1 + VTv[s][2] * vtsi2;
      const double vtn = VTv[s][0] * VTv[s][0] + VTv[s][1] * VTv[s][1] + VTv[s][2] * VTv[s][2];
      const double svbi = cs * sv1[s];
      const double Z = 2 * wn * (svbi - vtdot);
      // sumOfNormsVabSame_s and sumOfNormsVabOppo of the other spin
      const double sNS = GV ? nrmU[0][s] : cs * cs * nrmU[0][s];
      const double sNOb = GV ? nrmU[NT - 1][sb] : co * co * nrmU[0][sb];
      if (own) {
        vs[0 * nb + i] = wn * (faa[s] * Ai - fai[s]);
        vs[1 * nb + i] = wn * (VTv[s][0] * dAi[0] + VTv[s][1] * dAi[1] + VTv[s][2] * dAi[2] - cs * vai1[s]);
        vs[2 * nb + i] = wn * (vtn - sNS - sNOb) * Ai + Z;
        vs[3 * nb + i] = Z;
        vs[4 * nb + i] = vtsi0;
        vs[5 * nb + i] = vtsi1;
        vs[6 * nb + i] = vtsi2;
      }
      double* al = a.alpha + (n * 2 + s) * 2;
      if (lane == 0) {
        if (GV) {
          al[0] = -2 * wn * occ[s];
          al[1] = -2 * wn * occ[sb];
        } else {
          al[0] = -2 * wn * (occ[s] * cs * cs + occ[sb] * co * co);
        }
      }
      // G3's V-linear terms: h_same . V_same(j,i) + h_oppo . V_oppo(j,i)
      if (lane < 3) {
        const int c = lane;
        double* h = a.hvec + ((int64_t)n * 2 + s) * 6;
        if (GV) {
          h[c] = wn * (2 * occ[s] * VTv[s][c] - 2 * scvt[s][c]);
          h[3 + c] = wn * (2 * occ[sb] * VTv[sb][c] - 2 * scvt[sb][c]);
        } else {
          h[c] = wn * (2 * occ[s] * cs * VTv[s][c] - 2 * cs * scvt[s][c] + 2 * occ[sb] * co * VTv[sb][c] -
                       2 * co * scvt[sb][c]);
        }
      }
    }
    wave_sync_lds();
  }
}

/* ------------------------------------------------------------------------
 * Pass 1 in column layout (nb <= 32) for D_s = C_s C_s^T, the density every
 * SCF iteration forms from its occupied orbitals.  Then every F- and V-linear
 * intermediate of a point is a product M C of one of the point's matrices
 * M in {F_same, F_oppo, V_c, Vo_c} with the occupied orbitals (columns of
 * spin 0, then spin 1):
 *   X_s = C_s x_s, x_s = C_s^T A          kS_s,c = C_s y_s,c, y_s,c = C_s^T dA_c
 *   Faa_s = tr(C_s^T Fs C_s) + tr(C_sb^T Fo C_sb)     FaiS_s = (Fs C_s) x_s
 *   U_s,c = C_s^T V_c C_s:  V_c . D_s = tr(U_s,c),  sum |U|^2 as before
 *   Q_s,c = (V_c C_s) x_s    Vai_s = sum_c (V_c C_s) y_s,c
 *   E_s,c = D_s VTSI_s,c = C_s g_s,c with g_s,c = c_s U_s,c x_s, so
 *   V(k, i) . E_s = sum_c (V_c C_s) g_s,c and SumOverb_conjVTildeb_occb_c = x_s . g_s,c
 * One pass over j, uniform across the wave with C[j][.] in scalar registers,
 * accumulates P_M[i][col] = sum_j M[j][i] C[j][col]: lane h * 32 + i holds
 * column i of the matrices of half h (h = 0: Fs, Fo, V_0 [, Vo_0]; h = 1:
 * V_2, [Vo_2 |  -], V_1 [, Vo_1]).  The remaining cross-lane sums (U, x, y)
 * are dot products over l of C[l][a] with the columns of W = V C, A and dA,
 * one lane per entry, through LDS.  The G3 V-linear terms h . V(j, i) are
 * added here from the staged V rows, so the build reads V once.
 * ---------------------------------------------------------------------- */
struct ColArgs {
  const void* Fs;
  const void* Fo;
  const void* V;
  const void* Vo;
  const double* A;
  const double* dA;
  const double* w;
  double c_same, c_oppo;
  int64_t n1, npair;
  int nb, nmax, nocc0, nocc1, nvec, nent;
  double* vecs;   // [n1][2][nvec][nb]
  double* alpha;  // [n1][2][2]
  double* vpart;  // [gridDim.x][NS][npair]: V-linear partial sums per workgroup
};

// Cc[j][col] (ld noc): the occupied orbitals of spin 0 in columns [0, nocc0),
// of spin 1 (ns = 2) in [nocc0, nocc0 + nocc1), zero elsewhere; Dw[s][p]
// (ld kRowStride): D_s packed by pair p = i (i + 1) / 2 + j (j <= i), doubled
// off the diagonal, zero past npair.
constexpr int kDwStride = 576;
__global__ __launch_bounds__(256) void tc_pack_cols(const double* __restrict__ C, const double* __restrict__ D, int nb,
                                                    int nmax, int nocc0, int nocc1, int ns, int noc,
                                                    double* __restrict__ Cc, double* __restrict__ Dw) {
  for (int e = threadIdx.x; e < nb * noc; e += blockDim.x) {
    const int j = e / noc, col = e - j * noc;
    double v = 0;
    if (col < nocc0)
      v = C[(int64_t)j * nmax + col];
    else if (ns == 2 && col - nocc0 < nocc1)
      v = C[((int64_t)nb + j) * nmax + col - nocc0];
    Cc[e] = v;
  }
  const int np = nb * (nb + 1) / 2;
  for (int e = threadIdx.x; e < 2 * kDwStride; e += blockDim.x) {
    const int s = e / kDwStride, p = e - s * kDwStride;
    double v = 0;
    if (p < np) {
      int ii = (int)((sqrt(8.0 * p + 1.0) - 1.0) * 0.5);
      while (ii * (ii + 1) / 2 > p) --ii;
      while ((ii + 1) * (ii + 2) / 2 <= p) ++ii;
      const int jj = p - ii * (ii + 1) / 2;
      v = (jj == ii ? 1.0 : 2.0) * D[((int64_t)s * nb + jj) * nb + ii];
    }
    Dw[e] = v;
  }
}

// LDS layout of the column kernel, in doubles.  Per workgroup: Cb[32][NOC]
// (the orbitals), the dot-product entry table (nent ints) and the task table
// of the step after it (ntask ints).  Per wave: the staged rows (NROW rows of
// kRowStride ST; later in the point reused for the per-lane results
// Rq[NS][10][32]), Wt[nb][LDW] (per basis index: the W = V C columns, Fs C, A
// and dA), Res[nent] (dot products), Xg[NS][4][NOC] (x and y by column),
// Gs[NS][3][NOC] (g by column), Tn[NS][NT][3][2] (trace and sum of squares of
// U) and Sc[NS][4] (occ, x . g_c).
static_assert(kDwStride == 576, "");
constexpr int kRowStride = kDwStride;  // staged row stride (elements): 9 loads of 64 lanes, >= npair at nb = 32
static_assert(31 * 32 / 2 + 31 < kRowStride, "packed (i, j) for i, j < 32 lies inside a staged row");
template <int NOC, bool GV>
struct ColShape {
  static constexpr int NT = GV ? 2 : 1;
  static constexpr int NROW = GV ? 8 : 4;   // Fs, V_0, V_1, V_2 [, Vo_0..2 and one unused row]
  static constexpr int XF = 3 * NT * NOC;   // Wt column of Fs C
  static constexpr int XA = XF + NOC;       // A, then dA_0..2
  static constexpr int LDW = (XA + 4) | 1;  // odd row stride
};
__host__ __device__ inline int64_t col_rows_doubles(int nrow, int esz, int ns) {
  const int64_t rows = ((int64_t)nrow * kRowStride * esz + 7) / 8;
  return rows > ns * 10 * 32 ? rows : ns * 10 * 32;
}
__host__ __device__ inline int64_t col_wave_doubles(int nrow, int esz, int nb, int ldw, int nent, int ns, int nt,
                                                    int noc) {
  return col_rows_doubles(nrow, esz, ns) + (int64_t)nb * ldw + nent + ns * 7 * noc + ns * nt * 6 + ns * 4;
}
__host__ __device__ inline int col_tasks(int ns, int nt, int nc0, int nc1, int noc) {
  return 3 * (nc0 + (ns == 2 ? nc1 : 0)) + ns * nt * 3 + ns + ns * 4 * noc;
}
__host__ __device__ inline int64_t col_block_doubles(int noc, int nent, int ntask) {
  return 32 * noc + (nent + ntask + 1) / 2;
}

// Entries of the dot-product step, in order: for s, t, c, b < nc_s, a <= b:
// U^t_s,c[a][b] (Wt column (3 t + c) NOC + off_s + b); then per s: x_s[a]
// (column XA) and y_s,c[a] (column XA + 1 + c), a < nc_s.
__host__ __device__ inline int col_entries(int ns, int nt, int nc0, int nc1) {
  const int t0 = nc0 * (nc0 + 1) / 2, t1 = ns == 2 ? nc1 * (nc1 + 1) / 2 : 0;
  return nt * 3 * (t0 + t1) + 4 * (nc0 + (ns == 2 ? nc1 : 0));
}

template <class ST, int NS, bool GV, int NOC>
__global__ __launch_bounds__(256, 2) void tc_point_col_kernel(ColArgs a, const double* __restrict__ Cc,
                                                              const double* __restrict__ Dw) {
  using Sh = ColShape<NOC, GV>;
  constexpr int NT = Sh::NT, NROW = Sh::NROW, XF = Sh::XF, XA = Sh::XA, LDW = Sh::LDW;
  constexpr int K = GV ? 4 : 2;              // matrix slots per lane
  constexpr int NSTG = GV ? 7 : 4;           // rows staged per point
  constexpr int NL = kRowStride / 64;        // loads per lane per row
  constexpr int RS = kRowStride;
  extern __shared__ double smem[];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int h = lane >> 5, i = lane & 31;
  const int nb = a.nb, nent = a.nent;
  const int64_t np = a.npair;
  const int nc[2] = {a.nocc0, NS == 2 ? a.nocc1 : a.nocc0};
  const int off[2] = {0, NS == 2 ? a.nocc0 : 0};
  const double cs = a.c_same, co = a.c_oppo;
  const bool act = i < nb;
  const int ic = act ? i : nb - 1;
  const int Ti = ic * (ic + 1) / 2;
  const int tri0 = nc[0] * (nc[0] + 1) / 2, tri1 = NS == 2 ? nc[1] * (nc[1] + 1) / 2 : 0;
  const int nu0 = NT * 3 * tri0, nu = nu0 + NT * 3 * tri1;
  const int na = 3 * (nc[0] + (NS == 2 ? nc[1] : 0)), nbt = NS * NT * 3, ntask = na + nbt + NS + NS * 4 * NOC;
  auto ubase = [&](int s, int t, int c) { return (s ? nu0 : 0) + (3 * t + c) * (s ? tri1 : tri0); };
  auto xbase = [&](int s) { return nu + (s ? 4 * nc[0] : 0); };

  // LDS carving
  double* Cb = smem;                                 // [32][NOC]
  int* ent = reinterpret_cast<int*>(Cb + 32 * NOC);  // [nent]: acol | wcol << 8
  int* tsk = ent + nent;                             // [ntask]
  double* wbase = smem + col_block_doubles(NOC, nent, ntask) +
                  (int64_t)wave * col_wave_doubles(NROW, sizeof(ST), nb, LDW, nent, NS, NT, NOC);
  ST* rows = reinterpret_cast<ST*>(wbase);  // [NROW][RS]
  double* Rq = wbase;                       // [NS][10][32], once the rows are consumed
  double* Wt = wbase + col_rows_doubles(NROW, sizeof(ST), NS);  // [nb][LDW]
  double* Res = Wt + (int64_t)nb * LDW;     // [nent]
  double* Xg = Res + nent;                  // [NS][4][NOC]: x, y_0..2 by column
  double* Gs = Xg + NS * 4 * NOC;           // [NS][3][NOC]
  double* Tn = Gs + NS * 3 * NOC;           // [NS][NT][3][2]
  double* Sc = Tn + NS * NT * 6;            // [NS][4]: occ, x . g_c

  // workgroup set-up: C rows, the entry table and the task table
  for (int e = tid; e < 32 * NOC; e += blockDim.x) Cb[e] = e / NOC < nb ? Cc[e] : 0.0;
  for (int e = tid; e < nent; e += blockDim.x) {
    int acol, wcol;
    if (e < nu) {
      const int s = e < nu0 ? 0 : 1, tri = s ? tri1 : tri0;
      const int r0 = s ? e - nu0 : e, grp = r0 / tri, r = r0 - grp * tri;  // grp = 3 t + c
      int b = (int)((sqrt(8.0 * r + 1.0) - 1.0) * 0.5);
      while (b * (b + 1) / 2 > r) --b;
      while ((b + 1) * (b + 2) / 2 <= r) ++b;
      acol = off[s] + r - b * (b + 1) / 2;
      wcol = grp * NOC + off[s] + b;
    } else {
      const int r0 = e - nu, s = (NS == 2 && r0 >= 4 * nc[0]) ? 1 : 0, r = s ? r0 - 4 * nc[0] : r0;
      if (r < nc[s]) {
        acol = off[s] + r;
        wcol = XA;
      } else {
        const int q = r - nc[s];
        acol = off[s] + q / 3;
        wcol = XA + 1 + q % 3;
      }
    }
    ent[e] = acol | (wcol << 8);
  }
  // tasks: kind 0 g_s,c[b] | 1 trace and squares of U^t_s,c | 2 occ_s | 3 the
  // x / y copy of column col (q = 0 x, 1 + c y_c); packed as kind | s << 2 |
  // c << 3 | t << 5 | b (or col) << 6 | q << 11
  for (int e = tid; e < ntask; e += blockDim.x) {
    int v;
    if (e < na) {
      const int s = (NS == 2 && e >= 3 * nc[0]) ? 1 : 0, r = s ? e - 3 * nc[0] : e;
      const int c = r / nc[s], b = r - c * nc[s];
      v = 0 | (s << 2) | (c << 3) | (b << 6);
    } else if (e < na + nbt) {
      const int r = e - na, s = r / (NT * 3), t = (r / 3) % NT, c = r % 3;
      v = 1 | (s << 2) | (c << 3) | (t << 5);
    } else if (e < na + nbt + NS) {
      v = 2 | ((e - na - nbt) << 2);
    } else {
      const int r = e - na - nbt - NS, s = r / (4 * NOC), q = (r / NOC) % 4, col = r % NOC;
      v = 3 | (s << 2) | (col << 6) | (q << 11);
    }
    tsk[e] = v;
  }

  // the lane's matrix slots: without V_oppo half 0 Fs, V_0 and half 1 V_1,
  // V_2 (staged rows 0..3); with it half 0 Fs, V_0, Vo_0 and an unused row,
  // half 1 V_1, V_2, Vo_1, Vo_2 (rows 4..7)
  const int rbase = h * K * RS;
  double vl[NS][NL];
#pragma unroll
  for (int s = 0; s < NS; ++s)
#pragma unroll
    for (int u = 0; u < NL; ++u) vl[s][u] = 0;

  // the point's rows through registers, one point ahead: the staged rows and
  // F_oppo (used only in Faa, at staging)
  const ST* Fs = (const ST*)a.Fs;
  const ST* Fo = (const ST*)a.Fo;
  const ST* Vv = (const ST*)a.V;
  const ST* Vw = (const ST*)(GV ? a.Vo : a.V);
  ST pre[NSTG][NL], pfo[NL];
  double pa = 0, pw = 0, pd[3] = {0, 0, 0};
  auto load = [&](int64_t nn) {
    const int64_t n = uniform_i64(nn);
#pragma unroll
    for (int r = 0; r < NSTG; ++r) {
      // pre row r: 0 Fs, 1..3 V_0..2, 4..6 Vo_0..2
      const ST* src = r == 0 ? Fs + n * np : (r < 4 ? Vv + (n * 3 + r - 1) * np : Vw + (n * 3 + r - 4) * np);
      const __amdgpu_buffer_rsrc_t rv = buffer_view(src, np);
#pragma unroll
      for (int t = 0; t < NL; ++t) pre[r][t] = buffer_load<ST>(rv, (lane + 64 * t) * (int)sizeof(ST));
    }
    const __amdgpu_buffer_rsrc_t ro = buffer_view(Fo + n * np, np);
#pragma unroll
    for (int t = 0; t < NL; ++t) pfo[t] = buffer_load<ST>(ro, (lane + 64 * t) * (int)sizeof(ST));
    pa = a.A[n * nb + ic];
    pd[0] = a.dA[(n * nb + ic) * 3];
    pd[1] = a.dA[(n * nb + ic) * 3 + 1];
    pd[2] = a.dA[(n * nb + ic) * 3 + 2];
    pw = a.w[n];
  };
  // staged row of pre row r and of V^t_c
  //   without Vo: Fs 0, V_0..2 1..3;  with Vo: Fs 0, V_0 1, V_1 4, V_2 5, Vo_0 2, Vo_1 6, Vo_2 7
  auto srow = [](int r) {
    constexpr int g[7] = {0, 1, 4, 5, 2, 6, 7};
    return GV ? g[r] : r;
  };
  auto vrow = [&](int t, int c) { return srow(t ? 4 + c : 1 + c); };
  // packed D with doubled off-diagonal (Faa = sum_p Dw_p F_p), [2][kRowStride]
  // zero past npair
  const __amdgpu_buffer_rsrc_t dwv = buffer_view(Dw, 2 * kRowStride);
  const int64_t gw = (int64_t)blockIdx.x * 4 + wave, nw = (int64_t)gridDim.x * 4;  // 4 waves per workgroup
  if (gw < a.n1) load(gw);
  __syncthreads();

  for (int64_t n = gw; n < a.n1; n += nw) {
    // staged rows (each load past npair reads zero into the row's padding);
    // Faa_s = tr(Fs D_s) + tr(Fo D_sb) from the rows in registers
    double fa[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) fa[s] = 0;
#pragma unroll
    for (int t = 0; t < NL; ++t) {
      const double f = (double)pre[0][t], g = (double)pfo[t];
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        const int sb = NS == 1 ? s : 1 - s;
        const double d = buffer_load<double>(dwv, (s * RS + lane + 64 * t) * 8);
        const double db = buffer_load<double>(dwv, (sb * RS + lane + 64 * t) * 8);
        fa[s] += f * d + g * db;
      }
    }
#pragma unroll
    for (int r = 0; r < NSTG; ++r)
#pragma unroll
      for (int t = 0; t < NL; ++t) rows[srow(r) * RS + lane + 64 * t] = pre[r][t];
    const double Ai = act ? pa : 0.0, wn = pw;
    const double dAi[3] = {act ? pd[0] : 0.0, act ? pd[1] : 0.0, act ? pd[2] : 0.0};
    load(n + nw < a.n1 ? n + nw : n);
    double faa[2];
#pragma unroll
    for (int s = 0; s < NS; ++s) faa[s] = uniform(wave_sum(fa[s]));
    wave_sync_lds();

    {
      // P[k][col] = sum_j M_k[j][i] C[j][col]
      double acc[K][NOC];
#pragma unroll
      for (int k = 0; k < K; ++k)
#pragma unroll
        for (int col = 0; col < NOC; ++col) acc[k][col] = 0;
      const ST* lr = rows + rbase;
      // JR j per round: the lane's row elements and the C rows (LDS, the C
      // rows as broadcast reads: only LDS reads in flight, so each FMA waits
      // for its own operands) of a round issued before its FMAs.  A round
      // never passes j = 31 (nb <= 32); j past nb takes the zero rows of Cb
      // (its row elements, packed index <= 527, lie inside the staged rows)
      constexpr int JR = NOC <= 5 ? 4 : (NOC <= 8 ? 2 : 1);
#pragma unroll 1
      for (int j0 = 0; j0 < nb; j0 += JR) {
        const int jb = __builtin_amdgcn_readfirstlane(j0);
        double m[JR][K], c[JR][NOC];
#pragma unroll
        for (int u = 0; u < JR; ++u) {
          const int j = jb + u;
          // packed (i, j): Ti + j for j <= i, else T(j) + i (a select, no branch)
          const int below = -(int)(j <= ic);
          const int pj = ((Ti + j) & below) | ((((j * (j + 1)) >> 1) + ic) & ~below);
#pragma unroll
          for (int k = 0; k < K; ++k) m[u][k] = (double)lr[k * RS + pj];
          const double* cj = Cb + j * NOC;
#pragma unroll
          for (int col = 0; col < NOC; ++col) c[u][col] = cj[col];
        }
#pragma unroll
        for (int u = 0; u < JR; ++u)
#pragma unroll
          for (int col = 0; col < NOC; ++col)
#pragma unroll
            for (int k = 0; k < K; ++k) acc[k][col] += m[u][k] * c[u][col];
      }
      // row i of Wt: the W^t_c columns from the slots that hold them, half 0
      // also Fs C, A and dA
      if (act) {
        double* wr = Wt + (int64_t)i * LDW;
        if (!h) {
#pragma unroll
          for (int col = 0; col < NOC; ++col) {
            wr[XF + col] = acc[0][col];                  // Fs C
            wr[col] = acc[1][col];                       // W^0_0
            if (GV) wr[3 * NOC + col] = acc[2][col];     // W^1_0
          }
          wr[XA] = Ai;
          wr[XA + 1] = dAi[0];
          wr[XA + 2] = dAi[1];
          wr[XA + 3] = dAi[2];
        } else {
#pragma unroll
          for (int col = 0; col < NOC; ++col) {
            wr[NOC + col] = acc[0][col];                 // W^0_1
            wr[2 * NOC + col] = acc[1][col];             // W^0_2
            if (GV) {
              wr[4 * NOC + col] = acc[2][col];           // W^1_1
              wr[5 * NOC + col] = acc[K - 1][col];       // W^1_2
            }
          }
        }
      }
    }
    wave_sync_lds();
    // dot products over l: U, x, y
    for (int e = lane; e < nent; e += 64) {
      const int en = ent[e], acol = en & 0xff, wcol = en >> 8;
      const double* cb = Cb + acol;
      const double* wt = Wt + wcol;
      double v0 = 0, v1 = 0, v2 = 0, v3 = 0;  // four chains of loads in flight
      int l = 0;
      for (; l + 4 <= nb; l += 4) {
        v0 += cb[l * NOC] * wt[l * LDW];
        v1 += cb[(l + 1) * NOC] * wt[(l + 1) * LDW];
        v2 += cb[(l + 2) * NOC] * wt[(l + 2) * LDW];
        v3 += cb[(l + 3) * NOC] * wt[(l + 3) * LDW];
      }
      for (; l < nb; ++l) v0 += cb[l * NOC] * wt[l * LDW];
      Res[e] = (v0 + v1) + (v2 + v3);
    }
    wave_sync_lds();
    // per spin: g_s,c = c_s U^0_s,c x_s (by column); trace and sum of squares
    // of U^t_s,c (2 sum over the triangle - sum of the diagonal); occ_s =
    // |x_s|^2; x and y by column
    for (int e = lane; e < ntask; e += 64) {
      const int v = tsk[e], kind = v & 3, s = (v >> 2) & 1, c = (v >> 3) & 3, t = (v >> 5) & 1, b = (v >> 6) & 31;
      if (kind == 0) {
        const int ub = ubase(s, 0, c), xb = xbase(s);
        double g0 = 0, g1 = 0;
        for (int q = 0; q < nc[s]; ++q) {
          const int lo = q < b ? q : b, hi = q < b ? b : q;
          const double u = Res[ub + hi * (hi + 1) / 2 + lo] * Res[xb + q];
          if (q & 1) g1 += u; else g0 += u;
        }
        Gs[(s * 3 + c) * NOC + off[s] + b] = cs * (g0 + g1);
      } else if (kind == 1) {
        const int tri = s ? tri1 : tri0, ub = ubase(s, t, c);
        double sq0 = 0, sq1 = 0, tr = 0, dq = 0;
        int q = 0;
        for (; q + 2 <= tri; q += 2) {
          const double u0 = Res[ub + q], u1 = Res[ub + q + 1];
          sq0 += u0 * u0;
          sq1 += u1 * u1;
        }
        if (q < tri) sq0 += Res[ub + q] * Res[ub + q];
        for (int p = 0; p < nc[s]; ++p) {
          const double d = Res[ub + p * (p + 1) / 2 + p];
          tr += d;
          dq += d * d;
        }
        Tn[((s * NT + t) * 3 + c) * 2] = tr;
        Tn[((s * NT + t) * 3 + c) * 2 + 1] = 2 * (sq0 + sq1) - dq;
      } else if (kind == 2) {
        const int xb = xbase(s);
        double v0 = 0, v1 = 0;
        for (int q = 0; q < nc[s]; ++q) {
          const double x = Res[xb + q];
          if (q & 1) v1 += x * x; else v0 += x * x;
        }
        Sc[s * 4] = v0 + v1;  // occ_s = |x_s|^2
      } else {
        // x_s (q = 0) and y_s,c (q = 1 + c) at their columns, zero elsewhere
        const int col = b, q = v >> 11, bb = col - off[s];
        const bool in = bb >= 0 && bb < nc[s];
        const double xv = in ? Res[xbase(s) + (q == 0 ? bb : nc[s] + 3 * bb + q - 1)] : 0.0;
        Xg[(s * 4 + q) * NOC + col] = xv;
        if (q < 3 && !in) Gs[(s * 3 + q) * NOC + col] = 0.0;  // g outside the spin's columns
      }
    }
    wave_sync_lds();
    // x_s . g_s,c (SumOverb_conjVTildeb_occb) by column, one lane per (s, c)
    if (lane < 3 * NS) {
      const int s = lane / 3, c = lane % 3;
      double v = 0;
#pragma unroll
      for (int col = 0; col < NOC; ++col) v += Xg[s * 4 * NOC + col] * Gs[(s * 3 + c) * NOC + col];
      Sc[s * 4 + 1 + c] = v;
    }
    wave_sync_lds();
    // point-uniform values
    double vtv[2][3], occ[2], scvt[2][3], nr[NT][2];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      occ[s] = Sc[s * 4];
#pragma unroll
      for (int c = 0; c < 3; ++c) scvt[s][c] = Sc[s * 4 + 1 + c];
#pragma unroll
      for (int t = 0; t < NT; ++t)
        nr[t][s] = Tn[((s * NT + t) * 3) * 2 + 1] + Tn[((s * NT + t) * 3 + 1) * 2 + 1] + Tn[((s * NT + t) * 3 + 2) * 2 + 1];
    }
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const int sb = NS == 1 ? s : 1 - s;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const double vd = Tn[((s * NT + 0) * 3 + c) * 2], vdb = Tn[((sb * NT + 0) * 3 + c) * 2];
        vtv[s][c] = GV ? vd + Tn[((sb * NT + 1) * 3 + c) * 2] : cs * vd + co * vdb;  // VaaSamePlusOppo
      }
    }
    // G3's V-linear terms h_same . V_same(j, i) + h_oppo . V_oppo(j, i) over
    // the staged V rows (lanes past npair add the rows' zero padding)
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const int sb = NS == 1 ? s : 1 - s;
      double hv[NT][3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        if (GV) {
          hv[0][c] = wn * (2 * occ[s] * vtv[s][c] - 2 * scvt[s][c]);
          hv[NT - 1][c] = wn * (2 * occ[sb] * vtv[sb][c] - 2 * scvt[sb][c]);
        } else {
          hv[0][c] = wn * (2 * occ[s] * cs * vtv[s][c] - 2 * cs * scvt[s][c] + 2 * occ[sb] * co * vtv[sb][c] -
                           2 * co * scvt[sb][c]);
        }
      }
#pragma unroll
      for (int u = 0; u < NL; ++u) {
        const int p = lane + 64 * u;
        double v = 0;
#pragma unroll
        for (int t = 0; t < NT; ++t)
#pragma unroll
          for (int c = 0; c < 3; ++c) v += hv[t][c] * (double)rows[vrow(t, c) * RS + p];
        vl[s][u] += v;
      }
    }
    wave_sync_lds();  // the rows are consumed: Rq may overwrite them
    // per-lane dot products over the columns, five per lane, the same code in
    // both halves: half 0 Fs C . x, W_0 . x, W_1 . x, W_0 . y_0, W_0 . g_0;
    // half 1 W_2 . x, W_1 . y_1, W_2 . y_2, W_1 . g_1, W_2 . g_2
    if (act) {
      const double* wr = Wt + (int64_t)i * LDW;
      const int wsrc[5] = {h ? 2 * NOC : XF, h ? NOC : 0, h ? 2 * NOC : NOC, h ? NOC : 0, h ? 2 * NOC : 0};
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        const double* xg = Xg + s * 4 * NOC;
        const double* gs = Gs + s * 3 * NOC;
        const double* csrc[5] = {xg, h ? xg + 2 * NOC : xg, h ? xg + 3 * NOC : xg, h ? gs + NOC : xg + NOC,
                                 h ? gs + 2 * NOC : gs};
        double r[5] = {0, 0, 0, 0, 0};
#pragma unroll
        for (int col = 0; col < NOC; ++col)
#pragma unroll
          for (int k = 0; k < 5; ++k) r[k] += wr[wsrc[k] + col] * csrc[k][col];
        double* rq = Rq + s * 10 * 32;
#pragma unroll
        for (int k = 0; k < 5; ++k) rq[(5 * h + k) * 32 + i] = r[k];
      }
    }
    wave_sync_lds();
    // the point's vectors: half 0 u0..u4, half 1 u5, u6; the W vectors.  Rq
    // rows: 0 FaiS, 1 Q_0, 2 Q_1, 3 Vai y_0 part, 4 V.E g_0 part, 5 Q_2,
    // 6 Vai y_1 part, 7 Vai y_2 part, 8 V.E g_1 part, 9 V.E g_2 part
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const int sb = NS == 1 ? s : 1 - s;
      double* vs = a.vecs + ((int64_t)n * 2 + s) * a.nvec * nb;
      const double* rq = Rq + s * 10 * 32;
      if (act) {
        if (!h) {
          const double fai = rq[i], vt0 = cs * rq[32 + i], vt1 = cs * rq[64 + i], vt2 = cs * rq[160 + i];
          const double vai = rq[96 + i] + rq[192 + i] + rq[224 + i];
          const double sv1 = rq[128 + i] + rq[256 + i] + rq[288 + i];
          const double vtdot = vtv[s][0] * vt0 + vtv[s][1] * vt1 + vtv[s][2] * vt2;
          const double vtn = vtv[s][0] * vtv[s][0] + vtv[s][1] * vtv[s][1] + vtv[s][2] * vtv[s][2];
          const double Z = 2 * wn * (cs * sv1 - vtdot);
          const double sNS = GV ? nr[0][s] : cs * cs * nr[0][s];
          const double sNOb = GV ? nr[NT - 1][sb] : co * co * nr[0][sb];
          vs[0 * nb + i] = wn * (faa[s] * Ai - fai);
          vs[1 * nb + i] = wn * (vtv[s][0] * dAi[0] + vtv[s][1] * dAi[1] + vtv[s][2] * dAi[2] - cs * vai);
          vs[2 * nb + i] = wn * (vtn - sNS - sNOb) * Ai + Z;
          vs[3 * nb + i] = Z;
          vs[4 * nb + i] = vt0;
        } else {
          vs[5 * nb + i] = cs * rq[64 + i];
          vs[6 * nb + i] = cs * rq[160 + i];
        }
        // W^t_s,c[i][b] for b < nmax (zero past nc_s), rows 3 (t nmax + b) + c;
        // the halves take alternate b
        const double* wr = Wt + (int64_t)i * LDW;
#pragma unroll
        for (int t = 0; t < NT; ++t)
          for (int b = h; b < a.nmax; b += 2)
#pragma unroll
            for (int c = 0; c < 3; ++c)
              vs[(kVecFixed + 3 * (t * a.nmax + b) + c) * nb + i] =
                  b < nc[s] ? wr[(3 * t + c) * NOC + off[s] + b] : 0.0;
      }
      double* al = a.alpha + (n * 2 + s) * 2;
      if (lane == 0) {
        if (GV) {
          al[0] = -2 * wn * occ[s];
          al[1] = -2 * wn * occ[sb];
        } else {
          al[0] = -2 * wn * (occ[s] * cs * cs + occ[sb] * co * co);
        }
      }
    }
    wave_sync_lds();
  }
  // V-linear partials: the waves' sums added in wave order
  __syncthreads();
  double* red = smem;  // [4 waves][NS][npair] over the workgroup's LDS
#pragma unroll
  for (int s = 0; s < NS; ++s)
#pragma unroll
    for (int u = 0; u < NL; ++u) {
      const int p = lane + 64 * u;
      if (p < np) red[((int64_t)wave * NS + s) * np + p] = vl[s][u];
    }
  __syncthreads();
  for (int64_t e = tid; e < NS * np; e += blockDim.x) {
    double v = 0;
    for (int wv = 0; wv < 4; ++wv) v += red[(int64_t)wv * NS * np + e];
    a.vpart[(int64_t)blockIdx.x * NS * np + e] = v;
  }
}

struct RankArgs {
  const double* vecs;   // [n1][2][nvec][nb]
  const double* alpha;  // [n1][2 spins][2 tensors]
  const double* A;      // [n1][nb]
  const double* dA;     // [n1][nb][3]
  const double* w;      // [n1]
  int64_t n1;
  int nb, nvec, nocc0, nocc1, nmax, nwt;  // nwt: V tensors with W blocks (1 or 2)
  int ustart[7];        // workgroups of target 3 s + t: [ustart[st], ustart[st + 1])
  double* part;         // [gridDim.x][nb][nb]
  int pts[6];           // points per workgroup of target st (a multiple of 4)
};

// (point, entry) terms per point of target st = 3 s + t: 1, 4, 5 + 3 nwt nocc_s.
__host__ __device__ inline int rank_entries(int st, int nocc0, int nocc1, int nwt) {
  const int s = st / 3, tg = st - 3 * s;
  return tg == 0 ? 1 : (tg == 1 ? 4 : 5 + 3 * nwt * (s ? nocc1 : nocc0));
}

// Pass 2: G_{s,t}[j][i] = sum over (point, entry) of L[j] R[i] (the rank-one
// terms listed above) as an FP64 MFMA product, blockIdx.y = 3 s + t.  A
// workgroup takes pts[st] consecutive points of its target; one k-step is
// one entry of four consecutive points (lane group g: point 4 q + g), so the
// entry -- which vectors, offsets and scale -- is uniform across the wave and
// decoded in scalar registers.  The 4 waves take interleaved k-steps (entry
// fastest) over all NT x NT output tiles (16x16 each) and are summed in a
// fixed order through LDS.  v_mfma_f64_16x16x4: A[j][k] from lane 16 g + r =
// L_{k0+g}[16 m + r], B[k][i] = R_{k0+g}[16 t + r].
template <int NT>
__global__ __launch_bounds__(256) void tc_rank_kernel(RankArgs a) {
  __shared__ double red[NT * 16][NT * 16];
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int g = lane >> 4, r = lane & 15;
  const int u = blockIdx.x;
  int st = 0;
  while (st < 5 && u >= a.ustart[st + 1]) ++st;
  const int s = st / 3, tg = st - 3 * s;
  const int nb = a.nb;
  const int E = rank_entries(st, a.nocc0, a.nocc1, a.nwt);
  const int w3 = 3 * (s ? a.nocc1 : a.nocc0);  // W entries per tensor
  // this workgroup's points pb .. pb + np - 1, as nq quads
  const int P = a.pts[st];
  const int64_t pb = (int64_t)(u - a.ustart[st]) * P;
  const int np = (int)(a.n1 - pb < P ? a.n1 - pb : P);
  const int nq = (np + 3) / 4;
  const double* vs = a.vecs + s * a.nvec * nb;  // spin s of point 0
  // Operands of a k-step (lane (g, r): L[16 m + r] and R[16 m + r] of the
  // step's entry at point 4 q + g) in two halves: fetch issues the raw loads
  // into a ring slot (no arithmetic on loaded values, so nothing waits for
  // them there); use scales / masks them just before the MFMAs, three steps
  // later.  The fetch cursor (fq, fe) walks this wave's k-steps wave + 4 j.
  struct Slot {
    double l[NT], r[NT];
    double w, al;
    int sk;  // scale: 0 none, 1 -w, 2 2w, 3 alpha, 4 zero (past the points)
  };
  int fq = wave / E, fe = wave - (wave / E) * E;
  auto fetch = [&](Slot& o) {
    const int q = fq, e = fe;
    fe += 4;
    while (fe >= E) {
      fe -= E;
      ++fq;
    }
    const bool in = 4 * q + g < np;
    const int64_t n = pb + (in ? 4 * q + g : 0);
    // entry e of target tg (uniform): L / R source (0 the point's vectors, 1 A,
    // 2 dA), offset and stride, and the scale (W entries of tensor wt use alpha_wt)
    int ls = 1, lo = 0, rs = 0, ro = 0, rst = 1, sk = 0, wt = 0;
    if (tg == 1) {
      const bool z = e == 0;
      ls = z ? 1 : 0;
      lo = z ? 0 : (3 + e) * nb;  // (A, u1) | (-w t_c, dA_c), c = e - 1
      rs = z ? 0 : 2;
      ro = z ? nb : e - 1;
      rst = z ? 1 : 3;
      sk = z ? 0 : 1;
    } else if (tg == 2) {
      const bool z = e == 0, one = e == 1, tc3 = e >= 2 && e < 5;
      const int ww = e - 5;
      wt = ww >= w3 ? 1 : 0;
      ls = z ? 1 : 0;  // (A, u2) | (u3, A) | (2 w t_c, t_c) | (alpha_wt W^wt_bc, W^wt_bc)
      lo = z ? 0 : (one ? 3 * nb : (tc3 ? (2 + e) * nb : (kVecFixed + (wt ? 3 * a.nmax + ww - w3 : ww)) * nb));
      rs = one ? 1 : 0;
      ro = z ? 2 * nb : (one ? 0 : lo);
      sk = z || one ? 0 : (tc3 ? 2 : 3);
    }  // tg == 0: (A, u0)
    const double* L = (ls == 0 ? vs + n * 2 * a.nvec * nb : a.A + n * nb) + lo;
    const double* R = (rs == 0 ? vs + n * 2 * a.nvec * nb : (rs == 1 ? a.A + n * nb : a.dA + n * nb * 3)) + ro;
    o.w = a.w[n];
    o.al = a.alpha[(n * 2 + s) * 2 + wt];
    o.sk = in && q < nq ? sk : 4;
#pragma unroll
    for (int m = 0; m < NT; ++m) {
      const int row = 16 * m + r < nb ? 16 * m + r : nb - 1;  // in range; masked in use
      o.l[m] = L[row];
      o.r[m] = R[row * rst];
    }
  };
  d4 acc[NT][NT];
#pragma unroll
  for (int m = 0; m < NT; ++m)
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[m][t] = d4{0, 0, 0, 0};
  auto use = [&](const Slot& o) {
    const double fl = o.sk == 0 ? 1.0 : (o.sk == 1 ? -o.w : (o.sk == 2 ? 2 * o.w : (o.sk == 3 ? o.al : 0.0)));
    double av[NT], bv[NT];
#pragma unroll
    for (int m = 0; m < NT; ++m) {
      const bool in = 16 * m + r < nb;
      av[m] = in ? fl * o.l[m] : 0.0;
      bv[m] = in && o.sk != 4 ? o.r[m] : 0.0;
    }
#pragma unroll
    for (int m = 0; m < NT; ++m)
#pragma unroll
      for (int t = 0; t < NT; ++t) acc[m][t] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[m], bv[t], acc[m][t], 0, 0, 0);
  };
  // a ring of three slots: loads run three k-steps ahead of their MFMAs (steps
  // past the points load in-range data that use() zeroes)
  Slot s0, s1, s2;
  const int KS = nq * E;  // k-steps of the workgroup
  fetch(s0);
  fetch(s1);
  for (int k0 = wave; k0 < KS; k0 += 12) {
    fetch(s2);
    use(s0);
    fetch(s0);
    use(s1);
    fetch(s1);
    use(s2);
  }
  // fixed-order sum of the 4 waves' tiles
  for (int w = 0; w < 4; ++w) {
    if (wave == w)
#pragma unroll
      for (int m = 0; m < NT; ++m)
#pragma unroll
        for (int t = 0; t < NT; ++t)
#pragma unroll
          for (int ii = 0; ii < 4; ++ii) {
            const int row = 16 * m + g + 4 * ii, col = 16 * t + r;
            red[row][col] = (w == 0 ? 0.0 : red[row][col]) + acc[m][t][ii];
          }
    __syncthreads();
  }
  double* out = a.part + (int64_t)blockIdx.x * nb * nb;
  for (int e2 = threadIdx.x; e2 < nb * nb; e2 += 256) out[e2] = red[e2 / nb][e2 % nb];
}

// Pass 3: Rv[s][p] = sum_n sum_t sum_c h[n][s][t][c] V_t[n][c][p] for the NK
// spins formed and the NT tensors (V, or V_same and V_oppo).
template <class ST, int NK, int NT>
__global__ __launch_bounds__(256) void tc_vlin_kernel(const ST* __restrict__ V, const ST* __restrict__ Vo,
                                                      const double* __restrict__ hvec, int64_t n1, int64_t npair,
                                                      double* __restrict__ part) {
  const int64_t per = (n1 + gridDim.x - 1) / gridDim.x;
  const int64_t p0 = (int64_t)blockIdx.x * per;
  const int64_t p1 = p0 + per < n1 ? p0 + per : n1;
  for (int64_t p = threadIdx.x; p < npair; p += 256) {
    double acc[NK] = {};
#pragma unroll 4
    for (int64_t n = p0; n < p1; ++n) {  // unrolled: several points' loads in flight
      const double* h = hvec + n * 12;
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const ST* Vt = t ? Vo : V;
        const double v0 = Vt[(n * 3) * npair + p], v1 = Vt[(n * 3 + 1) * npair + p], v2 = Vt[(n * 3 + 2) * npair + p];
#pragma unroll
        for (int k = 0; k < NK; ++k) {
          const double* hk = h + (k * 2 + t) * 3;
          acc[k] += v0 * hk[0] + v1 * hk[1] + v2 * hk[2];
        }
      }
    }
#pragma unroll
    for (int k = 0; k < NK; ++k) part[((int64_t)blockIdx.x * NK + k) * npair + p] = acc[k];
  }
}

// out[st][e] = sum over the rank workgroups u of target st of part[u][e], e < len.
struct RankRanges {
  int ustart[7];
};
// Sum over the partials b in [b_begin, b_end) of part[b][e] for 16 consecutive
// elements e of a workgroup of 1024 threads: the 64 groups of 16 lanes (lane
// = element) each sum a contiguous slice of the partials (eight loads of one
// 128-byte line in flight), then the 64 slice sums are added in group order.
// Returns the sum in the threads of group 0.
__device__ inline double sum_partials_16(const double* __restrict__ part, int b_begin, int b_end, int64_t len,
                                         int64_t e) {
  __shared__ double red[64][16];
  const int el = threadIdx.x & 15, grp = threadIdx.x >> 4;
  const int per = (b_end - b_begin + 63) / 64;
  const int b0 = b_begin + grp * per, b1 = b0 + per < b_end ? b0 + per : b_end;
  double s = 0;
  if (e < len) {
    int b = b0;
    for (; b + 8 <= b1; b += 8) {
      double x[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) x[u] = part[(int64_t)(b + u) * len + e];
#pragma unroll
      for (int u = 0; u < 8; ++u) s += x[u];
    }
    for (; b < b1; ++b) s += part[(int64_t)b * len + e];
  }
  red[grp][el] = s;
  __syncthreads();
  double t = 0;
  if (grp == 0)
    for (int u = 0; u < 64; ++u) t += red[u][el];
  return t;
}

// out[st * len + e] = sum over the rank workgroups u of target st of part[u][e]:
// one workgroup of 1024 per (target, 16 consecutive elements).
__global__ __launch_bounds__(1024) void tc_reduce_ranges(const double* __restrict__ part, RankRanges r, int64_t len,
                                                         double* __restrict__ out) {
  const int64_t nblk = (len + 15) / 16;
  const int st = (int)(blockIdx.x / nblk);
  const int64_t e = (blockIdx.x - (int64_t)st * nblk) * 16 + (threadIdx.x & 15);
  const double t = sum_partials_16(part, r.ustart[st], r.ustart[st + 1], len, e);
  if ((threadIdx.x >> 4) == 0 && e < len) out[(int64_t)st * len + e] = t;
}

// out[e] = sum_b part[b][e], b < nparts: one workgroup of 1024 per 16 elements.
__global__ __launch_bounds__(1024) void tc_reduce(const double* __restrict__ part, int nparts, int64_t len,
                                                  double* __restrict__ out) {
  const int64_t e = (int64_t)blockIdx.x * 16 + (threadIdx.x & 15);
  const double t = sum_partials_16(part, 0, nparts, len, e);
  if ((threadIdx.x >> 4) == 0 && e < len) out[e] = t;
}

double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

inline int64_t pk(int j, int i) { return j <= i ? (int64_t)i * (i + 1) / 2 + j : (int64_t)j * (j + 1) / 2 + i; }

// The point-kernel instance for a storage type, staging choice, spin count and
// V form (independent V_same / V_oppo or one proportional tensor).
template <int NS, bool GV>
inline const void* point_kernel_ns(int f32, int stage) {
  if (f32)
    return stage ? (const void*)tc_point_kernel<float, true, NS, GV> : (const void*)tc_point_kernel<float, false, NS, GV>;
  return stage ? (const void*)tc_point_kernel<double, true, NS, GV> : (const void*)tc_point_kernel<double, false, NS, GV>;
}
inline const void* point_kernel(int f32, int stage, bool restricted, bool gv) {
  if (gv) return restricted ? point_kernel_ns<1, true>(f32, stage) : point_kernel_ns<2, true>(f32, stage);
  return restricted ? point_kernel_ns<1, false>(f32, stage) : point_kernel_ns<2, false>(f32, stage);
}

// The column-kernel instance (FP32 rows, one V tensor) for a spin count and
// column count noc (4, 5, 8 or 16).
template <int NS>
inline const void* col_kernel_g(int noc) {
  return noc <= 4 ? (const void*)tc_point_col_kernel<float, NS, false, 4>
                  : (noc <= 5 ? (const void*)tc_point_col_kernel<float, NS, false, 5>
                              : (noc <= 8 ? (const void*)tc_point_col_kernel<float, NS, false, 8>
                                          : (const void*)tc_point_col_kernel<float, NS, false, 16>));
}
inline const void* col_kernel(bool restricted, int noc) { return restricted ? col_kernel_g<1>(noc) : col_kernel_g<2>(noc); }
inline int col_ldw(int noc, int nt) { return (3 * nt * noc + noc + 4) | 1; }

}  // namespace tcs
}  // namespace tc

namespace tc {
using namespace tcs;

size_t tc_out_doubles(int nb) { return 6 * (size_t)nb * nb + 2 * (size_t)nb * (nb + 1) / 2; }

// (point, entry) terms per rank workgroup.  The workgroups of one build run in
// rounds of the device's resident slots (CUs x workgroups per CU); a chunk
// that leaves the last round mostly idle costs that round in full (Ne: 512
// terms gave 1.44 rounds of 1024 slots, 76 us; whole rounds measured as fast
// at twice the chunk).  So the chunk is the smallest multiple of 16 whose
// workgroups fill whole rounds: ceil(terms / (rounds x slots)), with as many
// rounds as keep it at or below 1024 terms.  TCSCF_RANK_CHUNK fixes it for
// A/B timing.
static int rank_chunk(int nup, int ndown, int nwt, int nb, int64_t n1) {
  if (const char* e = std::getenv("TCSCF_RANK_CHUNK")) {
    const int v = std::atoi(e);
    if (v >= 16 && v <= 65536) return v / 16 * 16;
  }
  static std::mutex mu;
  static std::map<std::pair<int, int>, int> slots_of;  // (device, NT) -> resident workgroups
  int dev = 0;
  (void)hipGetDevice(&dev);
  const int nt = nb <= 32 ? 2 : (nb <= 48 ? 3 : 4);
  int slots = 0;
  {
    std::lock_guard<std::mutex> lock(mu);
    auto f = slots_of.find({dev, nt});
    if (f != slots_of.end()) {
      slots = f->second;
    } else {
      int ncu = 0, per = 0;
      const void* fn = nt == 2 ? (const void*)tc_rank_kernel<2>
                               : (nt == 3 ? (const void*)tc_rank_kernel<3> : (const void*)tc_rank_kernel<4>);
      if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
          hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn, 256, 0) != hipSuccess || ncu <= 0 || per <= 0) {
        ncu = 256;
        per = 1;
      }
      slots = slots_of[{dev, nt}] = ncu * per;
    }
  }
  int64_t terms = 0;
  for (int st = 0; st < 6; ++st) terms += n1 * rank_entries(st, nup, ndown, nwt);
  for (int64_t rounds = 1;; ++rounds) {
    const int64_t c = (terms + rounds * slots - 1) / (rounds * slots);
    if (c <= 1024 || rounds >= 64) return (int)std::max<int64_t>(64, (c + 15) / 16 * 16);
  }
}

// Rank workgroups per target st = 3 s + t, as prefix offsets ustart[0..6],
// and the points per workgroup of each target (about chunk terms, a
// multiple of 4 points).
static void rank_units(int nup, int ndown, int nwt, int64_t n1, int chunk, int* ustart, int* pts) {
  ustart[0] = 0;
  for (int st = 0; st < 6; ++st) {
    const int E = rank_entries(st, nup, ndown, nwt);
    const int P = std::max(4, (chunk / E + 3) / 4 * 4);
    if (pts) pts[st] = P;
    ustart[st + 1] = ustart[st] + (int)((n1 + P - 1) / P);
  }
}

size_t tc_ws_doubles(int nup, int ndown, int nb, int64_t n1, bool general_v) {
  const int nmax = std::max(1, std::max(nup, ndown)), nwt = general_v ? 2 : 1;
  const size_t nb2 = (size_t)nb * nb, npair = (size_t)nb * (nb + 1) / 2, nvec = kVecFixed + 3 * nmax * nwt;
  int us[7];
  rank_units(nup, ndown, nwt, n1, rank_chunk(nup, ndown, nwt, nb, n1), us, nullptr);
  return 2 * nb2 + 2 * (size_t)nb * nmax + (size_t)n1 * 2 * nvec * nb + (size_t)n1 * 4 + (size_t)n1 * 12 +
         (size_t)us[6] * nb2 + (size_t)kBlocks * 2 * npair + tc_out_doubles(nb) + 32 * 16 + 2 * kDwStride;
}

// waves per workgroup of the unstaged point kernel (halved while the LDS does not fit)
constexpr int kUnstagedWaves = 8;

int tc_prepare(TcPlan& p, int nup, int ndown, int nb, const double* A1, const double* dA1, const double* w1,
               int64_t n1, const void* Fs, const void* Fo, const void* V, const void* Vo, int f_dtype,
               const tcscf_jastrow* jp, double* ws) {
  TC_CHECK(A1 && dA1 && w1 && Fs && Fo && V && ws, "tcscf_tc_fock: null");
  TC_CHECK(Vo || jp, "tcscf_tc_fock: one V tensor needs the Jastrow's c_same / c_oppo");
  TC_CHECK(f_dtype == TCSCF_F64 || f_dtype == TCSCF_F32, "tcscf_tc_fock: f_dtype must be TCSCF_F64 or TCSCF_F32");
  p.f32 = f_dtype == TCSCF_F32;
  TC_CHECK(nb > 0 && nb <= 64, "tcscf_tc_fock: 1 <= nb <= 64");
  TC_CHECK(nup >= 0 && ndown >= 0 && nup <= nb && ndown <= nb, "tcscf_tc_fock: bad electron counts");
  p.nup = nup;
  p.ndown = ndown;
  p.nb = nb;
  p.nmax = std::max(1, std::max(nup, ndown));
  p.gv = Vo != nullptr;
  p.nwt = p.gv ? 2 : 1;
  p.nvec = kVecFixed + 3 * p.nmax * p.nwt;
  p.n1 = n1;
  p.npair = (int64_t)nb * (nb + 1) / 2;
  p.A1 = A1;
  p.dA1 = dA1;
  p.w1 = w1;
  p.Fs = Fs;
  p.Fo = Fo;
  p.V = V;
  p.Vo = Vo;
  p.c_same = p.gv ? 1.0 : jp->c_same;
  p.c_oppo = p.gv ? 1.0 : jp->c_oppo;
  const int64_t nb2 = (int64_t)nb * nb;
  p.dD = ws;
  p.dC = p.dD + 2 * nb2;
  p.dVec = p.dC + 2 * (int64_t)nb * p.nmax;
  p.dAl = p.dVec + n1 * 2 * p.nvec * nb;
  p.dH = p.dAl + n1 * 4;
  p.dPG = p.dH + n1 * 12;
  p.rchunk = rank_chunk(nup, ndown, p.nwt, nb, n1);
  rank_units(nup, ndown, p.nwt, n1, p.rchunk, p.ustart, p.rpts);
  p.dPV = p.dPG + (int64_t)p.ustart[6] * nb2;
  p.dOut = p.dPV + (int64_t)kBlocks * 2 * p.npair;
  p.dCc = p.dOut + tc_out_doubles(nb);
  TC_CHECK(tc_launch_shape(p), "tcscf_tc_fock: basis / occupied orbitals too large for the LDS staging");
  for (int r = 0; r < 2; ++r) {
    if (p.sh1[r] > 64 * 1024) TC_HIP(lds_opt_in(point_kernel(p.f32, p.stage[r], r, p.gv)));
    if (p.col[r] && p.col_sh[r] > 64 * 1024) TC_HIP(lds_opt_in(col_kernel(r, p.col_noc[r])));
  }
  return 0;
}

// Point-kernel launch per spin count (index 0: two spins, 1: restricted), from
// nb, nmax, the storage type and the V form: the V rows are staged in LDS when
// 8 waves of a workgroup hold them (the gathers then come from LDS and 8 waves
// hide their latency), else the gathers go to memory with 8 waves per
// workgroup.  False when even one wave does not fit.
bool tc_launch_shape(TcPlan& p) {
  const int nb = p.nb;
  const int64_t nb2 = (int64_t)nb * nb;
  for (int r = 0; r < 2; ++r) {
    const int ns = r ? 1 : 2;
    const size_t shared = sizeof(double) * ns * (nb2 + (int64_t)nb * p.nmax);  // D and C
    const size_t staged_wave =
        sizeof(double) * (point_wave_elems(nb, ns) +
                          (p.f32 ? staged_doubles<float>(true, nb, p.nwt) : staged_doubles<double>(true, nb, p.nwt)));
    p.stage[r] = shared + 8 * staged_wave <= (size_t)kMaxLdsBytes ? 1 : 0;
    const size_t per_wave = p.stage[r] ? staged_wave : sizeof(double) * point_wave_elems(nb, ns);
    // staged: 8 waves per workgroup, 12 for the one-spin kernel on FP32 rows
    // of one V tensor (three waves per SIMD fit its registers) when LDS holds
    // them; unstaged: 8, so that one copy of D and C in LDS serves more waves
    p.wpw[r] = p.stage[r] ? (ns == 1 && p.f32 && !p.gv && shared + 12 * per_wave <= (size_t)kMaxLdsBytes ? 12 : 8)
                          : kUnstagedWaves;
    while (p.wpw[r] > 1 && shared + p.wpw[r] * per_wave > (size_t)kMaxLdsBytes) p.wpw[r] /= 2;
    p.sh1[r] = shared + p.wpw[r] * per_wave;
    if (p.sh1[r] > (size_t)kMaxLdsBytes) return false;
    p.pblocks[r] = std::max<int64_t>(1, std::min<int64_t>(2048, (p.n1 + p.wpw[r] - 1) / p.wpw[r]));
    // column kernel: nb <= 32 and at most 16 occupied columns (both spins)
    const int ncol = ns == 1 ? std::max(p.nup, p.ndown) : p.nup + p.ndown;
    const int noc = ncol <= 4 ? 4 : (ncol <= 5 ? 5 : (ncol <= 8 ? 8 : 16));
    const int nt = p.gv ? 2 : 1, nrow = p.gv ? 8 : 4;
    const int nc0 = ns == 1 ? std::max(p.nup, p.ndown) : p.nup;
    p.col_noc[r] = noc;
    p.col_nent[r] = col_entries(ns, nt, nc0, p.ndown);
    const int ntask = col_tasks(ns, nt, nc0, p.ndown, noc);
    p.col_sh[r] = sizeof(double) * (col_block_doubles(noc, p.col_nent[r], ntask) +
                                    4 * col_wave_doubles(nrow, p.f32 ? 4 : 8, nb, col_ldw(noc, nt), p.col_nent[r],
                                                         ns, nt, noc));
    // (FP32 rows of one V tensor: with FP64 rows or independent V_same / V_oppo
    // its staging registers spill and the D-based kernel is faster)
    p.col[r] = nb <= 32 && ncol <= 16 && p.f32 && !p.gv && (ns == 2 || p.nup == p.ndown) &&
               p.col_sh[r] <= (size_t)kMaxLdsBytes;
  }
  return true;
}

bool tc_restricted(const TcPlan& p, const double* D, const double* C) {
  const size_t nb2 = (size_t)p.nb * p.nb, nc = (size_t)p.nb * p.nmax;
  return p.nup == p.ndown && std::memcmp(D, D + nb2, nb2 * sizeof(double)) == 0 &&
         std::memcmp(C, C + nc, nc * sizeof(double)) == 0;
}

// D_s = C_s C_s^T for the occupied columns of each spin formed (to rounding:
// 1e-14 of the largest |D|), the condition of the column point kernel.
static bool density_from_orbitals(const TcPlan& p, const double* D, const double* C, bool restricted) {
  const int nb = p.nb;
  double dmax = 0;
  for (int64_t q = 0; q < 2 * (int64_t)nb * nb; ++q) dmax = std::max(dmax, std::fabs(D[q]));
  const double tol = 1e-14 * (1.0 + dmax);
  for (int s = 0; s < (restricted ? 1 : 2); ++s) {
    const int no = s ? p.ndown : p.nup;
    const double* Cs = C + (size_t)s * nb * p.nmax;
    for (int u = 0; u < nb; ++u)
      for (int v = 0; v < nb; ++v) {
        double t = 0;
        for (int b = 0; b < no; ++b) t += Cs[(size_t)u * p.nmax + b] * Cs[(size_t)v * p.nmax + b];
        if (!(std::fabs(t - D[((size_t)s * nb + u) * nb + v]) <= tol)) return false;
      }
  }
  return true;
}

bool tc_col_ok(const TcPlan& p, const double* D, const double* C, bool restricted) {
  return p.col[restricted ? 1 : 0] && density_from_orbitals(p, D, C, restricted);
}

// The build's kernels from D and C already in p.dD / p.dC; the result block
// (tc_out_doubles) goes to out_dev.
static int tc_build(const TcPlan& p, bool restricted, bool col, hipStream_t st, double* out_dev) {
  const int64_t nb2 = (int64_t)p.nb * p.nb;
  const int nst = restricted ? 3 : 6;  // rank targets (spin, term) formed
  const int r = restricted ? 1 : 0;
  if (col) {
    // orbitals packed by column, then the column point kernel (its V-linear
    // sums included) over kBlocks workgroups of 4 waves
    const int ns = restricted ? 1 : 2, noc = p.col_noc[r];
    double* dw = p.dCc + 32 * 16;
    hipLaunchKernelGGL(tc_pack_cols, dim3(1), dim3(256), 0, st, p.dC, p.dD, p.nb, p.nmax, p.nup, p.ndown, ns, noc,
                       p.dCc, dw);
    TC_HIP(hipGetLastError());
    ColArgs ca{p.Fs, p.Fo, p.V, p.Vo, p.A1, p.dA1, p.w1, p.c_same, p.c_oppo, p.n1, p.npair, p.nb, p.nmax,
               p.nup, p.ndown, p.nvec, p.col_nent[r], p.dVec, p.dAl, p.dPV};
    const double* cc = p.dCc;
    const double* dwc = dw;
    void* cargs[] = {&ca, &cc, &dwc};
    TC_HIP(hipLaunchKernel(col_kernel(restricted, noc), dim3(kBlocks), dim3(256), cargs, p.col_sh[r],
                           st));
  } else {
    PointArgs a{p.Fs, p.Fo, p.V, p.Vo, p.A1, p.dA1, p.w1, p.dD, p.dC, p.c_same, p.c_oppo, p.n1, p.npair, p.nb,
                p.nmax, p.nup, p.ndown, p.nvec, p.dVec, p.dAl, p.dH};
    void* pargs[] = {&a};
    TC_HIP(hipLaunchKernel(point_kernel(p.f32, p.stage[r], restricted, p.gv), dim3((unsigned)p.pblocks[r]),
                           dim3(64 * p.wpw[r]), pargs, p.sh1[r], st));
  }
  RankArgs ra{p.dVec, p.dAl, p.A1, p.dA1, p.w1, p.n1, p.nb, p.nvec, p.nup, p.ndown, p.nmax, p.nwt, {}, p.dPG, {}};
  RankRanges rr{};
  for (int q = 0; q < 7; ++q) ra.ustart[q] = rr.ustart[q] = p.ustart[q];
  for (int q = 0; q < 6; ++q) ra.pts[q] = p.rpts[q];
  const dim3 rgrid((unsigned)p.ustart[nst]);
  if (p.nb <= 32)
    hipLaunchKernelGGL(tc_rank_kernel<2>, rgrid, dim3(256), 0, st, ra);
  else if (p.nb <= 48)
    hipLaunchKernelGGL(tc_rank_kernel<3>, rgrid, dim3(256), 0, st, ra);
  else
    hipLaunchKernelGGL(tc_rank_kernel<4>, rgrid, dim3(256), 0, st, ra);
  TC_HIP(hipGetLastError());
  const int nk = restricted ? 1 : 2;  // spins of the V-linear sums
  if (!col) {
    const dim3 vg(kBlocks), vb(256);
#define TC_VLIN(ST, NK, NT)                                                                                       \
  hipLaunchKernelGGL((tc_vlin_kernel<ST, NK, NT>), vg, vb, 0, st, (const ST*)p.V, (const ST*)(p.gv ? p.Vo : p.V), \
                     p.dH, p.n1, p.npair, p.dPV)
    if (p.f32) {
      if (p.gv) { if (restricted) TC_VLIN(float, 1, 2); else TC_VLIN(float, 2, 2); }
      else { if (restricted) TC_VLIN(float, 1, 1); else TC_VLIN(float, 2, 1); }
    } else {
      if (p.gv) { if (restricted) TC_VLIN(double, 1, 2); else TC_VLIN(double, 2, 2); }
      else { if (restricted) TC_VLIN(double, 1, 1); else TC_VLIN(double, 2, 1); }
    }
#undef TC_VLIN
    TC_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(tc_reduce_ranges, dim3((unsigned)(nst * ((nb2 + 15) / 16))), dim3(1024), 0, st, p.dPG, rr, nb2,
                     out_dev);
  TC_HIP(hipGetLastError());
  hipLaunchKernelGGL(tc_reduce, dim3((unsigned)((nk * p.npair + 15) / 16)), dim3(1024), 0, st, p.dPV, kBlocks,
                     nk * p.npair, out_dev + 6 * nb2);
  TC_HIP(hipGetLastError());
  return 0;
}

int tc_enqueue(const TcPlan& p, const double* D_host, const double* C_host, double* out_host, bool restricted,
               hipStream_t st) {
  const int64_t nb2 = (int64_t)p.nb * p.nb;
  TC_HIP(hipMemcpyAsync(p.dD, D_host, sizeof(double) * 2 * nb2, hipMemcpyHostToDevice, st));
  TC_HIP(hipMemcpyAsync(p.dC, C_host, sizeof(double) * 2 * p.nb * p.nmax, hipMemcpyHostToDevice, st));
  const int rc = tc_build(p, restricted, tc_col_ok(p, D_host, C_host, restricted), st, p.dOut);
  if (rc) return rc;
  TC_HIP(hipMemcpyAsync(out_host, p.dOut, sizeof(double) * tc_out_doubles(p.nb), hipMemcpyDeviceToHost, st));
  return 0;
}

int tc_enqueue_mapped(const TcPlan& p, const double* DC_mapped, double* out_mapped, bool restricted, bool col,
                      hipStream_t st) {
  // p.dC follows p.dD in the workspace, as C follows D in the mapped buffer
  const int64_t n = 2 * (int64_t)p.nb * p.nb + 2 * (int64_t)p.nb * p.nmax;
  hipLaunchKernelGGL(copy_doubles, copy_grid(n), dim3(256), 0, st, DC_mapped, p.dD, n);
  TC_HIP(hipGetLastError());
  return tc_build(p, restricted, col, st, out_mapped);
}


void tc_assemble(const TcPlan& p, const double* out, bool restricted, double* G2, double* G3) {
  const int nb = p.nb;
  const int64_t nb2 = (int64_t)nb * nb;
  for (int s = 0; s < 2; ++s) {
    const int so = restricted ? 0 : s;  // the spin whose blocks were formed
    const double* Sc = out + (so * 3 + 0) * nb2;
    const double* Vc = out + (so * 3 + 1) * nb2;
    const double* S3 = out + (so * 3 + 2) * nb2;
    const double* Rg3 = out + 6 * nb2 + so * p.npair;
    for (int j = 0; j < nb; ++j)
      for (int i = 0; i < nb; ++i) {
        G2[s * nb2 + j * nb + i] = Sc[j * nb + i] / 2 + (Vc[j * nb + i] / 2 - Vc[i * nb + j] / 2);
        G3[s * nb2 + j * nb + i] = (S3[j * nb + i] + Rg3[pk(j, i)]) / 2;
      }
  }
}

TcIter::TcIter(int nup, int ndown, int nb_, const double* S_, const double* H_, const tcscf_scf_opts* o_)
    : nb(nb_), ne{nup, ndown}, nmax(std::max(1, std::max(nup, ndown))), S(S_), H(H_), o(o_),
      D(2 * (size_t)nb_ * nb_, 0.0), C(2 * (size_t)nb_ * std::max(1, std::max(nup, ndown)), 0.0),
      Fock(2 * (size_t)nb_ * nb_), pseudo(nup + ndown, 0.0), diis(o_->diis) {}

int TcIter::step(double* G2, double* G3, int it, tcscf_scf_result* res, double* trace, bool* done) {
  const int64_t nb2 = (int64_t)nb * nb;
  *done = false;
  if (o->mask)
    for (int q = 0; q < 2 * nb2; ++q)
      if (!(std::fabs(H[q % nb2]) > 0)) G2[q] = G3[q] = 0;
  double e = 0;
  for (int s = 0; s < 2; ++s)
    for (int j = 0; j < nb; ++j)
      for (int i = 0; i < nb; ++i) {
        const double g2 = G2[s * nb2 + j * nb + i], g3 = G3[s * nb2 + j * nb + i];
        Fock[s * nb2 + j * nb + i] = H[j * nb + i] + g2 - g3;
        e += D[s * nb2 + i * nb + j] * (H[j * nb + i] + g2 / 2 - g3 / 3);
      }
  if (trace) trace[it] = e;
  res->iterations = it + 1;
  if (std::fabs(e - energy) < o->tol) {
    res->energy = e;
    res->status = TCSCF_OK;
    *done = true;
    return 0;
  }
  energy = e;
  if (o->diis > 0 && it > 0) {  // iteration 0 has D = 0, whose error is 0 trivially
    dense::right_eigen_error(2, nb, Fock.data(), D.data(), S, err);
    diis.extrapolate(Fock, err);
  }
  for (double f : Fock)
    if (!std::isfinite(f)) {
      res->status = TCSCF_ESOLVER;
      set_error("TC SCF: non-finite Fock matrix (F / V inputs not finite?)");
      return TCSCF_ESOLVER;
    }
  // Closed shell with bitwise-equal spin operators: the spin-1 solve would
  // repeat spin 0's exactly, so its orbitals, density and pseudo-energies are
  // copied instead.
  const bool same = ne[0] == ne[1] && std::memcmp(Fock.data(), Fock.data() + nb2, sizeof(double) * nb2) == 0;
  for (int s = 0; s < (same ? 1 : 2); ++s) {
    std::vector<std::complex<double>> ev, X;
    if (!dense::gen_nonsym_eig(nb, Fock.data() + s * nb2, S, ev, X)) {
      res->status = TCSCF_ESOLVER;
      set_error("TC SCF: generalized eigen-solve failed");
      return TCSCF_ESOLVER;
    }
    // the ordering below needs finite real parts (a strict weak order)
    for (const auto& x : ev)
      if (!std::isfinite(x.real())) {
        res->status = TCSCF_ESOLVER;
        set_error("TC SCF: non-finite eigenvalue");
        return TCSCF_ESOLVER;
      }
    std::vector<int> idx(nb);
    std::iota(idx.begin(), idx.end(), 0);
    std::sort(idx.begin(), idx.end(), [&](int a, int b) { return ev[a].real() < ev[b].real(); });
    const int no = ne[s];
    std::vector<std::complex<double>> vec((size_t)no * nb);
    for (int b = 0; b < no; ++b) {
      pseudo[s * ne[0] + b] = ev[idx[b]].real();
      for (int j = 0; j < nb; ++j) vec[b * nb + j] = X[j * nb + idx[b]];
    }
    dense::mgs(no, nb, vec, S, o->mgs_mode == 0);
    // real basis: remove each orbital's global phase (density unchanged)
    for (int b = 0; b < no; ++b) {
      int jm = 0;
      for (int j = 1; j < nb; ++j)
        if (std::abs(vec[b * nb + j]) > std::abs(vec[b * nb + jm])) jm = j;
      const std::complex<double> ph =
          std::abs(vec[b * nb + jm]) > 0 ? std::conj(vec[b * nb + jm]) / std::abs(vec[b * nb + jm]) : 1.0;
      for (int j = 0; j < nb; ++j) C[((size_t)s * nb + j) * nmax + b] = (vec[b * nb + j] * ph).real();
    }
    for (int u = 0; u < nb; ++u)
      for (int v = 0; v < nb; ++v) {
        double t = 0;
        for (int b = 0; b < no; ++b) t += C[((size_t)s * nb + u) * nmax + b] * C[((size_t)s * nb + v) * nmax + b];
        D[s * nb2 + u * nb + v] = t;
      }
  }
  if (same) {
    std::copy(C.begin(), C.begin() + (size_t)nb * nmax, C.begin() + (size_t)nb * nmax);
    std::copy(D.begin(), D.begin() + nb2, D.begin() + nb2);
    for (int b = 0; b < ne[1]; ++b) pseudo[ne[0] + b] = pseudo[b];
  }
  return 0;
}

}  // namespace tc

using namespace tc;
using namespace tc::tcs;

extern "C" {

/* One TC Fock build on the device: from D (2 x nb x nb) and the occupied
 * orbitals C (2 x nb x nmax, column b of spin s valid for b < nocc_s) to
 * G2 and G3 (each 2 x nb x nb, host), before the one-electron mask. */
int tcscf_tc_fock(int nup, int ndown, int nb, const double* A1, const double* dA1, const double* w1, int64_t n1,
                  const void* Fs, const void* Fo, const void* V, const void* Vo, int f_dtype,
                  const tcscf_jastrow* jp, const double* D_host, const double* C_host, double* G2_host,
                  double* G3_host, void* stream) {
  TC_CHECK(D_host && C_host && G2_host && G3_host, "tcscf_tc_fock: null");
  hipStream_t st = (hipStream_t)stream;
  int rc = 0;
  double* ws = (double*)scratch(kSlotSweep, sizeof(double) * tc_ws_doubles(nup, ndown, nb, n1, Vo != nullptr), &rc);
  if (rc) return rc;
  TcPlan plan;
  if ((rc = tc_prepare(plan, nup, ndown, nb, A1, dA1, w1, n1, Fs, Fo, V, Vo, f_dtype, jp, ws))) return rc;
  std::vector<double> out(tc_out_doubles(nb));
  const bool restricted = tc_restricted(plan, D_host, C_host);
  if ((rc = tc_enqueue(plan, D_host, C_host, out.data(), restricted, st))) return rc;
  TC_HIP(hipStreamSynchronize(st));
  tc_assemble(plan, out.data(), restricted, G2_host, G3_host);
  return 0;
}

/* The point-kernel launch the TC Fock build picks: info[0..2] = V rows staged
 * in LDS, waves per workgroup, LDS bytes for two spins; info[3..5] the same
 * for the restricted closed-shell build. */
int tcscf_tc_launch_info(int nup, int ndown, int nb, int f_dtype, int general_v, int* info) {
  TC_CHECK(info && nb > 0 && nb <= 64 && nup >= 0 && ndown >= 0 && nup <= nb && ndown <= nb,
           "tcscf_tc_launch_info: bad arguments");
  TC_CHECK(f_dtype == TCSCF_F64 || f_dtype == TCSCF_F32, "tcscf_tc_launch_info: f_dtype must be TCSCF_F64 or TCSCF_F32");
  TcPlan p{};
  p.nb = nb;
  p.nmax = std::max(1, std::max(nup, ndown));
  p.f32 = f_dtype == TCSCF_F32;
  p.gv = general_v ? 1 : 0;
  p.nwt = p.gv ? 2 : 1;
  p.n1 = 1;
  TC_CHECK(tc_launch_shape(p), "tcscf_tc_launch_info: basis / occupied orbitals too large for the LDS staging");
  for (int r = 0; r < 2; ++r) {
    info[3 * r] = p.stage[r];
    info[3 * r + 1] = p.wpw[r];
    info[3 * r + 2] = (int)p.sh1[r];
  }
  return 0;
}

/* TCHartreeFock::compute (src/HartreeFock.cpp:1007-1362), real basis.
 * Fs, Fo: packed n1 x npair; Vs, Vo: packed n1 x 3 x npair V_same / V_oppo,
 * or Vo = NULL and Vs the c = 1 tensor scaled by jp's c_same / c_oppo (all
 * f_dtype: TCSCF_F32 stores them in FP32, the Fock build's arithmetic is
 * FP64); A1: n1 x nb; dA1: n1 x nb x 3.  D_out 2 x nb x nb, pseudo_out
 * nup + ndown, obs: the solver's observable members (may be NULL). */
int tcscf_tc_scf(int nup, int ndown, int nb, const double* S, const double* H, const double* A1, const double* dA1,
                 const double* w1, int64_t n1, const void* Fs, const void* Fo, const void* Vs, const void* Vo,
                 int f_dtype, const tcscf_jastrow* jp, const tcscf_scf_opts* o, tcscf_scf_result* res, double* trace,
                 double* D_out, double* pseudo_out, tcscf_scf_observables* obs, void* stream) {
  TC_CHECK(S && H && o && res && o->max_iter > 0, "tcscf_tc_scf: bad arguments");
  TC_CHECK(o->diis >= 0, "tcscf_tc_scf: diis depth must be >= 0");
  TC_CHECK(o->graph >= 0 && o->graph <= 2, "tcscf_tc_scf: graph must be 0 (default), 1 (graph) or 2 (direct)");
  TC_CHECK(o->diis == 0 || o->mgs_mode == 1, "tcscf_tc_scf: DIIS needs S-orthonormal orbitals (mgs_mode = 1)");
  hipStream_t st = (hipStream_t)stream;
  const int64_t nb2 = (int64_t)nb * nb;
  int rc = 0;
  IterContext& cx = iter_context(1, st);
  std::lock_guard<std::mutex> lock(cx.mu);
  {
    const size_t nmx = std::max(1, std::max(nup, ndown));
    const size_t ndc = 2 * (size_t)nb2 + 2 * (size_t)nb * nmx + tc_out_doubles(nb);  // D, C, result block
    if ((rc = cx.prepare(ndc, sizeof(double) * tc_ws_doubles(nup, ndown, nb, n1, Vo != nullptr)))) return rc;
  }
  double* ws = (double*)cx.ws;
  TcPlan plan;
  std::memset(&plan, 0, sizeof plan);  // its bytes key the iteration graph
  if ((rc = tc_prepare(plan, nup, ndown, nb, A1, dA1, w1, n1, Fs, Fo, Vs, Vo, f_dtype, jp, ws))) return rc;
  TcIter iter(nup, ndown, nb, S, H, o);
  std::vector<double> G2(2 * nb2), G3(2 * nb2);
  // D and C up, the result block down, through the device's coherent
  // page-locked buffer the device maps: each iteration replays one HIP graph
  // (upload kernel, the build's kernels, reductions writing into the mapped
  // buffer), captured again when the kernel choice changes (restricted build,
  // column kernel).  Where the mapped buffer cannot be had, or with
  // TCSCF_SCF_GRAPH=0, every iteration is enqueued on the caller's stream with
  // copies (pageable vectors if the page-locked allocation failed).
  const size_t nD = iter.D.size(), nC = iter.C.size(), nO = tc_out_doubles(nb);
  std::vector<double> stage;
  if (!cx.host) stage.resize(nD + nC + nO);
  double* Dh = cx.host ? cx.host : stage.data();
  double *Ch = Dh + nD, *outh = Ch + nC;
  const bool use_graph = IterContext::enabled(1, o->graph) && cx.dev;
  if (use_graph) TC_HIP(hipStreamSynchronize(st));  // the tensors come from the caller's stream
  res->status = TCSCF_ENOTCONV;
  res->energy = std::numeric_limits<double>::max();
  res->iterations = 0;
  res->t_sweep_ms = res->t_solve_ms = 0;
  res->graph_iterations = res->graph_captures = 0;
  for (int it = 0; it < o->max_iter; ++it) {
    double t0 = now_ms();
    const bool restricted = tc_restricted(plan, iter.D.data(), iter.C.data());
    std::copy(iter.D.begin(), iter.D.end(), Dh);
    std::copy(iter.C.begin(), iter.C.end(), Ch);
    if (use_graph) {
      const bool col = tc_col_ok(plan, Dh, Ch, restricted);
      std::vector<unsigned char> key;
      key_append(key, plan);
      key_append(key, cx.dev);
      key_append(key, (int)restricted);
      key_append(key, (int)col);
      if ((rc = cx.run(key, [&](hipStream_t s) {
             return tc_enqueue_mapped(plan, cx.dev, cx.dev + nD + nC, restricted, col, s);
           })))
        return rc;
      ++res->graph_iterations;
      res->graph_captures = cx.captures;
    } else {
      if ((rc = tc_enqueue(plan, Dh, Ch, outh, restricted, st))) return rc;
      TC_HIP(hipStreamSynchronize(st));
    }
    tc_assemble(plan, outh, restricted, G2.data(), G3.data());
    double t1 = now_ms();
    res->t_sweep_ms += t1 - t0;
    bool done = false;
    if ((rc = iter.step(G2.data(), G3.data(), it, res, trace, &done))) return rc;
    res->t_solve_ms += now_ms() - t1;
    if (done) break;
  }
  if (D_out) std::copy(iter.D.begin(), iter.D.end(), D_out);
  if (pseudo_out) std::copy(iter.pseudo.begin(), iter.pseudo.end(), pseudo_out);
  if (obs) {
    if (obs->fock) std::copy(iter.Fock.begin(), iter.Fock.end(), obs->fock);
    if (obs->occupied) std::copy(iter.C.begin(), iter.C.end(), obs->occupied);
  }
  if (res->status != TCSCF_OK) set_error("TC SCF: did not converge");
  return res->status;
}

}  // extern "C"
