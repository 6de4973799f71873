// Transcorrelated Fock build and SCF driver (TCHartreeFock::compute,
// src/HartreeFock.cpp:1007-1362), real basis.
//
// Inputs per r1 point n: packed F_same[n][p], F_oppo[n][p] (scalar TC kernel
// with c_same / c_oppo) and V[n][xyz][p] for c = 1 (V_same = c_same V,
// V_oppo = c_oppo V); basis values A[n][k] and gradients dA[n][k][xyz].
//
// Every term of the reference's G2 / G3 assembly (:1157-1224) is one of
//   (a) a rank-one product L_j R_i of per-point vectors      -> "rows"
//   (b) a per-point 3-vector h dotted into V(n, j, i)        -> "V-linear"
// Pass 1 (tc_point_kernel, one wave per point) forms every per-point
// intermediate of the reference (:1118-1148) and writes the rows and the h
// vectors; pass 2 (tc_rank_kernel) sums L^T R over points and rows per target
// matrix; pass 3 (tc_vlin_kernel) sums h . V over points.  Partials per
// workgroup are reduced in a fixed order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <limits>
#include <numeric>
#include <vector>

#include "common.hpp"
#include "dense.hpp"

namespace tc {
namespace tcs {

constexpr int kBlocks = 512;
constexpr int kMaxOcc = 8;  // occupied orbitals per spin handled by the point kernel

// Row layout per (point, spin): 2 vectors (L, R) of nb doubles each.
//   0      G2 : (A, w (faa A - FaiS))
//   1      VEC: (A, w (VT . dA - c_s Vai1))
//   2..4   VEC: (-w VTSI_c, dA_c)                  (Vja == VTSI for symmetric V)
//   5      G3 : (A, w (|VT|^2 - sNS - sNO_other) A + Z)
//   6      G3 : (Z, A),  Z = 2 w (SVbiVT - VTSI . VT)
//   7..9   G3 : (2 w VTSI_c, VTSI_c)
//   10..   G3 : (alpha W_bc, W_bc), b < nocc_s, alpha = -2 w (occ_s c_s^2 + occ_o c_o^2)
constexpr int kFixedRows = 10;
__host__ __device__ inline int target_of_row(int r) { return r == 0 ? 0 : (r < 5 ? 1 : 2); }

struct PointArgs {
  const double* Fs;   // [n1][npair]
  const double* Fo;   // [n1][npair]
  const double* V;    // [n1][3][npair]
  const double* A;    // [n1][nb]
  const double* dA;   // [n1][nb][3]
  const double* w;    // [n1]
  const double* D;    // [2][nb][nb]
  const double* C;    // [2][nb][nmax]
  double c_same, c_oppo;
  int64_t n1, npair;
  int nb, nmax, nocc0, nocc1, nrow;
  double* rows;       // [n1][2][nrow][2][nb]
  double* hvec;       // [n1][2 spins][2 targets (VEC, G3)][3]
};

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ inline void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Pass 1: one wave per r1 point, lane i = basis index (nb <= 64).
__global__ __launch_bounds__(256) void tc_point_kernel(PointArgs a) {
  extern __shared__ double smem[];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nb = a.nb, nmax = a.nmax;
  const int64_t np = a.npair, nb2 = (int64_t)nb * nb;
  // workgroup-shared: D [2][nb][nb], C [2][nb][nmax]
  double* Ds = smem;
  double* Cs = Ds + 2 * nb2;
  // per wave: f_s, f_o [np], v [3][np], A [nb], dA [nb][3], X [2][nb], kS [2][nb][3],
  //           VTSI [2][nb][3], E [2][nb][3], W [2][nb][nmax][3], U scratch
  const int64_t per_wave = 5 * np + nb + 3 * nb + 2 * nb + 6 * nb + 6 * nb + 6 * nb + 6 * (int64_t)nb * nmax;
  double* base = Cs + 2 * (int64_t)nb * nmax + wave * per_wave;
  double* fs = base;
  double* fo = fs + np;
  double* vv = fo + np;  // [3][np]
  double* As = vv + 3 * np;
  double* dAs = As + nb;
  double* Xs = dAs + 3 * nb;
  double* kS = Xs + 2 * nb;
  double* VT = kS + 6 * nb;  // VTSI [2][nb][3]
  double* Es = VT + 6 * nb;
  double* Ws = Es + 6 * nb;  // [2][nb][nmax][3]
  for (int64_t e = tid; e < 2 * nb2; e += 256) Ds[e] = a.D[e];
  for (int64_t e = tid; e < 2 * (int64_t)nb * nmax; e += 256) Cs[e] = a.C[e];
  __syncthreads();

  const int i = lane;
  const bool act = i < nb;
  const int tri_i = i * (i + 1) / 2;
  const double cs = a.c_same, co = a.c_oppo;
  const int nocc[2] = {a.nocc0, a.nocc1};
  const int64_t per = (a.n1 + gridDim.x - 1) / gridDim.x;
  const int64_t p0 = (int64_t)blockIdx.x * per;
  const int64_t p1 = p0 + per < a.n1 ? p0 + per : a.n1;

  for (int64_t n = p0 + wave; n < p1; n += 4) {
    for (int64_t p = lane; p < np; p += 64) {
      fs[p] = a.Fs[n * np + p];
      fo[p] = a.Fo[n * np + p];
      vv[p] = a.V[(n * 3 + 0) * np + p];
      vv[np + p] = a.V[(n * 3 + 1) * np + p];
      vv[2 * np + p] = a.V[(n * 3 + 2) * np + p];
    }
    if (act) {
      As[i] = a.A[n * nb + i];
      for (int c = 0; c < 3; ++c) dAs[i * 3 + c] = a.dA[(n * nb + i) * 3 + c];
    }
    wave_sync_lds();
    // X_s[i] = sum_k D_s[k][i] A[k];  kS_s[i] = sum_k D_s[k][i] dA[k]
    if (act)
      for (int s = 0; s < 2; ++s) {
        double x = 0, k0 = 0, k1 = 0, k2 = 0;
        for (int k = 0; k < nb; ++k) {
          const double d = Ds[s * nb2 + k * nb + i];
          x += d * As[k];
          k0 += d * dAs[k * 3];
          k1 += d * dAs[k * 3 + 1];
          k2 += d * dAs[k * 3 + 2];
        }
        Xs[s * nb + i] = x;
        kS[(s * nb + i) * 3] = k0;
        kS[(s * nb + i) * 3 + 1] = k1;
        kS[(s * nb + i) * 3 + 2] = k2;
      }
    wave_sync_lds();
    const double Ai = act ? As[i] : 0.0;
    double dAi[3] = {0, 0, 0};
    if (act)
      for (int c = 0; c < 3; ++c) dAi[c] = dAs[i * 3 + c];
    double occ[2], sg[2][3];
    for (int s = 0; s < 2; ++s) {
      const double x = act ? Xs[s * nb + i] : 0.0;
      occ[s] = wave_sum(x * Ai);  // calculateOccupiedElectronDensity
      for (int c = 0; c < 3; ++c) sg[s][c] = wave_sum(x * dAi[c]);  // calculateScaledGradients
    }
    // one pass over j: Faa, FaiS, V . D, Vai, Q (= VTSI / c_s), W (registers)
    double faa_p[2] = {0, 0}, fai[2] = {0, 0}, vd[2][3] = {{0, 0, 0}, {0, 0, 0}}, vai1[2] = {0, 0};
    double q[2][3] = {{0, 0, 0}, {0, 0, 0}};
    double wreg[2][kMaxOcc][3];
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
      for (int b = 0; b < kMaxOcc; ++b) wreg[s][b][0] = wreg[s][b][1] = wreg[s][b][2] = 0;
    if (act) {
      for (int j = 0; j < nb; ++j) {
        const int64_t pj = j <= i ? tri_i + j : j * (j + 1) / 2 + i;
        const double f_s = fs[pj], f_o = fo[pj];
        const double v0 = vv[pj], v1 = vv[np + pj], v2 = vv[2 * np + pj];
#pragma unroll
        for (int s = 0; s < 2; ++s) {
          const double dsj = Ds[s * nb2 + j * nb + i], dbj = Ds[(1 - s) * nb2 + j * nb + i];
          faa_p[s] += f_s * dsj + f_o * dbj;
          const double xj = Xs[s * nb + j];
          fai[s] += xj * f_s;
          vd[s][0] += v0 * dsj;
          vd[s][1] += v1 * dsj;
          vd[s][2] += v2 * dsj;
          const double* ks = kS + (s * nb + j) * 3;
          vai1[s] += v0 * ks[0] + v1 * ks[1] + v2 * ks[2];
          q[s][0] += xj * v0;
          q[s][1] += xj * v1;
          q[s][2] += xj * v2;
          const double* crow = Cs + ((int64_t)s * nb + j) * nmax;
#pragma unroll
          for (int b = 0; b < kMaxOcc; ++b)
            if (b < nocc[s]) {
              const double cb = crow[b];
              wreg[s][b][0] += cb * v0;
              wreg[s][b][1] += cb * v1;
              wreg[s][b][2] += cb * v2;
            }
        }
      }
#pragma unroll
      for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int b = 0; b < kMaxOcc; ++b)
          if (b < nocc[s]) {
            double* wrow = Ws + (((int64_t)(s * nb + i) * nmax) + b) * 3;
            wrow[0] = wreg[s][b][0];
            wrow[1] = wreg[s][b][1];
            wrow[2] = wreg[s][b][2];
          }
    }
    double faa[2], VTv[2][3];
    double vdsum[2][3];
    for (int s = 0; s < 2; ++s) {
      faa[s] = wave_sum(faa_p[s]);  // FaaSamePlusOppo
      for (int c = 0; c < 3; ++c) vdsum[s][c] = wave_sum(vd[s][c]);
    }
    for (int s = 0; s < 2; ++s)
      for (int c = 0; c < 3; ++c) VTv[s][c] = cs * vdsum[s][c] + co * vdsum[1 - s][c];  // VaaSamePlusOppo == VTilde
    if (act)
      for (int s = 0; s < 2; ++s)
        for (int c = 0; c < 3; ++c) VT[(s * nb + i) * 3 + c] = cs * q[s][c];  // VTildeSubI (== Vja)
    wave_sync_lds();
    // E_s[k] = sum_l D_s[l][k] VTSI_s[l]
    if (act)
      for (int s = 0; s < 2; ++s) {
        double e0 = 0, e1 = 0, e2 = 0;
        for (int l = 0; l < nb; ++l) {
          const double d = Ds[s * nb2 + l * nb + i];
          const double* t = VT + (s * nb + l) * 3;
          e0 += d * t[0];
          e1 += d * t[1];
          e2 += d * t[2];
        }
        Es[(s * nb + i) * 3] = e0;
        Es[(s * nb + i) * 3 + 1] = e1;
        Es[(s * nb + i) * 3 + 2] = e2;
      }
    wave_sync_lds();
    // SumOverb_Vbi_dot_VTildeb_s[i] = c_s sum_k V(k,i) . E_s[k]
    double sv1[2] = {0, 0};
    if (act)
      for (int k = 0; k < nb; ++k) {
        const int64_t pk = k <= i ? tri_i + k : k * (k + 1) / 2 + i;
        const double v0 = vv[pk], v1 = vv[np + pk], v2 = vv[2 * np + pk];
        for (int s = 0; s < 2; ++s) {
          const double* e = Es + (s * nb + k) * 3;
          sv1[s] += v0 * e[0] + v1 * e[1] + v2 * e[2];
        }
      }
    double scvt[2][3], nrmU[2];
    for (int s = 0; s < 2; ++s) {
      const double x = act ? Xs[s * nb + i] : 0.0;
      for (int c = 0; c < 3; ++c) scvt[s][c] = wave_sum(act ? x * VT[(s * nb + i) * 3 + c] : 0.0);  // SumOverb_conjVTildeb_occb
      // sum_ab |U_s[a][b]|^2, U_s[a][b][c] = sum_l C_s[l][a] W_s[l][b][c]  (calculateVab)
      double acc = 0;
      const int nu = nocc[s] * nocc[s] * 3;
      for (int t = lane; t < nu; t += 64) {
        const int aa = t / (nocc[s] * 3), rem = t - aa * nocc[s] * 3, bb = rem / 3, c = rem - bb * 3;
        double u = 0;
        for (int l = 0; l < nb; ++l) u += Cs[((int64_t)s * nb + l) * nmax + aa] * Ws[(((int64_t)s * nb + l) * nmax + bb) * 3 + c];
        acc += u * u;
      }
      nrmU[s] = wave_sum(acc);
    }
    // rows and h vectors
    const double wn = a.w[n];
    double* rbase = a.rows + (int64_t)n * 2 * a.nrow * 2 * nb;
    for (int s = 0; s < 2; ++s) {
      const int sb = 1 - s;
      double* rs = rbase + (int64_t)s * a.nrow * 2 * nb;
      auto put = [&](int r, double L, double R) {
        if (act) {
          rs[((int64_t)r * 2) * nb + i] = L;
          rs[((int64_t)r * 2 + 1) * nb + i] = R;
        }
      };
      const double* vts = act ? VT + (s * nb + i) * 3 : nullptr;
      const double vtsi0 = act ? vts[0] : 0, vtsi1 = act ? vts[1] : 0, vtsi2 = act ? vts[2] : 0;
      const double vtdot = VTv[s][0] * vtsi0 + VTv[s][1] * vtsi1 + VTv[s][2] * vtsi2;
      const double vtn = VTv[s][0] * VTv[s][0] + VTv[s][1] * VTv[s][1] + VTv[s][2] * VTv[s][2];
      const double svbi = cs * sv1[s];
      const double Z = 2 * wn * (svbi - vtdot);
      const double sNS = cs * cs * nrmU[s], sNOb = co * co * nrmU[sb];
      put(0, Ai, wn * (faa[s] * Ai - fai[s]));
      put(1, Ai, wn * (VTv[s][0] * dAi[0] + VTv[s][1] * dAi[1] + VTv[s][2] * dAi[2] - cs * vai1[s]));
      put(2, -wn * vtsi0, dAi[0]);
      put(3, -wn * vtsi1, dAi[1]);
      put(4, -wn * vtsi2, dAi[2]);
      put(5, Ai, wn * (vtn - sNS - sNOb) * Ai + Z);
      put(6, Z, Ai);
      put(7, 2 * wn * vtsi0, vtsi0);
      put(8, 2 * wn * vtsi1, vtsi1);
      put(9, 2 * wn * vtsi2, vtsi2);
      const double alpha = -2 * wn * (occ[s] * cs * cs + occ[sb] * co * co);
      for (int b = 0; b < nmax; ++b)
        for (int c = 0; c < 3; ++c) {
          const double wv = (act && b < nocc[s]) ? Ws[(((int64_t)s * nb + i) * nmax + b) * 3 + c] : 0.0;
          put(kFixedRows + b * 3 + c, alpha * wv, wv);
        }
      if (lane < 3) {
        const int c = lane;
        double* h = a.hvec + ((int64_t)n * 2 + s) * 6;
        h[c] = wn * (cs * sg[s][c] + co * sg[sb][c]);
        h[3 + c] = wn * (2 * occ[s] * cs * VTv[s][c] - 2 * cs * scvt[s][c] + 2 * occ[sb] * co * VTv[sb][c] -
                         2 * co * scvt[sb][c]);
      }
    }
    wave_sync_lds();
  }
}

// Pass 2: G_target,s[j][i] += sum_n sum_rows(target) L[j] R[i].  One
// workgroup row (blockIdx.y) per (spin, target); thread-owned elements
// (j, i) = (e / NBP, e % NBP), e = tid + 256 k; the rows of PB points are staged
// in LDS per step.
template <int NBP>
__global__ __launch_bounds__(256) void tc_rank_kernel(const double* __restrict__ rows, int64_t n1, int nb, int nrow,
                                                      double* __restrict__ part) {
  constexpr int KG = NBP * NBP / 256;
  constexpr int PB = 4;
  extern __shared__ double rl[];  // [PB][nr][2][nb]
  const int tid = threadIdx.x;
  const int st = blockIdx.y, s = st / 3, tg = st - 3 * s;
  const int r0 = tg == 0 ? 0 : (tg == 1 ? 1 : 5), r1 = tg == 0 ? 1 : (tg == 1 ? 5 : nrow);
  const int nr = r1 - r0;
  const int64_t nb2 = (int64_t)nb * nb;
  const int64_t per = (n1 + gridDim.x - 1) / gridDim.x;
  const int64_t p0 = (int64_t)blockIdx.x * per;
  const int64_t p1 = p0 + per < n1 ? p0 + per : n1;
  const int64_t rpt = (int64_t)2 * nrow * 2 * nb;  // doubles per point in `rows`
  const int64_t blk = (int64_t)nr * 2 * nb;        // doubles staged per point
  double acc[KG];
  int jj[KG], ii[KG];
#pragma unroll
  for (int k = 0; k < KG; ++k) {
    acc[k] = 0;
    const int e = tid + 256 * k;
    jj[k] = e / NBP;
    ii[k] = e % NBP;
  }
  for (int64_t n0 = p0; n0 < p1; n0 += PB) {
    const int pn = (int)((n0 + PB < p1) ? PB : p1 - n0);
    __syncthreads();
    for (int64_t e = tid; e < pn * blk; e += 256) {
      const int64_t pt = e / blk, off = e - pt * blk;
      rl[e] = rows[(n0 + pt) * rpt + ((int64_t)s * nrow + r0) * 2 * nb + off];
    }
    __syncthreads();
    for (int t = 0; t < pn * nr; ++t) {
      const double* L = rl + (int64_t)t * 2 * nb;
      const double* R = L + nb;
#pragma unroll
      for (int k = 0; k < KG; ++k)
        if (jj[k] < nb && ii[k] < nb) acc[k] += L[jj[k]] * R[ii[k]];
    }
  }
#pragma unroll
  for (int k = 0; k < KG; ++k)
    if (jj[k] < nb && ii[k] < nb) part[((int64_t)blockIdx.x * 6 + st) * nb2 + (int64_t)jj[k] * nb + ii[k]] = acc[k];
}

// Pass 3: Rv[s][target][p] = sum_n sum_c h[n][s][target][c] V[n][c][p].
__global__ __launch_bounds__(256) void tc_vlin_kernel(const double* __restrict__ V, const double* __restrict__ hvec,
                                                      int64_t n1, int64_t npair, double* __restrict__ part) {
  const int64_t per = (n1 + gridDim.x - 1) / gridDim.x;
  const int64_t p0 = (int64_t)blockIdx.x * per;
  const int64_t p1 = p0 + per < n1 ? p0 + per : n1;
  for (int64_t p = threadIdx.x; p < npair; p += 256) {
    double acc[4] = {0, 0, 0, 0};
    for (int64_t n = p0; n < p1; ++n) {
      const double v0 = V[(n * 3) * npair + p], v1 = V[(n * 3 + 1) * npair + p], v2 = V[(n * 3 + 2) * npair + p];
      const double* h = hvec + n * 12;
      for (int k = 0; k < 4; ++k) acc[k] += v0 * h[3 * k] + v1 * h[3 * k + 1] + v2 * h[3 * k + 2];
    }
    for (int k = 0; k < 4; ++k) part[((int64_t)blockIdx.x * 4 + k) * npair + p] = acc[k];
  }
}

__global__ __launch_bounds__(256) void tc_reduce(const double* __restrict__ part, int nparts, int64_t len,
                                                 double* __restrict__ out) {
  __shared__ double red[256];
  const int64_t e = blockIdx.x;
  double s = 0;
  for (int b = threadIdx.x; b < nparts; b += 256) s += part[(int64_t)b * len + e];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[e] = red[0];
}

double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

inline int64_t pk(int j, int i) { return j <= i ? (int64_t)i * (i + 1) / 2 + j : (int64_t)j * (j + 1) / 2 + i; }

}  // namespace tcs
}  // namespace tc

using namespace tc;
using namespace tc::tcs;

extern "C" {

/* One TC Fock build on the device: from D (2 x nb x nb) and the occupied
 * orbitals C (2 x nb x nmax, column b of spin s valid for b < nocc_s) to
 * G2 and G3 (each 2 x nb x nb, host), before the one-electron mask. */
int tcscf_tc_fock(int nup, int ndown, int nb, const double* A1, const double* dA1, const double* w1, int64_t n1,
                  const double* Fs, const double* Fo, const double* V, const tcscf_jastrow* jp, const double* D_host,
                  const double* C_host, double* G2_host, double* G3_host, void* stream) {
  TC_CHECK(A1 && dA1 && w1 && Fs && Fo && V && jp && D_host && C_host && G2_host && G3_host, "tcscf_tc_fock: null");
  TC_CHECK(nb > 0 && nb <= 64, "tcscf_tc_fock: 1 <= nb <= 64");
  TC_CHECK(nup >= 0 && ndown >= 0 && nup <= nb && ndown <= nb, "tcscf_tc_fock: bad electron counts");
  TC_CHECK(nup <= kMaxOcc && ndown <= kMaxOcc, "tcscf_tc_fock: at most 8 occupied orbitals per spin");
  hipStream_t st = (hipStream_t)stream;
  const int nmax = std::max(1, std::max(nup, ndown));
  const int nrow = kFixedRows + 3 * nmax;
  const int64_t npair = (int64_t)nb * (nb + 1) / 2, nb2 = (int64_t)nb * nb;
  // device scratch: D, C, rows, hvec, partials, results
  const size_t nD = 2 * nb2, nC = 2 * (size_t)nb * nmax, nRows = (size_t)n1 * 2 * nrow * 2 * nb, nH = (size_t)n1 * 12;
  const size_t nPartG = (size_t)kBlocks * 6 * nb2, nPartV = (size_t)kBlocks * 4 * npair, nOut = 6 * nb2 + 4 * npair;
  int rc = 0;
  double* ws = (double*)scratch(kSlotSweep, sizeof(double) * (nD + nC + nRows + nH + nPartG + nPartV + nOut), &rc);
  if (rc) return rc;
  double* dD = ws;
  double* dC = dD + nD;
  double* dRows = dC + nC;
  double* dH = dRows + nRows;
  double* dPG = dH + nH;
  double* dPV = dPG + nPartG;
  double* dOut = dPV + nPartV;
  TC_HIP(hipMemcpyAsync(dD, D_host, sizeof(double) * nD, hipMemcpyHostToDevice, st));
  TC_HIP(hipMemcpyAsync(dC, C_host, sizeof(double) * nC, hipMemcpyHostToDevice, st));
  PointArgs a{Fs, Fo, V, A1, dA1, w1, dD, dC, jp->c_same, jp->c_oppo, n1, npair, nb, nmax, nup, ndown, nrow, dRows, dH};
  const int64_t per_wave = 5 * npair + nb + 3 * nb + 2 * nb + 6 * nb + 6 * nb + 6 * nb + 6 * (int64_t)nb * nmax;
  const size_t sh1 = sizeof(double) * (2 * nb2 + 2 * (int64_t)nb * nmax + 4 * per_wave);
  TC_CHECK(sh1 <= 160 * 1024, "tcscf_tc_fock: nb too large for the LDS staging");
  if (sh1 > 64 * 1024) TC_HIP(hipFuncSetAttribute((const void*)tc_point_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh1));
  hipLaunchKernelGGL(tc_point_kernel, dim3(kBlocks), dim3(256), sh1, st, a);
  TC_HIP(hipGetLastError());
  const size_t sh2 = sizeof(double) * 4 * (nrow - 5) * 2 * nb;
  TC_CHECK(sh2 <= 64 * 1024, "tcscf_tc_fock: too many occupied orbitals for the rank kernel");
  if (nb <= 32)
    hipLaunchKernelGGL(tc_rank_kernel<32>, dim3(kBlocks, 6), dim3(256), sh2, st, dRows, n1, nb, nrow, dPG);
  else
    hipLaunchKernelGGL(tc_rank_kernel<64>, dim3(kBlocks, 6), dim3(256), sh2, st, dRows, n1, nb, nrow, dPG);
  TC_HIP(hipGetLastError());
  hipLaunchKernelGGL(tc_vlin_kernel, dim3(kBlocks), dim3(256), 0, st, V, dH, n1, npair, dPV);
  TC_HIP(hipGetLastError());
  hipLaunchKernelGGL(tc_reduce, dim3((unsigned)(6 * nb2)), dim3(256), 0, st, dPG, kBlocks, 6 * nb2, dOut);
  TC_HIP(hipGetLastError());
  hipLaunchKernelGGL(tc_reduce, dim3((unsigned)(4 * npair)), dim3(256), 0, st, dPV, kBlocks, 4 * npair, dOut + 6 * nb2);
  TC_HIP(hipGetLastError());
  std::vector<double> out(nOut);
  TC_HIP(hipMemcpyAsync(out.data(), dOut, sizeof(double) * nOut, hipMemcpyDeviceToHost, st));
  TC_HIP(hipStreamSynchronize(st));
  // G2 = scalar/2 + (vec/2 - vec^T/2), G3 = scalar3/2 (HartreeFock.cpp:1209-1222)
  for (int s = 0; s < 2; ++s) {
    const double* Sc = out.data() + (s * 3 + 0) * nb2;
    const double* Vc = out.data() + (s * 3 + 1) * nb2;
    const double* S3 = out.data() + (s * 3 + 2) * nb2;
    const double* Rvec = out.data() + 6 * nb2 + (s * 2 + 0) * npair;
    const double* Rg3 = out.data() + 6 * nb2 + (s * 2 + 1) * npair;
    for (int j = 0; j < nb; ++j)
      for (int i = 0; i < nb; ++i) {
        const double vji = Vc[j * nb + i] + Rvec[pk(j, i)];
        const double vij = Vc[i * nb + j] + Rvec[pk(i, j)];
        G2_host[s * nb2 + j * nb + i] = Sc[j * nb + i] / 2 + (vji / 2 - vij / 2);
        G3_host[s * nb2 + j * nb + i] = (S3[j * nb + i] + Rg3[pk(j, i)]) / 2;
      }
  }
  return 0;
}

/* TCHartreeFock::compute (src/HartreeFock.cpp:1007-1362), real basis.
 * Fs, Fo: packed n1 x npair (FP64); V: packed n1 x 3 x npair for c = 1;
 * A1: n1 x nb; dA1: n1 x nb x 3.  D_out 2 x nb x nb, pseudo_out nup + ndown. */
int tcscf_tc_scf(int nup, int ndown, int nb, const double* S, const double* H, const double* A1, const double* dA1,
                 const double* w1, int64_t n1, const double* Fs, const double* Fo, const double* V,
                 const tcscf_jastrow* jp, const tcscf_scf_opts* o, tcscf_scf_result* res, double* trace,
                 double* D_out, double* pseudo_out, void* stream) {
  TC_CHECK(S && H && o && res && o->max_iter > 0, "tcscf_tc_scf: bad arguments");
  const int64_t nb2 = (int64_t)nb * nb;
  const int ne[2] = {nup, ndown};
  const int nmax = std::max(1, std::max(nup, ndown));
  std::vector<double> D(2 * nb2, 0.0), C(2 * (size_t)nb * nmax, 0.0), G2(2 * nb2), G3(2 * nb2), Fock(2 * nb2);
  std::vector<double> pseudo(nup + ndown, 0.0);
  double energy = std::numeric_limits<double>::max();
  res->status = TCSCF_ENOTCONV;
  res->energy = std::numeric_limits<double>::max();
  res->iterations = 0;
  res->t_sweep_ms = res->t_solve_ms = 0;
  for (int it = 0; it < o->max_iter; ++it) {
    double t0 = now_ms();
    int rc = tcscf_tc_fock(nup, ndown, nb, A1, dA1, w1, n1, Fs, Fo, V, jp, D.data(), C.data(), G2.data(), G3.data(),
                           stream);
    if (rc) return rc;
    double t1 = now_ms();
    res->t_sweep_ms += t1 - t0;
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
      res->t_solve_ms += now_ms() - t1;
      break;
    }
    energy = e;
    for (int s = 0; s < 2; ++s) {
      std::vector<std::complex<double>> ev, X;
      if (!dense::gen_nonsym_eig(nb, Fock.data() + s * nb2, S, ev, X)) {
        res->status = TCSCF_ESOLVER;
        set_error("TC SCF: generalized eigen-solve failed");
        return TCSCF_ESOLVER;
      }
      std::vector<int> idx(nb);
      std::iota(idx.begin(), idx.end(), 0);
      std::sort(idx.begin(), idx.end(), [&](int a, int b) { return ev[a].real() < ev[b].real(); });
      const int no = ne[s];
      std::vector<std::complex<double>> vec((size_t)no * nb);
      for (int b = 0; b < no; ++b) {
        pseudo[s * nup + b] = ev[idx[b]].real();
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
    res->t_solve_ms += now_ms() - t1;
  }
  if (D_out) std::copy(D.begin(), D.end(), D_out);
  if (pseudo_out) std::copy(pseudo.begin(), pseudo.end(), pseudo_out);
  if (res->status != TCSCF_OK) set_error("TC SCF: did not converge");
  return res->status;
}

}  // extern "C"
