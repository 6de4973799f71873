// Per-iteration Fock build (HBM-bound sweeps over the stored packed F) and the
// RCS / UOS SCF drivers.  Each SCF iteration: upload D, one sweep kernel over
// the r1 grid, a deterministic reduction of per-workgroup partials, download
// G (nspin x nb x nb), then mask, Fock, energy and the FP64 generalized
// eigen-solve on the host.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <tuple>
#include <vector>

#include "common.hpp"
#include "dense.hpp"
#include "device_util.hpp"
#include "scf_steps.hpp"

namespace tc {

constexpr int kSweepBlocks = 512;
constexpr int kGPerThread = 32;  // NSPIN * nb * nb <= 256 * 32

// G_s[j][i] = sum_n w_n A[n,j] (c_self A[n,i] Faa_s[n] + c_other A[n,i] Faa_sbar[n] - Fai_s[n,i])
//   Faa_s[n]  = sum_{j,i} F[n,j,i] D_s[i,j]                   (calculateIaa, HartreeFock.cpp:183-213)
//   Fai_s[n,i] = sum_j (sum_k D_s[k,j] A[n,k]) F[n,j,i]       (calculateFai, HartreeFock.cpp:144-179)
// RCS (NSPIN = 1): c_self = 2, c_other = 0   (HartreeFock.cpp:828-842)
// UOS (NSPIN = 2): c_self = c_other = 1      (HartreeFock.cpp:940-955)
// D is symmetric (C C^T), so D[i][j] = D[j][i].
// This kernel serves 32 < nb <= 64 (NBP = 64: one point per wave, lane i).
// The workgroup stages D in LDS once, walks its r1 range in chunks of CH
// points, stages each point's packed F row and A row, forms
// Y_s[n,i] = w_n (...) for the chunk, then folds the chunk into register-held
// partial G with a rank-CH update.
template <class TF, int NSPIN, int NBP>
__global__ __launch_bounds__(256) void hf_sweep_kernel(const TF* __restrict__ F, int64_t npair,
                                                       const double* __restrict__ A, const double* __restrict__ w,
                                                       int64_t n1, int nb, const double* __restrict__ D,
                                                       double c_self, double c_other, double* __restrict__ Gpart) {
  static_assert(NBP == 64, "nb <= 32 runs hf_sweep_col_kernel");
  constexpr int PPW = 1;   // points per wave at a time
  constexpr int CH = 16;   // points per chunk
  constexpr int KG = NSPIN * NBP * NBP / 256;  // G elements per thread
  extern __shared__ double smem[];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int i = lane;
  const int64_t nb2 = (int64_t)nb * nb;
  double* Ds = smem;                                   // [NSPIN][nb][nb]
  double* Frow = Ds + NSPIN * nb2;                     // [4][npair]
  double* Abuf = Frow + 4 * PPW * npair;               // [CH][nb]
  double* Ybuf = Abuf + CH * nb;                       // [CH][NSPIN][nb]
  double* Xb = Ybuf + CH * NSPIN * nb;                 // [4][NSPIN][nb]
  for (int64_t e = tid; e < NSPIN * nb2; e += 256) Ds[e] = D[e];
  const int64_t per = (n1 + gridDim.x - 1) / gridDim.x;
  const int64_t p0 = (int64_t)blockIdx.x * per;
  const int64_t p1 = p0 + per < n1 ? p0 + per : n1;
  double gacc[KG];
#pragma unroll
  for (int t = 0; t < KG; ++t) gacc[t] = 0;
  const int tri_i = i * (i + 1) / 2;
  auto chunk_len = [&](int64_t cc0) { return (int)((cc0 + CH < p1) ? CH : p1 - cc0); };
  __syncthreads();

  for (int64_t c0 = p0; c0 < p1; c0 += CH) {
    const int cn = chunk_len(c0);
    for (int pt = wave; pt < cn; pt += 4) {
      double* fr = Frow + (int64_t)wave * npair;
      for (int64_t e = lane; e < npair; e += 64) fr[e] = (double)F[(c0 + pt) * npair + e];
      const bool live = i < nb;
      const int64_t n = c0 + pt;
      double* ab = Abuf + (int64_t)pt * nb;
      if (live) ab[i] = A[n * nb + i];
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
      double* xb = Xb + (int64_t)wave * NSPIN * nb;
      if (live) {
#pragma unroll
        for (int s = 0; s < NSPIN; ++s) {
          const double* Dd = Ds + s * nb2;
          double x = 0;
          for (int k = 0; k < nb; ++k) x += Dd[k * nb + i] * ab[k];
          xb[s * nb + i] = x;
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
      double faa[NSPIN], fai[NSPIN];
#pragma unroll
      for (int s = 0; s < NSPIN; ++s) faa[s] = fai[s] = 0;
      if (live) {
        const double* f = fr;
        for (int j = 0; j < nb; ++j) {
          const double fv = f[j <= i ? tri_i + j : j * (j + 1) / 2 + i];
#pragma unroll
          for (int s = 0; s < NSPIN; ++s) {
            faa[s] += fv * Ds[s * nb2 + j * nb + i];
            fai[s] += xb[s * nb + j] * fv;
          }
        }
      }
#pragma unroll
      for (int s = 0; s < NSPIN; ++s)
#pragma unroll
        for (int off = NBP / 2; off > 0; off >>= 1) faa[s] += __shfl_xor(faa[s], off, 64);
      if (live) {
        const double wn = w[n];
        double* yb = Ybuf + (int64_t)pt * NSPIN * nb;
#pragma unroll
        for (int s = 0; s < NSPIN; ++s) {
          const double other = NSPIN == 2 ? faa[1 - s] : 0.0;
          yb[s * nb + i] = wn * (ab[i] * (c_self * faa[s] + c_other * other) - fai[s]);
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < KG; ++t) {
      const int64_t e = tid + 256 * t;
      if (e < NSPIN * nb2) {
        const int s = (int)(e / nb2);
        const int64_t ji = e - s * nb2;
        const int j = (int)(ji / nb), ii = (int)(ji - (int64_t)j * nb);
        double acc = 0;
        for (int pt = 0; pt < cn; ++pt) acc += Abuf[pt * nb + j] * Ybuf[((int64_t)pt * NSPIN + s) * nb + ii];
        gacc[t] += acc;
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int t = 0; t < KG; ++t) {
    const int64_t e = tid + 256 * t;
    if (e < NSPIN * nb2) Gpart[(int64_t)blockIdx.x * NSPIN * nb2 + e] = gacc[t];
  }
}

template <class TF, int NBP>
constexpr int hf_col_ld() { return NBP == 32 ? (sizeof(TF) == 8 ? 33 : 34) : 20; }

// The same G for nb <= NBP in {16, 32}, one r1 point per wave step.  Lane
// h * NBP + i of a wave owns basis column i and the row slice
// j in [h J, h J + J) (H = 64 / NBP slices of J = NBP / H rows): it keeps that
// slice of column i of G in registers (RCS: of D too), so the per-point work is
// register FMAs against LDS reads that are either one contiguous row of the
// point's F matrix or a broadcast:
//   X_s[i]   = sum_k D_s[k][i] a[k]                (slice sums added over h)
//   Faa_s    = sum_{j,i} F[j][i] D_s[j][i]         (added over the 64 lanes)
//   Fai_s[i] = sum_j X_s[j] F[j][i]                (slice sums added over h)
//   G_s[j][i] += a[j] w (a[i] (c_self Faa_s + c_other Faa_sbar) - Fai_s[i])
// Each wave streams its points' packed F rows through registers, one step
// ahead, and unpacks them into its own NBP x NBP LDS matrix (zero outside
// nb x nb, so every loop runs to NBP unmasked).  Waves never wait for each
// other until the end, where the block sums its waves' G in a fixed order into
// one partial G.
template <class TF, int NSPIN, int NBP>
__global__ __launch_bounds__(256, NSPIN == 1 ? 3 : 2) void hf_sweep_col_kernel(const TF* __restrict__ F, int64_t npair,
                                                              const double* __restrict__ A,
                                                              const double* __restrict__ w, int64_t n1, int nb,
                                                              const double* __restrict__ D, double c_self,
                                                              double c_other, double* __restrict__ Gpart) {
  constexpr int H = 64 / NBP, J = NBP / H;
  // padded row stride of the unpacked matrix, so that neither the scattered
  // stores of a packed row nor the H row slices read at once share LDS banks
  constexpr int LD = hf_col_ld<TF, NBP>(), FMAT = NBP * LD;
  constexpr int NL = (NBP * (NBP + 1) / 2 + 63) / 64;  // loads per lane per point at nb = NBP
  extern __shared__ double smem[];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // uniform: point indices and row addresses stay scalar
  const int h = lane / NBP, i = lane % NBP, j0 = h * J;
  double* ab = smem + wave * NBP;                  // [4][NBP]: a of the wave's point
  double* xb = smem + 4 * NBP + wave * NSPIN * NBP;  // [4][NSPIN][NBP]: X of the wave's point
  TF* fm = reinterpret_cast<TF*>(smem + 4 * NBP * (1 + NSPIN)) + wave * (FMAT + 2);  // [4][NBP x LD + 2]

  // zero the unpacked matrix once: entries outside nb x nb are never written
  // (element FMAT takes the stores of lanes with no pair)
  for (int e = lane; e < FMAT + 2; e += 64) fm[e] = TF(0);
  double dreg[NSPIN][J];
#pragma unroll
  for (int s = 0; s < NSPIN; ++s)
#pragma unroll
    for (int jj = 0; jj < J; ++jj) {
      const int k = j0 + jj;
      const double d = D[((int64_t)s * nb + (k < nb ? k : 0)) * nb + (i < nb ? i : 0)];
      dreg[s][jj] = (k < nb && i < nb) ? d : 0.0;
    }
  // where element lane + 64 t of a point's packed row goes: pair (jj <= ii) of
  // the packed row ii (ii + 1) / 2 + jj, stored at both mirror positions (two
  // 16-bit offsets; FMAT for lanes past the row)
  uint32_t dst[NL];
#pragma unroll
  for (int t = 0; t < NL; ++t) {
    const int p = lane + 64 * t;
    dst[t] = (uint32_t)FMAT | ((uint32_t)FMAT << 16);
    if (p < npair) {
      int ii = (int)((sqrt(8.0 * p + 1.0) - 1.0) * 0.5);
      while (ii * (ii + 1) / 2 > p) --ii;
      while ((ii + 1) * (ii + 2) / 2 <= p) ++ii;
      const int jj = p - ii * (ii + 1) / 2;
      dst[t] = (uint32_t)(jj * LD + ii) | ((uint32_t)(ii * LD + jj) << 16);
    }
  }
  double g[NSPIN][J];
#pragma unroll
  for (int s = 0; s < NSPIN; ++s)
#pragma unroll
    for (int jj = 0; jj < J; ++jj) g[s][jj] = 0;

  const int64_t gw = (int64_t)blockIdx.x * 4 + wave, nw = (int64_t)gridDim.x * 4;
  TF pre[NL];
  double apre = 0, wpre = 0;
  // through a raw buffer view of the row: scalar base, one lane offset, and
  // elements past the row read as zero (no branches, no clamping)
  const int ic = i < nb ? i : 0;
  auto load = [&](int64_t n) {
    const __amdgpu_buffer_rsrc_t row = buffer_view(F + n * npair, npair);
#pragma unroll
    for (int t = 0; t < NL; ++t) pre[t] = buffer_load<TF>(row, (lane + 64 * t) * (int)sizeof(TF));
    apre = A[n * nb + ic];
    wpre = w[n];
  };
  if (gw < n1) load(gw);
  __syncthreads();
  for (int64_t n = gw; n < n1; n += nw) {
#pragma unroll
    for (int t = 0; t < NL; ++t) {
      fm[dst[t] & 0xffff] = pre[t];
      fm[dst[t] >> 16] = pre[t];
    }
    const double a = i < nb ? apre : 0.0, wn = wpre;
    ab[i] = a;  // the H lanes of column i store the same value
    load(n + nw < n1 ? n + nw : n);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
#pragma unroll
    for (int s = 0; s < NSPIN; ++s) {
      double x = 0;
#pragma unroll
      for (int jj = 0; jj < J; ++jj) x += dreg[s][jj] * ab[j0 + jj];
#pragma unroll
      for (int off = NBP; off < 64; off <<= 1) x += __shfl_xor(x, off, 64);
      xb[s * NBP + i] = x;  // equal in the H lanes of column i
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    double faa[NSPIN], fai[NSPIN];
#pragma unroll
    for (int s = 0; s < NSPIN; ++s) faa[s] = fai[s] = 0;
#pragma unroll
    for (int jj = 0; jj < J; ++jj) {
      const double f = (double)fm[(j0 + jj) * LD + i];
#pragma unroll
      for (int s = 0; s < NSPIN; ++s) {
        faa[s] += f * dreg[s][jj];
        fai[s] += xb[s * NBP + j0 + jj] * f;
      }
    }
#pragma unroll
    for (int s = 0; s < NSPIN; ++s) {
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) faa[s] += __shfl_xor(faa[s], off, 64);
#pragma unroll
      for (int off = NBP; off < 64; off <<= 1) fai[s] += __shfl_xor(fai[s], off, 64);
    }
    double y[NSPIN];
#pragma unroll
    for (int s = 0; s < NSPIN; ++s) {
      const double other = NSPIN == 2 ? faa[1 - s] : 0.0;
      y[s] = wn * (a * (c_self * faa[s] + c_other * other) - fai[s]);
    }
#pragma unroll
    for (int jj = 0; jj < J; ++jj) {
      const double aj = ab[j0 + jj];
#pragma unroll
      for (int s = 0; s < NSPIN; ++s) g[s][jj] += aj * y[s];
    }
  }
  // block sum in wave order, per spin
  __syncthreads();
  double* red = smem;  // [4][NBP][NBP] doubles over the block's LDS
#pragma unroll
  for (int s = 0; s < NSPIN; ++s) {
#pragma unroll
    for (int jj = 0; jj < J; ++jj) red[(wave * NBP + j0 + jj) * NBP + i] = g[s][jj];
    __syncthreads();
    const int64_t nb2 = (int64_t)nb * nb;
    for (int e = tid; e < nb2; e += 256) {
      const int j = (int)(e / nb), c = (int)(e - (int64_t)j * nb);
      Gpart[((int64_t)blockIdx.x * NSPIN + s) * nb2 + e] =
          ((red[j * NBP + c] + red[(NBP + j) * NBP + c]) + red[(2 * NBP + j) * NBP + c]) +
          red[(3 * NBP + j) * NBP + c];
    }
    __syncthreads();
  }
}

// Deterministic sum of per-workgroup partials: out[e] = sum_b part[b][e].
// Workgroup of 1024 (16 waves) per 64 consecutive elements; wave v sums the
// contiguous slice of partials v * per .. (v + 1) * per - 1 for its lanes'
// elements (eight loads in flight at a time), then the 16 slice sums are added
// in wave order.
__global__ __launch_bounds__(1024) void reduce_partials(const double* __restrict__ part, int nparts, int64_t len,
                                                        double* __restrict__ out) {
  __shared__ double red[16][64];
  const int lane = threadIdx.x & 63, v = threadIdx.x >> 6;
  const int64_t e = (int64_t)blockIdx.x * 64 + lane;
  const int per = (nparts + 15) / 16;
  const int b0 = v * per, b1 = b0 + per < nparts ? b0 + per : nparts;
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
  red[v][lane] = s;
  __syncthreads();
  if (v == 0 && e < len) {
    double t = 0;
    for (int u = 0; u < 16; ++u) t += red[u][lane];
    out[e] = t;
  }
}

namespace {

double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

void mask_by(int nb, double* G, const double* H) {
  for (int q = 0; q < nb * nb; ++q)
    if (!(std::fabs(H[q]) > 0)) G[q] = 0;
}

// D[u][v] = sum_{a < nocc} C[u][a] C[v][a]  (getNewDensity, HartreeFock.cpp:20-44)
void density_from(int nb, int nocc, const std::vector<double>& C, double* D) {
  for (int u = 0; u < nb; ++u)
    for (int v = 0; v < nb; ++v) {
      double t = 0;
      for (int a = 0; a < nocc; ++a) t += C[u * nb + a] * C[v * nb + a];
      D[u * nb + v] = t;
    }
}

}  // namespace

IterContext& iter_context(int kind, hipStream_t st) {
  static std::mutex mu;
  static std::map<std::tuple<int, int, hipStream_t>, IterContext*> all;
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::lock_guard<std::mutex> lock(mu);
  IterContext*& c = all[std::make_tuple(dev, kind, st)];
  if (!c) c = new IterContext();  // kept to the end of the process
  return *c;
}

size_t hf_ws_doubles(int nspin, int nb) { return (size_t)nspin * nb * nb * (kSweepBlocks + 2); }

int hf_prepare(HfPlan& p, int nspin, int nb, int64_t n1, const void* F, int f32, const double* A, const double* w,
               double* ws) {
  const int64_t nb2 = (int64_t)nb * nb;
  TC_CHECK(nb <= 64, "SCF sweep: nb > 64 is not supported by the sweep kernel");
  TC_CHECK(nspin * nb2 <= 256 * kGPerThread, "SCF sweep: nb too large for the register-held G partials");
  p.nspin = nspin;
  p.nb = nb;
  p.n1 = n1;
  p.npair = (int64_t)nb * (nb + 1) / 2;
  p.F = F;
  p.f32 = f32;
  p.A = A;
  p.w = w;
  p.dPart = ws;
  p.dD = p.dPart + (int64_t)kSweepBlocks * nspin * nb2;
  p.dG = p.dD + nspin * nb2;
  p.nbp = nb <= 16 ? 16 : nb <= 32 ? 32 : 64;
  if (p.nbp < 64) {
    // column-layout kernel: a and X of each wave's point, the unpacked F
    // matrices; the final block sum reuses the LDS as 4 FP64 matrices
    const int mat = p.nbp * p.nbp;
    const int ld = p.nbp == 32 ? (f32 ? hf_col_ld<float, 32>() : hf_col_ld<double, 32>()) : hf_col_ld<double, 16>();
    const size_t fsz = (size_t)4 * (p.nbp * ld + 2) * (f32 ? sizeof(float) : sizeof(double));
    p.shmem = std::max(sizeof(double) * 4 * p.nbp * (1 + nspin) + fsz, sizeof(double) * 4 * mat);
    p.blocks = (int)std::max<int64_t>(1, std::min<int64_t>(kSweepBlocks, (n1 + 3) / 4));
  } else {
    const int ppw = 64 / p.nbp, ch = 16;
    p.shmem = sizeof(double) * (nspin * nb2 + 4 * ppw * p.npair + ch * nb * (1 + nspin) + 4 * ppw * nspin * nb);
    p.blocks = kSweepBlocks;
  }
  TC_CHECK(p.shmem <= (size_t)kMaxLdsBytes, "SCF sweep: nb too large for the LDS staging");
  if (p.shmem > 64 * 1024) {
    // opt in to more than 64 KiB of dynamic LDS (gfx950 has 160 KiB per CU)
    const void* fns[12] = {
        (const void*)hf_sweep_kernel<float, 1, 64>,      (const void*)hf_sweep_kernel<double, 1, 64>,
        (const void*)hf_sweep_kernel<float, 2, 64>,      (const void*)hf_sweep_kernel<double, 2, 64>,
        (const void*)hf_sweep_col_kernel<float, 1, 32>,  (const void*)hf_sweep_col_kernel<double, 1, 32>,
        (const void*)hf_sweep_col_kernel<float, 2, 32>,  (const void*)hf_sweep_col_kernel<double, 2, 32>,
        (const void*)hf_sweep_col_kernel<float, 1, 16>,  (const void*)hf_sweep_col_kernel<double, 1, 16>,
        (const void*)hf_sweep_col_kernel<float, 2, 16>,  (const void*)hf_sweep_col_kernel<double, 2, 16>};
    for (const void* f : fns) TC_HIP(lds_opt_in(f));
  }
  return 0;
}

namespace {

// The sweep over F and the partial-sum kernel writing G (nspin x nb x nb) to
// G_dev, from the density already in p.dD.
int hf_sweep_reduce(const HfPlan& p, hipStream_t st, double* G_dev) {
  const int64_t nb2 = (int64_t)p.nb * p.nb;
  const double cs = p.nspin == 1 ? 2.0 : 1.0, co = p.nspin == 1 ? 0.0 : 1.0;
  dim3 grid(p.blocks), block(256);
#define TC_SWEEP(K, TF, NS, NBP)                                                                               \
  hipLaunchKernelGGL((K<TF, NS, NBP>), grid, block, p.shmem, st, (const TF*)p.F, p.npair, p.A, p.w, p.n1, p.nb, \
                     p.dD, cs, co, p.dPart)
#define TC_SWEEP_NS(K, NS, NBP) \
  do {                           \
    if (p.f32)                   \
      TC_SWEEP(K, float, NS, NBP); \
    else                         \
      TC_SWEEP(K, double, NS, NBP); \
  } while (0)
#define TC_SWEEP_ALL(K, NBP)                       \
  do {                                             \
    if (p.nspin == 1) TC_SWEEP_NS(K, 1, NBP);      \
    else TC_SWEEP_NS(K, 2, NBP);                   \
  } while (0)
  if (p.nbp == 16) TC_SWEEP_ALL(hf_sweep_col_kernel, 16);
  else if (p.nbp == 32) TC_SWEEP_ALL(hf_sweep_col_kernel, 32);
  else TC_SWEEP_ALL(hf_sweep_kernel, 64);
#undef TC_SWEEP_ALL
#undef TC_SWEEP_NS
#undef TC_SWEEP
  TC_HIP(hipGetLastError());
  const int64_t len = p.nspin * nb2;
  hipLaunchKernelGGL(reduce_partials, dim3((unsigned)((len + 63) / 64)), dim3(1024), 0, st, p.dPart, p.blocks, len,
                     G_dev);
  TC_HIP(hipGetLastError());
  return 0;
}

}  // namespace

int hf_enqueue(const HfPlan& p, const double* D_host, double* G_host, hipStream_t st, double* G_mapped) {
  const int64_t len = (int64_t)p.nspin * p.nb * p.nb;
  TC_HIP(hipMemcpyAsync(p.dD, D_host, sizeof(double) * len, hipMemcpyHostToDevice, st));
  int rc = hf_sweep_reduce(p, st, G_mapped ? G_mapped : p.dG);
  if (rc) return rc;
  if (!G_mapped) TC_HIP(hipMemcpyAsync(G_host, p.dG, sizeof(double) * len, hipMemcpyDeviceToHost, st));
  return 0;
}

int hf_enqueue_mapped(const HfPlan& p, const double* D_mapped, double* G_mapped, hipStream_t st) {
  const int64_t len = (int64_t)p.nspin * p.nb * p.nb;
  hipLaunchKernelGGL(copy_doubles, copy_grid(len), dim3(256), 0, st, D_mapped, p.dD, len);
  TC_HIP(hipGetLastError());
  return hf_sweep_reduce(p, st, G_mapped);
}

HfIter::HfIter(int nspin_, const int* nocc_, int nb_, const double* S_, const double* H_, const tcscf_scf_opts* o_)
    : nspin(nspin_), nb(nb_), nocc{nocc_[0], nspin_ > 1 ? nocc_[1] : 0}, S(S_), H(H_), o(o_),
      D((size_t)nspin_ * nb_ * nb_, 0.0), Fock((size_t)nspin_ * nb_ * nb_), evals((size_t)nspin_ * nb_, 0.0),
      evecs((size_t)nspin_ * nb_ * nb_, 0.0), diis(o_->diis) {}

int HfIter::step(double* G, int it, tcscf_scf_result* res, double* trace, bool* done) {
  const int64_t nb2 = (int64_t)nb * nb;
  *done = false;
  if (o->mask)
    for (int s = 0; s < nspin; ++s) mask_by(nb, G + s * nb2, H);
  // constructSCFOperator (HartreeFock.cpp:48-98); RCS energy x2 (:850)
  double e = 0;
  for (int s = 0; s < nspin; ++s)
    for (int j = 0; j < nb; ++j)
      for (int i = 0; i < nb; ++i) {
        const double g = G[s * nb2 + j * nb + i];
        Fock[s * nb2 + j * nb + i] = H[j * nb + i] + g;
        e += D[s * nb2 + i * nb + j] * (H[j * nb + i] + g / 2);
      }
  if (nspin == 1) e *= 2;
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
    dense::commutator_error(nspin, nb, Fock.data(), D.data(), S, err);
    diis.extrapolate(Fock, err);
  }
  for (double f : Fock)
    if (!std::isfinite(f)) {
      res->status = TCSCF_ESOLVER;
      set_error("SCF: non-finite Fock matrix (F input not finite?)");
      return TCSCF_ESOLVER;
    }
  for (int s = 0; s < nspin; ++s) {
    std::vector<double> ev;
    if (!dense::gen_sym_eig(nb, Fock.data() + s * nb2, S, ev, C)) {
      res->status = TCSCF_ESOLVER;
      set_error("SCF: generalized eigen-solve failed (overlap not positive definite?)");
      return TCSCF_ESOLVER;
    }
    std::copy(ev.begin(), ev.end(), evals.begin() + s * nb);
    std::copy(C.begin(), C.begin() + nb2, evecs.begin() + s * nb2);
    density_from(nb, nocc[s], C, D.data() + s * nb2);
  }
  return 0;
}

namespace {

int hf_scf(int nspin, const int* nocc, int nb, const double* S, const double* H, const double* A1, const double* w1,
           int64_t n1, const void* F, int f_dtype, const tcscf_scf_opts* o, tcscf_scf_result* res, double* trace,
           double* D_out, double* evals_out, tcscf_scf_observables* obs, hipStream_t st) {
  TC_CHECK(S && H && A1 && w1 && F && o && res && nb > 0 && n1 > 0, "SCF: bad arguments");
  TC_CHECK(o->max_iter > 0, "SCF: max_iter must be positive");
  TC_CHECK(o->diis >= 0, "SCF: diis depth must be >= 0");
  TC_CHECK(o->graph >= 0 && o->graph <= 2, "SCF: graph must be 0 (default), 1 (graph) or 2 (direct launches)");
  const int64_t nb2 = (int64_t)nb * nb;
  int rc = 0;
  IterContext& cx = iter_context(0, st);
  std::lock_guard<std::mutex> lock(cx.mu);
  const int64_t len = nspin * nb2;
  if ((rc = cx.prepare(2 * (size_t)len, sizeof(double) * hf_ws_doubles(nspin, nb)))) return rc;
  double* ws = (double*)cx.ws;
  HfPlan plan;
  std::memset(&plan, 0, sizeof plan);  // its bytes key the iteration graph
  if ((rc = hf_prepare(plan, nspin, nb, n1, F, f_dtype == TCSCF_F32, A1, w1, ws))) return rc;
  HfIter iter(nspin, nocc, nb, S, H, o);
  // D up and G down through the device's coherent page-locked buffers the
  // device maps (pageable vectors where that allocation failed).  Every
  // iteration is enqueued on the caller's stream (D copy, sweep, partial sums
  // writing G straight into the mapped buffer); with TCSCF_SCF_GRAPH=1 it
  // replays one HIP graph instead (D upload kernel, sweep, partial sums).
  std::vector<double> Dv, Gv;
  double *Dh, *Gh, *Dmap = nullptr, *Gmap = nullptr;
  if (cx.host) {
    Dh = cx.host;
    Gh = cx.host + len;
    Dmap = cx.dev;
    Gmap = cx.dev + len;
  } else {
    Dv.resize(len);
    Gv.resize(len);
    Dh = Dv.data();
    Gh = Gv.data();
  }
  const bool use_graph = IterContext::enabled(0, o->graph) && Dmap;
  std::vector<unsigned char> key;
  if (use_graph) {
    key_append(key, plan);
    key_append(key, Dmap);
    TC_HIP(hipStreamSynchronize(st));  // F, A, w come from the caller's stream
  }
  res->status = TCSCF_ENOTCONV;
  res->energy = std::numeric_limits<double>::max();
  res->iterations = 0;
  res->t_sweep_ms = res->t_solve_ms = 0;
  res->graph_iterations = res->graph_captures = 0;
  for (int it = 0; it < o->max_iter; ++it) {
    double t0 = now_ms();
    std::copy(iter.D.begin(), iter.D.end(), Dh);
    if (use_graph) {
      if ((rc = cx.run(key, [&](hipStream_t s) { return hf_enqueue_mapped(plan, Dmap, Gmap, s); }))) return rc;
      ++res->graph_iterations;
      res->graph_captures = cx.captures;
    } else {
      if ((rc = hf_enqueue(plan, Dh, Gh, st, Gmap))) return rc;
      TC_HIP(hipStreamSynchronize(st));
    }
    double t1 = now_ms();
    res->t_sweep_ms += t1 - t0;
    bool done = false;
    if ((rc = iter.step(Gh, it, res, trace, &done))) return rc;
    res->t_solve_ms += now_ms() - t1;
    if (done) break;
  }
  if (D_out) std::copy(iter.D.begin(), iter.D.end(), D_out);
  if (evals_out) std::copy(iter.evals.begin(), iter.evals.end(), evals_out);
  if (obs) {
    if (obs->fock) std::copy(iter.Fock.begin(), iter.Fock.end(), obs->fock);
    if (obs->eigenvectors) std::copy(iter.evecs.begin(), iter.evecs.end(), obs->eigenvectors);
  }
  if (res->status != TCSCF_OK) set_error("SCF: did not converge");
  return res->status;
}

}  // namespace
}  // namespace tc

extern "C" {

int tcscf_rcs_scf(int nelec, int nb, const double* S, const double* H, const double* A1, const double* w1, int64_t n1,
                  const void* F, int f_dtype, const tcscf_scf_opts* o, tcscf_scf_result* res, double* trace,
                  double* D_out, double* evals_out, tcscf_scf_observables* obs, void* stream) {
  TC_CHECK(nelec > 0 && nelec % 2 == 0, "RCS: the number of electrons must be even");
  const int nocc[1] = {nelec / 2};
  return tc::hf_scf(1, nocc, nb, S, H, A1, w1, n1, F, f_dtype, o, res, trace, D_out, evals_out, obs,
                    (hipStream_t)stream);
}

int tcscf_uos_scf(int nup, int ndown, int nb, const double* S, const double* H, const double* A1, const double* w1,
                  int64_t n1, const void* F, int f_dtype, const tcscf_scf_opts* o, tcscf_scf_result* res,
                  double* trace, double* D_out, double* evals_out, tcscf_scf_observables* obs, void* stream) {
  TC_CHECK(nup >= 0 && ndown >= 0 && nup <= nb && ndown <= nb, "UOS: bad electron counts");
  const int nocc[2] = {nup, ndown};
  return tc::hf_scf(2, nocc, nb, S, H, A1, w1, n1, F, f_dtype, o, res, trace, D_out, evals_out, obs,
                    (hipStream_t)stream);
}

}  // extern "C"
