// Per-iteration Fock build (HBM-bound sweeps over the stored packed F) and the
// RCS / UOS SCF drivers.  Each SCF iteration: upload D, one sweep kernel over
// the r1 grid, a deterministic reduction of per-workgroup partials, download
// G (nspin x nb x nb), then mask, Fock, energy and the FP64 generalized
// eigen-solve on the host.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <limits>
#include <vector>

#include "common.hpp"
#include "dense.hpp"
#include "scf_steps.hpp"

namespace tc {

constexpr int kSweepBlocks = 512;
constexpr int kGPerThread = 32;  // NSPIN * nb * nb <= 256 * 32

__device__ inline int64_t pidx(int j, int i) { return j <= i ? (int64_t)i * (i + 1) / 2 + j : (int64_t)j * (j + 1) / 2 + i; }

// G_s[j][i] = sum_n w_n A[n,j] (c_self A[n,i] Faa_s[n] + c_other A[n,i] Faa_sbar[n] - Fai_s[n,i])
//   Faa_s[n]  = sum_{j,i} F[n,j,i] D_s[i,j]                   (calculateIaa, HartreeFock.cpp:183-213)
//   Fai_s[n,i] = sum_j (sum_k D_s[k,j] A[n,k]) F[n,j,i]       (calculateFai, HartreeFock.cpp:144-179)
// RCS (NSPIN = 1): c_self = 2, c_other = 0   (HartreeFock.cpp:828-842)
// UOS (NSPIN = 2): c_self = c_other = 1      (HartreeFock.cpp:940-955)
// D is symmetric (C C^T), so D[i][j] = D[j][i].
// Lane layout: a wave works on PPW = 64 / NBP points at once, lane = q * NBP + i
// (point q of the wave, basis index i < nb <= NBP).  The workgroup stages D in
// LDS once, walks its r1 range in chunks of CH points, stages each point's
// packed F row and A row, forms Y_s[n,i] = w_n (...) for the chunk, then folds
// the chunk into register-held partial G with a rank-CH update.
template <class TF, int NSPIN, int NBP>
__global__ __launch_bounds__(256) void hf_sweep_kernel(const TF* __restrict__ F, int64_t npair,
                                                       const double* __restrict__ A, const double* __restrict__ w,
                                                       int64_t n1, int nb, const double* __restrict__ D,
                                                       double c_self, double c_other, double* __restrict__ Gpart) {
  constexpr int PPW = 64 / NBP;  // points per wave at a time
  constexpr int CH = NBP == 64 ? 16 : 32;  // points per chunk
  constexpr int KG = NSPIN * NBP * NBP / 256;  // G elements per thread
  extern __shared__ double smem[];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int q = lane / NBP, i = lane % NBP;
  const int64_t nb2 = (int64_t)nb * nb;
  double* Ds = smem;                                   // [NSPIN][nb][nb]
  double* Frow = Ds + NSPIN * nb2;                     // [4][PPW][npair]
  double* Abuf = Frow + 4 * PPW * npair;               // [CH][nb]
  double* Ybuf = Abuf + CH * nb;                       // [CH][NSPIN][nb]
  double* Xb = Ybuf + CH * NSPIN * nb;                 // [4][PPW][NSPIN][nb]
  for (int64_t e = tid; e < NSPIN * nb2; e += 256) Ds[e] = D[e];
  const int64_t per = (n1 + gridDim.x - 1) / gridDim.x;
  const int64_t p0 = (int64_t)blockIdx.x * per;
  const int64_t p1 = p0 + per < n1 ? p0 + per : n1;
  double gacc[KG];
#pragma unroll
  for (int t = 0; t < KG; ++t) gacc[t] = 0;
  const int tri_i = i * (i + 1) / 2;
  // F rows of a wave's step (points pb .. pb + PPW - 1 of chunk c0) travel
  // through registers: the next step's rows are loaded while this step computes.
  // (nb <= 32; the NBP = 64 build keeps synchronous staging, which needs no
  // extra registers)
  constexpr bool PF = NBP == 32;
  constexpr int NL = PF ? (PPW * (NBP * (NBP + 1) / 2) + 63) / 64 : 1;
  const int64_t rowlen = PPW * npair;
  TF pre[NL];
  double apre = 0;  // this lane's A value of the prefetched step
  auto load_rows = [&](int64_t cc0, int ccn, int cpb) {
    if (cpb + q < ccn && i < nb) apre = A[(cc0 + cpb + q) * nb + i];
#pragma unroll
    for (int t = 0; t < NL; ++t) {
      const int64_t e = lane + 64 * t;
      const int qq = e >= npair ? 1 : 0;
      if (e < rowlen && cpb + qq < ccn) pre[t] = F[(cc0 + cpb) * npair + e];
    }
  };
  auto chunk_len = [&](int64_t cc0) { return (int)((cc0 + CH < p1) ? CH : p1 - cc0); };
  if (PF && p0 < p1 && wave * PPW < chunk_len(p0)) load_rows(p0, chunk_len(p0), wave * PPW);
  __syncthreads();

  for (int64_t c0 = p0; c0 < p1; c0 += CH) {
    const int cn = chunk_len(c0);
    for (int pb = wave * PPW; pb < cn; pb += 4 * PPW) {
      // F rows of points pb .. pb+PPW-1 from registers into LDS, then the
      // next step's rows into registers
      double* fr = Frow + (int64_t)wave * PPW * npair;
      double acur = 0;
      if constexpr (PF) {
        acur = apre;
#pragma unroll
        for (int t = 0; t < NL; ++t) {
          const int64_t e = lane + 64 * t;
          const int qq = e >= npair ? 1 : 0;
          if (e < rowlen && pb + qq < cn) fr[e] = (double)pre[t];
        }
        if (pb + 4 * PPW < cn) load_rows(c0, cn, pb + 4 * PPW);
        else if (c0 + CH < p1 && wave * PPW < chunk_len(c0 + CH)) load_rows(c0 + CH, chunk_len(c0 + CH), wave * PPW);
      } else {
        for (int64_t e = lane; e < rowlen; e += 64)
          if (pb + (e >= npair ? 1 : 0) < cn) fr[e] = (double)F[(c0 + pb) * npair + e];
      }
      const int pt = pb + q;
      const bool live = pt < cn && i < nb;
      const int64_t n = c0 + pt;
      double* ab = Abuf + (int64_t)(live ? pt : 0) * nb;
      if (live) ab[i] = PF ? acur : A[n * nb + i];
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
      double* xb = Xb + ((int64_t)wave * PPW + q) * NSPIN * nb;
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
        const double* f = fr + q * npair;
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

// Deterministic sum of per-workgroup partials, one workgroup per element:
// out[e] = sum_b part[b][e] (fixed tree order).
__global__ __launch_bounds__(256) void reduce_partials(const double* __restrict__ part, int nparts, int64_t len,
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
  p.nbp = nb <= 32 ? 32 : 64;
  const int ppw = 64 / p.nbp, ch = p.nbp == 64 ? 16 : 32;
  p.shmem = sizeof(double) * (nspin * nb2 + 4 * ppw * p.npair + ch * nb * (1 + nspin) + 4 * ppw * nspin * nb);
  TC_CHECK(p.shmem <= (size_t)kMaxLdsBytes, "SCF sweep: nb too large for the LDS staging");
  if (p.shmem > 64 * 1024) {
    // opt in to more than 64 KiB of dynamic LDS (gfx950 has 160 KiB per CU)
    const void* fns[8] = {(const void*)hf_sweep_kernel<float, 1, 32>,  (const void*)hf_sweep_kernel<float, 1, 64>,
                          (const void*)hf_sweep_kernel<double, 1, 32>, (const void*)hf_sweep_kernel<double, 1, 64>,
                          (const void*)hf_sweep_kernel<float, 2, 32>,  (const void*)hf_sweep_kernel<float, 2, 64>,
                          (const void*)hf_sweep_kernel<double, 2, 32>, (const void*)hf_sweep_kernel<double, 2, 64>};
    for (const void* f : fns) TC_HIP(lds_opt_in(f));
  }
  return 0;
}

int hf_enqueue(const HfPlan& p, const double* D_host, double* G_host, hipStream_t st) {
  const int64_t nb2 = (int64_t)p.nb * p.nb;
  TC_HIP(hipMemcpyAsync(p.dD, D_host, sizeof(double) * p.nspin * nb2, hipMemcpyHostToDevice, st));
  const double cs = p.nspin == 1 ? 2.0 : 1.0, co = p.nspin == 1 ? 0.0 : 1.0;
  dim3 grid(kSweepBlocks), block(256);
#define TC_SWEEP(TF, NS, NBP)                                                                                   \
  hipLaunchKernelGGL((hf_sweep_kernel<TF, NS, NBP>), grid, block, p.shmem, st, (const TF*)p.F, p.npair, p.A, p.w, \
                     p.n1, p.nb, p.dD, cs, co, p.dPart)
  if (p.nspin == 1) {
    if (p.f32) { if (p.nbp == 32) TC_SWEEP(float, 1, 32); else TC_SWEEP(float, 1, 64); }
    else { if (p.nbp == 32) TC_SWEEP(double, 1, 32); else TC_SWEEP(double, 1, 64); }
  } else {
    if (p.f32) { if (p.nbp == 32) TC_SWEEP(float, 2, 32); else TC_SWEEP(float, 2, 64); }
    else { if (p.nbp == 32) TC_SWEEP(double, 2, 32); else TC_SWEEP(double, 2, 64); }
  }
#undef TC_SWEEP
  TC_HIP(hipGetLastError());
  const int64_t len = p.nspin * nb2;
  hipLaunchKernelGGL(reduce_partials, dim3((unsigned)len), dim3(256), 0, st, p.dPart, kSweepBlocks, len, p.dG);
  TC_HIP(hipGetLastError());
  TC_HIP(hipMemcpyAsync(G_host, p.dG, sizeof(double) * len, hipMemcpyDeviceToHost, st));
  return 0;
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
  const int64_t nb2 = (int64_t)nb * nb;
  int rc = 0;
  double* ws = (double*)scratch(kSlotSweep, sizeof(double) * hf_ws_doubles(nspin, nb), &rc);
  if (rc) return rc;
  HfPlan plan;
  if ((rc = hf_prepare(plan, nspin, nb, n1, F, f_dtype == TCSCF_F32, A1, w1, ws))) return rc;
  HfIter iter(nspin, nocc, nb, S, H, o);
  // D up and G down through page-locked buffers (pageable vectors if that
  // allocation fails)
  Pinned pin(2 * (size_t)nspin * nb2);
  std::vector<double> Dv, Gv;
  if (!pin.p) {
    Dv.resize(nspin * nb2);
    Gv.resize(nspin * nb2);
  }
  double* Dh = pin.p ? pin.p : Dv.data();
  double* Gh = pin.p ? pin.p + nspin * nb2 : Gv.data();
  res->status = TCSCF_ENOTCONV;
  res->energy = std::numeric_limits<double>::max();
  res->iterations = 0;
  res->t_sweep_ms = res->t_solve_ms = 0;
  for (int it = 0; it < o->max_iter; ++it) {
    double t0 = now_ms();
    std::copy(iter.D.begin(), iter.D.end(), Dh);
    if ((rc = hf_enqueue(plan, Dh, Gh, st))) return rc;
    TC_HIP(hipStreamSynchronize(st));
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
