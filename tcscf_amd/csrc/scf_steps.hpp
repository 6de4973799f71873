// The SCF drivers split into reusable pieces: a device plan per instance (its
// own workspace, asynchronous enqueue of one Fock build between pinned host
// buffers) and the host half of one iteration.  The single-instance C entry
// points and the batched driver (one graph of every instance's Fock build per
// iteration) share them.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <limits>
#include <vector>

#include "common.hpp"
#include "dense.hpp"

namespace tc {

// Page-locked host staging for the per-iteration uploads and downloads
// (hipHostMalloc with the given flags); p is null when the allocation failed.
struct Pinned {
  double* p = nullptr;
  explicit Pinned(size_t n, unsigned flags = hipHostMallocDefault) {
    if (n && hipHostMalloc((void**)&p, n * sizeof(double), flags) != hipSuccess) p = nullptr;
  }
  ~Pinned() {
    if (p) (void)hipHostFree(p);
  }
  Pinned(const Pinned&) = delete;
  Pinned& operator=(const Pinned&) = delete;
};

/* ---------------------------------------------------------------- RCS / UOS */
struct HfPlan {
  int nspin, nb;
  int64_t n1, npair;
  const void* F;
  int f32;
  const double* A;
  const double* w;
  double* dD;     // nspin x nb x nb
  double* dPart;  // per-workgroup partial G
  double* dG;     // nspin x nb x nb
  int nbp;     // lane group per point: 16 / 32 (column kernel) or 64
  int blocks;  // sweep workgroups (= partials)
  size_t shmem;
};
size_t hf_ws_doubles(int nspin, int nb);
// Checks sizes, points the plan at ws (hf_ws_doubles of it) and opts the sweep
// kernels in to their LDS size.
int hf_prepare(HfPlan& p, int nspin, int nb, int64_t n1, const void* F, int f32, const double* A, const double* w,
               double* ws);
// D (host, nspin x nb x nb) -> G (host), asynchronous on st.  With G_mapped
// (the device address of G_host, coherent page-locked memory) the partial-sum
// kernel writes G there directly and no copy follows.
int hf_enqueue(const HfPlan& p, const double* D_host, double* G_host, hipStream_t st, double* G_mapped = nullptr);

// Host half of one RCS / UOS iteration (src/HartreeFock.cpp:785-1003): mask,
// Fock, energy from the current density, convergence test, optional DIIS,
// generalized eigen-solve and new density.
struct HfIter {
  int nspin, nb;
  int nocc[2];
  const double* S;
  const double* H;
  const tcscf_scf_opts* o;
  std::vector<double> D, Fock, evals, C, err;
  std::vector<double> evecs;  // nspin x nb x nb: [s][j][a], eigenvector a of the last solve
  dense::Diis diis;
  double energy = std::numeric_limits<double>::max();
  HfIter(int nspin_, const int* nocc_, int nb_, const double* S_, const double* H_, const tcscf_scf_opts* o_);
  // G: nspin x nb x nb from the device (modified: the mask is applied in place).
  // Returns 0, or a negative status; *done is set when converged (res->energy set).
  int step(double* G, int it, tcscf_scf_result* res, double* trace, bool* done);
};

/* ------------------------------------------------------------------------ TC */
struct TcPlan {
  int nup, ndown, nb, nmax, nvec;
  int64_t n1, npair;
  const double *A1, *dA1, *w1;
  const void *Fs, *Fo, *V, *Vo;  // element type: FP32 when f32, else FP64
  int f32;                       // FP32 storage of F / V (per-point arithmetic in FP64)
  int gv, nwt;                   // V / Vo independent (V_same, V_oppo): gv = 1, nwt = 2 W tensors
  double c_same, c_oppo;         // scales of one c = 1 tensor V (1, 1 when gv)
  double *dD, *dC, *dVec, *dAl, *dH, *dPG, *dPV, *dOut;
  // point-kernel launch, index 0 for two spins, 1 for the restricted build:
  // waves per workgroup, V rows staged in LDS, LDS bytes, workgroups
  int wpw[2], stage[2];
  size_t sh1[2];
  int64_t pblocks[2];
  // column point kernel (nb <= 32, D = C C^T): usable, LDS bytes, column
  // count of its instance, dot-product entries; dCc the packed orbitals
  int col[2], col_noc[2], col_nent[2];
  size_t col_sh[2];
  double* dCc;
  int ustart[7];  // rank workgroups per target (prefix offsets)
};
size_t tc_ws_doubles(int nup, int ndown, int nb, int64_t n1, bool general_v);
size_t tc_out_doubles(int nb);  // the per-build result block copied back to the host
// Vo non-null: V and Vo are the reference's independent VjiSame / VjiOppo (jp
// unused); else V is the c = 1 tensor and jp gives c_same / c_oppo.
int tc_prepare(TcPlan& p, int nup, int ndown, int nb, const double* A1, const double* dA1, const double* w1,
               int64_t n1, const void* Fs, const void* Fo, const void* V, const void* Vo, int f_dtype,
               const tcscf_jastrow* jp, double* ws);
// The point kernel's launch (stage, wpw, sh1, pblocks) from nb, nmax, f32, gv
// and n1; false when the LDS cannot hold one wave.
bool tc_launch_shape(TcPlan& p);
// Closed shell whose two spins hold bitwise identical densities and orbitals:
// the spin-1 Fock terms are then the spin-0 numbers, formed once.
bool tc_restricted(const TcPlan& p, const double* D, const double* C);
// D (2 x nb x nb) and C (2 x nb x nmax) host -> out (host, tc_out_doubles), async;
// restricted (tc_restricted of D and C) forms the spin-0 blocks only.
int tc_enqueue(const TcPlan& p, const double* D_host, const double* C_host, double* out_host, bool restricted,
               hipStream_t st);
// out -> G2, G3 (2 x nb x nb each, before the mask): G2 = scalar/2 + (vec/2 - vec^T/2),
// G3 = scalar3/2 (src/HartreeFock.cpp:1209-1222); restricted copies spin 0 to spin 1.
void tc_assemble(const TcPlan& p, const double* out, bool restricted, double* G2, double* G3);

// Host half of one TC iteration (src/HartreeFock.cpp:1007-1362).
struct TcIter {
  int nb, ne[2], nmax;
  const double* S;
  const double* H;
  const tcscf_scf_opts* o;
  std::vector<double> D, C, Fock, pseudo, err;
  dense::Diis diis;
  double energy = std::numeric_limits<double>::max();
  TcIter(int nup, int ndown, int nb_, const double* S_, const double* H_, const tcscf_scf_opts* o_);
  int step(double* G2, double* G3, int it, tcscf_scf_result* res, double* trace, bool* done);
};

}  // namespace tc
