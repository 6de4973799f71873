// The SCF drivers split into reusable pieces: a device plan per instance (its
// own workspace, asynchronous enqueue of one Fock build between pinned host
// buffers) and the host half of one iteration.  The single-instance C entry
// points and the batched driver (one graph of every instance's Fock build per
// iteration) share them.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <limits>
#include <mutex>
#include <string>
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

// The per-call resources of one driver kind (0: RCS / UOS, 1: TC) on one
// device and caller stream: the sweep workspace, the coherent page-locked
// buffers the device maps (inputs in, results out), the capture / replay
// stream, and the last per-iteration HIP graph with the bytes it was captured
// from (the plan, the buffer address, the kernel choice).  Every iteration's
// device build replays that graph (kernel nodes only: the data move through
// the mapped buffers by kernels); a solve whose key matches replays it without
// capturing, another key captures anew.  Kept across calls, created on first
// use and never destroyed (the runtime may be gone at exit).  The mutex
// serializes the drivers' calls on one (device, stream); calls on different
// streams have separate resources and may run concurrently.  Without the
// graph the drivers enqueue every iteration directly (the same kernels in the
// same order: the same results).
struct IterContext {
  std::mutex mu;
  hipStream_t cs = nullptr;
  Pinned* buf = nullptr;  // coherent, mapped
  double* host = nullptr;
  double* dev = nullptr;
  size_t cap = 0;
  void* ws = nullptr;  // sweep workspace (device)
  size_t ws_bytes = 0;
  hipGraph_t g = nullptr;
  hipGraphExec_t ex = nullptr;
  std::vector<unsigned char> key;
  int captures = 0;  // captures during the current call

  // opt: tcscf_scf_opts.graph.  TCSCF_SCF_GRAPH=1: graphs for every driver,
  // 0: none; unset: where the graph was measured faster than direct launches.
  // TC (Ne, nb = 30): 298 against 310 us per iteration.  RCS / UOS, whose
  // iteration is three kernels, run 7 us slower per iteration from a graph
  // and enqueue directly.
  static bool enabled(int kind, int opt) {
    if (opt == 1) return true;
    if (opt == 2) return false;
    const char* e = std::getenv("TCSCF_SCF_GRAPH");
    if (e && *e) return std::atoi(e) != 0;
    return kind == 1;
  }
  // n doubles of mapped buffer (host / dev null when that fails), a workspace
  // of ws_need bytes (null with *rc set when that fails), the stream.
  int prepare(size_t n, size_t ws_need) {
    captures = 0;
    if (!cs) TC_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    if (n > cap) {
      drop_graph();
      delete buf;
      host = dev = nullptr;
      cap = 0;
      buf = new Pinned(n, hipHostMallocCoherent | hipHostMallocMapped);
      if (buf->p && hipHostGetDevicePointer((void**)&dev, buf->p, 0) == hipSuccess && dev) {
        host = buf->p;
        cap = n;
      } else {
        dev = nullptr;
      }
    }
    if (ws_need > ws_bytes) {
      drop_graph();
      if (ws) {
        TC_HIP(hipStreamSynchronize(cs));
        TC_HIP(hipFree(ws));
        ws = nullptr;
        ws_bytes = 0;
      }
      const size_t want = std::max(ws_need, (size_t)1 << 20);
      if (hipMalloc(&ws, want) != hipSuccess) {
        ws = nullptr;
        set_error("SCF: hipMalloc of " + std::to_string(want) + " bytes of workspace failed");
        return TCSCF_ENOMEM;
      }
      ws_bytes = want;
    }
    return 0;
  }
  // One iteration: captures enqueue(stream) when the key differs from the
  // last capture's, then replays the graph and waits for it.
  template <class F>
  int run(const std::vector<unsigned char>& k, F&& enqueue) {
    if (!ex || k != key) {
      drop_graph();
      TC_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
      const int rc = enqueue(cs);
      hipGraph_t graph = nullptr;
      const hipError_t e = hipStreamEndCapture(cs, &graph);
      if (rc || e != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        if (rc) return rc;
        TC_HIP(e);
      }
      g = graph;
      TC_HIP(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0));
      key = k;
      ++captures;
    }
    TC_HIP(hipGraphLaunch(ex, cs));
    TC_HIP(hipStreamSynchronize(cs));
    return 0;
  }
  void drop_graph() {
    if (ex) (void)hipGraphExecDestroy(ex);
    if (g) (void)hipGraphDestroy(g);
    ex = nullptr;
    g = nullptr;
    key.clear();
  }
};
// The context of driver kind 0 (RCS / UOS) or 1 (TC) on the current device
// for calls on stream st.
IterContext& iter_context(int kind, hipStream_t st);

// The bytes of a trivially copyable value, appended to a graph key.
template <class T>
inline void key_append(std::vector<unsigned char>& k, const T& v) {
  const unsigned char* b = reinterpret_cast<const unsigned char*>(&v);
  k.insert(k.end(), b, b + sizeof(T));
}

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
// The same build between mapped page-locked buffers (device addresses): D
// uploaded by a kernel, G written by the partial-sum kernel; kernels only.
int hf_enqueue_mapped(const HfPlan& p, const double* D_mapped, double* G_mapped, hipStream_t st);

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
  int rchunk;     // (point, entry) terms per rank workgroup
  int rpts[6];    // points per rank workgroup of each target
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
// Whether the column point kernel serves D and C (host): nb, the storage and
// the occupied columns allow it, and D = C C^T for every spin formed.
bool tc_col_ok(const TcPlan& p, const double* D, const double* C, bool restricted);
// The same build between mapped page-locked buffers (device addresses): D then
// C (2 nb^2 + 2 nb nmax doubles) uploaded by a kernel, the result block
// (tc_out_doubles) written by the reductions; kernels only.  col: tc_col_ok
// of the uploaded D and C.
int tc_enqueue_mapped(const TcPlan& p, const double* DC_mapped, double* out_mapped, bool restricted, bool col,
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
