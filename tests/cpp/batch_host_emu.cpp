// CPU emulation of tcscf_batch_scf for sanitizer runs.  The real batch driver
// (batch.cpp) and the real host halves of the SCF iterations (scf.hip,
// tc_scf.hip compiled host-only, dense.cpp) are linked against host stand-ins
// for the HIP runtime: device memory is host memory, a captured graph is the
// list of its copies, and replaying it plays each instance's Fock build with a
// small CPU model of D (and C) in place of the kernels.  The model is not the
// physics; it gives every instance a contraction that converges in a few
// iterations, so the driver goes through its whole life: captures, restricted
// TC closed shells, instances finishing at different iterations, DIIS, the
// worker threads and the last instance on the calling thread.
//
//   batch_host_emu REPS      runs REPS x (DIIS 0, DIIS 6) batches, rc 0 = all ok
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <random>
#include <string>
#include <vector>

#include "tcscf_hip.h"

namespace tc {
thread_local std::string g_err;
void set_error(const std::string& m) { g_err = m; }
const std::string& last_error() { return g_err; }
void* scratch(int, size_t, int* rc) { *rc = -1; return nullptr; }
}  // namespace tc

// Launch configuration of the kernel being launched (hipLaunchKernelGGL goes
// through push / pop, then hipLaunchKernel), per thread as in the runtime.
struct CallCfg { dim3 grid, block; size_t shmem; hipStream_t st; };
static thread_local CallCfg g_cfg;

extern "C" {
void** __hipRegisterFatBinary(const void*) { static void* m[4]; return m; }
void __hipRegisterFunction(void**, const void*, char*, const char*, unsigned int, void*, void*, void*, void*, int*) {}
void __hipUnregisterFatBinary(void**) {}
hipError_t __hipPushCallConfiguration(dim3 g, dim3 b, size_t sh, hipStream_t st) {
  g_cfg = CallCfg{g, b, sh, st};
  return hipSuccess;
}
hipError_t __hipPopCallConfiguration(dim3* g, dim3* b, size_t* sh, hipStream_t* st) {
  *g = g_cfg.grid;
  *b = g_cfg.block;
  *sh = g_cfg.shmem;
  *st = g_cfg.st;
  return hipSuccess;
}
}

// Per-kernel MaxDynamicSharedMemorySize as the runtime keeps it (process-wide,
// last write wins); every launch's dynamic LDS is checked against it (64 KiB
// when never set), as the device enforces it.
static std::mutex g_attr_mu;
static std::vector<std::pair<const void*, int>> g_lds_cap;
static int g_lds_bad = 0;

struct Op { void* dst; const void* src; size_t n; int kind; };
static bool g_capturing = false;
static std::vector<Op>* g_cap = nullptr;
struct Alloc { char* p; size_t n; int idx; };
static std::vector<Alloc> g_allocs;
static std::mutex g_mu;
static int g_nalloc = 0;

// per allocation index: instance description for the model
struct Model { int kind, nb, nspin, nup, ndown; double eps; std::vector<double> T; };
std::vector<Model> g_models;

static const Alloc* find_alloc(const void* p) {
  for (auto& a : g_allocs)
    if ((const char*)p >= a.p && (const char*)p < a.p + a.n) return &a;
  return nullptr;
}

static double T4(const Model& m, int i, int j, int k, int l) {
  const int nb = m.nb;
  return m.T[((size_t)(i * nb + j) * nb + k) * nb + l];
}

// G / out block from D (and C), written to dst
static void play_model(const Alloc& a, const void* src, void* dst, size_t n) {
  const Model& m = g_models[a.idx];
  const int nb = m.nb;
  const size_t nb2 = (size_t)nb * nb;
  double* out = (double*)dst;
  if (m.kind != TCSCF_SCF_TC) {
    const double* D = (const double*)src - m.nspin * nb2;  // dD precedes dG
    if (n != sizeof(double) * m.nspin * nb2) { std::printf("HF size mismatch\n"); std::abort(); }
    std::vector<double> Dt(nb2, 0.0);
    for (int s = 0; s < m.nspin; ++s)
      for (size_t e = 0; e < nb2; ++e) Dt[e] += D[s * nb2 + e] * (m.nspin == 1 ? 2.0 : 1.0);
    for (int s = 0; s < m.nspin; ++s)
      for (int i = 0; i < nb; ++i)
        for (int j = 0; j < nb; ++j) {
          double g = 0;
          for (int k = 0; k < nb; ++k)
            for (int l = 0; l < nb; ++l) g += T4(m, i, j, k, l) * Dt[k * nb + l] - T4(m, i, k, j, l) * D[s * nb2 + k * nb + l];
          out[s * nb2 + i * nb + j] = m.eps * g;
        }
    return;
  }
  const double* D = (const double*)a.p;  // TC: dD at the workspace start
  const int nmax = std::max(1, std::max(m.nup, m.ndown));
  const double* C = D + 2 * nb2;
  const size_t npair = (size_t)nb * (nb + 1) / 2;
  if (n != sizeof(double) * (6 * nb2 + 2 * npair)) { std::printf("TC size mismatch %zu\n", n); std::abort(); }
  std::vector<double> Dt(nb2, 0.0);
  for (int s = 0; s < 2; ++s)
    for (size_t e = 0; e < nb2; ++e) Dt[e] += D[s * nb2 + e];
  for (int s = 0; s < 2; ++s) {
    double* Sc = out + (s * 3 + 0) * nb2;
    double* Vc = out + (s * 3 + 1) * nb2;
    double* S3 = out + (s * 3 + 2) * nb2;
    for (int i = 0; i < nb; ++i)
      for (int j = 0; j < nb; ++j) {
        double g = 0, v = 0, g3 = 0;
        for (int k = 0; k < nb; ++k)
          for (int l = 0; l < nb; ++l) {
            g += T4(m, i, j, k, l) * Dt[k * nb + l] - T4(m, i, k, j, l) * D[s * nb2 + k * nb + l];
            v += T4(m, i, k, l, j) * D[s * nb2 + k * nb + l] * 0.3;
          }
        for (int b = 0; b < nmax; ++b) g3 += C[(s * nb + i) * nmax + b] * C[(s * nb + j) * nmax + b] * 0.01;
        Sc[i * nb + j] = 2 * m.eps * g;
        Vc[i * nb + j] = m.eps * v;
        S3[i * nb + j] = m.eps * g3;
      }
  }
  double* R = out + 6 * nb2;
  for (size_t e = 0; e < 2 * npair; ++e) R[e] = m.eps * 1e-3 * std::sin(0.37 * e + a.idx);
}

static void run_op(const Op& o) {
  if (o.kind == hipMemcpyHostToDevice || o.kind == hipMemcpyHostToHost) {
    std::memcpy(o.dst, o.src, o.n);
  } else {
    const Alloc* a = find_alloc(o.src);
    if (!a) { std::printf("D2H from unknown memory\n"); std::abort(); }
    play_model(*a, o.src, o.dst, o.n);
  }
}

extern "C" {
hipError_t hipMalloc(void** p, size_t n) {
  std::lock_guard<std::mutex> l(g_mu);
  char* q = (char*)std::malloc(n);
  *p = q;
  g_allocs.push_back({q, n, g_nalloc++});
  return hipSuccess;
}
hipError_t hipFree(void* p) {
  std::lock_guard<std::mutex> l(g_mu);
  for (size_t i = 0; i < g_allocs.size(); ++i)
    if (g_allocs[i].p == p) { g_allocs.erase(g_allocs.begin() + i); break; }
  std::free(p);
  return hipSuccess;
}
hipError_t hipHostMalloc(void** p, size_t n, unsigned int) { *p = std::malloc(n); return hipSuccess; }
hipError_t hipHostFree(void* p) { std::free(p); return hipSuccess; }
hipError_t hipHostGetDevicePointer(void** d, void* h, unsigned int) { *d = h; return hipSuccess; }
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned int) { *s = (hipStream_t)new int(1); return hipSuccess; }
hipError_t hipStreamDestroy(hipStream_t s) { delete (int*)s; return hipSuccess; }
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { *e = (hipEvent_t)new int(2); return hipSuccess; }
hipError_t hipEventDestroy(hipEvent_t e) { delete (int*)e; return hipSuccess; }
hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned int) { return hipSuccess; }
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) {
  g_capturing = true;
  g_cap = new std::vector<Op>();
  return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t* g) {
  g_capturing = false;
  *g = (hipGraph_t)g_cap;
  g_cap = nullptr;
  return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t g, hipGraphNode_t*, char*, size_t) {
  *e = (hipGraphExec_t) new std::vector<Op>(*(std::vector<Op>*)g);
  return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g) { delete (std::vector<Op>*)g; return hipSuccess; }
hipError_t hipGraphExecDestroy(hipGraphExec_t e) { delete (std::vector<Op>*)e; return hipSuccess; }
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t) {
  for (const Op& o : *(std::vector<Op>*)e) run_op(o);
  return hipSuccess;
}
hipError_t hipMemcpyAsync(void* dst, const void* src, size_t n, hipMemcpyKind k, hipStream_t) {
  Op o{dst, src, n, (int)k};
  if (g_capturing) g_cap->push_back(o);
  else run_op(o);
  return hipSuccess;
}
hipError_t hipLaunchKernel(const void* fn, dim3, dim3, void**, size_t shmem, hipStream_t) {
  std::lock_guard<std::mutex> l(g_attr_mu);
  int cap = 64 * 1024;
  for (auto& e : g_lds_cap)
    if (e.first == fn) cap = e.second;
  if (shmem > (size_t)cap) {
    std::printf("launch with %zu bytes of dynamic LDS above the kernel's cap %d\n", shmem, cap);
    ++g_lds_bad;
  }
  return hipSuccess;
}
hipError_t hipFuncSetAttribute(const void* fn, hipFuncAttribute attr, int v) {
  if (attr != hipFuncAttributeMaxDynamicSharedMemorySize) return hipSuccess;
  std::lock_guard<std::mutex> l(g_attr_mu);
  for (auto& e : g_lds_cap)
    if (e.first == fn) {
      e.second = v;
      return hipSuccess;
    }
  g_lds_cap.push_back({fn, v});
  return hipSuccess;
}
hipError_t hipGetDevice(int* d) {
  *d = 0;
  return hipSuccess;
}
hipError_t hipDeviceGetAttribute(int* v, hipDeviceAttribute_t attr, int) {
  *v = attr == hipDeviceAttributeMultiprocessorCount ? 256 : 0;
  return hipSuccess;
}
hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int* n, const void*, int, size_t) {
  *n = 4;
  return hipSuccess;
}
hipError_t hipGetLastError(void) { return hipSuccess; }
const char* hipGetErrorString(hipError_t) { return "mock"; }
}

// ------------------------------------------------------------------ driver
static void sym_rand(std::mt19937_64& r, int nb, std::vector<double>& T, double scale) {
  std::uniform_real_distribution<double> u(-1, 1);
  const size_t n = (size_t)nb * nb * nb * nb;
  std::vector<double> raw(n);
  for (auto& x : raw) x = u(r);
  T.assign(n, 0.0);
  auto at = [&](int i, int j, int k, int l) { return raw[((size_t)(i * nb + j) * nb + k) * nb + l]; };
  for (int i = 0; i < nb; ++i)
    for (int j = 0; j < nb; ++j)
      for (int k = 0; k < nb; ++k)
        for (int l = 0; l < nb; ++l)
          T[((size_t)(i * nb + j) * nb + k) * nb + l] =
              scale * (at(i, j, k, l) + at(j, i, k, l) + at(i, j, l, k) + at(j, i, l, k) + at(k, l, i, j) +
                       at(l, k, i, j) + at(k, l, j, i) + at(l, k, j, i)) / 8;
}

int main(int argc, char** argv) {
  const int reps = argc > 1 ? std::atoi(argv[1]) : 4;
  setenv("TCSCF_BATCH_SCF", "1", 1);
  std::mt19937_64 rng(7);
  struct Case { int kind, nup, ndown, nb; double eps; };
  const std::vector<Case> cases = {{TCSCF_SCF_RCS, 1, 1, 5, 0.05},  {TCSCF_SCF_RCS, 2, 2, 9, 0.04},
                                   {TCSCF_SCF_UOS, 2, 1, 9, 0.05},  {TCSCF_SCF_TC, 2, 2, 11, 0.03},
                                   {TCSCF_SCF_TC, 3, 2, 15, 0.03},  {TCSCF_SCF_TC, 5, 4, 40, 0.02},
                                   {TCSCF_SCF_TC, 4, 3, 36, 0.02}};
  // the last two: open-shell TC Fock builds above 64 KiB of LDS, the larger
  // prepared first (a cap set per plan would then be the smaller one's)
  int bad = 0;
  for (int rep = 0; rep < reps; ++rep)
    for (int diis : {0, 6}) {
      const int n = (int)cases.size();
      std::vector<tcscf_scf_item> items(n);
      std::vector<std::vector<double>> S(n), H(n), tr(n), Dout(n);
      std::vector<double> dummy(4096, 0.5);
      g_models.clear();
      g_nalloc = 0;
      for (int k = 0; k < n; ++k) {
        const Case& c = cases[k];
        const int nb = c.nb;
        S[k].assign(nb * nb, 0.0);
        H[k].assign(nb * nb, 0.0);
        std::uniform_real_distribution<double> u(-1, 1);
        for (int i = 0; i < nb; ++i)
          for (int j = 0; j <= i; ++j) {
            const double s = i == j ? 1.0 : 0.1 * u(rng) / nb;
            S[k][i * nb + j] = S[k][j * nb + i] = s;
            const double h = i == j ? -2.0 + 0.3 * i : 0.2 * u(rng);
            H[k][i * nb + j] = H[k][j * nb + i] = h;
          }
        tr[k].assign(200, 0.0);
        Dout[k].assign((c.kind == TCSCF_SCF_RCS ? 1 : 2) * nb * nb, 0.0);
        tcscf_scf_item& it = items[k];
        std::memset(&it, 0, sizeof(it));
        it.kind = c.kind;
        it.nup = c.nup;
        it.ndown = c.ndown;
        it.nb = nb;
        it.S = S[k].data();
        it.H = H[k].data();
        it.A1_dev = it.dA1_dev = it.w1_dev = dummy.data();
        it.n1 = 64;
        it.F_dev = it.Fs_dev = it.Fo_dev = it.V_dev = dummy.data();
        it.f_dtype = TCSCF_F64;
        it.jastrow.c_same = 0.5;
        it.jastrow.c_oppo = 0.25;
        it.jastrow.a12 = 1.0;
        it.energy_trace = tr[k].data();
        it.D_out = Dout[k].data();
        Model m{c.kind, nb, c.kind == TCSCF_SCF_RCS ? 1 : 2, c.nup, c.ndown, c.eps, {}};
        sym_rand(rng, nb, m.T, 1.0);
        g_models.push_back(m);
      }
      tcscf_scf_opts o{1e-10, 200, 0, 1, diis};
      // the same batch twice: results may not depend on how the host threads ran
      std::vector<tcscf_scf_result> res[2] = {std::vector<tcscf_scf_result>(n), std::vector<tcscf_scf_result>(n)};
      std::vector<std::vector<double>> tr0;
      int builds = 0;
      for (int pass = 0; pass < 2; ++pass) {
        g_nalloc = 0;
        const int rc = tcscf_batch_scf(items.data(), n, &o, res[pass].data(), &builds, nullptr);
        if (pass == 0) tr0 = tr;
        std::printf("rep %d diis %d pass %d rc %d builds %d :", rep, diis, pass, rc, builds);
        for (int k = 0; k < n; ++k)
          std::printf(" [%d %d %.6f]", res[pass][k].status, res[pass][k].iterations, res[pass][k].energy);
        std::printf("\n");
        if (rc || builds < 2) ++bad;
        for (int k = 0; k < n; ++k)
          if (res[pass][k].status != TCSCF_OK) ++bad;
      }
      for (int k = 0; k < n; ++k)
        if (res[0][k].energy != res[1][k].energy || res[0][k].iterations != res[1][k].iterations || tr0[k] != tr[k]) {
          std::printf("instance %d differs between passes\n", k);
          ++bad;
        }
    }
  bad += g_lds_bad;
  std::printf("%d failure(s)\n", bad);
  return bad ? 1 : 0;
}
