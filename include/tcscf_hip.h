/*
 * tcscf_hip.h -- C ABI of the MI355X (gfx950) grid-integral library.
 *
 * The reference (corbett5/tcscf) exposes this path as templated C++ calls
 * (src/HartreeFock.hpp, src/integration/integrateAll.hpp); lambdas cannot cross
 * an ABI, so kernels become an enum plus parameters.  Each entry point names the
 * reference interface it replaces.
 *
 * Conventions
 *   - Return 0 on success, a negative tcscf_status otherwise; tcscf_last_error()
 *     gives the message of the calling thread's last failure.  No exceptions
 *     cross the ABI.
 *   - Pointers documented "dev" are device (HBM) pointers; "host" are host memory.
 *   - Grid points are SoA 3 x n (x row, y row, z row); a 2-D grid of the
 *     reference (QMCGrid<REAL,2>, points in the xz-plane) is passed with y == 0
 *     and scale = 2*pi.  Weights carry every Jacobian (r^2 sin(theta) ...).
 *   - Basis tables are row-major n x nb; gradients n x nb x 3.
 *   - Pair-packed tensors: p(j,i) = i*(i+1)/2 + j for j <= i, npair = nb(nb+1)/2.
 *     F_packed[n1][npair] holds F(r1, j, i) = F(r1, i, j) (real basis).
 *   - `stream` is a hipStream_t (NULL = default stream).  Entry points are
 *     asynchronous unless documented otherwise; calls on one stream are ordered.
 */
#ifndef TCSCF_HIP_H
#define TCSCF_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum tcscf_status {
  TCSCF_OK = 0,
  TCSCF_EINVAL = -1,   /* bad argument / shape (reference: LVARRAY_ERROR_IF_NE abort) */
  TCSCF_EHIP = -2,     /* HIP runtime error */
  TCSCF_ENOTCONV = -3, /* SCF did not converge (reference: ConvergenceError) */
  TCSCF_ESOLVER = -4,  /* eigen-solver failure (reference: LvArray::dense::Error) */
  TCSCF_ENOMEM = -5
};

enum tcscf_kernel {
  TCSCF_K_COULOMB = 0,      /* 1/r12 */
  TCSCF_K_TC_SCALAR = 1,    /* 2/r12 - |grad1 u(r1,r2)|^2 - |grad1 u(r2,r1)|^2 */
  TCSCF_K_JASTROW_GRAD = 2, /* grad1 u(r1,r2), 3 components, for c = 1 */
  TCSCF_K_TC_SMOOTH = 3     /* 1/(r12 + a12)^4, the smooth part of TCSCF_K_TC_SCALAR */
};

enum tcscf_precision {
  TCSCF_FP64 = 0,   /* v_mfma_f64_16x16x4_f64, FP64 kernel generation */
  TCSCF_FP32 = 1,   /* v_mfma_f32_16x16x4_f32, FP64 cross-chunk sums */
  TCSCF_BF16X3 = 2, /* v_mfma_f32_16x16x32_bf16: K = Kh + Kl, P = Ph + Pm;
                       Kh Ph + Kh Pm + Kl Ph, FP64 sums per 32-point step */
  TCSCF_BF16X5 = 3, /* as BF16X3 with P = Ph + Pm + Pl and + Kh Pl + Kl Pm */
  TCSCF_BF16X4 = 4  /* as BF16X5 without Kl Pm; FP64 sums per 64-point pair of
                       steps, each step's products issued smallest first */
};

enum tcscf_dtype { TCSCF_F64 = 0, TCSCF_F32 = 1 };

enum tcscf_radial_map {
  TCSCF_RADIAL_M4 = 0,        /* Treutler-Ahlrichs M4(a, eps) */
  TCSCF_RADIAL_BECKE = 1,     /* r = (1+x)/(1-x) */
  TCSCF_RADIAL_XOVER1MX = 2   /* r = x/(1-x) */
};

/* One Slater-type orbital N r^(n-1) e^(-zeta r) S_lm (real harmonics, l <= 3). */
typedef struct tcscf_sto {
  int n, l, m;
  double zeta;
} tcscf_sto;

/* One function of the reference's atomic bases: g(r) S_lm(x, y, z), with S_lm
 * = r^l Y_lm the real solid harmonic (l <= 3) and g
 *   TCSCF_BASIS_STO       N r^(n-1-l) e^(-alpha r), n >= 1           (SlaterTypeOrbital.hpp:15-113)
 *   TCSCF_BASIS_OCHI      N L_n^(2l+2)(2 alpha r) e^(-alpha r), n >= 0 (OchiBasis.hpp:15-110)
 *   TCSCF_BASIS_HYDROGEN  N L_(n-l-1)^(2l+1)(2 b r) e^(-b r), b = alpha / n, with alpha the
 *                         function's nuclear charge Z, n > l             (HydrogenLikeBasis.hpp:12-61)
 * (L the associated Laguerre polynomials) and the reference's normalizations N
 * (the hydrogen-like one keeps its negative sign). */
enum tcscf_basis_kind { TCSCF_BASIS_STO = 0, TCSCF_BASIS_OCHI = 1, TCSCF_BASIS_HYDROGEN = 2 };
typedef struct tcscf_basis_fn {
  int kind;
  int n, l, m;
  double alpha; /* STO: zeta; Ochi: alpha; hydrogen-like: its Z */
} tcscf_basis_fn;

/* Jastrow u(r12) = c * r12 / (r12 + a12): the r12-only term of the reference's
 * Ochi form (src/jastrowFunctions.hpp:9-115) with S = {(1,0,0)}. */
typedef struct tcscf_jastrow {
  double a12;
  double c_same;
  double c_oppo;
} tcscf_jastrow;

/* The reference's general Ochi Jastrow, jastrowFunctions::Ochi{a, a12, c, S}
 * (src/jastrowFunctions.hpp:9-115):
 *   u(r1, r2) = sum_t c_t (r12/(r12+a12))^i_t (r1/(r1+a))^j_t (r2/(r2+a))^k_t
 * with S(t, :) = (i_t, j_t, k_t) and c(t, sameSpin) = c_same[t] or c_oppo[t]. */
#define TCSCF_OCHI_MAX_TERMS 8
typedef struct tcscf_ochi_jastrow {
  double a;
  double a12;
  int nterms; /* 0 .. TCSCF_OCHI_MAX_TERMS */
  int i[TCSCF_OCHI_MAX_TERMS], j[TCSCF_OCHI_MAX_TERMS], k[TCSCF_OCHI_MAX_TERMS]; /* 0 .. 16 */
  double c_same[TCSCF_OCHI_MAX_TERMS];
  double c_oppo[TCSCF_OCHI_MAX_TERMS];
} tcscf_ochi_jastrow;

typedef struct tcscf_scf_opts {
  double tol;     /* |E_new - E_old| < tol ends the loop (reference: 1e-5) */
  int max_iter;   /* reference: 50 RCS, 100 UOS, 200 TC */
  int mask;       /* respectOneElectronSymmetry: zero G where H == 0 */
  int mgs_mode;   /* TC: 0 = reference MGS (Euclidean norm for i >= 1), 1 = metric */
  int diis;       /* 0: the reference's plain Roothaan loop; k > 0: Pulay DIIS over the
                     last k Fock operators (FP64, host), from the second iteration on.
                     No reference counterpart (the benchmark workloads ask for it);
                     TC needs mgs_mode = 1 with it. */
  int graph;      /* per-iteration HIP graph: 0 the driver's default (TC: a graph; RCS / UOS:
                     direct launches; TCSCF_SCF_GRAPH=1 / 0 overrides it), 1 a graph, 2 direct
                     launches (e.g. for solves running concurrently on several streams) */
} tcscf_scf_opts;

typedef struct tcscf_scf_result {
  double energy;
  int iterations;   /* numberOfConvergenceLoops() */
  int status;       /* TCSCF_OK or TCSCF_ENOTCONV / TCSCF_ESOLVER */
  double t_sweep_ms;  /* time in the per-iteration device sweeps */
  double t_solve_ms;  /* time in the FP64 eigen-solves and density updates */
  int graph_iterations; /* iterations whose device build replayed a HIP graph (0: enqueued directly) */
  int graph_captures;   /* captures of that graph during this call: 0 when a graph kept from an earlier
                           call with the same inputs served, more than 1 when the kernel choice changed */
} tcscf_scf_result;

/* The solver objects' observable members after compute (host buffers, each
 * may be NULL).  fock: fockOperator (src/HartreeFock.hpp:286, :342, :405),
 * nspin x nb x nb, [s][j][i], the Fock operator H + G (TC: H + G2 - G3, after
 * the mask) of the last iteration: at convergence the one built from the
 * converged density, as the reference leaves it.  eigenvectors (RCS / UOS,
 * :288, :344): nspin x nb x nb, [s][j][a] = coefficient of basis function j in
 * eigenvector a (ascending eigenvalues) of the last eigen-solve, the one that
 * gave the final density.  occupied (TC, occupiedOrbitals :408): 2 x nb x
 * max(nup, ndown), [s][j][b], the Gram-Schmidt-orthonormalized occupied
 * orbitals of the last solve (entries b >= nocc_s zero). */
typedef struct tcscf_scf_observables {
  double* fock;
  double* eigenvectors;
  double* occupied;
} tcscf_scf_observables;

/* ---------------------------------------------------------------- misc */
const char* tcscf_last_error(void);
int tcscf_version(void);

/* ---------------------------------------------------------------- host grids */
/* ChebyshevGauss + createGrid(map): src/integration/ChebyshevGauss.hpp:10-48,
 * src/integration/changeOfVariables.hpp:100-162, src/integration/quadrature.hpp:76-102 */
int tcscf_radial_grid(int n, int map, double a, double eps, double* r_host, double* w_host);
/* getLebedevGrid: src/integration/Lebedev.cpp:23-41 ({theta, phi, w}, sum w = 4 pi) */
int tcscf_lebedev_npts(int order);
int tcscf_lebedev(int order, double* theta_host, double* phi_host, double* w_host);
/* TreutlerAhlrichsLebedev packed into the QMCGrid shape
 * (src/integration/TreutlerAhlrichsLebedev.hpp:19-73, integrateAll.hpp:41-64);
 * point k = ia*nrad + ir, weight = w_r * w_ang * r^2.  Returns n or < 0. */
int tcscf_product_grid(int nrad, int map, double a, double eps, int order, double* pts_host, double* w_host);

/* QMCGrid(gridSize) (src/integration/integrateAll.hpp:21-65, QMCCache
 * src/integration/qmcWrapper.hpp:95-141): nshifts randomly shifted copies of a
 * rank-1 lattice (n = first prime >= min_points, Korobov generating vector
 * chosen by the P_2 criterion; the reference takes mppmu/qmc's tables, whose
 * exact points are not reproducible here), mapped by TreutlerAhlrichsM4(0, 0.9)
 * x Azimuthal x Linear(2 pi) with weight *= r^2 sin(theta).  ndim 2 stores
 * (r sin(theta), 0, r cos(theta)).  pts 3 x N SoA; returns N (or < 0). */
int64_t tcscf_qmc_lattice_npts(int64_t min_points, int nshifts);
int64_t tcscf_qmc_lattice_grid(int64_t min_points, int ndim, int nshifts, uint64_t seed, double* pts_host,
                               double* w_host);
/* The point sets of the reference's per-element QMC integrals,
 * sphericalCoordinates{3,5,6}DIntegral (src/integration/qmcWrapper.hpp:146-276):
 * a randomly shifted Korobov lattice in [0,1)^ndim mapped to r = x/(1-x),
 * theta = pi x, phi = 2 pi x per electron (ndim 5: phi2 = 0), weights with the
 * Jacobians, the prefactor (2 pi^2; 4 pi^4 for ndim 5, 6) and 1/N.  pts1 (and
 * pts2 for ndim 5, 6): 3 x N Cartesian SoA on the host; N = nshifts * (first
 * prime >= min_points), returned. */
int64_t tcscf_qmc_spherical_points(int64_t min_points, int ndim, int nshifts, uint64_t seed, double* pts1_host,
                                   double* pts2_host, double* w_host);
/* No reference counterpart (the QMC workload's point set): Sobol points in
 * [0, 1)^dim (dim <= 3, Joe-Kuo directions, Gray-code order), optionally
 * scrambled (linear matrix scrambling + digital shift, seeded); u n x dim. */
int tcscf_sobol_points(int64_t n, int dim, int scramble, uint64_t seed, double* u_host);
/* Importance-sampled grid from n scrambled Sobol points: r = -ln(1 - u1)/2
 * (density 2 e^{-2r}), uniform directions, w = 4 pi r^2 / (2 n e^{-2r}). */
int tcscf_sobol_grid(int64_t n, uint64_t seed, double* pts_host, double* w_host);

/* loadSTO (src/SlaterTypeOrbital.cpp:24-76): reads the BASIS block of the ADF
 * file <directory>/<basis_set>/<element> (basis_set one of TZ3P, QZ6P, ATZ3P,
 * AQZ6P), expands each shell to m = -l..l and skips l > max_l.  Returns the
 * function count (writing at most `capacity` to out_host; out_host may be NULL
 * with capacity 0 to query), or < 0.  A missing element file gives 0 functions,
 * as in the reference. */
int tcscf_load_sto(const char* directory, const char* basis_set, const char* element, int max_l, tcscf_sto* out_host,
                   int capacity);

/* ---------------------------------------------------------------- host one-electron */
/* overlap / coreMatrixElement: src/SlaterTypeOrbital.hpp:125-171.  nb x nb row-major. */
int tcscf_sto_overlap_core(const tcscf_sto* fns, int nb, int Z, double* S_host, double* H_host);

/* overlap / coreMatrixElement for one kind of basis (all functions of one
 * kind): STO as tcscf_sto_overlap_core; Ochi (OchiBasis.hpp:117-155, all
 * functions of one alpha): S = delta, H = c I - delta_nn' alpha^2 / 2 with
 * c = alpha (l+1)(1 + 2 min(n, n') / (2l+3)) - Z and I = int fnl fnl' / r dr,
 * evaluated exactly from the Laguerre coefficients (the reference integrates
 * it on a 1000-point radial grid); hydrogen-like (HydrogenLikeBasis.hpp:67-80):
 * S = delta, H = -Z^2 / (2 n^2) delta.  nb x nb row-major. */
int tcscf_basis_one_electron(const tcscf_basis_fn* fns, int nb, int Z, double* S_host, double* H_host);

/* ---------------------------------------------------------------- device basis */
/* QMCGrid::setBasisFunctions: src/integration/integrateAll.hpp:88-137, for
 * functions of any tcscf_basis_fn kinds (mixed kinds allowed).
 * values_dev: n x nb; grads_dev: n x nb x 3 or NULL. */
int tcscf_basis_eval_fns(const double* pts_dev, int64_t n, const tcscf_basis_fn* fns_host, int nb, double* values_dev,
                         double* grads_dev, void* stream);
/* The same for a list of STOs. */
int tcscf_basis_eval(const double* pts_dev, int64_t n, const tcscf_sto* fns_host, int nb, double* values_dev,
                     double* grads_dev, void* stream);

/* ---------------------------------------------------------------- device F / V */
/* precomputeIntegrand + computeF / computeV fused: src/HartreeFock.hpp:23-229.
 *   F[n1, c, p] = scale * sum_n2 K_c(r1, r2) w2[n2] A2[n2, j] A2[n2, i]
 * with K generated per tile and never stored; coincident pairs (r12 == 0) give 0.
 * ncomp = 3 for TCSCF_K_JASTROW_GRAD (layout F[n1][3][npair]), else 1
 * (layout F[n1][npair]).  For TCSCF_K_TC_SCALAR, same_spin selects c_same or
 * c_oppo; TCSCF_K_JASTROW_GRAD uses c = 1 (scale by c_same / c_oppo).
 * ord2_dev (n2 int32, or NULL) is the order in which the r2 points are summed;
 * any permutation gives the same F in exact arithmetic, and summing shell by
 * shell (points sorted by radius) keeps the dynamic range inside each
 * 32-point FP32 step small.  out_dtype: TCSCF_F64 or TCSCF_F32.  Allocates and
 * caches a device workspace for the packed pair densities
 * (tcscf_F_workspace_bytes). */
int tcscf_compute_F(const double* pts1_dev, int64_t n1, const double* pts2_dev, const double* w2_dev,
                    const int32_t* ord2_dev, int64_t n2, const double* A2_dev, int nb, int kernel,
                    const tcscf_jastrow* jastrow, int same_spin, int precision, double scale, void* F_dev,
                    int out_dtype, void* stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops that must not call
 * hipMalloc): tcscf_compute_F with the same arguments, results and stream
 * semantics, but its workspace (the packed pair densities and padded
 * coordinates, tcscf_F_workspace_bytes(n2, nb, precision)) lives in a fixed
 * 128 MiB device array of the library instead of memory it allocates; it
 * calls no hipMalloc / hipFree.  A contraction whose workspace exceeds
 * 128 MiB is rejected with TCSCF_EINVAL (the Ne grid of 30,200 points and 30
 * functions fits at every precision).  The first call on a device sets the
 * array up (looks up its device address; the runtime maps it then): make that
 * call outside graph capture.  Every call on one GPU shares that GPU's array:
 * calls that overlap on one GPU (different streams or threads without
 * ordering) race; calls ordered on one stream do not. */
int tcscf_compute_F_noalloc(const double* pts1_dev, int64_t n1, const double* pts2_dev, const double* w2_dev,
                            const int32_t* ord2_dev, int64_t n2, const double* A2_dev, int nb, int kernel,
                            const tcscf_jastrow* jastrow, int same_spin, int precision, double scale, void* F_dev,
                            int out_dtype, void* stream);
size_t tcscf_F_workspace_bytes(int64_t n2, int nb, int precision);

/* tcscf_compute_F for the general Ochi Jastrow (the integrand lambdas of
 * src/unitTests/HartreeFockTest.cpp:276-316 with any jastrowFunctions::Ochi):
 *   kernel TCSCF_K_TC_SCALAR:   2/r12 - |grad1 u(r1,r2)|^2 - |grad1 u(r2,r1)|^2
 *   kernel TCSCF_K_JASTROW_GRAD: grad1 u(r1,r2), 3 components
 * with the coefficients of same_spin (the reference's VjiSame / VjiOppo, not
 * the c = 1 tensor of the r12-only form).  Coincident pairs give 0; at r = 0
 * the r1 / r2 terms of the gradient contribute 0.  Other arguments and layouts
 * as tcscf_compute_F. */
int tcscf_compute_F_ochi(const double* pts1_dev, int64_t n1, const double* pts2_dev, const double* w2_dev,
                         const int32_t* ord2_dev, int64_t n2, const double* A2_dev, int nb, int kernel,
                         const tcscf_ochi_jastrow* jastrow, int same_spin, int precision, double scale, void* F_dev,
                         int out_dtype, void* stream);

/* F_same and F_oppo of TCSCF_K_TC_SCALAR together (the two computeF calls of
 * src/unitTests/HartreeFockTest.cpp:276-296): the kernel is
 * 2/r12 - 2 c^2 a12^2 / (r12 + a12)^4, so F_spin = 2 F_C - 2 c_spin^2 a12^2 G
 * with F_C the 1/r12 contraction (at precision) and G the contraction of the
 * smooth, bounded 1/(r12 + a12)^4 (at smooth_precision, which may be a cheaper
 * split than the one the 1/r12 cusp needs).  Fs_dev, Fo_dev: n1 x npair each,
 * out_dtype; other arguments as tcscf_compute_F. */
int tcscf_compute_F_tc_pair(const double* pts1_dev, int64_t n1, const double* pts2_dev, const double* w2_dev,
                            const int32_t* ord2_dev, int64_t n2, const double* A2_dev, int nb,
                            const tcscf_jastrow* jastrow, int precision, int smooth_precision, double scale,
                            void* Fs_dev, void* Fo_dev, int out_dtype, void* stream);

/* One instance of a batched contraction: the arguments of tcscf_compute_F
 * (FP64 output, packed n1 x ncomp x npair; scale = 2 pi for a 2-D r2 grid). */
typedef struct tcscf_batch_item {
  const double* pts1_dev;
  int64_t n1;
  const double* pts2_dev;
  const double* w2_dev;
  const int32_t* ord2_dev;
  int64_t n2;
  const double* A2_dev;
  int nb;
  int kernel;
  tcscf_jastrow jastrow;
  int same_spin;
  double scale;
  void* F_dev;   /* output, tcscf_compute_F layout, of out_dtype */
  int out_dtype; /* TCSCF_F64 or TCSCF_F32 (FP32 storage of the mixed-precision path) */
} tcscf_batch_item;

/* No single reference counterpart: computeF / computeV (src/HartreeFock.hpp:
 * 88-229) for many instances of uneven size (atoms, bases, grids, kernels) at
 * once.  BF16X3 / BF16X4 / BF16X5: one persistent launch over a work queue of
 * (instance, row block, 256-pair block, component) tiles, largest instances
 * first; the rows per tile (16 or 32) come from an online autotuner that times
 * both on the head of the queue the first time a precision is used in the
 * process.  Results equal tcscf_compute_F's bit for bit (for a V item, when
 * its r1 points x r3 points x pairs stay below 2e12, above which the single
 * call folds its FP32 partials every step).  FP64 / FP32: one launch per
 * instance.  Synchronizes the stream before returning. */
int tcscf_batch_compute_F(const tcscf_batch_item* items_host, int count, int precision, void* stream);
/* Rows per tile the autotuner chose (16 or 32; 0 before the first batch). */
int tcscf_batch_tuned_rows(int precision);

/* integrateAllR1R2<false> (src/integration/integrateAll.hpp:221-317) from a
 * computed F: ERI[p(b1,b3)][p(b2,b4)] = sum_n1 w1 A1[n1,b1] A1[n1,b3] F[n1,p(b2,b4)]
 * (chemist (b1 b3 | b2 b4); the reference's answer(b1,b2,b3,b4)).  F_dev packed
 * n1 x npair FP64 (the 2 pi of a 2-D r2 grid is already in F); eri_dev
 * npair x npair.  FP64 MFMA GEMM split over r1. */
int tcscf_eri(const double* A1_dev, const double* w1_dev, int64_t n1, int nb, const void* F_dev, double* eri_dev,
              void* stream);

/* fillR12Array (src/AtomicBasis.hpp:167-177) with r12MatrixElement (:30-48),
 * the per-element 6-D QMC route, in the real basis: for the point pairs
 * (r1_k, r2_k) and weights w_k of tcscf_qmc_spherical_points(ndim = 6),
 *   eri[p(a,c)][p(b,d)] = sum_k w_k phi_a(r1_k) phi_c(r1_k) phi_b(r2_k) phi_d(r2_k) / r12_k
 * (coincident pairs give 0), the layout of tcscf_eri.  A1_dev / A2_dev: the basis
 * values at the r1 / r2 points (n x nb). */
int tcscf_eri_qmc6d(const double* pts1_dev, const double* pts2_dev, const double* w_dev, int64_t n,
                    const double* A1_dev, const double* A2_dev, int nb, double* eri_dev, void* stream);

/* atomicCoulombOperator and atomicExchangeOperator (src/AtomicBasis.hpp:59-140),
 * the per-element 5-D QMC route, for every quadruple at once, with the
 * reference's complex functions g = R(r) sph_legendre(l, |m|, theta) e^{i m phi}:
 * for the point pairs and weights of tcscf_qmc_spherical_points(ndim = 5) (r2 on
 * the phi = 0 half-plane, the phi2 integral in the weights),
 *   eri[2e], eri[2e+1] = Re, Im of sum_k w_k conj(g_a(r1_k)) g_c(r1_k) g_b(r2_k) g_d(r2_k) / r12_k,
 *   e = ((a nb + b) nb + c) nb + d  (complex nb^4, device),
 * which is the Coulomb integrand where (l, m) of a, c and of b, d agree and the
 * exchange integrand (phase e^{i (m_c - m_a) phi1}) where a, d and b, c agree.
 * g is formed on the device from the real table A1_dev / A2_dev (n x nb, as
 * tcscf_basis_eval_fns writes it): g_{+u} = (-1)^u (S_{+u} + i S_{-u}) / sqrt2,
 * g_{-u} = (-1)^u (S_{+u} - i S_{-u}) / sqrt2.  m_host[a] is the m of column a and
 * partner_host[a] the column with the same function and -m (a itself for
 * m = 0); the basis must hold both.  nb <= 32. */
int tcscf_eri_qmc5d_atomic(const double* pts1_dev, const double* pts2_dev, const double* w_dev, int64_t n,
                           const double* A1_dev, const double* A2_dev, int nb, const int* m_host,
                           const int* partner_host, double* eri_dev, void* stream);

/* integrateAllR1R2<true> (src/integration/integrateAll.hpp:184-215, 221-317),
 * real basis, for the vector kernel f = c_same grad_1 u(r1, r2) (kernel must be
 * TCSCF_K_JASTROW_GRAD):
 *   answer(b2, b1, b4, b3) = sum_r1 w1 A1_b1 A1_b3 scale sum_r2 w2 A2_b2 f.grad A2_b4
 * written as eri_dev[p(b1, b3)][b2 nb + b4] (npair x nb^2 FP64).  dA2_dev is
 * n2 x nb x 3; scale = 2 pi for the reference's 2-D r2 grid (points with y = 0).
 * The inner integral is one FP64 MFMA GEMM with K = 3 n2 (point x component). */
int tcscf_eri_gradient(const double* pts1_dev, const double* A1_dev, const double* w1_dev, int64_t n1,
                       const double* pts2_dev, const double* w2_dev, int64_t n2, const double* A2_dev,
                       const double* dA2_dev, int nb, int kernel, const tcscf_jastrow* jastrow, double scale,
                       double* eri_dev, void* stream);

/* threeElectronIntegrals (src/integration/integrateAll.hpp:369-442), real
 * basis: Q(p,q,r,s,t,u) = sum_r1 w1 A1_p A1_s V(q,t).V(r,u), with V = c V_dev
 * (V_dev the packed c = 1 n1 x 3 x npair tensor of TCSCF_K_JASTROW_GRAD; the
 * reference evaluates the same-spin gradient, so c = c_same).  Q_dev: nb^6
 * FP64, index order (p, q, r, s, t, u).  Validation-sized: nb <= 16. */
int tcscf_three_electron(const double* A1_dev, const double* w1_dev, int64_t n1, const double* V_dev, int nb, double c,
                         double* Q_dev, void* stream);

/* ---------------------------------------------------------------- SCF drivers */
/* RCSHartreeFock::compute (src/HartreeFock.cpp:785-886).  S, H host nb x nb
 * row-major (S symmetric: the reference's column-major overlap view holds the
 * same numbers); A1_dev n1 x nb; w1_dev n1; F_dev packed n1 x npair (f_dtype).
 * Synchronous.  energy_trace (max_iter), D_out (nb x nb), evals_out (nb) and
 * obs (fockOperator / eigenvectors, nspin = 1) may be NULL. */
int tcscf_rcs_scf(int nelec, int nb, const double* S, const double* H, const double* A1_dev, const double* w1_dev,
                  int64_t n1, const void* F_dev, int f_dtype, const tcscf_scf_opts* opts, tcscf_scf_result* res,
                  double* energy_trace, double* D_out, double* evals_out, tcscf_scf_observables* obs, void* stream);

/* UOSHartreeFock::compute (src/HartreeFock.cpp:890-1003).  D_out 2 x nb x nb,
 * evals_out 2 x nb, obs nspin = 2. */
int tcscf_uos_scf(int nup, int ndown, int nb, const double* S, const double* H, const double* A1_dev,
                  const double* w1_dev, int64_t n1, const void* F_dev, int f_dtype, const tcscf_scf_opts* opts,
                  tcscf_scf_result* res, double* energy_trace, double* D_out, double* evals_out,
                  tcscf_scf_observables* obs, void* stream);

/* TCHartreeFock::compute(overlap, oneElectronTerms, r1Grid, FjiSame, FjiOppo,
 * VjiSame, VjiOppo, respectOneElectronSymmetry) (src/HartreeFock.hpp:386-394,
 * src/HartreeFock.cpp:1007-1362), real basis.  Fs_dev, Fo_dev: packed
 * n1 x npair F_same / F_oppo (TCSCF_K_TC_SCALAR).  Vs_dev, Vo_dev: packed
 * n1 x 3 x npair VjiSame / VjiOppo (TCSCF_K_JASTROW_GRAD or the general Ochi
 * gradient for each spin), any two tensors; or Vo_dev = NULL, Vs_dev the c = 1
 * r12 Jastrow gradient and jastrow's c_same / c_oppo its scales (V_same =
 * c_same V, V_oppo = c_oppo V: one tensor read instead of two; jastrow is
 * unused when Vo_dev is given).  All of f_dtype: TCSCF_F32 stores them in
 * FP32, the Fock build's arithmetic is FP64 either way.  A1_dev n1 x nb;
 * dA1_dev n1 x nb x 3; S, H host.  The Fock build runs on the device, the FP64
 * generalized non-Hermitian eigen-solve (role of LAPACK ggev, per symmetry
 * block), the Re(lambda) sort and the MGS of the occupied orbitals on the
 * host.  Synchronous.  D_out 2 x nb x nb; pseudo_out nup + ndown occupied
 * pseudo-energies; obs (fockOperator, occupiedOrbitals) may be NULL. */
int tcscf_tc_scf(int nup, int ndown, int nb, const double* S, const double* H, const double* A1_dev,
                 const double* dA1_dev, const double* w1_dev, int64_t n1, const void* Fs_dev, const void* Fo_dev,
                 const void* Vs_dev, const void* Vo_dev, int f_dtype, const tcscf_jastrow* jastrow,
                 const tcscf_scf_opts* opts, tcscf_scf_result* res, double* energy_trace, double* D_out,
                 double* pseudo_out, tcscf_scf_observables* obs, void* stream);

/* One TC Fock build (the per-iteration sweeps and assembly of
 * src/HartreeFock.cpp:1118-1224) from D_host (2 x nb x nb) and the occupied
 * orbitals C_host (2 x nb x max(nup, ndown), column b of spin s used for
 * b < nocc_s): writes G2 and G3 (2 x nb x nb each, host) before the mask.
 * V tensors as in tcscf_tc_scf.  When D is C C^T (as in every SCF
 * iteration), nb <= 32 and the rows are FP32 of one V tensor, the sweep works
 * from C alone; any other D takes the D-based sweep.  The two agree to FP64
 * rounding. */
int tcscf_tc_fock(int nup, int ndown, int nb, const double* A1_dev, const double* dA1_dev, const double* w1_dev,
                  int64_t n1, const void* Fs_dev, const void* Fo_dev, const void* Vs_dev, const void* Vo_dev,
                  int f_dtype, const tcscf_jastrow* jastrow, const double* D_host, const double* C_host,
                  double* G2_host, double* G3_host, void* stream);

/* No reference counterpart (a diagnostic): the point-kernel launch the TC
 * Fock build picks for these sizes, f_dtype and V form (general_v: VjiSame
 * and VjiOppo as two tensors).  info[0..2]: V rows staged in LDS (0 / 1), waves
 * per workgroup, LDS bytes of the two-spin build; info[3..5] the same for the
 * restricted closed-shell build. */
int tcscf_tc_launch_info(int nup, int ndown, int nb, int f_dtype, int general_v, int* info);

/* ---------------------------------------------------------------- batched SCF */
enum tcscf_scf_kind { TCSCF_SCF_RCS = 0, TCSCF_SCF_UOS = 1, TCSCF_SCF_TC = 2 };

/* One instance of a batched solve: the arguments of tcscf_rcs_scf /
 * tcscf_uos_scf (kind RCS / UOS: F_dev, f_dtype; RCS uses nup + ndown electrons)
 * or tcscf_tc_scf (kind TC: Fs_dev, Fo_dev, V_dev, Vo_dev, f_dtype, dA1_dev, jastrow). */
typedef struct tcscf_scf_item {
  int kind;
  int nup, ndown, nb;
  const double* S;
  const double* H;
  const double* A1_dev;
  const double* dA1_dev;
  const double* w1_dev;
  int64_t n1;
  const void* F_dev;
  int f_dtype; /* of F_dev, or of Fs_dev, Fo_dev and V_dev */
  const void* Fs_dev;
  const void* Fo_dev;
  const void* V_dev;
  const void* Vo_dev; /* VjiOppo with V_dev = VjiSame, or NULL (V_dev c = 1, scaled by jastrow) */
  tcscf_jastrow jastrow;
  double* energy_trace; /* host, opts->max_iter entries, or NULL */
  double* D_out;        /* host, nspin x nb x nb (RCS 1, else 2), or NULL */
} tcscf_scf_item;

/* No single reference counterpart: RCS/UOS/TCHartreeFock::compute
 * (src/HartreeFock.cpp:785-1362) for many instances at once, with the same
 * options.  Each iteration replays one HIP graph holding every active
 * instance's Fock build (spread over 4 streams); the FP64 host halves of the
 * iterations run on a pool of host threads.  The graph is captured again when
 * an instance converges.  results[k].status is TCSCF_OK / ENOTCONV / ESOLVER per
 * instance; t_sweep_ms / t_solve_ms hold the batch's device / host totals.
 * Returns 0 once every instance has finished, < 0 on a set-up error.
 * graph_builds (may be NULL): number of graph captures. */
int tcscf_batch_scf(const tcscf_scf_item* items, int count, const tcscf_scf_opts* opts, tcscf_scf_result* results,
                    int* graph_builds, void* stream);

/* ---------------------------------------------------------------- diagnostics */
/* No reference counterpart: runs one MFMA fragment (which = 0: f64 16x16x4,
 * 1: f32 16x16x4, 2: bf16 16x16x32) on A_dev (16 x K) and B_dev (K x 16),
 * row-major FP64, with lane l = 16 g + r supplying A[r][k], B[k][r] for its k
 * slice, and writes the raw accumulator registers to D_dev[4 l + i] (256
 * doubles), so tests can pin the fragment layouts the contraction kernels use. */
int tcscf_selftest_mfma(int which, const double* A_dev, const double* B_dev, double* D_dev, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* TCSCF_HIP_H */
