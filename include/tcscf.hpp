/* C++ host facade: the reference's call surface over the tcscf_hip C ABI.
 *
 * Mirrors the names, argument meaning and error behaviour of
 *   src/SlaterTypeOrbital.hpp      SlaterTypeOrbital, overlap, coreMatrixElement, loadSTO
 *   src/OchiBasis.hpp, src/HydrogenLikeBasis.hpp
 *                                  OchiBasisFunction, HydrogenLikeBasisFunction
 *   src/integration/integrateAll.hpp  integration::QMCGrid, integrateAllR1R2,
 *                                  threeElectronIntegrals
 *   src/HartreeFock.hpp            computeF / computeV (fused with precomputeIntegrand),
 *                                  RCSHartreeFock, UOSHartreeFock, TCHartreeFock
 * so a driver written against the reference (src/unitTests/HartreeFockTest.cpp)
 * ports by swapping includes.  Differences, all forced by the boundary:
 *   - the integrand lambdas f(r1, r2) become kernel descriptors (Coulomb, the TC
 *     scalar kernel and the Jastrow gradient of the r12 Ochi term), since a
 *     lambda cannot cross into device code compiled elsewhere;
 *   - grids, basis tables and F / V live in HBM (DeviceArray); F and V are
 *     packed over pairs j <= i; TCHartreeFock takes VjiSame and VjiOppo as the
 *     reference does, or one c = 1 tensor V with the r12 Jastrow's scales
 *     (V_same = c_same V, V_oppo = c_oppo V), which reads one tensor instead of two;
 *   - fockOperator, eigenvectors and occupiedOrbitals are row-major host
 *     vectors indexed as the reference's accessors: (spin, j, i), (spin, j, a),
 *     [spin](j, b);
 *   - the basis is real (real solid harmonics), a unitary change of the
 *     reference's complex Y_lm that leaves every energy unchanged;
 *   - a non-converged RCS solve throws ConvergenceError (the reference aborts).
 * Header-only; link with -ltcscf_hip and the HIP runtime. */
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <complex>
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <array>
#include <vector>

#include "tcscf_hip.h"

namespace tcscf {

using IndexType = std::ptrdiff_t;

/* Errors: TCSCF_ENOTCONV -> ConvergenceError, TCSCF_ESOLVER -> SolverError
 * (the reference's LvArray::dense::Error), everything else -> Error. */
class Error : public std::runtime_error {
 public:
  using std::runtime_error::runtime_error;
};
class ConvergenceError : public Error {
 public:
  using Error::Error;
};
class SolverError : public Error {
 public:
  using Error::Error;
};

namespace detail {
inline int check(int rc) {
  if (rc >= 0) return rc;
  const std::string msg = tcscf_last_error();
  if (rc == TCSCF_ENOTCONV) throw ConvergenceError(msg);
  if (rc == TCSCF_ESOLVER) throw SolverError(msg);
  throw Error(msg);
}
inline void hip(hipError_t e) {
  if (e != hipSuccess) throw Error(std::string("HIP: ") + hipGetErrorString(e));
}
}  // namespace detail

/* Owning device (HBM) array. */
template <class T>
class DeviceArray {
 public:
  DeviceArray() = default;
  explicit DeviceArray(std::size_t n) { resize(n); }
  DeviceArray(const DeviceArray&) = delete;
  DeviceArray& operator=(const DeviceArray&) = delete;
  DeviceArray(DeviceArray&& o) noexcept : p_(std::exchange(o.p_, nullptr)), n_(std::exchange(o.n_, 0)) {}
  DeviceArray& operator=(DeviceArray&& o) noexcept {
    if (this != &o) {
      release();
      p_ = std::exchange(o.p_, nullptr);
      n_ = std::exchange(o.n_, 0);
    }
    return *this;
  }
  ~DeviceArray() { release(); }

  void resize(std::size_t n) {
    release();
    if (n) detail::hip(hipMalloc(&p_, n * sizeof(T)));
    n_ = n;
  }
  void upload(const T* h, std::size_t n) {
    if (n != n_) resize(n);
    if (n) detail::hip(hipMemcpy(p_, h, n * sizeof(T), hipMemcpyHostToDevice));
  }
  std::vector<T> download() const {
    std::vector<T> h(n_);
    if (n_) detail::hip(hipMemcpy(h.data(), p_, n_ * sizeof(T), hipMemcpyDeviceToHost));
    return h;
  }
  T* data() { return p_; }
  const T* data() const { return p_; }
  std::size_t size() const { return n_; }
  bool empty() const { return n_ == 0; }

 private:
  void release() {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    n_ = 0;
  }
  T* p_ = nullptr;
  std::size_t n_ = 0;
};

/* ------------------------------------------------------------------ STO basis */
/* SlaterTypeOrbital (src/SlaterTypeOrbital.hpp:15-113): N r^(n-1) e^(-alpha r) S_lm,
 * N = (2 alpha)^n sqrt(2 alpha / (2n)!); l <= 3 on the device. */
template <class REAL>
struct SlaterTypeOrbital {
  using Real = REAL;
  static constexpr bool isBasisReal = true;
  static constexpr bool isOrthonormal = false;

  SlaterTypeOrbital(Real alphaP, int nP, int lP, int mP) : alpha{alphaP}, n{nP}, l{lP}, m{mP} {
    if (n < 1 || n < l || l < std::abs(m) || l > 3) throw Error("SlaterTypeOrbital: need n >= 1, n >= l >= |m|, l <= 3");
    normalization = norm();
  }
  void resetOrbitalExponent(Real newAlpha) {
    alpha = newAlpha;
    normalization = norm();
  }
  tcscf_basis_fn abi() const { return tcscf_basis_fn{TCSCF_BASIS_STO, n, l, m, static_cast<double>(alpha)}; }

  Real alpha;
  int n, l, m;
  Real normalization;

 private:
  Real norm() const { return std::pow(2 * alpha, n) * std::sqrt(2 * alpha / std::tgamma(2.0 * n + 1)); }
};

/* OchiBasisFunction (src/OchiBasis.hpp:15-155): N r^l L_n^(2l+2)(2 alpha r)
 * e^(-alpha r) S_lm, N = (2 alpha)^(l + 3/2) / sqrt((n + 2l + 2)! / n!).  The
 * functions of one alpha are orthonormal; l <= 3 on the device. */
template <class REAL>
struct OchiBasisFunction {
  using Real = REAL;
  static constexpr bool isBasisReal = true;
  static constexpr bool isOrthonormal = true;

  OchiBasisFunction(Real alphaP, int nP, int lP, int mP) : alpha{alphaP}, n{nP}, l{lP}, m{mP} {
    if (n < 0 || l < std::abs(m) || l > 3 || !(alpha > 0))
      throw Error("OchiBasisFunction: need n >= 0, l >= |m|, l <= 3, alpha > 0");
  }
  tcscf_basis_fn abi() const { return tcscf_basis_fn{TCSCF_BASIS_OCHI, n, l, m, static_cast<double>(alpha)}; }

  Real alpha;
  int n, l, m;
};

/* HydrogenLikeBasisFunction (src/HydrogenLikeBasis.hpp:12-57): the bound
 * eigenfunctions of -1/2 lap - Z/r, N r^l L_(n-l-1)^(2l+1)(2 beta r) e^(-beta r)
 * S_lm with beta = Z / n and the reference's negative normalization.  The
 * device also tabulates their gradients (the reference leaves them out). */
template <class REAL>
struct HydrogenLikeBasisFunction {
  using Real = REAL;
  static constexpr bool isBasisReal = true;
  static constexpr bool isOrthonormal = true;

  HydrogenLikeBasisFunction(int Z, int nP, int lP, int mP) : n{nP}, l{lP}, m{mP}, beta{Z / Real(nP)}, Z_{Z} {
    if (Z <= 0 || n <= 0 || l >= n || l < std::abs(m) || l > 3)
      throw Error("HydrogenLikeBasisFunction: need Z > 0, n > l >= |m|, l <= 3");
  }
  tcscf_basis_fn abi() const { return tcscf_basis_fn{TCSCF_BASIS_HYDROGEN, n, l, m, static_cast<double>(Z_)}; }

  int n, l, m;
  Real beta;

 private:
  int Z_;
};

namespace detail {
template <class B>
std::vector<tcscf_basis_fn> abi(const std::vector<B>& fns) {
  std::vector<tcscf_basis_fn> out;
  out.reserve(fns.size());
  for (const B& b : fns) out.push_back(b.abi());
  return out;
}
template <class B>
void oneElectronPair(int Z, const B& b1, const B& b2, double& S, double& H) {
  const tcscf_basis_fn f[2] = {b1.abi(), b2.abi()};
  double s[4], h[4];
  detail::check(tcscf_basis_one_electron(f, 2, Z, s, h));
  S = s[1];
  H = h[1];
}
}  // namespace detail

/* overlap / coreMatrixElement: src/SlaterTypeOrbital.hpp:125-171 (closed
 * forms), src/OchiBasis.hpp:118-153 (the reference's coefficient times the
 * radial integral, here in closed form), src/HydrogenLikeBasis.hpp:67-80. */
template <class B>
typename B::Real overlap(const B& b1, const B& b2) {
  double S, H;
  detail::oneElectronPair(0, b1, b2, S, H);
  return static_cast<typename B::Real>(S);
}
template <class B>
typename B::Real coreMatrixElement(int Z, const B& b1, const B& b2) {
  double S, H;
  detail::oneElectronPair(Z, b1, b2, S, H);
  return static_cast<typename B::Real>(H);
}
/* Whole matrices at once (row-major nb x nb; S is symmetric, so it is also the
 * reference's column-major overlap view). */
template <class B>
void oneElectronMatrices(const std::vector<B>& fns, int Z, std::vector<double>& S, std::vector<double>& H) {
  const auto f = detail::abi(fns);
  S.assign(f.size() * f.size(), 0.0);
  H.assign(f.size() * f.size(), 0.0);
  detail::check(tcscf_basis_one_electron(f.data(), static_cast<int>(f.size()), Z, S.data(), H.data()));
}

/* loadSTO (src/SlaterTypeOrbital.cpp:24-76). */
inline std::vector<SlaterTypeOrbital<double>> loadSTO(const std::string& directory, const std::string& basisSet,
                                                      const std::string& element, int maxL = INT_MAX) {
  const int n = detail::check(tcscf_load_sto(directory.c_str(), basisSet.c_str(), element.c_str(), maxL, nullptr, 0));
  std::vector<tcscf_sto> f(static_cast<std::size_t>(n));
  detail::check(tcscf_load_sto(directory.c_str(), basisSet.c_str(), element.c_str(), maxL, f.data(), n));
  std::vector<SlaterTypeOrbital<double>> out;
  out.reserve(f.size());
  for (const tcscf_sto& s : f) out.emplace_back(s.zeta, s.n, s.l, s.m);
  return out;
}

/* --------------------------------------------------------------------- grids */
namespace integration {

/* QuadratureGrid (src/integration/quadrature.hpp:26-35): points 3 x N (SoA),
 * weights N, in HBM; host copies kept for set-up and checks. */
template <class REAL>
struct QuadratureGrid {
  DeviceArray<REAL> points;
  DeviceArray<REAL> weights;
  std::vector<REAL> hostPoints;
  std::vector<REAL> hostWeights;
};

/* QMCGrid (src/integration/integrateAll.hpp:12-142).  NDIM == 2 keeps y = 0
 * (the reference's axially symmetric r2 grid; computeF multiplies by 2 pi). */
template <class REAL, int NDIM>
struct QMCGrid {
  static_assert(std::is_same_v<REAL, double>, "the device path computes in FP64 / mixed precision from FP64 grids");
  static_assert(NDIM == 2 || NDIM == 3);

  /* QMCGrid(gridSize): a randomly shifted rank-1 lattice (2 shifts) through the
   * TreutlerAhlrichsM4(0, 0.9) x Azimuthal x Linear maps, as the reference. */
  explicit QMCGrid(IndexType gridSize, std::uint64_t seed = 0) {
    const int64_t n = tcscf_qmc_lattice_npts(gridSize, 2);
    if (n < 0) detail::check(static_cast<int>(n));
    std::vector<REAL> p(3 * n), w(n);
    const int64_t got = tcscf_qmc_lattice_grid(gridSize, NDIM, 2, seed, p.data(), w.data());
    if (got < 0) detail::check(static_cast<int>(got));
    setPoints(std::move(p), std::move(w));
  }
  /* Explicit points (3 x N SoA) and weights (every Jacobian included). */
  QMCGrid(std::vector<REAL> points, std::vector<REAL> weights) { setPoints(std::move(points), std::move(weights)); }

  /* Gauss-Chebyshev radial x Lebedev (TreutlerAhlrichsLebedev in QMCGrid shape). */
  static QMCGrid radialLebedev(int nRadial, int lebedevOrder, int radialMap = TCSCF_RADIAL_M4, double a = 1.0,
                               double eps = 1.0) {
    static_assert(NDIM == 3);
    const int na = detail::check(tcscf_lebedev_npts(lebedevOrder));
    const std::size_t n = static_cast<std::size_t>(nRadial) * na;
    std::vector<REAL> p(3 * n), w(n);
    detail::check(tcscf_product_grid(nRadial, radialMap, a, eps, lebedevOrder, p.data(), w.data()));
    return QMCGrid(std::move(p), std::move(w));
  }
  /* n scrambled-Sobol points, r ~ 2 e^{-2r}, uniform directions, importance weights. */
  static QMCGrid sobol(IndexType n, std::uint64_t seed) {
    static_assert(NDIM == 3);
    std::vector<REAL> p(3 * n), w(n);
    detail::check(tcscf_sobol_grid(n, seed, p.data(), w.data()));
    return QMCGrid(std::move(p), std::move(w));
  }

  /* setBasisFunctions (src/integration/integrateAll.hpp:88-137). */
  template <class BASIS>
  void setBasisFunctions(const std::vector<BASIS>& basisFunctions, bool storeGradients) {
    const auto f = detail::abi(basisFunctions);
    nb_ = static_cast<IndexType>(f.size());
    basisValues.resize(static_cast<std::size_t>(nGrid() * nb_));
    if (storeGradients) basisGradients.resize(static_cast<std::size_t>(nGrid() * nb_ * 3));
    else basisGradients.resize(0);
    detail::check(tcscf_basis_eval_fns(quadratureGrid.points.data(), nGrid(), f.data(), static_cast<int>(nb_),
                                       basisValues.data(), storeGradients ? basisGradients.data() : nullptr, stream));
    detail::hip(hipStreamSynchronize(stream));
  }

  IndexType nBasis() const { return nb_; }
  IndexType nGrid() const { return static_cast<IndexType>(quadratureGrid.hostWeights.size()); }
  bool gradientsStored() const { return !basisGradients.empty(); }

  QuadratureGrid<REAL> quadratureGrid;
  DeviceArray<REAL> basisValues;     /* N x nb */
  DeviceArray<REAL> basisGradients;  /* N x nb x 3 (Cartesian) */
  DeviceArray<int32_t> radialOrder;  /* summation order of contractions over this grid */
  hipStream_t stream = nullptr;

 private:
  void setPoints(std::vector<REAL> p, std::vector<REAL> w) {
    const std::size_t n = w.size();
    if (p.size() != 3 * n) throw Error("QMCGrid: points must be 3 x N");
    std::vector<std::pair<REAL, int32_t>> r(n);
    for (std::size_t k = 0; k < n; ++k)
      r[k] = {std::sqrt(p[k] * p[k] + p[n + k] * p[n + k] + p[2 * n + k] * p[2 * n + k]), static_cast<int32_t>(k)};
    std::stable_sort(r.begin(), r.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    std::vector<int32_t> ord(n);
    for (std::size_t k = 0; k < n; ++k) ord[k] = r[k].second;
    quadratureGrid.points.upload(p.data(), p.size());
    quadratureGrid.weights.upload(w.data(), w.size());
    radialOrder.upload(ord.data(), ord.size());
    quadratureGrid.hostPoints = std::move(p);
    quadratureGrid.hostWeights = std::move(w);
  }
  IndexType nb_ = 0;
};

}  // namespace integration

/* ------------------------------------------------------------- the integrals */
/* Jastrow u(r12) = c r12 / (r12 + a12) per spin pair (the r12 Ochi term of
 * src/unitTests/HartreeFockTest.cpp:351-360); fromB(b) gives r12 / (2 (1 + b r12)). */
struct Jastrow {
  double a12, c_same, c_oppo;
  static Jastrow fromB(double b) { return Jastrow{1.0 / b, 0.25 / b, 0.5 / b}; }
  tcscf_jastrow abi() const { return tcscf_jastrow{a12, c_same, c_oppo}; }
};

/* jastrowFunctions::Ochi{a, a12, c, S} (src/jastrowFunctions.hpp:9-115): terms
 * S[t] = (i, j, k) of (r12/(r12+a12))^i (r1/(r1+a))^j (r2/(r2+a))^k with
 * coefficients c[t] = {c(t, false), c(t, true)} = {opposite spin, same spin}. */
struct OchiJastrow {
  double a, a12;
  std::vector<std::array<double, 2>> c;
  std::vector<std::array<int, 3>> S;

  tcscf_ochi_jastrow abi() const {
    if (c.size() != S.size() || S.size() > TCSCF_OCHI_MAX_TERMS)
      throw Error("OchiJastrow: c and S need one entry per term, at most 8 terms");
    tcscf_ochi_jastrow u{};
    u.a = a;
    u.a12 = a12;
    u.nterms = static_cast<int>(S.size());
    for (std::size_t t = 0; t < S.size(); ++t) {
      u.i[t] = S[t][0];
      u.j[t] = S[t][1];
      u.k[t] = S[t][2];
      u.c_oppo[t] = c[t][0];
      u.c_same[t] = c[t][1];
    }
    return u;
  }
  /* The TC Fock build takes one V with V_same = c_same V and V_oppo = c_oppo V.
   * With V = VjiSame that is {a12, 1, rho} when c_oppo = rho c_same term by term. */
  Jastrow tcScales() const {
    double rho = 0;
    bool any = false;
    for (const auto& ct : c) {
      if (ct[1] == 0) {
        if (ct[0] != 0) throw Error("OchiJastrow: TC needs c_oppo = rho * c_same term by term");
        continue;
      }
      const double r = ct[0] / ct[1];
      if (any && std::fabs(r - rho) > 1e-14 * std::fabs(rho)) throw Error("OchiJastrow: TC needs c_oppo = rho * c_same term by term");
      rho = r;
      any = true;
    }
    if (!any) throw Error("OchiJastrow: TC needs a nonzero same-spin coefficient");
    return Jastrow{a12, 1.0, rho};
  }
};

enum class Precision : int {
  FP64 = TCSCF_FP64,
  FP32 = TCSCF_FP32,
  BF16X3 = TCSCF_BF16X3,
  BF16X4 = TCSCF_BF16X4,
  BF16X5 = TCSCF_BF16X5
};

namespace detail {
/* tcscf_dtype of a device array's element type (FP64, or FP32 storage of the
 * mixed-precision path). */
template <class FT>
constexpr int dtype() {
  static_assert(std::is_same_v<FT, double> || std::is_same_v<FT, float>, "F / V elements are double or float");
  return std::is_same_v<FT, float> ? TCSCF_F32 : TCSCF_F64;
}

template <class FT, class REAL, int N1, int N2>
DeviceArray<FT> contract(const integration::QMCGrid<REAL, N1>& r1, const integration::QMCGrid<REAL, N2>& r2,
                         int kernel, const tcscf_jastrow* jp, int sameSpin, Precision prec) {
  if (r1.nBasis() != r2.nBasis() && kernel != TCSCF_K_COULOMB) throw Error("computeF: different bases");
  const int nb = static_cast<int>(r2.nBasis());
  if (nb == 0) throw Error("computeF: setBasisFunctions was not called on the inner grid");
  const IndexType npair = static_cast<IndexType>(nb) * (nb + 1) / 2;
  const int ncomp = kernel == TCSCF_K_JASTROW_GRAD ? 3 : 1;
  DeviceArray<FT> F(static_cast<std::size_t>(r1.nGrid() * ncomp * npair));
  const double scale = N2 == 2 ? 2 * 3.14159265358979323846 : 1.0;
  check(tcscf_compute_F(r1.quadratureGrid.points.data(), r1.nGrid(), r2.quadratureGrid.points.data(),
                        r2.quadratureGrid.weights.data(), r2.radialOrder.data(), r2.nGrid(), r2.basisValues.data(), nb,
                        kernel, jp, sameSpin, static_cast<int>(prec), scale, F.data(), dtype<FT>(), r1.stream));
  hip(hipStreamSynchronize(r1.stream));
  return F;
}

template <class FT, class REAL, int N1, int N2>
DeviceArray<FT> contractOchi(const integration::QMCGrid<REAL, N1>& r1, const integration::QMCGrid<REAL, N2>& r2,
                             int kernel, const OchiJastrow& jastrow, bool sameSpin, Precision prec) {
  if (r1.nBasis() != r2.nBasis()) throw Error("computeF: different bases");
  const int nb = static_cast<int>(r2.nBasis());
  if (nb == 0) throw Error("computeF: setBasisFunctions was not called on the inner grid");
  const IndexType npair = static_cast<IndexType>(nb) * (nb + 1) / 2;
  const int ncomp = kernel == TCSCF_K_JASTROW_GRAD ? 3 : 1;
  DeviceArray<FT> F(static_cast<std::size_t>(r1.nGrid() * ncomp * npair));
  const double scale = N2 == 2 ? 2 * 3.14159265358979323846 : 1.0;
  const tcscf_ochi_jastrow u = jastrow.abi();
  check(tcscf_compute_F_ochi(r1.quadratureGrid.points.data(), r1.nGrid(), r2.quadratureGrid.points.data(),
                             r2.quadratureGrid.weights.data(), r2.radialOrder.data(), r2.nGrid(),
                             r2.basisValues.data(), nb, kernel, &u, sameSpin ? 1 : 0, static_cast<int>(prec), scale,
                             F.data(), dtype<FT>(), r1.stream));
  hip(hipStreamSynchronize(r1.stream));
  return F;
}
}  // namespace detail

/* precomputeIntegrand + computeF (src/HartreeFock.hpp:23-180) for f = 1/r12:
 * packed F[n1][p(j, i)] = scale sum_n2 f(r1, r2) w2 phi_j(r2) phi_i(r2).  FT:
 * the stored element type (float: the FP32 storage of the mixed-precision path). */
template <class FT = double, class REAL, int N1, int N2>
DeviceArray<FT> computeF(const integration::QMCGrid<REAL, N1>& r1Grid, const integration::QMCGrid<REAL, N2>& r2Grid,
                         Precision precision = Precision::FP64) {
  return detail::contract<FT>(r1Grid, r2Grid, TCSCF_K_COULOMB, nullptr, 1, precision);
}
/* The TC scalar kernel 2/r12 - |grad_1 u12|^2 - |grad_2 u12|^2 of one spin pair
 * (src/unitTests/HartreeFockTest.cpp:276-296): FjiSame / FjiOppo. */
template <class FT = double, class REAL, int N1, int N2>
DeviceArray<FT> computeF(const integration::QMCGrid<REAL, N1>& r1Grid, const integration::QMCGrid<REAL, N2>& r2Grid,
                         const Jastrow& jastrow, bool sameSpin, Precision precision = Precision::FP64) {
  const tcscf_jastrow j = jastrow.abi();
  return detail::contract<FT>(r1Grid, r2Grid, TCSCF_K_TC_SCALAR, &j, sameSpin ? 1 : 0, precision);
}
/* The same for the general Ochi Jastrow (r12, r1 and r2 terms). */
template <class FT = double, class REAL, int N1, int N2>
DeviceArray<FT> computeF(const integration::QMCGrid<REAL, N1>& r1Grid, const integration::QMCGrid<REAL, N2>& r2Grid,
                         const OchiJastrow& jastrow, bool sameSpin, Precision precision = Precision::FP64) {
  return detail::contractOchi<FT>(r1Grid, r2Grid, TCSCF_K_TC_SCALAR, jastrow, sameSpin, precision);
}
/* FjiSame and FjiOppo together: the TC scalar kernel is 2/r12 - 2 c^2 a12^2 / (r12 + a12)^4,
 * so both follow from the 1/r12 contraction (at precision) and the contraction of the smooth
 * part (at smoothPrecision). */
template <class FT = double, class REAL, int N1, int N2>
std::pair<DeviceArray<FT>, DeviceArray<FT>> computeFPair(const integration::QMCGrid<REAL, N1>& r1Grid,
                                                         const integration::QMCGrid<REAL, N2>& r2Grid,
                                                         const Jastrow& jastrow, Precision precision = Precision::FP64,
                                                         Precision smoothPrecision = Precision::FP64) {
  if (r1Grid.nBasis() != r2Grid.nBasis()) throw Error("computeFPair: different bases");
  const int nb = static_cast<int>(r2Grid.nBasis());
  if (nb == 0) throw Error("computeFPair: setBasisFunctions was not called on the inner grid");
  const std::size_t len = static_cast<std::size_t>(r1Grid.nGrid()) * (static_cast<std::size_t>(nb) * (nb + 1) / 2);
  DeviceArray<FT> Fs(len), Fo(len);
  const tcscf_jastrow j = jastrow.abi();
  const double scale = N2 == 2 ? 2 * 3.14159265358979323846 : 1.0;
  detail::check(tcscf_compute_F_tc_pair(r1Grid.quadratureGrid.points.data(), r1Grid.nGrid(),
                                        r2Grid.quadratureGrid.points.data(), r2Grid.quadratureGrid.weights.data(),
                                        r2Grid.radialOrder.data(), r2Grid.nGrid(), r2Grid.basisValues.data(), nb, &j,
                                        static_cast<int>(precision), static_cast<int>(smoothPrecision), scale,
                                        Fs.data(), Fo.data(), detail::dtype<FT>(), r1Grid.stream));
  detail::hip(hipStreamSynchronize(r1Grid.stream));
  return {std::move(Fs), std::move(Fo)};
}
/* computeV (src/HartreeFock.hpp:185-229) for c = 1: packed V[n1][xyz][p(j, i)]. */
template <class FT = double, class REAL, int N1, int N3>
DeviceArray<FT> computeV(const integration::QMCGrid<REAL, N1>& r1Grid, const integration::QMCGrid<REAL, N3>& r3Grid,
                         const Jastrow& jastrow, Precision precision = Precision::FP64) {
  const tcscf_jastrow j = jastrow.abi();
  return detail::contract<FT>(r1Grid, r3Grid, TCSCF_K_JASTROW_GRAD, &j, 1, precision);
}
/* computeV for the general Ochi Jastrow: the reference's VjiSame (sameSpin) or
 * VjiOppo, packed V[n1][xyz][p(j, i)]. */
template <class FT = double, class REAL, int N1, int N3>
DeviceArray<FT> computeV(const integration::QMCGrid<REAL, N1>& r1Grid, const integration::QMCGrid<REAL, N3>& r3Grid,
                         const OchiJastrow& jastrow, bool sameSpin, Precision precision = Precision::FP64) {
  return detail::contractOchi<FT>(r1Grid, r3Grid, TCSCF_K_JASTROW_GRAD, jastrow, sameSpin, precision);
}

/* integrateAllR1R2<false> (src/integration/integrateAll.hpp:221-317) from F:
 * answer(b1, b2, b3, b4), nb^4 row-major (host). */
template <class REAL, int N1>
std::vector<REAL> integrateAllR1R2(const integration::QMCGrid<REAL, N1>& r1Grid, const DeviceArray<REAL>& F) {
  const int nb = static_cast<int>(r1Grid.nBasis());
  const IndexType np = static_cast<IndexType>(nb) * (nb + 1) / 2;
  DeviceArray<REAL> E(static_cast<std::size_t>(np * np));
  detail::check(tcscf_eri(r1Grid.basisValues.data(), r1Grid.quadratureGrid.weights.data(), r1Grid.nGrid(), nb,
                          F.data(), E.data(), r1Grid.stream));
  const std::vector<REAL> e = E.download();
  auto pk = [](int j, int i) { return j <= i ? static_cast<IndexType>(i) * (i + 1) / 2 + j : static_cast<IndexType>(j) * (j + 1) / 2 + i; };
  std::vector<REAL> out(static_cast<std::size_t>(nb) * nb * nb * nb);
  for (int b1 = 0; b1 < nb; ++b1)
    for (int b2 = 0; b2 < nb; ++b2)
      for (int b3 = 0; b3 < nb; ++b3)
        for (int b4 = 0; b4 < nb; ++b4)
          out[((static_cast<std::size_t>(b1) * nb + b2) * nb + b3) * nb + b4] = e[pk(b1, b3) * np + pk(b2, b4)];
  return out;
}

/* ------------------------------------------- per-element QMC route (AtomicBasis) */
namespace detail {
/* fillTwoElectronSymmetricHermitianArray (src/mathFunctions.hpp:367-433): value
 * on the canonical quadruples only, copied to their images; A (nb^4) accumulates. */
template <class F>
void fillTwoElectron(int N, bool real, F value, std::vector<std::complex<double>>& A) {
  auto at = [&](int a, int b, int c, int d) -> std::complex<double>& {
    return A[((static_cast<std::size_t>(a) * N + b) * N + c) * N + d];
  };
  for (int a = 0; a < N; ++a)
    for (int b = a; b < N; ++b)
      for (int c = a; c < N; ++c)
        for (int d = real ? b : a; d < N; ++d) {
          const std::complex<double> v = value(a, b, c, d), vc = std::conj(v);
          const bool sym = a != b, herm = a != c && a != d, rl = real && d != b;
          at(a, b, c, d) += v;
          if (sym) at(b, a, d, c) += v;
          if (herm || (real && a != c)) at(c, d, a, b) += vc;
          if (herm && sym) at(d, c, b, a) += vc;
          if (rl) at(a, d, c, b) += v;
          if (rl && sym) at(d, a, b, c) += v;
          if (rl && herm) at(c, b, a, d) += v;
          if (rl && herm && sym) at(b, c, d, a) += v;
        }
}

/* Every quadruple of tcscf_eri_qmc6d (ndim 6, real) or tcscf_eri_qmc5d_atomic
 * (ndim 5, the reference's complex functions) for basis f, as a complex nb^4
 * array; for ndim 5 missing -m partners are tabulated alongside and dropped. */
inline std::vector<std::complex<double>> qmcQuadruples(std::vector<tcscf_basis_fn> f, int ndim, IndexType minPoints,
                                                       int nshifts, std::uint64_t seed, hipStream_t stream) {
  const int n0 = static_cast<int>(f.size());
  std::vector<int> partner(f.size());
  if (ndim == 5) {
    auto same = [](const tcscf_basis_fn& x, const tcscf_basis_fn& y, int m) {
      return x.kind == y.kind && x.n == y.n && x.l == y.l && y.m == m && x.alpha == y.alpha;
    };
    for (std::size_t a = 0; a < f.size(); ++a) {
      if (f[a].m == 0) {
        partner[a] = static_cast<int>(a);
        continue;
      }
      std::size_t q = 0;
      while (q < f.size() && !same(f[a], f[q], -f[a].m)) ++q;
      if (q == f.size()) {
        tcscf_basis_fn g = f[a];
        g.m = -g.m;
        f.push_back(g);
        partner.push_back(0);
      }
      partner[a] = static_cast<int>(q);
    }
    for (std::size_t a = static_cast<std::size_t>(n0); a < f.size(); ++a) {
      std::size_t q = 0;
      while (!same(f[a], f[q], -f[a].m)) ++q;
      partner[a] = static_cast<int>(q);
    }
  }
  const int nb = static_cast<int>(f.size());
  const int64_t n = tcscf_qmc_lattice_npts(minPoints, nshifts);
  if (n < 0) check(static_cast<int>(n));
  std::vector<double> p1(3 * n), p2(3 * n), w(n);
  const int64_t got = tcscf_qmc_spherical_points(minPoints, ndim, nshifts, seed, p1.data(), p2.data(), w.data());
  if (got < 0) check(static_cast<int>(got));
  DeviceArray<double> P1, P2, W, A1(static_cast<std::size_t>(n * nb)), A2(static_cast<std::size_t>(n * nb));
  P1.upload(p1.data(), p1.size());
  P2.upload(p2.data(), p2.size());
  W.upload(w.data(), w.size());
  check(tcscf_basis_eval_fns(P1.data(), n, f.data(), nb, A1.data(), nullptr, stream));
  check(tcscf_basis_eval_fns(P2.data(), n, f.data(), nb, A2.data(), nullptr, stream));
  std::vector<std::complex<double>> out(static_cast<std::size_t>(n0) * n0 * n0 * n0);
  auto o = [&](int a, int b, int c, int d) -> std::complex<double>& {
    return out[((static_cast<std::size_t>(a) * n0 + b) * n0 + c) * n0 + d];
  };
  if (ndim == 6) {
    const IndexType np = static_cast<IndexType>(nb) * (nb + 1) / 2;
    DeviceArray<double> E(static_cast<std::size_t>(np * np));
    check(tcscf_eri_qmc6d(P1.data(), P2.data(), W.data(), n, A1.data(), A2.data(), nb, E.data(), stream));
    const std::vector<double> e = E.download();
    auto pk = [](int j, int i) { return j <= i ? static_cast<IndexType>(i) * (i + 1) / 2 + j : static_cast<IndexType>(j) * (j + 1) / 2 + i; };
    for (int a = 0; a < n0; ++a)
      for (int b = 0; b < n0; ++b)
        for (int c = 0; c < n0; ++c)
          for (int d = 0; d < n0; ++d) o(a, b, c, d) = e[pk(a, c) * np + pk(b, d)];
    return out;
  }
  std::vector<int> ms(f.size());
  for (std::size_t a = 0; a < f.size(); ++a) ms[a] = f[a].m;
  DeviceArray<double> E(static_cast<std::size_t>(nb) * nb * nb * nb * 2);
  check(tcscf_eri_qmc5d_atomic(P1.data(), P2.data(), W.data(), n, A1.data(), A2.data(), nb, ms.data(), partner.data(),
                               E.data(), stream));
  const std::vector<double> e = E.download();
  for (int a = 0; a < n0; ++a)
    for (int b = 0; b < n0; ++b)
      for (int c = 0; c < n0; ++c)
        for (int d = 0; d < n0; ++d) {
          const std::size_t k = ((static_cast<std::size_t>(a) * nb + b) * nb + c) * nb + d;
          o(a, b, c, d) = {e[2 * k], e[2 * k + 1]};
        }
  return out;
}
}  // namespace detail

/* fillR12Array (src/AtomicBasis.hpp:167-177) with r12MatrixElement (:30-48):
 * every (ab|cd) as one 6-D QMC integral over the point set of
 * tcscf_qmc_spherical_points(ndim = 6), real basis, filled as the reference
 * fills it (canonical quadruples copied to their images); nb^4 row-major. */
template <class B>
std::vector<double> fillR12Array(const std::vector<B>& basis, IndexType minPoints, int nshifts = 2,
                                 std::uint64_t seed = 0, hipStream_t stream = nullptr) {
  const auto E = detail::qmcQuadruples(detail::abi(basis), 6, minPoints, nshifts, seed, stream);
  const int N = static_cast<int>(basis.size());
  std::vector<std::complex<double>> A(E.size());
  detail::fillTwoElectron(N, true, [&](int a, int b, int c, int d) {
    return E[((static_cast<std::size_t>(a) * N + b) * N + c) * N + d]; }, A);
  std::vector<double> out(A.size());
  for (std::size_t k = 0; k < A.size(); ++k) out[k] = A[k].real();
  return out;
}

/* fillAtomicR12Array (src/AtomicBasis.hpp:197-248): the 5-D Coulomb integrals
 * ((l, m) of a, c and of b, d equal) and, elsewhere, the exchange integrals
 * ((l, m) of a, d and of b, c equal) of the reference's complex functions,
 * 0 otherwise; complex nb^4 row-major. */
template <class B>
std::vector<std::complex<double>> fillAtomicR12Array(const std::vector<B>& basis, IndexType minPoints,
                                                     int nshifts = 2, std::uint64_t seed = 0,
                                                     hipStream_t stream = nullptr) {
  const auto E = detail::qmcQuadruples(detail::abi(basis), 5, minPoints, nshifts, seed, stream);
  const int N = static_cast<int>(basis.size());
  auto e = [&](int a, int b, int c, int d) { return E[((static_cast<std::size_t>(a) * N + b) * N + c) * N + d]; };
  auto lm = [&](int a, int b) { return basis[a].l == basis[b].l && basis[a].m == basis[b].m; };
  std::vector<std::complex<double>> A(E.size());
  detail::fillTwoElectron(N, true, [&](int a, int b, int c, int d) {
    return (lm(a, c) && lm(b, d)) ? std::complex<double>(e(a, b, c, d).real()) : std::complex<double>(0.0); }, A);
  detail::fillTwoElectron(N, false, [&](int a, int b, int c, int d) {
    return (lm(a, d) && lm(b, c) && !(lm(a, c) && lm(b, d))) ? e(a, b, c, d) : std::complex<double>(0.0); }, A);
  return A;
}

/* atomicCoulombOperator (src/AtomicBasis.hpp:59-98); b1, b3 and b2, b4 must share (l, m). */
template <class B>
double atomicCoulombOperator(const B& b1, const B& b2, const B& b3, const B& b4, IndexType minPoints,
                             int nshifts = 2, std::uint64_t seed = 0, hipStream_t stream = nullptr) {
  if (b1.l != b3.l || b1.m != b3.m || b2.l != b4.l || b2.m != b4.m)
    throw Error("atomicCoulombOperator: b1, b3 and b2, b4 need equal (l, m)");
  const auto E = detail::qmcQuadruples({b1.abi(), b2.abi(), b3.abi(), b4.abi()}, 5, minPoints, nshifts, seed, stream);
  return E[((0 * 4 + 1) * 4 + 2) * 4 + 3].real();
}

/* atomicExchangeOperator (src/AtomicBasis.hpp:103-140); b1, b4 and b2, b3 must share (l, m). */
template <class B>
std::complex<double> atomicExchangeOperator(const B& b1, const B& b2, const B& b3, const B& b4, IndexType minPoints,
                                            int nshifts = 2, std::uint64_t seed = 0, hipStream_t stream = nullptr) {
  if (b1.l != b4.l || b1.m != b4.m || b2.l != b3.l || b2.m != b3.m)
    throw Error("atomicExchangeOperator: b1, b4 and b2, b3 need equal (l, m)");
  const auto E = detail::qmcQuadruples({b1.abi(), b2.abi(), b3.abi(), b4.abi()}, 5, minPoints, nshifts, seed, stream);
  return E[((0 * 4 + 1) * 4 + 2) * 4 + 3];
}

/* threeElectronIntegrals (src/integration/integrateAll.hpp:369-442), same-spin
 * Jastrow: Q(p,q,r,s,t,u), nb^6 row-major (host); nb <= 16. */
template <class REAL, int N1, int N2>
std::vector<REAL> threeElectronIntegrals(const integration::QMCGrid<REAL, N1>& r1Grid,
                                         const integration::QMCGrid<REAL, N2>& r2Grid, const Jastrow& jastrow) {
  const int nb = static_cast<int>(r1Grid.nBasis());
  DeviceArray<REAL> V = computeV(r1Grid, r2Grid, jastrow);
  const std::size_t n6 = static_cast<std::size_t>(nb) * nb * nb * nb * nb * nb;
  DeviceArray<REAL> Q(n6);
  detail::check(tcscf_three_electron(r1Grid.basisValues.data(), r1Grid.quadratureGrid.weights.data(), r1Grid.nGrid(),
                                     V.data(), nb, jastrow.c_same, Q.data(), r1Grid.stream));
  return Q.download();
}

/* --------------------------------------------------------------- SCF solvers */
namespace detail {
inline tcscf_scf_opts opts(double tol, int maxIter, bool mask, int mgsMode, int diis) {
  return tcscf_scf_opts{tol, maxIter, mask ? 1 : 0, mgsMode, diis};
}
inline void checkSquare(const std::vector<double>& M, int nb, const char* what) {
  if (M.size() != static_cast<std::size_t>(nb) * nb) throw Error(std::string(what) + ": expected nb x nb");
}
}  // namespace detail

/* RCSHartreeFock (src/HartreeFock.hpp:233-293, src/HartreeFock.cpp:785-886). */
template <class T>
class RCSHartreeFock {
 public:
  RCSHartreeFock(int nElectrons, int nBasis) : nElectrons_(nElectrons), basisSize(nBasis) {
    if (nElectrons % 2 != 0) throw Error("RCSHartreeFock: the number of electrons must be even");
    density.assign(static_cast<std::size_t>(nBasis) * nBasis, 0);
    fockOperator.assign(static_cast<std::size_t>(nBasis) * nBasis, 0);
    eigenvalues.assign(nBasis, 0);
    eigenvectors.assign(static_cast<std::size_t>(nBasis) * nBasis, 0);
  }
  static constexpr bool needsGradients() { return false; }

  template <int N1, class FT>
  T compute(const std::vector<double>& overlap, const std::vector<double>& oneElectronTerms,
            const integration::QMCGrid<T, N1>& r1Grid, const DeviceArray<FT>& Fji, bool respectOneElectronSymmetry,
            double tol = 1e-5, int maxIter = 50, int diis = 0) {
    detail::checkSquare(overlap, basisSize, "overlap");
    detail::checkSquare(oneElectronTerms, basisSize, "oneElectronTerms");
    tcscf_scf_opts o = detail::opts(tol, maxIter, respectOneElectronSymmetry, 0, diis);
    tcscf_scf_result res{};
    energyTrace.assign(maxIter, 0);
    tcscf_scf_observables obs{fockOperator.data(), eigenvectors.data(), nullptr};
    const int rc = tcscf_rcs_scf(nElectrons_, basisSize, overlap.data(), oneElectronTerms.data(), r1Grid.basisValues.data(),
                                 r1Grid.quadratureGrid.weights.data(), r1Grid.nGrid(), Fji.data(), detail::dtype<FT>(), &o, &res,
                                 energyTrace.data(), density.data(), eigenvalues.data(), &obs, r1Grid.stream);
    iter_ = res.iterations;
    energyTrace.resize(iter_ > 0 ? iter_ : 0);
    detail::check(rc);
    return res.energy;
  }
  int numberOfConvergenceLoops() const { return iter_; }
  T highestOccupiedOrbitalEnergy() const { return eigenvalues[nElectrons_ / 2 - 1]; }

  std::vector<T> density;       /* nb x nb (one spin) */
  std::vector<T> fockOperator;  /* nb x nb: (j, i), built from the converged density */
  std::vector<T> eigenvalues;   /* ascending */
  std::vector<T> eigenvectors;  /* nb x nb: (j, a), column a = eigenvector a */
  std::vector<T> energyTrace;

 private:
  int nElectrons_;

 public:
  const int basisSize;

 private:
  int iter_ = 0;
};

/* UOSHartreeFock (src/HartreeFock.hpp:296-350, src/HartreeFock.cpp:890-1003). */
template <class T>
class UOSHartreeFock {
 public:
  UOSHartreeFock(int nUp, int nDown, int nBasis) : nElectrons{nUp, nDown}, basisSize(nBasis) {
    density.assign(2 * static_cast<std::size_t>(nBasis) * nBasis, 0);
    fockOperator.assign(2 * static_cast<std::size_t>(nBasis) * nBasis, 0);
    eigenvalues.assign(2 * static_cast<std::size_t>(nBasis), 0);
    eigenvectors.assign(2 * static_cast<std::size_t>(nBasis) * nBasis, 0);
  }
  static constexpr bool needsGradients() { return false; }

  template <int N1, class FT>
  T compute(const std::vector<double>& overlap, const std::vector<double>& oneElectronTerms,
            const integration::QMCGrid<T, N1>& r1Grid, const DeviceArray<FT>& Fji, bool respectOneElectronSymmetry,
            double tol = 1e-5, int maxIter = 100, int diis = 0) {
    detail::checkSquare(overlap, basisSize, "overlap");
    detail::checkSquare(oneElectronTerms, basisSize, "oneElectronTerms");
    tcscf_scf_opts o = detail::opts(tol, maxIter, respectOneElectronSymmetry, 0, diis);
    tcscf_scf_result res{};
    energyTrace.assign(maxIter, 0);
    tcscf_scf_observables obs{fockOperator.data(), eigenvectors.data(), nullptr};
    const int rc = tcscf_uos_scf(nElectrons[0], nElectrons[1], basisSize, overlap.data(), oneElectronTerms.data(),
                                 r1Grid.basisValues.data(), r1Grid.quadratureGrid.weights.data(), r1Grid.nGrid(),
                                 Fji.data(), detail::dtype<FT>(), &o, &res, energyTrace.data(), density.data(),
                                 eigenvalues.data(), &obs, r1Grid.stream);
    iter_ = res.iterations;
    energyTrace.resize(iter_ > 0 ? iter_ : 0);
    detail::check(rc);
    return res.energy;
  }
  int numberOfConvergenceLoops() const { return iter_; }
  T highestOccupiedOrbitalEnergy() const {
    T e = -INFINITY;
    for (int s = 0; s < 2; ++s)
      if (nElectrons[s] > 0) e = std::max(e, eigenvalues[static_cast<std::size_t>(s) * basisSize + nElectrons[s] - 1]);
    return e;
  }

  const int nElectrons[2];
  const int basisSize;
  std::vector<T> density;       /* 2 x nb x nb */
  std::vector<T> fockOperator;  /* 2 x nb x nb: (spin, j, i) */
  std::vector<T> eigenvalues;   /* 2 x nb, ascending per spin */
  std::vector<T> eigenvectors;  /* 2 x nb x nb: (spin, j, a) */
  std::vector<T> energyTrace;

 private:
  int iter_ = 0;
};

/* TCHartreeFock (src/HartreeFock.hpp:353-418, src/HartreeFock.cpp:1007-1362).
 * compute takes VjiSame and VjiOppo as the reference does; the overload with
 * one c = 1 tensor V and a Jastrow's c_same / c_oppo (V_same = c_same V,
 * V_oppo = c_oppo V) is the same Fock build reading one tensor. */
template <class T>
class TCHartreeFock {
 public:
  TCHartreeFock(int nUp, int nDown, int nBasis) : nElectrons{nUp, nDown}, basisSize(nBasis) {
    density.assign(2 * static_cast<std::size_t>(nBasis) * nBasis, 0);
    fockOperator.assign(2 * static_cast<std::size_t>(nBasis) * nBasis, 0);
    occupiedOrbitalPseudoEnergy.assign(static_cast<std::size_t>(nUp + nDown), 0);
    occupiedOrbitals[0].assign(static_cast<std::size_t>(nBasis) * nUp, 0);
    occupiedOrbitals[1].assign(static_cast<std::size_t>(nBasis) * nDown, 0);
  }
  static constexpr bool needsGradients() { return true; }

  /* The reference's signature (src/HartreeFock.hpp:386-394): F_same / F_oppo
   * and V_same / V_oppo, e.g. from computeF / computeV with an OchiJastrow for
   * each spin.  mgsMode 0: the reference's Gram-Schmidt (Euclidean norm after
   * the first vector, src/orthogonalization.hpp:152); 1: the overlap metric
   * throughout.  FT = float: FP32 storage of F / V (FP64 arithmetic). */
  template <int N1, class FT>
  T compute(const std::vector<double>& overlap, const std::vector<double>& oneElectronTerms,
            const integration::QMCGrid<T, N1>& r1Grid, const DeviceArray<FT>& FjiSame, const DeviceArray<FT>& FjiOppo,
            const DeviceArray<FT>& VjiSame, const DeviceArray<FT>& VjiOppo, bool respectOneElectronSymmetry,
            double tol = 1e-5, int maxIter = 200, int mgsMode = 0, int diis = 0) {
    return run(overlap, oneElectronTerms, r1Grid, FjiSame, FjiOppo, VjiSame, &VjiOppo, nullptr,
               respectOneElectronSymmetry, tol, maxIter, mgsMode, diis);
  }
  /* One c = 1 tensor Vji (computeV with a Jastrow) and the Jastrow's scales. */
  template <int N1, class FT>
  T compute(const std::vector<double>& overlap, const std::vector<double>& oneElectronTerms,
            const integration::QMCGrid<T, N1>& r1Grid, const DeviceArray<FT>& FjiSame, const DeviceArray<FT>& FjiOppo,
            const DeviceArray<FT>& Vji, const Jastrow& jastrow, bool respectOneElectronSymmetry, double tol = 1e-5,
            int maxIter = 200, int mgsMode = 0, int diis = 0) {
    const tcscf_jastrow j = jastrow.abi();
    return run(overlap, oneElectronTerms, r1Grid, FjiSame, FjiOppo, Vji, static_cast<const DeviceArray<FT>*>(nullptr), &j,
               respectOneElectronSymmetry, tol, maxIter, mgsMode, diis);
  }
  /* With a general Ochi Jastrow whose opposite-spin coefficients are one
   * multiple rho of the same-spin ones: Vji is VjiSame and VjiOppo = rho VjiSame
   * (OchiJastrow::tcScales); otherwise pass both tensors. */
  template <int N1, class FT>
  T compute(const std::vector<double>& overlap, const std::vector<double>& oneElectronTerms,
            const integration::QMCGrid<T, N1>& r1Grid, const DeviceArray<FT>& FjiSame, const DeviceArray<FT>& FjiOppo,
            const DeviceArray<FT>& VjiSame, const OchiJastrow& jastrow, bool respectOneElectronSymmetry,
            double tol = 1e-5, int maxIter = 200, int mgsMode = 0, int diis = 0) {
    return compute(overlap, oneElectronTerms, r1Grid, FjiSame, FjiOppo, VjiSame, jastrow.tcScales(),
                   respectOneElectronSymmetry, tol, maxIter, mgsMode, diis);
  }
  int numberOfConvergenceLoops() const { return iter_; }
  T highestOccupiedOrbitalEnergy() const {
    T e = -INFINITY;
    if (nElectrons[0] > 0) e = std::max(e, occupiedOrbitalPseudoEnergy[nElectrons[0] - 1]);
    if (nElectrons[1] > 0) e = std::max(e, occupiedOrbitalPseudoEnergy[nElectrons[0] + nElectrons[1] - 1]);
    return e;
  }

  const int nElectrons[2];
  const int basisSize;
  std::vector<T> density;                      /* 2 x nb x nb */
  std::vector<T> fockOperator;                 /* 2 x nb x nb: (spin, j, i), H + G2 - G3 */
  std::vector<T> occupiedOrbitalPseudoEnergy;  /* nUp then nDown, ascending Re(lambda) */
  std::vector<T> occupiedOrbitals[2];          /* per spin nb x nocc_s: (j, b), S-metric Gram-Schmidt */
  std::vector<T> energyTrace;

 private:
  template <int N1, class FT>
  T run(const std::vector<double>& overlap, const std::vector<double>& oneElectronTerms,
        const integration::QMCGrid<T, N1>& r1Grid, const DeviceArray<FT>& FjiSame, const DeviceArray<FT>& FjiOppo,
        const DeviceArray<FT>& V, const DeviceArray<FT>* Vo, const tcscf_jastrow* j, bool respectOneElectronSymmetry,
        double tol, int maxIter, int mgsMode, int diis) {
    detail::checkSquare(overlap, basisSize, "overlap");
    detail::checkSquare(oneElectronTerms, basisSize, "oneElectronTerms");
    if (!r1Grid.gradientsStored()) throw Error("TCHartreeFock: the r1 grid needs basis gradients");
    const std::size_t np = static_cast<std::size_t>(basisSize) * (basisSize + 1) / 2, n1 = r1Grid.nGrid();
    if (FjiSame.size() != n1 * np || FjiOppo.size() != n1 * np || V.size() != 3 * n1 * np ||
        (Vo && Vo->size() != 3 * n1 * np))
      throw Error("TCHartreeFock: F must be n1 x npair and V n1 x 3 x npair (packed)");
    tcscf_scf_opts o = detail::opts(tol, maxIter, respectOneElectronSymmetry, mgsMode, diis);
    tcscf_scf_result res{};
    energyTrace.assign(maxIter, 0);
    const int nmax = std::max(1, std::max(nElectrons[0], nElectrons[1]));
    std::vector<T> occ(2 * static_cast<std::size_t>(basisSize) * nmax, 0);
    tcscf_scf_observables obs{fockOperator.data(), nullptr, occ.data()};
    const int rc = tcscf_tc_scf(nElectrons[0], nElectrons[1], basisSize, overlap.data(), oneElectronTerms.data(),
                                r1Grid.basisValues.data(), r1Grid.basisGradients.data(),
                                r1Grid.quadratureGrid.weights.data(), r1Grid.nGrid(), FjiSame.data(), FjiOppo.data(),
                                V.data(), Vo ? Vo->data() : nullptr, detail::dtype<FT>(), j, &o, &res,
                                energyTrace.data(), density.data(), occupiedOrbitalPseudoEnergy.data(), &obs,
                                r1Grid.stream);
    for (int s = 0; s < 2; ++s)
      for (int jj = 0; jj < basisSize; ++jj)
        for (int b = 0; b < nElectrons[s]; ++b)
          occupiedOrbitals[s][static_cast<std::size_t>(jj) * nElectrons[s] + b] =
              occ[(static_cast<std::size_t>(s) * basisSize + jj) * nmax + b];
    iter_ = res.iterations;
    energyTrace.resize(iter_ > 0 ? iter_ : 0);
    detail::check(rc);
    return res.energy;
  }
  int iter_ = 0;
};

}  // namespace tcscf
