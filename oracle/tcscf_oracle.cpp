/*
 * tcscf_oracle.cpp -- FP64 CPU restatement of the tcscf grid-integral path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity checker for the HIP
 * library in tcscf_amd/ and the timed CPU baseline in bench.py.  Nothing in the
 * product links, loads or calls it.  Every function cites the reference
 * (corbett5/tcscf) source it restates.
 *
 * Conventions shared with the product (so both see the same inputs):
 *   - grid points are SoA: pts[0*n + k] = x_k, pts[1*n + k] = y_k, pts[2*n + k] = z_k;
 *     a "2-D" grid (the reference's QMCGrid<REAL,2>) is passed with y == 0;
 *   - grid weights already contain every Jacobian (r^2 sin(theta) etc.);
 *   - basis tables are row-major N x nb (basis index fastest), gradients N x nb x 3;
 *   - packed pair index p(j,i) for j <= i:  p = i*(i+1)/2 + j.
 *
 * Parity pins (tests/ hold the checks):
 *   grids           -> reference KATs (IntegrationTest.cpp), sum(w)=4pi, Y_lm orthonormality
 *   STO values/S/H  -> reference KAT (SlaterTypeOrbitalTest.cpp), closed forms
 *   STO gradients   -> central differences (no reference test exists)
 *   Jastrow         -> reference KAT (gradientsTest.cpp), closed form of the r12 term
 *   computeF        -> closed-form 1s Coulomb integrals, brute force at tiny N
 *   SCF drivers     -> invariants (Tr DS, idempotency, UOS==RCS closed shell,
 *                      TC(c=0)==UOS), He single-zeta closed form
 *   eigen-solvers   -> residual checks; two independent algorithms
 *   threeElectron/computeQ/assembleG -> cross-checked against each other only
 *   TC two-body vector term -> the Fock matrix of the operator grad_1 u . grad_1
 *                      from its four-index integrals over point pairs (grad u
 *                      from the Jastrow's closed form), direct minus exchange
 */
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#include <omp.h>

namespace {

using cplx = std::complex<double>;
constexpr double kPi = 3.141592653589793238462643383279502884;
constexpr double kLn2 = 0.693147180559945309417232121458176568;

#include "../tcscf_amd/data/lebedev_tables.inc"

thread_local std::string g_err;

/* Planted mistakes for the tests that show a check can fail (0 = none).  Set
 * only by orc_set_plant; read by tc_fock_build. */
int g_plant = 0;

inline int64_t pidx(int j, int i) { return (int64_t)i * (i + 1) / 2 + j; }  // j <= i

double factorial(int n) { return std::tgamma(n + 1.0); }

/* ------------------------------------------------------------------------ */
/* Grids                                                                     */
/* ------------------------------------------------------------------------ */

/* src/integration/ChebyshevGauss.hpp:32-45 (node i: theta=(i+1)pi/(n+1), x=cos,
 * w=sin) with the grid weight pi/(n+1) folded in as createGrid does
 * (src/integration/quadrature.hpp:98). */
void chebyshev_gauss(int n, double* x, double* w) {
  double const freq = kPi / (n + 1);
  for (int i = 0; i < n; ++i) {
    double const theta = (i + 1) * freq;
    x[i] = std::cos(theta);
    w[i] = std::sin(theta) * freq;
  }
}

/* Radial maps.
 *  kind 0: TreutlerAhlrichsM4(a, eps)  src/integration/changeOfVariables.hpp:126-162
 *  kind 1: r = (1+x)/(1-x), dr/dx = 2/(1-x)^2 (Becke 1988; the reference's
 *          XOver1mX, changeOfVariables.hpp:100-120, after x -> (1+x)/2)
 *  kind 2: XOver1mX  r = x/(1-x), dr/dx = (1-x)^-2 */
void radial_map(int kind, double a, double eps, double x, double* r, double* jac) {
  if (kind == 0) {
    double const e = eps / kLn2;
    double const alpha = 0.6;
    *r = e * std::pow(x + a, alpha) * std::log((1 + a) / (1 - x));
    double const logPart = std::log((1 + a) / (1 - x));
    double const p1 = std::pow(x + a, alpha - 1);
    double const p = p1 * (x + a);
    *jac = e * (alpha * p1 * logPart + p / (1 - x));
  } else if (kind == 1) {
    *r = (1 + x) / (1 - x);
    *jac = 2.0 / ((1 - x) * (1 - x));
  } else {
    *r = x / (1 - x);
    *jac = std::pow(1 - x, -2);
  }
}

const LebedevEntry* find_lebedev(int order) {
  for (int k = 0; k < kLebedevTableSize; ++k)
    if (kLebedevTable[k].order == order) return &kLebedevTable[k];
  return nullptr;
}

}  // namespace

/* ======================================================================== */
/* C ABI                                                                     */
/* ======================================================================== */
extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

void orc_set_threads(int n) { if (n > 0) omp_set_num_threads(n); }
void orc_set_plant(int p) { g_plant = p; }
int orc_get_threads() { return omp_get_max_threads(); }

void orc_chebyshev_gauss(int n, double* x, double* w) { chebyshev_gauss(n, x, w); }

/* createGrid(ChebyshevGauss(n), map) -> radial nodes and weights
 * (src/integration/quadrature.hpp:76-102). */
void orc_radial_grid(int n, int kind, double a, double eps, double* r, double* w) {
  std::vector<double> x(n), wx(n);
  chebyshev_gauss(n, x.data(), wx.data());
  for (int i = 0; i < n; ++i) {
    double rr, jac;
    radial_map(kind, a, eps, x[i], &rr, &jac);
    r[i] = rr;
    w[i] = wx[i] * jac;
  }
}

int orc_lebedev_npts(int order) {
  const LebedevEntry* e = find_lebedev(order);
  return e ? e->npts : -1;
}

/* getLebedevGrid (src/integration/Lebedev.cpp:23-41): {theta, phi, w}. */
int orc_lebedev(int order, double* theta, double* phi, double* w) {
  const LebedevEntry* e = find_lebedev(order);
  if (!e) { g_err = "Lebedev order not available"; return -1; }
  for (int k = 0; k < e->npts; ++k) {
    theta[k] = e->data[3 * k + 0];
    phi[k] = e->data[3 * k + 1];
    w[k] = e->data[3 * k + 2];
  }
  return e->npts;
}

/* Radial x Lebedev product grid packed in the QMCGrid shape
 * (src/integration/integrateAll.hpp:41-64: Cartesian points, weight *= r^2;
 * the Lebedev weight already carries the solid angle).  Point order follows
 * the reference's product integrator, angular outer / radial inner
 * (src/integration/TreutlerAhlrichsLebedev.hpp:158-167): k = ia*nrad + ir. */
int orc_product_grid(int nrad, int kind, double a, double eps, int order, double* pts, double* w) {
  const LebedevEntry* e = find_lebedev(order);
  if (!e) { g_err = "Lebedev order not available"; return -1; }
  std::vector<double> r(nrad), wr(nrad);
  orc_radial_grid(nrad, kind, a, eps, r.data(), wr.data());
  int64_t const n = (int64_t)nrad * e->npts;
  for (int ia = 0; ia < e->npts; ++ia) {
    double const th = e->data[3 * ia], ph = e->data[3 * ia + 1], wa = e->data[3 * ia + 2];
    for (int ir = 0; ir < nrad; ++ir) {
      int64_t const k = (int64_t)ia * nrad + ir;
      /* Spherical::x/y/z, src/mathFunctions.hpp:213-220 */
      pts[0 * n + k] = r[ir] * std::cos(ph) * std::sin(th);
      pts[1 * n + k] = r[ir] * std::sin(ph) * std::sin(th);
      pts[2 * n + k] = r[ir] * std::cos(th);
      w[k] = wr[ir] * wa * r[ir] * r[ir];
    }
  }
  return (int)n;
}

/* ------------------------------------------------------------------------ */
/* Slater-type orbitals                                                      */
/* ------------------------------------------------------------------------ */

/* fns: 4 doubles per function {n, l, m, zeta}.
 * complex values, reference form: N r^(n-1) e^(-zeta r) sph_legendre(l,|m|,theta) e^(i m phi)
 *   src/SlaterTypeOrbital.hpp:27-58, src/mathFunctions.hpp:126-139.
 * gradient: the reference's spherical-basis formula, src/SlaterTypeOrbital.hpp:63-106;
 *   on the z axis, where that formula is 0/0 (NaN in the reference), its limit.
 * out: values N x nb complex (interleaved), grads N x nb x 3 complex or NULL.  */
void orc_sto_complex(const double* pts, int64_t n, const double* fns, int nb, double* values, double* grads) {
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < n; ++k) {
    double const x = pts[k], y = pts[n + k], z = pts[2 * n + k];
    double const r = std::hypot(x, y, z);
    double const theta = std::acos(z / r);
    double const phi = std::atan2(y, x);
    for (int b = 0; b < nb; ++b) {
      int const nq = (int)fns[4 * b], l = (int)fns[4 * b + 1], m = (int)fns[4 * b + 2];
      double const alpha = fns[4 * b + 3];
      double const norm = std::pow(2 * alpha, nq) * std::sqrt(2 * alpha / factorial(2 * nq));
      double const radial = norm * std::pow(r, nq - 1) * std::exp(-alpha * r);
      cplx const ylm = std::sph_legendre(l, std::abs(m), theta) * cplx(std::cos(m * phi), std::sin(m * phi));
      cplx const v = radial * ylm;
      values[2 * (k * nb + b)] = v.real();
      values[2 * (k * nb + b) + 1] = v.imag();
      if (grads) {
        double const dRadial = ((nq - 1) / r - alpha) * radial;
        cplx ylmp1 = 0;
        if (l != m) {
          ylmp1 = std::sqrt((double)((l - m) * (l + m + 1))) * cplx(std::cos(-phi), std::sin(-phi)) *
                  (std::sph_legendre(l, std::abs(m + 1), theta) *
                   cplx(std::cos((m + 1) * phi), std::sin((m + 1) * phi)));
        }
        cplx const rHat = dRadial * ylm;
        cplx thetaHat, phiHat;
        if (std::fabs(std::sin(theta)) > 1e-12) {
          thetaHat = ((double)m * ylm / std::tan(theta) + ylmp1) * radial / r;
          phiHat = cplx(0, 1) * (double)m * ylm * radial / (r * std::sin(theta));
        } else {
          /* The reference's formula is 0/0 on the z axis (Lebedev grids have
           * points there; at theta = pi, sph_legendre returns an exact 0 and the
           * term is lost).  Its limit: Y_lm / sin(theta) -> 0 for |m| != 1 and,
           * for |m| = 1, -sqrt((2l+1)/(4 pi) (l-1)!/(l+1)!) P_l'(cos theta) e^{i m phi}
           * with P_l'(1) = l(l+1)/2. */
          cplx yos = 0;
          if (std::abs(m) == 1) {
            double const pl = 0.5 * l * (l + 1) * (std::cos(theta) > 0 ? 1.0 : ((l + 1) % 2 == 0 ? 1.0 : -1.0));
            yos = -std::sqrt((2 * l + 1) / (4 * kPi) * factorial(l - 1) / factorial(l + 1)) * pl *
                  cplx(std::cos(m * phi), std::sin(m * phi));
          }
          thetaHat = ((double)m * std::cos(theta) * yos + ylmp1) * radial / r;
          phiHat = cplx(0, 1) * (double)m * yos * radial / r;
        }
        double const rc[3] = {std::cos(phi) * std::sin(theta), std::sin(phi) * std::sin(theta), std::cos(theta)};
        double const tc[3] = {std::cos(phi) * std::cos(theta), std::sin(phi) * std::cos(theta), -std::sin(theta)};
        double const pc[3] = {-std::sin(phi), std::cos(phi), 0};
        for (int d = 0; d < 3; ++d) {
          cplx const g = rHat * rc[d] + thetaHat * tc[d] + phiHat * pc[d];
          grads[2 * ((k * nb + b) * 3 + d)] = g.real();
          grads[2 * ((k * nb + b) * 3 + d) + 1] = g.imag();
        }
      }
    }
  }
}

/* Real STOs: the unitary combination of the reference's complex functions,
 *   m > 0:  (-1)^m sqrt2 Re phi(l, m)      m < 0:  (-1)^m sqrt2 Im phi(l, |m|)
 * i.e. the standard real spherical harmonics without the Condon-Shortley
 * phase (p_{+1} ~ x, p_{-1} ~ y, p_0 ~ z).  Values and gradients (from the
 * reference's gradient formula) are formed from orc_sto_complex. */
void orc_sto_real(const double* pts, int64_t n, const double* fns, int nb, double* values, double* grads) {
  std::vector<double> f2(4 * nb);
  for (int b = 0; b < nb; ++b) {
    f2[4 * b] = fns[4 * b];
    f2[4 * b + 1] = fns[4 * b + 1];
    f2[4 * b + 2] = std::abs(fns[4 * b + 2]);
    f2[4 * b + 3] = fns[4 * b + 3];
  }
  int64_t const chunk = 4096;
  for (int64_t k0 = 0; k0 < n; k0 += chunk) {
    int64_t const cn = std::min(chunk, n - k0);
    std::vector<double> p(3 * cn), cv(2 * cn * nb), cg(grads ? 6 * cn * nb : 0);
    for (int d = 0; d < 3; ++d)
      for (int64_t k = 0; k < cn; ++k) p[d * cn + k] = pts[d * n + k0 + k];
    orc_sto_complex(p.data(), cn, f2.data(), nb, cv.data(), grads ? cg.data() : nullptr);
    for (int64_t k = 0; k < cn; ++k) {
      for (int b = 0; b < nb; ++b) {
        int const m = (int)fns[4 * b + 2];
        double const s = (m % 2 == 0) ? 1.0 : -1.0;
        double f = 1.0;
        int part = 0;
        if (m > 0) { f = s * std::sqrt(2.0); part = 0; }
        if (m < 0) { f = s * std::sqrt(2.0); part = 1; }
        values[(k0 + k) * nb + b] = f * cv[2 * (k * nb + b) + part];
        if (grads)
          for (int d = 0; d < 3; ++d) grads[((k0 + k) * nb + b) * 3 + d] = f * cg[2 * ((k * nb + b) * 3 + d) + part];
      }
    }
  }
}

/* overlap / coreMatrixElement, src/SlaterTypeOrbital.hpp:125-171.  S, H are
 * nb x nb row-major. */
void orc_sto_S_H(const double* fns, int nb, int Z, double* S, double* H) {
  auto nn = [](int sn, double sa) { return std::tgamma(sn + 1.0) / std::pow(sa, sn + 1); };
  for (int i = 0; i < nb; ++i)
    for (int j = 0; j < nb; ++j) {
      int const n1 = (int)fns[4 * i], l1 = (int)fns[4 * i + 1], m1 = (int)fns[4 * i + 2];
      int const n2 = (int)fns[4 * j], l2 = (int)fns[4 * j + 1], m2 = (int)fns[4 * j + 2];
      double const a1 = fns[4 * i + 3], a2 = fns[4 * j + 3];
      double const N1 = std::pow(2 * a1, n1) * std::sqrt(2 * a1 / factorial(2 * n1));
      double const N2 = std::pow(2 * a2, n2) * std::sqrt(2 * a2 / factorial(2 * n2));
      if (l1 != l2 || m1 != m2) { S[i * nb + j] = 0; H[i * nb + j] = 0; continue; }
      int const nSum = n1 + n2;
      double const aSum = a1 + a2;
      S[i * nb + j] = N1 * N2 * nn(nSum, aSum);
      H[i * nb + j] = 0.5 * N1 * N2 *
                      ((l2 * (l2 + 1) - n2 * (n2 - 1)) * nn(nSum - 2, aSum) + 2 * (n2 * a2 - Z) * nn(nSum - 1, aSum) -
                       a2 * a2 * nn(nSum, aSum));
    }
}

/* ------------------------------------------------------------------------ */
/* Ochi Jastrow factor, src/jastrowFunctions.hpp:9-115                       */
/* prm: {a, a12, nterms, then per term: i, j, k, c_oppo, c_same}             */
/* ------------------------------------------------------------------------ */
double orc_ochi_value(const double* prm, double r1, double r2, double r12, int same) {
  double const a = prm[0], a12 = prm[1];
  int const nt = (int)prm[2];
  double const r12T = r12 / (r12 + a12), r1T = r1 / (r1 + a), r2T = r2 / (r2 + a);
  double result = 0;
  for (int t = 0; t < nt; ++t) {
    const double* q = prm + 3 + 5 * t;
    double const c = same ? q[4] : q[3];
    result = result + c * std::pow(r12T, (int)q[0]) * std::pow(r1T, (int)q[1]) * std::pow(r2T, (int)q[2]);
  }
  return result;
}

void orc_ochi_gradient(const double* prm, const double* r1c, const double* r2c, int same, double* g) {
  double const a = prm[0], a12 = prm[1];
  int const nt = (int)prm[2];
  double const d[3] = {r1c[0] - r2c[0], r1c[1] - r2c[1], r1c[2] - r2c[2]};
  double const r1 = std::hypot(r1c[0], r1c[1], r1c[2]);
  double const r2 = std::hypot(r2c[0], r2c[1], r2c[2]);
  double const r12 = std::hypot(d[0], d[1], d[2]);
  double const r1T = r1 / (r1 + a), r2T = r2 / (r2 + a), r12T = r12 / (r12 + a12);
  double const r1Scale = a / (std::pow(r1, 2) * (r1 + a));
  double const r12Scale = a12 / (std::pow(r12, 2) * (r12 + a12));
  double r1Length = 0, r12Length = 0;
  for (int t = 0; t < nt; ++t) {
    const double* q = prm + 3 + 5 * t;
    int const i = (int)q[0], j = (int)q[1], k = (int)q[2];
    double const c = same ? q[4] : q[3];
    double const factor = c * std::pow(r12T, i) * std::pow(r1T, j) * std::pow(r2T, k);
    r1Length += factor * j * r1Scale;
    r12Length += factor * i * r12Scale;
  }
  for (int x = 0; x < 3; ++x) g[x] = r12Length * d[x] + r1Length * r1c[x];
}

/* ------------------------------------------------------------------------ */
/* precomputeIntegrand + computeF / computeV                                 */
/* src/HartreeFock.hpp:23-83 (Kw[n1,n2] = f(r1,r2) w2[n2]), :88-180 (F),     */
/* :185-229 (V).  The reference stores Kw for all (n1, n2); here one row of  */
/* Kw is formed per r1 point (storage only; the arithmetic is unchanged).    */
/* kernel: 0 = 1/r12                                                         */
/*         1 = 2/r12 - |grad1 u(r1,r2)|^2 - |grad1 u(r2,r1)|^2 (TC scalar)   */
/*             src/unitTests/HartreeFockTest.cpp:276-296                      */
/*         2 = grad1 u(r1,r2) (3 components), HartreeFockTest.cpp:300-316   */
/* Coincident pairs (r12 == 0) contribute 0.                                 */
/* F out: full N1 x nb x nb (ncomp = 1) or N1 x nb x nb x 3 (kernel 2),      */
/* scaled by `scale` (2 pi for the reference's 2-D r2 grid, 1 for 3-D).      */
/* ------------------------------------------------------------------------ */
static void kernel_row(int kernel, const double* prm, int same, const double* p1, int64_t n1, int64_t k1,
                       const double* p2, const double* w2, int64_t n2, double* kw) {
  double const r1c[3] = {p1[k1], p1[n1 + k1], p1[2 * n1 + k1]};
  for (int64_t k2 = 0; k2 < n2; ++k2) {
    double const r2c[3] = {p2[k2], p2[n2 + k2], p2[2 * n2 + k2]};
    double const dx = r1c[0] - r2c[0], dy = r1c[1] - r2c[1], dz = r1c[2] - r2c[2];
    double const r12 = std::hypot(dx, dy, dz);
    if (r12 == 0) {
      for (int c = 0; c < (kernel == 2 ? 3 : 1); ++c) kw[c * n2 + k2] = 0;
      continue;
    }
    if (kernel == 0) {
      kw[k2] = 1 / r12 * w2[k2];
    } else if (kernel == 1) {
      double g12[3], g21[3];
      orc_ochi_gradient(prm, r1c, r2c, same, g12);
      orc_ochi_gradient(prm, r2c, r1c, same, g21);
      double const dd = g12[0] * g12[0] + g12[1] * g12[1] + g12[2] * g12[2];
      double const ee = g21[0] * g21[0] + g21[1] * g21[1] + g21[2] * g21[2];
      kw[k2] = (2 / r12 - dd - ee) * w2[k2];
    } else {
      double g[3];
      orc_ochi_gradient(prm, r1c, r2c, same, g);
      for (int c = 0; c < 3; ++c) kw[c * n2 + k2] = g[c] * w2[k2];
    }
  }
}

/* r1 rows [row0, row1) only (the bounded CPU-baseline sample uses a
 * contiguous slice; the full call passes 0..n1). */
int orc_compute_F(int kernel, const double* prm, int same, const double* p1, int64_t n1, int64_t row0, int64_t row1,
                  const double* p2, const double* w2, int64_t n2, const double* A2, int nb, double scale, double* F) {
  int const ncomp = kernel == 2 ? 3 : 1;
#pragma omp parallel
  {
    std::vector<double> kw(ncomp * n2);
#pragma omp for schedule(dynamic, 4)
    for (int64_t k1 = row0; k1 < row1; ++k1) {
      kernel_row(kernel, prm, same, p1, n1, k1, p2, w2, n2, kw.data());
      double* Fr = F + (k1 - row0) * nb * nb * ncomp;
      for (int j = 0; j < nb; ++j)
        for (int i = j; i < nb; ++i) {
          double ans[3] = {0, 0, 0};
          for (int64_t k2 = 0; k2 < n2; ++k2) {
            double const s = A2[k2 * nb + j] * A2[k2 * nb + i];
            for (int c = 0; c < ncomp; ++c) ans[c] += s * kw[c * n2 + k2];
          }
          for (int c = 0; c < ncomp; ++c) {
            Fr[(j * nb + i) * ncomp + c] = scale * ans[c];
            Fr[(i * nb + j) * ncomp + c] = scale * ans[c];
          }
        }
    }
  }
  return 0;
}

/* Complex-basis computeF (src/HartreeFock.hpp:136-180), used only to show
 * that the real basis is a unitary change of the reference's complex one.
 * A2c: N2 x nb complex interleaved; F out complex N1 x nb x nb. */
int orc_compute_F_complex(const double* p1, int64_t n1, const double* p2, const double* w2, int64_t n2,
                          const double* A2c, int nb, double* Fc) {
  const cplx* A = reinterpret_cast<const cplx*>(A2c);
  cplx* F = reinterpret_cast<cplx*>(Fc);
#pragma omp parallel
  {
    std::vector<double> kw(n2);
#pragma omp for schedule(dynamic, 4)
    for (int64_t k1 = 0; k1 < n1; ++k1) {
      kernel_row(0, nullptr, 0, p1, n1, k1, p2, w2, n2, kw.data());
      for (int j = 0; j < nb; ++j)
        for (int i = j; i < nb; ++i) {
          cplx ans = 0;
          for (int64_t k2 = 0; k2 < n2; ++k2) ans += std::conj(A[k2 * nb + j]) * kw[k2] * A[k2 * nb + i];
          F[(k1 * nb + j) * nb + i] = ans;
          F[(k1 * nb + i) * nb + j] = std::conj(ans);
        }
    }
  }
  return 0;
}

/* integrateAllR1R2<false> (src/integration/integrateAll.hpp:221-317) given F
 * from computeF: ERI[b1,b2,b3,b4] = sum_r1 w1 A1[b1] A1[b3] F[r1,b2,b4].
 * out: nb^4 row-major. */
void orc_eri_from_F(const double* A1, const double* w1, int64_t n1, const double* F, int nb, double* eri) {
  int64_t const nb2 = (int64_t)nb * nb;
  std::fill(eri, eri + nb2 * nb2, 0.0);
#pragma omp parallel
  {
    std::vector<double> loc(nb2 * nb2, 0.0);
#pragma omp for schedule(static)
    for (int64_t k = 0; k < n1; ++k)
      for (int b1 = 0; b1 < nb; ++b1)
        for (int b3 = 0; b3 < nb; ++b3) {
          double const s = w1[k] * A1[k * nb + b1] * A1[k * nb + b3];
          for (int b2 = 0; b2 < nb; ++b2)
            for (int b4 = 0; b4 < nb; ++b4)
              loc[((b1 * nb + b2) * nb + b3) * nb + b4] += s * F[(k * nb + b2) * nb + b4];
        }
#pragma omp critical
    for (int64_t q = 0; q < nb2 * nb2; ++q) eri[q] += loc[q];
  }
}

/* threeElectronIntegrals (src/integration/integrateAll.hpp:328-442), real
 * basis, given V from computeV (N1 x nb x nb x 3, no r2 scale):
 * Q[p,q,r,s,t,u] = sum_r1 w1 A1[p] A1[s] V[r1,q,t].V[r1,r,u].  nb^6 out. */
void orc_three_electron(const double* A1, const double* w1, int64_t n1, const double* V, int nb, double* Q) {
  int64_t const nb3 = (int64_t)nb * nb * nb;
  std::fill(Q, Q + nb3 * nb3, 0.0);
  for (int64_t k = 0; k < n1; ++k)
    for (int p = 0; p < nb; ++p)
      for (int q = 0; q < nb; ++q)
        for (int r = 0; r < nb; ++r)
          for (int s = 0; s < nb; ++s)
            for (int t = 0; t < nb; ++t)
              for (int u = 0; u < nb; ++u) {
                const double* vqt = V + ((k * nb + q) * nb + t) * 3;
                const double* vru = V + ((k * nb + r) * nb + u) * 3;
                double const dot = vqt[0] * vru[0] + vqt[1] * vru[1] + vqt[2] * vru[2];
                Q[((((p * nb + q) * nb + r) * nb + s) * nb + t) * nb + u] +=
                    w1[k] * A1[k * nb + p] * dot * A1[k * nb + s];
              }
}

/* integrateAllR1R2<true>, src/integration/integrateAll.hpp:184-317, real basis:
 * answer(b2, b1, b4, b3) = sum_r1 w1 A1_b1 A1_b3 * scale * sum_r2 w2 A2_b2
 * f(r1, r2).grad A2_b4, with f = grad_1 u (the Ochi gradient of prm, spin
 * `same`; coincident pairs give 0 as in orc_compute_F).  out: nb^4 in the
 * reference's (b2, b1, b4, b3) order.  Points SoA 3 x n. */
int orc_eri_gradient(const double* prm, int same, const double* p1, const double* A1, const double* w1, int64_t n1,
                     const double* p2, const double* w2, int64_t n2, const double* A2, const double* dA2, int nb,
                     double scale, double* out) {
  int64_t const nb2 = (int64_t)nb * nb;
  std::fill(out, out + nb2 * nb2, 0.0);
#pragma omp parallel
  {
    std::vector<double> loc(nb2 * nb2, 0.0), inner(nb2);
#pragma omp for schedule(static)
    for (int64_t k = 0; k < n1; ++k) {
      double const r1[3] = {p1[k], p1[n1 + k], p1[2 * n1 + k]};
      std::fill(inner.begin(), inner.end(), 0.0);
      for (int64_t q = 0; q < n2; ++q) {
        double const r2[3] = {p2[q], p2[n2 + q], p2[2 * n2 + q]};
        double const dx = r1[0] - r2[0], dy = r1[1] - r2[1], dz = r1[2] - r2[2];
        if (dx * dx + dy * dy + dz * dz == 0) continue;
        double f[3];
        orc_ochi_gradient(prm, r1, r2, same, f);
        for (int b4 = 0; b4 < nb; ++b4) {
          const double* g = dA2 + (q * nb + b4) * 3;
          double const fg = w2[q] * (f[0] * g[0] + f[1] * g[1] + f[2] * g[2]);
          for (int b2 = 0; b2 < nb; ++b2) inner[b2 * nb + b4] += A2[q * nb + b2] * fg;
        }
      }
      for (int b2 = 0; b2 < nb; ++b2)
        for (int b1 = 0; b1 < nb; ++b1)
          for (int b4 = 0; b4 < nb; ++b4)
            for (int b3 = 0; b3 < nb; ++b3)
              loc[((b2 * nb + b1) * nb + b4) * nb + b3] +=
                  w1[k] * A1[k * nb + b1] * A1[k * nb + b3] * scale * inner[b2 * nb + b4];
    }
#pragma omp critical
    for (int64_t q = 0; q < nb2 * nb2; ++q) out[q] += loc[q];
  }
  return 0;
}

}  // extern "C"

/* ======================================================================== */
/* Dense FP64 linear algebra (textbook methods, written for this oracle)     */
/* ======================================================================== */
namespace {

/* Cholesky S = L L^T (row-major, lower).  Returns false if not SPD. */
bool cholesky(int n, const double* S, std::vector<double>& L) {
  L.assign((size_t)n * n, 0.0);
  for (int j = 0; j < n; ++j) {
    double d = S[j * n + j];
    for (int k = 0; k < j; ++k) d -= L[j * n + k] * L[j * n + k];
    if (!(d > 0)) return false;
    L[j * n + j] = std::sqrt(d);
    for (int i = j + 1; i < n; ++i) {
      double s = S[i * n + j];
      for (int k = 0; k < j; ++k) s -= L[i * n + k] * L[j * n + k];
      L[i * n + j] = s / L[j * n + j];
    }
  }
  return true;
}

/* C = L^{-1} A L^{-T} for a general (not necessarily symmetric) A. */
void reduce_to_standard(int n, const std::vector<double>& L, const double* A, std::vector<double>& C) {
  std::vector<double> T((size_t)n * n);
  /* T = L^{-1} A : forward substitution column by column */
  for (int c = 0; c < n; ++c)
    for (int i = 0; i < n; ++i) {
      double s = A[i * n + c];
      for (int k = 0; k < i; ++k) s -= L[i * n + k] * T[k * n + c];
      T[i * n + c] = s / L[i * n + i];
    }
  /* C = T L^{-T}  <=>  C^T = L^{-1} T^T */
  C.assign((size_t)n * n, 0.0);
  for (int r = 0; r < n; ++r)
    for (int i = 0; i < n; ++i) {
      double s = T[r * n + i];
      for (int k = 0; k < i; ++k) s -= L[i * n + k] * C[r * n + k];
      C[r * n + i] = s / L[i * n + i];
    }
}

/* x = L^{-T} y (back substitution), y, x length n, complex. */
void back_transform(int n, const std::vector<double>& L, const cplx* y, cplx* x) {
  for (int i = n - 1; i >= 0; --i) {
    cplx s = y[i];
    for (int k = i + 1; k < n; ++k) s -= L[k * n + i] * x[k];
    x[i] = s / L[i * n + i];
  }
}

/* Cyclic Jacobi for a real symmetric matrix.  Eigenvalues ascending,
 * eigenvectors as columns of V (row-major n x n). */
void jacobi_sym(int n, std::vector<double> A, std::vector<double>& evals, std::vector<double>& V) {
  V.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) V[i * n + i] = 1;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0, tot = 0;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        tot += A[i * n + j] * A[i * n + j];
        if (i != j) off += A[i * n + j] * A[i * n + j];
      }
    if (off <= 1e-30 * tot || off == 0) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        double const apq = A[p * n + q];
        if (apq == 0) continue;
        double const theta = (A[q * n + q] - A[p * n + p]) / (2 * apq);
        double const t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1));
        double const c = 1 / std::sqrt(t * t + 1), s = t * c;
        for (int k = 0; k < n; ++k) {
          double const akp = A[k * n + p], akq = A[k * n + q];
          A[k * n + p] = c * akp - s * akq;
          A[k * n + q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {
          double const apk = A[p * n + k], aqk = A[q * n + k];
          A[p * n + k] = c * apk - s * aqk;
          A[q * n + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {
          double const vkp = V[k * n + p], vkq = V[k * n + q];
          V[k * n + p] = c * vkp - s * vkq;
          V[k * n + q] = s * vkp + c * vkq;
        }
      }
  }
  std::vector<int> idx(n);
  std::iota(idx.begin(), idx.end(), 0);
  std::vector<double> d(n);
  for (int i = 0; i < n; ++i) d[i] = A[i * n + i];
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return d[a] < d[b]; });
  evals.resize(n);
  std::vector<double> V2((size_t)n * n);
  for (int c = 0; c < n; ++c) {
    evals[c] = d[idx[c]];
    for (int r = 0; r < n; ++r) V2[r * n + c] = V[r * n + idx[c]];
  }
  V.swap(V2);
}

/* Eigenvalues of a general real matrix: Householder reduction to Hessenberg
 * form, then complex single-shift QR with Wilkinson shifts and deflation
 * (Golub & Van Loan, Algorithm 7.5.x in complex arithmetic).  No Schur
 * vectors; eigenvectors come from inverse iteration below. */
std::vector<cplx> eig_values_general(int n, const std::vector<double>& A0) {
  std::vector<cplx> H((size_t)n * n);
  for (int i = 0; i < n * n; ++i) H[i] = A0[i];
  for (int k = 0; k < n - 2; ++k) {
    double alpha = 0;
    for (int i = k + 1; i < n; ++i) alpha += std::norm(H[i * n + k]);
    alpha = std::sqrt(alpha);
    if (alpha == 0) continue;
    std::vector<cplx> v(n, 0.0);
    cplx const x0 = H[(k + 1) * n + k];
    cplx const ph = std::abs(x0) == 0 ? cplx(1) : x0 / std::abs(x0);
    for (int i = k + 1; i < n; ++i) v[i] = H[i * n + k];
    v[k + 1] += ph * alpha;
    double vn = 0;
    for (int i = k + 1; i < n; ++i) vn += std::norm(v[i]);
    if (vn == 0) continue;
    for (int c = 0; c < n; ++c) {
      cplx s = 0;
      for (int i = k + 1; i < n; ++i) s += std::conj(v[i]) * H[i * n + c];
      s *= 2.0 / vn;
      for (int i = k + 1; i < n; ++i) H[i * n + c] -= v[i] * s;
    }
    for (int r = 0; r < n; ++r) {
      cplx s = 0;
      for (int i = k + 1; i < n; ++i) s += H[r * n + i] * v[i];
      s *= 2.0 / vn;
      for (int i = k + 1; i < n; ++i) H[r * n + i] -= s * std::conj(v[i]);
    }
  }
  std::vector<cplx> ev(n);
  int hi = n - 1, iter = 0;
  while (hi >= 0) {
    if (hi == 0) { ev[0] = H[0]; break; }
    int lo = hi;
    while (lo > 0) {
      double const s = std::abs(H[(lo - 1) * n + lo - 1]) + std::abs(H[lo * n + lo]);
      if (std::abs(H[lo * n + lo - 1]) <= 1e-16 * (s == 0 ? 1.0 : s)) { H[lo * n + lo - 1] = 0; break; }
      --lo;
    }
    if (lo == hi) { ev[hi] = H[hi * n + hi]; --hi; iter = 0; continue; }
    if (++iter > 1000) throw std::runtime_error("QR did not converge");
    cplx const a = H[(hi - 1) * n + hi - 1], b = H[(hi - 1) * n + hi], c = H[hi * n + hi - 1], d = H[hi * n + hi];
    cplx const tr = a + d, det = a * d - b * c;
    cplx const disc = std::sqrt(tr * tr / 4.0 - det);
    cplx const l1 = tr / 2.0 + disc, l2 = tr / 2.0 - disc;
    cplx mu = std::abs(l1 - d) < std::abs(l2 - d) ? l1 : l2;
    if (iter % 11 == 10) mu += std::abs(H[hi * n + hi - 1]);  // exceptional shift
    for (int i = lo; i <= hi; ++i) H[i * n + i] -= mu;
    std::vector<cplx> cs(hi - lo), sn(hi - lo);
    for (int k = lo; k < hi; ++k) {
      cplx const x = H[k * n + k], y = H[(k + 1) * n + k];
      double const r = std::sqrt(std::norm(x) + std::norm(y));
      cplx cc = 1, ss = 0;
      if (r != 0) { cc = x / r; ss = y / r; }
      cs[k - lo] = cc; sn[k - lo] = ss;
      for (int col = k; col < n; ++col) {
        cplx const h1 = H[k * n + col], h2 = H[(k + 1) * n + col];
        H[k * n + col] = std::conj(cc) * h1 + std::conj(ss) * h2;
        H[(k + 1) * n + col] = -ss * h1 + cc * h2;
      }
    }
    for (int k = lo; k < hi; ++k) {
      cplx const cc = cs[k - lo], ss = sn[k - lo];
      for (int row = 0; row <= std::min(k + 2, hi); ++row) {
        cplx const h1 = H[row * n + k], h2 = H[row * n + k + 1];
        H[row * n + k] = h1 * cc + h2 * ss;
        H[row * n + k + 1] = -h1 * std::conj(ss) + h2 * std::conj(cc);
      }
    }
    for (int i = lo; i <= hi; ++i) H[i * n + i] += mu;
  }
  return ev;
}

/* Right eigenvector of general real A for eigenvalue lam: two steps of
 * inverse iteration with a complex LU (partial pivoting). */
std::vector<cplx> inverse_iteration(int n, const std::vector<double>& A, cplx lam) {
  double anorm = 0;
  for (int i = 0; i < n * n; ++i) anorm = std::max(anorm, std::fabs(A[i]));
  cplx const shift = lam + cplx(1e-13 * (anorm > 0 ? anorm : 1.0), 0);
  std::vector<cplx> M((size_t)n * n);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) M[i * n + j] = A[i * n + j] - (i == j ? shift : cplx(0));
  std::vector<int> piv(n);
  for (int k = 0; k < n; ++k) {
    int p = k;
    for (int i = k + 1; i < n; ++i)
      if (std::abs(M[i * n + k]) > std::abs(M[p * n + k])) p = i;
    piv[k] = p;
    if (p != k)
      for (int j = 0; j < n; ++j) std::swap(M[k * n + j], M[p * n + j]);
    if (std::abs(M[k * n + k]) == 0) M[k * n + k] = 1e-300;
    for (int i = k + 1; i < n; ++i) {
      M[i * n + k] /= M[k * n + k];
      for (int j = k + 1; j < n; ++j) M[i * n + j] -= M[i * n + k] * M[k * n + j];
    }
  }
  std::vector<cplx> x(n, cplx(1.0, 0.0));
  for (int it = 0; it < 3; ++it) {
    for (int k = 0; k < n; ++k) std::swap(x[k], x[piv[k]]);
    for (int i = 0; i < n; ++i)
      for (int k = 0; k < i; ++k) x[i] -= M[i * n + k] * x[k];
    for (int i = n - 1; i >= 0; --i) {
      for (int k = i + 1; k < n; ++k) x[i] -= M[i * n + k] * x[k];
      x[i] /= M[i * n + i];
    }
    double nrm = 0;
    for (auto& v : x) nrm += std::norm(v);
    nrm = std::sqrt(nrm);
    for (auto& v : x) v /= nrm;
  }
  return x;
}

/* Connected components of the nonzero pattern of (A, B): the (l,m) symmetry
 * blocks when the one-electron mask is applied. */
std::vector<std::vector<int>> blocks_of(int n, const double* A, const double* B) {
  std::vector<int> comp(n, -1);
  std::vector<std::vector<int>> out;
  for (int s = 0; s < n; ++s) {
    if (comp[s] >= 0) continue;
    std::vector<int> stack{s}, mem;
    comp[s] = (int)out.size();
    while (!stack.empty()) {
      int const u = stack.back();
      stack.pop_back();
      mem.push_back(u);
      for (int v = 0; v < n; ++v)
        if (comp[v] < 0 && (A[u * n + v] != 0 || A[v * n + u] != 0 || B[u * n + v] != 0)) {
          comp[v] = comp[s];
          stack.push_back(v);
        }
    }
    std::sort(mem.begin(), mem.end());
    out.push_back(mem);
  }
  return out;
}

/* Generalized symmetric-definite eigenproblem F x = lam S x (the role of
 * LAPACK hegvd, src/HartreeFock.cpp:864-879), solved per symmetry block.
 * Eigenvalues ascending; eigenvectors S-orthonormal, columns of C (n x n). */
bool gen_sym_eig(int n, const double* F, const double* S, std::vector<double>& evals, std::vector<double>& C) {
  auto blocks = blocks_of(n, F, S);
  std::vector<std::pair<double, std::vector<double>>> pairs;
  for (auto& blk : blocks) {
    int const m = (int)blk.size();
    std::vector<double> Fb((size_t)m * m), Sb((size_t)m * m), L, A, ev, V;
    for (int a = 0; a < m; ++a)
      for (int b = 0; b < m; ++b) {
        Fb[a * m + b] = 0.5 * (F[blk[a] * n + blk[b]] + F[blk[b] * n + blk[a]]);
        Sb[a * m + b] = S[blk[a] * n + blk[b]];
      }
    if (!cholesky(m, Sb.data(), L)) return false;
    reduce_to_standard(m, L, Fb.data(), A);
    for (int a = 0; a < m; ++a)
      for (int b = a + 1; b < m; ++b) A[a * m + b] = A[b * m + a] = 0.5 * (A[a * m + b] + A[b * m + a]);
    jacobi_sym(m, A, ev, V);
    for (int c = 0; c < m; ++c) {
      std::vector<cplx> y(m), x(m);
      for (int r = 0; r < m; ++r) y[r] = V[r * m + c];
      back_transform(m, L, y.data(), x.data());
      std::vector<double> full(n, 0.0);
      for (int r = 0; r < m; ++r) full[blk[r]] = x[r].real();
      pairs.push_back({ev[c], full});
    }
  }
  std::stable_sort(pairs.begin(), pairs.end(), [](auto& a, auto& b) { return a.first < b.first; });
  evals.resize(n);
  C.assign((size_t)n * n, 0.0);
  for (int c = 0; c < n; ++c) {
    evals[c] = pairs[c].first;
    for (int r = 0; r < n; ++r) C[r * n + c] = pairs[c].second[r];
  }
  return true;
}

/* Generalized nonsymmetric eigenproblem F x = lam S x (the role of LAPACK
 * ggev, src/HartreeFock.cpp:1277-1291) with S SPD, per symmetry block:
 * reduce with Cholesky, eigenvalues by QR, eigenvectors by inverse iteration. */
bool gen_nonsym_eig(int n, const double* F, const double* S, std::vector<cplx>& evals, std::vector<cplx>& X) {
  auto blocks = blocks_of(n, F, S);
  evals.clear();
  std::vector<std::vector<cplx>> vecs;
  for (auto& blk : blocks) {
    int const m = (int)blk.size();
    std::vector<double> Fb((size_t)m * m), Sb((size_t)m * m), L, A;
    for (int a = 0; a < m; ++a)
      for (int b = 0; b < m; ++b) {
        Fb[a * m + b] = F[blk[a] * n + blk[b]];
        Sb[a * m + b] = S[blk[a] * n + blk[b]];
      }
    if (!cholesky(m, Sb.data(), L)) return false;
    reduce_to_standard(m, L, Fb.data(), A);
    std::vector<cplx> ev = eig_values_general(m, A);
    for (int c = 0; c < m; ++c) {
      std::vector<cplx> y = inverse_iteration(m, A, ev[c]), x(m);
      back_transform(m, L, y.data(), x.data());
      std::vector<cplx> full(n, 0.0);
      for (int r = 0; r < m; ++r) full[blk[r]] = x[r];
      evals.push_back(ev[c]);
      vecs.push_back(full);
    }
  }
  X.assign((size_t)n * n, 0.0);
  for (int c = 0; c < n; ++c)
    for (int r = 0; r < n; ++r) X[r * n + c] = vecs[c][r];
  return true;
}

/* ------------------------------------------------------------------------ */
/* SCF helpers, src/HartreeFock.cpp:20-98, 756-780                            */
/* ------------------------------------------------------------------------ */

/* getNewDensity: D[u,v] = sum_{a<nocc} C[u,a] C[v,a]  (src/HartreeFock.cpp:20-44) */
void new_density(int nb, int nocc, const std::vector<double>& C, int ldc, double* D) {
  for (int u = 0; u < nb; ++u)
    for (int v = 0; v < nb; ++v) {
      double t = 0;
      for (int a = 0; a < nocc; ++a) t += C[u * ldc + a] * C[v * ldc + a];
      D[u * nb + v] = t;
    }
}

/* enforceOneElectronSymmetry (src/HartreeFock.cpp:756-780) */
void mask_by(int nb, double* G, const double* H) {
  for (int q = 0; q < nb * nb; ++q)
    if (!(std::fabs(H[q]) > 0)) G[q] = 0;
}

/* Modified Gram-Schmidt with a metric, src/orthogonalization.hpp:253-271:
 * vector 0 normalized in the metric, later vectors with the Euclidean norm
 * (mode 0 = reference behaviour) or the metric norm (mode 1).  vecs: k x n. */
void mgs_metric(int k, int n, std::vector<cplx>& v, const double* S, int mode) {
  auto ip = [&](int a, int b) {
    cplx d = 0;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) d = d + std::conj(v[a * n + i]) * S[i * n + j] * v[b * n + j];
    return d;
  };
  auto mnorm = [&](int a) { return std::sqrt(ip(a, a).real()); };
  auto enorm = [&](int a) {
    double d = 0;
    for (int i = 0; i < n; ++i) d = d + std::norm(v[a * n + i]);
    return std::sqrt(d);
  };
  if (k == 0) return;
  {
    double const nv = mnorm(0);
    for (int i = 0; i < n; ++i) v[i] *= 1 / nv;
  }
  for (int i = 1; i < k; ++i) {
    for (int j = i; j < k; ++j) {
      cplx const sc = ip(i - 1, j);
      for (int q = 0; q < n; ++q) v[j * n + q] = v[j * n + q] + (-sc) * v[(i - 1) * n + q];
    }
    double const nv = mode == 0 ? enorm(i) : mnorm(i);
    for (int q = 0; q < n; ++q) v[i * n + q] *= 1 / nv;
  }
}

}  // namespace

/* ======================================================================== */
/* SCF drivers (C ABI)                                                        */
/* ======================================================================== */
extern "C" {

struct orc_scf_opts {
  double tol;       /* |E_new - E_old| < tol ends the loop (reference: 1e-5) */
  int max_iter;     /* reference: 50 (RCS), 100 (UOS), 200 (TC) */
  int mask;         /* respectOneElectronSymmetry */
  int mgs_mode;     /* TC only: 0 = reference (Euclidean for i>=1), 1 = metric */
};

struct orc_scf_out {
  double energy;
  int iterations;   /* numberOfConvergenceLoops() = _iter + 1 */
  int status;       /* 0 ok, -1 not converged, -2 solver failure */
};

void orc_mgs(int k, int n, double* vecs_cplx, const double* S, int mode) {
  std::vector<cplx> v(reinterpret_cast<cplx*>(vecs_cplx), reinterpret_cast<cplx*>(vecs_cplx) + (size_t)k * n);
  mgs_metric(k, n, v, S, mode);
  std::copy(v.begin(), v.end(), reinterpret_cast<cplx*>(vecs_cplx));
}

int orc_gen_sym_eig(int n, const double* F, const double* S, double* evals, double* C) {
  std::vector<double> e, c;
  if (!gen_sym_eig(n, F, S, e, c)) return -2;
  std::copy(e.begin(), e.end(), evals);
  std::copy(c.begin(), c.end(), C);
  return 0;
}

int orc_gen_nonsym_eig(int n, const double* F, const double* S, double* evals_c, double* X_c) {
  std::vector<cplx> e, x;
  try {
    if (!gen_nonsym_eig(n, F, S, e, x)) return -2;
  } catch (std::exception const& ex) {
    g_err = ex.what();
    return -2;
  }
  std::copy(e.begin(), e.end(), reinterpret_cast<cplx*>(evals_c));
  std::copy(x.begin(), x.end(), reinterpret_cast<cplx*>(X_c));
  return 0;
}

/* RCSHartreeFock::compute, src/HartreeFock.cpp:785-886.
 * S, H: nb x nb row-major; A1: N1 x nb; w1: N1; F: N1 x nb x nb.
 * energy_trace[it] receives the energy of iteration it (may be NULL).
 * D_out: nb x nb (density C_occ C_occ^T), evals_out: nb (may be NULL). */
int orc_rcs_scf(int nelec, int nb, const double* S, const double* H, const double* A1, const double* w1, int64_t n1,
                const double* F, const orc_scf_opts* o, orc_scf_out* out, double* energy_trace, double* D_out,
                double* evals_out) {
  int64_t const nb2 = (int64_t)nb * nb;
  std::vector<double> D(nb2, 0.0), G(nb2), Fock(nb2), evals(nb, 0.0), C;
  std::vector<double> Faa(n1), Fai(n1 * nb);
  double energy = std::numeric_limits<double>::max();
  out->status = -1;
  for (int it = 0; it < o->max_iter; ++it) {
    /* calculateIaa (HartreeFock.cpp:183-213), calculateFai (:144-179) */
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < n1; ++k) {
      const double* Fk = F + k * nb2;
      double res = 0;
      for (int j = 0; j < nb; ++j)
        for (int i = 0; i < nb; ++i) res += Fk[j * nb + i] * D[i * nb + j];
      Faa[k] = res;
      for (int i = 0; i < nb; ++i) Fai[k * nb + i] = 0;
      for (int j = 0; j < nb; ++j) {
        double ks = 0;
        for (int q = 0; q < nb; ++q) ks += D[q * nb + j] * A1[k * nb + q];
        for (int i = 0; i < nb; ++i) Fai[k * nb + i] += ks * Fk[j * nb + i];
      }
    }
    /* G loop, HartreeFock.cpp:828-842 */
#pragma omp parallel for schedule(static)
    for (int64_t ji = 0; ji < nb2; ++ji) {
      int const j = (int)(ji / nb), i = (int)(ji % nb);
      double t = 0;
      for (int64_t k = 0; k < n1; ++k)
        t += w1[k] * A1[k * nb + j] * (2 * A1[k * nb + i] * Faa[k] - Fai[k * nb + i]);
      G[ji] = t;
    }
    if (o->mask) mask_by(nb, G.data(), H);
    /* constructSCFOperator, HartreeFock.cpp:48-98, x2 for RCS (:850) */
    double e = 0;
    for (int j = 0; j < nb; ++j)
      for (int i = 0; i < nb; ++i) {
        Fock[j * nb + i] = H[j * nb + i] + G[j * nb + i];
        e += D[i * nb + j] * (H[j * nb + i] + G[j * nb + i] / 2);
      }
    double const newEnergy = 2 * e;
    if (energy_trace) energy_trace[it] = newEnergy;
    out->iterations = it + 1;
    if (std::fabs(newEnergy - energy) < o->tol) {
      out->energy = newEnergy;
      out->status = 0;
      break;
    }
    energy = newEnergy;
    if (!gen_sym_eig(nb, Fock.data(), S, evals, C)) { out->status = -2; return -2; }
    new_density(nb, nelec / 2, C, nb, D.data());
  }
  if (out->status != 0) out->energy = std::numeric_limits<double>::max();
  if (D_out) std::copy(D.begin(), D.end(), D_out);
  if (evals_out) std::copy(evals.begin(), evals.end(), evals_out);
  return out->status;
}

/* UOSHartreeFock::compute, src/HartreeFock.cpp:890-1003.  D_out: 2 x nb x nb. */
int orc_uos_scf(int nup, int ndown, int nb, const double* S, const double* H, const double* A1, const double* w1,
                int64_t n1, const double* F, const orc_scf_opts* o, orc_scf_out* out, double* energy_trace,
                double* D_out, double* evals_out) {
  int64_t const nb2 = (int64_t)nb * nb;
  int const ne[2] = {nup, ndown};
  std::vector<double> D(2 * nb2, 0.0), G(2 * nb2), Fock(2 * nb2), evals(2 * nb, 0.0), C[2];
  std::vector<double> Faa(2 * n1), Fai(2 * n1 * nb);
  double energy = std::numeric_limits<double>::max();
  out->status = -1;
  for (int it = 0; it < o->max_iter; ++it) {
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < n1; ++k) {
      const double* Fk = F + k * nb2;
      for (int s = 0; s < 2; ++s) {
        const double* Ds = D.data() + s * nb2;
        double res = 0;
        for (int j = 0; j < nb; ++j)
          for (int i = 0; i < nb; ++i) res += Fk[j * nb + i] * Ds[i * nb + j];
        Faa[s * n1 + k] = res;
        double* fa = Fai.data() + (s * n1 + k) * nb;
        for (int i = 0; i < nb; ++i) fa[i] = 0;
        for (int j = 0; j < nb; ++j) {
          double ks = 0;
          for (int q = 0; q < nb; ++q) ks += Ds[q * nb + j] * A1[k * nb + q];
          for (int i = 0; i < nb; ++i) fa[i] += ks * Fk[j * nb + i];
        }
      }
    }
    for (int s = 0; s < 2; ++s) {
#pragma omp parallel for schedule(static)
      for (int64_t ji = 0; ji < nb2; ++ji) {
        int const j = (int)(ji / nb), i = (int)(ji % nb);
        double t = 0;
        for (int64_t k = 0; k < n1; ++k)
          t += w1[k] * A1[k * nb + j] *
               (A1[k * nb + i] * (Faa[s * n1 + k] + Faa[(1 - s) * n1 + k]) - Fai[(s * n1 + k) * nb + i]);
        G[s * nb2 + ji] = t;
      }
    }
    if (o->mask) {
      mask_by(nb, G.data(), H);
      mask_by(nb, G.data() + nb2, H);
    }
    double e = 0;
    for (int s = 0; s < 2; ++s)
      for (int j = 0; j < nb; ++j)
        for (int i = 0; i < nb; ++i) {
          Fock[s * nb2 + j * nb + i] = H[j * nb + i] + G[s * nb2 + j * nb + i];
          e += D[s * nb2 + i * nb + j] * (H[j * nb + i] + G[s * nb2 + j * nb + i] / 2);
        }
    if (energy_trace) energy_trace[it] = e;
    out->iterations = it + 1;
    if (std::fabs(e - energy) < o->tol) {
      out->energy = e;
      out->status = 0;
      break;
    }
    energy = e;
    for (int s = 0; s < 2; ++s) {
      std::vector<double> ev;
      if (!gen_sym_eig(nb, Fock.data() + s * nb2, S, ev, C[s])) { out->status = -2; return -2; }
      std::copy(ev.begin(), ev.end(), evals.begin() + s * nb);
    }
    for (int s = 0; s < 2; ++s) new_density(nb, ne[s], C[s], nb, D.data() + s * nb2);
  }
  if (out->status != 0) out->energy = std::numeric_limits<double>::max();
  if (D_out) std::copy(D.begin(), D.end(), D_out);
  if (evals_out) std::copy(evals.begin(), evals.end(), evals_out);
  return out->status;
}

/* One TC Fock build, src/HartreeFock.cpp:1118-1224, real basis: the per-point
 * intermediates of :1118-1148 (the reference's internal::calculate* sweeps,
 * :102-609) and the (j, i) assembly of :1150-1224.  D: 2 x nb x nb (density(s,
 * i, j) at [s][i][j]); Cocc[s]: nb x ne[s] row-major.  G2, G3: 2 x nb x nb,
 * before the mask. */
static void tc_fock_build(const int ne[2], int nb, const double* A1, const double* dA1, const double* w1, int64_t n1,
                          const double* Fs, const double* Fo, const double* Vs, const double* Vo, const double* D,
                          const double* const Cocc[2], double* G2, double* G3) {
  int64_t const nb2 = (int64_t)nb * nb;
  int const nmax = std::max(ne[0], ne[1]);
  std::vector<double> FaaSPO(2 * n1), FaiS(2 * n1 * nb), VaaSPO(2 * n1 * 3), Vai(2 * n1 * nb), Vja(2 * n1 * nb * 3),
      sg(2 * n1 * 3), sNVabS(2 * n1), sNVabO(2 * n1), VTSI(2 * n1 * nb * 3), SVbiVT(2 * n1 * nb), occDen(2 * n1),
      SCVT(2 * n1 * 3);
  std::vector<double> VjbS(2 * n1 * nb * std::max(nmax, 1) * 3), VjbO(2 * n1 * nb * std::max(nmax, 1) * 3);
  std::vector<double> VC(2 * nb2);
  auto V3 = [&](const double* V, int64_t k, int j, int i) { return V + ((k * nb + j) * nb + i) * 3; };
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < n1; ++k) {
    const double* A = A1 + k * nb;
    const double* dA = dA1 + k * nb * 3;
    for (int s = 0; s < 2; ++s) {
      const double* Ds = D + s * nb2;
      const double* Db = D + (1 - s) * nb2;
      /* calculateIaa x2 (F and V), :1120-1126 */
      double faa = 0, vaa[3] = {0, 0, 0};
      for (int j = 0; j < nb; ++j)
        for (int i = 0; i < nb; ++i) {
          faa += Fs[(k * nb + j) * nb + i] * Ds[i * nb + j];
          for (int c = 0; c < 3; ++c) vaa[c] += V3(Vs, k, j, i)[c] * Ds[i * nb + j];
        }
      for (int j = 0; j < nb; ++j)
        for (int i = 0; i < nb; ++i) {
          faa += Fo[(k * nb + j) * nb + i] * Db[i * nb + j];
          for (int c = 0; c < 3; ++c) vaa[c] += V3(Vo, k, j, i)[c] * (g_plant == 2 ? Ds : Db)[i * nb + j];
        }
      FaaSPO[s * n1 + k] = faa;
      for (int c = 0; c < 3; ++c) VaaSPO[(s * n1 + k) * 3 + c] = vaa[c];
      /* calculateFai, :1123 */
      double* fai = FaiS.data() + (s * n1 + k) * nb;
      for (int i = 0; i < nb; ++i) fai[i] = 0;
      for (int j = 0; j < nb; ++j) {
        double ks = 0;
        for (int q = 0; q < nb; ++q) ks += Ds[q * nb + j] * A[q];
        for (int i = 0; i < nb; ++i) fai[i] += ks * Fs[(k * nb + j) * nb + i];
      }
      /* calculateVai (:217-252) */
      double* vai = Vai.data() + (s * n1 + k) * nb;
      for (int i = 0; i < nb; ++i) vai[i] = 0;
      for (int j = 0; j < nb; ++j) {
        double ks[3] = {0, 0, 0};
        const double* Dv = g_plant == 4 ? Db : Ds;
        for (int q = 0; q < nb; ++q)
          for (int c = 0; c < 3; ++c) ks[c] += Dv[q * nb + j] * dA[q * 3 + c];
        for (int i = 0; i < nb; ++i) {
          const double* v = V3(Vs, k, j, i);
          vai[i] += v[0] * ks[0] + v[1] * ks[1] + v[2] * ks[2];
        }
      }
      /* calculateVja (:256-291) */
      double* vja = Vja.data() + (s * n1 + k) * nb * 3;
      for (int q = 0; q < nb * 3; ++q) vja[q] = 0;
      for (int i = 0; i < nb; ++i) {
        double ks = 0;
        for (int q = 0; q < nb; ++q) ks += Ds[i * nb + q] * A[q];
        for (int j = 0; j < nb; ++j)
          for (int c = 0; c < 3; ++c) vja[j * 3 + c] += ks * V3(Vs, k, j, i)[c];
      }
      /* calculateScaledGradients (:102-140) */
      double g[3] = {0, 0, 0};
      for (int j = 0; j < nb; ++j)
        for (int i = 0; i < nb; ++i)
          for (int c = 0; c < 3; ++c) g[c] += Ds[i * nb + j] * A[j] * dA[i * 3 + c];
      for (int c = 0; c < 3; ++c) sg[(s * n1 + k) * 3 + c] = g[c];
      /* calculateVjb (:295-337), calculateVab (:341-393) with the spin's own orbitals */
      const double* C = Cocc[s];
      int const no = ne[s];
      double* vjbs = VjbS.data() + (s * n1 + k) * nb * nmax * 3;
      double* vjbo = VjbO.data() + (s * n1 + k) * nb * nmax * 3;
      for (int j = 0; j < nb; ++j)
        for (int b = 0; b < no; ++b) {
          double rs[3] = {0, 0, 0}, ro[3] = {0, 0, 0};
          for (int i = 0; i < nb; ++i)
            for (int c = 0; c < 3; ++c) {
              rs[c] += C[i * no + b] * V3(Vs, k, j, i)[c];
              ro[c] += C[i * no + b] * V3(Vo, k, j, i)[c];
            }
          for (int c = 0; c < 3; ++c) {
            vjbs[(j * nmax + b) * 3 + c] = rs[c];
            vjbo[(j * nmax + b) * 3 + c] = ro[c];
          }
        }
      double nS = 0, nO = 0;
      for (int a = 0; a < no; ++a)
        for (int b = 0; b < no; ++b) {
          double rs[3] = {0, 0, 0}, ro[3] = {0, 0, 0};
          for (int j = 0; j < nb; ++j)
            for (int c = 0; c < 3; ++c) {
              rs[c] += C[j * no + a] * vjbs[(j * nmax + b) * 3 + c];
              ro[c] += C[j * no + a] * vjbo[(j * nmax + b) * 3 + c];
            }
          nS += rs[0] * rs[0] + rs[1] * rs[1] + rs[2] * rs[2];
          nO += ro[0] * ro[0] + ro[1] * ro[1] + ro[2] * ro[2];
        }
      sNVabS[s * n1 + k] = nS;
      sNVabO[s * n1 + k] = nO;
      /* calculateVTildeSubI (:438-480) */
      double* vtsi = VTSI.data() + (s * n1 + k) * nb * 3;
      for (int q = 0; q < nb * 3; ++q) vtsi[q] = 0;
      for (int kk = 0; kk < nb; ++kk) {
        double es = 0;
        for (int l = 0; l < nb; ++l) es += Ds[l * nb + kk] * A[l];
        for (int i = 0; i < nb; ++i)
          for (int c = 0; c < 3; ++c) vtsi[i * 3 + c] += es * V3(Vs, k, kk, i)[c];
      }
      /* calculateOccupiedElectronDensity (:530-567) */
      double od = 0;
      for (int kk = 0; kk < nb; ++kk)
        for (int l = 0; l < nb; ++l) od += Ds[l * nb + kk] * A[kk] * A[l];
      occDen[s * n1 + k] = od;
    }
    for (int s = 0; s < 2; ++s) {
      const double* Ds = D + s * nb2;
      const double* vtsi = VTSI.data() + (s * n1 + k) * nb * 3;
      /* calculateSumOverb_Vbi_dot_VTildeb (:484-526) */
      double* sv = SVbiVT.data() + (s * n1 + k) * nb;
      for (int i = 0; i < nb; ++i) sv[i] = 0;
      for (int kk = 0; kk < nb; ++kk) {
        double es[3] = {0, 0, 0};
        for (int l = 0; l < nb; ++l)
          for (int c = 0; c < 3; ++c) es[c] += Ds[l * nb + kk] * vtsi[l * 3 + c];
        for (int i = 0; i < nb; ++i) {
          const double* v = V3(Vs, k, kk, i);
          sv[i] += v[0] * es[0] + v[1] * es[1] + v[2] * es[2];
        }
      }
      /* calculateSumOverb_conjVTildeb_occb (:571-609) */
      double cs[3] = {0, 0, 0};
      for (int kk = 0; kk < nb; ++kk)
        for (int l = 0; l < nb; ++l)
          for (int c = 0; c < 3; ++c) cs[c] += Ds[l * nb + kk] * A1[k * nb + l] * vtsi[kk * 3 + c];
      for (int c = 0; c < 3; ++c) SCVT[(s * n1 + k) * 3 + c] = cs[c];
    }
  }
  /* Fock assembly, :1150-1224 */
  for (int s = 0; s < 2; ++s) {
    int const sb = 1 - s;
#pragma omp parallel for schedule(static)
    for (int64_t ji = 0; ji < nb2; ++ji) {
      int const j = (int)(ji / nb), i = (int)(ji % nb);
      double scalar = 0, vector = 0, scalar3 = 0;
      for (int64_t k = 0; k < n1; ++k) {
        const double* A = A1 + k * nb;
        const double* dA = dA1 + k * nb * 3;
        auto dot3 = [](const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; };
        const double* vaa = &VaaSPO[(s * n1 + k) * 3];
        const double* vt = vaa;       /* VTilde == VaaSamePlusOppo */
        const double* vtb = &VaaSPO[(sb * n1 + k) * 3];
        const double* vsji = V3(Vs, k, j, i);
        const double* voji = V3(Vo, k, j, i);
        double const sa = A[j] * (FaaSPO[s * n1 + k] * A[i] - FaiS[(s * n1 + k) * nb + i]);
        scalar += w1[k] * sa;
        double va = A[j] * (dot3(vaa, dA + i * 3) - Vai[(s * n1 + k) * nb + i]);
        va += dot3(vsji, &sg[(s * n1 + k) * 3]);
        va += dot3(voji, &sg[(sb * n1 + k) * 3]);
        if (g_plant == 1)
          va -= dot3(&Vja[((s * n1 + k) * nb + i) * 3], dA + j * 3);
        else
          va -= dot3(&Vja[((s * n1 + k) * nb + j) * 3], dA + i * 3);
        vector += w1[k] * va;
        const double* vtsi_i = &VTSI[((s * n1 + k) * nb + i) * 3];
        const double* vtsi_j = &VTSI[((s * n1 + k) * nb + j) * 3];
        double s3 = A[j] * (dot3(vt, vt) - sNVabS[s * n1 + k] - sNVabO[sb * n1 + k]) * A[i];
        s3 += 2 * A[j] * (SVbiVT[(s * n1 + k) * nb + i] - dot3(vt, vtsi_i));
        s3 += 2 * (SVbiVT[(s * n1 + k) * nb + j] - dot3(vtsi_j, vt)) * A[i];
        double sbS = 0, sbO = 0;
        const double* vjbs = &VjbS[(s * n1 + k) * nb * nmax * 3];
        const double* vjbo = &VjbO[(s * n1 + k) * nb * nmax * 3];
        for (int b = 0; b < ne[s]; ++b) {
          sbS += dot3(vjbs + (j * nmax + b) * 3, vjbs + (i * nmax + b) * 3);
          sbO += dot3(vjbo + (j * nmax + b) * 3, vjbo + (i * nmax + b) * 3);
        }
        s3 += 2 * occDen[s * n1 + k] * (dot3(vsji, vt) - sbS);
        s3 += 2 * (dot3(vtsi_i, vtsi_j) - dot3(vsji, &SCVT[(s * n1 + k) * 3]));
        s3 += 2 * occDen[sb * n1 + k] * (dot3(voji, vtb) - sbO);
        s3 -= 2 * dot3(voji, &SCVT[(sb * n1 + k) * 3]);
        scalar3 += w1[k] * s3;
      }
      G2[s * nb2 + ji] = scalar / 2;
      VC[s * nb2 + ji] = vector / 2;
      G3[s * nb2 + ji] = scalar3 / 2;
    }
    for (int j = 0; j < nb; ++j)
      for (int i = 0; i < nb; ++i)
        G2[s * nb2 + j * nb + i] += VC[s * nb2 + j * nb + i] - (g_plant == 3 ? -1 : 1) * VC[s * nb2 + i * nb + j];
  }
}

/* TCHartreeFock::compute, src/HartreeFock.cpp:1007-1362, real basis.
 * Fs, Fo: N1 x nb x nb; Vs, Vo: N1 x nb x nb x 3; A1: N1 x nb; dA1: N1 x nb x 3.
 * D_out: 2 x nb x nb, pseudo_out: nup + ndown occupied pseudo-energies. */
static int tc_scf_impl(int nup, int ndown, int nb, const double* S, const double* H, const double* A1, const double* dA1,
               const double* w1, int64_t n1, const double* Fs, const double* Fo, const double* Vs, const double* Vo,
               const orc_scf_opts* o, orc_scf_out* out, double* energy_trace, double* D_out, double* pseudo_out) {
  int64_t const nb2 = (int64_t)nb * nb;
  int const ne[2] = {nup, ndown};
  std::vector<double> D(2 * nb2, 0.0);
  std::vector<double> Cocc[2] = {std::vector<double>((size_t)nb * nup, 0.0), std::vector<double>((size_t)nb * ndown, 0.0)};
  std::vector<double> pseudo(nup + ndown, 0.0);
  std::vector<double> G2(2 * nb2), G3(2 * nb2), Fock(2 * nb2);
  double energy = std::numeric_limits<double>::max();
  out->status = -1;
  for (int it = 0; it < o->max_iter; ++it) {
    const double* Cp[2] = {Cocc[0].data(), Cocc[1].data()};
    tc_fock_build(ne, nb, A1, dA1, w1, n1, Fs, Fo, Vs, Vo, D.data(), Cp, G2.data(), G3.data());
    if (o->mask)
      for (int s = 0; s < 2; ++s) {
        mask_by(nb, G2.data() + s * nb2, H);
        mask_by(nb, G3.data() + s * nb2, H);
      }
    double e = 0;
    for (int s = 0; s < 2; ++s)
      for (int j = 0; j < nb; ++j)
        for (int i = 0; i < nb; ++i) {
          double const g2 = G2[s * nb2 + j * nb + i], g3 = G3[s * nb2 + j * nb + i];
          Fock[s * nb2 + j * nb + i] = H[j * nb + i] + g2 - g3;
          e += D[s * nb2 + i * nb + j] * (H[j * nb + i] + g2 / 2 - g3 / 3);
        }
    if (energy_trace) energy_trace[it] = e;
    out->iterations = it + 1;
    if (std::fabs(e - energy) < o->tol) {
      out->energy = e;
      out->status = 0;
      break;
    }
    energy = e;
    /* ggev per spin, sort by Re(lambda), MGS with the overlap metric, :1270-1357 */
    for (int s = 0; s < 2; ++s) {
      std::vector<cplx> ev, X;
      if (!gen_nonsym_eig(nb, Fock.data() + s * nb2, S, ev, X)) { out->status = -2; return -2; }
      std::vector<int> idx(nb);
      std::iota(idx.begin(), idx.end(), 0);
      std::sort(idx.begin(), idx.end(), [&](int a, int b) { return ev[a].real() < ev[b].real(); });
      int const no = ne[s];
      std::vector<cplx> vec((size_t)no * nb);
      for (int b = 0; b < no; ++b) {
        pseudo[s * nup + b] = ev[idx[b]].real();
        for (int j = 0; j < nb; ++j) vec[b * nb + j] = X[j * nb + idx[b]];
      }
      mgs_metric(no, nb, vec, S, o->mgs_mode);
      /* Real basis: remove each orbital's global phase so it is real; the
       * density and every phase-invariant sum are unchanged. */
      for (int b = 0; b < no; ++b) {
        int jm = 0;
        for (int j = 1; j < nb; ++j)
          if (std::abs(vec[b * nb + j]) > std::abs(vec[b * nb + jm])) jm = j;
        cplx const ph = std::abs(vec[b * nb + jm]) > 0 ? std::conj(vec[b * nb + jm]) / std::abs(vec[b * nb + jm]) : 1.0;
        for (int j = 0; j < nb; ++j) Cocc[s][j * no + b] = (vec[b * nb + j] * ph).real();
      }
      new_density(nb, no, Cocc[s], no, D.data() + s * nb2);
    }
  }
  if (out->status != 0) out->energy = std::numeric_limits<double>::max();
  if (D_out) std::copy(D.begin(), D.end(), D_out);
  if (pseudo_out) std::copy(pseudo.begin(), pseudo.end(), pseudo_out);
  return out->status;
}

/* The drivers report solver failures (including a QR iteration that does not
 * converge) as status -2 instead of letting exceptions cross the C ABI. */
int orc_tc_scf(int nup, int ndown, int nb, const double* S, const double* H, const double* A1, const double* dA1,
               const double* w1, int64_t n1, const double* Fs, const double* Fo, const double* Vs, const double* Vo,
               const orc_scf_opts* o, orc_scf_out* out, double* energy_trace, double* D_out, double* pseudo_out) {
  try {
    return tc_scf_impl(nup, ndown, nb, S, H, A1, dA1, w1, n1, Fs, Fo, Vs, Vo, o, out, energy_trace, D_out, pseudo_out);
  } catch (std::exception const& e) {
    g_err = e.what();
    out->status = -2;
    return -2;
  }
}

/* One Fock build of the TC loop (tc_fock_build) for a given density and
 * occupied orbitals: C: 2 x nb x cmax row-major, column b of spin s used for
 * b < ne[s]. */
int orc_tc_fock(int nup, int ndown, int nb, const double* A1, const double* dA1, const double* w1, int64_t n1,
                const double* Fs, const double* Fo, const double* Vs, const double* Vo, const double* D,
                const double* C, int cmax, double* G2, double* G3) {
  int const ne[2] = {nup, ndown};
  if (cmax < std::max(nup, ndown)) return -1;
  std::vector<double> Cs[2];
  for (int s = 0; s < 2; ++s) {
    Cs[s].assign((size_t)nb * std::max(ne[s], 1), 0.0);
    for (int i = 0; i < nb; ++i)
      for (int b = 0; b < ne[s]; ++b) Cs[s][i * ne[s] + b] = C[((size_t)s * nb + i) * cmax + b];
  }
  const double* Cp[2] = {Cs[0].data(), Cs[1].data()};
  tc_fock_build(ne, nb, A1, dA1, w1, n1, Fs, Fo, Vs, Vo, D, Cp, G2, G3);
  return 0;
}

/* computeQ, src/HartreeFock.cpp:613-682, real basis: the explicit six-index
 * three-body tensors, Q(j, p, r, i, q, s) = sum_r1 w1 [A_j A_i X12(p,q).X13(r,s)
 * + A_p A_q X12(j,i).X23(r,s) + A_r A_s X13(j,i).X23(p,q)], electron 1 = (j, i),
 * 2 = (p, q), 3 = (r, s), X = V_same or V_oppo by the pair's spins.
 * labelled = 0: the reference's choice (QSS: all same; QSO: X12 = same,
 * X13 = X23 = oppo except the third term, which reads V_same(j,i).V_oppo(p,q);
 * QOO: all oppo).  labelled = 1: X per pair from the spins (s,s,s), (s,s,t),
 * (s,t,t): QSO X12 = same, X13 = X23 = oppo; QOO X12 = X13 = oppo, X23 = same.
 * Vs, Vo: N1 x nb x nb x 3.  Each Q: nb^6. */
void orc_compute_Q(const double* A1, const double* w1, int64_t n1, const double* Vs, const double* Vo, int nb,
                   int labelled, double* QSS, double* QSO, double* QOO) {
  int64_t const n6 = (int64_t)nb * nb * nb * nb * nb * nb;
  std::fill(QSS, QSS + n6, 0.0);
  std::fill(QSO, QSO + n6, 0.0);
  std::fill(QOO, QOO + n6, 0.0);
  auto V3 = [&](const double* V, int64_t k, int j, int i) { return V + ((k * nb + j) * nb + i) * 3; };
  auto dot3 = [](const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; };
  /* pair tensors (X12, X13, X23) per Q */
  const double* const X[3][3] = {{Vs, Vs, Vs}, {Vs, Vo, Vo}, {Vo, Vo, labelled ? Vs : Vo}};
  double* const Q[3] = {QSS, QSO, QOO};
  for (int64_t k = 0; k < n1; ++k) {
    const double* A = A1 + k * nb;
    double const wk = w1[k];
    int64_t idx = 0;
    for (int j = 0; j < nb; ++j)
      for (int p = 0; p < nb; ++p)
        for (int r = 0; r < nb; ++r)
          for (int i = 0; i < nb; ++i)
            for (int q = 0; q < nb; ++q)
              for (int t = 0; t < nb; ++t, ++idx) {
                /* the reference's loop letters (j, p, r, i, q, s), with s -> t here */
                double const a123 = A[j] * A[i], a213 = A[p] * A[q], a312 = A[r] * A[t];
                for (int m = 0; m < 3; ++m) {
                  const double *x12 = X[m][0], *x13 = X[m][1], *x23 = X[m][2];
                  /* the reference's so312 term reads V_same(j,i) where the pair (1,3) is opposite */
                  const double* x13_third = (!labelled && m == 1) ? Vs : x13;
                  Q[m][idx] += wk * (a123 * dot3(V3(x12, k, p, q), V3(x13, k, r, t)) +
                                     a213 * dot3(V3(x12, k, j, i), V3(x23, k, r, t)) +
                                     a312 * dot3(V3(x13_third, k, j, i), V3(x23, k, p, q)));
                }
              }
  }
}

/* assembleG, src/HartreeFock.cpp:686-750: G3 from the explicit Q tensors and the
 * density (D[s][i][j] = density(s, i, j)).  G3: 2 x nb x nb. */
void orc_assemble_G3(int nb, const double* QSS, const double* QSO, const double* QOO, const double* D, double* G3) {
  int64_t const nb2 = (int64_t)nb * nb;
  auto Q = [nb](const double* T, int a, int b, int c, int d, int e, int f) {
    return T[(((((int64_t)a * nb + b) * nb + c) * nb + d) * nb + e) * nb + f];
  };
  for (int j = 0; j < nb; ++j)
    for (int i = 0; i < nb; ++i) {
      double value[2] = {0, 0};
      for (int p = 0; p < nb; ++p)
        for (int q = 0; q < nb; ++q)
          for (int r = 0; r < nb; ++r)
            for (int t = 0; t < nb; ++t)
              for (int s = 0; s < 2; ++s) {
                const double* Ds = D + s * nb2;
                const double* Db = D + (1 - s) * nb2;
                value[s] += Ds[q * nb + p] * Ds[t * nb + r] *
                            (Q(QSS, j, p, r, i, q, t) - Q(QSS, j, p, r, i, t, q) + 2 * Q(QSS, j, p, r, q, t, i) -
                             2 * Q(QSS, j, p, r, t, q, i));
                value[s] += 2 * Ds[q * nb + p] * Db[t * nb + r] * (Q(QSO, j, p, r, i, q, t) - Q(QSO, j, p, r, q, i, t));
                value[s] += Db[q * nb + p] * Db[t * nb + r] * (Q(QOO, j, p, r, i, q, t) - Q(QOO, j, p, r, i, t, q));
              }
      G3[j * nb + i] = value[0] / 2;
      G3[nb2 + j * nb + i] = value[1] / 2;
    }
}

}  // extern "C"
