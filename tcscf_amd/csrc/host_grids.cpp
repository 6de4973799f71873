// Host-side quadrature grids and analytic one-electron matrices (FP64).
//
// Grids are built once on the host in FP64 and uploaded, so every consumer
// (device kernels, the CPU reference path) integrates over identical points.
#include <cmath>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "common.hpp"

namespace {

#include "../data/lebedev_tables.inc"

constexpr double kPi = 3.141592653589793238462643383279502884;
constexpr double kLn2 = 0.693147180559945309417232121458176568;

// One radial change of variables x in (-1, 1) -> r, with dr/dx.
struct RadialMap {
  int kind;
  double a, eps;
  void apply(double x, double* r, double* drdx) const {
    switch (kind) {
      case TCSCF_RADIAL_M4: {
        // Treutler-Ahlrichs M4 (src/integration/changeOfVariables.hpp:126-162)
        const double c = eps / kLn2, alpha = 0.6;
        const double lg = std::log((1 + a) / (1 - x));
        const double pm1 = std::pow(x + a, alpha - 1);
        *r = c * std::pow(x + a, alpha) * lg;
        *drdx = c * (alpha * pm1 * lg + pm1 * (x + a) / (1 - x));
        break;
      }
      case TCSCF_RADIAL_BECKE:
        *r = (1 + x) / (1 - x);
        *drdx = 2.0 / ((1 - x) * (1 - x));
        break;
      default:
        *r = x / (1 - x);
        *drdx = std::pow(1 - x, -2);
    }
  }
};

const LebedevEntry* lebedev_entry(int order) {
  for (int k = 0; k < kLebedevTableSize; ++k)
    if (kLebedevTable[k].order == order) return &kLebedevTable[k];
  return nullptr;
}

double fact(int n) { return std::tgamma(n + 1.0); }

double sto_norm(int n, double zeta) { return std::pow(2 * zeta, n) * std::sqrt(2 * zeta / fact(2 * n)); }

}  // namespace

namespace tc {

thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }

}  // namespace tc

extern "C" {

const char* tcscf_last_error(void) { return tc::g_last_error.c_str(); }

int tcscf_version(void) { return 1; }

// Gauss-Chebyshev (second kind) nodes x_i = cos((i+1) pi/(n+1)) with weights
// sin((i+1) pi/(n+1)) * pi/(n+1), mapped to r.
int tcscf_radial_grid(int n, int map, double a, double eps, double* r, double* w) {
  TC_CHECK(n > 0 && r && w, "tcscf_radial_grid: bad arguments");
  TC_CHECK(map >= 0 && map <= 2, "tcscf_radial_grid: unknown map");
  const RadialMap rm{map, a, eps};
  const double h = kPi / (n + 1);
  for (int i = 0; i < n; ++i) {
    const double t = (i + 1) * h;
    double rr, jac;
    rm.apply(std::cos(t), &rr, &jac);
    r[i] = rr;
    w[i] = std::sin(t) * h * jac;
  }
  return 0;
}

int tcscf_lebedev_npts(int order) {
  const LebedevEntry* e = lebedev_entry(order);
  if (!e) {
    tc::set_error("Lebedev order " + std::to_string(order) + " not supported");
    return TCSCF_EINVAL;
  }
  return e->npts;
}

int tcscf_lebedev(int order, double* theta, double* phi, double* w) {
  const LebedevEntry* e = lebedev_entry(order);
  TC_CHECK(e, "Lebedev order not supported");
  for (int k = 0; k < e->npts; ++k) {
    theta[k] = e->data[3 * k];
    phi[k] = e->data[3 * k + 1];
    w[k] = e->data[3 * k + 2];
  }
  return e->npts;
}

int tcscf_product_grid(int nrad, int map, double a, double eps, int order, double* pts, double* w) {
  const LebedevEntry* e = lebedev_entry(order);
  TC_CHECK(e, "Lebedev order not supported");
  std::vector<double> r(nrad), wr(nrad);
  int rc = tcscf_radial_grid(nrad, map, a, eps, r.data(), wr.data());
  if (rc) return rc;
  const int64_t n = (int64_t)nrad * e->npts;
  for (int ia = 0; ia < e->npts; ++ia) {
    const double th = e->data[3 * ia], ph = e->data[3 * ia + 1], wa = e->data[3 * ia + 2];
    const double st = std::sin(th), ct = std::cos(th), sp = std::sin(ph), cp = std::cos(ph);
    for (int ir = 0; ir < nrad; ++ir) {
      const int64_t k = (int64_t)ia * nrad + ir;
      pts[k] = r[ir] * cp * st;
      pts[n + k] = r[ir] * sp * st;
      pts[2 * n + k] = r[ir] * ct;
      w[k] = wr[ir] * wa * r[ir] * r[ir];
    }
  }
  return (int)n;
}

// S_ab = N_a N_b (n_a+n_b)!/(z_a+z_b)^(n_a+n_b+1) delta_{l l'} delta_{m m'};
// H_ab = (T + V_nuc)_ab from the radial closed forms, evaluated on b.
int tcscf_sto_overlap_core(const tcscf_sto* f, int nb, int Z, double* S, double* H) {
  TC_CHECK(nb > 0 && f && S && H, "tcscf_sto_overlap_core: bad arguments");
  auto I = [](int k, double z) { return std::tgamma(k + 1.0) / std::pow(z, k + 1); };
  for (int a = 0; a < nb; ++a) {
    TC_CHECK(f[a].n >= 1 && f[a].n >= f[a].l && f[a].l >= std::abs(f[a].m) && f[a].l <= 3,
             "tcscf_sto_overlap_core: invalid (n, l, m)");
    for (int b = 0; b < nb; ++b) {
      if (f[a].l != f[b].l || f[a].m != f[b].m) {
        S[a * nb + b] = 0;
        H[a * nb + b] = 0;
        continue;
      }
      const double Na = sto_norm(f[a].n, f[a].zeta), Nb = sto_norm(f[b].n, f[b].zeta);
      const int ns = f[a].n + f[b].n;
      const double zs = f[a].zeta + f[b].zeta;
      const int n = f[b].n, l = f[b].l;
      const double z = f[b].zeta;
      S[a * nb + b] = Na * Nb * I(ns, zs);
      H[a * nb + b] = 0.5 * Na * Nb *
                      ((l * (l + 1) - n * (n - 1)) * I(ns - 2, zs) + 2 * (n * z - Z) * I(ns - 1, zs) - z * z * I(ns, zs));
    }
  }
  return 0;
}

}  // extern "C"
