// Host-side quadrature grids and analytic one-electron matrices (FP64).
//
// Grids are built once on the host in FP64 and uploaded, so every consumer
// (device kernels, the CPU reference path) integrates over identical points.
#include <cmath>
#include <cstdint>
#include <filesystem>
#include <fstream>
#include <mutex>
#include <sstream>
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


// ------------------------------------------------------------ QMC point sets
// SplitMix64 (Steele, Lea, Flood 2014): seeds the lattice shifts, the
// generating-vector search and the Sobol scrambling.
struct SplitMix64 {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }  // [0, 1)
};

bool is_prime(int64_t n) {
  if (n < 2) return false;
  for (int64_t d = 2; d * d <= n; ++d)
    if (n % d == 0) return false;
  return true;
}

int64_t next_prime(int64_t n) {
  while (!is_prime(n)) ++n;
  return n;
}

// Korobov rank-1 lattice z = (1, a, a^2 mod n): a minimises the worst-case
// error P_2 = -1 + (1/n) sum_k prod_j (1 + 2 pi^2 B_2({k z_j / n})), B_2(x) =
// x^2 - x + 1/6 (Sloan & Joe, Lattice Methods for Multiple Integration, 1994),
// over a seeded sample of candidates.
std::vector<int64_t> korobov_vector(int64_t n, int dim, uint64_t seed) {
  SplitMix64 rng{seed ^ 0x6C61747469636521ull};
  const int ncand = n > 8 ? 256 : 1;
  double best = 1e300;
  std::vector<int64_t> z(dim, 1), zb(dim, 1);
  for (int c = 0; c < ncand; ++c) {
    const int64_t a = n > 3 ? 2 + (int64_t)(rng.next() % (uint64_t)(n - 3)) : 1;
    for (int j = 1; j < dim; ++j) z[j] = (int64_t)((__int128)z[j - 1] * a % n);
    double sum = 0;
    for (int64_t k = 0; k < n; ++k) {
      double prod = 1;
      for (int j = 0; j < dim; ++j) {
        const double x = (double)((__int128)k * z[j] % n) / (double)n;
        prod *= 1 + 2 * kPi * kPi * (x * x - x + 1.0 / 6.0);
      }
      sum += prod;
    }
    const double p2 = sum / (double)n - 1;
    if (p2 < best) {
      best = p2;
      zb = z;
    }
  }
  return zb;
}

// Sobol direction numbers for dimensions 1-3 (Joe & Kuo 2008, new-joe-kuo
// table: dim 2 s=1 a=0 m=(1); dim 3 s=2 a=1 m=(1,3)), 32-bit, MSB first.
void sobol_directions(int dim, uint32_t v[32]) {
  if (dim == 0) {
    for (int k = 0; k < 32; ++k) v[k] = 1u << (31 - k);
    return;
  }
  const int s = dim == 1 ? 1 : 2;
  const unsigned a = dim == 1 ? 0 : 1;
  uint32_t m[33];
  m[1] = 1;
  if (s == 2) m[2] = 3;
  for (int k = s + 1; k <= 32; ++k) {
    uint32_t x = m[k - s] ^ (m[k - s] << s);
    for (int i = 1; i < s; ++i)
      if ((a >> (s - 1 - i)) & 1u) x ^= m[k - i] << i;
    m[k] = x;
  }
  for (int k = 1; k <= 32; ++k) v[k - 1] = m[k] << (32 - k);
}

// Sobol points in Gray-code order, x_i = XOR of v_k over the bits of i ^ (i >> 1);
// scrambled: random lower-triangular (unit diagonal) linear matrix scrambling
// of the digits plus a random digital shift (Matousek 1998), per dimension.
void sobol_points(int64_t n, int dim, int scramble, uint64_t seed, double* u) {
  for (int d = 0; d < dim; ++d) {
    uint32_t v[32];
    sobol_directions(d, v);
    uint32_t shift = 0;
    if (scramble) {
      SplitMix64 rng{seed * 0x9E3779B97F4A7C15ull + 0x736F626F6Cull + (uint64_t)d};
      uint32_t L[32];  // row r: bits c <= r (bit 31 - c), diagonal set
      for (int r = 0; r < 32; ++r) {
        uint32_t row = 1u << (31 - r);
        for (int c = 0; c < r; ++c)
          if (rng.next() & 1ull) row |= 1u << (31 - c);
        L[r] = row;
      }
      for (int k = 0; k < 32; ++k) {
        uint32_t out = 0;
        for (int r = 0; r < 32; ++r)
          if (__builtin_popcount(L[r] & v[k]) & 1) out |= 1u << (31 - r);
        v[k] = out;
      }
      shift = (uint32_t)(rng.next() >> 32);
    }
    for (int64_t i = 0; i < n; ++i) {
      const uint64_t g = (uint64_t)i ^ ((uint64_t)i >> 1);
      uint32_t x = shift;
      for (int k = 0; k < 32 && (g >> k); ++k)
        if ((g >> k) & 1ull) x ^= v[k];
      u[i * dim + d] = (double)x * 0x1.0p-32;
    }
  }
}

double fact(int n) { return std::tgamma(n + 1.0); }

double sto_norm(int n, double zeta) { return std::pow(2 * zeta, n) * std::sqrt(2 * zeta / fact(2 * n)); }

}  // namespace

namespace tc {

thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }
const std::string& last_error() { return g_last_error; }

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


// QMCGrid(gridSize) (src/integration/integrateAll.hpp:21-65) point count:
// nshifts copies of a rank-1 lattice of n = the first prime >= min_points.
int64_t tcscf_qmc_lattice_npts(int64_t min_points, int nshifts) {
  if (min_points < 1 || nshifts < 1) return TCSCF_EINVAL;
  return (int64_t)nshifts * next_prime(min_points);
}

// The QMCGrid(gridSize) construction: lattice points x in [0, 1)^ndim (shift
// s, point i: frac(i z / n + Delta_s), shift-major), mapped by
// TreutlerAhlrichsM4(a = 0, eps = 0.9) for r, Azimuthal (theta = acos(1 - 2x),
// dtheta/dx = 1/sqrt(x(1-x))) and Linear(2 pi) for phi
// (src/integration/changeOfVariables.hpp:46-94, 126-162), weight
// = jacobian / (nshifts n) * r^2 sin(theta).  ndim = 2 drops phi and stores
// (r sin(theta), 0, r cos(theta)), as the reference's 2-D grid does.
// pts: 3 x N SoA, w: N.  Returns N.
int64_t tcscf_qmc_lattice_grid(int64_t min_points, int ndim, int nshifts, uint64_t seed, double* pts, double* w) {
  TC_CHECK(min_points >= 1 && nshifts >= 1 && (ndim == 2 || ndim == 3) && pts && w,
           "tcscf_qmc_lattice_grid: bad arguments");
  const int64_t n = next_prime(min_points), N = n * nshifts;
  const std::vector<int64_t> z = korobov_vector(n, ndim, seed);
  SplitMix64 rng{seed ^ 0x7368696674733131ull};
  const double c = 0.9 / kLn2, alpha = 0.6;
  for (int sh = 0; sh < nshifts; ++sh) {
    double delta[3];
    for (int j = 0; j < ndim; ++j) delta[j] = rng.uniform();
    for (int64_t i = 0; i < n; ++i) {
      double x[3] = {0, 0, 0};
      for (int j = 0; j < ndim; ++j) {
        const double t = (double)((__int128)i * z[j] % n) / (double)n + delta[j];
        x[j] = t - std::floor(t);
      }
      const double lg = std::log(1.0 / (1 - x[0]));
      const double pm1 = std::pow(x[0], alpha - 1);
      const double r = c * std::pow(x[0], alpha) * lg;
      const double jr = c * (alpha * pm1 * lg + pm1 * x[0] / (1 - x[0]));
      const double th = std::acos(1 - 2 * x[1]);
      const double jt = 1 / std::sqrt(x[1] * (1 - x[1]));
      const double ph = ndim == 3 ? 2 * kPi * x[2] : 0.0;
      const double jp = ndim == 3 ? 2 * kPi : 1.0;
      const int64_t k = (int64_t)sh * n + i;
      const double st = std::sin(th);
      pts[k] = r * st * std::cos(ph);
      pts[N + k] = ndim == 3 ? r * st * std::sin(ph) : 0.0;
      pts[2 * N + k] = r * std::cos(th);
      w[k] = jr * jt * jp / (double)N * r * r * st;
    }
  }
  return N;
}

// The point sets of sphericalCoordinates{3,5,6}DIntegral (src/integration/
// qmcWrapper.hpp:146-276) on a randomly shifted Korobov lattice in [0, 1)^ndim:
// r = x0 / (1 - x0), theta = pi x1, phi = 2 pi x2 for electron 1 and the same
// from x3, x4, x5 for electron 2 (ndim 5 fixes phi2 = 0, its integral being
// the prefactor's 2 pi).  Weight = prefactor / N * prod_e r_e^2 sin(theta_e) /
// (1 - x_{3e})^2, prefactor 2 pi^2 (ndim 3) or 4 pi^4 (ndim 5, 6).  pts1 (and
// pts2 for ndim 5, 6): 3 x N Cartesian SoA; w: N.  Returns N.
int64_t tcscf_qmc_spherical_points(int64_t min_points, int ndim, int nshifts, uint64_t seed, double* pts1,
                                   double* pts2, double* w) {
  TC_CHECK(min_points >= 1 && nshifts >= 1 && (ndim == 3 || ndim == 5 || ndim == 6) && pts1 && w &&
               (ndim == 3 || pts2),
           "tcscf_qmc_spherical_points: bad arguments");
  const int64_t n = next_prime(min_points), N = n * nshifts;
  const std::vector<int64_t> z = korobov_vector(n, ndim, seed);
  SplitMix64 rng{seed ^ 0x7370686572696361ull};
  const double pref = ndim == 3 ? 2 * kPi * kPi : 4 * kPi * kPi * kPi * kPi;
  for (int sh = 0; sh < nshifts; ++sh) {
    double delta[6];
    for (int j = 0; j < ndim; ++j) delta[j] = rng.uniform();
    for (int64_t i = 0; i < n; ++i) {
      double x[6] = {0, 0, 0, 0, 0, 0};
      for (int j = 0; j < ndim; ++j) {
        const double t = (double)((__int128)i * z[j] % n) / (double)n + delta[j];
        x[j] = t - std::floor(t);
      }
      const int64_t k = (int64_t)sh * n + i;
      double wk = pref / (double)N;
      for (int e = 0; e < (ndim == 3 ? 1 : 2); ++e) {
        const double* xe = x + 3 * e;
        const double r = xe[0] / (1 - xe[0]), th = kPi * xe[1];
        const double ph = (e == 1 && ndim == 5) ? 0.0 : 2 * kPi * xe[2];
        const double st = std::sin(th);
        double* p = e == 0 ? pts1 : pts2;
        p[k] = r * st * std::cos(ph);
        p[N + k] = r * st * std::sin(ph);
        p[2 * N + k] = r * std::cos(th);
        wk *= r * r * st / ((1 - xe[0]) * (1 - xe[0]));
      }
      w[k] = wk;
    }
  }
  return N;
}

// Raw Sobol points (dim <= 3, n <= 2^32) in [0, 1), Gray-code order, u: n x dim.
int tcscf_sobol_points(int64_t n, int dim, int scramble, uint64_t seed, double* u) {
  TC_CHECK(n >= 1 && n <= (1ll << 32) && dim >= 1 && dim <= 3 && u, "tcscf_sobol_points: bad arguments");
  sobol_points(n, dim, scramble, seed, u);
  return 0;
}

// Importance-sampled 3-D grid from n scrambled Sobol points: r = -ln(1 - u1)/2
// (density 2 e^{-2r}), theta = acos(1 - 2 u2), phi = 2 pi u3 (uniform on the
// sphere), w = 4 pi r^2 / (n 2 e^{-2r}).  pts: 3 x n SoA.
int tcscf_sobol_grid(int64_t n, uint64_t seed, double* pts, double* w) {
  TC_CHECK(n >= 1 && n <= (1ll << 32) && pts && w, "tcscf_sobol_grid: bad arguments");
  std::vector<double> u((size_t)n * 3);
  sobol_points(n, 3, 1, seed, u.data());
  for (int64_t i = 0; i < n; ++i) {
    const double r = -std::log1p(-u[3 * i]) / 2;
    const double ct = 1 - 2 * u[3 * i + 1], st = std::sqrt(std::max(0.0, 1 - ct * ct));
    const double ph = 2 * kPi * u[3 * i + 2];
    pts[i] = r * st * std::cos(ph);
    pts[n + i] = r * st * std::sin(ph);
    pts[2 * n + i] = r * ct;
    w[i] = 4 * kPi * r * r / ((double)n * 2 * std::exp(-2 * r));
  }
  return 0;
}

// loadSTO (src/SlaterTypeOrbital.cpp:24-76): the BASIS ... END block of an ADF
// basis file <directory>/<basis_set>/<element>, lines "<n><SPDF> <alpha>", each
// shell expanded to m = -l..l; shells with l > max_l are skipped.  Returns the
// number of functions (writes at most `capacity` of them to out).
int tcscf_load_sto(const char* directory, const char* basis_set, const char* element, int max_l, tcscf_sto* out,
                   int capacity) {
  TC_CHECK(directory && basis_set && element && capacity >= 0 && (out || capacity == 0), "tcscf_load_sto: bad arguments");
  const std::string set = basis_set;
  TC_CHECK(set == "TZ3P" || set == "QZ6P" || set == "ATZ3P" || set == "AQZ6P",
           "tcscf_load_sto: '" + set + "' is not a valid basis set");
  const std::filesystem::path dir = std::filesystem::path(directory) / set;
  TC_CHECK(std::filesystem::is_directory(dir), "tcscf_load_sto: '" + set + "' not found in '" + directory + "'");
  std::ifstream in(dir / element);
  int count = 0;
  bool in_block = false;
  for (std::string line; std::getline(in, line);) {
    if (!in_block) {
      in_block = line.rfind("BASIS", 0) == 0;
      continue;
    }
    if (line.rfind("END", 0) == 0) return count;
    std::istringstream ls(line);
    int n;
    char type;
    double alpha;
    if (ls >> n >> type >> alpha) {
      const int l = type == 'S' ? 0 : type == 'P' ? 1 : type == 'D' ? 2 : type == 'F' ? 3 : -1;
      TC_CHECK(l >= 0, std::string("tcscf_load_sto: unrecognized type: ") + type);
      if (l > max_l) continue;
      for (int m = -l; m <= l; ++m, ++count)
        if (count < capacity) out[count] = tcscf_sto{n, l, m, alpha};
    } else {
      TC_CHECK(line.empty(), "tcscf_load_sto: invalid line in file: " + line);
    }
  }
  TC_CHECK(!in_block, "tcscf_load_sto: BASIS block without END");
  return count;  // no BASIS block (or no such file): no functions, as the reference returns
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

int tcscf_basis_one_electron(const tcscf_basis_fn* f, int nb, int Z, double* S, double* H) {
  TC_CHECK(nb > 0 && f && S && H, "tcscf_basis_one_electron: bad arguments");
  const int kind = f[0].kind;
  for (int a = 0; a < nb; ++a) {
    TC_CHECK(f[a].kind == kind, "tcscf_basis_one_electron: all functions must be of one kind");
    TC_CHECK(f[a].l >= std::abs(f[a].m) && f[a].l <= 3 && f[a].alpha > 0, "tcscf_basis_one_electron: invalid function");
  }
  if (kind == TCSCF_BASIS_STO) {
    std::vector<tcscf_sto> sto(nb);
    for (int a = 0; a < nb; ++a) sto[a] = tcscf_sto{f[a].n, f[a].l, f[a].m, f[a].alpha};
    return tcscf_sto_overlap_core(sto.data(), nb, Z, S, H);
  }
  TC_CHECK(kind == TCSCF_BASIS_OCHI || kind == TCSCF_BASIS_HYDROGEN, "tcscf_basis_one_electron: unknown basis kind");
  for (int a = 0; a < nb; ++a)
    for (int b = 0; b < nb; ++b) {
      const bool lm = f[a].l == f[b].l && f[a].m == f[b].m;
      S[a * nb + b] = lm && f[a].n == f[b].n ? 1.0 : 0.0;  // orthonormal sets
      H[a * nb + b] = 0;
      if (!lm) continue;
      if (kind == TCSCF_BASIS_HYDROGEN) {
        TC_CHECK(f[a].n > f[a].l, "tcscf_basis_one_electron: hydrogen-like function needs n > l");
        if (f[a].n == f[b].n) H[a * nb + b] = -double(Z) * Z / (2.0 * f[a].n * f[a].n);
        continue;
      }
      TC_CHECK(f[a].alpha == f[b].alpha, "tcscf_basis_one_electron: Ochi functions of one alpha only");
      TC_CHECK(f[a].n >= 0, "tcscf_basis_one_electron: Ochi function needs n >= 0");
      const int l = f[a].l, n1 = f[a].n, n2 = f[b].n, mn = std::min(n1, n2), k = 2 * l + 2;
      const double alpha = f[a].alpha;
      // I = int fnl_1 fnl_2 / r dr with fnl = N r^(l+1) L_n^(k)(2 alpha r) e^(-alpha r); with
      // x = 2 alpha r and int e^-x x^(k-1) L_m^(k) L_n^(k) dx = Gamma(k + 1 + min) / (k min!):
      const double lnI = std::lgamma(k + 1.0 + mn) - std::lgamma(mn + 1.0) - std::log((double)k) -
                         0.5 * (std::lgamma(n1 + k + 1.0) - std::lgamma(n1 + 1.0) + std::lgamma(n2 + k + 1.0) -
                                std::lgamma(n2 + 1.0));
      const double I = 2 * alpha * std::exp(lnI);
      const double c = alpha * (l + 1) * (1 + 2.0 * mn / (2 * l + 3)) - Z;
      H[a * nb + b] = c * I - (n1 == n2 ? alpha * alpha / 2 : 0.0);
    }
  return 0;
}

}  // extern "C"
