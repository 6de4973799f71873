// Drives the C++ host facade (include/tcscf.hpp) the way the reference's
// drivers drive src/HartreeFock.hpp (src/unitTests/HartreeFockTest.cpp:238-447).
//
//   facade_test cpu <basis-dir>   host-only checks (no GPU needed)
//   facade_test gpu <basis-dir>   He / Ne / TC runs on the device; prints one
//                                 "key value" line per result for the pytest
//                                 wrapper to compare with the Python path
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "tcscf.hpp"

using tcscf::SlaterTypeOrbital;
using Grid = tcscf::integration::QMCGrid<double, 3>;

static int failures = 0;
#define EXPECT(cond)                                                   \
  do {                                                                 \
    if (!(cond)) {                                                     \
      std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond);      \
      ++failures;                                                      \
    }                                                                  \
  } while (0)

static std::vector<SlaterTypeOrbital<double>> he_config1() {
  std::vector<SlaterTypeOrbital<double>> b;
  for (int k = 0; k < 5; ++k) b.emplace_back(0.5 * std::pow(2.0, k), 1, 0, 0);
  return b;
}

static int cpu_checks(const std::string& dir) {
  // loadSTO: TZ3P Ne = 1S x2, 2S x3, 2P x3, 3D x2, 4F x1 -> 5 + 9 + 10 + 7 = 31
  const auto ne = tcscf::loadSTO(dir, "TZ3P", "Ne");
  EXPECT(ne.size() == 31);
  if (ne.size() == 31) {
    EXPECT(ne[0].n == 1 && ne[0].l == 0 && std::fabs(ne[0].alpha - 12.45) < 1e-15);
    EXPECT(ne[5].n == 2 && ne[5].l == 1 && ne[5].m == -1 && std::fabs(ne[5].alpha - 5.50) < 1e-15);
    EXPECT(ne[30].n == 4 && ne[30].l == 3 && ne[30].m == 3 && std::fabs(ne[30].alpha - 3.00) < 1e-15);
  }
  EXPECT(tcscf::loadSTO(dir, "TZ3P", "Ne", 1).size() == 14);
  EXPECT(tcscf::loadSTO(dir, "QZ6P", "Ne").size() == 8 + 15 + 15 + 21);
  EXPECT(tcscf::loadSTO(dir, "TZ3P", "NoSuchElement").empty());
  bool threw = false;
  try {
    tcscf::loadSTO(dir, "DZP", "Ne");
  } catch (const tcscf::Error&) {
    threw = true;
  }
  EXPECT(threw);
  // overlap / coreMatrixElement closed forms for 1s functions
  const SlaterTypeOrbital<double> a(0.8, 1, 0, 0), b(1.7, 1, 0, 0);
  const double Sab = 8 * std::pow(0.8 * 1.7, 1.5) / std::pow(0.8 + 1.7, 3);
  EXPECT(std::fabs(tcscf::overlap(a, b) - Sab) < 1e-14);
  EXPECT(std::fabs(tcscf::overlap(a, a) - 1) < 1e-14);
  // <1s|h|1s> for one exponent zeta, nucleus Z: zeta^2/2 - Z zeta
  EXPECT(std::fabs(tcscf::coreMatrixElement(2, a, a) - (0.32 - 1.6)) < 1e-14);
  // Ochi n = 0, l = 0 is the 1s STO of zeta = alpha; the set is orthonormal
  const tcscf::OchiBasisFunction<double> o0(1.34, 0, 0, 0), o1(1.34, 1, 0, 0), p0(1.34, 0, 1, 0);
  EXPECT(tcscf::overlap(o0, o0) == 1.0 && tcscf::overlap(o0, o1) == 0.0 && tcscf::overlap(o0, p0) == 0.0);
  EXPECT(std::fabs(tcscf::coreMatrixElement(2, o0, o0) - (1.34 * 1.34 / 2 - 2 * 1.34)) < 1e-14);
  EXPECT(tcscf::coreMatrixElement(2, o0, p0) == 0.0);
  // hydrogen-like functions of the nucleus: -Z^2 / (2 n^2)
  const tcscf::HydrogenLikeBasisFunction<double> h2p(3, 2, 1, 0), h3s(3, 3, 0, 0);
  EXPECT(std::fabs(tcscf::coreMatrixElement(3, h2p, h2p) + 9.0 / 8) < 1e-15);
  EXPECT(tcscf::overlap(h2p, h3s) == 0.0 && tcscf::coreMatrixElement(3, h2p, h3s) == 0.0);
  // OchiJastrow: the one-V TC form needs c_oppo = rho c_same term by term
  const tcscf::OchiJastrow prop{1.0, 1.0, {{0.5, 0.25}, {-0.2, -0.1}}, {{1, 0, 0}, {2, 1, 0}}};
  EXPECT(prop.tcScales().c_same == 1.0 && prop.tcScales().c_oppo == 2.0);
  threw = false;
  try {
    const tcscf::OchiJastrow np{1.0, 1.0, {{0.5, 0.25}, {0.3, 0.1}}, {{1, 0, 0}, {2, 1, 0}}};
    (void)np.tcScales();
  } catch (const tcscf::Error&) {
    threw = true;
  }
  EXPECT(threw);
  threw = false;
  try {
    tcscf::HydrogenLikeBasisFunction<double> bad(2, 1, 1, 0);  // l >= n
  } catch (const tcscf::Error&) {
    threw = true;
  }
  EXPECT(threw);
  threw = false;
  try {
    SlaterTypeOrbital<double> bad(1.0, 1, 2, 0);  // n < l
  } catch (const tcscf::Error&) {
    threw = true;
  }
  EXPECT(threw);
  threw = false;
  try {
    tcscf::RCSHartreeFock<double> odd(3, 5);
  } catch (const tcscf::Error&) {
    threw = true;
  }
  EXPECT(threw);
  std::printf("cpu checks: %d failure(s)\n", failures);
  return failures ? 1 : 0;
}

static int gpu_runs(const std::string& dir) {
  // He, config 1: RHF and UHF, FP64, mask off, tol 1e-10
  {
    const auto basis = he_config1();
    Grid g = Grid::radialLebedev(25, 11, TCSCF_RADIAL_BECKE);
    g.setBasisFunctions(basis, false);
    std::vector<double> S, H;
    tcscf::oneElectronMatrices(basis, 2, S, H);
    auto F = tcscf::computeF(g, g);
    tcscf::RCSHartreeFock<double> rhf(2, 5);
    const double e = rhf.compute(S, H, g, F, false, 1e-10);
    tcscf::UOSHartreeFock<double> uhf(1, 1, 5);
    const double eu = uhf.compute(S, H, g, F, false, 1e-10);
    std::printf("he_rhf %.15f\nhe_uhf %.15f\nhe_rhf_iters %d\n", e, eu, rhf.numberOfConvergenceLoops());
    auto eri = tcscf::integrateAllR1R2(g, F);
    std::printf("he_eri_0000 %.15f\nhe_eri_0123 %.15f\n", eri[0], eri[((0 * 5 + 1) * 5 + 2) * 5 + 3]);
  }
  // Ne with the TZ3P STOs (loadSTO, s/p/d/f), 40 x 302 grid: FP64 and BF16x5
  {
    const auto basis = tcscf::loadSTO(dir, "TZ3P", "Ne");
    Grid g = Grid::radialLebedev(40, 29, TCSCF_RADIAL_M4);
    g.setBasisFunctions(basis, false);
    std::vector<double> S, H;
    tcscf::oneElectronMatrices(basis, 10, S, H);
    const int nb = static_cast<int>(basis.size());
    for (auto prec : {tcscf::Precision::FP64, tcscf::Precision::BF16X5}) {
      auto F = tcscf::computeF(g, g, prec);
      tcscf::RCSHartreeFock<double> rhf(10, nb);
      const double e = rhf.compute(S, H, g, F, true, 1e-10, 100, 6);
      std::printf("ne_tz3p_%s %.15f\n", prec == tcscf::Precision::FP64 ? "fp64" : "bf16x5", e);
    }
  }
  // TC Be (2 + 2 electrons) on small grids, b = 1, metric MGS
  {
    std::vector<SlaterTypeOrbital<double>> basis;
    for (int k = 0; k < 5; ++k) basis.emplace_back(0.6 * std::pow(2.2, k), 1, 0, 0);
    for (int k = 0; k < 2; ++k)
      for (int m = -1; m <= 1; ++m) basis.emplace_back(0.6 * std::pow(2.2, k), 2, 1, m);
    Grid g1 = Grid::radialLebedev(14, 11, TCSCF_RADIAL_M4);
    Grid g3 = Grid::radialLebedev(10, 7, TCSCF_RADIAL_M4);
    g1.setBasisFunctions(basis, true);
    g3.setBasisFunctions(basis, true);
    const auto jas = tcscf::Jastrow::fromB(1.0);
    auto Fs = tcscf::computeF(g1, g1, jas, true);
    auto Fo = tcscf::computeF(g1, g1, jas, false);
    auto V = tcscf::computeV(g1, g3, jas);
    std::vector<double> S, H;
    tcscf::oneElectronMatrices(basis, 4, S, H);
    tcscf::TCHartreeFock<double> tc(2, 2, static_cast<int>(basis.size()));
    const double e = tc.compute(S, H, g1, Fs, Fo, V, jas, true, 1e-10, 200, 1);
    std::printf("be_tc %.15f\nbe_tc_iters %d\n", e, tc.numberOfConvergenceLoops());
    // mixed precision: F pair (1/r12 at BF16x5, the smooth part at BF16x3), V at BF16x3, FP32 storage
    auto F32 = tcscf::computeFPair<float>(g1, g1, jas, tcscf::Precision::BF16X5, tcscf::Precision::BF16X3);
    auto& Fs32 = F32.first;
    auto& Fo32 = F32.second;
    auto V32 = tcscf::computeV<float>(g1, g3, jas, tcscf::Precision::BF16X3);
    tcscf::TCHartreeFock<double> tcm(2, 2, static_cast<int>(basis.size()));
    std::printf("be_tc_mixed %.15f\n", tcm.compute(S, H, g1, Fs32, Fo32, V32, jas, true, 1e-10, 200, 1, 6));
    // the same Jastrow as a general Ochi parameter set (S = {(1, 0, 0)}, c = {c_oppo, c_same})
    const tcscf::OchiJastrow u{1.0, jas.a12, {{jas.c_oppo, jas.c_same}}, {{1, 0, 0}}};
    auto Fsu = tcscf::computeF(g1, g1, u, true);
    auto Fou = tcscf::computeF(g1, g1, u, false);
    auto Vsu = tcscf::computeV(g1, g3, u, true);
    tcscf::TCHartreeFock<double> tcu(2, 2, static_cast<int>(basis.size()));
    std::printf("be_tc_ochi %.15f\n", tcu.compute(S, H, g1, Fsu, Fou, Vsu, u, true, 1e-10, 200, 1));
    // the reference's signature, VjiSame and VjiOppo as two tensors: the same Jastrow ...
    auto Vou = tcscf::computeV(g1, g3, u, false);
    tcscf::TCHartreeFock<double> tcr(2, 2, static_cast<int>(basis.size()));
    std::printf("be_tc_ref_signature %.15f\n", tcr.compute(S, H, g1, Fsu, Fou, Vsu, Vou, true, 1e-10, 200, 1));
    // ... and a general Ochi Jastrow whose spin coefficients are not proportional
    const tcscf::OchiJastrow ug{1.1, 1.0, {{0.5, 0.25}, {-0.3, 0.2}, {0.15, -0.4}}, {{1, 0, 0}, {2, 1, 0}, {1, 1, 1}}};
    auto Fsg = tcscf::computeF(g1, g1, ug, true);
    auto Fog = tcscf::computeF(g1, g1, ug, false);
    auto Vsg = tcscf::computeV(g1, g3, ug, true);
    auto Vog = tcscf::computeV(g1, g3, ug, false);
    tcscf::TCHartreeFock<double> tcg(2, 2, static_cast<int>(basis.size()));
    std::printf("be_tc_general %.15f\nbe_tc_general_iters %d\n", tcg.compute(S, H, g1, Fsg, Fog, Vsg, Vog, true, 1e-10, 200, 1),
                tcg.numberOfConvergenceLoops());
    // observables: occupiedOrbitals (S-orthonormal under the metric Gram-Schmidt),
    // density = C C^T, fockOperator
    const int nb = static_cast<int>(basis.size());
    double orth = 0, dens = 0;
    for (int s = 0; s < 2; ++s) {
      const auto& C = tcg.occupiedOrbitals[s];
      for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) {
          double t = 0;
          for (int j = 0; j < nb; ++j)
            for (int k = 0; k < nb; ++k) t += C[j * 2 + a] * S[j * nb + k] * C[k * 2 + b];
          orth = std::max(orth, std::fabs(t - (a == b ? 1.0 : 0.0)));
        }
      for (int j = 0; j < nb; ++j)
        for (int k = 0; k < nb; ++k)
          dens = std::max(dens, std::fabs(tcg.density[(s * nb + j) * nb + k] -
                                          (C[j * 2] * C[k * 2] + C[j * 2 + 1] * C[k * 2 + 1])));
    }
    std::printf("be_tc_general_orth_err %.3e\nbe_tc_general_density_err %.3e\n", orth, dens);
    std::printf("be_tc_general_fock_00 %.15f\nbe_tc_general_fock_1_23 %.15f\n", tcg.fockOperator[0],
                tcg.fockOperator[(nb + 2) * nb + 3]);
  }
  // the per-element QMC route (AtomicBasis): 5-D atomic and 6-D arrays, hydrogen-like Z = 2
  {
    using HL = tcscf::HydrogenLikeBasisFunction<double>;
    const std::vector<HL> bz = {HL(2, 1, 0, 0), HL(2, 2, 0, 0), HL(2, 2, 1, 0), HL(2, 2, 1, 1)};
    const auto A5 = tcscf::fillAtomicR12Array(bz, 20000);
    const auto A6 = tcscf::fillR12Array(bz, 20000);
    auto ix = [](int a, int b, int c, int d) { return ((a * 4 + b) * 4 + c) * 4 + d; };
    std::printf("atomic_j_2p0_2p1 %.15e\n", A5[ix(2, 3, 2, 3)].real());
    std::printf("atomic_k_2p0_2p1_re %.15e\natomic_k_2p0_2p1_im %.15e\n", A5[ix(2, 3, 3, 2)].real(),
                A5[ix(2, 3, 3, 2)].imag());
    std::printf("r12_0101 %.15e\n", A6[ix(0, 1, 0, 1)]);
    std::printf("coulomb_op %.15e\n", tcscf::atomicCoulombOperator(bz[2], bz[3], bz[2], bz[3], 20000));
    std::printf("exchange_op_re %.15e\n", tcscf::atomicExchangeOperator(bz[2], bz[3], bz[3], bz[2], 20000).real());
  }
  return 0;
}

int main(int argc, char** argv) {
  if (argc < 3) {
    std::printf("usage: facade_test cpu|gpu <basis-dir>\n");
    return 2;
  }
  const std::string mode = argv[1], dir = argv[2];
  try {
    return mode == "cpu" ? cpu_checks(dir) : gpu_runs(dir);
  } catch (const std::exception& e) {
    std::printf("exception: %s\n", e.what());
    return 1;
  }
}
