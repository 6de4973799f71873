"""tcscf_amd: MI355X (gfx950) implementation of tcscf's grid-integral hot path.

The compute lives in libtcscf_hip.so (hand-written HIP, C ABI in
include/tcscf_hip.h); this package is the host-side mirror of the reference's
QMCGrid / computeF / computeV / *HartreeFock call surface.
"""
from ._lib import (BF16X3, BF16X4, BF16X5, FP32, FP64, K_COULOMB, K_JASTROW_GRAD, K_TC_SCALAR, RADIAL_BECKE, RADIAL_M4,
                   RADIAL_XOVER1MX, ConvergenceError, SolverError, TcscfError, lib)
from .api import (HydrogenLikeBasisFunction, Jastrow, OchiBasisFunction, OchiJastrow, batch_scf, QMCGrid, QuadratureGrid, RCSHartreeFock,
                  SlaterTypeOrbital, TCHartreeFock, UOSHartreeFock,
                  coreMatrixElement, computeF, computeF_batch, fillR12Array, fillAtomicR12Array, atomicCoulombOperator, atomicExchangeOperator, qmc_overlap, qmc_spherical_points, computeF_tc_pair, computeV, eri_full, eri_full_gradient, even_tempered,
                  integrateAllR1R2, integrateAllR1R2Gradient, lebedev, loadSTO, npair, one_electron_matrices, overlap,
                  pair_index, product_grid, qmc_lattice_grid, radial_grid, sobol_grid, sobol_points,
                  threeElectronIntegrals, unpack)

__all__ = [
    "BF16X3", "BF16X4", "BF16X5", "FP32", "FP64", "K_COULOMB", "K_JASTROW_GRAD", "K_TC_SCALAR", "RADIAL_BECKE", "RADIAL_M4",
    "RADIAL_XOVER1MX", "ConvergenceError", "SolverError", "TcscfError", "lib", "HydrogenLikeBasisFunction", "Jastrow",
    "OchiBasisFunction", "OchiJastrow", "QMCGrid",
    "QuadratureGrid", "batch_scf", "RCSHartreeFock", "SlaterTypeOrbital", "TCHartreeFock", "UOSHartreeFock", "coreMatrixElement",
    "computeF", "computeF_batch", "fillR12Array", "fillAtomicR12Array", "atomicCoulombOperator", "atomicExchangeOperator", "qmc_overlap", "qmc_spherical_points", "computeF_tc_pair", "computeV", "eri_full", "eri_full_gradient", "even_tempered", "integrateAllR1R2",
    "integrateAllR1R2Gradient", "lebedev", "loadSTO", "npair", "one_electron_matrices", "overlap", "pair_index", "product_grid",
    "qmc_lattice_grid", "radial_grid", "sobol_grid", "sobol_points", "threeElectronIntegrals", "unpack",
]
