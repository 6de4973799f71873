"""The benchmark workloads as concrete inputs (atoms, STO bases, grids)."""
from .api import even_tempered
from . import _lib


def he_config1():
    """He, Z=2, RHF: 5 x 1s STOs, zeta_k = 0.5 * 2**k; 25 Gauss-Chebyshev radial
    points mapped by r = (1+x)/(1-x) x Lebedev order 11 (50) = 1,250 points."""
    basis = even_tempered({0: 5}, 0.5, 2.0)
    return dict(name="He", Z=2, nup=1, ndown=1, basis=basis, grid=(25, 11, _lib.RADIAL_BECKE))


def ne_config2():
    """Ne, Z=10, RHF: 30 STOs, zeta = 0.3 * 1.8**k per l with s k=0..12,
    p k=0..3, d k=0; 100 radial (Gauss-Chebyshev, Treutler-Ahlrichs M4 map,
    a = 1, eps = 1) x Lebedev order 29 (302) = 30,200 points."""
    basis = even_tempered({0: 13, 1: 4, 2: 1}, 0.3, 1.8)
    assert len(basis) == 30
    return dict(name="Ne", Z=10, nup=5, ndown=5, basis=basis, grid=(100, 29, _lib.RADIAL_M4))


def be_config3():
    """Be, Z=4, TC: 20 STOs (s k=0..10, p k=0..2), zeta = 0.3 * 1.8**k."""
    basis = even_tempered({0: 11, 1: 3}, 0.3, 1.8)
    assert len(basis) == 20
    return dict(name="Be", Z=4, nup=2, ndown=2, basis=basis, grid=(100, 29, _lib.RADIAL_M4),
                grid3=(50, 23, _lib.RADIAL_M4))


def ne_config3():
    c = ne_config2()
    c["grid3"] = (50, 23, _lib.RADIAL_M4)
    return c


def ne_qmc_config4():
    """Ne, Z=10, TC on scrambled-Sobol grids: 40 STOs (s k=0..14, p k=0..4,
    d k=0..1, zeta = 0.3 * 1.8**k); 2**16 points per electron coordinate (r1,
    r2, r3 with seeds 1, 2, 3), r from the inverse CDF of 2 e^{-2r}, uniform
    directions, importance weights; Jastrow b in {0.5, 1, 2}."""
    basis = even_tempered({0: 15, 1: 5, 2: 2}, 0.3, 1.8)
    assert len(basis) == 40
    return dict(name="Ne", Z=10, nup=5, ndown=5, basis=basis, npts=2 ** 16, seeds=(1, 2, 3), b_values=(0.5, 1.0, 2.0))
