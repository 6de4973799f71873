"""The benchmark workloads as concrete inputs (atoms, STO bases, grids)."""
import numpy as np

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


BATCH_SEED = 20251203
LEBEDEV_ORDERS = {11: 50, 17: 110, 23: 194, 29: 302}


def batch_instances(count=64, seed=BATCH_SEED):
    """The batched-serving workload: `count` uneven instances drawn from a fixed
    seed.  Atoms and singly charged cations, Z = 2..18; even-tempered STO bases
    (zeta = 0.3 * 1.8**k per l, n = l + 1) of 5..60 functions (_batch_shells);
    radial x Lebedev grids of
    1,250..30,200 points (25/50/75/100 radial points, M4 map, Lebedev 11/17/23/
    29); instances alternate plain HF (RCS if closed shell, else UOS) and TC
    with b ~ U[0.5, 2] (the r3 grid is the r1 grid).  Returns plain dicts."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        Z = int(rng.integers(2, 19))
        charge = int(rng.integers(0, 2)) if Z > 2 else 0
        ne = Z - charge
        nup, ndown = (ne + 1) // 2, ne // 2
        nb_target = int(rng.integers(max(5, nup + 2), 61))
        shells = _batch_shells(Z, nb_target)
        nrad = int(rng.choice([25, 50, 75, 100]))
        order = int(rng.choice(sorted(LEBEDEV_ORDERS)))
        method = "tc" if k % 2 else "hf"
        b = float(rng.uniform(0.5, 2.0)) if method == "tc" else None
        out.append(dict(index=k, Z=Z, charge=charge, nup=nup, ndown=ndown, shells={int(l): int(c) for l, c in shells.items()},
                        nb=sum(c * (2 * l + 1) for l, c in shells.items()), grid=[nrad, order],
                        npts=nrad * LEBEDEV_ORDERS[order], method=method, b=b))
    return out


def _batch_shells(Z, nb):
    """Shells for about nb functions: ~40% s (3..15 exponents, so the tightest
    stays within what the grids resolve), then up to 6 p, 4 d and 2 f shells
    (l <= 1 for Z <= 2, <= 2 for Z <= 10, <= 3 beyond); functions that fit no
    shell go to s (up to 15) or are dropped."""
    lmax = 1 if Z <= 2 else (2 if Z <= 10 else 3)
    caps = {0: 15, 1: 6, 2: 4, 3: 2}
    ns = min(15, max(3, int(round(0.4 * nb))))
    rem = nb - ns
    shells = {0: ns}
    for l in (1, 2, 3):
        if l > lmax:
            break
        c = min(caps[l], rem // (2 * l + 1))
        if c:
            shells[l] = c
            rem -= c * (2 * l + 1)
    shells[0] = min(15, shells[0] + rem)
    return shells


def batch_basis(instance):
    return even_tempered({int(l): int(c) for l, c in instance["shells"].items()}, 0.3, 1.8)
