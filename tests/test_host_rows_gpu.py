"""The host-side rows of the path checked in the GPU run: the product's grids,
one-electron matrices and basis-file reader against the reference's own
known-answer tests and the oracle, with the basis tabulated on the device.

The same checks of the oracle live in test_oracle_cpu.py; these run them on
the product (tcscf_amd), marked gpu so that the run on the device host sees
them next to the device parity tests."""
import math

import numpy as np
import pytest

import test_abi_cpu as cpu_checks
from conftest import fns_of

pytestmark = pytest.mark.gpu


def test_radial_rules_kats_and_oracle(tc, O):
    """src/unitTests/IntegrationTest.cpp:19-139 on the product's radial rules:
    the M4 map (a = 1) integrates e^{-r} to 1 (1e-7) and (r^2 - 3r) e^{-3r/2} to
    -20/27 (1e-10) at n = 100; r = (1+x)/(1-x) needs n = 1000 for 1e-6 / 1e-12;
    nodes and weights equal the oracle's to 1e-15."""
    from tcscf_amd import api
    r, w = api.radial_grid(100, tc.RADIAL_M4, 1.0, 1.0)
    assert abs(np.sum(w * np.exp(-r)) - 1.0) < 1e-7
    assert abs(np.sum(w * (r * r - 3 * r) * np.exp(-1.5 * r)) + 20 / 27) < 1e-10
    r, w = api.radial_grid(1000, tc.RADIAL_BECKE, 1.0, 1.0)
    assert abs(np.sum(w * np.exp(-r)) - 1.0) < 1e-6
    assert abs(np.sum(w * (r * r - 3 * r) * np.exp(-1.5 * r)) + 20 / 27) < 1e-12
    cpu_checks.test_radial_and_lebedev_match_oracle(tc, O)


@pytest.mark.parametrize("order", [3, 11, 17, 23, 29, 41])
def test_lebedev_kats_with_device_harmonics(tc, order, gpu):
    """src/unitTests/IntegrationTest.cpp:142-197 on the product's Lebedev
    tables: sum w = 4 pi, and the real harmonics the device tabulates (STOs
    with n = l + 1, zeta = 1 at r = 1) are orthonormal to (l1 + l2 + order)
    3e-15 for l1 + l2 <= order (l <= 3)."""
    from tcscf_amd import SlaterTypeOrbital as STO
    th, ph, w = tc.lebedev(order)
    assert abs(w.sum() / (4 * math.pi) - 1) < order * 3e-15
    pts = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)])
    fns = [STO(1.0, l + 1, l, m) for l in range(4) for m in range(-l, l + 1)]
    g = tc.QMCGrid(pts, w)
    g.setBasisFunctions(fns, False)
    Y = g.basisValues.cpu().numpy()
    nrm = np.array([2.0 ** (f.l + 1) * math.sqrt(2.0 / math.factorial(2 * f.l + 2)) * math.exp(-1.0) for f in fns])
    Y = Y / nrm
    G = np.array([[np.sum(w * Y[:, a] * Y[:, b]) for b in range(Y.shape[1])] for a in range(Y.shape[1])])
    ls = np.array([f.l for f in fns])
    for a in range(len(fns)):
        for b in range(len(fns)):
            if ls[a] + ls[b] <= order:
                assert abs(G[a, b] - (a == b)) < (ls[a] + ls[b] + order) * 3e-15 + 1e-14, (order, a, b)


def test_product_grid_and_sto_overlap_kat(tc, O, gpu):
    """The product grid equals the oracle's, and the reference's STO KAT
    (src/unitTests/SlaterTypeOrbitalTest.cpp:14-47): on the 1000 x Lebedev-11
    M4 grid the device-tabulated basis gives the analytic overlaps of the
    test's pairs to 1e-14 (each sum over the 50,000 points taken pairwise)."""
    from tcscf_amd import SlaterTypeOrbital as STO
    cpu_checks.test_product_grid_matches_oracle(tc, O)
    g = tc.QMCGrid.radial_lebedev(1000, 11, tc.RADIAL_M4, 1.0, 1.0)
    # the reference's pairs: (n, l, m, 1.524) with (n', l, m, 2.532), n' <= n < order / 2
    pairs = [(STO(1.524, n, l, m), STO(2.532, npr, l, m)) for n in range(1, 5) for npr in range(1, n + 1)
             for l in range(0, npr) for m in range(-l, l + 1)]
    fns = [f for pr in pairs for f in pr]
    g.setBasisFunctions(fns, False)
    A = g.basisValues.cpu().numpy()
    S, _ = tc.one_electron_matrices(fns, 1)
    for k in range(len(pairs)):
        num = np.sum(g.weights_host * A[:, 2 * k] * A[:, 2 * k + 1])
        assert abs(num - S[2 * k, 2 * k + 1]) < 1e-14, (k, num, S[2 * k, 2 * k + 1])


def test_one_electron_matrices_closed_forms_and_oracle(tc, O):
    """S and H of the product (src/SlaterTypeOrbital.hpp:125-181): 1s overlaps
    (2 sqrt(za zb)/(za + zb))^3, hydrogenic <h> = -Z^2/2 (1s, zeta = Z) and
    -Z^2/8 (2p, zeta = Z/2), single-zeta <h> = zeta^2/2 - Z zeta; the workload
    bases against the oracle."""
    from tcscf_amd import SlaterTypeOrbital as STO
    z = [0.7, 1.3, 2.9]
    S, H = tc.one_electron_matrices([STO(q, 1, 0, 0) for q in z], 3)
    for a in range(3):
        for b in range(3):
            assert abs(S[a, b] - (2 * math.sqrt(z[a] * z[b]) / (z[a] + z[b])) ** 3) < 1e-14
        assert abs(H[a, a] - (z[a] ** 2 / 2 - 3 * z[a])) < 1e-13
    _, H = tc.one_electron_matrices([STO(3.0, 1, 0, 0)], 3)
    assert abs(H[0, 0] + 4.5) < 1e-14
    _, H = tc.one_electron_matrices([STO(1.5, 2, 1, 0)], 3)
    assert abs(H[0, 0] + 9 / 8) < 1e-14
    cpu_checks.test_one_electron_matrices_match_oracle(tc, O)


def test_load_sto_on_reference_files(tc):
    """loadSTO (src/SlaterTypeOrbital.cpp:24-76) on copies of the reference's
    ADF basis files."""
    cpu_checks.test_load_sto(tc)


def test_qmc_grids_properties(tc):
    """The QMC point sets (lattice and scrambled Sobol): the properties that
    pin them (the exact mppmu/qmc sets are third-party and unpinned)."""
    cpu_checks.test_sobol_points_match_scipy_and_keep_the_net_property(tc)
    cpu_checks.test_sobol_importance_grid_integrates(tc)
    cpu_checks.test_qmc_lattice_grid(tc)
