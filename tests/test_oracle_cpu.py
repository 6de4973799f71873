"""Pins the FP64 CPU oracle against the reference's own known-answer tests,
closed forms and LAPACK.  (The reference's tests are gtest C++; the ones
restated here are cited by file:line.)"""
import math

import numpy as np
import pytest
import scipy.linalg as sl

from conftest import rel_err


# ----------------------------------------------------------------- quadrature
def test_chebyshev_gauss_kats(O):
    """src/unitTests/IntegrationTest.cpp:19-73."""
    x, w = O.chebyshev_gauss(100)
    assert abs(np.sum(w * x)) < 3e-16
    x, w = O.chebyshev_gauss(1000)
    assert abs(np.sum(w * x * x) - 2 / 3) < 1e-5
    assert abs(np.sum(w * (x ** 3 + 3 * x ** 2 - 4 * x)) - 2.0) < 1e-5


def test_radial_maps_kats(O):
    """src/unitTests/IntegrationTest.cpp:76-139 integrate e^{-r} = 1 and
    (r^2 - 3r) e^{-3r/2} = -20/27 after a radial map (the test's one-parameter
    map is not in the tree): the M4 map (a = 1) meets the test's tolerances at
    n = 100; the Becke map r = (1+x)/(1-x) with the same nodes needs n = 1000
    for 1e-6 / 1e-12."""
    r, w = O.radial_grid(100, O.RADIAL_M4, 1.0, 1.0)
    assert abs(np.sum(w * np.exp(-r)) - 1.0) < 1e-7
    assert abs(np.sum(w * (r * r - 3 * r) * np.exp(-1.5 * r)) + 20 / 27) < 1e-10
    r, w = O.radial_grid(1000, O.RADIAL_BECKE, 1.0, 1.0)
    assert abs(np.sum(w * np.exp(-r)) - 1.0) < 1e-6
    assert abs(np.sum(w * (r * r - 3 * r) * np.exp(-1.5 * r)) + 20 / 27) < 1e-12


def test_radial_jacobian_is_derivative(O):
    """The stored weight is sin(theta) pi/(n+1) * dr/dx: compare dr/dx with a
    central difference of the map r(x) through a shifted grid."""
    x, wx = O.chebyshev_gauss(64)
    for kind, a, eps in ((O.RADIAL_M4, 1.0, 0.9), (O.RADIAL_BECKE, 1.0, 1.0), (O.RADIAL_XOVER1MX, 1.0, 1.0)):
        r, w = O.radial_grid(64, kind, a, eps)
        jac = w / wx
        if kind == O.RADIAL_M4:
            f = lambda t: eps / math.log(2) * (t + a) ** 0.6 * np.log((1 + a) / (1 - t))
        elif kind == O.RADIAL_BECKE:
            f = lambda t: (1 + t) / (1 - t)
        else:
            f = lambda t: t / (1 - t)
        inner = (x > -0.99) & (x < 0.99)
        h = 1e-6
        fd = (f(x + h) - f(x - h)) / (2 * h)
        assert np.max(np.abs(jac[inner] - fd[inner]) / np.abs(fd[inner])) < 1e-7
        assert np.allclose(r, f(x), rtol=1e-14, atol=0)


@pytest.mark.parametrize("order", [3, 5, 7, 9, 11, 13, 15, 17, 19, 21, 23, 25, 27, 29, 31, 35, 41])
def test_lebedev_sum_and_orthonormality(O, order):
    """src/unitTests/IntegrationTest.cpp:142-197: sum w = 4 pi and Y_lm
    orthonormality to (l1 + l2 + order) * 3e-15 for l1 + l2 <= order."""
    th, ph, w = O.lebedev(order)
    # Y_00 normalisation (the l1 = l2 = 0 case of the test); order 3 also to 3e-14 absolute
    assert abs(w.sum() / (4 * math.pi) - 1) < order * 3e-15
    if order == 3:
        assert abs(w.sum() - 4 * math.pi) < 3e-14
    pts = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)])
    rng = np.random.default_rng(order)
    # Y_lm through the oracle's STO at r = 1: phi / (N e^{-zeta}) with n = l + 1
    for _ in range(20):
        l1 = int(rng.integers(0, min(order, 6) + 1))
        l2 = int(rng.integers(0, min(order - l1, 6) + 1))
        m1 = int(rng.integers(-l1, l1 + 1))
        m2 = int(rng.integers(-l2, l2 + 1))
        v = O.sto_complex(pts, [(l1 + 1, l1, m1, 1.0), (l2 + 1, l2, m2, 1.0)])
        nrm = [(2.0) ** (l + 1) * math.sqrt(2.0 / math.factorial(2 * l + 2)) * math.exp(-1.0) for l in (l1, l2)]
        y1, y2 = v[:, 0] / nrm[0], v[:, 1] / nrm[1]
        val = np.sum(w * np.conj(y1) * y2)
        delta = float(l1 == l2 and m1 == m2)
        assert abs(val - delta) < (l1 + l2 + order) * 3e-15 + 1e-14


# ----------------------------------------------------------------- STO
def test_sto_overlap_kat(O):
    """src/unitTests/SlaterTypeOrbitalTest.cpp:14-47: numerical overlap on
    TreutlerAhlrichsLebedev(1.0, 1000, 11) equals the analytic overlap to 1e-14."""
    pts, w = O.product_grid(1000, 11, O.RADIAL_M4, 1.0, 1.0)
    order = 11
    for n in range(1, order // 2):
        for npr in range(1, n + 1):
            for l in range(0, npr + 1):
                for m in range(-l, l + 1):
                    f1 = (n, l, m, 1.524)
                    f2 = (npr, l, m, 2.532)
                    if l > n:
                        continue
                    v = O.sto_complex(pts, [f1, f2])
                    num = np.sum(w * np.conj(v[:, 0]) * v[:, 1])
                    S, _ = O.sto_S_H([f1, f2], 1)
                    assert abs(num.real - S[0, 1]) < 1e-14, (f1, f2)
                    assert abs(num.imag) < 1e-14


def test_sto_S_H_closed_forms(O):
    """1s STOs: S_ab = (2 sqrt(za zb)/(za + zb))^3; a hydrogenic 1s with
    zeta = Z has <h> = -Z^2/2; single-zeta <h> = zeta^2/2 - Z zeta."""
    z = [0.7, 1.3, 2.9]
    S, H = O.sto_S_H([(1, 0, 0, q) for q in z], 3)
    for a in range(3):
        for b in range(3):
            assert abs(S[a, b] - (2 * math.sqrt(z[a] * z[b]) / (z[a] + z[b])) ** 3) < 1e-14
        assert abs(H[a, a] - (z[a] ** 2 / 2 - 3 * z[a])) < 1e-13
    S, H = O.sto_S_H([(1, 0, 0, 3.0)], 3)
    assert abs(H[0, 0] + 4.5) < 1e-14
    # 2p hydrogenic with zeta = Z/2 is an eigenfunction: <h> = -Z^2/8
    S, H = O.sto_S_H([(2, 1, 0, 1.5)], 3)
    assert abs(H[0, 0] + 9 / 8) < 1e-14
    # H_ab symmetric although the formula is evaluated on b
    fns = [(1, 0, 0, 0.8), (2, 0, 0, 1.7), (3, 0, 0, 2.2), (2, 1, 1, 1.1), (3, 1, 1, 0.6), (3, 2, -1, 1.4)]
    S, H = O.sto_S_H(fns, 5)
    assert np.allclose(H, H.T, atol=1e-13)
    assert np.allclose(S, S.T, atol=1e-15)


def test_sto_core_matrix_numerically(O):
    """H_ab = <a| -1/2 lap - Z/r |b> checked by quadrature using the oracle's
    own gradients: <grad a . grad b>/2 - Z <a|1/r|b>.  The reference's
    spherical gradient formula is singular at the poles (src/SlaterTypeOrbital.hpp:73-74),
    which Lebedev grids contain, so the grid is rotated first (a rotated
    Lebedev rule integrates the same harmonics exactly)."""
    fns = [(1, 0, 0, 0.9), (2, 0, 0, 1.6), (2, 1, 1, 1.2), (3, 1, 1, 0.7), (3, 2, 0, 1.1), (3, 2, 0, 2.3)]
    Z = 3
    pts, w = O.product_grid(400, 17, O.RADIAL_M4, 1.0, 1.0)
    Q, _ = np.linalg.qr(np.random.default_rng(11).normal(size=(3, 3)))
    pts = np.ascontiguousarray(Q @ pts)
    v, g = O.sto_real(pts, fns, grads=True)
    r = np.linalg.norm(pts, axis=0)
    T = 0.5 * np.einsum("n,nad,nbd->ab", w, g, g)
    V = -Z * np.einsum("n,na,nb->ab", w / r, v, v)
    S, H = O.sto_S_H(fns, Z)
    assert np.max(np.abs(T + V - H)) < 1e-9
    assert np.max(np.abs(np.einsum("n,na,nb->ab", w, v, v) - S)) < 1e-10


def test_sto_gradient_finite_difference(O):
    """The reference has no STO-gradient test; central differences as in
    src/unitTests/gradientsTest.cpp:10-20 (h = 2^-15)."""
    rng = np.random.default_rng(1)
    fns = [(n, l, m, z) for (n, l, z) in ((1, 0, 1.3), (2, 1, 0.8), (3, 2, 1.7), (4, 3, 1.1)) for m in range(-l, l + 1)]
    pts = rng.uniform(-2, 2, size=(3, 40))
    v, g = O.sto_real(pts, fns, grads=True)
    h = 2.0 ** -15
    for d in range(3):
        e = np.zeros((3, 1))
        e[d] = h
        fd = (O.sto_real(pts + e, fns) - O.sto_real(pts - e, fns)) / (2 * h)
        assert np.max(np.abs(fd - g[:, :, d])) < 1e-8


def test_sto_gradient_on_the_z_axis(O):
    """Where the reference's spherical gradient is 0/0 (theta = 0) the oracle
    uses its limit; central differences agree there too."""
    fns = [(n, l, m, z) for (n, l, z) in ((1, 0, 1.3), (2, 1, 0.8), (3, 2, 1.7), (4, 3, 1.1)) for m in range(-l, l + 1)]
    pts = np.array([[0.0, 0.0, 0.0, 0.0], [0.0, 0.0, 0.0, 0.0], [0.3, 1.1, -0.7, -1.9]])
    v, g = O.sto_real(pts, fns, grads=True)
    assert np.all(np.isfinite(g))
    for d in range(3):  # Richardson-extrapolated central differences
        fd = []
        for h in (2.0 ** -10, 2.0 ** -11):
            e = np.zeros((3, 1))
            e[d] = h
            fd.append((O.sto_real(pts + e, fns) - O.sto_real(pts - e, fns)) / (2 * h))
        fdx = (4 * fd[1] - fd[0]) / 3
        assert np.max(np.abs(fdx - g[:, :, d])) < 1e-9


def test_real_basis_is_unitary_change_of_complex(O):
    """The real STOs are (-1)^m (phi_m +/- phi_-m)/sqrt2 combinations of the
    reference's complex ones (phi_-m = conj phi_m)."""
    rng = np.random.default_rng(2)
    pts = rng.uniform(-2, 2, size=(3, 30))
    for l in (1, 2, 3):
        fc = [(l + 1, l, m, 1.2) for m in range(-l, l + 1)]
        vc = O.sto_complex(pts, fc)
        vr = O.sto_real(pts, fc)
        for m in range(1, l + 1):
            s = (-1) ** m
            plus = s * (vc[:, l + m] + vc[:, l - m]) / math.sqrt(2)
            minus = s * (vc[:, l + m] - vc[:, l - m]) / (math.sqrt(2) * 1j)
            assert np.allclose(plus.real, vr[:, l + m], atol=1e-14) and np.allclose(plus.imag, 0, atol=1e-14)
            assert np.allclose(minus.real, vr[:, l - m], atol=1e-14) and np.allclose(minus.imag, 0, atol=1e-14)


# ----------------------------------------------------------------- Jastrow
def test_ochi_gradient_kat(O):
    """src/unitTests/gradientsTest.cpp:59-93: FD (h = 2^-15) vs analytic to 1e-8,
    a = a12 = 1.5, c = (a12/2, a12/4), r1 = (1,1,1), r2 ~ U(-5, 5)^3."""
    a = 1.5
    prm = O.ochi_params(a, a, [(1, 0, 0, a / 2, a / 4)])
    rng = np.random.default_rng(5489)
    r1 = np.array([1.0, 1.0, 1.0])
    h = 2.0 ** -15
    for _ in range(100):
        r2 = rng.uniform(-5, 5, 3)
        f = lambda p: O.ochi_value(prm, np.linalg.norm(p), np.linalg.norm(r2), np.linalg.norm(p - r2), False)
        fd = np.array([(f(r1 + h * e) - f(r1 - h * e)) / (2 * h) for e in np.eye(3)])
        assert np.max(np.abs(O.ochi_gradient(prm, r1, r2, False) - fd)) < 1e-8


def test_ochi_r12_term_closed_form(O):
    """u = c r12/(r12 + a12) gives grad1 u = c a12 / (r12 (r12 + a12)^2) (r1 - r2);
    the general terms (r1, r2 powers) also pass the FD check."""
    a12, c = 0.8, 0.37
    prm = O.ochi_params(1.1, a12, [(1, 0, 0, c, c / 2)])
    rng = np.random.default_rng(3)
    for _ in range(50):
        r1, r2 = rng.uniform(-3, 3, 3), rng.uniform(-3, 3, 3)
        r12 = np.linalg.norm(r1 - r2)
        g = O.ochi_gradient(prm, r1, r2, False)
        assert np.allclose(g, c * a12 / (r12 * (r12 + a12) ** 2) * (r1 - r2), rtol=1e-14, atol=1e-15)
    prm = O.ochi_params(1.1, a12, [(1, 0, 0, c, c / 2), (2, 1, 0, 0.2, 0.1), (0, 2, 1, -0.3, 0.4)])
    h = 2.0 ** -15
    for _ in range(20):
        r1, r2 = rng.uniform(-3, 3, 3), rng.uniform(-3, 3, 3)
        f = lambda p: O.ochi_value(prm, np.linalg.norm(p), np.linalg.norm(r2), np.linalg.norm(p - r2), True)
        fd = np.array([(f(r1 + h * e) - f(r1 - h * e)) / (2 * h) for e in np.eye(3)])
        assert np.max(np.abs(O.ochi_gradient(prm, r1, r2, True) - fd)) < 1e-8


# ----------------------------------------------------------------- computeF / ERI
def _eri_1s(za, zb, zc, zd):
    """(ab|cd) for normalized 1s STOs: N^4 (1/4pi)^2 32 pi^2 (p^2+3pq+q^2)/(p^2 q^2 (p+q)^3)."""
    N = [2 * z ** 1.5 for z in (za, zb, zc, zd)]
    p, q = za + zb, zc + zd
    return np.prod(N) / (4 * math.pi) ** 2 * 32 * math.pi ** 2 * (p * p + 3 * p * q + q * q) / (p * p * q * q * (p + q) ** 3)


def test_compute_F_against_closed_form_eri(O):
    """Grid ERI (ab|cd) = sum_r1 w1 phi_a phi_b F[r1, c, d] converges to the
    1s closed form as the grid grows; the error shrinks with the grid."""
    z = [0.8, 1.6]
    fns = [(1, 0, 0, q) for q in z]
    errs = []
    for nrad, order in ((20, 7), (40, 11), (60, 17)):
        pts, w = O.product_grid(nrad, order)
        A = O.sto_real(pts, fns)
        F = O.compute_F(pts, pts, w, A)
        eri = O.eri_from_F(A, w, F)
        exact = np.array([[[[_eri_1s(z[a], z[b], z[c], z[d]) for d in range(2)] for c in range(2)] for b in range(2)]
                          for a in range(2)])
        # eri index order (b1, b2, b3, b4) = (a, c, b, d) in chemist notation
        errs.append(np.max(np.abs(eri.transpose(0, 2, 1, 3) - exact)) / np.max(exact))
    # a dropped 2 pi, a missing weight or a transposed index is an O(1) error;
    # the coincident-pair-excluded double sum converges slowly (~1/N^(1/3))
    assert errs[-1] < 5e-3
    assert errs[0] > errs[1] > errs[2]


def test_compute_F_scale_and_symmetry(O):
    """F is symmetric in (j, i); scale multiplies; the vector kernel is odd
    under exchanging the grids' roles; coincident pairs contribute nothing."""
    rng = np.random.default_rng(4)
    pts = rng.uniform(-2, 2, (3, 60))
    w = rng.uniform(0.1, 1.0, 60)
    fns = [(1, 0, 0, 1.0), (2, 1, 1, 0.7), (2, 1, -1, 0.7)]
    A = O.sto_real(pts, fns)
    F = O.compute_F(pts, pts, w, A)
    assert np.array_equal(F, F.transpose(0, 2, 1))
    F2 = O.compute_F(pts, pts, w, A, scale=2 * math.pi)
    assert np.allclose(F2, 2 * math.pi * F, rtol=1e-15)
    # explicit double loop with the diagonal excluded
    d = np.linalg.norm(pts[:, :, None] - pts[:, None, :], axis=0)
    K = np.where(d > 0, 1 / np.where(d > 0, d, 1), 0.0) * w[None, :]
    ref = np.einsum("ab,bj,bi->aji", K, A, A)
    assert rel_err(F, ref) < 1e-13


def test_tc_scalar_kernel_reduces_to_coulomb(O):
    """With c = 0 the TC scalar kernel is exactly 2/r12."""
    rng = np.random.default_rng(5)
    pts = rng.uniform(-2, 2, (3, 40))
    w = rng.uniform(0.1, 1.0, 40)
    A = O.sto_real(pts, [(1, 0, 0, 1.0), (2, 0, 0, 0.5)])
    F = O.compute_F(pts, pts, w, A)
    prm = O.ochi_params(1.0, 1.0, [(1, 0, 0, 0.0, 0.0)])
    Ft = O.compute_F(pts, pts, w, A, kernel=O.K_TC_SCALAR, prm=prm)
    assert rel_err(Ft, 2 * F) < 1e-14


# ----------------------------------------------------------------- dense algebra
def _rand_spd(n, rng):
    M = rng.normal(size=(n, n))
    return M @ M.T + n * np.eye(n)


def test_gen_sym_eig_matches_lapack(O):
    """The oracle's Cholesky + Jacobi solver against LAPACK dsygvd (scipy),
    the role hegvd plays in the reference (src/HartreeFock.cpp:869-876)."""
    rng = np.random.default_rng(6)
    for n in (1, 5, 17, 40):
        S = _rand_spd(n, rng)
        F = rng.normal(size=(n, n))
        F = F + F.T
        e, C = O.gen_sym_eig(F, S)
        e2 = sl.eigh(F, S, eigvals_only=True)
        assert np.allclose(e, e2, rtol=1e-12, atol=1e-12)
        assert np.allclose(C.T @ S @ C, np.eye(n), atol=1e-11)
        assert np.allclose(F @ C, S @ C @ np.diag(e), atol=1e-10)


def test_gen_nonsym_eig_matches_lapack(O):
    """The oracle's QR + inverse-iteration solver against LAPACK dggev
    (scipy.linalg.eig(F, S)), the role ggev plays (src/HartreeFock.cpp:1277-1291)."""
    rng = np.random.default_rng(7)
    for n in (1, 4, 12, 30):
        S = _rand_spd(n, rng)
        F = rng.normal(size=(n, n)) + 3 * np.diag(np.arange(n))
        e, X = O.gen_nonsym_eig(F, S)
        e2 = sl.eig(F, S, right=False)
        left = list(e2)
        for x in e:  # greedy nearest match (conjugate pairs sort unstably)
            k = int(np.argmin([abs(x - y) for y in left]))
            assert abs(x - left[k]) < 1e-10 * max(1.0, abs(x))
            left.pop(k)
        for k in range(n):
            r = F @ X[:, k] - e[k] * (S @ X[:, k])
            assert np.linalg.norm(r) < 1e-9 * np.linalg.norm(X[:, k]) * (np.abs(F).max() + abs(e[k]) * np.abs(S).max())


def test_block_solve_gives_block_pure_vectors(O):
    """With a block-diagonal pencil the eigenvectors stay inside one block even
    for degenerate eigenvalues (two identical blocks)."""
    rng = np.random.default_rng(8)
    S1 = _rand_spd(3, rng)
    F1 = rng.normal(size=(3, 3))
    F1 = F1 + F1.T
    S = sl.block_diag(S1, S1)
    F = sl.block_diag(F1, F1)
    e, C = O.gen_sym_eig(F, S)
    for k in range(6):
        top, bot = np.abs(C[:3, k]).max(), np.abs(C[3:, k]).max()
        assert min(top, bot) == 0.0


def test_mgs_modes(O):
    """src/unitTests/orthogonalizationTest.cpp:48-62 (plain MGS: off-diagonal
    dots below size * 1e-15), and the metric variant's quirk: vector 0 is
    metric-normalized, later ones Euclidean-normalized (src/orthogonalization.hpp:253-271)."""
    rng = np.random.default_rng(9)
    for k, n in ((5, 5), (10, 10), (2, 5), (8, 20)):
        v = rng.uniform(-10, 10, (k, n)).astype(complex)
        out = O.mgs(v, np.eye(n), mode=0)
        G = np.conj(out) @ out.T
        assert np.max(np.abs(G - np.diag(np.diag(G)))) < n * 1e-15 * 10
        assert np.allclose(np.diag(G), 1, atol=1e-14)
    S = _rand_spd(6, rng) / 6
    v = rng.normal(size=(3, 6)).astype(complex)
    out = O.mgs(v, S, mode=0)
    assert abs(np.conj(out[0]) @ S @ out[0] - 1) < 1e-13
    for i in (1, 2):
        assert abs(np.linalg.norm(out[i]) - 1) < 1e-13
    out = O.mgs(v, S, mode=1)
    M = np.conj(out) @ S @ out.T
    assert np.allclose(M, np.eye(3), atol=1e-12)


# ----------------------------------------------------------------- SCF
def test_he_single_zeta_closed_form(O):
    """E(zeta) = zeta^2 - 2 Z zeta + 5 zeta / 8 for one 1s STO; the grid SCF
    converges to -(27/16)^2 as the grid grows."""
    z = 27 / 16
    errs = []
    for nrad, order in ((25, 11), (50, 23)):
        pts, w = O.product_grid(nrad, order)
        A = O.sto_real(pts, [(1, 0, 0, z)])
        S, H = O.sto_S_H([(1, 0, 0, z)], 2)
        F = O.compute_F(pts, pts, w, A)
        r = O.rcs_scf(2, S, H, A, w, F, tol=1e-12)
        assert r["status"] == 0
        errs.append(abs(r["energy"] + z * z))
    assert errs[1] < errs[0] < 0.03


def test_he_ochi_fixture_exact_energies():
    """The grid-free RHF energies of He in the Ochi s basis (tests/golden/
    make_he_ochi.py) against the reference's recorded QMC-grid runs
    (convergence.txt): each within 2 of that run's standard deviations; and
    variational: decreasing with the basis size and above the HF limit."""
    import json
    import os
    data = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "he_ochi_convergence.json")))
    rows = data["rows"]
    assert [r["nBasis"] for r in rows] == list(range(2, 11))
    for r in rows:
        assert abs(r["energy"] - r["exact_energy"]) < 2 * r["stddev"], r
        assert r["exact_energy"] > data["hf_limit"] - 1e-7
    e = [r["exact_energy"] for r in rows]
    assert all(a > b for a, b in zip(e, e[1:]))
    assert e[-1] - data["hf_limit"] < 1e-6


def _he_config1(O):
    fns = [(1, 0, 0, 0.5 * 2 ** k) for k in range(5)]
    pts, w = O.product_grid(25, 11)
    A = O.sto_real(pts, fns)
    S, H = O.sto_S_H(fns, 2)
    return fns, pts, w, A, S, H, O.compute_F(pts, pts, w, A)


def test_he_config1_rcs_uos_agree_and_invariants(O):
    fns, pts, w, A, S, H, F = _he_config1(O)
    for mask in (True, False):
        r = O.rcs_scf(2, S, H, A, w, F, tol=1e-10, mask=mask)
        u = O.uos_scf(1, 1, S, H, A, w, F, tol=1e-10, mask=mask)
        assert r["status"] == 0 and u["status"] == 0
        assert abs(r["energy"] - u["energy"]) < 1e-10
        D = r["D"]
        assert abs(np.trace(D @ S) - 1) < 1e-12
        assert np.allclose(D @ S @ D, D, atol=1e-11)
        assert np.allclose(u["D"][0], u["D"][1], atol=1e-12)
    # the reference's symmetry mask zeroes G[3,3] because H[3,3] == 0 exactly
    # (zeta = 4 = 2 Z), which collapses the variational problem
    assert H[3, 3] == 0.0
    r_mask = O.rcs_scf(2, S, H, A, w, F, tol=1e-10, mask=True)
    r_free = O.rcs_scf(2, S, H, A, w, F, tol=1e-10, mask=False)
    assert r_mask["energy"] < -30 and -2.9 < r_free["energy"] < -2.85


def test_scf_nonconvergence_and_solver_errors(O):
    """Failure paths forced from the host: an iteration cap of 1 gives
    ConvergenceError (-1 here); a duplicated basis function makes S singular."""
    fns, pts, w, A, S, H, F = _he_config1(O)
    r = O.uos_scf(1, 1, S, H, A, w, F, max_iter=1)
    assert r["status"] == -1
    fns2 = fns + [fns[0]]
    S2, H2 = O.sto_S_H(fns2, 2)
    A2 = O.sto_real(pts, fns2)
    F2 = O.compute_F(pts, pts, w, A2)
    r = O.rcs_scf(2, S2, H2, A2, w, F2)
    assert r["status"] == -2


def test_tc_with_zero_jastrow_equals_uos(O):
    """c_same = c_oppo = 0: the TC scalar kernel is 2/r12 and V = 0, so
    TCHartreeFock reduces to UOSHartreeFock (the /2 at HartreeFock.cpp:1209
    restores 1/r12)."""
    fns = [(1, 0, 0, 0.6), (1, 0, 0, 1.4), (2, 0, 0, 1.0), (2, 1, 1, 1.2), (2, 1, -1, 1.2), (2, 1, 0, 1.2)]
    pts, w = O.product_grid(20, 11)
    A, dA = O.sto_real(pts, fns, grads=True)
    S, H = O.sto_S_H(fns, 4)
    F = O.compute_F(pts, pts, w, A)
    u = O.uos_scf(2, 2, S, H, A, w, F, tol=1e-10)
    V = np.zeros(F.shape + (3,))
    for mgs_mode in (0, 1):
        t = O.tc_scf(2, 2, S, H, A, dA, w, 2 * F, 2 * F, V, V, tol=1e-10, mgs_mode=mgs_mode)
        assert t["status"] == 0
        if mgs_mode == 1:
            assert abs(t["energy"] - u["energy"]) < 1e-9
    # He-like (one electron per spin): both MGS modes coincide
    t0 = O.tc_scf(1, 1, S, H, A, dA, w, 2 * F, 2 * F, V, V, tol=1e-10, mgs_mode=0)
    t1 = O.tc_scf(1, 1, S, H, A, dA, w, 2 * F, 2 * F, V, V, tol=1e-10, mgs_mode=1)
    u1 = O.uos_scf(1, 1, S, H, A, w, F, tol=1e-10)
    assert abs(t0["energy"] - t1["energy"]) < 1e-12
    assert abs(t0["energy"] - u1["energy"]) < 1e-9


# ------------------------------------------- three-body: factorized vs explicit
def _tc_small_case(O, nb_fns, npts_r1=(4, 7), npts_r3=(5, 5), seed=7, nocc=(2, 1)):
    """Small TC inputs: r1 grid, V_same / V_oppo from the Jastrow-gradient
    kernel on an r3 grid (c_oppo = 2 c_same, as the reference's ratio), random
    occupied orbitals C and D = C C^T per spin."""
    pts1, w1 = O.product_grid(*npts_r1)
    pts3, w3 = O.product_grid(*npts_r3)
    A1, dA1 = O.sto_real(pts1, nb_fns, grads=True)
    A3 = O.sto_real(pts3, nb_fns)
    prm = O.ochi_params(1.0, 1.0, [(1, 0, 0, 0.5, 0.25)])
    Vs = O.compute_F(pts1, pts3, w3, A3, kernel=O.K_JASTROW_GRAD, prm=prm, same=True)
    Vo = O.compute_F(pts1, pts3, w3, A3, kernel=O.K_JASTROW_GRAD, prm=prm, same=False)
    Fs = O.compute_F(pts1, pts1, w1, A1, kernel=O.K_TC_SCALAR, prm=prm, same=True)
    Fo = O.compute_F(pts1, pts1, w1, A1, kernel=O.K_TC_SCALAR, prm=prm, same=False)
    nb = len(nb_fns)
    rng = np.random.default_rng(seed)
    C = np.zeros((2, nb, max(nocc)))
    D = np.zeros((2, nb, nb))
    for s in range(2):
        C[s, :, :nocc[s]] = rng.normal(size=(nb, nocc[s])) * 0.5
        D[s] = C[s, :, :nocc[s]] @ C[s, :, :nocc[s]].T
    return dict(A1=A1, dA1=dA1, w1=w1, Fs=Fs, Fo=Fo, Vs=Vs, Vo=Vo, C=C, D=D, nocc=nocc)


@pytest.mark.parametrize("nocc", [(1, 1), (2, 1), (2, 2)])
def test_factorized_G3_equals_explicit_Q(O, nocc):
    """The reference's own disabled cross-check (src/HartreeFock.cpp:1098-1116,
    1227-1247) of the factorized three-body term G3 against assembleG over the
    explicit six-index tensors of computeQ (src/HartreeFock.cpp:613-750).

    - With the pair tensors taken from the electrons' spins (QSO = (s,s,t),
      QOO = (s,t,t)), assembleG equals the factorized G3 for arbitrary,
      independent V_same and V_oppo and different orbitals per spin.
    - computeQ as the reference writes it labels the pairs differently (its
      so312 term reads V_same(j,i), QOO is all V_oppo): it agrees only when
      V_oppo = V_same, and differs at the reference's c_oppo = 2 c_same, which
      is why that check stays disabled there.
    - The identities assembleG asserts (2 QSS = QSO, 4 QSS = QOO,
      Q(j,p,r,i,q,s) = Q(j,r,p,i,s,q)) hold for the reference's computeQ at
      c_oppo = 2 c_same."""
    fns = [(1, 0, 0, 1.3), (2, 1, 0, 0.9), (2, 0, 0, 0.7)]
    c = _tc_small_case(O, fns, nocc=nocc)
    args = (c["A1"], c["dA1"], c["w1"], c["Fs"], c["Fo"])
    Vs, Vo, D, C = c["Vs"], c["Vo"], c["D"], c["C"]

    QSS, QSO, QOO = O.compute_Q(c["A1"], c["w1"], Vs, Vo)
    scale = np.max(np.abs(QSS))
    assert np.max(np.abs(2 * QSS - QSO)) < 1e-13 * scale
    assert np.max(np.abs(4 * QSS - QOO)) < 1e-13 * scale
    assert np.max(np.abs(QSS - QSS.transpose(0, 2, 1, 3, 5, 4))) < 1e-13 * scale

    for Vo_case in (Vo, 0.7 * Vo, Vs):
        _, G3 = O.tc_fock(*nocc, *args, Vs, Vo_case, D, C)
        assert np.max(np.abs(G3)) > 1e-7
        G3q = O.assemble_G3(*O.compute_Q(c["A1"], c["w1"], Vs, Vo_case, labelled=True), D)
        assert rel_err(G3, G3q) < 1e-12
        G3r = O.assemble_G3(*O.compute_Q(c["A1"], c["w1"], Vs, Vo_case, labelled=False), D)
        if Vo_case is Vs:
            assert rel_err(G3, G3r) < 1e-12
        else:
            assert rel_err(G3, G3r) > 1e-2

    # planted mistakes in the factorized inputs are visible
    G3q = O.assemble_G3(*O.compute_Q(c["A1"], c["w1"], Vs, Vo, labelled=True), D)
    _, G3x = O.tc_fock(*nocc, *args, Vs, 0 * Vo, D, C)
    assert rel_err(G3x, G3q) > 1e-2
    Cx = C.copy()
    Cx[0] *= -1.1          # orbitals inconsistent with D
    _, G3x = O.tc_fock(*nocc, *args, Vs, Vo, D, Cx)
    assert rel_err(G3x, G3q) > 1e-3


def _ochi_grad1_closed_form(terms, a, a12, same, p1, p2):
    """grad_1 u(r1, r2) for every point pair (n1, n2, 3), by the chain rule on
    u = sum_t c_t rb12^i rb1^j rb2^k with rb = r / (r + a) (rb12 with a12):
    du/dr12 (r1 - r2) / r12 + du/dr1 r1 / |r1|."""
    d = p1.T[:, None, :] - p2.T[None, :, :]
    r12 = np.linalg.norm(d, axis=-1)
    r1 = np.linalg.norm(p1.T, axis=-1)[:, None]
    r2 = np.linalg.norm(p2.T, axis=-1)[None, :]
    t12, t1, t2 = r12 / (r12 + a12), r1 / (r1 + a), r2 / (r2 + a)
    du12 = np.zeros_like(r12)
    du1 = np.zeros_like(r12)
    for (i, j, k, c_oppo, c_same) in terms:
        c = c_same if same else c_oppo
        if i:
            du12 += c * i * t12 ** (i - 1) * a12 / (r12 + a12) ** 2 * t1 ** j * t2 ** k
        if j:
            du1 += c * j * t1 ** (j - 1) * a / (r1 + a) ** 2 * t12 ** i * t2 ** k
    return du12[..., None] * d / r12[..., None] + (du1 / r1)[..., None] * p1.T[:, None, :]


@pytest.mark.parametrize("jastrow", ["r12_ratio2", "general"])
def test_tc_two_body_vector_term_brute_force(O, jastrow):
    """The two-body gradient term of the TC Fock build (HartreeFock.cpp:1168-1181,
    antisymmetrized at :1215-1222) restated without the per-point
    intermediates: the mean field of the operator grad_1 u(1, 2) . grad_1 over
    its four-index integrals
        M_X[a, b, c, d] = sum_{n1, n3} w1 w3 A_a(1) A_b(3) g_X(1, 3) . grad A_c(1) A_d(3)
    with g_X = grad_1 u from the Jastrow's closed form per point pair (not
    from V).  With (j, i) on electron 1: direct over both spins' densities
    (same-spin kernel for D_s, opposite-spin for D_s') minus same-spin
    exchange; with (j, i) on electron 2 likewise:
        vec_s(j,i) = sum_kl D_s(l,k) [M_S(j,k,i,l) - M_S(j,k,l,i) + M_S(k,j,l,i) - M_S(k,j,i,l)]
                   + sum_kl D_s'(l,k) [M_O(j,k,i,l) + M_O(k,j,l,i)]
    and G2 = vec/2 - vec^T/2 with F = 0.  (The terms with (j, i) on electron 2
    and the gradient on the density, V(j,i) . scaledGradients at :1172-1174,
    are symmetric in (j, i) and cancel in that antisymmetrization.)  Cases:
    the reference's r12 Jastrow
    (c_oppo = 2 c_same) and a general Ochi Jastrow with r1 / r12 r2 terms, whose
    V_same and V_oppo are independent.  Each planted mistake in the oracle's
    factorized build (a transposed operand, a wrong spin, a wrong sign, the
    wrong density) fails the comparison."""
    fns = [(1, 0, 0, 1.3), (2, 1, 0, 0.9), (2, 0, 0, 0.7), (2, 1, 1, 1.1)]
    pts1, w1 = O.product_grid(4, 7)
    pts3, w3 = O.product_grid(5, 5)
    A1, dA1 = O.sto_real(pts1, fns, grads=True)
    A3 = O.sto_real(pts3, fns)
    if jastrow == "r12_ratio2":
        a, a12, terms = 1.0, 1.0, [(1, 0, 0, 0.5, 0.25)]
    else:
        a, a12, terms = 0.8, 1.3, [(1, 0, 0, 0.5, 0.25), (0, 1, 0, -0.3, 0.2), (1, 0, 1, 0.15, -0.4),
                                   (2, 2, 0, 0.1, 0.35)]
    prm = O.ochi_params(a, a12, terms)
    Vs = O.compute_F(pts1, pts3, w3, A3, kernel=O.K_JASTROW_GRAD, prm=prm, same=True)
    Vo = O.compute_F(pts1, pts3, w3, A3, kernel=O.K_JASTROW_GRAD, prm=prm, same=False)
    nb = len(fns)
    nocc = (2, 1)
    rng = np.random.default_rng(11)
    C = np.zeros((2, nb, 2))
    D = np.zeros((2, nb, nb))
    for s in range(2):
        C[s, :, :nocc[s]] = rng.normal(size=(nb, nocc[s])) * 0.5
        D[s] = C[s, :, :nocc[s]] @ C[s, :, :nocc[s]].T

    M = {}
    for same in (True, False):
        g = _ochi_grad1_closed_form(terms, a, a12, same, pts1, pts3)
        M[same] = np.einsum("pa,qb,pqx,pcx,qd,p,q->abcd", A1, A3, g, dA1, A3, w1, w3, optimize=True)
    # the closed-form kernel reproduces the oracle's V (so g is the same function)
    assert rel_err(np.einsum("pjiq,->pjiq", Vs, 1.0),
                   np.einsum("qj,qi,pqx,q->pjix", A3, A3, _ochi_grad1_closed_form(terms, a, a12, True, pts1, pts3),
                             w3)) < 1e-13
    if jastrow == "general":  # V_oppo is not one multiple of V_same
        ratio = np.sum(Vo * Vs) / np.sum(Vs * Vs)
        assert rel_err(Vo, ratio * Vs) > 0.05
    MS, MO = M[True], M[False]
    G2b = np.zeros((2, nb, nb))
    for s in range(2):
        Ds, Db = D[s], D[1 - s]
        vec = (np.einsum("lk,jkil->ji", Ds, MS) - np.einsum("lk,jkli->ji", Ds, MS)
               + np.einsum("lk,kjli->ji", Ds, MS) - np.einsum("lk,kjil->ji", Ds, MS)
               + np.einsum("lk,jkil->ji", Db, MO) + np.einsum("lk,kjli->ji", Db, MO))
        G2b[s] = vec / 2 - vec.T / 2
    Z = np.zeros((A1.shape[0], nb, nb))
    G2, _ = O.tc_fock(*nocc, A1, dA1, w1, Z, Z, Vs, Vo, D, C)
    assert np.max(np.abs(G2b)) > 1e-3
    assert rel_err(G2, G2b) < 1e-12
    # the vector term alone is antisymmetric
    assert np.max(np.abs(G2 + G2.transpose(0, 2, 1))) < 1e-13 * np.max(np.abs(G2))
    # planted mistakes in the oracle, and swapped / sign-flipped inputs, fail
    try:
        for plant in (1, 2, 3, 4):
            O.set_plant(plant)
            G2x, _ = O.tc_fock(*nocc, A1, dA1, w1, Z, Z, Vs, Vo, D, C)
            assert rel_err(G2x, G2b) > 1e-3, plant
    finally:
        O.set_plant(0)
    G2x, _ = O.tc_fock(*nocc, A1, dA1, w1, Z, Z, -Vs, -Vo, D, C)
    assert rel_err(G2x, G2b) > 1e-3
    if jastrow == "general":
        G2x, _ = O.tc_fock(*nocc, A1, dA1, w1, Z, Z, Vo, Vs, D, C)
        assert rel_err(G2x, G2b) > 1e-3


def test_three_electron_integrals_are_the_unsymmetrized_Q(O):
    """threeElectronIntegrals (src/integration/integrateAll.hpp:369-442) is the
    one-term kernel T(p,q,r,s,t,u) = sum w A_p A_s V(q,t).V(r,u) whose three
    electron placements make computeQ's QSameSame (src/HartreeFock.cpp:613-682):
    QSS(j,p,r,i,q,s) = T(j,p,r,i,q,s) + T(p,j,r,q,i,s) + T(r,j,p,s,i,q)."""
    fns = [(1, 0, 0, 1.3), (2, 1, 1, 0.9), (2, 0, 0, 0.7)]
    c = _tc_small_case(O, fns)
    T = O.three_electron(c["A1"], c["w1"], c["Vs"])
    QSS, _, _ = O.compute_Q(c["A1"], c["w1"], c["Vs"], c["Vo"])
    Qt = T + np.einsum("pjrqis->jpriqs", T) + np.einsum("rjpsiq->jpriqs", T)
    assert np.max(np.abs(QSS)) > 1e-8
    assert rel_err(Qt, QSS) < 1e-13
    # T(p,q,r,s,t,u) = T(p,r,q,s,u,t) (the two V factors commute), and a
    # transposed V index breaks nothing for real orbitals: V(q,t) = V(t,q)
    assert rel_err(T.transpose(0, 2, 1, 3, 5, 4), T) < 1e-14
    assert rel_err(T.transpose(0, 4, 2, 3, 1, 5), T) < 1e-14


def test_gradient_eri_reduces_to_eri_of_V(O):
    """integrateAllR1R2<true> (src/integration/integrateAll.hpp:184-317) with
    'gradients' replaced by the values times a unit vector e_c is the plain ERI
    contraction of the c-th component of computeV; the result sits at
    (b2, b1, b4, b3) (src/integration/integrateAll.hpp:302)."""
    fns = [(1, 0, 0, 1.1), (2, 1, 1, 0.8), (2, 0, 0, 0.6)]
    rng = np.random.default_rng(2)
    pts1, w1 = O.product_grid(5, 7)
    pts2, w2 = O.product_grid(6, 5)
    pts1 = pts1 + rng.uniform(-0.2, 0.2, pts1.shape)  # no parity zeros
    pts2 = pts2 + rng.uniform(-0.2, 0.2, pts2.shape)
    A1 = O.sto_real(pts1, fns)
    A2 = O.sto_real(pts2, fns)
    prm = O.ochi_params(1.0, 1.0, [(1, 0, 0, 0.5, 0.25)])
    V = O.compute_F(pts1, pts2, w2, A2, kernel=O.K_JASTROW_GRAD, prm=prm, same=True)
    for c in range(3):
        dA2 = np.zeros(A2.shape + (3,))
        dA2[..., c] = A2
        G = O.eri_gradient(prm, True, pts1, A1, w1, pts2, w2, A2, dA2, scale=1.5)
        E = O.eri_from_F(A1, w1, np.ascontiguousarray(V[..., c]))
        assert np.max(np.abs(E)) > 1e-6
        assert rel_err(G.transpose(1, 0, 3, 2), 1.5 * E) < 1e-13


def test_atomic_qmc5d_hydrogenic_closed_forms(O):
    """atomicCoulombOperator / atomicExchangeOperator (src/AtomicBasis.hpp:59-140)
    as the oracle restates them, on the 5-D point set, for hydrogen-like
    functions of Z = 1, against the textbook Slater-Condon values:
    F0(1s,1s) = 5/8, J(1s,2s) = 17/81, K(1s,2s) = 16/729, and with
    F0(2p,2p) = 93/512, F2(2p,2p) = 45/512: J(2p0,2p0) = F0 + 4/25 F2,
    J(2p0,2p1) = F0 - 2/25 F2, K(2p0,2p1) = 3/25 F2 (its imaginary part
    integrates to 0).  The reference's own test prints J(2p0,2p0) = 0.196207 +-
    0.0018 and J(2p0,2p1) = 0.174825 +- 0.000862
    (src/unitTests/IntegrationTest.cpp:397-398, 444-446).  QMC error: the
    tolerances are in the assertions."""
    import tcscf_amd as tc
    fns = [(2, 1, 0, 0, 1.0), (2, 2, 0, 0, 1.0), (2, 2, 1, 0, 1.0), (2, 2, 1, 1, 1.0)]
    p1, p2, w = tc.qmc_spherical_points(100000, 5, 2, 0)
    E = O.eri_qmc5d_atomic(p1, p2, w, fns)
    F0, F2 = 93 / 512, 45 / 512
    checks = [((0, 0, 0, 0), 5 / 8, 1e-3), ((0, 1, 0, 1), 17 / 81, 5e-4), ((0, 1, 1, 0), 16 / 729, 2e-4),
              ((2, 2, 2, 2), F0 + 4 / 25 * F2, 1e-3), ((2, 3, 2, 3), F0 - 2 / 25 * F2, 5e-4),
              ((2, 3, 3, 2), 3 / 25 * F2, 3e-4)]
    for q, exact, tol in checks:
        assert abs(E[q].real - exact) < tol, (q, E[q], exact)
        assert abs(E[q].imag) < 1e-4, (q, E[q])
    assert abs(E[2, 2, 2, 2].real - 0.196207) < 0.0018 and abs(E[2, 3, 2, 3].real - 0.174825) < 0.000862
    # conj(g_a) g_c at r1: answer(c, b, a, d) = conj(answer(a, b, c, d)) on every point set
    assert np.allclose(E.transpose(2, 1, 0, 3), E.conj(), rtol=0, atol=1e-15)


def test_fill_two_electron_symmetries(O):
    """fillTwoElectronSymmetricHermitianArray (src/mathFunctions.hpp:367-433) as
    restated in the oracle and in the product (tcscf_amd.api._fill_two_electron):
    on tensors with the symmetries it assumes (8-fold for real functions;
    electron swap and Hermitian for complex ones) it reproduces the tensor
    exactly, so every image is written once; on a tensor without them, the
    canonical quadruples (a <= b, a <= c, d >= b or d >= a) keep their own
    values and the two restatements agree entry for entry."""
    from tcscf_amd.api import _fill_two_electron
    rng = np.random.default_rng(4)
    N = 5
    R = rng.normal(size=(N,) * 4)
    R = R + R.transpose(1, 0, 3, 2)
    R = R + R.transpose(2, 3, 0, 1)
    R = R + R.transpose(0, 3, 2, 1)
    R = R + R.transpose(2, 1, 0, 3)
    X = rng.normal(size=(N,) * 4) + 1j * rng.normal(size=(N,) * 4)
    C = X + X.transpose(1, 0, 3, 2)
    C = C + C.transpose(2, 3, 0, 1).conj()
    for fill in (O.fill_two_electron, _fill_two_electron):
        assert np.array_equal(fill(N, True, lambda a, b, c, d: R[a, b, c, d]), R.astype(complex))
        assert np.allclose(fill(N, False, lambda a, b, c, d: C[a, b, c, d]), C, rtol=0, atol=1e-13)
    for real in (True, False):
        A = O.fill_two_electron(N, real, lambda a, b, c, d: X[a, b, c, d])
        assert np.array_equal(A, _fill_two_electron(N, real, lambda a, b, c, d: X[a, b, c, d]))
        for a in range(N):
            for b in range(a + 1, N):
                for c in range(a + 1, N):
                    for d in range(b + 1 if real else a + 1, N):
                        assert A[a, b, c, d] == X[a, b, c, d]


def test_qmc_spherical_points_and_eri_qmc6d(O):
    """The per-element QMC route (src/AtomicBasis.hpp, src/integration/
    qmcWrapper.hpp:146-276): the point maps integrate e^{-2r} to pi (3-D) and
    e^{-2 r1 - 2 r2} to pi^2 (5-D, 6-D); the oracle's 6-D r12 matrix elements
    of 1s STOs approach the closed form (ac|bd) (QMC error, tolerance in the
    assertion); the 3-D overlap approaches the analytic one."""
    import tcscf_amd as tc
    p1, _, w = tc.qmc_spherical_points(20000, 3, 2, 7)
    assert abs((w * np.exp(-2 * np.sqrt((p1 ** 2).sum(0)))).sum() - math.pi) < 1e-6
    for nd in (5, 6):
        p1, p2, w = tc.qmc_spherical_points(20000, nd, 2, 7)
        r1, r2 = np.sqrt((p1 ** 2).sum(0)), np.sqrt((p2 ** 2).sum(0))
        assert abs((w * np.exp(-2 * r1 - 2 * r2)).sum() - math.pi ** 2) < 1e-3
        if nd == 5:
            assert np.all(p2[1] == 0)  # phi2 = 0
    z = [0.8, 1.7]
    fns = [(1, 0, 0, zz) for zz in z]
    p1, p2, w = tc.qmc_spherical_points(20000, 6, 2, 3)
    E = O.eri_qmc6d(p1, p2, w, O.sto_real(p1, fns), O.sto_real(p2, fns))
    ex = np.array([[[[_eri_1s(z[a], z[c], z[b], z[d]) for d in range(2)] for c in range(2)] for b in range(2)]
                   for a in range(2)])
    assert np.max(np.abs(E - ex) / np.abs(ex)) < 3e-3
    p1, _, w = tc.qmc_spherical_points(20000, 3, 2, 5)
    A = O.sto_real(p1, fns)
    S = (A * w[:, None]).T @ A
    Sx, _ = O.sto_S_H([(1, 0, 0, zz) for zz in z], 1)
    assert np.max(np.abs(S - Sx)) < 1e-5
