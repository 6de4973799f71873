"""ctypes front end of the FP64 CPU oracle (oracle/tcscf_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as the
checker and by bench.py for its CPU-baseline leg.  The product (tcscf_amd)
never imports this module.

All arrays are numpy float64, C-contiguous.  Grid points are SoA (3, N).
"""
import ctypes
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libtcscf_oracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)


class ScfOpts(ctypes.Structure):
    _fields_ = [("tol", ctypes.c_double), ("max_iter", ctypes.c_int), ("mask", ctypes.c_int),
                ("mgs_mode", ctypes.c_int)]


class ScfOut(ctypes.Structure):
    _fields_ = [("energy", ctypes.c_double), ("iterations", ctypes.c_int), ("status", ctypes.c_int)]


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
    return _lib


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous, (a.dtype, a.flags)
    return a.ctypes.data_as(_dp)


def set_threads(n):
    lib().orc_set_threads(int(n))


def get_threads():
    return lib().orc_get_threads()


def set_plant(p):
    """Plant a mistake in the TC Fock build (tests only; 0 removes it):
    1 transposed operand in the Vja term, 2 the same spin's density against
    V_oppo in VaaSamePlusOppo, 3 wrong sign of the antisymmetrization, 4 Vai
    from the other spin's density."""
    lib().orc_set_plant(int(p))


def chebyshev_gauss(n):
    x = np.zeros(n)
    w = np.zeros(n)
    lib().orc_chebyshev_gauss(n, _p(x), _p(w))
    return x, w


RADIAL_M4, RADIAL_BECKE, RADIAL_XOVER1MX = 0, 1, 2


def radial_grid(n, kind=RADIAL_BECKE, a=1.0, eps=1.0):
    r = np.zeros(n)
    w = np.zeros(n)
    lib().orc_radial_grid(n, kind, ctypes.c_double(a), ctypes.c_double(eps), _p(r), _p(w))
    return r, w


def lebedev(order):
    n = lib().orc_lebedev_npts(order)
    if n < 0:
        raise ValueError(f"Lebedev order {order} not available")
    t, p, w = np.zeros(n), np.zeros(n), np.zeros(n)
    lib().orc_lebedev(order, _p(t), _p(p), _p(w))
    return t, p, w


def product_grid(nrad, order, kind=RADIAL_BECKE, a=1.0, eps=1.0):
    n = nrad * lib().orc_lebedev_npts(order)
    pts = np.zeros((3, n))
    w = np.zeros(n)
    rc = lib().orc_product_grid(nrad, kind, ctypes.c_double(a), ctypes.c_double(eps), order, _p(pts), _p(w))
    if rc < 0:
        raise ValueError("bad Lebedev order")
    return pts, w


def _fns(fns):
    return np.ascontiguousarray(np.asarray(fns, dtype=np.float64).reshape(-1, 4))


def sto_complex(pts, fns, grads=False):
    f = _fns(fns)
    n = pts.shape[1]
    v = np.zeros((n, len(f), 2))
    g = np.zeros((n, len(f), 3, 2)) if grads else None
    lib().orc_sto_complex(_p(np.ascontiguousarray(pts)), ctypes.c_int64(n), _p(f), len(f), _p(v), _p(g))
    vc = v[..., 0] + 1j * v[..., 1]
    return (vc, g[..., 0] + 1j * g[..., 1]) if grads else vc


def sto_real(pts, fns, grads=False):
    f = _fns(fns)
    n = pts.shape[1]
    v = np.zeros((n, len(f)))
    g = np.zeros((n, len(f), 3)) if grads else None
    lib().orc_sto_real(_p(np.ascontiguousarray(pts)), ctypes.c_int64(n), _p(f), len(f), _p(v), _p(g))
    return (v, g) if grads else v


def _laguerre_coeffs(q, k):
    """L_q^(k)(x) = sum_j c_j x^j with c_j = (-1)^j C(q+k, q-j) / j! (the explicit
    series; the device code uses the three-term recurrence)."""
    return [(-1) ** j * math.comb(q + k, q - j) / math.factorial(j) for j in range(q + 1)]


def atomic_as_stos(fn):
    """An Ochi (src/OchiBasis.hpp:29-59) or hydrogen-like
    (src/HydrogenLikeBasis.hpp:24-50) function as an exact linear combination of
    normalized STOs of one exponent: N r^l L_q^(k)(2 a r) e^(-a r) Y_lm =
    sum_j N c_j (2a)^j / N_sto(l+1+j, a) STO(l+1+j, a).  fn: (kind, n, l, m, alpha);
    kind 1 Ochi (q = n, k = 2l+2, a = alpha), 2 hydrogen-like (q = n-l-1,
    k = 2l+1, a = Z/n, alpha = Z).  Returns [(weight, (n', l, m, a)), ...]."""
    kind, n, l, m, alpha = fn
    if kind == 1:
        q, k, a = n, 2 * l + 2, alpha
        N = (2 * a) ** (l + 1.5) / math.sqrt(math.factorial(n + 2 * l + 2) / math.factorial(n))
    else:
        q, k, a = n - l - 1, 2 * l + 1, alpha / n
        N = -(2 * a) ** (l + 1.5) / math.sqrt(2 * n * math.factorial(n + l) / math.factorial(n - l - 1))
    out = []
    for j, c in enumerate(_laguerre_coeffs(q, k)):
        ns = l + 1 + j
        nsto = (2 * a) ** ns * math.sqrt(2 * a / math.factorial(2 * ns))
        out.append((N * c * (2 * a) ** j / nsto, (ns, l, m, a)))
    return out


def eri_qmc6d(pts1, pts2, w, A1, A2):
    """fillR12Array with r12MatrixElement (src/AtomicBasis.hpp:30-48, 167-177)
    on given point pairs (r1_k, r2_k) with weights w_k (the 6-D QMC route of
    src/integration/qmcWrapper.hpp:249-276), real basis:
    answer[a, b, c, d] = sum_k w_k A1[k, a] A2[k, b] A1[k, c] A2[k, d] / r12_k,
    coincident pairs excluded."""
    r12 = np.sqrt(((np.asarray(pts1) - np.asarray(pts2)) ** 2).sum(0))
    inv = np.where(r12 > 0, 1.0 / np.where(r12 > 0, r12, 1.0), 0.0)
    return np.einsum("k,ka,kb,kc,kd->abcd", np.asarray(w) * inv, A1, A2, A1, A2, optimize=True)


def atomic_real(pts, fns, grads=False):
    """Real-harmonic values (and gradients) of (kind, n, l, m, alpha) functions:
    kind 0 STO (sto_real), 1 Ochi, 2 hydrogen-like (via atomic_as_stos)."""
    terms, owner = [], []
    for b, f in enumerate(fns):
        if f[0] == 0:
            terms.append((1.0, (f[1], f[2], f[3], f[4])))
            owner.append(b)
        else:
            for w, sto in atomic_as_stos(f):
                terms.append((w, sto))
                owner.append(b)
    r = sto_real(pts, [t[1] for t in terms], grads)
    v, g = (r if grads else (r, None))
    n = pts.shape[1]
    vo = np.zeros((n, len(fns)))
    go = np.zeros((n, len(fns), 3)) if grads else None
    for t, (w, _) in enumerate(terms):
        vo[:, owner[t]] += w * v[:, t]
        if grads:
            go[:, owner[t]] += w * g[:, t]
    return (vo, go) if grads else vo


def atomic_complex(pts, fns):
    """The reference's complex functions R(r) sph_legendre(l, |m|, theta) e^{i m phi}
    (src/SlaterTypeOrbital.hpp, src/OchiBasis.hpp, src/HydrogenLikeBasis.hpp over
    src/mathFunctions.hpp:126-139) of (kind, n, l, m, alpha) functions: (n, nb)."""
    terms, owner = [], []
    for b, f in enumerate(fns):
        for w, sto in ([(1.0, (f[1], f[2], f[3], f[4]))] if f[0] == 0 else atomic_as_stos(f)):
            terms.append((w, sto))
            owner.append(b)
    v = sto_complex(pts, [t[1] for t in terms])
    vo = np.zeros((pts.shape[1], len(fns)), complex)
    for t, (w, _) in enumerate(terms):
        vo[:, owner[t]] += w * v[:, t]
    return vo


def eri_qmc5d_atomic(pts1, pts2, w, fns):
    """atomicCoulombOperator / atomicExchangeOperator (src/AtomicBasis.hpp:59-140)
    for every quadruple on the 5-D point set (r2 on the phi = 0 half-plane, the
    phi2 integral in the weights), the reference's complex functions g:
    answer[a, b, c, d] = sum_k w_k conj(g_a(r1_k) g_b(r2_k)) g_c(r1_k) g_d(r2_k) / r12_k.
    At phi2 = 0 the r2 factors are the reference's sph_legendre magnitudes, so
    the Coulomb entries (a, c and b, d of equal l, m) are its Coulomb integrand
    and the exchange entries (a, d and b, c equal) its exchange integrand with
    the phase e^{i (m_c - m_a) phi1}."""
    g1, g2 = atomic_complex(pts1, fns), atomic_complex(pts2, fns)
    r12 = np.sqrt(((np.asarray(pts1) - np.asarray(pts2)) ** 2).sum(0))
    inv = np.where(r12 > 0, 1.0 / np.where(r12 > 0, r12, 1.0), 0.0)
    return np.einsum("k,ka,kb,kc,kd->abcd", np.asarray(w) * inv, g1.conj(), g2.conj(), g1, g2, optimize=True)


def fill_two_electron(N, real, value):
    """fillTwoElectronSymmetricHermitianArray (src/mathFunctions.hpp:367-433):
    value(a, b, c, d) on the canonical quadruples only (a <= b, a <= c, and
    d >= b for real functions, else d >= a), each copied to its images under the
    electron swap, Hermitian conjugation and, for real functions, the b <-> d
    swap; the array accumulates (+=).  Pure Python, for small N."""
    A = np.zeros((N,) * 4, complex)
    for a in range(N):
        for b in range(a, N):
            for c in range(a, N):
                for d in range(b if real else a, N):
                    v = complex(value(a, b, c, d))
                    A[a, b, c, d] += v
                    sym, herm, rl = a != b, a != c and a != d, real and d != b
                    if sym:
                        A[b, a, d, c] += v
                    if herm or (real and a != c):
                        A[c, d, a, b] += v.conjugate()
                    if herm and sym:
                        A[d, c, b, a] += v.conjugate()
                    if rl:
                        A[a, d, c, b] += v
                    if rl and sym:
                        A[d, a, b, c] += v
                    if rl and herm:
                        A[c, b, a, d] += v
                    if rl and herm and sym:
                        A[b, c, d, a] += v
    return A


def sto_S_H(fns, Z):
    f = _fns(fns)
    nb = len(f)
    S = np.zeros((nb, nb))
    H = np.zeros((nb, nb))
    lib().orc_sto_S_H(_p(f), nb, int(Z), _p(S), _p(H))
    return S, H


def ochi_params(a, a12, terms):
    """terms: list of (i, j, k, c_oppo, c_same)."""
    out = [a, a12, len(terms)]
    for t in terms:
        out += list(map(float, t))
    return np.array(out, dtype=np.float64)


def ochi_value(prm, r1, r2, r12, same):
    f = lib().orc_ochi_value
    f.restype = ctypes.c_double
    return f(_p(prm), ctypes.c_double(r1), ctypes.c_double(r2), ctypes.c_double(r12), int(same))


def ochi_gradient(prm, r1, r2, same):
    g = np.zeros(3)
    lib().orc_ochi_gradient(_p(prm), _p(np.asarray(r1, dtype=np.float64)), _p(np.asarray(r2, dtype=np.float64)),
                            int(same), _p(g))
    return g


K_COULOMB, K_TC_SCALAR, K_JASTROW_GRAD = 0, 1, 2


def compute_F(pts1, pts2, w2, A2, kernel=K_COULOMB, prm=None, same=True, scale=1.0, rows=None):
    """Full F: (n1 rows, nb, nb) or (n1 rows, nb, nb, 3) for the vector kernel."""
    n1 = pts1.shape[1]
    n2 = pts2.shape[1]
    nb = A2.shape[1]
    r0, r1 = (0, n1) if rows is None else rows
    ncomp = 3 if kernel == K_JASTROW_GRAD else 1
    F = np.zeros((r1 - r0, nb, nb, ncomp))
    if prm is None:
        prm = ochi_params(1.0, 1.0, [])
    lib().orc_compute_F(kernel, _p(prm), int(same), _p(np.ascontiguousarray(pts1)), ctypes.c_int64(n1),
                        ctypes.c_int64(r0), ctypes.c_int64(r1), _p(np.ascontiguousarray(pts2)), _p(w2),
                        ctypes.c_int64(n2), _p(np.ascontiguousarray(A2)), nb, ctypes.c_double(scale), _p(F))
    return F if ncomp == 3 else F[..., 0]


def compute_F_complex(pts1, pts2, w2, A2c):
    n1, n2, nb = pts1.shape[1], pts2.shape[1], A2c.shape[1]
    a = np.ascontiguousarray(np.stack([A2c.real, A2c.imag], axis=-1))
    F = np.zeros((n1, nb, nb, 2))
    lib().orc_compute_F_complex(_p(np.ascontiguousarray(pts1)), ctypes.c_int64(n1), _p(np.ascontiguousarray(pts2)),
                                _p(w2), ctypes.c_int64(n2), _p(a), nb, _p(F))
    return F[..., 0] + 1j * F[..., 1]


def eri_from_F(A1, w1, F):
    nb = A1.shape[1]
    out = np.zeros((nb, nb, nb, nb))
    lib().orc_eri_from_F(_p(np.ascontiguousarray(A1)), _p(w1), ctypes.c_int64(A1.shape[0]),
                         _p(np.ascontiguousarray(F)), nb, _p(out))
    return out


def eri_gradient(prm, same, pts1, A1, w1, pts2, w2, A2, dA2, scale=1.0):
    """integrateAllR1R2<true>: nb^4 in the reference's (b2, b1, b4, b3) order."""
    nb = A1.shape[1]
    out = np.zeros((nb,) * 4)
    c = np.ascontiguousarray
    lib().orc_eri_gradient(_p(prm), int(same), _p(c(pts1)), _p(c(A1)), _p(w1), ctypes.c_int64(A1.shape[0]),
                           _p(c(pts2)), _p(w2), ctypes.c_int64(A2.shape[0]), _p(c(A2)), _p(c(dA2)), nb,
                           ctypes.c_double(scale), _p(out))
    return out


def three_electron(A1, w1, V):
    nb = A1.shape[1]
    Q = np.zeros((nb,) * 6)
    lib().orc_three_electron(_p(np.ascontiguousarray(A1)), _p(w1), ctypes.c_int64(A1.shape[0]),
                             _p(np.ascontiguousarray(V)), nb, _p(Q))
    return Q


def gen_sym_eig(F, S):
    n = F.shape[0]
    e = np.zeros(n)
    C = np.zeros((n, n))
    rc = lib().orc_gen_sym_eig(n, _p(np.ascontiguousarray(F)), _p(np.ascontiguousarray(S)), _p(e), _p(C))
    if rc:
        raise RuntimeError("solver failure")
    return e, C


def gen_nonsym_eig(F, S):
    n = F.shape[0]
    e = np.zeros((n, 2))
    X = np.zeros((n, n, 2))
    rc = lib().orc_gen_nonsym_eig(n, _p(np.ascontiguousarray(F)), _p(np.ascontiguousarray(S)), _p(e), _p(X))
    if rc:
        raise RuntimeError("solver failure")
    return e[:, 0] + 1j * e[:, 1], X[..., 0] + 1j * X[..., 1]


def mgs(vecs, S, mode=0):
    v = np.ascontiguousarray(np.stack([vecs.real, vecs.imag], axis=-1).astype(np.float64))
    k, n = vecs.shape
    lib().orc_mgs(k, n, _p(v), _p(np.ascontiguousarray(S)), int(mode))
    return v[..., 0] + 1j * v[..., 1]


def _opts(tol, max_iter, mask, mgs_mode=0):
    return ScfOpts(tol, max_iter, int(mask), int(mgs_mode))


def rcs_scf(nelec, S, H, A1, w1, F, tol=1e-5, max_iter=50, mask=True):
    nb = S.shape[0]
    out = ScfOut()
    trace = np.zeros(max_iter)
    D = np.zeros((nb, nb))
    ev = np.zeros(nb)
    o = _opts(tol, max_iter, mask)
    lib().orc_rcs_scf(int(nelec), nb, _p(S), _p(H), _p(np.ascontiguousarray(A1)), _p(w1),
                      ctypes.c_int64(A1.shape[0]), _p(np.ascontiguousarray(F)), ctypes.byref(o), ctypes.byref(out),
                      _p(trace), _p(D), _p(ev))
    return dict(energy=out.energy, iterations=out.iterations, status=out.status,
                trace=trace[:max(out.iterations, 0)], D=D, evals=ev)


def uos_scf(nup, ndown, S, H, A1, w1, F, tol=1e-5, max_iter=100, mask=True):
    nb = S.shape[0]
    out = ScfOut()
    trace = np.zeros(max_iter)
    D = np.zeros((2, nb, nb))
    ev = np.zeros((2, nb))
    o = _opts(tol, max_iter, mask)
    lib().orc_uos_scf(int(nup), int(ndown), nb, _p(S), _p(H), _p(np.ascontiguousarray(A1)), _p(w1),
                      ctypes.c_int64(A1.shape[0]), _p(np.ascontiguousarray(F)), ctypes.byref(o), ctypes.byref(out),
                      _p(trace), _p(D), _p(ev))
    return dict(energy=out.energy, iterations=out.iterations, status=out.status,
                trace=trace[:max(out.iterations, 0)], D=D, evals=ev)


def tc_scf(nup, ndown, S, H, A1, dA1, w1, Fs, Fo, Vs, Vo, tol=1e-5, max_iter=200, mask=True, mgs_mode=0):
    nb = S.shape[0]
    out = ScfOut()
    trace = np.zeros(max_iter)
    D = np.zeros((2, nb, nb))
    ps = np.zeros(nup + ndown)
    o = _opts(tol, max_iter, mask, mgs_mode)
    c = np.ascontiguousarray
    lib().orc_tc_scf(int(nup), int(ndown), nb, _p(S), _p(H), _p(c(A1)), _p(c(dA1)), _p(w1),
                     ctypes.c_int64(A1.shape[0]), _p(c(Fs)), _p(c(Fo)), _p(c(Vs)), _p(c(Vo)), ctypes.byref(o),
                     ctypes.byref(out), _p(trace), _p(D), _p(ps))
    return dict(energy=out.energy, iterations=out.iterations, status=out.status,
                trace=trace[:max(out.iterations, 0)], D=D, pseudo=ps)


def tc_fock(nup, ndown, A1, dA1, w1, Fs, Fo, Vs, Vo, D, C):
    """One TC Fock build at density D (2, nb, nb) and occupied orbitals C
    (2, nb, cmax): returns (G2, G3), each (2, nb, nb), before the mask."""
    nb = A1.shape[1]
    c = np.ascontiguousarray
    C = c(C, dtype=np.float64)
    G2 = np.zeros((2, nb, nb))
    G3 = np.zeros((2, nb, nb))
    rc = lib().orc_tc_fock(int(nup), int(ndown), nb, _p(c(A1)), _p(c(dA1)), _p(w1), ctypes.c_int64(A1.shape[0]),
                           _p(c(Fs)), _p(c(Fo)), _p(c(Vs)), _p(c(Vo)), _p(c(D, dtype=np.float64)), _p(C),
                           int(C.shape[2]), _p(G2), _p(G3))
    if rc:
        raise ValueError("tc_fock: C has fewer columns than occupied orbitals")
    return G2, G3


def compute_Q(A1, w1, Vs, Vo, labelled=False):
    """computeQ: the three (nb,)*6 tensors QSameSame, QSameOppo, QOppoOppo;
    labelled=True takes V_same / V_oppo per electron pair from the spins."""
    nb = A1.shape[1]
    Q = [np.zeros((nb,) * 6) for _ in range(3)]
    c = np.ascontiguousarray
    lib().orc_compute_Q(_p(c(A1)), _p(w1), ctypes.c_int64(A1.shape[0]), _p(c(Vs)), _p(c(Vo)), nb,
                        int(bool(labelled)), _p(Q[0]), _p(Q[1]), _p(Q[2]))
    return Q


def assemble_G3(QSS, QSO, QOO, D):
    """assembleG: G3 (2, nb, nb) from the explicit Q tensors."""
    nb = QSS.shape[0]
    G3 = np.zeros((2, nb, nb))
    lib().orc_assemble_G3(nb, _p(QSS), _p(QSO), _p(QOO), _p(np.ascontiguousarray(D, dtype=np.float64)), _p(G3))
    return G3
