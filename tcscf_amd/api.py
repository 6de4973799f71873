"""Host-side mirror of the reference's call surface for the grid-integral path.

Names, argument meaning and error behaviour follow corbett5/tcscf:
  integration::QMCGrid              src/integration/integrateAll.hpp:12-142
  SlaterTypeOrbital, overlap, ...   src/SlaterTypeOrbital.hpp:15-171
  computeF / computeV               src/HartreeFock.hpp:88-229
  RCS/UOS/TCHartreeFock::compute    src/HartreeFock.hpp:233-418, src/HartreeFock.cpp
Device memory is held in torch tensors (plumbing only); every computation runs
in libtcscf_hip.so through the C ABI.
"""
import ctypes
import math

import numpy as np

from . import _lib
from ._lib import check, lib

_dp = ctypes.POINTER(ctypes.c_double)


def _torch():
    import torch
    return torch


def _dev():
    torch = _torch()
    if not torch.cuda.is_available():
        raise RuntimeError("tcscf_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


def _stream():
    torch = _torch()
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _np(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


def npair(nb):
    return nb * (nb + 1) // 2


def pair_index(j, i):
    """Packed index of (j, i); symmetric in its arguments."""
    if j > i:
        j, i = i, j
    return i * (i + 1) // 2 + j


def unpack(Fp, nb):
    """Packed (..., npair) -> full (..., nb, nb) (numpy or torch)."""
    idx = np.array([[pair_index(j, i) for i in range(nb)] for j in range(nb)])
    if isinstance(Fp, np.ndarray):
        return Fp[..., idx]
    torch = _torch()
    return Fp[..., torch.as_tensor(idx, device=Fp.device)]


# --------------------------------------------------------------------------- basis
class SlaterTypeOrbital:
    """N r^(n-1) e^(-alpha r) S_lm with real solid harmonics S_lm (l <= 3)."""
    kind = _lib.BASIS_STO

    def __init__(self, alpha, n, l, m):
        if n < 1 or n < l or l < abs(m) or l > 3:
            raise ValueError(f"invalid STO (n={n}, l={l}, m={m})")
        self.alpha, self.n, self.l, self.m = float(alpha), int(n), int(l), int(m)

    @property
    def normalization(self):
        a, n = self.alpha, self.n
        return (2 * a) ** n * math.sqrt(2 * a / math.factorial(2 * n))

    def as_tuple(self):
        return (self.n, self.l, self.m, self.alpha)

    def __repr__(self):
        return f"STO(n={self.n}, l={self.l}, m={self.m}, zeta={self.alpha:g})"


class OchiBasisFunction:
    """OchiBasisFunction (src/OchiBasis.hpp:15-110): N r^l L_n^(2l+2)(2 alpha r)
    e^(-alpha r) Y_lm, n >= 0, N = (2 alpha)^(l+3/2) / sqrt((n+2l+2)! / n!);
    orthonormal for one alpha.  Real harmonics here (l <= 3)."""
    kind = _lib.BASIS_OCHI

    def __init__(self, alpha, n, l, m):
        if n < 0 or l < abs(m) or l > 3 or alpha <= 0:
            raise ValueError(f"invalid Ochi function (n={n}, l={l}, m={m}, alpha={alpha})")
        self.alpha, self.n, self.l, self.m = float(alpha), int(n), int(l), int(m)

    @property
    def normalization(self):
        n, l = self.n, self.l
        return (2 * self.alpha) ** (l + 1.5) / math.sqrt(math.factorial(n + 2 * l + 2) / math.factorial(n))

    def __repr__(self):
        return f"Ochi(n={self.n}, l={self.l}, m={self.m}, alpha={self.alpha:g})"


class HydrogenLikeBasisFunction:
    """HydrogenLikeBasisFunction (src/HydrogenLikeBasis.hpp:12-61): the bound
    eigenfunction (n, l, m) of charge Z, N r^l L_(n-l-1)^(2l+1)(2 b r) e^(-b r)
    Y_lm with b = Z / n and the reference's (negative) normalization."""
    kind = _lib.BASIS_HYDROGEN

    def __init__(self, Z, n, l, m):
        if Z <= 0 or n <= l or l < abs(m) or l > 3:
            raise ValueError(f"invalid hydrogen-like function (Z={Z}, n={n}, l={l}, m={m})")
        self.Z, self.n, self.l, self.m = int(Z), int(n), int(l), int(m)
        self.alpha = float(Z)

    @property
    def beta(self):
        return self.Z / self.n

    def __repr__(self):
        return f"HydrogenLike(Z={self.Z}, n={self.n}, l={self.l}, m={self.m})"


def _fn_array(basis):
    arr = (_lib.BasisFn * len(basis))()
    for k, b in enumerate(basis):
        arr[k].kind, arr[k].n, arr[k].l, arr[k].m, arr[k].alpha = b.kind, b.n, b.l, b.m, b.alpha
    return arr


def _sto_array(basis):
    arr = (_lib.Sto * len(basis))()
    for k, b in enumerate(basis):
        arr[k].n, arr[k].l, arr[k].m, arr[k].zeta = b.n, b.l, b.m, b.alpha
    return arr


def loadSTO(directory, basisSet, element, maxL=2 ** 31 - 1):
    """loadSTO (src/SlaterTypeOrbital.cpp:24-76): the STOs of an ADF basis file
    <directory>/<basisSet>/<element> (TZ3P, QZ6P, ATZ3P or AQZ6P), shells
    expanded to m = -l..l, l > maxL skipped."""
    args = (str(directory).encode(), str(basisSet).encode(), str(element).encode(), int(min(maxL, 2 ** 31 - 1)))
    n = check(lib().tcscf_load_sto(*args, None, 0))
    arr = (_lib.Sto * max(n, 1))()
    check(lib().tcscf_load_sto(*args, arr, n))
    return [SlaterTypeOrbital(arr[k].zeta, arr[k].n, arr[k].l, arr[k].m) for k in range(n)]


def even_tempered(shells, zeta0, ratio):
    """Even-tempered STOs: shells = {l: count}; zeta_k = zeta0 * ratio**k per l,
    n = l + 1, m = -l..l (all m of one exponent adjacent)."""
    basis = []
    for l in sorted(shells):
        for k in range(shells[l]):
            for m in range(-l, l + 1):
                basis.append(SlaterTypeOrbital(zeta0 * ratio ** k, l + 1, l, m))
    return basis


def one_electron_matrices(basis, Z):
    """Overlap S and core Hamiltonian H of one kind of basis: STOs
    (src/SlaterTypeOrbital.hpp:125-171), Ochi functions (src/OchiBasis.hpp:117-155)
    or hydrogen-like functions (src/HydrogenLikeBasis.hpp:67-80)."""
    nb = len(basis)
    S = np.zeros((nb, nb))
    H = np.zeros((nb, nb))
    check(lib().tcscf_basis_one_electron(_fn_array(basis), nb, int(Z), S.ctypes.data_as(_dp), H.ctypes.data_as(_dp)))
    return S, H


def overlap(b1, b2):
    return one_electron_matrices([b1, b2], 0)[0][0, 1]


def coreMatrixElement(Z, b1, b2):
    return one_electron_matrices([b1, b2], Z)[1][0, 1]


# --------------------------------------------------------------------------- grids
def radial_grid(n, radial_map=_lib.RADIAL_BECKE, a=1.0, eps=1.0):
    r = np.zeros(n)
    w = np.zeros(n)
    check(lib().tcscf_radial_grid(n, radial_map, ctypes.c_double(a), ctypes.c_double(eps), r.ctypes.data_as(_dp),
                                  w.ctypes.data_as(_dp)))
    return r, w


def lebedev(order):
    n = check(lib().tcscf_lebedev_npts(order))
    t, p, w = np.zeros(n), np.zeros(n), np.zeros(n)
    check(lib().tcscf_lebedev(order, t.ctypes.data_as(_dp), p.ctypes.data_as(_dp), w.ctypes.data_as(_dp)))
    return t, p, w


def product_grid(nrad, order, radial_map=_lib.RADIAL_BECKE, a=1.0, eps=1.0):
    """Radial x Lebedev points (3, N) SoA and weights (N,) incl. r^2, on the host."""
    n = nrad * check(lib().tcscf_lebedev_npts(order))
    pts = np.zeros((3, n))
    w = np.zeros(n)
    check(lib().tcscf_product_grid(nrad, radial_map, ctypes.c_double(a), ctypes.c_double(eps), order,
                                   pts.ctypes.data_as(_dp), w.ctypes.data_as(_dp)))
    return pts, w


def qmc_lattice_grid(gridSize, ndim=3, nshifts=2, seed=0):
    """The QMCGrid(gridSize) point set (src/integration/integrateAll.hpp:21-65)
    on the host: points (3, N) and weights (N,), N = nshifts * (first prime >=
    gridSize)."""
    n = check(lib().tcscf_qmc_lattice_npts(int(gridSize), int(nshifts)))
    pts = np.zeros((3, n))
    w = np.zeros(n)
    check(lib().tcscf_qmc_lattice_grid(int(gridSize), int(ndim), int(nshifts), int(seed) & (2 ** 64 - 1),
                                       pts.ctypes.data, w.ctypes.data))
    return pts, w


def qmc_spherical_points(min_points, ndim=6, nshifts=2, seed=0):
    """The point sets of sphericalCoordinates{3,5,6}DIntegral
    (src/integration/qmcWrapper.hpp:146-276) on a shifted Korobov lattice:
    (pts1 (3, N), pts2 (3, N) or None for ndim 3, weights (N,)), weights with
    the Jacobians, the prefactor and 1/N."""
    n = check(lib().tcscf_qmc_lattice_npts(int(min_points), int(nshifts)))
    p1 = np.zeros((3, n))
    p2 = np.zeros((3, n)) if ndim != 3 else None
    w = np.zeros(n)
    check(lib().tcscf_qmc_spherical_points(int(min_points), int(ndim), int(nshifts), int(seed) & (2 ** 64 - 1),
                                           p1.ctypes.data, p2.ctypes.data if p2 is not None else None, w.ctypes.data))
    return p1, p2, w


def sobol_points(n, dim=3, scramble=True, seed=0):
    """Sobol points (n, dim) in [0, 1), Gray-code order."""
    u = np.zeros((int(n), int(dim)))
    check(lib().tcscf_sobol_points(int(n), int(dim), int(bool(scramble)), int(seed) & (2 ** 64 - 1), u.ctypes.data))
    return u


def sobol_grid(n, seed=0):
    """Importance-sampled grid from n scrambled Sobol points (r density
    2 e^{-2r}, uniform directions): points (3, n), weights (n,)."""
    pts = np.zeros((3, int(n)))
    w = np.zeros(int(n))
    check(lib().tcscf_sobol_grid(int(n), int(seed) & (2 ** 64 - 1), pts.ctypes.data, w.ctypes.data))
    return pts, w


class QuadratureGrid:
    """points (3, N) and weights (N,), device-resident FP64."""

    def __init__(self, points, weights):
        self.points = points
        self.weights = weights


class QMCGrid:
    """Mirror of integration::QMCGrid<double, NDIM>: grid points + weights (with
    every Jacobian) and the basis tabulated on them, all in HBM.

    ndim == 2 reproduces the reference's r2 grid trick (points in the xz-plane,
    computeF multiplies by 2 pi, src/HartreeFock.hpp:123-124)."""

    def __init__(self, points, weights, ndim=3):
        torch = _torch()
        pts = np.ascontiguousarray(points, dtype=np.float64)
        if pts.ndim != 2 or pts.shape[0] != 3:
            raise ValueError("points must be (3, N)")
        if ndim == 2 and np.any(pts[1] != 0):
            raise ValueError("a 2-D grid keeps y == 0")
        w = np.ascontiguousarray(weights, dtype=np.float64)
        if w.shape != (pts.shape[1],):
            raise ValueError("weights must be (N,)")
        dev = _dev()
        self.ndim = ndim
        self.quadratureGrid = QuadratureGrid(torch.from_numpy(pts).to(dev), torch.from_numpy(w).to(dev))
        # summation order for contractions over this grid: shell by shell
        # (by radius), so each 32-point FP32 step of the mixed-precision
        # kernels sees pair densities of similar magnitude
        order = np.argsort(np.sqrt((pts * pts).sum(0)), kind="stable").astype(np.int32)
        self.radial_order = torch.from_numpy(order).to(dev)
        self.points_host = pts
        self.weights_host = w
        self.basisValues = None
        self.basisGradients = None
        self.basis = None

    @classmethod
    def radial_lebedev(cls, nrad, order, radial_map=_lib.RADIAL_BECKE, a=1.0, eps=1.0):
        pts, w = product_grid(nrad, order, radial_map, a, eps)
        return cls(pts, w, 3)

    @classmethod
    def lattice(cls, gridSize, ndim=3, nshifts=2, seed=0):
        """QMCGrid(gridSize) (src/integration/integrateAll.hpp:21-65)."""
        pts, w = qmc_lattice_grid(gridSize, ndim, nshifts, seed)
        return cls(pts, w, ndim)

    @classmethod
    def sobol(cls, n, seed=0):
        """Scrambled-Sobol grid, exponential radial importance sampling."""
        pts, w = sobol_grid(n, seed)
        return cls(pts, w, 3)

    def nGrid(self):
        return self.quadratureGrid.weights.shape[0]

    def nBasis(self):
        return 0 if self.basisValues is None else self.basisValues.shape[1]

    def gradientsStored(self):
        return self.basisGradients is not None

    def setBasisFunctions(self, basisFunctions, storeGradients):
        """QMCGrid::setBasisFunctions (src/integration/integrateAll.hpp:88-137)."""
        torch = _torch()
        n, nb = self.nGrid(), len(basisFunctions)
        dev = self.quadratureGrid.points.device
        self.basisValues = torch.empty((n, nb), dtype=torch.float64, device=dev)
        self.basisGradients = torch.empty((n, nb, 3), dtype=torch.float64, device=dev) if storeGradients else None
        check(lib().tcscf_basis_eval_fns(_ptr(self.quadratureGrid.points), ctypes.c_int64(n), _fn_array(basisFunctions),
                                         nb, _ptr(self.basisValues),
                                         _ptr(self.basisGradients) if storeGradients else None, _stream()))
        self.basis = list(basisFunctions)


# --------------------------------------------------------------------------- F / V
class Jastrow:
    """u(r12) = c r12 / (r12 + a12) per spin pair: the r12-only Ochi term
    (src/jastrowFunctions.hpp:9-115, S = {(1,0,0)}).  BASELINE's
    u = r12 / (2 (1 + b r12)) is a12 = 1/b, c = a12/2."""

    def __init__(self, a12, c_same, c_oppo):
        self.a12, self.c_same, self.c_oppo = float(a12), float(c_same), float(c_oppo)

    @classmethod
    def from_b(cls, b, same_over_oppo=0.5):
        a12 = 1.0 / b
        return cls(a12, same_over_oppo * a12 / 2, a12 / 2)

    def c(self, same):
        return self.c_same if same else self.c_oppo

    def _struct(self):
        return _lib.Jastrow(self.a12, self.c_same, self.c_oppo)

    def _tc_struct(self):
        return self._struct()


class OchiJastrow:
    """The reference's general Ochi Jastrow, jastrowFunctions::Ochi{a, a12, c, S}
    (src/jastrowFunctions.hpp:9-115): u = sum_t c(t, sameSpin) (r12/(r12+a12))^i_t
    (r1/(r1+a))^j_t (r2/(r2+a))^k_t with S = [(i, j, k), ...] and
    c = [(c_oppo, c_same), ...] (the reference's c(t, false), c(t, true)); at
    most 8 terms, exponents 0..16."""

    def __init__(self, a, a12, c, S):
        c = np.asarray(c, dtype=np.float64).reshape(-1, 2)
        S = np.asarray(S, dtype=np.int64).reshape(-1, 3)
        if len(c) != len(S) or len(S) > _lib.OCHI_MAX_TERMS:
            raise ValueError("OchiJastrow: c and S need one row per term, at most 8 terms")
        if a <= 0 or a12 <= 0 or (S < 0).any() or (S > 16).any():
            raise ValueError("OchiJastrow: need a > 0, a12 > 0 and exponents in 0..16")
        self.a, self.a12, self.c, self.S = float(a), float(a12), c, S

    @classmethod
    def from_jastrow(cls, jastrow, a=1.0):
        """The r12-only Jastrow as an Ochi parameter set with S = [(1, 0, 0)]."""
        return cls(a, jastrow.a12, [(jastrow.c_oppo, jastrow.c_same)], [(1, 0, 0)])

    def _struct(self):
        u = _lib.OchiJastrow()
        u.a, u.a12, u.nterms = self.a, self.a12, len(self.S)
        for t, ((i, j, k), (co, cs)) in enumerate(zip(self.S, self.c)):
            u.i[t], u.j[t], u.k[t] = int(i), int(j), int(k)
            u.c_same[t], u.c_oppo[t] = float(cs), float(co)
        return u

    def spin_ratio(self):
        """rho with c_oppo = rho c_same for every term, or None if there is none
        (a same-spin coefficient that is zero must have a zero partner)."""
        cs, co = self.c[:, 1], self.c[:, 0]
        nz = cs != 0
        if not nz.any() or (co[~nz] != 0).any():
            return None
        rho = co[nz] / cs[nz]
        return float(rho[0]) if np.allclose(rho, rho[0], rtol=1e-14, atol=0) else None

    def _tc_struct(self):
        """The TC Fock build takes one V and (c_same, c_oppo) with V_same =
        c_same V and V_oppo = c_oppo V: for V = VjiSame that is (1, rho)."""
        rho = self.spin_ratio()
        if rho is None:
            raise ValueError("TC with an Ochi Jastrow needs c_oppo = rho * c_same term by term (one V tensor); "
                             "pass VjiSame as Vji")
        return _lib.Jastrow(self.a12, 1.0, rho)


def _compute_ochi(r1Grid, r2Grid, kernel, jastrow, same_spin, precision, out_dtype, ordered=True):
    torch = _torch()
    if r1Grid.nBasis() != r2Grid.nBasis() or r2Grid.basisValues is None:
        raise ValueError("computeF: both grids need the same tabulated basis")
    prec = _lib.PRECISIONS[precision] if isinstance(precision, str) else int(precision)
    nb = r2Grid.nBasis()
    n1, n2 = r1Grid.nGrid(), r2Grid.nGrid()
    dtype = torch.float64 if out_dtype in (torch.float64, "fp64", _lib.F64) else torch.float32
    shape = (n1, npair(nb)) if kernel == _lib.K_TC_SCALAR else (n1, 3, npair(nb))
    F = torch.empty(shape, dtype=dtype, device=r1Grid.quadratureGrid.points.device)
    scale = 2 * math.pi if r2Grid.ndim == 2 else 1.0
    u = jastrow._struct()
    check(lib().tcscf_compute_F_ochi(_ptr(r1Grid.quadratureGrid.points), ctypes.c_int64(n1),
                                     _ptr(r2Grid.quadratureGrid.points), _ptr(r2Grid.quadratureGrid.weights),
                                     _ptr(r2Grid.radial_order) if ordered else None, ctypes.c_int64(n2),
                                     _ptr(r2Grid.basisValues), nb, kernel, ctypes.byref(u), int(bool(same_spin)), prec,
                                     ctypes.c_double(scale), _ptr(F), _lib.F32 if dtype == torch.float32 else _lib.F64,
                                     _stream()))
    return F


def _compute(r1Grid, r2Grid, kernel, jastrow, same_spin, precision, out_dtype, ordered=True):
    torch = _torch()
    if r1Grid.nBasis() != r2Grid.nBasis() or r2Grid.basisValues is None:
        raise ValueError("computeF: both grids need the same tabulated basis")
    prec = _lib.PRECISIONS[precision] if isinstance(precision, str) else int(precision)
    nb = r2Grid.nBasis()
    n1, n2 = r1Grid.nGrid(), r2Grid.nGrid()
    ncomp = 3 if kernel == _lib.K_JASTROW_GRAD else 1
    dtype = torch.float64 if out_dtype in (torch.float64, "fp64", _lib.F64) else torch.float32
    shape = (n1, npair(nb)) if ncomp == 1 else (n1, 3, npair(nb))
    F = torch.empty(shape, dtype=dtype, device=r1Grid.quadratureGrid.points.device)
    scale = 2 * math.pi if r2Grid.ndim == 2 else 1.0
    jp = jastrow._struct() if jastrow is not None else None
    check(lib().tcscf_compute_F(_ptr(r1Grid.quadratureGrid.points), ctypes.c_int64(n1),
                                _ptr(r2Grid.quadratureGrid.points), _ptr(r2Grid.quadratureGrid.weights),
                                _ptr(r2Grid.radial_order) if ordered else None, ctypes.c_int64(n2), _ptr(r2Grid.basisValues), nb, kernel,
                                ctypes.byref(jp) if jp is not None else None, int(bool(same_spin)), prec,
                                ctypes.c_double(scale), _ptr(F), _lib.F32 if dtype == torch.float32 else _lib.F64,
                                _stream()))
    return F


def computeF(r1Grid, r2Grid, kernel="coulomb", jastrow=None, same_spin=True, precision="fp64",
             out_dtype="fp64", ordered=True):
    """precomputeIntegrand + computeF fused (src/HartreeFock.hpp:23-131):
    packed F[n1, p(j,i)] = scale * sum_n2 K(r1, r2) w2 phi_j(r2) phi_i(r2).
    kernel: 'coulomb' (1/r12) or 'tc' (2/r12 - |grad u12|^2 - |grad u21|^2,
    with a Jastrow or an OchiJastrow)."""
    k = {"coulomb": _lib.K_COULOMB, "tc": _lib.K_TC_SCALAR}[kernel]
    if isinstance(jastrow, OchiJastrow):
        if k != _lib.K_TC_SCALAR:
            raise ValueError("computeF: an OchiJastrow goes with kernel='tc'")
        return _compute_ochi(r1Grid, r2Grid, k, jastrow, same_spin, precision, out_dtype, ordered)
    return _compute(r1Grid, r2Grid, k, jastrow, same_spin, precision, out_dtype, ordered)


def computeF_tc_pair(r1Grid, r2Grid, jastrow, precision="fp64", smooth_precision=None, out_dtype="fp64",
                     ordered=True):
    """F_same and F_oppo of the TC scalar kernel together (the two computeF
    calls of src/unitTests/HartreeFockTest.cpp:276-296).  The kernel is
    2/r12 - 2 c^2 a12^2 / (r12 + a12)^4, so F_spin = 2 F_C - 2 c_spin^2 a12^2 G:
    F_C (1/r12) at `precision`, G (the smooth part) at `smooth_precision`
    (default: `precision`).  Returns (FjiSame, FjiOppo)."""
    torch = _torch()
    if not isinstance(jastrow, Jastrow):
        raise TypeError("computeF_tc_pair: needs the r12-only Jastrow (its kernel splits into 1/r12 and a smooth part)")
    if r1Grid.nBasis() != r2Grid.nBasis() or r2Grid.basisValues is None:
        raise ValueError("computeF_tc_pair: both grids need the same tabulated basis")
    prec = _lib.PRECISIONS[precision] if isinstance(precision, str) else int(precision)
    sp = prec if smooth_precision is None else (
        _lib.PRECISIONS[smooth_precision] if isinstance(smooth_precision, str) else int(smooth_precision))
    nb = r2Grid.nBasis()
    n1, n2 = r1Grid.nGrid(), r2Grid.nGrid()
    dtype = torch.float64 if out_dtype in (torch.float64, "fp64", _lib.F64) else torch.float32
    dev = r1Grid.quadratureGrid.points.device
    Fs = torch.empty((n1, npair(nb)), dtype=dtype, device=dev)
    Fo = torch.empty((n1, npair(nb)), dtype=dtype, device=dev)
    scale = 2 * math.pi if r2Grid.ndim == 2 else 1.0
    jp = jastrow._struct()
    check(lib().tcscf_compute_F_tc_pair(_ptr(r1Grid.quadratureGrid.points), ctypes.c_int64(n1),
                                        _ptr(r2Grid.quadratureGrid.points), _ptr(r2Grid.quadratureGrid.weights),
                                        _ptr(r2Grid.radial_order) if ordered else None, ctypes.c_int64(n2),
                                        _ptr(r2Grid.basisValues), nb, ctypes.byref(jp), prec, sp,
                                        ctypes.c_double(scale), _ptr(Fs), _ptr(Fo),
                                        _lib.F32 if dtype == torch.float32 else _lib.F64, _stream()))
    return Fs, Fo


def computeV(r1Grid, r3Grid, jastrow, precision="fp64", out_dtype="fp64", ordered=True, same_spin=True):
    """computeV (src/HartreeFock.hpp:185-229): packed V[n1, xyz, p].  For the
    r12-only Jastrow the c = 1 tensor (V_same = c_same V, V_oppo = c_oppo V);
    for an OchiJastrow the reference's VjiSame (same_spin) or VjiOppo."""
    if isinstance(jastrow, OchiJastrow):
        return _compute_ochi(r1Grid, r3Grid, _lib.K_JASTROW_GRAD, jastrow, same_spin, precision, out_dtype, ordered)
    return _compute(r1Grid, r3Grid, _lib.K_JASTROW_GRAD, jastrow, True, precision, out_dtype, ordered)


def _tc_dtype(Fs, Fo, V, n1, Vo=None):
    """The f_dtype of the TC tensors: all float64 or all float32 (FP32
    storage; the Fock build's arithmetic is FP64 either way), one row per r1
    point, each V the size of the first."""
    torch = _torch()
    ts = (Fs, Fo, V) + ((Vo,) if Vo is not None else ())
    dts = {T.dtype for T in ts}
    if len(dts) != 1 or dts.pop() not in (torch.float64, torch.float32) or any(T.shape[0] != n1 for T in ts):
        raise ValueError("TC: F_same, F_oppo and V must all be float64 or all float32, one row per r1 point")
    if Vo is not None and Vo.numel() != V.numel():
        raise ValueError("TC: VjiSame and VjiOppo must have the same shape")
    return _lib.F32 if Fs.dtype == torch.float32 else _lib.F64


def _observables(fock, eigenvectors, occupied):
    """tcscf_scf_observables over numpy float64 arrays (each may be None)."""
    def ptr(a):
        if a is None:
            return None
        assert a.dtype == np.float64 and a.flags.c_contiguous
        return a.ctypes.data
    return _lib.ScfObservables(ptr(fock), ptr(eigenvectors), ptr(occupied))


def computeF_batch(specs, precision="bf16x5", out_dtype="fp64"):
    """computeF / computeV for many instances in one call (BF16: one persistent
    launch over a work queue of tiles of every instance, results equal to the
    single calls bit for bit).  specs: list of dicts with keys r1, r2, kernel
    ('coulomb', 'tc' or 'grad'), jastrow (TC kernels) and same_spin.  Returns
    the packed tensors (out_dtype 'fp64' or 'fp32' storage) in spec order."""
    torch = _torch()
    prec = _lib.PRECISIONS[precision] if isinstance(precision, str) else int(precision)
    f32 = out_dtype in (torch.float32, "fp32", _lib.F32)
    kinds = {"coulomb": _lib.K_COULOMB, "tc": _lib.K_TC_SCALAR, "grad": _lib.K_JASTROW_GRAD}
    items = (_lib.BatchItem * len(specs))()
    outs = []
    for k, sp in enumerate(specs):
        r1, r2 = sp["r1"], sp["r2"]
        kind = kinds[sp.get("kernel", "coulomb")]
        if r1.nBasis() != r2.nBasis() or r2.basisValues is None:
            raise ValueError("computeF_batch: both grids need the same tabulated basis")
        nb, n1 = r2.nBasis(), r1.nGrid()
        shape = (n1, npair(nb)) if kind != _lib.K_JASTROW_GRAD else (n1, 3, npair(nb))
        F = torch.empty(shape, dtype=torch.float32 if f32 else torch.float64, device=r1.quadratureGrid.points.device)
        outs.append(F)
        it = items[k]
        it.pts1_dev, it.n1 = r1.quadratureGrid.points.data_ptr(), n1
        it.pts2_dev, it.w2_dev = r2.quadratureGrid.points.data_ptr(), r2.quadratureGrid.weights.data_ptr()
        it.ord2_dev, it.n2 = r2.radial_order.data_ptr(), r2.nGrid()
        it.A2_dev, it.nb, it.kernel = r2.basisValues.data_ptr(), nb, kind
        jas = sp.get("jastrow")
        if jas is not None:
            it.jastrow = jas._struct()
        it.same_spin = int(bool(sp.get("same_spin", True)))
        it.scale = 2 * math.pi if r2.ndim == 2 else 1.0
        it.F_dev = F.data_ptr()
        it.out_dtype = _lib.F32 if f32 else _lib.F64
    check(lib().tcscf_batch_compute_F(items, len(specs), prec, _stream()))
    return outs


def fillR12Array(basisFunctions, min_points, nshifts=2, seed=0):
    """fillR12Array (src/AtomicBasis.hpp:167-177) with r12MatrixElement
    (:30-48): every (ab|cd) as one 6-D QMC integral over (r1, r2) on the point
    set of qmc_spherical_points(ndim=6), in the real basis.  Every quadruple
    comes from one device pass; as the reference fills the array
    (fillTwoElectronSymmetricHermitianArray with real functions), the canonical
    quadruples' values are copied to their symmetry images, so entries that
    differ only by the electron swap are equal.  Returns the (nb, nb, nb, nb)
    host array answer[a, b, c, d]."""
    torch = _torch()
    p1, p2, w = qmc_spherical_points(min_points, 6, nshifts, seed)
    g1, g2 = QMCGrid(p1, w), QMCGrid(p2, w)
    g1.setBasisFunctions(basisFunctions, False)
    g2.setBasisFunctions(basisFunctions, False)
    nb, n = g1.nBasis(), g1.nGrid()
    E = torch.empty((npair(nb), npair(nb)), dtype=torch.float64, device=g1.basisValues.device)
    check(lib().tcscf_eri_qmc6d(_ptr(g1.quadratureGrid.points), _ptr(g2.quadratureGrid.points),
                                _ptr(g1.quadratureGrid.weights), ctypes.c_int64(n), _ptr(g1.basisValues),
                                _ptr(g2.basisValues), nb, _ptr(E), _stream()))
    full = eri_full(E, nb)
    return _fill_two_electron(nb, True, lambda a, b, c, d: full[a, b, c, d]).real


def qmc_overlap(basisFunctions, min_points, nshifts=2, seed=0):
    """overlap (src/AtomicBasis.hpp:15-26) for every pair as a 3-D QMC integral
    on the point set of qmc_spherical_points(ndim=3), real basis: (nb, nb)."""
    p1, _, w = qmc_spherical_points(min_points, 3, nshifts, seed)
    g = QMCGrid(p1, w)
    g.setBasisFunctions(basisFunctions, False)
    A = g.basisValues
    return (A.t() @ (A * g.quadratureGrid.weights[:, None])).cpu().numpy()


def _fill_two_electron(N, real, value):
    """fillTwoElectronSymmetricHermitianArray (src/mathFunctions.hpp:367-433):
    value(a, b, c, d) is asked for the canonical quadruples only (a <= b,
    a <= c, d >= b for real functions, else d >= a) and copied to their images
    under the electron swap (b, a, d, c), Hermitian conjugation (c, d, a, b),
    (d, c, b, a) and, for real functions, b <-> d; entries accumulate."""
    A = np.zeros((N,) * 4, complex)
    for a in range(N):
        for b in range(a, N):
            for c in range(a, N):
                for d in range(b if real else a, N):
                    v = complex(value(a, b, c, d))
                    if v == 0:
                        continue
                    vc = v.conjugate()
                    sym, herm, rl = a != b, a != c and a != d, real and d != b
                    A[a, b, c, d] += v
                    if sym:
                        A[b, a, d, c] += v
                    if herm or (real and a != c):
                        A[c, d, a, b] += vc
                    if herm and sym:
                        A[d, c, b, a] += vc
                    if rl:
                        A[a, d, c, b] += v
                        if sym:
                            A[d, a, b, c] += v
                        if herm:
                            A[c, b, a, d] += v
                        if herm and sym:
                            A[b, c, d, a] += v
    return A


def _atomic_eri5d(basisFunctions, min_points, nshifts, seed):
    """Every quadruple of tcscf_eri_qmc5d_atomic for the reference's complex
    functions (l, m) of basisFunctions: complex (nb,)*4 numpy array.  Missing -m
    partners are tabulated alongside (the complex functions are formed from
    real pairs) and dropped from the result."""
    import copy
    torch = _torch()
    basis = list(basisFunctions)
    key = lambda b: (type(b).__name__, b.kind, b.n, b.l, b.m, b.alpha)
    index = {key(b): k for k, b in enumerate(basis)}
    for b in list(basis):
        if b.m != 0:
            q = copy.copy(b)
            q.m = -b.m
            if key(q) not in index:
                index[key(q)] = len(basis)
                basis.append(q)
    nb = len(basis)
    if nb > 32:
        raise ValueError("atomic QMC integrals: at most 32 functions (with their -m partners)")
    partner = np.array([k if b.m == 0 else index[(type(b).__name__, b.kind, b.n, b.l, -b.m, b.alpha)]
                        for k, b in enumerate(basis)], dtype=np.int32)
    ms = np.array([b.m for b in basis], dtype=np.int32)
    p1, p2, w = qmc_spherical_points(min_points, 5, nshifts, seed)
    g1, g2 = QMCGrid(p1, w), QMCGrid(p2, w)
    g1.setBasisFunctions(basis, False)
    g2.setBasisFunctions(basis, False)
    E = torch.empty((nb, nb, nb, nb, 2), dtype=torch.float64, device=g1.basisValues.device)
    check(lib().tcscf_eri_qmc5d_atomic(_ptr(g1.quadratureGrid.points), _ptr(g2.quadratureGrid.points),
                                       _ptr(g1.quadratureGrid.weights), ctypes.c_int64(g1.nGrid()),
                                       _ptr(g1.basisValues), _ptr(g2.basisValues), nb, ms.ctypes.data,
                                       partner.ctypes.data, _ptr(E), _stream()))
    E = E.cpu().numpy()
    n0 = len(basisFunctions)
    return (E[..., 0] + 1j * E[..., 1])[:n0, :n0, :n0, :n0]


def atomicCoulombOperator(b1, b2, b3, b4, min_points, nshifts=2, seed=0):
    """atomicCoulombOperator (src/AtomicBasis.hpp:59-98): the 5-D QMC integral
    of |Y_l1m1(theta1)|^2 |Y_l2m2(theta2)|^2 R1 R3(r1) R2 R4(r2) / r12 on the
    point set of qmc_spherical_points(ndim=5); b1, b3 and b2, b4 must share
    (l, m).  Returns a float."""
    if (b1.l, b1.m) != (b3.l, b3.m) or (b2.l, b2.m) != (b4.l, b4.m):
        raise ValueError("atomicCoulombOperator: b1, b3 and b2, b4 need equal (l, m)")
    E = _atomic_eri5d([b1, b2, b3, b4], min_points, nshifts, seed)
    return float(E[0, 1, 2, 3].real)


def atomicExchangeOperator(b1, b2, b3, b4, min_points, nshifts=2, seed=0):
    """atomicExchangeOperator (src/AtomicBasis.hpp:103-140): the 5-D QMC
    integral of R1 P_l1m1(theta1) R2 P_l2m2(theta2) R3 P_l2m2(theta1)
    R4 P_l1m1(theta2) e^{i (m2 - m1) phi1} / r12; b1, b4 and b2, b3 must share
    (l, m).  Returns a complex."""
    if (b1.l, b1.m) != (b4.l, b4.m) or (b2.l, b2.m) != (b3.l, b3.m):
        raise ValueError("atomicExchangeOperator: b1, b4 and b2, b3 need equal (l, m)")
    E = _atomic_eri5d([b1, b2, b3, b4], min_points, nshifts, seed)
    return complex(E[0, 1, 2, 3])


def fillAtomicR12Array(basisFunctions, min_points, nshifts=2, seed=0):
    """fillAtomicR12Array (src/AtomicBasis.hpp:197-248): the (nb,)*4 complex
    array of the reference's complex functions holding the Coulomb integrals
    ((l, m) of a, c and of b, d equal; real-function symmetry) and, elsewhere,
    the exchange integrals ((l, m) of a, d and of b, c equal; Hermitian
    symmetry), 0 otherwise.  Every integral comes from one device pass over all
    quadruples on one 5-D point set; the reference's canonical quadruples and
    their symmetry copies are kept as it fills them."""
    E = _atomic_eri5d(basisFunctions, min_points, nshifts, seed)
    lm = [(b.l, b.m) for b in basisFunctions]
    N = len(lm)
    coul = lambda a, b, c, d: lm[a] == lm[c] and lm[b] == lm[d]
    exch = lambda a, b, c, d: lm[a] == lm[d] and lm[b] == lm[c]
    A = _fill_two_electron(N, True, lambda a, b, c, d: E[a, b, c, d].real if coul(a, b, c, d) else 0)
    A += _fill_two_electron(N, False, lambda a, b, c, d: E[a, b, c, d] if exch(a, b, c, d) and not coul(a, b, c, d)
                            else 0)
    return A


def integrateAllR1R2(r1Grid, F):
    """Grid ERI tensor (src/integration/integrateAll.hpp:221-317, USE_GRADIENT =
    false) from F = computeF(r1Grid, r2Grid, ...): returns the packed
    (npair, npair) device tensor E with E[p(b1,b3), p(b2,b4)] =
    answer(b1, b2, b3, b4).  eri_full() expands it to nb^4."""
    torch = _torch()
    nb = r1Grid.nBasis()
    if F.dtype != torch.float64 or F.shape != (r1Grid.nGrid(), npair(nb)):
        raise ValueError("integrateAllR1R2: F must be packed FP64 (n1, npair)")
    E = torch.empty((npair(nb), npair(nb)), dtype=torch.float64, device=F.device)
    check(lib().tcscf_eri(_ptr(r1Grid.basisValues), _ptr(r1Grid.quadratureGrid.weights),
                          ctypes.c_int64(r1Grid.nGrid()), nb, _ptr(F), _ptr(E), _stream()))
    return E


def integrateAllR1R2Gradient(r1Grid, r2Grid, jastrow):
    """integrateAllR1R2<true> (src/integration/integrateAll.hpp:184-317) with the
    vector kernel f = c_same grad_1 u(r1, r2): the packed (npair, nb^2) device
    tensor E with E[p(b1,b3), b2 nb + b4] = answer(b2, b1, b4, b3) =
    sum_r1 w1 phi_b1 phi_b3 scale sum_r2 w2 phi_b2 f.grad phi_b4 (scale = 2 pi
    for a 2-D r2 grid).  eri_full_gradient() expands it to nb^4."""
    torch = _torch()
    nb = r1Grid.nBasis()
    if r2Grid.nBasis() != nb or not r2Grid.gradientsStored():
        raise ValueError("integrateAllR1R2Gradient: the r2 grid needs the same basis with gradients stored")
    E = torch.empty((npair(nb), nb * nb), dtype=torch.float64, device=r1Grid.basisValues.device)
    scale = 2 * math.pi if r2Grid.ndim == 2 else 1.0
    jp = jastrow._struct()
    check(lib().tcscf_eri_gradient(_ptr(r1Grid.quadratureGrid.points), _ptr(r1Grid.basisValues),
                                   _ptr(r1Grid.quadratureGrid.weights), ctypes.c_int64(r1Grid.nGrid()),
                                   _ptr(r2Grid.quadratureGrid.points), _ptr(r2Grid.quadratureGrid.weights),
                                   ctypes.c_int64(r2Grid.nGrid()), _ptr(r2Grid.basisValues),
                                   _ptr(r2Grid.basisGradients), nb, _lib.K_JASTROW_GRAD, ctypes.byref(jp),
                                   ctypes.c_double(scale), _ptr(E), _stream()))
    return E


def eri_full_gradient(E, nb):
    """Packed gradient ERI -> answer(b2, b1, b4, b3) (numpy, index order
    (b2, b1, b4, b3) as the reference stores it)."""
    idx = np.array([[pair_index(j, i) for i in range(nb)] for j in range(nb)])
    E = E.cpu().numpy() if not isinstance(E, np.ndarray) else E
    b = np.arange(nb)
    # out[b2, b1, b4, b3] = E[idx[b1, b3], b2 * nb + b4]
    return E[idx[None, :, None, :], (b[:, None, None, None] * nb + b[None, None, :, None])]


def threeElectronIntegrals(r1Grid, r2Grid, jastrow, precision="fp64"):
    """threeElectronIntegrals (src/integration/integrateAll.hpp:369-442): the
    (nb,)*6 device tensor Q(p,q,r,s,t,u) = sum_r1 w1 phi_p phi_s(r1)
    V(r1,q,t).V(r1,r,u), with V(r1,q,t) = sum_r2 w2 phi_q phi_t(r2) grad_1 u(r1, r2)
    for the same-spin Jastrow.  Validation-sized (nb <= 16)."""
    torch = _torch()
    nb = r1Grid.nBasis()
    if r2Grid.nBasis() != nb:
        raise ValueError("threeElectronIntegrals: the grids carry different bases")
    V = computeV(r1Grid, r2Grid, jastrow, precision=precision)
    Q = torch.empty((nb,) * 6, dtype=torch.float64, device=V.device)
    check(lib().tcscf_three_electron(_ptr(r1Grid.basisValues), _ptr(r1Grid.quadratureGrid.weights),
                                     ctypes.c_int64(r1Grid.nGrid()), _ptr(V), nb, ctypes.c_double(jastrow.c_same),
                                     _ptr(Q), _stream()))
    return Q


def batch_scf(instances, tol=1e-5, max_iter=200, respectOneElectronSymmetry=True, mgs_mode=1, diis=0):
    """Many SCF solves at once (tcscf_batch_scf): every iteration replays one
    HIP graph with the Fock builds of all active instances.  instances: dicts
    with kind ('rcs', 'uos' or 'tc'), nup, ndown, S, H, r1 (QMCGrid with the
    basis tabulated; gradients for TC) and F (RCS/UOS) or Fs, Fo, V, jastrow
    (TC).  Returns (list of dicts energy/iterations/status/trace/D, graph builds,
    device ms, host ms)."""
    torch = _torch()
    n = len(instances)
    items = (_lib.ScfItem * n)()
    keep = []
    traces = [np.zeros(max_iter) for _ in range(n)]
    Ds = []
    kinds = {"rcs": 0, "uos": 1, "tc": 2}
    for k, x in enumerate(instances):
        it = items[k]
        g = x["r1"]
        nb = g.nBasis()
        S, _ = _np(x["S"])
        H, _ = _np(x["H"])
        keep += [S, H]
        it.kind, it.nup, it.ndown, it.nb = kinds[x["kind"]], int(x["nup"]), int(x["ndown"]), nb
        it.S, it.H = S.ctypes.data, H.ctypes.data
        it.A1_dev, it.w1_dev, it.n1 = g.basisValues.data_ptr(), g.quadratureGrid.weights.data_ptr(), g.nGrid()
        if x["kind"] == "tc":
            if not g.gradientsStored():
                raise ValueError("batch_scf: TC instances need basis gradients on r1")
            it.dA1_dev = g.basisGradients.data_ptr()
            it.f_dtype = _tc_dtype(x["Fs"], x["Fo"], x["V"], g.nGrid())
            it.Fs_dev, it.Fo_dev, it.V_dev = x["Fs"].data_ptr(), x["Fo"].data_ptr(), x["V"].data_ptr()
            it.Vo_dev = None
            it.jastrow = x["jastrow"]._struct()
        else:
            F = x["F"]
            it.F_dev = F.data_ptr()
            it.f_dtype = _lib.F32 if F.dtype == torch.float32 else _lib.F64
        it.energy_trace = traces[k].ctypes.data
        D = np.zeros((1 if x["kind"] == "rcs" else 2, nb, nb))
        Ds.append(D)
        it.D_out = D.ctypes.data
    res = (_lib.ScfResult * n)()
    builds = ctypes.c_int(0)
    check(lib().tcscf_batch_scf(items, n, ctypes.byref(_opts(tol, max_iter, respectOneElectronSymmetry, mgs_mode,
                                                              diis)), res, ctypes.byref(builds), _stream()))
    out = [dict(energy=res[k].energy, iterations=res[k].iterations, status=res[k].status,
                trace=traces[k][:max(res[k].iterations, 0)].copy(), D=Ds[k]) for k in range(n)]
    return out, builds.value, (res[0].t_sweep_ms if n else 0.0), (res[0].t_solve_ms if n else 0.0)


def eri_full(E, nb):
    """Packed ERI -> answer(b1, b2, b3, b4) (numpy)."""
    idx = np.array([[pair_index(j, i) for i in range(nb)] for j in range(nb)])
    E = E.cpu().numpy() if not isinstance(E, np.ndarray) else E
    return E[idx[:, None, :, None], idx[None, :, None, :]]


# --------------------------------------------------------------------------- SCF
def _opts(tol, max_iter, mask, mgs_mode=0, diis=0, graph=None):
    """graph: None the driver's default, True a per-iteration HIP graph, False
    direct launches (tcscf_scf_opts.graph 0 / 1 / 2)."""
    g = 0 if graph is None else (1 if graph else 2)
    return _lib.ScfOpts(float(tol), int(max_iter), int(bool(mask)), int(mgs_mode), int(diis), g)


def _f_args(r1Grid, F):
    torch = _torch()
    if F.dtype not in (torch.float64, torch.float32):
        raise TypeError("F must be float64 or float32")
    if F.shape[0] != r1Grid.nGrid():
        raise ValueError("F rows must match the r1 grid")
    return _ptr(F), (_lib.F32 if F.dtype == torch.float32 else _lib.F64)


class _HFBase:
    def numberOfConvergenceLoops(self):
        return self._iters

    def _finish(self, rc, res, trace):
        self._iters = res.iterations
        self.energy = res.energy
        self.energy_trace = trace[:max(res.iterations, 0)].copy()
        self.t_sweep_ms = res.t_sweep_ms
        self.t_solve_ms = res.t_solve_ms
        self.graph_iterations = res.graph_iterations  # iterations replayed from a HIP graph
        self.graph_captures = res.graph_captures  # captures of that graph during this call
        check(rc)
        return res.energy


class RCSHartreeFock(_HFBase):
    """RCSHartreeFock (src/HartreeFock.hpp:233-293, src/HartreeFock.cpp:785-886)."""

    def __init__(self, nElectrons, nBasis):
        if nElectrons % 2:
            raise ValueError("RCS needs an even number of electrons")
        self.nElectrons, self.basisSize = nElectrons, nBasis
        self.density = np.zeros((1, nBasis, nBasis))
        self.fockOperator = np.zeros((1, nBasis, nBasis))
        self.eigenvalues = np.zeros(nBasis)
        self.eigenvectors = np.zeros((1, nBasis, nBasis))
        self._iters = 0

    @staticmethod
    def needsGradients():
        return False

    def compute(self, overlap, oneElectronTerms, r1Grid, Fji, respectOneElectronSymmetry=True, tol=1e-5,
                max_iter=50, diis=0, graph=None):
        nb = self.basisSize
        S, Sp = _np(overlap)
        H, Hp = _np(oneElectronTerms)
        if S.shape != (nb, nb) or H.shape != (nb, nb) or r1Grid.nBasis() != nb:
            raise ValueError("RCS: shape mismatch")
        fp, fd = _f_args(r1Grid, Fji)
        res = _lib.ScfResult()
        trace = np.zeros(max_iter)
        D = np.zeros((nb, nb))
        ev = np.zeros(nb)
        obs = _observables(self.fockOperator, self.eigenvectors, None)
        rc = lib().tcscf_rcs_scf(self.nElectrons, nb, Sp, Hp, _ptr(r1Grid.basisValues),
                                 _ptr(r1Grid.quadratureGrid.weights), ctypes.c_int64(r1Grid.nGrid()), fp, fd,
                                 ctypes.byref(_opts(tol, max_iter, respectOneElectronSymmetry, 0, diis, graph)), ctypes.byref(res),
                                 trace.ctypes.data_as(_dp), D.ctypes.data_as(_dp), ev.ctypes.data_as(_dp),
                                 ctypes.byref(obs), _stream())
        self.density[0] = D
        self.eigenvalues[:] = ev
        return self._finish(rc, res, trace)

    def highestOccupiedOrbitalEnergy(self):
        return self.eigenvalues[self.nElectrons // 2 - 1]


class UOSHartreeFock(_HFBase):
    """UOSHartreeFock (src/HartreeFock.hpp:297-349, src/HartreeFock.cpp:890-1003)."""

    def __init__(self, numSpinUp, numSpinDown, nBasis):
        self.nElectrons = (numSpinUp, numSpinDown)
        self.basisSize = nBasis
        self.density = np.zeros((2, nBasis, nBasis))
        self.fockOperator = np.zeros((2, nBasis, nBasis))
        self.eigenvalues = np.zeros((2, nBasis))
        self.eigenvectors = np.zeros((2, nBasis, nBasis))
        self._iters = 0

    @staticmethod
    def needsGradients():
        return False

    def compute(self, overlap, oneElectronTerms, r1Grid, Fji, respectOneElectronSymmetry=True, tol=1e-5,
                max_iter=100, diis=0, graph=None):
        nb = self.basisSize
        S, Sp = _np(overlap)
        H, Hp = _np(oneElectronTerms)
        if S.shape != (nb, nb) or H.shape != (nb, nb) or r1Grid.nBasis() != nb:
            raise ValueError("UOS: shape mismatch")
        fp, fd = _f_args(r1Grid, Fji)
        res = _lib.ScfResult()
        trace = np.zeros(max_iter)
        D = np.zeros((2, nb, nb))
        ev = np.zeros((2, nb))
        obs = _observables(self.fockOperator, self.eigenvectors, None)
        rc = lib().tcscf_uos_scf(self.nElectrons[0], self.nElectrons[1], nb, Sp, Hp, _ptr(r1Grid.basisValues),
                                 _ptr(r1Grid.quadratureGrid.weights), ctypes.c_int64(r1Grid.nGrid()), fp, fd,
                                 ctypes.byref(_opts(tol, max_iter, respectOneElectronSymmetry, 0, diis, graph)), ctypes.byref(res),
                                 trace.ctypes.data_as(_dp), D.ctypes.data_as(_dp), ev.ctypes.data_as(_dp),
                                 ctypes.byref(obs), _stream())
        self.density[:] = D
        self.eigenvalues[:] = ev
        return self._finish(rc, res, trace)

    def highestOccupiedOrbitalEnergy(self):
        return max(self.eigenvalues[0, self.nElectrons[0] - 1], self.eigenvalues[1, self.nElectrons[1] - 1])


class TCHartreeFock(_HFBase):
    """TCHartreeFock (src/HartreeFock.hpp:353-418, src/HartreeFock.cpp:1007-1362).

    compute(S, H, r1Grid, FjiSame, FjiOppo, VjiSame, VjiOppo, ...) takes the
    packed F_same / F_oppo (computeF with kernel='tc', or an OchiJastrow per
    spin) and V_same / V_oppo as the reference does.  In place of VjiOppo a
    Jastrow (or an OchiJastrow with c_oppo = rho c_same) may be given: VjiSame is
    then the c = 1 tensor V of computeV and V_same = c_same V, V_oppo = c_oppo V
    are applied inside the Fock build, which reads one tensor instead of two.
    mgs_mode 0 keeps the reference's Gram-Schmidt (Euclidean norm after the
    first vector), 1 uses the overlap metric throughout.  After compute:
    density, fockOperator (2, nb, nb), occupiedOrbitals [spin] (nb, nocc_s),
    occupiedOrbitalPseudoEnergy."""

    def __init__(self, numSpinUp, numSpinDown, nBasis):
        self.nElectrons = (numSpinUp, numSpinDown)
        self.basisSize = nBasis
        self.density = np.zeros((2, nBasis, nBasis))
        self.fockOperator = np.zeros((2, nBasis, nBasis))
        self.occupiedOrbitalPseudoEnergy = [np.zeros(numSpinUp), np.zeros(numSpinDown)]
        self.occupiedOrbitals = [np.zeros((nBasis, numSpinUp)), np.zeros((nBasis, numSpinDown))]
        self._iters = 0

    @staticmethod
    def needsGradients():
        return True

    @staticmethod
    def _v_args(VjiSame, VjiOppo, n1):
        """(Vs, Vo or None, jastrow struct or None) and the f_dtype check."""
        if isinstance(VjiOppo, (Jastrow, OchiJastrow)):
            return VjiSame, None, VjiOppo._tc_struct()
        return VjiSame, VjiOppo, None

    def compute(self, overlap, oneElectronTerms, r1Grid, FjiSame, FjiOppo, VjiSame, VjiOppo,
                respectOneElectronSymmetry=True, tol=1e-5, max_iter=200, mgs_mode=0, diis=0, graph=None):
        nb = self.basisSize
        S, Sp = _np(overlap)
        H, Hp = _np(oneElectronTerms)
        if S.shape != (nb, nb) or H.shape != (nb, nb) or r1Grid.nBasis() != nb or not r1Grid.gradientsStored():
            raise ValueError("TC: shape mismatch or basis gradients not stored on the r1 grid")
        n1 = r1Grid.nGrid()
        Vs, Vo, jp = self._v_args(VjiSame, VjiOppo, n1)
        fdt = _tc_dtype(FjiSame, FjiOppo, Vs, n1, Vo)
        res = _lib.ScfResult()
        trace = np.zeros(max_iter)
        D = np.zeros((2, nb, nb))
        ps = np.zeros(sum(self.nElectrons))
        nmax = max(1, max(self.nElectrons))
        occ = np.zeros((2, nb, nmax))
        obs = _observables(self.fockOperator, None, occ)
        rc = lib().tcscf_tc_scf(self.nElectrons[0], self.nElectrons[1], nb, Sp, Hp, _ptr(r1Grid.basisValues),
                                _ptr(r1Grid.basisGradients), _ptr(r1Grid.quadratureGrid.weights), ctypes.c_int64(n1),
                                _ptr(FjiSame), _ptr(FjiOppo), _ptr(Vs), _ptr(Vo) if Vo is not None else None, fdt,
                                ctypes.byref(jp) if jp is not None else None,
                                ctypes.byref(_opts(tol, max_iter, respectOneElectronSymmetry, mgs_mode, diis, graph)),
                                ctypes.byref(res), trace.ctypes.data_as(_dp), D.ctypes.data_as(_dp),
                                ps.ctypes.data_as(_dp), ctypes.byref(obs), _stream())
        self.density[:] = D
        self.occupiedOrbitalPseudoEnergy = [ps[:self.nElectrons[0]].copy(), ps[self.nElectrons[0]:].copy()]
        self.occupiedOrbitals = [occ[s, :, :self.nElectrons[s]].copy() for s in range(2)]
        return self._finish(rc, res, trace)

    def highestOccupiedOrbitalEnergy(self):
        return max(self.occupiedOrbitalPseudoEnergy[0][-1], self.occupiedOrbitalPseudoEnergy[1][-1])

    def fockBuild(self, r1Grid, FjiSame, FjiOppo, VjiSame, VjiOppo, density, occupiedOrbitals):
        """One Fock build of the TC loop (src/HartreeFock.cpp:1118-1224) at a
        given density (2, nb, nb) and occupied orbitals (2, nb, max nocc):
        returns the two-electron term G2 and the three-electron term G3, each
        (2, nb, nb), before the one-electron-symmetry mask.  VjiOppo as in
        compute (a tensor, or a Jastrow with VjiSame the c = 1 tensor)."""
        nb = self.basisSize
        nmax = max(self.nElectrons)
        D = np.ascontiguousarray(density, dtype=np.float64)
        C = np.ascontiguousarray(occupiedOrbitals, dtype=np.float64)
        if D.shape != (2, nb, nb) or C.shape != (2, nb, nmax) or r1Grid.nBasis() != nb:
            raise ValueError("fockBuild: density must be (2, nb, nb) and orbitals (2, nb, max nocc)")
        Vs, Vo, jp = self._v_args(VjiSame, VjiOppo, r1Grid.nGrid())
        fdt = _tc_dtype(FjiSame, FjiOppo, Vs, r1Grid.nGrid(), Vo)
        G2 = np.zeros((2, nb, nb))
        G3 = np.zeros((2, nb, nb))
        check(lib().tcscf_tc_fock(self.nElectrons[0], self.nElectrons[1], nb, _ptr(r1Grid.basisValues),
                                  _ptr(r1Grid.basisGradients), _ptr(r1Grid.quadratureGrid.weights),
                                  ctypes.c_int64(r1Grid.nGrid()), _ptr(FjiSame), _ptr(FjiOppo), _ptr(Vs),
                                  _ptr(Vo) if Vo is not None else None, fdt,
                                  ctypes.byref(jp) if jp is not None else None, D.ctypes.data_as(_dp),
                                  C.ctypes.data_as(_dp), G2.ctypes.data_as(_dp), G3.ctypes.data_as(_dp), _stream()))
        return G2, G3
