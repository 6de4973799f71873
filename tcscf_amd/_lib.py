"""ctypes binding of libtcscf_hip.so (the C ABI declared in include/tcscf_hip.h).

The shared library is built in-tree by tcscf_amd.build; importing this module
never falls back to anything else: a missing library raises.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtcscf_hip.so")

# enums (include/tcscf_hip.h)
OK, EINVAL, EHIP, ENOTCONV, ESOLVER, ENOMEM = 0, -1, -2, -3, -4, -5
K_COULOMB, K_TC_SCALAR, K_JASTROW_GRAD, K_TC_SMOOTH = 0, 1, 2, 3
FP64, FP32, BF16X3, BF16X5, BF16X4 = 0, 1, 2, 3, 4
F64, F32 = 0, 1
RADIAL_M4, RADIAL_BECKE, RADIAL_XOVER1MX = 0, 1, 2

PRECISIONS = {"fp64": FP64, "fp32": FP32, "bf16x3": BF16X3, "bf16x4": BF16X4, "bf16x5": BF16X5}

# every symbol include/tcscf_hip.h declares
EXPORTS = [
    "tcscf_last_error", "tcscf_version", "tcscf_radial_grid", "tcscf_lebedev_npts", "tcscf_lebedev",
    "tcscf_product_grid", "tcscf_sto_overlap_core", "tcscf_basis_eval", "tcscf_compute_F", "tcscf_compute_F_noalloc",
    "tcscf_F_workspace_bytes", "tcscf_compute_F_tc_pair", "tcscf_compute_F_ochi", "tcscf_eri_qmc6d", "tcscf_eri_qmc5d_atomic", "tcscf_qmc_spherical_points", "tcscf_basis_eval_fns", "tcscf_basis_one_electron", "tcscf_rcs_scf", "tcscf_uos_scf", "tcscf_tc_scf", "tcscf_tc_fock", "tcscf_tc_launch_info", "tcscf_eri", "tcscf_eri_gradient", "tcscf_three_electron",
    "tcscf_qmc_lattice_npts", "tcscf_qmc_lattice_grid", "tcscf_sobol_points", "tcscf_sobol_grid", "tcscf_load_sto",
    "tcscf_batch_compute_F", "tcscf_batch_tuned_rows", "tcscf_batch_scf",
    "tcscf_selftest_mfma",
]


class TcscfError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"tcscf error {code}: {msg}")
        self.code = code


class ConvergenceError(TcscfError):
    """Raised when an SCF loop hits its iteration cap (reference: ConvergenceError)."""


class SolverError(TcscfError):
    """Raised when the generalized eigen-solve fails (reference: LvArray::dense::Error)."""


class Sto(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("l", ctypes.c_int), ("m", ctypes.c_int), ("zeta", ctypes.c_double)]


class Jastrow(ctypes.Structure):
    _fields_ = [("a12", ctypes.c_double), ("c_same", ctypes.c_double), ("c_oppo", ctypes.c_double)]


OCHI_MAX_TERMS = 8


class OchiJastrow(ctypes.Structure):
    _fields_ = [("a", ctypes.c_double), ("a12", ctypes.c_double), ("nterms", ctypes.c_int),
                ("i", ctypes.c_int * OCHI_MAX_TERMS), ("j", ctypes.c_int * OCHI_MAX_TERMS),
                ("k", ctypes.c_int * OCHI_MAX_TERMS), ("c_same", ctypes.c_double * OCHI_MAX_TERMS),
                ("c_oppo", ctypes.c_double * OCHI_MAX_TERMS)]


BASIS_STO, BASIS_OCHI, BASIS_HYDROGEN = 0, 1, 2


class BasisFn(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("n", ctypes.c_int), ("l", ctypes.c_int), ("m", ctypes.c_int),
                ("alpha", ctypes.c_double)]


class BatchItem(ctypes.Structure):
    _fields_ = [("pts1_dev", ctypes.c_void_p), ("n1", ctypes.c_int64), ("pts2_dev", ctypes.c_void_p),
                ("w2_dev", ctypes.c_void_p), ("ord2_dev", ctypes.c_void_p), ("n2", ctypes.c_int64),
                ("A2_dev", ctypes.c_void_p), ("nb", ctypes.c_int), ("kernel", ctypes.c_int), ("jastrow", Jastrow),
                ("same_spin", ctypes.c_int), ("scale", ctypes.c_double), ("F_dev", ctypes.c_void_p),
                ("out_dtype", ctypes.c_int)]


class ScfOpts(ctypes.Structure):
    _fields_ = [("tol", ctypes.c_double), ("max_iter", ctypes.c_int), ("mask", ctypes.c_int),
                ("mgs_mode", ctypes.c_int), ("diis", ctypes.c_int), ("graph", ctypes.c_int)]


class ScfItem(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("nup", ctypes.c_int), ("ndown", ctypes.c_int), ("nb", ctypes.c_int),
                ("S", ctypes.c_void_p), ("H", ctypes.c_void_p), ("A1_dev", ctypes.c_void_p),
                ("dA1_dev", ctypes.c_void_p), ("w1_dev", ctypes.c_void_p), ("n1", ctypes.c_int64),
                ("F_dev", ctypes.c_void_p), ("f_dtype", ctypes.c_int), ("Fs_dev", ctypes.c_void_p),
                ("Fo_dev", ctypes.c_void_p), ("V_dev", ctypes.c_void_p), ("Vo_dev", ctypes.c_void_p),
                ("jastrow", Jastrow),
                ("energy_trace", ctypes.c_void_p), ("D_out", ctypes.c_void_p)]


class ScfObservables(ctypes.Structure):
    _fields_ = [("fock", ctypes.c_void_p), ("eigenvectors", ctypes.c_void_p), ("occupied", ctypes.c_void_p)]


class ScfResult(ctypes.Structure):
    _fields_ = [("energy", ctypes.c_double), ("iterations", ctypes.c_int), ("status", ctypes.c_int),
                ("t_sweep_ms", ctypes.c_double), ("t_solve_ms", ctypes.c_double), ("graph_iterations", ctypes.c_int),
                ("graph_captures", ctypes.c_int)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built: run `python -m tcscf_amd.build` (hipcc, gfx950)")
        # torch ships its own libamdhip64 (SONAME libamdhip64.so.7).  Importing it
        # first lets the loader satisfy this library's libamdhip64.so.7 with the
        # copy already mapped, so the process runs one HIP runtime; loading ours
        # first would map /opt/rocm's copy beside torch's and the two runtimes
        # then fight over the device.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        L.tcscf_last_error.restype = ctypes.c_char_p
        L.tcscf_F_workspace_bytes.restype = ctypes.c_size_t
        L.tcscf_F_workspace_bytes.argtypes = [ctypes.c_int64, ctypes.c_int, ctypes.c_int]
        L.tcscf_qmc_lattice_npts.restype = ctypes.c_int64
        L.tcscf_qmc_lattice_npts.argtypes = [ctypes.c_int64, ctypes.c_int]
        L.tcscf_qmc_lattice_grid.restype = ctypes.c_int64
        L.tcscf_qmc_lattice_grid.argtypes = [ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_uint64,
                                             ctypes.c_void_p, ctypes.c_void_p]
        L.tcscf_sobol_points.argtypes = [ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_void_p]
        L.tcscf_qmc_spherical_points.restype = ctypes.c_int64
        L.tcscf_qmc_spherical_points.argtypes = [ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_uint64,
                                                 ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.tcscf_eri_qmc6d.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                      ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        L.tcscf_eri_qmc5d_atomic.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                             ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                             ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.tcscf_sobol_grid.argtypes = [ctypes.c_int64, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p]
        _lib = L
    return _lib


def last_error():
    return lib().tcscf_last_error().decode()


def check(rc):
    if rc >= 0:
        return rc
    msg = last_error()
    if rc == ENOTCONV:
        raise ConvergenceError(rc, msg)
    if rc == ESOLVER:
        raise SolverError(rc, msg)
    raise TcscfError(rc, msg)
