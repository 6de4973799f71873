"""CPU checks of the C-ABI library: it loads, exports every symbol the header
declares, and its host-side pieces (grids, one-electron matrices) match the
oracle.  No device calls."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT, fns_of


def _declared():
    text = open(os.path.join(ROOT, "include", "tcscf_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tcscf_\w+)\s*\(", text)))


def test_header_and_exports_agree(tc):
    from tcscf_amd import _lib
    declared = _declared()
    assert sorted(_lib.EXPORTS) == declared
    L = tc.lib()
    for name in declared:
        assert getattr(L, name) is not None


def test_library_has_gfx950_code():
    """The in-tree library carries a gfx950 code object (no CPU fallback)."""
    from tcscf_amd import _lib
    blob = open(_lib.LIB_PATH, "rb").read()
    assert b"gfx950" in blob


def test_radial_and_lebedev_match_oracle(tc, O):
    from tcscf_amd import api
    for kind in (0, 1, 2):
        for n in (25, 100):
            r, w = api.radial_grid(n, kind, 1.0, 0.9)
            r2, w2 = O.radial_grid(n, kind, 1.0, 0.9)
            # the oracle is built with FMA contraction, the library without: a few ulp
            assert np.allclose(r, r2, rtol=1e-15, atol=0) and np.allclose(w, w2, rtol=1e-15, atol=0)
    for order in (3, 11, 23, 29, 41):
        a = api.lebedev(order)
        b = O.lebedev(order)
        for x, y in zip(a, b):
            assert np.array_equal(x, y)


def test_product_grid_matches_oracle(tc, O):
    from tcscf_amd import api
    for nrad, order, kind in ((25, 11, 1), (100, 29, 0), (50, 23, 0)):
        p, w = api.product_grid(nrad, order, kind, 1.0, 1.0)
        p2, w2 = O.product_grid(nrad, order, kind, 1.0, 1.0)
        assert np.allclose(p, p2, rtol=1e-15, atol=1e-300)
        assert np.max(np.abs(w - w2) / np.abs(w2)) <= 1e-15


def test_one_electron_matrices_match_oracle(tc, O):
    from tcscf_amd import api, configs
    for cfg in (configs.he_config1(), configs.ne_config2(), configs.be_config3()):
        S, H = api.one_electron_matrices(cfg["basis"], cfg["Z"])
        S2, H2 = O.sto_S_H(fns_of(cfg["basis"]), cfg["Z"])
        assert np.allclose(S, S2, rtol=1e-14, atol=1e-15)
        assert np.allclose(H, H2, rtol=1e-13, atol=1e-13)


def test_error_paths(tc):
    from tcscf_amd import _lib, api
    L = tc.lib()
    assert L.tcscf_lebedev_npts(4) == _lib.EINVAL
    assert "not supported" in _lib.last_error()
    with pytest.raises(ValueError):
        api.SlaterTypeOrbital(1.0, 1, 2, 0)
    bad = (_lib.Sto * 1)()
    bad[0].n, bad[0].l, bad[0].m, bad[0].zeta = 2, 1, 3, 1.0
    S = np.zeros(1)
    H = np.zeros(1)
    dp = ctypes.POINTER(ctypes.c_double)
    assert L.tcscf_sto_overlap_core(bad, 1, 1, S.ctypes.data_as(dp), H.ctypes.data_as(dp)) == _lib.EINVAL


def test_single_hip_runtime_in_process(tc):
    """torch bundles its own libamdhip64; the package loads torch first so this
    library binds to the same runtime (two runtimes in one process do not see
    the device consistently)."""
    paths = {line.split()[-1] for line in open("/proc/self/maps") if "libamdhip64" in line}
    assert len(paths) == 1, paths
