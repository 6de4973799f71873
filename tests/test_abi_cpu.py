"""CPU checks of the C-ABI library: it loads, exports every symbol the header
declares, and its host-side pieces (grids, one-electron matrices) match the
oracle.  No device calls."""
import ctypes
import math
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


# ------------------------------------------------------------------ QMC grids
def test_sobol_points_match_scipy_and_keep_the_net_property(tc):
    """Unscrambled Sobol points (Joe-Kuo directions, Gray-code order) equal
    scipy's bit for bit; scrambling (linear matrix + digital shift) keeps the
    (0, m, 1)-net property of every coordinate and the (0, m, 2)-net property
    of the first two; the seed selects the scramble."""
    from scipy.stats import qmc
    u = tc.sobol_points(4096, 3, scramble=False)
    assert np.array_equal(u, qmc.Sobol(d=3, scramble=False).random(4096))
    m = 10
    us = tc.sobol_points(2 ** m, 3, scramble=True, seed=7)
    assert np.all((us >= 0) & (us < 1))
    for d in range(3):
        assert np.all(np.bincount((us[:, d] * 2 ** m).astype(int), minlength=2 ** m) == 1)
    for a in (3, 5, 7):
        b = m - a
        box = (us[:, 0] * 2 ** a).astype(int) * 2 ** b + (us[:, 1] * 2 ** b).astype(int)
        assert np.all(np.bincount(box, minlength=2 ** m) == 1)
    assert np.array_equal(us, tc.sobol_points(2 ** m, 3, scramble=True, seed=7))
    assert not np.array_equal(us, tc.sobol_points(2 ** m, 3, scramble=True, seed=8))


def test_sobol_importance_grid_integrates(tc):
    """r ~ 2 e^{-2r}, uniform directions: the 1s (zeta = 1) density integrates
    to 1 and e^{-2r} to pi, with QMC accuracy; the weights are the importance
    ratios 4 pi r^2 / (2 n e^{-2r})."""
    n = 2 ** 14
    pts, w = tc.sobol_grid(n, seed=3)
    r = np.sqrt((pts ** 2).sum(0))
    assert np.allclose(w, 4 * math.pi * r * r / (2 * n * np.exp(-2 * r)), rtol=1e-13)
    assert abs((w * np.exp(-2 * r) / math.pi).sum() - 1) < 2e-3
    assert abs((w * np.exp(-2 * r)).sum() - math.pi) < 2e-3 * math.pi
    # odd integrand 2 pi r^2 z per unit density (mean |.| ~ 4.7): zero to QMC accuracy
    assert abs((w * np.exp(-2 * r) * pts[2]).sum()) < 0.05


def test_qmc_lattice_grid(tc):
    """QMCGrid(gridSize) shape (src/integration/integrateAll.hpp:21-65): two
    shifted copies of a prime-size lattice; M4(0, 0.9) x Azimuthal x Linear maps
    with weight r^2 sin(theta) integrate a 1s density to 1 and an odd function
    to 0; the 2-D variant keeps y = 0 and, times 2 pi, integrates axially
    symmetric functions.  Exact point sets of mppmu/qmc are not reproducible
    (parity unpinned); these properties are what is checked."""
    pts, w = tc.qmc_lattice_grid(5000, 3, 2, seed=1)
    assert len(w) == 2 * 5003  # 5003 is the first prime >= 5000
    assert np.all(w > 0)
    r = np.sqrt((pts ** 2).sum(0))
    assert abs((w * np.exp(-2 * r) / math.pi).sum() - 1) < 1e-6
    z2 = 1.7
    assert abs((w * np.exp(-2 * z2 * r) * z2 ** 3 / math.pi).sum() - 1) < 1e-6
    assert abs((w * np.exp(-2 * r) * pts[0]).sum()) < 1e-4  # int |x| e^{-2r} = 3 pi/8
    p2, w2 = tc.qmc_lattice_grid(5000, 2, 2, seed=1)
    assert np.all(p2[1] == 0)
    r2 = np.sqrt((p2 ** 2).sum(0))
    assert abs(2 * math.pi * (w2 * np.exp(-2 * r2) / math.pi).sum() - 1) < 1e-6
    a, _ = tc.qmc_lattice_grid(5000, 3, 2, seed=2)
    assert not np.array_equal(a, pts)


# ------------------------------------------------------ loadSTO and C++ facade
BASIS_DIR = os.path.join(ROOT, "tests", "golden", "basisFunctions")


def test_load_sto(tc):
    """loadSTO (src/SlaterTypeOrbital.cpp:24-76) on copies of the reference's
    ADF basis files: shells expanded to m = -l..l in file order, maxL filter,
    invalid set names rejected, a missing element gives no functions."""
    ne = tc.loadSTO(BASIS_DIR, "TZ3P", "Ne")
    shells = [(1, "S", 12.45), (1, "S", 8.70), (2, "S", 3.65), (2, "S", 2.20), (2, "S", 1.55), (2, "P", 5.50),
              (2, "P", 2.85), (2, "P", 1.45), (3, "D", 4.00), (3, "D", 2.00), (4, "F", 3.00)]
    expect = [(n, "SPDF".index(t), m, a) for n, t, a in shells for m in range(-"SPDF".index(t), "SPDF".index(t) + 1)]
    assert [(b.n, b.l, b.m, b.alpha) for b in ne] == expect
    assert len(tc.loadSTO(BASIS_DIR, "TZ3P", "Ne", maxL=1)) == 14
    assert len(tc.loadSTO(BASIS_DIR, "QZ6P", "Ne")) == 59  # indented lines and ".950" parse
    he = tc.loadSTO(BASIS_DIR, "TZ3P", "He")
    assert [(b.n, b.l) for b in he][:3] == [(1, 0)] * 3 and len(he) == 3 + 6 + 5
    assert tc.loadSTO(BASIS_DIR, "TZ3P", "Xx") == []
    with pytest.raises(tc.TcscfError):
        tc.loadSTO(BASIS_DIR, "DZP", "Ne")


def test_cpp_facade_host_checks(tc):
    """The C++ host facade (include/tcscf.hpp) compiles against the C ABI and
    its host-side pieces behave (loadSTO, STO checks, S/H closed forms, RCS
    electron-count check)."""
    import subprocess
    from tcscf_amd import build
    exe = os.path.join(ROOT, "tcscf_amd", "_build", "facade_test")
    srcs = [os.path.join(ROOT, "tests", "cpp", "facade_test.cpp"), os.path.join(ROOT, "include", "tcscf.hpp")]
    if not os.path.exists(exe) or os.path.getmtime(exe) < max(os.path.getmtime(s) for s in srcs):
        build.build_facade_test()
    r = subprocess.run([exe, "cpu", BASIS_DIR], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failure(s)" in r.stdout


@pytest.mark.parametrize("sanitizer", ["address", "thread"])
def test_batch_scf_host_logic_under_sanitizers(tmp_path, sanitizer):
    """tcscf_batch_scf's host side on the CPU (tests/cpp/batch_host_emu.cpp):
    the real batch driver and the real host halves of the RCS / UOS / TC
    iterations, linked against host stand-ins for the HIP runtime that replay
    each captured graph with a CPU model of the Fock builds, under
    AddressSanitizer and ThreadSanitizer.  Every instance converges, the graph
    is captured again as instances finish, and two runs of one batch agree
    bitwise (energies, iteration counts, traces)."""
    import glob
    import re
    import shutil
    import subprocess
    clang = "/opt/rocm/llvm/bin/clang++"
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not (os.path.exists(clang) and os.path.exists(hipcc)):
        pytest.skip("ROCm clang / hipcc not found")
    csrc = os.path.join(ROOT, "tcscf_amd", "csrc")
    inc = ["-I", csrc, "-I", os.path.join(ROOT, "include")]
    flags = ["-O1", "-g", "-std=c++17", "-fno-omit-frame-pointer"]
    objs = []
    for f in ("scf.hip", "tc_scf.hip"):  # host halves only; kernels are never launched here
        o = str(tmp_path / (f + ".o"))
        subprocess.run([hipcc, "-x", "hip", "--offload-host-only", "-fno-gpu-sanitize", "-Xarch_host",
                        "-fsanitize=" + sanitizer, *flags, *inc, "-c", os.path.join(csrc, f), "-o", o],
                       check=True, capture_output=True)
        objs.append(o)
    for f in (os.path.join(csrc, "batch.cpp"), os.path.join(csrc, "dense.cpp"),
              os.path.join(ROOT, "tests", "cpp", "batch_host_emu.cpp")):
        o = str(tmp_path / (os.path.basename(f) + ".o"))
        subprocess.run([clang, "-fsanitize=" + sanitizer, "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", *flags,
                        *inc, "-c", f, "-o", o], check=True, capture_output=True)
        objs.append(o)
    # the host-only objects still name their (absent) device code objects
    names = set()
    for o in objs[:2]:
        nm = subprocess.run(["nm", "-u", o], check=True, capture_output=True, text=True).stdout
        names |= set(re.findall(r"__hip_fatbin_[0-9a-f]+", nm))
    stub = tmp_path / "fatbin_stubs.cpp"
    stub.write_text("".join('extern "C" char %s[64]; char %s[64];\n' % (n, n) for n in sorted(names)))
    subprocess.run([clang, "-c", str(stub), "-o", str(stub) + ".o"], check=True)
    exe = str(tmp_path / "batch_host_emu")
    subprocess.run([clang, "-no-pie", "-fsanitize=" + sanitizer, "-o", exe, *objs, str(stub) + ".o", "-lpthread"],
                   check=True, capture_output=True)
    r = subprocess.run([exe, "2"], capture_output=True, text=True, timeout=600)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "0 failure(s)" in r.stdout and "Sanitizer" not in out, out[-4000:]
    assert len(glob.glob(str(tmp_path / "*.o"))) == 6


def test_batch_workload_generator_is_pinned(tc):
    """The batched-serving instance list is reproducible from its seed and
    matches the committed fixture; it spans the workload's ranges."""
    import json
    from tcscf_amd import configs
    with open(os.path.join(ROOT, "tests", "golden", "batch64_instances.json")) as f:
        fixture = json.load(f)
    drawn = json.loads(json.dumps(configs.batch_instances()))
    assert drawn == fixture
    assert len(drawn) == 64
    assert {x["Z"] for x in drawn} <= set(range(2, 19))
    assert min(x["npts"] for x in drawn) == 1250 and max(x["npts"] for x in drawn) == 30200
    assert sum(x["method"] == "tc" for x in drawn) == 32
    assert all(x["nb"] == len(configs.batch_basis(x)) and x["nb"] >= x["nup"] + 2 for x in drawn)
    assert all(0.5 <= x["b"] <= 2 for x in drawn if x["method"] == "tc")


def _one_electron_numeric(O, fns, Z, pts, w):
    """S_ab = sum w phi_a phi_b and H_ab = sum w (1/2 grad phi_a . grad phi_b
    - Z phi_a phi_b / r) from the oracle's values and gradients."""
    v, g = O.atomic_real(pts, fns, grads=True)
    r = np.sqrt((pts ** 2).sum(0))
    S = (v * w[:, None]).T @ v
    T = 0.5 * np.einsum("n,nac,nbc->ab", w, g, g)
    V = -Z * (v * (w / r)[:, None]).T @ v
    return S, T + V


def test_ochi_and_hydrogen_bases_one_electron(tc, O):
    """Ochi functions (src/OchiBasis.hpp) and hydrogen-like functions
    (src/HydrogenLikeBasis.hpp): the oracle's exact STO expansion against
    scipy's Laguerre polynomials; orthonormality on a fine grid; the hydrogen-like
    core matrix -Z^2/(2 n^2) as the expectation value of -1/2 lap - Z/r; and the
    library's closed-form Ochi H (the reference's coefficient times the exact
    Laguerre integral) against the numerically integrated <a| -1/2 lap - Z/r |b>."""
    from scipy.special import eval_genlaguerre
    pts, w = tc.product_grid(240, 17, tc.RADIAL_M4, 1.0, 1.0)
    r = np.sqrt((pts ** 2).sum(0))
    # s functions: the expansion against the closed form (Y_00 = 1/sqrt(4 pi))
    for kind, n, alpha in ((1, 0, 1.3), (1, 3, 1.3), (1, 6, 0.9), (2, 1, 2.0), (2, 3, 2.0), (2, 4, 3.0)):
        v = O.atomic_real(pts, [(kind, n, 0, 0, alpha)])[:, 0]
        if kind == 1:
            N = (2 * alpha) ** 1.5 / math.sqrt(math.factorial(n + 2) / math.factorial(n))
            ref = N * eval_genlaguerre(n, 2, 2 * alpha * r) * np.exp(-alpha * r)
        else:
            b = alpha / n
            N = -(2 * b) ** 1.5 / math.sqrt(2 * n * math.factorial(n) / math.factorial(n - 1))
            ref = N * eval_genlaguerre(n - 1, 1, 2 * b * r) * np.exp(-b * r)
        assert np.max(np.abs(v - ref / math.sqrt(4 * math.pi))) < 1e-12 * max(1.0, np.max(np.abs(ref)))
    # Ochi: s and p functions of one alpha are orthonormal; H from the library vs quadrature
    ochi = [tc.OchiBasisFunction(1.4, n, l, m) for n in range(4) for l in (0, 1) for m in range(-l, l + 1)]
    fns = [(1, b.n, b.l, b.m, b.alpha) for b in ochi]
    Sn, Hn = _one_electron_numeric(O, fns, 2, pts, w)
    S, H = tc.one_electron_matrices(ochi, 2)
    assert np.max(np.abs(Sn - np.eye(len(ochi)))) < 1e-9
    assert np.array_equal(S, np.eye(len(ochi)))
    assert np.max(np.abs(H - Hn)) < 1e-8
    # hydrogen-like functions of the atom's Z: H = -Z^2 / (2 n^2) on the diagonal
    hyd = [tc.HydrogenLikeBasisFunction(3, n, l, 0) for n in (1, 2, 3) for l in range(n) if l <= 1]
    fns = [(2, b.n, b.l, b.m, b.alpha) for b in hyd]
    Sn, Hn = _one_electron_numeric(O, fns, 3, pts, w)
    S, H = tc.one_electron_matrices(hyd, 3)
    assert np.max(np.abs(Sn - np.eye(len(hyd)))) < 1e-9
    assert np.max(np.abs(Hn - H)) < 1e-8
    assert np.allclose(np.diag(H), [-9 / (2 * b.n ** 2) for b in hyd], rtol=0, atol=1e-15)


def test_ochi_jastrow_host_logic(tc):
    """OchiJastrow (src/jastrowFunctions.hpp:9-115 parameters): the C struct it
    passes, the one-V TC scales (c_oppo = rho c_same term by term, else an
    error) and argument checks."""
    u = tc.OchiJastrow(1.2, 0.9, [(0.5, 0.25), (-0.2, -0.1), (0.0, 0.0)], [(1, 0, 0), (2, 1, 0), (0, 3, 2)])
    s = u._struct()
    assert (s.a, s.a12, s.nterms) == (1.2, 0.9, 3)
    assert list(s.i)[:3] == [1, 2, 0] and list(s.j)[:3] == [0, 1, 3] and list(s.k)[:3] == [0, 0, 2]
    assert list(s.c_oppo)[:3] == [0.5, -0.2, 0.0] and list(s.c_same)[:3] == [0.25, -0.1, 0.0]
    assert u.spin_ratio() == 2.0
    t = u._tc_struct()
    assert (t.a12, t.c_same, t.c_oppo) == (0.9, 1.0, 2.0)
    for c in ([(0.5, 0.25), (0.3, 0.1)], [(0.5, 0.25), (0.1, 0.0)], [(0.5, 0.0), (0.1, 0.0)]):
        v = tc.OchiJastrow(1.0, 1.0, c, [(1, 0, 0), (0, 1, 0)])
        assert v.spin_ratio() is None
        with pytest.raises(ValueError):
            v._tc_struct()
    j = tc.Jastrow.from_b(2.0)
    w = tc.OchiJastrow.from_jastrow(j)
    assert w.S.tolist() == [[1, 0, 0]] and w.c.tolist() == [[j.c_oppo, j.c_same]] and w.spin_ratio() == 2.0
    for bad in (dict(a=0.0), dict(S=[(1, 0, 17)]), dict(c=[(1, 1)] * 9, S=[(1, 0, 0)] * 9)):
        kw = dict(a=1.0, a12=1.0, c=[(1, 1)], S=[(1, 0, 0)])
        kw.update(bad)
        with pytest.raises(ValueError):
            tc.OchiJastrow(**kw)


def test_tc_launch_info_table(tc):
    """The TC point kernel's launch choice (no GPU call): V rows staged in LDS
    with 12 waves per workgroup for the restricted FP32 Ne build (nb 30,
    5 + 5 electrons), 8 unstaged waves once the rows do not fit (FP32 two spins
    at nb 44), the independent-V build staging both tensors at nb 30, at
    least one wave for every nb <= 64 and an error past it."""
    import ctypes
    from tcscf_amd import _lib
    def info(nup, ndown, nb, f32, gv):
        out = (ctypes.c_int * 6)()
        _lib.check(tc.lib().tcscf_tc_launch_info(nup, ndown, nb, _lib.F32 if f32 else _lib.F64, gv, out))
        return list(out)
    assert info(5, 5, 30, True, 0)[3:5] == [1, 12]
    assert info(5, 5, 30, True, 0)[0:2] == [1, 8]
    assert info(6, 5, 44, True, 0)[0:2] == [0, 8]
    assert info(5, 5, 30, True, 1)[0:2] == [1, 8]
    assert all(x[2] <= 160 * 1024 and x[5] <= 160 * 1024 for x in [info(9, 9, 60, True, 1), info(9, 9, 60, False, 0)])
    assert info(60, 60, 60, False, 0)[1] >= 1
    with pytest.raises(tc.TcscfError):
        info(5, 5, 65, False, 0)
