"""GPU parity: the HIP path (through the C ABI) against the FP64 oracle on the
same seeded inputs.  Tolerances are written per test; FP64 paths are held to
~1e-13 relative, SCF energies to 1e-10 Eh."""
import math
import os

import numpy as np
import pytest

from conftest import fns_of, rel_err

pytestmark = pytest.mark.gpu


def _grid(tc, nrad, order, kind):
    return tc.QMCGrid.radial_lebedev(nrad, order, kind, 1.0, 1.0)


def _mixed_basis():
    from tcscf_amd import SlaterTypeOrbital as STO
    b = [STO(0.9, 1, 0, 0), STO(2.3, 1, 0, 0), STO(1.1, 2, 0, 0)]
    b += [STO(1.2, 2, 1, m) for m in (-1, 0, 1)] + [STO(0.6, 3, 1, m) for m in (-1, 0, 1)]
    b += [STO(1.4, 3, 2, m) for m in range(-2, 3)] + [STO(0.9, 4, 3, m) for m in range(-3, 4)]
    return b


# Output layouts the contraction kernels assume: register i of lane 16 g + r
# holds D[row(g, i)][r].
D_ROW = {0: lambda g, i: g + 4 * i, 1: lambda g, i: 4 * g + i, 2: lambda g, i: 4 * g + i}


def test_mfma_fragment_layouts(tc, gpu):
    """Pins the lane layouts of the three MFMA instructions the contraction
    kernels use, against numpy matmul (raw accumulator registers)."""
    torch = gpu
    import ctypes
    rng = np.random.default_rng(0)
    bad = []
    for which, K in ((0, 4), (1, 4), (2, 32)):
        A = rng.integers(-8, 8, size=(16, K)).astype(np.float64)
        B = rng.integers(-8, 8, size=(K, 16)).astype(np.float64)
        dA = torch.from_numpy(A).cuda()
        dB = torch.from_numpy(B).cuda()
        dD = torch.zeros(256, dtype=torch.float64, device="cuda")
        rc = tc.lib().tcscf_selftest_mfma(which, ctypes.c_void_p(dA.data_ptr()), ctypes.c_void_p(dB.data_ptr()),
                                          ctypes.c_void_p(dD.data_ptr()), None)
        assert rc == 0
        torch.cuda.synchronize()
        raw = dD.cpu().numpy().reshape(64, 4)
        C = A @ B
        want = np.array([[C[D_ROW[which](l // 16, i)][l % 16] for i in range(4)] for l in range(64)])
        if not np.array_equal(raw, want):
            found = [name for name, f in (("4g+i", lambda g, i: 4 * g + i), ("g+4i", lambda g, i: g + 4 * i))
                     if np.array_equal(raw, [[C[f(l // 16, i)][l % 16] for i in range(4)] for l in range(64)])]
            bad.append((which, found))
    assert not bad, bad


def test_basis_values_and_gradients(tc, O, gpu):
    basis = _mixed_basis()
    rng = np.random.default_rng(1)
    pts = rng.uniform(-3, 3, (3, 500))
    pts[:, :4] = [[0, 0, 0, 0], [0, 0, 0, 0], [0.5, -0.5, 2.0, -1.5]]  # z axis
    g = tc.QMCGrid(pts, np.ones(500))
    g.setBasisFunctions(basis, True)
    v = g.basisValues.cpu().numpy()
    d = g.basisGradients.cpu().numpy()
    v2, d2 = O.sto_real(pts, fns_of(basis), grads=True)
    assert rel_err(v, v2) < 1e-13
    assert np.max(np.abs(d - d2)) < 1e-12 * np.max(np.abs(d2))
    # also on a Lebedev product grid (poles included)
    g = _grid(tc, 30, 17, tc.RADIAL_M4)
    g.setBasisFunctions(basis, True)
    v2, d2 = O.sto_real(g.points_host, fns_of(basis), grads=True)
    assert rel_err(g.basisValues.cpu().numpy(), v2) < 1e-13
    assert np.max(np.abs(g.basisGradients.cpu().numpy() - d2)) < 1e-11 * np.max(np.abs(d2))


def _F_oracle(O, g1, g2, kernel=0, prm=None, same=True, scale=1.0, rows=None):
    A2 = g2.basisValues.cpu().numpy()
    return O.compute_F(g1.points_host, g2.points_host, g2.weights_host, A2, kernel=kernel, prm=prm, same=same,
                       scale=scale, rows=rows)


def test_F_fp64_he_config1(tc, O, gpu):
    from tcscf_amd import configs
    c = configs.he_config1()
    g = _grid(tc, *c["grid"])
    g.setBasisFunctions(c["basis"], False)
    F = tc.computeF(g, g)
    ref = _F_oracle(O, g, g)
    assert rel_err(tc.unpack(F.cpu().numpy(), 5), ref) < 1e-13


def test_F_fp64_mixed_l_two_grids(tc, O, gpu):
    """Different r1 and r2 grids, s..f functions, row count not a multiple of
    any tile size."""
    basis = _mixed_basis()
    g1 = _grid(tc, 13, 11, tc.RADIAL_M4)  # 650 points
    g2 = _grid(tc, 17, 7, tc.RADIAL_BECKE)  # 442 points
    g1.setBasisFunctions(basis, False)
    g2.setBasisFunctions(basis, False)
    F = tc.computeF(g1, g2)
    ref = _F_oracle(O, g1, g2)
    assert rel_err(tc.unpack(F.cpu().numpy(), len(basis)), ref) < 1e-13


def test_F_fp64_2d_grid_scale(tc, O, gpu):
    """The reference's 2-D r2 grid (points in the xz-plane, F x 2 pi)."""
    from tcscf_amd import SlaterTypeOrbital as STO
    basis = [STO(0.8, 1, 0, 0), STO(1.7, 1, 0, 0), STO(1.1, 2, 0, 0)]
    rng = np.random.default_rng(3)
    p2 = rng.uniform(0, 3, (3, 300))
    p2[1] = 0
    g1 = _grid(tc, 10, 11, tc.RADIAL_M4)
    g2 = tc.QMCGrid(p2, rng.uniform(0.1, 1, 300), ndim=2)
    g1.setBasisFunctions(basis, False)
    g2.setBasisFunctions(basis, False)
    F = tc.computeF(g1, g2)
    ref = _F_oracle(O, g1, g2, scale=2 * math.pi)
    assert rel_err(tc.unpack(F.cpu().numpy(), 3), ref) < 1e-13


def test_F_fp64_tc_kernels(tc, O, gpu):
    """TC scalar kernel (same and opposite spin) and the Jastrow-gradient
    vector kernel against the oracle's general Ochi evaluation."""
    basis = _mixed_basis()[:9]
    g1 = _grid(tc, 11, 11, tc.RADIAL_M4)
    g3 = _grid(tc, 9, 7, tc.RADIAL_M4)
    for g in (g1, g3):
        g.setBasisFunctions(basis, True)
    jas = tc.Jastrow.from_b(1.0)  # a12 = 1, c_oppo = 1/2, c_same = 1/4
    for same in (True, False):
        prm = O.ochi_params(1.0, jas.a12, [(1, 0, 0, jas.c_oppo, jas.c_same)])
        F = tc.computeF(g1, g1, kernel="tc", jastrow=jas, same_spin=same)
        ref = _F_oracle(O, g1, g1, kernel=O.K_TC_SCALAR, prm=prm, same=same)
        assert rel_err(tc.unpack(F.cpu().numpy(), 9), ref) < 1e-12
    V = tc.computeV(g1, g3, jas).cpu().numpy()  # (n1, 3, npair), c = 1
    prm = O.ochi_params(1.0, jas.a12, [(1, 0, 0, 1.0, 1.0)])
    ref = _F_oracle(O, g1, g3, kernel=O.K_JASTROW_GRAD, prm=prm)  # (n1, nb, nb, 3)
    full = tc.unpack(V, 9).transpose(0, 2, 3, 1)
    assert rel_err(full, ref) < 1e-12


def test_F_tc_pair_vs_oracle(tc, O, gpu):
    """computeF_tc_pair (F_C and the smooth-part contraction G combined) at
    FP64 against the oracle's TC scalar F from the general Ochi gradients, for
    two Jastrow parameter sets; and the mixed form (BF16x5 F_C, BF16x3 G,
    FP32 storage) against the FP64 pair (tolerance in the assertion)."""
    basis = _mixed_basis()[:9]
    g1 = _grid(tc, 11, 11, tc.RADIAL_M4)
    g1.setBasisFunctions(basis, True)
    for b in (1.0, 0.6):
        jas = tc.Jastrow.from_b(b)
        Fs, Fo = tc.computeF_tc_pair(g1, g1, jas)
        for same, F in ((True, Fs), (False, Fo)):
            prm = O.ochi_params(1.0, jas.a12, [(1, 0, 0, jas.c_oppo, jas.c_same)])
            ref = _F_oracle(O, g1, g1, kernel=O.K_TC_SCALAR, prm=prm, same=same)
            assert rel_err(tc.unpack(F.cpu().numpy(), 9), ref) < 1e-12
        for prec in ("bf16x5", "bf16x4"):
            Ms, Mo = tc.computeF_tc_pair(g1, g1, jas, precision=prec, smooth_precision="bf16x3", out_dtype="fp32")
            assert Ms.dtype == gpu.float32
            for M, F in ((Ms, Fs), (Mo, Fo)):
                assert rel_err(M.double().cpu().numpy(), F.cpu().numpy()) < 3e-6, prec


@pytest.mark.parametrize("precision,tol", [("fp32", 3e-6), ("bf16x3", 3e-5), ("bf16x4", 3e-6), ("bf16x5", 3e-6)])
def test_F_mixed_precision_error(tc, gpu, precision, tol):
    """Mixed-precision F against the FP64 drop-in on the Ne basis (nb = 30):
    max element error relative to max |F| (tolerance in the parameters)."""
    from tcscf_amd import configs
    c = configs.ne_config2()
    g = _grid(tc, 40, 17, tc.RADIAL_M4)
    g.setBasisFunctions(c["basis"], False)
    F64 = tc.computeF(g, g, precision="fp64").cpu().numpy()
    F = tc.computeF(g, g, precision=precision).cpu().numpy()
    err = np.max(np.abs(F - F64)) / np.max(np.abs(F64))
    assert err < tol, err


def test_F_split_last_round_matches_unsplit(tc, gpu, monkeypatch):
    """The pipelined BF16 kernel splits the last, partly filled round of its
    tiles along r2 (one run of steps per resident workgroup, pieces added by
    a fix-up pass).  On the full Ne grid (944 tiles of 64 rows: 3 full rounds
    and 176 split tiles at 256 resident workgroups) the results equal the
    unsplit launch up to the FP64 association of the pieces' sums: the split
    pieces start on even steps, so the two-step FP32 partials of BF16x4 pair
    the steps exactly as the unsplit tiles do.  Tolerances in the asserts:
    1e-13 of max |F| for FP64 storage, one FP32 rounding for FP32 storage;
    and the RHF energy to 1e-10 Eh.  The plain tile (V, the TC scalar
    kernel) splits its last round the same way."""
    import torch
    from tcscf_amd import configs
    c = configs.ne_config2()
    nrad, order, kind = c["grid"]
    g = tc.QMCGrid.radial_lebedev(nrad, order, kind, 1.0, 1.0)
    g.setBasisFunctions(c["basis"], False)
    jas = tc.Jastrow.from_b(1.0)
    cases = [("bf16x4", "fp64"), ("bf16x5", "fp64"), ("bf16x4", "fp32")]
    for prec, odt in cases:
        monkeypatch.setenv("TCSCF_BF16_SPLIT", "0")
        a = tc.computeF(g, g, precision=prec, out_dtype=odt)
        monkeypatch.setenv("TCSCF_BF16_SPLIT", "1")
        b = tc.computeF(g, g, precision=prec, out_dtype=odt)
        torch.cuda.synchronize()
        scale = a.double().abs().max().item()
        d = (a.double() - b.double()).abs().max().item() / scale
        print(f"{prec} {odt}: max diff {d:.2e} of max |F|, {(a != b).sum().item()} elements differ")
        assert d < (1e-13 if odt == "fp64" else 6e-8), (prec, odt, d)
    monkeypatch.setenv("TCSCF_BF16_SPLIT", "0")
    Pa = tc.computeF_tc_pair(g, g, jas, precision="bf16x4", smooth_precision="bf16x3")
    monkeypatch.setenv("TCSCF_BF16_SPLIT", "1")
    Pb = tc.computeF_tc_pair(g, g, jas, precision="bf16x4", smooth_precision="bf16x3")
    for x, y in zip(Pa, Pb):
        assert ((x - y).abs().max() / x.abs().max()).item() < 1e-13
    # the plain tile: V (three components, 16-row tiles over the 9,700-point
    # r3 grid of the TC workload) and the TC scalar kernel
    g.setBasisFunctions(c["basis"], True)
    g3 = tc.QMCGrid.radial_lebedev(50, 23, kind, 1.0, 1.0)
    g3.setBasisFunctions(c["basis"], True)
    for fn in (lambda: tc.computeV(g, g3, jas, precision="bf16x3"),
               lambda: tc.computeF(g, g3, kernel="tc", jastrow=jas, precision="bf16x4")):
        monkeypatch.setenv("TCSCF_BF16_SPLIT", "0")
        x = fn()
        monkeypatch.setenv("TCSCF_BF16_SPLIT", "1")
        y = fn()
        assert ((x - y).abs().max() / x.abs().max()).item() < 1e-13
    S, H = tc.one_electron_matrices(c["basis"], c["Z"])
    e = []
    for sp in ("0", "1"):
        monkeypatch.setenv("TCSCF_BF16_SPLIT", sp)
        F = tc.computeF(g, g, precision="bf16x4")
        e.append(tc.RCSHartreeFock(10, 30).compute(S, H, g, F, True, tol=1e-10, max_iter=100))
    assert abs(e[0] - e[1]) < 1e-10, e


def test_F_summation_order_is_a_permutation(tc, gpu):
    """Summing r2 in radius order or in grid order gives the same FP64 F."""
    from tcscf_amd import configs
    c = configs.ne_config2()
    g = _grid(tc, 20, 17, tc.RADIAL_M4)
    g.setBasisFunctions(c["basis"], False)
    a = tc.computeF(g, g, ordered=True).cpu().numpy()
    b = tc.computeF(g, g, ordered=False).cpu().numpy()
    assert rel_err(a, b) < 1e-13


def test_mixed_precision_energy_ne(tc, gpu):
    """Ne RHF on a 50 x 194-point grid: the BF16x5, BF16x4 and FP32 energies
    lie within 1e-6 Eh of the FP64 drop-in (BF16x3 is reported, not held)."""
    from tcscf_amd import configs
    c = configs.ne_config2()
    g, S, H = _scf_inputs(tc, c, (50, 23, tc.RADIAL_M4))
    e = {}
    for prec in ("fp64", "fp32", "bf16x5", "bf16x4", "bf16x3"):
        F = tc.computeF(g, g, precision=prec)
        e[prec] = tc.RCSHartreeFock(10, 30).compute(S, H, g, F, True, tol=1e-11, max_iter=100)
    print({k: v - e["fp64"] for k, v in e.items()})
    assert abs(e["bf16x5"] - e["fp64"]) < 1e-6
    assert abs(e["bf16x4"] - e["fp64"]) < 1e-6
    assert abs(e["fp32"] - e["fp64"]) < 1e-6


# ----------------------------------------------------------------- SCF
def _scf_inputs(tc, cfg, grid=None):
    g = _grid(tc, *(grid or cfg["grid"]))
    g.setBasisFunctions(cfg["basis"], False)
    S, H = tc.one_electron_matrices(cfg["basis"], cfg["Z"])
    return g, S, H


def test_rcs_uos_he_config1_vs_oracle(tc, O, gpu):
    from tcscf_amd import configs
    c = configs.he_config1()
    g, S, H = _scf_inputs(tc, c)
    F = tc.computeF(g, g)
    A = g.basisValues.cpu().numpy()
    w = g.weights_host
    Fo = _F_oracle(O, g, g)
    for mask in (True, False):
        hf = tc.RCSHartreeFock(2, 5)
        e = hf.compute(S, H, g, F, mask, tol=1e-10)
        ref = O.rcs_scf(2, S, H, A, w, Fo, tol=1e-10, mask=mask)
        assert abs(e - ref["energy"]) < 1e-10
        assert hf.numberOfConvergenceLoops() == ref["iterations"]
        assert np.max(np.abs(hf.energy_trace - ref["trace"])) < 1e-10
        uhf = tc.UOSHartreeFock(1, 1, 5)
        eu = uhf.compute(S, H, g, F, mask, tol=1e-10)
        refu = O.uos_scf(1, 1, S, H, A, w, Fo, tol=1e-10, mask=mask)
        assert abs(eu - refu["energy"]) < 1e-10
        assert abs(eu - e) < 1e-10


def test_rcs_ne_small_grid_end_to_end(tc, O, gpu):
    """Ne basis (30 STOs, s/p/d) on a 40 x 302-point grid: GPU basis + F + SCF
    against the all-oracle path."""
    from tcscf_amd import configs
    c = configs.ne_config2()
    g, S, H = _scf_inputs(tc, c, (40, 29, tc.RADIAL_M4))
    F = tc.computeF(g, g)
    A = O.sto_real(g.points_host, fns_of(c["basis"]))
    Fo = O.compute_F(g.points_host, g.points_host, g.weights_host, A)
    hf = tc.RCSHartreeFock(10, 30)
    e = hf.compute(S, H, g, F, True, tol=1e-10)
    ref = O.rcs_scf(10, S, H, A, g.weights_host, Fo, tol=1e-10)
    assert ref["status"] == 0
    assert abs(e - ref["energy"]) < 1e-10, (e, ref["energy"])
    # Pulay DIIS (an option the reference does not have) reaches the same fixed
    # point, in fewer iterations, for RCS and UOS
    hd = tc.RCSHartreeFock(10, 30)
    ed = hd.compute(S, H, g, F, True, tol=1e-10, diis=6)
    assert abs(ed - ref["energy"]) < 1e-9, (ed, ref["energy"])
    assert hd.numberOfConvergenceLoops() < hf.numberOfConvergenceLoops()
    eu = tc.UOSHartreeFock(5, 5, 30).compute(S, H, g, F, True, tol=1e-10, diis=6)
    assert abs(eu - ref["energy"]) < 1e-9
    # observables (src/HartreeFock.hpp:286-288, :342-344): the eigenvalues are
    # the oracle's; the eigenvectors are S-orthonormal, ordered by eigenvalue and
    # give the density; the Fock operator is the one the energy was read from
    # (RCS E = sum D (H + F), the reference's 2 sum D (H + G/2))
    assert np.max(np.abs(hf.eigenvalues - ref["evals"])) < 1e-8
    Cv = hf.eigenvectors[0]
    assert np.max(np.abs(Cv.T @ S @ Cv - np.eye(30))) < 1e-10
    assert np.max(np.abs(Cv[:, :5] @ Cv[:, :5].T - hf.density[0])) < 1e-12
    assert abs(np.sum(hf.density[0].T * (H + hf.fockOperator[0])) - e) < 1e-10 * abs(e)
    assert np.allclose(hf.fockOperator[0], hf.fockOperator[0].T, atol=1e-12)
    hu = tc.UOSHartreeFock(5, 5, 30)
    eu = hu.compute(S, H, g, F, True, tol=1e-10)
    for s_ in range(2):
        Cv = hu.eigenvectors[s_]
        assert np.max(np.abs(Cv.T @ S @ Cv - np.eye(30))) < 1e-10
        assert np.max(np.abs(Cv[:, :5] @ Cv[:, :5].T - hu.density[s_])) < 1e-12
    assert abs(0.5 * sum(np.sum(hu.density[s_].T * (H + hu.fockOperator[s_])) for s_ in range(2)) - eu) < 1e-10 * abs(eu)


def test_rcs_ne_config2_full_size(tc, O, gpu):
    """Ne config 2 at full size (30,200 points): F rows checked against the
    oracle on a contiguous slice; the SCF on the device F checked against the
    oracle SCF on the same F."""
    from tcscf_amd import configs
    c = configs.ne_config2()
    g, S, H = _scf_inputs(tc, c)
    F = tc.computeF(g, g)
    Fh = F.cpu().numpy()
    rows = (15000, 15064)
    ref_rows = _F_oracle(O, g, g, rows=rows)
    assert rel_err(tc.unpack(Fh[rows[0]:rows[1]], 30), ref_rows) < 1e-13
    hf = tc.RCSHartreeFock(10, 30)
    e = hf.compute(S, H, g, F, True, tol=1e-10)
    ref = O.rcs_scf(10, S, H, g.basisValues.cpu().numpy(), g.weights_host, tc.unpack(Fh, 30), tol=1e-10)
    assert abs(e - ref["energy"]) < 1e-10, (e, ref["energy"])


@pytest.mark.parametrize("shells,Z,nup,ndown", [({0: 10, 1: 6, 2: 2}, 10, 5, 5), ({0: 8, 1: 3}, 3, 2, 1)])
def test_scf_wide_basis_and_open_shell(tc, O, gpu, shells, Z, nup, ndown):
    """nb = 38 (> 32: one r1 point per wave in the sweep) RCS + UOS, and an
    open-shell UOS (Li, 2 up / 1 down) against the oracle."""
    basis = tc.even_tempered(shells, 0.4, 2.1)
    g = _grid(tc, 30, 17, tc.RADIAL_M4)
    g.setBasisFunctions(basis, False)
    S, H = tc.one_electron_matrices(basis, Z)
    F = tc.computeF(g, g)
    A = g.basisValues.cpu().numpy()
    Fo = tc.unpack(F.cpu().numpy(), len(basis))
    if nup == ndown:
        e = tc.RCSHartreeFock(nup + ndown, len(basis)).compute(S, H, g, F, True, tol=1e-10)
        ref = O.rcs_scf(nup + ndown, S, H, A, g.weights_host, Fo, tol=1e-10)
        assert abs(e - ref["energy"]) < 1e-10
    u = tc.UOSHartreeFock(nup, ndown, len(basis))
    eu = u.compute(S, H, g, F, True, tol=1e-10, max_iter=200)
    refu = O.uos_scf(nup, ndown, S, H, A, g.weights_host, Fo, tol=1e-10, max_iter=200)
    assert refu["status"] == 0
    assert abs(eu - refu["energy"]) < 1e-10
    assert np.max(np.abs(u.density - refu["D"])) < 1e-8


@pytest.mark.parametrize("shells,nb", [({0: 4, 1: 4}, 16), ({0: 5, 1: 4}, 17), ({0: 9, 1: 6, 2: 1}, 32)])
@pytest.mark.parametrize("dtype", ["fp64", "fp32"])
def test_scf_sweep_lane_groups_vs_oracle(tc, O, gpu, shells, nb, dtype):
    """The Fock-build sweep at the edges of its lane groups (nb = 16: four
    16-lane slices per column; 17 and 32: two 32-lane slices), F stored in FP64
    and in FP32: closed-shell Be and open-shell B (3 up / 2 down) against the
    oracle on the same F values, 1e-10 Eh over the energy traces."""
    basis = tc.even_tempered(shells, 0.4, 2.1)
    assert len(basis) == nb
    g = _grid(tc, 24, 11, tc.RADIAL_M4)
    g.setBasisFunctions(basis, False)
    A = g.basisValues.cpu().numpy()
    F = tc.computeF(g, g)
    if dtype == "fp32":
        F = F.float()
    Fo = tc.unpack(F.double().cpu().numpy(), nb)
    S, H = tc.one_electron_matrices(basis, 4)
    hf = tc.RCSHartreeFock(4, nb)
    e = hf.compute(S, H, g, F, True, tol=1e-10, max_iter=200)
    ref = O.rcs_scf(4, S, H, A, g.weights_host, Fo, tol=1e-10, max_iter=200)
    assert ref["status"] == 0
    assert len(hf.energy_trace) == len(ref["trace"])
    assert np.max(np.abs(hf.energy_trace - ref["trace"])) < 1e-10
    # open shell: the first 15 iterations of the plain loop (tol = 0 never
    # converges), traces compared step by step
    S, H = tc.one_electron_matrices(basis, 5)
    u = tc.UOSHartreeFock(3, 2, nb)
    with pytest.raises(tc.ConvergenceError):
        u.compute(S, H, g, F, True, tol=0.0, max_iter=15)
    refu = O.uos_scf(3, 2, S, H, A, g.weights_host, Fo, tol=0.0, max_iter=15)
    assert len(u.energy_trace) == len(refu["trace"]) == 15
    assert np.max(np.abs(u.energy_trace - refu["trace"])) < 1e-10


def test_scf_error_paths(tc, gpu):
    from tcscf_amd import configs
    c = configs.he_config1()
    g, S, H = _scf_inputs(tc, c)
    F = tc.computeF(g, g)
    with pytest.raises(tc.ConvergenceError):
        tc.UOSHartreeFock(1, 1, 5).compute(S, H, g, F, True, max_iter=1)
    basis = c["basis"] + [c["basis"][0]]
    g2, S2, H2 = _scf_inputs(tc, dict(c, basis=basis))
    F2 = tc.computeF(g2, g2)
    with pytest.raises(tc.SolverError):
        tc.RCSHartreeFock(2, 6).compute(S2, H2, g2, F2, True)
    # a non-finite F reaches the Fock matrix from the first density on: a
    # solver error, for RCS and for TC (whose eigenvalue ordering needs finite
    # values; nb = 19 takes the sort past its insertion-sort cutoff)
    Fn = F.clone()
    Fn[3, 2] = float("nan")
    with pytest.raises(tc.SolverError):
        tc.RCSHartreeFock(2, 5).compute(S, H, g, Fn, True, tol=1e-10)
    basis = tc.even_tempered({0: 7, 1: 4}, 0.5, 2.0)
    g1, g3, jas, Fs, Fo, V = _tc_case(tc, basis, (12, 11, tc.RADIAL_M4), (8, 7, tc.RADIAL_M4))
    St, Ht = tc.one_electron_matrices(basis, 5)
    Vn = V.clone()
    Vn[5, 1, 7] = float("inf")
    with pytest.raises(tc.SolverError):
        tc.TCHartreeFock(3, 2, len(basis)).compute(St, Ht, g1, Fs, Fo, Vn, jas, True, tol=1e-10, mgs_mode=1)


# ----------------------------------------------------------------- TC
def _tc_case(tc, basis, grid1, grid3, b=1.0):
    g1 = _grid(tc, *grid1)
    g3 = _grid(tc, *grid3)
    for g in (g1, g3):
        g.setBasisFunctions(basis, True)
    jas = tc.Jastrow.from_b(b)
    Fs = tc.computeF(g1, g1, kernel="tc", jastrow=jas, same_spin=True)
    Fo = tc.computeF(g1, g1, kernel="tc", jastrow=jas, same_spin=False)
    V = tc.computeV(g1, g3, jas)
    return g1, g3, jas, Fs, Fo, V


@pytest.mark.parametrize("Z,nup,ndown,mgs_mode", [(4, 2, 2, 0), (4, 2, 2, 1), (3, 2, 1, 1), (2, 1, 1, 0)])
def test_tc_scf_vs_oracle(tc, O, gpu, Z, nup, ndown, mgs_mode):
    """TCHartreeFock on the device Fock build against the oracle's restatement,
    fed the same F_same / F_oppo / V (full per-iteration energy trace)."""
    basis = tc.even_tempered({0: 5, 1: 2}, 0.6, 2.2)
    nb = len(basis)
    g1, g3, jas, Fs, Fo, V = _tc_case(tc, basis, (14, 11, tc.RADIAL_M4), (10, 7, tc.RADIAL_M4))
    S, H = tc.one_electron_matrices(basis, Z)
    hf = tc.TCHartreeFock(nup, ndown, nb)
    e = hf.compute(S, H, g1, Fs, Fo, V, jas, True, tol=1e-10, max_iter=200, mgs_mode=mgs_mode)
    A = g1.basisValues.cpu().numpy()
    dA = g1.basisGradients.cpu().numpy()
    Vf = tc.unpack(V.cpu().numpy(), nb).transpose(0, 2, 3, 1)  # (n1, nb, nb, 3), c = 1
    ref = O.tc_scf(nup, ndown, S, H, A, dA, g1.weights_host, tc.unpack(Fs.cpu().numpy(), nb),
                   tc.unpack(Fo.cpu().numpy(), nb), np.ascontiguousarray(jas.c_same * Vf),
                   np.ascontiguousarray(jas.c_oppo * Vf), tol=1e-10, max_iter=200, mgs_mode=mgs_mode)
    assert ref["status"] == 0
    n = min(len(hf.energy_trace), len(ref["trace"]))
    assert np.max(np.abs(hf.energy_trace[:n] - ref["trace"][:n])) < 1e-9
    assert abs(e - ref["energy"]) < 1e-9, (e, ref["energy"])
    assert np.max(np.abs(hf.density - ref["D"])) < 1e-7
    if mgs_mode == 1:
        # DIIS on the right-eigenvector error reaches the same TC fixed point
        hd = tc.TCHartreeFock(nup, ndown, nb)
        ed = hd.compute(S, H, g1, Fs, Fo, V, jas, True, tol=1e-10, max_iter=200, mgs_mode=1, diis=6)
        assert abs(ed - ref["energy"]) < 1e-8, (ed, ref["energy"])
    else:
        with pytest.raises(tc.TcscfError):
            tc.TCHartreeFock(nup, ndown, nb).compute(S, H, g1, Fs, Fo, V, jas, True, tol=1e-10, mgs_mode=0, diis=6)


def test_scf_iteration_graph_matches_direct_enqueue(tc, gpu, monkeypatch):
    """Each SCF iteration's device build replays one HIP graph (kernel nodes
    only: the D / C upload from mapped page-locked memory, the sweeps, the
    reductions writing into mapped memory); TCSCF_SCF_GRAPH=0 enqueues the same
    kernels directly.  Both give bitwise-equal traces, densities and energies
    for RCS, UOS, restricted TC on FP32 tensors (column kernel), open-shell TC
    and TC with independent V_same / V_oppo.  Every iteration replays the
    graph; a solve captures it at most a few times (once per kernel choice),
    and the same solve again replays the graph kept from the first without a
    capture."""
    from tcscf_amd import configs

    def both(run):
        out = {}
        for mode in ("1", "0", "1"):  # graphs for every driver, none, every driver again
            monkeypatch.setenv("TCSCF_SCF_GRAPH", mode)
            out[mode + str(len(out))] = run()
        monkeypatch.delenv("TCSCF_SCF_GRAPH")
        g1, d0, g2 = out["10"], out["01"], out["12"]
        for a in (g1, g2):
            assert a[0] == d0[0] and np.array_equal(a[1], d0[1]) and np.array_equal(a[2], d0[2])
        its = len(d0[1])
        assert g1[3] == its and 1 <= g1[4] <= 3, g1[3:]  # every iteration replayed; few captures
        assert d0[3] == 0 and d0[4] == 0, d0[3:]
        assert g2[3] == its and g2[4] == 0, g2[3:]  # the kept graph served the repeat

    def obs(hf, e):
        return e, hf.energy_trace, hf.density.copy(), hf.graph_iterations, hf.graph_captures

    c = configs.he_config1()
    g, S, H = _scf_inputs(tc, c)
    F = tc.computeF(g, g)
    both(lambda: (lambda hf: obs(hf, hf.compute(S, H, g, F, False, tol=1e-10)))(tc.RCSHartreeFock(2, 5)))
    basis = tc.even_tempered({0: 6, 1: 2}, 0.5, 2.0)
    g2 = _grid(tc, 14, 11, tc.RADIAL_M4)
    g2.setBasisFunctions(basis, False)
    S2, H2 = tc.one_electron_matrices(basis, 5)
    F2 = tc.computeF(g2, g2, precision="bf16x5", out_dtype="fp32")
    both(lambda: (lambda hf: obs(hf, hf.compute(S2, H2, g2, F2, True, tol=1e-10, diis=4)))(
        tc.UOSHartreeFock(3, 2, len(basis))))
    tb = tc.even_tempered({0: 5, 1: 2}, 0.6, 2.2)
    nb = len(tb)
    g1, g3, jas, Fs, Fo, V = _tc_case(tc, tb, (14, 11, tc.RADIAL_M4), (10, 7, tc.RADIAL_M4))
    Fs32, Fo32, V32 = Fs.float(), Fo.float(), V.float()
    for (nup, ndown, Z, tensors) in ((2, 2, 4, (Fs32, Fo32, V32, jas)), (3, 2, 5, (Fs, Fo, V, jas)),
                                     (2, 2, 4, (Fs, Fo, jas.c_same * V, jas.c_oppo * V))):
        St, Ht = tc.one_electron_matrices(tb, Z)
        both(lambda: (lambda hf: obs(hf, hf.compute(St, Ht, g1, *tensors, True, tol=1e-10, max_iter=100,
                                                    mgs_mode=1, diis=6)))(tc.TCHartreeFock(nup, ndown, nb)))
    # default: the TC driver replays its graph, RCS / UOS enqueue directly
    monkeypatch.delenv("TCSCF_SCF_GRAPH", raising=False)
    hf = tc.TCHartreeFock(2, 2, nb)
    hf.compute(St, Ht, g1, Fs32, Fo32, V32, jas, True, tol=1e-10, max_iter=100, mgs_mode=1, diis=6)
    assert hf.graph_iterations == hf.numberOfConvergenceLoops() > 0
    hr = tc.RCSHartreeFock(2, 5)
    hr.compute(S, H, g, F, False, tol=1e-10)
    assert hr.graph_iterations == 0


_OCHI_TERMS = [(1, 0, 0), (2, 1, 0), (0, 2, 1), (1, 1, 1)]


def test_F_general_ochi_vs_oracle(tc, O, gpu):
    """The general Ochi Jastrow (src/jastrowFunctions.hpp:9-115: r12, r1 and r2
    powers, coefficients per spin that are not proportional) in the TC scalar
    kernel and computeV, both spins, at FP64 against the oracle's general
    evaluation (1e-12); the FP32 and BF16x5 contractions against FP64
    (tolerances in the assertions); and S = [(1, 0, 0)] against the r12-only
    kernels."""
    basis = _mixed_basis()[:9]
    g1 = _grid(tc, 11, 11, tc.RADIAL_M4)
    g3 = _grid(tc, 9, 7, tc.RADIAL_M4)
    for g in (g1, g3):
        g.setBasisFunctions(basis, True)
    c = [(0.5, 0.25), (-0.3, 0.2), (0.15, -0.4), (0.1, 0.05)]
    u = tc.OchiJastrow(1.3, 0.8, c, _OCHI_TERMS)
    prm = O.ochi_params(1.3, 0.8, [(*S, co, cs) for S, (co, cs) in zip(_OCHI_TERMS, c)])
    for same in (True, False):
        F = tc.computeF(g1, g1, kernel="tc", jastrow=u, same_spin=same)
        ref = _F_oracle(O, g1, g1, kernel=O.K_TC_SCALAR, prm=prm, same=same)
        assert rel_err(tc.unpack(F.cpu().numpy(), 9), ref) < 1e-12
        V = tc.computeV(g1, g3, u, same_spin=same).cpu().numpy()
        ref = _F_oracle(O, g1, g3, kernel=O.K_JASTROW_GRAD, prm=prm, same=same)
        assert rel_err(tc.unpack(V, 9).transpose(0, 2, 3, 1), ref) < 1e-12
        for prec, tol in (("fp32", 3e-6), ("bf16x5", 3e-6)):
            M = tc.computeF(g1, g1, kernel="tc", jastrow=u, same_spin=same, precision=prec)
            assert rel_err(M.cpu().numpy(), F.cpu().numpy()) < tol, prec
    jas = tc.Jastrow.from_b(1.3)
    u1 = tc.OchiJastrow.from_jastrow(jas)
    for same in (True, False):
        a = tc.computeF(g1, g1, kernel="tc", jastrow=u1, same_spin=same).cpu().numpy()
        b = tc.computeF(g1, g1, kernel="tc", jastrow=jas, same_spin=same).cpu().numpy()
        assert rel_err(a, b) < 1e-13
    V1 = tc.computeV(g1, g3, u1, same_spin=True).cpu().numpy()
    assert rel_err(V1, jas.c_same * tc.computeV(g1, g3, jas).cpu().numpy()) < 1e-13
    with pytest.raises(TypeError):
        tc.computeF_tc_pair(g1, g1, u)


def test_tc_scf_general_ochi_vs_oracle(tc, O, gpu):
    """TCHartreeFock with a general Ochi Jastrow whose opposite-spin
    coefficients are twice the same-spin ones term by term (as in the
    reference's r12-only driver, HartreeFockTest.cpp:351-360) against the
    oracle fed VjiSame and VjiOppo separately; a Jastrow without that
    proportionality is rejected by the device Fock build's one-V interface."""
    basis = tc.even_tempered({0: 5, 1: 2}, 0.6, 2.2)
    nb = len(basis)
    g1 = _grid(tc, 14, 11, tc.RADIAL_M4)
    g3 = _grid(tc, 10, 7, tc.RADIAL_M4)
    for g in (g1, g3):
        g.setBasisFunctions(basis, True)
    cs = [0.25, -0.1, 0.05, 0.08]
    u = tc.OchiJastrow(1.1, 1.0, [(2 * x, x) for x in cs], _OCHI_TERMS)
    assert u.spin_ratio() == 2.0
    Fs = tc.computeF(g1, g1, kernel="tc", jastrow=u, same_spin=True)
    Fo = tc.computeF(g1, g1, kernel="tc", jastrow=u, same_spin=False)
    Vs = tc.computeV(g1, g3, u, same_spin=True)
    Vo = tc.computeV(g1, g3, u, same_spin=False)
    assert rel_err(Vo.cpu().numpy(), 2 * Vs.cpu().numpy()) < 1e-14
    S, H = tc.one_electron_matrices(basis, 4)
    hf = tc.TCHartreeFock(2, 2, nb)
    e = hf.compute(S, H, g1, Fs, Fo, Vs, u, True, tol=1e-10, max_iter=200, mgs_mode=1)
    A = g1.basisValues.cpu().numpy()
    dA = g1.basisGradients.cpu().numpy()
    full = lambda V: np.ascontiguousarray(tc.unpack(V.cpu().numpy(), nb).transpose(0, 2, 3, 1))
    ref = O.tc_scf(2, 2, S, H, A, dA, g1.weights_host, tc.unpack(Fs.cpu().numpy(), nb),
                   tc.unpack(Fo.cpu().numpy(), nb), full(Vs), full(Vo), tol=1e-10, max_iter=200, mgs_mode=1)
    assert ref["status"] == 0
    assert abs(e - ref["energy"]) < 1e-9, (e, ref["energy"])
    assert np.max(np.abs(hf.density - ref["D"])) < 1e-7
    # the reference's signature with both tensors: the same fixed point
    e2 = tc.TCHartreeFock(2, 2, nb).compute(S, H, g1, Fs, Fo, Vs, Vo, True, tol=1e-10, max_iter=200, mgs_mode=1)
    assert abs(e2 - e) < 1e-10
    bad = tc.OchiJastrow(1.1, 1.0, [(0.5, 0.25), (0.3, 0.1)], _OCHI_TERMS[:2])
    with pytest.raises(ValueError):
        tc.TCHartreeFock(2, 2, nb).compute(S, H, g1, Fs, Fo, Vs, bad, True, tol=1e-10)


@pytest.mark.parametrize("nup,ndown,mgs_mode,dtype", [(2, 2, 1, "fp64"), (3, 2, 1, "fp64"), (2, 2, 0, "fp64"),
                                                      (3, 2, 1, "fp32")])
def test_tc_scf_independent_V_vs_oracle(tc, O, gpu, nup, ndown, mgs_mode, dtype):
    """TCHartreeFock::compute with the reference's signature
    (src/HartreeFock.hpp:386-394): F_same, F_oppo, V_same and V_oppo from a
    general Ochi Jastrow whose spin coefficients are not proportional, so the
    two V tensors are independent; against the oracle fed the same tensors
    (energy trace 1e-9 Eh; FP32 storage against the oracle fed the rounded
    values).  The observables: fockOperator is H + G2 - G3 of the final
    density (mask applied), the occupied orbitals give the density and are
    S-orthonormal under the metric Gram-Schmidt."""
    torch = gpu
    basis = tc.even_tempered({0: 5, 1: 2}, 0.6, 2.2)
    nb = len(basis)
    g1 = _grid(tc, 14, 11, tc.RADIAL_M4)
    g3 = _grid(tc, 10, 7, tc.RADIAL_M4)
    for g in (g1, g3):
        g.setBasisFunctions(basis, True)
    c = [(0.5, 0.25), (-0.3, 0.2), (0.15, -0.4), (0.1, 0.05)]
    u = tc.OchiJastrow(1.1, 1.0, c, _OCHI_TERMS)
    assert u.spin_ratio() is None
    Fs = tc.computeF(g1, g1, kernel="tc", jastrow=u, same_spin=True)
    Fo = tc.computeF(g1, g1, kernel="tc", jastrow=u, same_spin=False)
    Vs = tc.computeV(g1, g3, u, same_spin=True)
    Vo = tc.computeV(g1, g3, u, same_spin=False)
    if dtype == "fp32":
        Fs, Fo, Vs, Vo = (T.float() for T in (Fs, Fo, Vs, Vo))
    S, H = tc.one_electron_matrices(basis, nup + ndown)
    hf = tc.TCHartreeFock(nup, ndown, nb)
    e = hf.compute(S, H, g1, Fs, Fo, Vs, Vo, True, tol=1e-10, max_iter=200, mgs_mode=mgs_mode)
    A = g1.basisValues.cpu().numpy()
    dA = g1.basisGradients.cpu().numpy()
    full = lambda V: np.ascontiguousarray(tc.unpack(V.double().cpu().numpy(), nb).transpose(0, 2, 3, 1))
    Fso, Foo = tc.unpack(Fs.double().cpu().numpy(), nb), tc.unpack(Fo.double().cpu().numpy(), nb)
    ref = O.tc_scf(nup, ndown, S, H, A, dA, g1.weights_host, Fso, Foo, full(Vs), full(Vo), tol=1e-10, max_iter=200,
                   mgs_mode=mgs_mode)
    assert ref["status"] == 0
    n = min(len(hf.energy_trace), len(ref["trace"]))
    assert np.max(np.abs(hf.energy_trace[:n] - ref["trace"][:n])) < 1e-9
    assert abs(e - ref["energy"]) < 1e-9, (e, ref["energy"])
    assert np.max(np.abs(hf.density - ref["D"])) < 1e-7
    # observables
    nocc = (nup, ndown)
    for s_ in range(2):
        C = hf.occupiedOrbitals[s_]
        assert C.shape == (nb, nocc[s_])
        assert np.max(np.abs(C @ C.T - hf.density[s_])) < 1e-12
        if mgs_mode == 1:
            assert np.max(np.abs(C.T @ S @ C - np.eye(nocc[s_]))) < 1e-10
    Cpad = np.zeros((2, nb, max(nocc)))
    for s_ in range(2):
        Cpad[s_, :, :nocc[s_]] = hf.occupiedOrbitals[s_]
    G2, G3 = O.tc_fock(nup, ndown, A, dA, g1.weights_host, Fso, Foo, full(Vs), full(Vo), hf.density, Cpad)
    mask = np.abs(H) > 0
    fock = H[None] + np.where(mask[None], G2, 0) - np.where(mask[None], G3, 0)
    assert rel_err(hf.fockOperator, fock) < 1e-9


def test_tc_zero_jastrow_equals_uos(tc, gpu):
    """c_same = c_oppo = 0 on the device path reproduces UOSHartreeFock."""
    basis = tc.even_tempered({0: 5, 1: 2}, 0.6, 2.2)
    nb = len(basis)
    g1 = _grid(tc, 14, 11, tc.RADIAL_M4)
    g1.setBasisFunctions(basis, True)
    jas = tc.Jastrow(1.0, 0.0, 0.0)
    Fs = tc.computeF(g1, g1, kernel="tc", jastrow=jas, same_spin=True)
    V = tc.computeV(g1, g1, jas)
    S, H = tc.one_electron_matrices(basis, 4)
    e = tc.TCHartreeFock(2, 2, nb).compute(S, H, g1, Fs, Fs, V, jas, True, tol=1e-10, mgs_mode=1)
    F = tc.computeF(g1, g1)
    eu = tc.UOSHartreeFock(2, 2, nb).compute(S, H, g1, F, True, tol=1e-10)
    assert abs(e - eu) < 1e-9


@pytest.mark.parametrize("nup,ndown,vform", [(2, 2, "one"), (2, 1, "one"), (1, 1, "one"), (2, 2, "two"),
                                              (2, 1, "two"), (1, 1, "two")])
@pytest.mark.parametrize("d_off", [False, True])
@pytest.mark.parametrize("f32", [False, True])
def test_tc_fock_build_vs_oracle_and_explicit_Q(tc, O, gpu, nup, ndown, vform, d_off, f32):
    """One device TC Fock build at a given density and orbitals (different per
    spin) against the oracle's factorized restatement (G2 and G3), and G3
    against assembleG over the spin-labelled explicit three-body tensors
    (src/HartreeFock.cpp:613-750).  vform 'one': one c = 1 tensor with the r12
    Jastrow's scales; 'two': independent V_same / V_oppo of a general Ochi
    Jastrow (the reference's VjiSame / VjiOppo).  f32: the tensors stored in
    FP32 (the oracle fed the same values); with D = C C^T and one V tensor this
    runs the column point kernel."""
    from tcscf_amd import SlaterTypeOrbital as STO
    basis = [STO(1.3, 1, 0, 0), STO(0.9, 2, 1, 0), STO(0.7, 2, 0, 0)]
    nb = len(basis)
    g1, g3, jas, Fs, Fo, V = _tc_case(tc, basis, (6, 7, tc.RADIAL_M4), (8, 7, tc.RADIAL_M4))
    if f32:
        Fs, Fo, V = Fs.float(), Fo.float(), V.float()
    rng = np.random.default_rng(11)
    C = np.zeros((2, nb, max(nup, ndown)))
    D = np.zeros((2, nb, nb))
    for s, no in enumerate((nup, ndown)):
        C[s, :, :no] = rng.normal(size=(nb, no)) * 0.5
        D[s] = C[s, :, :no] @ C[s, :, :no].T
    if d_off:
        # a density that is not C C^T: the build then runs the D-based point
        # kernel instead of the column kernel (which needs D = C C^T)
        D = D + 0.01 * np.eye(nb)[None]
    A = g1.basisValues.cpu().numpy()
    dA = g1.basisGradients.cpu().numpy()
    if vform == "one":
        G2, G3 = tc.TCHartreeFock(nup, ndown, nb).fockBuild(g1, Fs, Fo, V, jas, D, C)
        Vf = tc.unpack(V.double().cpu().numpy(), nb).transpose(0, 2, 3, 1)
        Vs, Vo = np.ascontiguousarray(jas.c_same * Vf), np.ascontiguousarray(jas.c_oppo * Vf)
    else:
        u = tc.OchiJastrow(1.1, 1.0, [(0.5, 0.25), (-0.3, 0.2), (0.15, -0.4)], _OCHI_TERMS[:3])
        Vsd, Vod = tc.computeV(g1, g3, u, same_spin=True), tc.computeV(g1, g3, u, same_spin=False)
        if f32:
            Vsd, Vod = Vsd.float(), Vod.float()
        G2, G3 = tc.TCHartreeFock(nup, ndown, nb).fockBuild(g1, Fs, Fo, Vsd, Vod, D, C)
        Vs, Vo = (np.ascontiguousarray(tc.unpack(T.double().cpu().numpy(), nb).transpose(0, 2, 3, 1)) for T in (Vsd, Vod))
    G2o, G3o = O.tc_fock(nup, ndown, A, dA, g1.weights_host, tc.unpack(Fs.double().cpu().numpy(), nb),
                         tc.unpack(Fo.double().cpu().numpy(), nb), Vs, Vo, D, C)
    assert rel_err(G2, G2o) < 1e-12
    assert rel_err(G3, G3o) < 1e-12
    if not d_off:  # the explicit three-body form holds for D = C C^T
        G3q = O.assemble_G3(*O.compute_Q(A, g1.weights_host, Vs, Vo, labelled=True), D)
        assert rel_err(G3, G3q) < 1e-11


@pytest.mark.parametrize("nbasis,nup,ndown,same,dtype,vform,launch",
                         [(12, 7, 6, False, "fp64", "one", None), (20, 6, 7, False, "fp64", "one", None),
                          (36, 9, 8, False, "fp64", "one", (0, 8)), (20, 6, 6, True, "fp64", "one", None),
                          (36, 7, 7, True, "fp64", "one", None),
                          # FP32 storage, unstaged 8-wave launches: two spins at nb >= 40, restricted at nb 56
                          (44, 6, 5, False, "fp32", "one", (0, 8)), (56, 7, 7, True, "fp32", "one", (0, 8)),
                          (20, 6, 6, True, "fp32", "one", (1, 12)),
                          # independent V_same / V_oppo: W blocks of both tensors
                          (20, 7, 6, False, "fp64", "two", None), (36, 7, 7, True, "fp32", "two", None),
                          (40, 6, 5, False, "fp32", "two", (0, 8))])
def test_tc_fock_build_many_orbitals_vs_oracle(tc, O, gpu, nbasis, nup, ndown, same, dtype, vform, launch):
    """The device TC Fock build with more occupied orbitals than one register
    block of W = V C holds (several blocks and LDS chunks for |U|^2), for each
    lane split of the basis (16, 32 and 64 lanes per basis row), against the
    oracle's factorized restatement.  same: a closed shell with identical spin
    densities and orbitals, which forms spin 0 once; both spin blocks must then
    be bitwise equal.  dtype fp32: F / V stored in FP32, the oracle fed the same
    rounded values (the build's arithmetic is FP64, so 1e-12 still holds).
    launch: the (staged, waves per workgroup) the build must pick, from
    tcscf_tc_launch_info."""
    import ctypes
    basis = (_mixed_basis() * 4)[:nbasis]
    nb = len(basis)
    g1, g3, jas, Fs, Fo, V = _tc_case(tc, basis, (5, 7, tc.RADIAL_M4), (6, 7, tc.RADIAL_M4))
    if vform == "two":
        u = tc.OchiJastrow(1.1, 1.0, [(0.5, 0.25), (-0.3, 0.2)], _OCHI_TERMS[:2])
        Vd, Vod = tc.computeV(g1, g3, u, same_spin=True), tc.computeV(g1, g3, u, same_spin=False)
    else:
        Vd, Vod = V, None
    if dtype == "fp32":
        Fs, Fo, Vd = Fs.float(), Fo.float(), Vd.float()
        Vod = Vod.float() if Vod is not None else None
    from tcscf_amd import _lib
    info = (ctypes.c_int * 6)()
    _lib.check(tc.lib().tcscf_tc_launch_info(nup, ndown, nb, _lib.F32 if dtype == "fp32" else _lib.F64,
                                             int(vform == "two"), info))
    if launch is not None:
        r = 3 if same else 0
        assert (info[r], info[r + 1]) == launch, list(info)
    rng = np.random.default_rng(nbasis)
    C = np.zeros((2, nb, max(nup, ndown)))
    D = np.zeros((2, nb, nb))
    C0 = rng.normal(size=(nb, max(nup, ndown))) * 0.3
    for s, no in enumerate((nup, ndown)):
        C[s, :, :no] = C0[:, :no] if same else rng.normal(size=(nb, no)) * 0.3
        D[s] = C[s, :, :no] @ C[s, :, :no].T
    hf = tc.TCHartreeFock(nup, ndown, nb)
    G2, G3 = hf.fockBuild(g1, Fs, Fo, Vd, jas if Vod is None else Vod, D, C)
    if same:
        assert np.array_equal(G2[0], G2[1]) and np.array_equal(G3[0], G3[1])
    A = g1.basisValues.cpu().numpy()
    dA = g1.basisGradients.cpu().numpy()
    full = lambda T: tc.unpack(T.double().cpu().numpy(), nb).transpose(0, 2, 3, 1)
    if Vod is None:
        Vs, Vo = np.ascontiguousarray(jas.c_same * full(Vd)), np.ascontiguousarray(jas.c_oppo * full(Vd))
    else:
        Vs, Vo = np.ascontiguousarray(full(Vd)), np.ascontiguousarray(full(Vod))
    G2o, G3o = O.tc_fock(nup, ndown, A, dA, g1.weights_host, tc.unpack(Fs.double().cpu().numpy(), nb),
                         tc.unpack(Fo.double().cpu().numpy(), nb), Vs, Vo, D, C)
    assert rel_err(G2, G2o) < 1e-12
    assert rel_err(G3, G3o) < 1e-12


def test_grid_eri_vs_oracle(tc, O, gpu):
    """integrateAllR1R2 (src/integration/integrateAll.hpp:221-317) on the
    device: against the oracle's contraction of the same F, the 8-fold
    permutational symmetry of real orbitals, and the 1s closed form."""
    from tcscf_amd import configs
    c = configs.he_config1()
    g = _grid(tc, *c["grid"])
    g.setBasisFunctions(c["basis"], False)
    F = tc.computeF(g, g)
    E = tc.eri_full(tc.integrateAllR1R2(g, F), 5)
    A = g.basisValues.cpu().numpy()
    ref = O.eri_from_F(A, g.weights_host, tc.unpack(F.cpu().numpy(), 5))
    assert rel_err(E, ref) < 1e-13
    chem = E.transpose(0, 2, 1, 3)  # (b1 b3 | b2 b4)
    for perm in ((1, 0, 2, 3), (0, 1, 3, 2), (2, 3, 0, 1)):
        assert rel_err(chem.transpose(perm), chem) < 1e-12
    # several r1 chunks, a different r2 grid, a basis with l > 0
    basis = _mixed_basis()[:9]
    g1 = _grid(tc, 23, 11, tc.RADIAL_M4)
    g2 = _grid(tc, 17, 7, tc.RADIAL_BECKE)
    g1.setBasisFunctions(basis, False)
    g2.setBasisFunctions(basis, False)
    F = tc.computeF(g1, g2)
    E = tc.eri_full(tc.integrateAllR1R2(g1, F), 9)
    ref = O.eri_from_F(g1.basisValues.cpu().numpy(), g1.weights_host, tc.unpack(F.cpu().numpy(), 9))
    assert rel_err(E, ref) < 1e-13


def test_three_electron_integrals_vs_oracle(tc, O, gpu):
    """threeElectronIntegrals on the device against the oracle, from the same
    device V and end to end from the oracle's own V."""
    from tcscf_amd import SlaterTypeOrbital as STO
    basis = [STO(1.3, 1, 0, 0), STO(0.9, 2, 1, 1), STO(0.7, 2, 0, 0), STO(1.1, 2, 1, -1)]
    nb = len(basis)
    g1 = _grid(tc, 7, 7, tc.RADIAL_M4)
    g2 = _grid(tc, 9, 11, tc.RADIAL_M4)
    for g in (g1, g2):
        g.setBasisFunctions(basis, False)
    jas = tc.Jastrow.from_b(1.0)
    Q = tc.threeElectronIntegrals(g1, g2, jas).cpu().numpy()
    A1 = g1.basisValues.cpu().numpy()
    V = tc.computeV(g1, g2, jas)
    Vf = np.ascontiguousarray(jas.c_same * tc.unpack(V.cpu().numpy(), nb).transpose(0, 2, 3, 1))
    ref = O.three_electron(A1, g1.weights_host, Vf)
    assert np.max(np.abs(ref)) > 1e-8
    assert rel_err(Q, ref) < 1e-12
    prm = O.ochi_params(1.0, 1.0, [(1, 0, 0, jas.c_oppo, jas.c_same)])
    Vo = O.compute_F(g1.points_host, g2.points_host, g2.weights_host, g2.basisValues.cpu().numpy(),
                     kernel=O.K_JASTROW_GRAD, prm=prm, same=True)
    assert rel_err(Q, O.three_electron(A1, g1.weights_host, Vo)) < 1e-11


def test_gradient_eri_vs_oracle(tc, O, gpu):
    """integrateAllR1R2<true> on the device against the oracle's brute-force
    restatement, on a 3-D r2 grid and on the reference's 2-D r2 grid (2 pi);
    and with synthetic gradients A e_c it equals the plain ERI of V_c."""
    import torch
    from tcscf_amd import SlaterTypeOrbital as STO
    basis = [STO(1.1, 1, 0, 0), STO(0.8, 2, 1, 1), STO(0.6, 2, 0, 0), STO(0.9, 3, 2, -1), STO(1.4, 2, 1, 0)]
    nb = len(basis)
    jas = tc.Jastrow.from_b(1.3)
    prm = O.ochi_params(1.0, jas.a12, [(1, 0, 0, jas.c_oppo, jas.c_same)])
    rng = np.random.default_rng(5)
    # scattered points: no parity zeros in any component
    g1 = tc.QMCGrid(rng.uniform(-2.5, 2.5, (3, 700)), rng.uniform(0.01, 0.05, 700))
    g1.setBasisFunctions(basis, False)
    p2d = rng.uniform(0.05, 3, (3, 230))
    p2d[1] = 0
    for g2, scale in ((_grid(tc, 11, 7, tc.RADIAL_M4), 1.0),
                      (tc.QMCGrid(p2d, rng.uniform(0.1, 1, 230), ndim=2), 2 * math.pi)):
        g2.setBasisFunctions(basis, True)
        G = tc.eri_full_gradient(tc.integrateAllR1R2Gradient(g1, g2, jas), nb)
        A2 = g2.basisValues.cpu().numpy()
        ref = O.eri_gradient(prm, True, g1.points_host, g1.basisValues.cpu().numpy(), g1.weights_host,
                             g2.points_host, g2.weights_host, A2, g2.basisGradients.cpu().numpy(), scale=scale)
        assert np.max(np.abs(ref)) > 1e-6
        assert rel_err(G, ref) < 1e-12
    g2 = tc.QMCGrid(rng.uniform(-3, 3, (3, 500)), rng.uniform(0.01, 0.05, 500))
    g2.setBasisFunctions(basis, True)
    V = tc.computeV(g1, g2, jas)  # c = 1
    for c in range(3):
        dA = torch.zeros_like(g2.basisGradients)
        dA[..., c] = g2.basisValues
        g2.basisGradients = dA
        G = tc.eri_full_gradient(tc.integrateAllR1R2Gradient(g1, g2, jas), nb)
        E = tc.eri_full(tc.integrateAllR1R2(g1, jas.c_same * V[:, c, :].contiguous()), nb)
        assert np.max(np.abs(E)) > 1e-6
        assert rel_err(G.transpose(1, 0, 3, 2), E) < 1e-12


def test_tc_scf_on_sobol_grids_vs_oracle(tc, O, gpu):
    """TC SCF on importance-sampled Sobol grids (separate seeds for r1 and r3)
    against the oracle on the same points, small size."""
    basis = tc.even_tempered({0: 4, 1: 1}, 0.8, 2.0)
    nb = len(basis)
    g1 = tc.QMCGrid.sobol(2 ** 11, seed=1)
    g3 = tc.QMCGrid.sobol(2 ** 10, seed=3)
    for g in (g1, g3):
        g.setBasisFunctions(basis, True)
    jas = tc.Jastrow.from_b(1.0)
    Fs = tc.computeF(g1, g1, kernel="tc", jastrow=jas, same_spin=True)
    Fo = tc.computeF(g1, g1, kernel="tc", jastrow=jas, same_spin=False)
    V = tc.computeV(g1, g3, jas)
    S, H = tc.one_electron_matrices(basis, 4)
    hf = tc.TCHartreeFock(2, 2, nb)
    e = hf.compute(S, H, g1, Fs, Fo, V, jas, True, tol=1e-10, mgs_mode=1)
    prm = O.ochi_params(1.0, jas.a12, [(1, 0, 0, jas.c_oppo, jas.c_same)])
    A = g1.basisValues.cpu().numpy()
    Fs_o = O.compute_F(g1.points_host, g1.points_host, g1.weights_host, A, kernel=O.K_TC_SCALAR, prm=prm, same=True)
    Fo_o = O.compute_F(g1.points_host, g1.points_host, g1.weights_host, A, kernel=O.K_TC_SCALAR, prm=prm, same=False)
    A3 = g3.basisValues.cpu().numpy()
    Vs = O.compute_F(g1.points_host, g3.points_host, g3.weights_host, A3, kernel=O.K_JASTROW_GRAD, prm=prm, same=True)
    Vo = O.compute_F(g1.points_host, g3.points_host, g3.weights_host, A3, kernel=O.K_JASTROW_GRAD, prm=prm, same=False)
    ref = O.tc_scf(2, 2, S, H, A, g1.basisGradients.cpu().numpy(), g1.weights_host, Fs_o, Fo_o, Vs, Vo, tol=1e-10,
                   mgs_mode=1)
    assert ref["status"] == 0
    assert abs(e - ref["energy"]) < 1e-9, (e, ref["energy"])


def test_cpp_facade_on_device(tc, O, gpu):
    """The C++ facade (include/tcscf.hpp, driven by tests/cpp/facade_test.cpp)
    against the oracle and the Python mirror: He config 1 RHF/UHF, the grid
    ERI, Ne with the TZ3P basis from loadSTO (FP64 and BF16x5) and TC Be."""
    import os
    import subprocess
    from conftest import ROOT
    from tcscf_amd import SlaterTypeOrbital as STO, configs
    exe = os.path.join(ROOT, "tcscf_amd", "_build", "facade_test")
    basis_dir = os.path.join(ROOT, "tests", "golden", "basisFunctions")
    print("step: facade_test gpu", flush=True)
    r = subprocess.run([exe, "gpu", basis_dir], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr
    out = {k: float(v) for k, v in (line.split() for line in r.stdout.splitlines() if len(line.split()) == 2)}
    c = configs.he_config1()
    g = _grid(tc, *c["grid"])
    g.setBasisFunctions(c["basis"], False)
    S, H = tc.one_electron_matrices(c["basis"], c["Z"])
    A = O.sto_real(g.points_host, fns_of(c["basis"]))
    Fo = O.compute_F(g.points_host, g.points_host, g.weights_host, A)
    ref = O.rcs_scf(2, S, H, A, g.weights_host, Fo, tol=1e-10, mask=False)
    assert abs(out["he_rhf"] - ref["energy"]) < 1e-10
    assert abs(out["he_uhf"] - out["he_rhf"]) < 1e-10
    eri = O.eri_from_F(A, g.weights_host, Fo)
    assert abs(out["he_eri_0000"] - eri[0, 0, 0, 0]) < 1e-12 * abs(eri[0, 0, 0, 0])
    assert abs(out["he_eri_0123"] - eri[0, 1, 2, 3]) < 1e-12 * abs(eri[0, 0, 0, 0])
    # Ne TZ3P through the Python mirror
    basis = tc.loadSTO(basis_dir, "TZ3P", "Ne")
    gn = _grid(tc, 40, 29, tc.RADIAL_M4)
    gn.setBasisFunctions(basis, False)
    Sn, Hn = tc.one_electron_matrices(basis, 10)
    e64 = tc.RCSHartreeFock(10, len(basis)).compute(Sn, Hn, gn, tc.computeF(gn, gn), True, tol=1e-10, max_iter=100,
                                                     diis=6)
    assert abs(out["ne_tz3p_fp64"] - e64) < 1e-9
    assert abs(out["ne_tz3p_bf16x5"] - e64) < 1e-6
    # TC Be
    bb = [STO(0.6 * 2.2 ** k, 1, 0, 0) for k in range(5)]
    bb += [STO(0.6 * 2.2 ** k, 2, 1, m) for k in range(2) for m in (-1, 0, 1)]
    g1, g3, jas, Fs, Fo2, V = _tc_case(tc, bb, (14, 11, tc.RADIAL_M4), (10, 7, tc.RADIAL_M4))
    Sb, Hb = tc.one_electron_matrices(bb, 4)
    etc = tc.TCHartreeFock(2, 2, len(bb)).compute(Sb, Hb, g1, Fs, Fo2, V, jas, True, tol=1e-10, mgs_mode=1)
    assert abs(out["be_tc"] - etc) < 1e-9
    assert abs(out["be_tc_mixed"] - etc) < 1e-6
    # the general Ochi path with S = {(1, 0, 0)}: the same kernel up to rounding
    assert abs(out["be_tc_ochi"] - etc) < 1e-9
    # the reference's signature (VjiSame, VjiOppo as two tensors) on the same Jastrow
    assert abs(out["be_tc_ref_signature"] - out["be_tc_ochi"]) < 1e-10
    # ... and on a general Ochi Jastrow with independent V_same / V_oppo, against the oracle
    ug = tc.OchiJastrow(1.1, 1.0, [(0.5, 0.25), (-0.3, 0.2), (0.15, -0.4)], [(1, 0, 0), (2, 1, 0), (1, 1, 1)])
    Fsg = tc.computeF(g1, g1, kernel="tc", jastrow=ug, same_spin=True)
    Fog = tc.computeF(g1, g1, kernel="tc", jastrow=ug, same_spin=False)
    full = lambda V: np.ascontiguousarray(tc.unpack(V.cpu().numpy(), len(bb)).transpose(0, 2, 3, 1))
    Vsg, Vog = full(tc.computeV(g1, g3, ug, same_spin=True)), full(tc.computeV(g1, g3, ug, same_spin=False))
    refg = O.tc_scf(2, 2, Sb, Hb, g1.basisValues.cpu().numpy(), g1.basisGradients.cpu().numpy(), g1.weights_host,
                    tc.unpack(Fsg.cpu().numpy(), len(bb)), tc.unpack(Fog.cpu().numpy(), len(bb)), Vsg, Vog, tol=1e-10,
                    max_iter=200, mgs_mode=1)
    assert refg["status"] == 0
    assert abs(out["be_tc_general"] - refg["energy"]) < 1e-9
    assert out["be_tc_general_orth_err"] < 1e-10 and out["be_tc_general_density_err"] < 1e-12
    # the facade's per-element QMC route gives the Python route's numbers (same points, same kernels)
    H = tc.HydrogenLikeBasisFunction
    bz = [H(2, 1, 0, 0), H(2, 2, 0, 0), H(2, 2, 1, 0), H(2, 2, 1, 1)]
    A5, A6 = tc.fillAtomicR12Array(bz, 20000), tc.fillR12Array(bz, 20000)
    for key, val in [("atomic_j_2p0_2p1", A5[2, 3, 2, 3].real), ("atomic_k_2p0_2p1_re", A5[2, 3, 3, 2].real),
                     ("r12_0101", A6[0, 1, 0, 1]),
                     ("coulomb_op", tc.atomicCoulombOperator(bz[2], bz[3], bz[2], bz[3], 20000)),
                     ("exchange_op_re", tc.atomicExchangeOperator(bz[2], bz[3], bz[3], bz[2], 20000).real)]:
        assert abs(out[key] - val) < 1e-12 * abs(val), (key, out[key], val)
    assert abs(out["atomic_k_2p0_2p1_im"] - A5[2, 3, 3, 2].imag) < 1e-14


def test_batched_contraction_matches_single_calls(tc, gpu):
    """One persistent BF16 launch over uneven instances (different grids,
    bases, kernels, a 2-D r2 grid, V with 3 components) gives exactly the
    single-instance results; the autotuner picked a tile height."""
    import torch
    from tcscf_amd import SlaterTypeOrbital as STO
    rng = np.random.default_rng(4)
    specs, singles = [], []
    jas = tc.Jastrow.from_b(0.7)
    for k, (nrad, order, shells) in enumerate([(20, 11, {0: 4}), (31, 17, {0: 6, 1: 2}), (12, 7, {0: 3, 1: 1, 2: 1}),
                                               (45, 23, {0: 9, 1: 3})]):
        basis = tc.even_tempered(shells, 0.5, 2.0)
        g1 = _grid(tc, nrad, order, tc.RADIAL_M4)
        g1.setBasisFunctions(basis, False)
        kind = ["coulomb", "tc", "grad", "tc"][k]
        specs.append(dict(r1=g1, r2=g1, kernel=kind, jastrow=jas, same_spin=k % 2 == 0))
    p2d = rng.uniform(0.05, 3, (3, 333))
    p2d[1] = 0
    g2d = tc.QMCGrid(p2d, rng.uniform(0.1, 1, 333), ndim=2)
    gq = _grid(tc, 9, 11, tc.RADIAL_M4)
    b2 = [STO(0.8, 1, 0, 0), STO(1.7, 1, 0, 0), STO(1.1, 2, 0, 0)]
    for g in (g2d, gq):
        g.setBasisFunctions(b2, False)
    specs.append(dict(r1=gq, r2=g2d, kernel="coulomb"))
    for prec in ("bf16x5", "bf16x4", "bf16x3"):
        outs = tc.computeF_batch(specs, precision=prec)
        for sp, F in zip(specs, outs):
            if sp["kernel"] == "grad":
                ref = tc.computeV(sp["r1"], sp["r2"], sp["jastrow"], precision=prec)
            else:
                ref = tc.computeF(sp["r1"], sp["r2"], kernel=sp["kernel"], jastrow=sp.get("jastrow"),
                                  same_spin=sp.get("same_spin", True), precision=prec)
            assert torch.equal(F, ref), (prec, sp["kernel"])
        assert tc.lib().tcscf_batch_tuned_rows(tc._lib.PRECISIONS[prec]) in (16, 32)
    # FP32 storage: the same rounding of the same sums as the single calls
    outs = tc.computeF_batch(specs, precision="bf16x5", out_dtype="fp32")
    for sp, F in zip(specs, outs):
        assert F.dtype == torch.float32
        if sp["kernel"] == "grad":
            ref = tc.computeV(sp["r1"], sp["r2"], sp["jastrow"], precision="bf16x5", out_dtype="fp32")
        else:
            ref = tc.computeF(sp["r1"], sp["r2"], kernel=sp["kernel"], jastrow=sp.get("jastrow"),
                              same_spin=sp.get("same_spin", True), precision="bf16x5", out_dtype="fp32")
        assert torch.equal(F, ref), ("fp32", sp["kernel"])
    # FP64 goes instance by instance and equals the single calls too
    outs = tc.computeF_batch(specs[:2], precision="fp64")
    assert torch.equal(outs[0], tc.computeF(specs[0]["r1"], specs[0]["r2"]))


@pytest.mark.skip(reason="tcscf_batch_scf is switched off: an intermittent segmentation fault on the calling "
                         "thread inside it was seen in two runs of this test and its cause is not found yet")
def test_batched_scf_matches_single_solves(tc, gpu):
    """tcscf_batch_scf (a HIP graph of every active instance's Fock build per
    iteration) gives the single-instance drivers' energies and iteration counts
    for a mix of RCS, UOS and TC instances of different sizes, and recaptures
    its graph as instances converge."""
    from tcscf_amd import SlaterTypeOrbital as STO
    cases = []
    for Z, nup, ndown, kind, shells, grid in [(2, 1, 1, "rcs", {0: 5}, (20, 11)), (4, 2, 2, "rcs", {0: 6, 1: 1}, (30, 17)),
                                              (3, 2, 1, "uos", {0: 6, 1: 1}, (25, 11)), (4, 2, 2, "tc", {0: 5, 1: 2}, (14, 11)),
                                              (5, 3, 2, "tc", {0: 7, 1: 2}, (20, 17))]:
        basis = tc.even_tempered(shells, 0.5, 2.0)
        g = _grid(tc, *grid, tc.RADIAL_M4)
        g.setBasisFunctions(basis, kind == "tc")
        S, H = tc.one_electron_matrices(basis, Z)
        d = dict(kind=kind, nup=nup, ndown=ndown, S=S, H=H, r1=g)
        if kind == "tc":
            jas = tc.Jastrow.from_b(1.0)
            d.update(Fs=tc.computeF(g, g, kernel="tc", jastrow=jas, same_spin=True),
                     Fo=tc.computeF(g, g, kernel="tc", jastrow=jas, same_spin=False), V=tc.computeV(g, g, jas),
                     jastrow=jas)
        else:
            d.update(F=tc.computeF(g, g))
        cases.append(d)
    for diis in (0, 6):
        res, builds, _, _ = tc.batch_scf(cases, tol=1e-10, max_iter=200, mgs_mode=1, diis=diis)
        assert builds >= 2
        for d, r in zip(cases, res):
            assert r["status"] == 0, (d["kind"], r)
            nb = d["r1"].nBasis()
            if d["kind"] == "tc":
                hf = tc.TCHartreeFock(d["nup"], d["ndown"], nb)
                e = hf.compute(d["S"], d["H"], d["r1"], d["Fs"], d["Fo"], d["V"], d["jastrow"], True, tol=1e-10,
                               max_iter=200, mgs_mode=1, diis=diis)
            elif d["kind"] == "rcs":
                hf = tc.RCSHartreeFock(d["nup"] + d["ndown"], nb)
                e = hf.compute(d["S"], d["H"], d["r1"], d["F"], True, tol=1e-10, max_iter=200, diis=diis)
            else:
                hf = tc.UOSHartreeFock(d["nup"], d["ndown"], nb)
                e = hf.compute(d["S"], d["H"], d["r1"], d["F"], True, tol=1e-10, max_iter=200, diis=diis)
            assert r["energy"] == e, (d["kind"], r["energy"], e)
            assert r["iterations"] == hf.numberOfConvergenceLoops()
            assert np.array_equal(r["trace"], hf.energy_trace)


def test_atomic_qmc5d_operators_vs_oracle(tc, O, gpu):
    """atomicCoulombOperator, atomicExchangeOperator and fillAtomicR12Array
    (src/AtomicBasis.hpp:59-140, 197-248) on the device: every quadruple of the
    reference's complex functions, formed on the device from the real table,
    against the oracle's sum over the same 5-D point set (1e-12); a basis with
    m != 0 functions whose -m partners are missing (tabulated alongside).  Then
    the reference's own check (src/unitTests/AtomicBasisTest.cpp:222-267):
    hydrogen-like Z = 2 {1s, 2s, 2p0, 3p0}, fillAtomicR12Array against
    fillR12Array (the 6-D route) within 5e-3 on the Coulomb and exchange
    entries."""
    H = tc.HydrogenLikeBasisFunction
    basis = [H(2, 1, 0, 0), H(2, 2, 1, 1), H(2, 3, 2, -2), tc.SlaterTypeOrbital(1.3, 2, 1, -1),
             tc.OchiBasisFunction(1.1, 1, 1, 0)]
    fns = [(b.kind, b.n, b.l, b.m, b.alpha) for b in basis]
    p1, p2, w = tc.qmc_spherical_points(20000, 5, 2, 11)
    ref = O.eri_qmc5d_atomic(p1, p2, w, fns)
    lm = [(b.l, b.m) for b in basis]
    coul = lambda a, b, c, d: lm[a] == lm[c] and lm[b] == lm[d]
    exch = lambda a, b, c, d: lm[a] == lm[d] and lm[b] == lm[c]
    A = tc.fillAtomicR12Array(basis, 20000, 2, 11)
    want = O.fill_two_electron(5, True, lambda a, b, c, d: ref[a, b, c, d].real if coul(a, b, c, d) else 0)
    want += O.fill_two_electron(5, False, lambda a, b, c, d: ref[a, b, c, d]
                                if exch(a, b, c, d) and not coul(a, b, c, d) else 0)
    assert np.max(np.abs(A - want)) < 1e-12 * np.abs(want).max()
    assert np.count_nonzero(A) > 0 and np.count_nonzero(A) < A.size
    j = tc.atomicCoulombOperator(basis[1], basis[2], basis[1], basis[2], 20000, 2, 11)
    k = tc.atomicExchangeOperator(basis[1], basis[3], basis[3], basis[1], 20000, 2, 11)
    assert abs(j - ref[1, 2, 1, 2].real) < 1e-12 * abs(j) and abs(k - ref[1, 3, 3, 1]) < 1e-12 * abs(k)
    assert abs(k.imag) > 0  # m_c - m_a = -2: the phase leaves a (QMC-sized) imaginary part
    with pytest.raises(ValueError):
        tc.atomicCoulombOperator(basis[0], basis[1], basis[2], basis[1], 20000)
    # the reference's parity test, 5-D against 6-D
    bz = [H(2, 1, 0, 0), H(2, 2, 0, 0), H(2, 2, 1, 0), H(2, 3, 1, 0)]
    A5 = tc.fillAtomicR12Array(bz, 50000)
    A6 = tc.fillR12Array(bz, 50000)
    lm = [(b.l, b.m) for b in bz]
    worst = max(abs(A5[a, b, c, d] - A6[a, b, c, d]) for a in range(4) for b in range(a, 4) for c in range(a, 4)
                for d in range(a, 4) if (lm[a] == lm[c] and lm[b] == lm[d]) or (lm[a] == lm[d] and lm[b] == lm[c]))
    assert worst < 5e-3, worst


def test_ochi_basis_on_device_and_reference_he_runs(tc, O, gpu):
    """Ochi (src/OchiBasis.hpp) and hydrogen-like (src/HydrogenLikeBasis.hpp)
    functions tabulated on the device against the oracle's STO expansion (values
    and gradients); then He RCS with the orthonormal Ochi s basis n = 0..nBasis-1
    at the exponents the reference's recorded runs ended at
    (tests/golden/he_ochi_convergence.json, from its convergence.txt/data.txt).
    As in those runs, r1 and r2 are distinct QMC lattice grids and the energy is
    the mean over grids (here 8 pairs of 2x20011 / 2x10007 points).  The mean
    must lie within 4 standard errors + 2e-6 Eh of the grid-free RHF energy of
    the same basis (exact_energy) and within 2 combined standard deviations of
    the reference's recorded energy."""
    import json
    g = _grid(tc, 30, 11, tc.RADIAL_M4)
    mixed = [tc.OchiBasisFunction(1.2, n, l, m) for n in (0, 2, 5) for l in (0, 1, 2) for m in range(-l, l + 1)]
    hyd = [tc.HydrogenLikeBasisFunction(4, n, l, m) for n in (1, 2, 3) for l in range(min(n, 3)) for m in (-l, 0, l)]
    for basis in (mixed, hyd):
        g.setBasisFunctions(basis, True)
        v, gr = O.atomic_real(g.points_host, [(b.kind, b.n, b.l, b.m, b.alpha) for b in basis], grads=True)
        assert rel_err(g.basisValues.cpu().numpy(), v) < 1e-12
        assert rel_err(g.basisGradients.cpu().numpy(), gr) < 1e-12
    data = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "he_ochi_convergence.json")))
    grids = [(tc.QMCGrid.lattice(20000, 3, 2, s), tc.QMCGrid.lattice(10000, 3, 2, 1000 + s)) for s in range(8)]
    for row in data["rows"]:
        basis = [tc.OchiBasisFunction(row["alpha"], n, 0, 0) for n in range(row["nBasis"])]
        S, H = tc.one_electron_matrices(basis, 2)
        es = []
        for g1, g2 in grids:
            g1.setBasisFunctions(basis, False)
            g2.setBasisFunctions(basis, False)
            es.append(tc.RCSHartreeFock(2, len(basis)).compute(S, H, g1, tc.computeF(g1, g2), True, tol=1e-11,
                                                                max_iter=200))
        mean, sem = np.mean(es), np.std(es, ddof=1) / np.sqrt(len(es))
        print(f"nBasis {row['nBasis']}: E = {mean:.9f} +/- {sem:.1e}, exact {row['exact_energy']:.9f}, "
              f"reference {row['energy']:.9f} +/- {row['stddev']:.1e}")
        assert abs(mean - row["exact_energy"]) < 4 * sem + 2e-6
        assert abs(mean - row["energy"]) < 2 * np.hypot(row["stddev"], sem)


def test_qmc6d_r12_array_and_overlap_vs_oracle(tc, O, gpu):
    """The per-element QMC route (src/AtomicBasis.hpp): fillR12Array with
    r12MatrixElement (:30-48, 167-177) on the device, over the 6-D point pairs
    of qmc_spherical_points, against the oracle's sum over the same pairs
    (1e-12, s, p and d functions); and overlap (:15-26) by the 3-D QMC route
    against the analytic STO overlaps (QMC error, tolerance in the assertion)."""
    from tcscf_amd import SlaterTypeOrbital as STO
    basis = [STO(0.8, 1, 0, 0), STO(1.7, 1, 0, 0), STO(0.9, 2, 0, 0), STO(1.1, 2, 1, 1), STO(1.1, 2, 1, 0),
             STO(1.3, 3, 2, -2)]
    E = tc.fillR12Array(basis, 30000, 2, 3)
    p1, p2, w = tc.qmc_spherical_points(30000, 6, 2, 3)
    fns = fns_of(basis)
    full = O.eri_qmc6d(p1, p2, w, O.sto_real(p1, fns), O.sto_real(p2, fns))
    ref = O.fill_two_electron(len(basis), True, lambda a, b, c, d: full[a, b, c, d]).real
    assert rel_err(E, ref) < 1e-12
    # the canonical quadruples (a < b, a < c, d > b) hold their own values
    for a, b, c, d in [(0, 1, 2, 3), (0, 2, 1, 4), (1, 2, 3, 5)]:
        assert abs(E[a, b, c, d] - full[a, b, c, d]) < 1e-12 * np.abs(full).max()
    S = tc.qmc_overlap(basis, 30000, 2, 5)
    Sx, _ = tc.one_electron_matrices(basis, 1)
    assert np.max(np.abs(S - Sx)) < 2e-4
