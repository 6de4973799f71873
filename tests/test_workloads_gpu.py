"""The benchmark workloads on the device at their own sizes: TC-HF Be (20
STOs) and Ne (30 STOs) on r1 = r2 = 30,200 / r3 = 9,700 points, the QMC Ne
workload for every Jastrow b, and the 64-instance batch from its committed
fixture.  Reference drivers: src/unitTests/HartreeFockTest.cpp:324-447."""
import json
import os

import numpy as np
import pytest

from conftest import ROOT, rel_err

pytestmark = pytest.mark.gpu


def _grid(tc, spec):
    nrad, order, kind = spec
    return tc.QMCGrid.radial_lebedev(nrad, order, kind, 1.0, 1.0)


def _full(tc, T, nb):
    """Packed (n1, 3, npair) device V -> the oracle's (n1, nb, nb, 3)."""
    return np.ascontiguousarray(tc.unpack(T.double().cpu().numpy(), nb).transpose(0, 2, 3, 1))


@pytest.mark.parametrize("mgs_mode", [0, 1])
def test_tc_be_config3_basis_reduced_grids_vs_oracle(tc, O, gpu, mgs_mode):
    """TC-HF Be with the workload's 20-function basis on reduced grids
    (r1 = r2 = 20 x Lebedev-11, r3 = 12 x Lebedev-7): the device SCF against the
    oracle's on the same F / V (the reference's plain loop, both Gram-Schmidt
    modes), 1e-9 Eh over the whole energy trace."""
    from tcscf_amd import configs
    c = configs.be_config3()
    basis, nb = c["basis"], len(c["basis"])
    g1 = _grid(tc, (20, 11, tc.RADIAL_M4))
    g3 = _grid(tc, (12, 7, tc.RADIAL_M4))
    for g in (g1, g3):
        g.setBasisFunctions(basis, True)
    jas = tc.Jastrow.from_b(1.0)
    Fs = tc.computeF(g1, g1, kernel="tc", jastrow=jas, same_spin=True)
    Fo = tc.computeF(g1, g1, kernel="tc", jastrow=jas, same_spin=False)
    V = tc.computeV(g1, g3, jas)
    S, H = tc.one_electron_matrices(basis, c["Z"])
    hf = tc.TCHartreeFock(c["nup"], c["ndown"], nb)
    e = hf.compute(S, H, g1, Fs, Fo, V, jas, True, tol=1e-10, max_iter=200, mgs_mode=mgs_mode)
    Vf = _full(tc, V, nb)
    ref = O.tc_scf(c["nup"], c["ndown"], S, H, g1.basisValues.cpu().numpy(), g1.basisGradients.cpu().numpy(),
                   g1.weights_host, tc.unpack(Fs.cpu().numpy(), nb), tc.unpack(Fo.cpu().numpy(), nb),
                   np.ascontiguousarray(jas.c_same * Vf), np.ascontiguousarray(jas.c_oppo * Vf), tol=1e-10,
                   max_iter=200, mgs_mode=mgs_mode)
    assert ref["status"] == 0
    assert len(hf.energy_trace) == len(ref["trace"])
    assert np.max(np.abs(hf.energy_trace - ref["trace"])) < 1e-9
    assert abs(e - ref["energy"]) < 1e-9


@pytest.mark.parametrize("atom", ["Be", "Ne"])
def test_tc_config3_full_size(tc, O, gpu, atom):
    """The TC workload at full size (r1 = r2 = 30,200, r3 = 9,700, b = 1):
    FP64 F_same / F_oppo rows and V rows against the oracle on a slice of r1
    (1e-13); one FP64 Fock build at the converged density against the oracle's
    over all 30,200 points (1e-12) and the fockOperator it leaves; the
    mixed-precision run (F pair BF16x5 / BF16x3, V BF16x3, FP32 storage) within
    1e-6 Eh of the FP64 one (same SCF settings: DIIS 6, metric Gram-Schmidt)."""
    import torch
    from tcscf_amd import configs
    c = configs.be_config3() if atom == "Be" else configs.ne_config3()
    basis, nb = c["basis"], len(c["basis"])
    print(f"step: TC {atom} full size", flush=True)
    g1 = _grid(tc, c["grid"])
    g3 = _grid(tc, c["grid3"])
    for g in (g1, g3):
        g.setBasisFunctions(basis, True)
    assert (g1.nGrid(), g3.nGrid()) == (30200, 9700)
    jas = tc.Jastrow.from_b(1.0)
    prm = O.ochi_params(1.0, jas.a12, [(1, 0, 0, jas.c_oppo, jas.c_same)])
    Fs = tc.computeF(g1, g1, kernel="tc", jastrow=jas, same_spin=True)
    Fo = tc.computeF(g1, g1, kernel="tc", jastrow=jas, same_spin=False)
    V = tc.computeV(g1, g3, jas)
    rows = (15000, 15032)
    A1 = g1.basisValues.cpu().numpy()
    A3 = g3.basisValues.cpu().numpy()
    for T, kernel, same, g2, A2 in ((Fs, O.K_TC_SCALAR, True, g1, A1), (Fo, O.K_TC_SCALAR, False, g1, A1)):
        ref = O.compute_F(g1.points_host, g2.points_host, g2.weights_host, A2, kernel=kernel, prm=prm, same=same,
                          rows=rows)
        assert rel_err(tc.unpack(T[rows[0]:rows[1]].cpu().numpy(), nb), ref) < 1e-13
    ref = O.compute_F(g1.points_host, g3.points_host, g3.weights_host, A3, kernel=O.K_JASTROW_GRAD, prm=prm,
                      same=True, rows=rows)  # c_same V
    assert rel_err(jas.c_same * _full(tc, V[rows[0]:rows[1]], nb), ref) < 1e-13
    S, H = tc.one_electron_matrices(basis, c["Z"])
    hf = tc.TCHartreeFock(c["nup"], c["ndown"], nb)
    e64 = hf.compute(S, H, g1, Fs, Fo, V, jas, True, tol=1e-9, max_iter=200, mgs_mode=1, diis=6)
    # the Fock build at the converged density, device against the oracle
    nmax = max(c["nup"], c["ndown"])
    C = np.zeros((2, nb, nmax))
    for s in range(2):
        C[s, :, :hf.nElectrons[s]] = hf.occupiedOrbitals[s]
    G2, G3 = hf.fockBuild(g1, Fs, Fo, V, jas, hf.density, C)
    Vf = _full(tc, V, nb)
    G2o, G3o = O.tc_fock(c["nup"], c["ndown"], A1, g1.basisGradients.cpu().numpy(), g1.weights_host,
                         tc.unpack(Fs.cpu().numpy(), nb), tc.unpack(Fo.cpu().numpy(), nb),
                         np.ascontiguousarray(jas.c_same * Vf), np.ascontiguousarray(jas.c_oppo * Vf), hf.density, C)
    del Vf
    assert rel_err(G2, G2o) < 1e-12
    assert rel_err(G3, G3o) < 1e-12
    mask = (np.abs(H) > 0)[None]
    assert rel_err(hf.fockOperator, H[None] + np.where(mask, G2o, 0) - np.where(mask, G3o, 0)) < 1e-12
    del Fs, Fo, V
    torch.cuda.empty_cache()
    V32 = tc.computeV(g1, g3, jas, precision="bf16x3", out_dtype="fp32")
    for prec in ("bf16x5", "bf16x4"):
        Fs32, Fo32 = tc.computeF_tc_pair(g1, g1, jas, precision=prec, smooth_precision="bf16x3", out_dtype="fp32")
        em = tc.TCHartreeFock(c["nup"], c["ndown"], nb).compute(S, H, g1, Fs32, Fo32, V32, jas, True, tol=1e-9,
                                                                max_iter=200, mgs_mode=1, diis=6)
        del Fs32, Fo32
        print(f"TC {atom} ({prec} F pair): E(fp64) = {e64:.10f}, E(mixed) = {em:.10f}, dE = {em - e64:.2e}")
        assert abs(em - e64) < 1e-6, prec


@pytest.mark.parametrize("b", [0.5, 1.0, 2.0])
def test_tc_qmc_ne_config4_mixed_vs_fp64(tc, gpu, b):
    """The QMC workload at full size (Ne, 40 STOs, 2**16 scrambled-Sobol points
    per coordinate) for each Jastrow b: the BF16x5 and BF16x4 precomputes give
    the FP64 TC energy to 1e-6 Eh (same grids, same SCF).  With this basis the plain
    Roothaan loop oscillates between two states on these grids (so does plain
    RHF), so the solve uses DIIS (depth 6)."""
    from tcscf_amd import configs
    c = configs.ne_qmc_config4()
    assert b in c["b_values"]
    basis, nb = c["basis"], len(c["basis"])
    g1 = tc.QMCGrid.sobol(c["npts"], seed=c["seeds"][0])
    g2 = tc.QMCGrid.sobol(c["npts"], seed=c["seeds"][1])
    g3 = tc.QMCGrid.sobol(c["npts"], seed=c["seeds"][2])
    for g in (g1, g2, g3):
        g.setBasisFunctions(basis, True)
    S, H = tc.one_electron_matrices(basis, c["Z"])
    jas = tc.Jastrow.from_b(b)
    E = {}
    for prec in ("bf16x5", "bf16x4", "fp64"):
        Fs = tc.computeF(g1, g2, kernel="tc", jastrow=jas, same_spin=True, precision=prec)
        Fo = tc.computeF(g1, g2, kernel="tc", jastrow=jas, same_spin=False, precision=prec)
        V = tc.computeV(g1, g3, jas, precision=prec)
        E[prec] = tc.TCHartreeFock(c["nup"], c["ndown"], nb).compute(S, H, g1, Fs, Fo, V, jas, True, tol=1e-9,
                                                                      mgs_mode=1, diis=6)
        del Fs, Fo, V
    print(f"QMC Ne TC b = {b}: E(bf16x5) = {E['bf16x5']:.10f}, E(bf16x4) = {E['bf16x4']:.10f}, "
          f"E(fp64) = {E['fp64']:.10f}")
    assert abs(E["bf16x5"] - E["fp64"]) < 1e-6
    assert abs(E["bf16x4"] - E["fp64"]) < 1e-6


@pytest.mark.parametrize("prec", ["bf16x5", "bf16x4"])
def test_batch64_fixture_contraction(tc, gpu, prec, monkeypatch):
    """The 64-instance serving workload from its committed fixture
    (tests/golden/batch64_instances.json): the one persistent BF16 launch over
    every F / V of every instance equals the single-instance calls bit for bit
    (FP32 storage, as the workload stores them; the single calls with the
    last-round split off, which the batched launch does not do), and each
    tensor is within 3e-6 relative of the FP64 contraction."""
    import torch
    monkeypatch.setenv("TCSCF_BF16_SPLIT", "0")
    from tcscf_amd import batch
    with open(os.path.join(ROOT, "tests", "golden", "batch64_instances.json")) as f:
        instances = json.load(f)
    assert len(instances) == 64
    prepared = batch.prepare(instances)
    specs = batch.contraction_specs(prepared)
    assert len(specs) == 32 + 32 * 3
    outs = tc.computeF_batch(specs, precision=prec, out_dtype="fp32")
    worst = 0.0
    for k, (sp, F) in enumerate(zip(specs, outs)):
        if sp["kernel"] == "grad":
            single = tc.computeV(sp["r1"], sp["r2"], sp["jastrow"], precision=prec, out_dtype="fp32")
            ref = tc.computeV(sp["r1"], sp["r2"], sp["jastrow"])
        else:
            kw = dict(kernel=sp["kernel"], jastrow=sp.get("jastrow"), same_spin=sp.get("same_spin", True))
            single = tc.computeF(sp["r1"], sp["r2"], precision=prec, out_dtype="fp32", **kw)
            ref = tc.computeF(sp["r1"], sp["r2"], **kw)
        assert torch.equal(F, single), k
        err = float((F.double() - ref).abs().max() / ref.abs().max())
        worst = max(worst, err)
        assert err < 3e-6, (k, err)
    print(f"batch64 ({prec}): worst relative tensor error vs FP64 {worst:.2e}")


@pytest.mark.parametrize("prec", ["bf16x5", "bf16x4"])
def test_batch_accurate_policy_subset(tc, gpu, prec):
    """The batch's 'accurate' policy on a subset of the fixture (the three
    lightest and the three heaviest instances): the light ones stay BF16 and
    are within 1e-6 Eh of FP64, the heavy ones (Z >= FP64_FROM_Z) contract in
    FP64 and equal the FP64 run bitwise."""
    import torch
    from tcscf_amd import batch
    with open(os.path.join(ROOT, "tests", "golden", "batch64_instances.json")) as f:
        instances = json.load(f)
    by_z = sorted(instances, key=lambda x: (x["Z"], x["npts"]))
    sub = by_z[:3] + by_z[-3:]
    prep = batch.prepare(sub)
    res_a, t_a = batch.run(prep, prec, synchronize=torch.cuda.synchronize, policy="accurate")
    res_d, _ = batch.run(prep, "fp64", synchronize=torch.cuda.synchronize)
    assert t_a["fp64_instances"] == sum(x["Z"] >= batch.FP64_FROM_Z for x in sub) == 3
    for x, a, d in zip(sub, res_a, res_d):
        assert a["status"] == d["status"] == 0
        if x["Z"] >= batch.FP64_FROM_Z:
            assert a["energy"] == d["energy"]
        else:
            assert abs(a["energy"] - d["energy"]) < 1e-6, (x["Z"], a["energy"] - d["energy"])


def test_batch_concurrent_solves_match_one_at_a_time(tc, gpu):
    """The batch's SCF half on four host threads, each on a stream of its own
    (the drivers' per-stream resources, direct launches), gives every instance
    bitwise the result of the same solves one after another on one stream, and
    of a solve alone with the default per-iteration graph.  Twelve instances of
    the fixture: RCS, UOS and TC, open and closed shells, nb from 7 to 53."""
    import torch
    from tcscf_amd import batch
    with open(os.path.join(ROOT, "tests", "golden", "batch64_instances.json")) as f:
        instances = json.load(f)
    sub = [x for x in instances if x["npts"] <= 5500][:12]
    assert {x["method"] for x in sub} >= {"tc", "hf"} and max(x["nb"] for x in sub) > 32
    prep = batch.prepare(sub)
    specs = batch.contraction_specs(prep)
    tensors = tc.computeF_batch(specs, precision="bf16x5", out_dtype="fp32")
    torch.cuda.synchronize()
    items, k = [], 0
    for p in prep:
        x = p["inst"]
        d = dict(nup=x["nup"], ndown=x["ndown"], S=p["S"], H=p["H"], r1=p["grid"])
        if x["method"] == "tc":
            d.update(kind="tc", Fs=tensors[k], Fo=tensors[k + 1], V=tensors[k + 2], jastrow=p["jastrow"])
            k += 3
        else:
            d.update(kind="rcs" if x["nup"] == x["ndown"] else "uos", F=tensors[k])
            k += 1
        items.append(d)
    par = batch.solve_each(items, tol=1e-8, max_iter=60, diis=6, workers=4)
    seq = batch.solve_each(items, tol=1e-8, max_iter=60, diis=6, workers=1)
    for x, a, b in zip(items, par, seq):
        assert a["status"] == b["status"] and a["iterations"] == b["iterations"]
        assert a["energy"] == b["energy"] and np.array_equal(a["trace"], b["trace"]) and np.array_equal(a["D"], b["D"])
    # the default (graph) path of a TC solve alone gives the same bits
    q = next(i for i, x in enumerate(items) if x["kind"] == "tc")
    x = items[q]
    hf = tc.TCHartreeFock(x["nup"], x["ndown"], x["r1"].nBasis())
    e = hf.compute(x["S"], x["H"], x["r1"], x["Fs"], x["Fo"], x["V"], x["jastrow"], True, tol=1e-8, max_iter=60,
                   mgs_mode=1, diis=6)
    assert hf.graph_iterations == hf.numberOfConvergenceLoops()
    assert e == par[q]["energy"] and np.array_equal(hf.energy_trace, par[q]["trace"])
