"""Benchmark: Ne RHF (30 STOs, 30,200-point grid) SCF on one MI355X.

One step = one full SCF solve on inputs already resident in HBM: STO basis
tabulation, the fused kernel-generation + contraction F = Kw . P (every
(r1, r2) grid pair, 1/r12 kernel), and the RHF iterations to convergence.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision bf16x4|bf16x5|bf16x3|fp32|fp64]

Prints one JSON line.  `value` is grid-pair contractions per second over the
whole step (pairs x kernel components / step time); the FP64 drop-in path is
timed beside it for the speed-up and the energy deviation, and the FP64 CPU
oracle (OpenMP, this host's cores) is timed on a bounded slice of the same
contraction for `cpu_baseline`.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "SCF wall time + grid-pair contractions/s: mixed-prec MFMA vs FP64 GPU @1e-6 Eh"
PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3, "bf16x3": 2500.0, "bf16x4": 2500.0, "bf16x5": 2500.0}
PRODUCTS = {"fp64": 1, "fp32": 1, "bf16x3": 3, "bf16x4": 4, "bf16x5": 5}
# HBM bytes per computeF of the contraction kernels on this workload (BF16: the
# full-round launch, the split last round and its fix-up together; BF16x4 with
# the FP32 F storage of the mixed RHF leg, BF16x5 and FP64 with FP64 F), from
# rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs of `bench.py --pmc-probe`, each
# scaled by the reading for a 512 MiB copy (FETCH x2.0, WRITE x1.0);
# profiles/bf16_fgemm_ne30_traffic.md
TRAFFIC_BYTES = {"bf16x4": 2.205e9 + 1.122e8, "bf16x5": 2.207e9 + 1.694e8, "fp64": 1.016e9 + 1.156e8}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--precision", default="bf16x4", choices=["bf16x4", "bf16x5", "bf16x3", "fp32", "fp64"])
    p.add_argument("--fp64-steps", type=int, default=2)
    p.add_argument("--storage", default="fp32", choices=["fp64", "fp32"],
                   help="F storage of the mixed-precision RHF leg (the FP64 reference leg stores FP64); FP32 "
                        "halves the sweep's bytes, the SCF arithmetic stays FP64")
    p.add_argument("--tol", type=float, default=1e-9)
    p.add_argument("--cpu-rows", type=int, default=0, help="rows of the CPU-baseline slice (0 = auto)")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--no-tc", action="store_true", help="skip the TC-HF Ne side measurement")
    p.add_argument("--tc-steps", type=int, default=5, help="timed TC-HF Ne solves per precision")
    p.add_argument("--tc-mgs", type=int, default=1, help="TC Gram-Schmidt: 0 reference, 1 metric")
    p.add_argument("--tc-diis", type=int, default=6, help="TC DIIS depth (the TC workload asks for 6; 0 = plain)")
    p.add_argument("--tc-smooth-precision", default="bf16x3", choices=["bf16x4", "bf16x5", "bf16x3", "fp32", "fp64"],
                   help="mixed TC: precision of the smooth 1/(r12 + a12)^4 part of F_same / F_oppo")
    p.add_argument("--tc-v-precision", default="bf16x3", choices=["bf16x4", "bf16x5", "bf16x3", "fp32", "fp64"],
                   help="mixed TC: precision of the Jastrow-gradient tensor V")
    p.add_argument("--no-batch", action="store_true", help="skip the batched 64-instance side measurement")
    p.add_argument("--no-qmc", action="store_true", help="skip the QMC-grid TC-HF Ne side measurement")
    p.add_argument("--qmc-repeats", type=int, default=5,
                   help="QMC workload: independently scrambled grid sets per Jastrow b (the error bar)")
    p.add_argument("--batch-precision", default="bf16x5", choices=["bf16x4", "bf16x5", "bf16x3", "fp32"],
                   help="the batch's mixed precision (BF16x5: its small grids average the dropped Kl*Pm term "
                        "less, and its tile does not gain from four products)")
    p.add_argument("--pmc-probe", action="store_true",
                   help="for rocprofv3 --pmc runs: BF16x5, BF16x4 and FP64 computeF on the Ne grid, one RHF "
                        "solve, the TC-Ne mixed contractions, then a 512 MiB device copy whose counter reading "
                        "calibrates the byte counters")
    return p.parse_args()


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this bench measures one GPU; a multi-GPU run would shard the r1 rows "
                                   "(one all-reduce of nb^2 values per SCF iteration), which is not built",
                          "n_gpus": a.gpus}))
        sys.exit(2)
    import torch
    import tcscf_amd as tc
    from tcscf_amd import configs

    assert torch.cuda.is_available(), "bench.py needs a ROCm GPU"
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    cfg = configs.ne_config2()
    if a.pmc_probe:
        pmc_probe(tc, cfg, torch)
        return
    nrad, order, kind = cfg["grid"]
    basis = cfg["basis"]
    nb = len(basis)
    npair = tc.npair(nb)
    grid = tc.QMCGrid.radial_lebedev(nrad, order, kind, 1.0, 1.0)
    S, H = tc.one_electron_matrices(basis, cfg["Z"])
    n = grid.nGrid()
    pairs = n * n - n  # coincident pairs excluded
    stream = torch.cuda.current_stream()

    def step(precision, ev=None):
        grid.setBasisFunctions(basis, False)
        if ev:
            ev[0].record(stream)
        F = tc.computeF(grid, grid, precision=precision,
                        out_dtype=a.storage if precision != "fp64" else "fp64")
        if ev:
            ev[1].record(stream)
        torch.cuda.synchronize()  # the SCF's own sweep timers then see the sweeps only
        hf = tc.RCSHartreeFock(cfg["nup"] + cfg["ndown"], nb)
        e = hf.compute(S, H, grid, F, True, tol=a.tol, max_iter=100)
        return e, hf

    def run(precision, steps, warmup):
        for _ in range(warmup):
            step(precision)
        torch.cuda.synchronize()
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
        t0 = time.perf_counter()
        out = None
        for k in range(steps):
            out = step(precision, evs[k])
        torch.cuda.synchronize()
        t = (time.perf_counter() - t0) / steps
        kern_ms = sum(e0.elapsed_time(e1) for e0, e1 in evs) / steps
        return t, kern_ms, out

    t_mix, k_mix, (e_mix, hf_mix) = run(a.precision, a.steps, a.warmup)
    if a.precision != "fp64":
        t64, k64, (e64, hf64) = run("fp64", a.fp64_steps, 1)
    else:
        t64, k64, e64, hf64 = t_mix, k_mix, e_mix, hf_mix

    flops = 2.0 * npair * pairs  # algorithmic: one MAC per pair density per grid pair
    fbytes32 = a.storage == "fp32" and a.precision != "fp64"
    achieved = flops / (k_mix * 1e-3) / 1e12
    peak = PEAK_TFLOPS[a.precision]
    roofline = {"bound": "mfma", "achieved": round(achieved, 2), "peak": peak, "unit": "TFLOP/s",
                "frac": round(achieved / peak, 4), "traffic": TRAFFIC_BYTES.get(a.precision),
                "traffic_unit": "bytes per computeF (HBM, rocprofv3 FETCH_SIZE/WRITE_SIZE, copy-calibrated)",
                "issued_frac": round(achieved * PRODUCTS[a.precision] / peak, 4),
                "kernel": "fgemm_bf16_pipe (fused 1/r12 generation + contraction: full rounds, split last round, "
                          "fix-up), computeF wall incl. packing",
                "kernel_ms": round(k_mix, 4)}

    blocks = block_errors(tc, grid, S, H, cfg, a, hf_mix, hf64, e_mix, e64) if a.precision != "fp64" else None

    tc_res = None
    if not a.no_tc:
        tc_res = run_tc(a, tc, configs, torch)
    batch_res = None
    if not a.no_batch:
        batch_res = run_batch(a, tc, configs, torch)
    qmc_res = None
    if not a.no_qmc:
        qmc_res = run_qmc(a, tc, configs, torch)

    cpu = None
    if not a.no_cpu_baseline:
        from oracle import oracle as O
        O.set_threads(int(os.environ.get("OMP_NUM_THREADS", 0)) or len(os.sched_getaffinity(0)) or 1)
        threads = O.get_threads()
        pts = grid.points_host
        A = grid.basisValues.cpu().numpy()

        def cpu_rows(rows):
            t0 = time.perf_counter()
            O.compute_F(pts, pts, grid.weights_host, A, rows=(0, rows))
            return time.perf_counter() - t0

        if a.cpu_rows:
            rows = a.cpu_rows
        else:  # size the sample to 10-20 s of CPU work
            probe = max(8, 2 * threads)
            rows = int(min(n, max(probe, probe * 20.0 / max(cpu_rows(probe), 1e-3))))
        tc_cpu = cpu_rows(rows)
        cpu = {"value": rows * n / tc_cpu, "unit": "grid-pair contractions/s", "cores": threads, "kind": "port",
               "cpu_model": cpu_model(),
               "sample": f"oracle computeF (FP64, OpenMP, {threads} threads) on r1 rows 0..{rows} x all {n} r2 "
                         f"points, nb={nb}: {tc_cpu:.2f} s; the full precompute extrapolates to "
                         f"{tc_cpu * n / rows:.0f} s"}

    res = {
        "metric": METRIC,
        "value": pairs / t_mix,
        "unit": "grid-pair contractions/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": t_mix * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"bf16x4": "bf16 split (K 2 terms, P 3 terms, 4 products: Kl*Pm dropped), f32 MFMA per two 32-pt steps, "
                            "f64 sums + f64 SCF",
                  "bf16x5": "bf16 split (K 2 terms, P 3 terms, 5 products), f32 MFMA per 32-pt step, f64 sums + f64 SCF",
                  "bf16x3": "bf16 split (K 2 terms, P 2 terms, 3 products), f32 MFMA per 32-pt step, f64 sums + f64 SCF",
                  "fp32": "f32 MFMA, f64 cross-chunk + f64 SCF", "fp64": "f64"}[a.precision],
        "data": "synthetic (even-tempered STO basis, deterministic radial x Lebedev grid)",
        "config": {"workload": f"Ne RHF, {nb} STOs (13s,4p,1d even-tempered 0.3*1.8^k), {n}-point grid "
                               f"({nrad} Gauss-Chebyshev/M4 radial x Lebedev-{order}), 1/r12, tol {a.tol:g}",
                   "precision": a.precision, "storage": a.storage, "grid_points": n, "nbasis": nb, "npair": npair},
        "scf_wall_ms": t_mix * 1e3,
        "scf_iterations": hf_mix.numberOfConvergenceLoops(),
        "energy": e_mix,
        "fp64_scf_wall_ms": t64 * 1e3,
        "fp64_energy": e64,
        "energy_dev_vs_fp64": e_mix - e64,
        "speedup_vs_fp64": t64 / t_mix,
        "computeF_ms": {a.precision: k_mix, "fp64": k64},
        "sweep_ms_per_scf": hf_mix.t_sweep_ms,
        "solve_ms_per_scf": hf_mix.t_solve_ms,
        "block_errors_vs_fp64": blocks,
        "sweep_roofline": hbm_roofline(n * npair * (4 if fbytes32 else 8) + n * nb * 8,
                                       hf_mix.t_sweep_ms / max(1, hf_mix.numberOfConvergenceLoops()),
                                       "%s: one read of the packed %s F and of A per iteration; time per "
                                       "iteration as the SCF measures it (D upload, sweep, reduction, G readback)"
                                       % ("hf_sweep_col_kernel" if nb <= 32 else "hf_sweep_kernel",
                                          "FP32" if fbytes32 else "FP64")),
        "roofline": roofline,
        "cpu_baseline": cpu,
        "tc_ne": tc_res,
        "batch64": batch_res,
        "qmc_ne": qmc_res,
    }
    print(json.dumps(res))


def rel_max(x, ref):
    """max |x - ref| / max |ref| over a block."""
    return float((x.double() - ref.double()).abs().max() / ref.double().abs().max())


def block_errors(tc, grid, S, H, cfg, a, hf_mix, hf64, e_mix, e64):
    """Error of the mixed-precision leg against FP64 per integral block (F, and
    the grid ERI integrateAllR1R2 formed from each F), per orbital energy and
    for the total energy (outside the timed region)."""
    F_mix = tc.computeF(grid, grid, precision=a.precision, out_dtype=a.storage)
    F64 = tc.computeF(grid, grid, precision="fp64")
    E_mix = tc.integrateAllR1R2(grid, F_mix.double())
    E64 = tc.integrateAllR1R2(grid, F64)
    nocc = cfg["nup"]
    de = np.abs(hf_mix.eigenvalues - hf64.eigenvalues)
    return {"F_max_rel": rel_max(F_mix, F64), "eri_max_rel": rel_max(E_mix, E64),
            "orbital_energy_max_abs": float(de.max()), "occupied_orbital_energy_max_abs": float(de[:nocc].max()),
            "energy": e_mix - e64}


def hbm_roofline(nbytes, ms, note):
    """Achieved HBM bandwidth of a per-iteration pass: algorithmic bytes / time."""
    gbps = nbytes / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    return {"bound": "hbm", "algorithmic_bytes": nbytes, "ms": round(ms, 4), "achieved": round(gbps, 1),
            "peak": 8000.0, "unit": "GB/s", "frac": round(gbps / 8000.0, 4), "note": note}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def run_tc(a, tc, configs, torch):
    """Transcorrelated HF on Ne (30 STOs, b = 1, c_same = c_oppo / 2): r1 = r2 =
    30,200 points, r3 = 50 x Lebedev-194 = 9,700 points; per precision one
    untimed solve (workspaces, code objects) then --tc-steps timed solves,
    each precompute (F_same, F_oppo, V) + SCF.  The FP64 path computes F_same
    and F_oppo with two computeF calls as the reference's driver does; the
    mixed path computes them together from the 1/r12 contraction (at
    --precision) and the contraction of the smooth part 1/(r12 + a12)^4 (at
    --tc-smooth-precision), V at --tc-v-precision, all three stored in FP32
    (the Fock build's arithmetic is FP64).  The F and V contractions are timed
    with HIP events on the launching stream for their MFMA roofline."""
    cfg = configs.ne_config3()
    basis = cfg["basis"]
    nb = len(basis)
    npair = tc.npair(nb)
    g1 = tc.QMCGrid.radial_lebedev(*cfg["grid"], 1.0, 1.0)
    g3 = tc.QMCGrid.radial_lebedev(*cfg["grid3"], 1.0, 1.0)
    S, H = tc.one_electron_matrices(basis, cfg["Z"])
    jas = tc.Jastrow.from_b(1.0)
    mixed = a.precision != "fp64"
    stream = torch.cuda.current_stream()
    n1, n3 = g1.nGrid(), g3.nGrid()

    def tensors(prec, ev):
        """(F_same, F_oppo, V); ev: HIP events around the F and the V launches."""
        ev[0].record(stream)
        if prec == "fp64":
            Fs = tc.computeF(g1, g1, kernel="tc", jastrow=jas, same_spin=True, precision="fp64")
            Fo = tc.computeF(g1, g1, kernel="tc", jastrow=jas, same_spin=False, precision="fp64")
        else:
            Fs, Fo = tc.computeF_tc_pair(g1, g1, jas, precision=prec, smooth_precision=a.tc_smooth_precision,
                                         out_dtype="fp32")
        ev[1].record(stream)
        V = tc.computeV(g1, g3, jas, precision="fp64" if prec == "fp64" else a.tc_v_precision,
                        out_dtype="fp64" if prec == "fp64" else "fp32")
        ev[2].record(stream)
        return Fs, Fo, V

    def solve(prec):
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        g1.setBasisFunctions(basis, True)
        g3.setBasisFunctions(basis, True)
        Fs, Fo, V = tensors(prec, ev)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        hf = tc.TCHartreeFock(cfg["nup"], cfg["ndown"], nb)
        try:
            e = hf.compute(S, H, g1, Fs, Fo, V, jas, True, tol=a.tol, max_iter=200, mgs_mode=a.tc_mgs,
                           diis=a.tc_diis)
            status = "ok"
        except tc.TcscfError as err:
            e, status = float("nan"), str(err)
        t2 = time.perf_counter()
        return dict(e=e, status=status, pre=(t1 - t0) * 1e3, scf=(t2 - t1) * 1e3, tot=(t2 - t0) * 1e3,
                    F_ms=ev[0].elapsed_time(ev[1]), V_ms=ev[1].elapsed_time(ev[2]), hf=hf), (Fs, Fo, V)

    out, keep = {}, {}
    legs = (a.precision, "fp64") if mixed else ("fp64",)
    for prec in legs:
        solve(prec)  # warm-up
        runs = []
        for k in range(max(1, a.tc_steps)):
            r, T = solve(prec)
            runs.append(r)
            if k == 0:
                keep[prec] = (r["hf"], [x.clone() for x in T] if prec == a.precision else T)
            del T
        hf = runs[-1]["hf"]
        tot = np.array([r["tot"] for r in runs])
        fb_ms = hf.t_sweep_ms / max(1, hf.numberOfConvergenceLoops())
        esz = 8 if prec == "fp64" else 4
        fb_bytes = 5 * n1 * npair * esz + 4 * n1 * nb * 8  # F_same, F_oppo, V (3) + A, dA (3), once each
        F_ms = float(np.mean([r["F_ms"] for r in runs]))
        V_ms = float(np.mean([r["V_ms"] for r in runs]))
        # algorithmic work: F_same and F_oppo are 2 kernel components over the r1 x r2 pairs, V 3 over r1 x r3
        fl_F, fl_V = 2 * 2.0 * npair * n1 * (n1 - 1), 3 * 2.0 * npair * n1 * n3
        peak = PEAK_TFLOPS["fp64" if prec == "fp64" else "bf16x5"]
        ach = (fl_F + fl_V) / ((F_ms + V_ms) * 1e-3) / 1e12
        out[prec] = {"energy": runs[-1]["e"], "status": runs[-1]["status"], "steps": len(runs),
                     "total_ms": float(tot.mean()), "total_ms_min": float(tot.min()), "total_ms_std": float(tot.std()),
                     "precompute_ms": float(np.mean([r["pre"] for r in runs])),
                     "scf_ms": float(np.mean([r["scf"] for r in runs])),
                     "F_ms": F_ms, "V_ms": V_ms, "iterations": hf.numberOfConvergenceLoops(), "fock_build_ms": fb_ms,
                     "scf_iteration_graph": {"iterations_replayed": hf.graph_iterations,
                                             "captures_in_last_solve": hf.graph_captures},
                     "precompute_roofline": {"bound": "mfma", "achieved": round(ach, 2), "peak": peak,
                                             "unit": "TFLOP/s", "frac": round(ach / peak, 4),
                                             "algorithmic_flop": fl_F + fl_V,
                                             "kernel": "F pair (F_same, F_oppo) + V contractions, HIP events on "
                                                       "the launching stream"},
                     "fock_build_roofline": hbm_roofline(fb_bytes, fb_ms, "TC Fock build (point, rank, V-linear "
                                                         "passes): one read of F_same, F_oppo, V, A and dA per "
                                                         "iteration; time per iteration as the SCF measures it")}
    res = {"workload": f"TC-HF Ne, {nb} STOs, b=1, r1=r2={n1}, r3={n3}, MGS {'metric' if a.tc_mgs else 'reference'}, "
                       f"DIIS {a.tc_diis}",
           "pairs_x_components": 2 * n1 * (n1 - 1) + 3 * n1 * n3, "fp64": out["fp64"]}
    if mixed:
        mix, f64 = out[a.precision], out["fp64"]
        hm, (Fs, Fo, V) = keep[a.precision]
        h64, (Fs64, Fo64, V64) = keep["fp64"]
        ps = np.concatenate(hm.occupiedOrbitalPseudoEnergy) - np.concatenate(h64.occupiedOrbitalPseudoEnergy)
        blocks = {"F_same_max_rel": rel_max(Fs, Fs64), "F_oppo_max_rel": rel_max(Fo, Fo64),
                  "V_max_rel": rel_max(V, V64), "pseudo_energy_max_abs": float(np.abs(ps).max()),
                  "energy": mix["energy"] - f64["energy"]}
        del Fs, Fo, V, Fs64, Fo64, V64
        # the FP64 path with F_same / F_oppo from the pair form as well (F only, for the second ratio)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        Fs, Fo = tc.computeF_tc_pair(g1, g1, jas, precision="fp64")
        torch.cuda.synchronize()
        pair64 = (time.perf_counter() - t0) * 1e3
        del Fs, Fo
        res.update({a.precision: mix, "mixed_tensors": {"F_coulomb": a.precision, "F_smooth": a.tc_smooth_precision,
                                                        "V": a.tc_v_precision, "storage": "fp32"},
                    "block_errors_vs_fp64": blocks,
                    "energy_dev_vs_fp64": mix["energy"] - f64["energy"],
                    "speedup_vs_fp64": f64["total_ms"] / mix["total_ms"],
                    "fp64_pair_F_ms": pair64,
                    "speedup_vs_fp64_with_pair_F": (f64["total_ms"] - f64["F_ms"] + pair64) / mix["total_ms"]})
    return res


def run_qmc(a, tc, configs, torch):
    """The QMC workload: TC-HF Ne (40 STOs) on scrambled-Sobol grids of 2**16
    points per electron coordinate, Jastrow b in {0.5, 1, 2}.  Per b,
    --qmc-repeats independently scrambled grid sets (r1, r2, r3 seeds 3k + 1,
    3k + 2, 3k + 3): the mixed precompute (as the TC-Ne leg: F pair at
    --precision / --tc-smooth-precision, V at --tc-v-precision, FP32 storage)
    and the SCF (DIIS 6, metric MGS), whose energies give the QMC error bar
    (mean +- standard deviation over the grid sets, as the reference's driver
    reports it); and on the first grid set the FP64 precompute and SCF, whose
    energy gives the FP64 deviation.  A grid set whose SCF does not converge
    is drawn again (the next seeds), at most three times in a row."""
    import numpy as np
    cfg = configs.ne_qmc_config4()
    basis, nb = cfg["basis"], len(cfg["basis"])
    S, H = tc.one_electron_matrices(basis, cfg["Z"])

    def solve(jas, k, prec):
        g = [tc.QMCGrid.sobol(cfg["npts"], seed=3 * k + q) for q in (1, 2, 3)]
        for x in g:
            x.setBasisFunctions(basis, True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        if prec == "fp64":
            Fs = tc.computeF(g[0], g[1], kernel="tc", jastrow=jas, same_spin=True, precision="fp64")
            Fo = tc.computeF(g[0], g[1], kernel="tc", jastrow=jas, same_spin=False, precision="fp64")
            V = tc.computeV(g[0], g[2], jas, precision="fp64")
        else:
            Fs, Fo = tc.computeF_tc_pair(g[0], g[1], jas, precision=prec, smooth_precision=a.tc_smooth_precision,
                                         out_dtype="fp32")
            V = tc.computeV(g[0], g[2], jas, precision=a.tc_v_precision, out_dtype="fp32")
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        hf = tc.TCHartreeFock(cfg["nup"], cfg["ndown"], nb)
        try:
            e = hf.compute(S, H, g[0], Fs, Fo, V, jas, True, tol=a.tol, max_iter=200, mgs_mode=1, diis=6)
        except tc.ConvergenceError:
            e = None
        t2 = time.perf_counter()
        del Fs, Fo, V
        return e, (t1 - t0) * 1e3, (t2 - t1) * 1e3, hf.numberOfConvergenceLoops()

    out = {"workload": f"TC-HF Ne, {nb} STOs, scrambled-Sobol grids of {cfg['npts']} points per coordinate "
                       f"(r1, r2, r3), DIIS 6, metric MGS",
           "mixed_tensors": {"F_coulomb": a.precision, "F_smooth": a.tc_smooth_precision, "V": a.tc_v_precision,
                             "storage": "fp32"}}
    solve(tc.Jastrow.from_b(1.0), 0, a.precision)  # warm-up (workspaces, code objects)
    for b in cfg["b_values"]:
        jas = tc.Jastrow.from_b(b)
        energies, pre, scf, its, k, fails, k0, sets = [], [], [], [], 0, 0, None, []
        while len(energies) < a.qmc_repeats and fails < 3:
            e, tp, ts, n_it = solve(jas, k, a.precision)
            k += 1
            if e is None:
                fails += 1
                continue
            fails = 0
            if k0 is None:
                k0 = k - 1
            sets.append(k - 1)
            energies.append(e)
            pre.append(tp)
            scf.append(ts)
            its.append(n_it)
        e64, tp64, ts64, _ = solve(jas, k0 or 0, "fp64")  # the first grid set that converged
        E = np.array(energies)
        # the grid set farthest from the median, also in FP64: whether its energy
        # is the grid's (an SCF state of that quadrature) or the precision's
        far = None
        if len(E) > 2:
            j = int(np.argmax(np.abs(E - np.median(E))))
            if j > 0:
                ef, _, _, _ = solve(jas, sets[j], "fp64")
                far = {"grid_set": sets[j], "energy": float(E[j]), "fp64_energy": ef,
                       "energy_dev_vs_fp64": (float(E[j]) - ef) if ef is not None else None}
        out[f"b={b:g}"] = {
            "energies": [float(x) for x in E],
            "energy_mean": float(E.mean()) if len(E) else None,
            "energy_median": float(np.median(E)) if len(E) else None,
            "qmc_error_bar_std": float(E.std(ddof=1)) if len(E) > 1 else None,
            "grid_sets": len(E), "grid_sets_redrawn": k - len(E),
            "mixed_ms_per_solve": {"precompute": float(np.mean(pre)) if pre else None,
                                   "scf": float(np.mean(scf)) if scf else None,
                                   "iterations": float(np.mean(its)) if its else None},
            "fp64_ms_per_solve": {"precompute": tp64, "scf": ts64},
            "fp64_energy_first_grid_set": e64,
            "energy_dev_vs_fp64_first_grid_set": (energies[0] - e64) if energies and e64 is not None else None,
            "speedup_vs_fp64": (tp64 + ts64) / (pre[0] + scf[0]) if pre else None,
            "farthest_from_median": far,
        }
    return out


def run_batch(a, tc, configs, torch):
    """The batched-serving workload: 64 uneven instances (Z 2..18, atoms and
    cations, 7..53 STOs, 1,250..30,200 points, half HF / half TC with b in
    [0.5, 2]) from a fixed seed.  One batched contraction call (a persistent
    launch for BF16) and the SCF of every instance (DIIS 6, tol 1e-8, at most
    60 iterations) per precision; the FP64 run gives the per-instance error."""
    import numpy as np
    from tcscf_amd import batch
    inst = configs.batch_instances()
    prep = batch.prepare(inst)
    torch.cuda.synchronize()
    prec = a.batch_precision
    batch.run(prep, prec, synchronize=torch.cuda.synchronize)  # warm-up (and the tile autotuner)
    res_m, t_m = batch.run(prep, prec, synchronize=torch.cuda.synchronize)
    res_d, t_d = batch.run(prep, "fp64", synchronize=torch.cuda.synchronize)
    res_a, t_a = batch.run(prep, prec, synchronize=torch.cuda.synchronize, policy="accurate")

    def errors(res):
        both = [(m, d) for m, d in zip(res, res_d) if m["status"] == 0 and d["status"] == 0]
        err = np.array([abs(m["energy"] - d["energy"]) for m, d in both])
        rel = np.array([abs(m["energy"] - d["energy"]) / abs(d["energy"]) for m, d in both])
        return {"max": float(err.max()), "median": float(np.median(err)), "max_relative": float(rel.max()),
                "instances": len(both), "within_1e-6": int((err <= 1e-6).sum())}

    units = batch.pair_units(prep)
    return {"workload": f"{len(inst)} instances, seed {configs.BATCH_SEED}: Z 2..18 atoms/cations, "
                        f"{min(x['nb'] for x in inst)}..{max(x['nb'] for x in inst)} STOs, "
                        f"{min(x['npts'] for x in inst)}..{max(x['npts'] for x in inst)} points, "
                        f"{sum(x['method'] == 'tc' for x in inst)} TC",
            "pair_units": units,
            "value": units / (t_m["total_ms"] * 1e-3), "unit": "grid-pair contractions/s",
            "precision": prec, prec: t_m, "fp64": t_d, "speedup_vs_fp64": t_d["total_ms"] / t_m["total_ms"],
            "converged": {prec: sum(r["status"] == 0 for r in res_m), "fp64": sum(r["status"] == 0 for r in res_d)},
            "energy_err_vs_fp64": errors(res_m),
            # every instance within 1e-6 Eh: FP64 contractions from Z = batch.FP64_FROM_Z up
            "accurate_policy": {"fp64_from_Z": batch.FP64_FROM_Z, "fp64_instances": t_a["fp64_instances"],
                                "total_ms": t_a["total_ms"], "value": units / (t_a["total_ms"] * 1e-3),
                                "speedup_vs_fp64": t_d["total_ms"] / t_a["total_ms"],
                                "energy_err_vs_fp64": errors(res_a)},
            "tile_rows_autotuned": tc.lib().tcscf_batch_tuned_rows(tc._lib.PRECISIONS[prec])
            if prec.startswith("bf16") else None}


def pmc_probe(tc, cfg, torch):
    """Kernels to read hardware counters on: the contraction at BF16x4,
    BF16x5 and FP64, one RHF solve (the HBM-bound sweep), the TC-Ne mixed
    contractions (F pair at BF16x4 / BF16x3, V at BF16x3, FP32 storage) and a
    plain 512 MiB copy of known traffic (read 512 MiB, write 512 MiB) to scale
    FETCH_SIZE/WRITE_SIZE."""
    g = tc.QMCGrid.radial_lebedev(*cfg["grid"], 1.0, 1.0)
    g.setBasisFunctions(cfg["basis"], True)
    F = tc.computeF(g, g, precision="bf16x5")
    torch.cuda.synchronize()
    F = tc.computeF(g, g, precision="bf16x4", out_dtype="fp32")  # as the bench's mixed RHF leg stores it
    torch.cuda.synchronize()
    F64 = tc.computeF(g, g, precision="fp64")
    torch.cuda.synchronize()
    S, H = tc.one_electron_matrices(cfg["basis"], cfg["Z"])
    e = tc.RCSHartreeFock(10, len(cfg["basis"])).compute(S, H, g, F, True, tol=1e-9)
    g3 = tc.QMCGrid.radial_lebedev(50, 23, cfg["grid"][2], 1.0, 1.0)
    g3.setBasisFunctions(cfg["basis"], True)
    jas = tc.Jastrow.from_b(1.0)
    tc.computeF_tc_pair(g, g, jas, precision="bf16x4", smooth_precision="bf16x3", out_dtype="fp32")
    tc.computeV(g, g3, jas, precision="bf16x3", out_dtype="fp32")
    torch.cuda.synchronize()
    x = torch.empty(64 * 2 ** 20, dtype=torch.float64, device="cuda").fill_(1.0)
    y = torch.empty_like(x)
    y.copy_(x)
    torch.cuda.synchronize()
    del F64
    print(json.dumps({"pmc_probe": "done", "energy": e}), flush=True)


if __name__ == "__main__":
    main()
