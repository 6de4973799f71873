"""The batched-serving workload end to end: grids and bases for every instance,
one batched contraction launch for all F / V tensors, then the SCF of every
instance."""
import time

from . import api, configs
from ._lib import TcscfError


def prepare(instances):
    """Device grids (basis values, gradients for TC) and S / H per instance."""
    out = []
    for x in instances:
        basis = configs.batch_basis(x)
        nrad, order = x["grid"]
        g = api.QMCGrid.radial_lebedev(nrad, order, configs._lib.RADIAL_M4, 1.0, 1.0)
        g.setBasisFunctions(basis, x["method"] == "tc")
        S, H = api.one_electron_matrices(basis, x["Z"])
        out.append(dict(inst=x, grid=g, S=S, H=H, jastrow=api.Jastrow.from_b(x["b"]) if x["b"] else None))
    return out


def contraction_specs(prepared):
    """One spec per F / V tensor: HF needs F; TC needs F_same, F_oppo and V."""
    specs = []
    for p in prepared:
        g = p["grid"]
        if p["inst"]["method"] == "tc":
            j = p["jastrow"]
            specs += [dict(r1=g, r2=g, kernel="tc", jastrow=j, same_spin=True),
                      dict(r1=g, r2=g, kernel="tc", jastrow=j, same_spin=False),
                      dict(r1=g, r2=g, kernel="grad", jastrow=j)]
        else:
            specs.append(dict(r1=g, r2=g, kernel="coulomb"))
    return specs


def pair_units(prepared):
    """Grid-pair contractions (pairs x distinct kernel components, coincident
    pairs excluded): 1 per pair for HF, 2 (F_same / F_oppo combine linearly) + 3
    (V) for TC, as the workload's algorithmic count."""
    total = 0
    for p in prepared:
        n = p["grid"].nGrid()
        total += (n * n - n) * (5 if p["inst"]["method"] == "tc" else 1)
    return total


# Atomic number from which the accuracy policy contracts in FP64: an FP32-class
# contraction (BF16 split or FP32 products, relative element error ~1e-7)
# moves the SCF energy by an amount that grows with the two-electron energy;
# on the 64-instance fixture every instance above 1e-6 Eh has Z >= 8 (the
# largest below: 7.7e-7 Eh at Z = 7).
FP64_FROM_Z = 8


def instance_precision(inst, precision, policy):
    """The contraction precision of one instance: `precision`, or under the
    'accurate' policy FP64 for Z >= FP64_FROM_Z."""
    if policy == "accurate" and inst["Z"] >= FP64_FROM_Z:
        return "fp64"
    return precision


def run(prepared, precision, tol=1e-8, max_iter=60, diis=6, synchronize=None, out_dtype=None, policy=None,
        workers=4):
    """Precompute (one batched call per contraction precision) + batched SCF.
    policy None: every instance at `precision`; 'accurate': FP64 for the
    instances instance_precision sends there.  F / V storage: FP32 for the BF16
    contractions (the Fock builds keep FP64 arithmetic), FP64 otherwise,
    unless out_dtype says.  workers: host threads (streams) of the SCF half.
    Returns (results, timings dict)."""
    specs = contraction_specs(prepared)
    owner = [k for k, p in enumerate(prepared) for _ in range(3 if p["inst"]["method"] == "tc" else 1)]
    precs = [instance_precision(prepared[k]["inst"], precision, policy) for k in owner]
    t0 = time.perf_counter()
    tensors = [None] * len(specs)
    for prec in dict.fromkeys(precs):  # each distinct precision once, in first-use order
        idx = [q for q, pr in enumerate(precs) if pr == prec]
        od = out_dtype or ("fp32" if prec.startswith("bf16") else "fp64")
        for q, T in zip(idx, api.computeF_batch([specs[q] for q in idx], precision=prec, out_dtype=od)):
            tensors[q] = T
    if synchronize:
        synchronize()
    t1 = time.perf_counter()
    items, k = [], 0
    for p in prepared:
        x = p["inst"]
        d = dict(nup=x["nup"], ndown=x["ndown"], S=p["S"], H=p["H"], r1=p["grid"])
        if x["method"] == "tc":
            d.update(kind="tc", Fs=tensors[k], Fo=tensors[k + 1], V=tensors[k + 2], jastrow=p["jastrow"])
            k += 3
        else:
            d.update(kind="rcs" if x["nup"] == x["ndown"] else "uos", F=tensors[k])
            k += 1
        items.append(d)
    res = solve_each(items, tol=tol, max_iter=max_iter, diis=diis, workers=workers)
    t2 = time.perf_counter()
    del tensors
    return res, dict(precompute_ms=(t1 - t0) * 1e3, scf_ms=(t2 - t1) * 1e3, total_ms=(t2 - t0) * 1e3,
                     scf=f"the single-instance drivers, {workers} host threads on streams of their own, largest first",
                     fp64_instances=sum(instance_precision(p["inst"], precision, policy) == "fp64" for p in prepared))


_STREAMS = {}


def _worker_streams(n):
    """n streams of the current device (default priority), kept for the process:
    the drivers keep their per-call resources per stream, so a fixed set bounds
    them."""
    import torch
    dev = torch.cuda.current_device()
    have = _STREAMS.setdefault(dev, [])
    while len(have) < n:
        have.append(torch.cuda.Stream())
    return have[:n]


def _cost(x):
    """Relative cost of an instance's SCF: points x basis pairs, five tensors'
    worth for TC."""
    nb = x["r1"].nBasis()
    return x["r1"].nGrid() * nb * nb * (5 if x["kind"] == "tc" else 1)


def solve_each(items, tol, max_iter, diis, workers=4):
    """The SCF of every instance with the single-instance drivers.  `workers`
    host threads, each on a stream of its own, take the instances largest first
    and launch each iteration's kernels directly (graph=False); the drivers'
    per-stream resources let the solves run side by side, and every solve's
    arithmetic is that of a solve alone.  (The graph-batched tcscf_batch_scf is
    switched off.)  Returns the per-instance dicts energy / iterations / status
    / trace / D (and C, the occupied orbitals, for TC)."""
    import itertools
    import threading

    import torch

    out = [None] * len(items)
    order = sorted(range(len(items)), key=lambda k: -_cost(items[k]))
    nxt = itertools.count()
    failures = []

    def solve(x):
        nb = x["r1"].nBasis()
        if x["kind"] == "tc":
            hf = api.TCHartreeFock(x["nup"], x["ndown"], nb)
            call = lambda: hf.compute(x["S"], x["H"], x["r1"], x["Fs"], x["Fo"], x["V"], x["jastrow"], True, tol=tol,
                                      max_iter=max_iter, mgs_mode=1, diis=diis, graph=False)
        elif x["kind"] == "rcs":
            hf = api.RCSHartreeFock(x["nup"] + x["ndown"], nb)
            call = lambda: hf.compute(x["S"], x["H"], x["r1"], x["F"], True, tol=tol, max_iter=max_iter, diis=diis,
                                      graph=False)
        else:
            hf = api.UOSHartreeFock(x["nup"], x["ndown"], nb)
            call = lambda: hf.compute(x["S"], x["H"], x["r1"], x["F"], True, tol=tol, max_iter=max_iter, diis=diis,
                                      graph=False)
        try:
            call()
            status = 0
        except TcscfError as err:
            status = err.code
        return dict(energy=hf.energy, iterations=hf.numberOfConvergenceLoops(), status=status,
                    trace=hf.energy_trace, D=hf.density, C=getattr(hf, "occupiedOrbitals", None))

    def worker(stream):
        try:
            with torch.cuda.stream(stream):
                while True:
                    q = next(nxt)  # atomic under the interpreter lock
                    if q >= len(order):
                        return
                    out[order[q]] = solve(items[order[q]])
        except BaseException as err:  # surfaces in the caller
            failures.append(err)

    torch.cuda.synchronize()  # the tensors come from the contraction's stream
    streams = _worker_streams(max(1, workers))
    threads = [threading.Thread(target=worker, args=(s,)) for s in streams]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    if failures:
        raise failures[0]
    return out
