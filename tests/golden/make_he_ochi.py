"""Writes he_ochi_convergence.json.

Two kinds of data per row, He RCS in the orthonormal Ochi s basis
(n = 0..nBasis-1, l = 0) at one exponent alpha:

* the reference's recorded result, parsed from its convergence.txt (the
  energies of data.txt with the exponent each run ended at).  Each run averaged
  10 QMC grids (2,000 r1 points, 1,000 r2 points); stddev is that spread;
* ``exact_energy``: the RHF energy of the same basis with every integral done
  to ~1e-13 by one-dimensional radial quadrature, independent of any grid code
  in this repository.  For s functions the angular average of 1/r12 is 1/r_>,
  so (ab|cd) = int f_ab(r) I_cd(r)/r dr + (ab <-> cd) with f_ab = R_a R_b r^2
  and I_cd(r) = int_0^r f_cd; T_ab = 1/2 int R_a' R_b' r^2 and the nuclear
  term -Z int R_a R_b r.  The Roothaan equations are iterated to 1e-14.

Run from the repository root with the reference checkout at /root/reference
(only its convergence.txt is read, as data)."""
import json
import math
import re
import sys

import numpy as np
from scipy import special


def ochi_radial(n, alpha, r):
    """R_n(r) = N L_n^(2)(2 alpha r) e^{-alpha r}, N = (2 alpha)^1.5 / sqrt((n+2)!/n!),
    and its derivative (src/OchiBasis.hpp:31-60 at l = 0)."""
    N = (2 * alpha) ** 1.5 / math.sqrt(math.factorial(n + 2) / math.factorial(n))
    L = special.eval_genlaguerre(n, 2, 2 * alpha * r)
    dL = -2 * alpha * special.eval_genlaguerre(n - 1, 3, 2 * alpha * r) if n > 0 else 0 * r
    e = np.exp(-alpha * r)
    return N * L * e, N * e * (dL - alpha * L)


def exact_rhf(nbasis, alpha, Z=2, panels=50, inner=160):
    g, wg = special.roots_legendre(16)
    g, wg = 0.5 * (g + 1), 0.5 * wg
    edges = np.linspace(0.0, 60.0 / alpha, panels + 1)
    r = (edges[:-1, None] + np.diff(edges)[:, None] * g[None, :]).ravel()
    wr = (np.diff(edges)[:, None] * wg[None, :]).ravel()
    t, wt = special.roots_legendre(inner)
    t, wt = 0.5 * (t + 1), 0.5 * wt
    R, dR = zip(*(ochi_radial(n, alpha, r) for n in range(nbasis)))
    R, dR = np.array(R), np.array(dR)
    rt = r[:, None] * t[None, :]
    Rt = np.array([ochi_radial(n, alpha, rt)[0] for n in range(nbasis)])
    f = np.einsum("aq,bq->abq", R, R) * r ** 2
    I = np.einsum("cqi,dqi,qi,i->cdq", Rt, Rt, rt ** 2, wt) * r
    A = np.einsum("abq,cdq,q->abcd", f, I / r, wr)
    eri = A + A.transpose(2, 3, 0, 1)
    S = np.einsum("abq,q->ab", f, wr)
    assert np.abs(S - np.eye(nbasis)).max() < 1e-12
    H = 0.5 * np.einsum("aq,bq,q->ab", dR, dR, r ** 2 * wr) - Z * np.einsum("abq,q->ab", f / r, wr)
    C = np.linalg.eigh(H)[1][:, :1]
    e_old = 0.0
    for _ in range(500):
        P = 2 * C @ C.T
        F = H + np.einsum("pqrs,rs->pq", eri, P) - 0.5 * np.einsum("prqs,rs->pq", eri, P)
        C = np.linalg.eigh(F)[1][:, :1]
        P = 2 * C @ C.T
        F = H + np.einsum("pqrs,rs->pq", eri, P) - 0.5 * np.einsum("prqs,rs->pq", eri, P)
        e = 0.5 * np.sum(P * (H + F))
        if abs(e - e_old) < 1e-14:
            return float(e)
        e_old = e
    raise RuntimeError("Roothaan iteration did not converge")


src = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/convergence.txt"
rows, nb = [], None
for line in open(src):
    m = re.search(r"nMax = (\d+), lMax = (\d+), nBasis = (\d+)", line)
    if m:
        nb = (int(m.group(1)), int(m.group(2)), int(m.group(3)))
        continue
    m = re.search(r"energy = (\S+) \+/- (\S+) Ht, error = (\S+) Ht, alpha = (\S+)", line)
    if m and nb:
        rows.append({"nMax": nb[0], "lMax": nb[1], "nBasis": nb[2], "energy": float(m.group(1)),
                     "stddev": float(m.group(2)), "error_vs_hf_limit": float(m.group(3)), "alpha": float(m.group(4))})
        nb = None
for row in rows:
    row["exact_energy"] = exact_rhf(row["nBasis"], row["alpha"])
    # the finer rule must agree: the quadrature is converged
    assert abs(exact_rhf(row["nBasis"], row["alpha"], panels=65, inner=220) - row["exact_energy"]) < 1e-12
json.dump({"source": "reference convergence.txt / data.txt (He, RCS, Ochi s basis, QMC grids 2000 x 1000); "
           "exact_energy from radial quadrature of every integral",
           "hf_limit": -2.8616800, "rows": rows}, open("tests/golden/he_ochi_convergence.json", "w"), indent=1)
print(len(rows), "rows")
