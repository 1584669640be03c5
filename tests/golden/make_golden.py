"""Writes tests/golden/lapack_golden.npz: input/output vectors of the LAPACK
routines on the reference path (DLAED4, DLAMRG, DLAED2, DLAED3, DSTEQR, DSTEDC),
computed with the LAPACK inside scipy's OpenBLAS 0.3.28.

The reference repository ships no golden vectors of its own (its TEST/ programs
only print); these fixtures pin the oracle's restatement of the LAPACK calls the
reference makes (mdlaed3.f90:208 DLAED4, mdlaed1.f90:199 DLAED3,
mdlaed1.f90:221 DLAMRG, mdlaed0.f90:263 DSTEQR; MDLAED2 follows DLAED2 through
mdlaed2.f90:409).  Run:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import lapack_ref as L  # noqa: E402
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from hsspack_amd.workloads import splitmix_uniform, tridiag_cfg  # noqa: E402


def merge_inputs(d, e, n1):
    """Eigen-decompose the two torn halves of T (mdlaed0.f90:213-219 tear) with
    LAPACK DSTEQR and return the DLAED2 inputs of their merge."""
    n = len(d)
    rho = e[n1 - 1]
    d = d.copy()
    d[n1 - 1] -= abs(rho)
    d[n1] -= abs(rho)
    w1, q1, _ = L.dsteqr(d[:n1], e[:n1 - 1])
    w2, q2, _ = L.dsteqr(d[n1:], e[n1:n - 1])
    q = np.zeros((n, n), order="F")
    q[:n1, :n1] = q1
    q[n1:, n1:] = q2
    z = np.concatenate([q1[n1 - 1, :], q2[0, :]])
    dd = np.concatenate([w1, w2])
    indxq = np.concatenate([np.arange(1, n1 + 1), np.arange(1, n - n1 + 1)]).astype(np.int32)
    return dd, q, indxq, rho, z


def main():
    out = {}
    rng = np.random.default_rng(20261202)
    # ---- DLAED4 ----
    fam = []
    for t in range(24):
        n = [1, 2, 3, 4, 7, 16, 33, 64][t % 8]
        if t < 8:
            d = np.sort(rng.standard_normal(n))
        elif t < 16:
            d = np.sort(rng.uniform(0, 1, n))
            if n > 2:
                d[1] = d[0] + 1e-10          # clustered poles
        else:
            d = np.sort(rng.uniform(-1, 1, n))
        z = rng.standard_normal(n)
        if t >= 16 and n > 3:
            z[n // 2] = 1e-9                 # tiny weight
        z /= np.linalg.norm(z)
        rho = float(abs(rng.standard_normal()) + 0.05)
        lams, deltas = [], []
        for j in range(1, n + 1):
            lam, delta, info = L.dlaed4(d, z, rho, j)
            assert info == 0
            lams.append(lam)
            deltas.append(delta)
        fam.append((d, z, rho, np.array(lams), np.array(deltas)))
    for i, (d, z, rho, lams, deltas) in enumerate(fam):
        out["laed4_%d_d" % i] = d
        out["laed4_%d_z" % i] = z
        out["laed4_%d_rho" % i] = np.array([rho])
        out["laed4_%d_lam" % i] = lams
        out["laed4_%d_delta" % i] = deltas
    out["laed4_count"] = np.array([len(fam)])
    # ---- DLAMRG ----
    for i, (n1, n2, s1, s2) in enumerate([(5, 7, 1, 1), (8, 8, 1, -1), (13, 4, -1, 1), (6, 6, -1, -1)]):
        a1 = np.sort(rng.integers(0, 6, n1).astype(np.float64))
        a2 = np.sort(rng.integers(0, 6, n2).astype(np.float64))
        if s1 < 0:
            a1 = a1[::-1]
        if s2 < 0:
            a2 = a2[::-1]
        a = np.concatenate([a1, a2])
        out["lamrg_%d_a" % i] = a
        out["lamrg_%d_args" % i] = np.array([n1, n2, s1, s2], np.int32)
        out["lamrg_%d_idx" % i] = L.dlamrg(a, n1, n2, s1, s2)
    out["lamrg_count"] = np.array([4])
    # ---- DLAED2 + DLAED3 on real merges ----
    cases = [
        ("rand", tridiag_cfg(2, 96, 11)),                  # U(-1,1): small-z deflation
        ("toep", tridiag_cfg(1, 80, 12)),                  # near-Toeplitz: little deflation
        ("exact", (np.full(64, 2.0), np.full(63, 1.0))),   # Testedc.f90:34-37: rotations
        ("neg", (np.linspace(-1, 1, 50), -np.abs(splitmix_uniform(13, 49)) - 0.1)),
    ]
    for i, (name, (d, e)) in enumerate(cases):
        n = len(d)
        n1 = n // 2
        dd, q, indxq, rho, z = merge_inputs(d, e, n1)
        r = L.dlaed2(dd, q, indxq, rho, z, n1)
        assert r["info"] == 0
        pre = dict(d=dd, q=q, indxq=indxq, rho=np.array([rho]), z=z, n1=np.array([n1]))
        for k, v in pre.items():
            out["laed2_%d_in_%s" % (i, k)] = v
        for k in ("k", "rho", "info"):
            out["laed2_%d_out_%s" % (i, k)] = np.array([r[k]])
        for k in ("d", "q", "indxq", "z", "dlamda", "w", "q2", "indx", "indxc", "indxp", "coltyp"):
            out["laed2_%d_out_%s" % (i, k)] = r[k]
        k = r["k"]
        if k > 0:
            d3, q3, info3 = L.dlaed3(k, n, n1, r["q"], r["rho"], r["dlamda"], r["q2"], r["indxc"],
                                      r["coltyp"][:4], r["w"])
            assert info3 == 0
            out["laed3_%d_d" % i] = d3
            out["laed3_%d_q" % i] = q3
    out["laed2_count"] = np.array([len(cases)])
    # ---- DSTEQR ----
    for i, n in enumerate([1, 2, 5, 16, 25]):
        d = rng.uniform(-1, 1, n)
        e = rng.uniform(-1, 1, max(n - 1, 0))
        w, zz, info = L.dsteqr(d, e)
        assert info == 0
        out["steqr_%d_d" % i] = d
        out["steqr_%d_e" % i] = e
        out["steqr_%d_w" % i] = w
    out["steqr_count"] = np.array([5])
    # ---- DSTEDC ----
    for i, (cfg, n, seed) in enumerate([(2, 64, 31), (1, 100, 32), (2, 300, 33), (1, 257, 34)]):
        d, e = tridiag_cfg(cfg, n, seed)
        w, zz, info = L.dstedc(d, e)
        assert info == 0
        out["stedc_%d_d" % i] = d
        out["stedc_%d_e" % i] = e
        out["stedc_%d_w" % i] = w
    out["stedc_count"] = np.array([4])
    np.savez_compressed(os.path.join(HERE, "lapack_golden.npz"), **out)
    print("wrote", os.path.join(HERE, "lapack_golden.npz"))


if __name__ == "__main__":
    main()
