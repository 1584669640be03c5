"""The C oracle against the reference's own Fortran (oracle/_ref, built by
oracle/build_ref.py from /root/reference/SRC with scipy's OpenBLAS as its LAPACK; every
reference call runs in a child process, tests/ref_runner.py).

* Whole solves, merge by merge in call order:
  - MDSTEDC (mdstedc.f90:139): order N, deflated size K, branch (HSS or dense) and
    deflation rotations of every merge identical; eigenvalues within 4 n eps |T|, vectors
    within 1e-9 (the near-Toeplitz solves reach the HSS branch);
  - MDBDSDC (mdbdsdc.f90:1): N, K and rotations of every merge identical; singular
    values within 4 n eps |B|, vectors within 1e-9 (a near-Toeplitz solve reaches the
    SVD's HSS branch, K >= 2000);
  - MDSBEVD (mdsbevd.f90:183), KD <= 12: N, K and the Q rotations of every rank-one
    update (MDSBED2 / MDSBED22) identical; eigenvalues within 8 n eps |A| of the
    oracle's.  The reference splits to leaves of at most SMLSIZ = 25 rows whatever KD
    (mdsbevd.f90:256): with KD >= 13 its leaves are narrower than the coupling blocks and
    DSB2FLC reads past AB (SFtransf.f:288), so the oracle and the device take leaves of
    at least 2 KD - 1 rows; for KD <= 12 both trees are the reference's.
* Constructions: the tree and every node's U/V/B/D dimensions identical for
  - DHSSEVC's structured RRLU construction (Cauchy2hssEig_omp, CauchyHssEig.f90:1945),
    and rank, pivot permutations and swap count of 240 single RRLUs, against
    libhssref_fix.so: the reference with its row pivot test's stale read corrected
    (:676, :678 read F(PL(k)), a node-local array at a global pole index; the column
    test, :829, :831, reads FF).  Against the unmodified reference the column RRLU is
    identical, the tree and leaves are identical and the product agrees to 1e-13;
  - mat2hssvd (ConstructHssd.f90:371, DHSSMM's construction) on the reference's testfmm
    matrix and a diagonal-plus-low-rank matrix (the order of compression differs: see
    the test); DHSSMM's product H B within the truncation error;
  - cauchy2hss (CauchyHss_VP2.f90:387) on the points of testCauchyHss_serial and on
    separated points; CauchyHssMM's product within 1e-12.
"""
import numpy as np
import pytest
import scipy.linalg as sl

import oracle_py as O
import ref_runner as R
from hsspack_amd.workloads import band_cfg, band_dense, bidiag_cfg, bidiag_toeplitz, tridiag_cfg
from test_oracle_cauchy import interlaced, separated
from test_oracle_dhssmm import fmm_matrix

EPS = 2.0 ** -53

pytestmark = pytest.mark.skipif(not R.available(), reason="reference library not built "
                                "(oracle/build_ref.py needs /root/reference)")


def align(a, b):
    """b's columns with the signs of a's."""
    s = np.sign(np.sum(a * b, axis=0))
    s[s == 0] = 1.0
    return b * s


@pytest.mark.parametrize("cfg,n,seed", [(2, 300, 1), (2, 1000, 2), (2, 4096, 3), (1, 2100, 4), (1, 4096, 1)])
def test_stedc_trace_matches_reference(cfg, n, seed):
    d, e = tridiag_cfg(cfg, n, seed)
    (ref,) = R.run([("stedc", dict(d=d, e=e))])
    assert ref["info"] == 0
    r = O.stedc(d, e, prm=O.params(tol=1e-17), trace_cap=1 << 16)
    assert r["info"] == 0
    tr, to = ref["trace"], r["trace"]
    assert len(tr["k"]) == len(to["k"]) > 0
    for key in ("n", "k", "hss", "nrot"):
        assert np.array_equal(tr[key], to[key]), key
    if cfg == 1:
        assert tr["hss"].sum() > 0  # the near-Toeplitz solves reach the HSS branch
    T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    nrm = np.abs(sl.eigvalsh_tridiagonal(d, e)).max()
    lam, z = ref["d"], ref["z"]
    assert np.abs(lam - r["d"]).max() <= 4 * n * EPS * nrm
    assert np.abs(z - align(z, r["z"])).max() <= 1e-9
    assert np.abs(z.T @ z - np.eye(n)).max() <= 1e-12
    assert np.abs(T @ z - z * lam).max() <= 1e-12 * nrm


@pytest.mark.parametrize("name,n", [("cfg5", 300), ("cfg5", 2000), ("toeplitz", 600), ("toeplitz", 2100)])
def test_bdsdc_trace_matches_reference(name, n):
    d, e = bidiag_cfg(n) if name == "cfg5" else bidiag_toeplitz(n)
    (ref,) = R.run([("bdsdc", dict(d=d, e=e))])
    assert ref["info"] == 0
    r = O.bdsdc(d, e, trace_cap=1 << 16)
    assert r["info"] == 0
    tr, to = ref["trace"], r["trace"]
    assert len(tr["k"]) == len(to["k"]) > 0
    assert np.array_equal(tr["n"], to["n"]) and np.array_equal(tr["k"], to["k"])
    assert np.array_equal(tr["nrot"], to["nrot"])
    if name == "toeplitz" and n >= 2100:
        assert tr["hss"].sum() > 0 and np.all((to["rank_max"] > 0) == (tr["hss"] > 0))
    B = np.diag(d) + np.diag(e, 1)
    nrm = np.abs(r["d"]).max()
    assert np.abs(ref["d"] - r["d"]).max() <= 4 * n * EPS * nrm
    assert np.abs(ref["u"] - align(ref["u"], r["u"])).max() <= 1e-9
    assert np.abs(ref["vt"].T - align(ref["vt"].T, r["vt"].T)).max() <= 1e-9
    assert np.abs(B - (ref["u"] * ref["d"]) @ ref["vt"]).max() <= 1e-12 * nrm


@pytest.mark.parametrize("n,kd", [(60, 1), (200, 3), (300, 8), (1000, 5), (700, 12), (4100, 2)])
def test_sbevd_trace_matches_reference(n, kd):
    ab = band_cfg(n, kd, seed=6, sub=n)
    (ref,) = R.run([("sbevd", dict(ab=ab, kd=kd))])
    assert ref["info"] == 0
    r = O.sbevd(ab, kd, smlsiz=25, trace_cap=1 << 16)
    assert r["info"] == 0
    tr, to = ref["trace"], r["trace"]
    assert len(tr["k"]) == len(to["k"]) > 0
    assert np.array_equal(tr["n"], to["n"]) and np.array_equal(tr["k"], to["k"])
    assert np.array_equal(tr["nrot"], to["nrot"])
    A = band_dense(ab)
    nrm = np.abs(sl.eig_banded(ab, lower=False, eigvals_only=True)).max()
    assert np.abs(ref["d"] - r["d"]).max() <= 8 * n * EPS * nrm
    q = ref["q"]
    assert np.abs(q.T @ q - np.eye(n)).max() <= 1e-11
    assert np.abs(A @ q - q * ref["d"]).max() <= 1e-11 * nrm


def secular_generators(kind, k, seed):
    """(f, d0, u, v, difl, difr) of one merge: MDLAED3's roots, Loewner z, alpha."""
    if kind == "random":
        rng = np.random.default_rng(seed)
        dl = np.sort(rng.uniform(0, 1, k))
        z = rng.standard_normal(k)
        rho = 1.0
    else:  # a near-Toeplitz tridiagonal torn in half (the HSS-branch merges)
        d, e = tridiag_cfg(1, k, seed)
        n1 = k // 2
        rho = e[n1 - 1]
        d1 = d[:n1].copy()
        d1[-1] -= rho
        d2 = d[n1:].copy()
        d2[0] -= rho
        w1, q1 = sl.eigh_tridiagonal(d1, e[:n1 - 1])
        w2, q2 = sl.eigh_tridiagonal(d2, e[n1:])
        dl = np.r_[w1, w2]
        z = np.r_[q1[-1, :], q2[0, :]]
        o = np.argsort(dl)
        dl, z = dl[o], z[o]
        keep = np.abs(z) > 1e-9
        keep[1:] &= np.diff(dl) > 1e-12
        dl, z = dl[keep], z[keep]
        rho = abs(rho)
    z = z / np.linalg.norm(z)
    g = O.mdlaed3_vectors(dl, z, rho)
    return g["dlamda"], g["d"], g["z"], g["alpha"], g["difl"], g["difr"]


def oracle_dims(H):
    ltr = H.ltr
    us, vs = O.hss_ranks(H)
    return dict(tr=np.ctypeslib.as_array(H.tr, (ltr,)).copy(), usz=us, vsz=vs,
                bsz=np.ctypeslib.as_array(H.bsz, (2 * ltr,)).reshape(-1, 2).copy(),
                dsz=np.ctypeslib.as_array(H.dsz, (ltr,)).copy())


def same_dims(a, b):
    return all(np.array_equal(a[k], b[k]) for k in ("tr", "usz", "vsz", "bsz", "dsz"))


@pytest.mark.parametrize("kind,k,seed", [("random", 1000, 1), ("random", 3000, 2), ("toeplitz", 2000, 4),
                                         ("toeplitz", 4096, 1)])
def test_hss_construction_matches_reference(kind, k, seed):
    gens = secular_generators(kind, k, seed)
    ni, tol = 200, 1e-17
    x = np.random.default_rng(3).standard_normal((48, len(gens[0])))
    H = O.hss_build(*gens, ni, tol)
    try:
        od = oracle_dims(H)
        xo = O.hss_mmL(H, x)
    finally:
        O.hss_free(H)
    fix, ref0 = R.run([("hss_build", dict(gens=gens, ni=ni, tol=tol, x=x, fixed=True)),
                       ("hss_build", dict(gens=gens, ni=ni, tol=tol, x=x))])
    # corrected reference: every node's generator dimensions identical
    assert fix["info"] == 0 and same_dims(fix, od)
    assert np.abs(fix["x"] - xo).max() <= 1e-13 * np.abs(xo).max()
    # unmodified reference: the same tree and leaves; the product agrees to rounding
    assert np.array_equal(ref0["tr"], od["tr"]) and np.array_equal(ref0["dsz"], od["dsz"])
    assert np.abs(ref0["x"] - xo).max() <= 1e-13 * np.abs(xo).max()


def rrlu_case(rowcol, rows, cols, gens, tol):
    """Arguments of one RRLU on a node, as ComprcauchyEig passes them
    (CauchyHssEig.f90:2377, 2409); F has K entries, the node's values first."""
    FF, D0, Z, AL, difl, difr = gens
    if rowcol == 0:
        F, D, A, V = FF[rows], D0[cols], Z[rows], AL[cols]
        P1, P2, M, N = rows + 1, cols + 1, len(rows), len(cols)
    else:
        F, D, A, V = FF[rows], D0[cols], AL[cols], Z[rows]
        P1, P2, M, N = cols + 1, rows + 1, len(cols), len(rows)
    Fb = np.zeros(len(FF) + 1)
    Fb[:len(F)] = F
    return (rowcol, Fb, D.copy(), A.copy(), V.copy(), difl, difr, tol, FF, P1.astype(np.int32), P2.astype(np.int32),
            M, N)


def oracle_rrlu(case):
    import ctypes
    rowcol, F, D, A, V, difl, difr, tol, FF, P1, P2, M, N = case
    F, D, A, V, P1, P2 = [np.array(a) for a in (F, D, A, V, P1, P2)]
    U, W, PQ = np.zeros(M + 1), np.zeros(min(M, N) + 1), np.zeros(M + 1, np.int32)
    rk, nf, ns = ctypes.c_int(0), ctypes.c_longlong(0), ctypes.c_longlong(0)
    O.lib().orc_rrlu(rowcol, O.dp(F), O.dp(D), O.dp(A), O.dp(V), O.dp(difl), O.dp(difr), ctypes.c_double(tol),
                     O.dp(U), O.dp(W), O.dp(FF), O.ip(P1), O.ip(P2), M, N, ctypes.byref(rk), O.ip(PQ),
                     ctypes.byref(nf), ctypes.byref(ns))
    return rk.value, PQ[:M].copy(), P1, P2, ns.value


def test_rrlu_matches_corrected_reference():
    gens = secular_generators("random", 800, 5)
    K = len(gens[0])
    rng = np.random.default_rng(7)
    cases = []
    for trial in range(240):
        M, N = int(rng.integers(5, 120)), int(rng.integers(5, 300))
        rows = rng.choice(K, M, replace=False)
        cols = rng.choice(np.setdiff1d(np.arange(K), rows), N, replace=False)
        if trial % 2 == 0:
            rows, cols = np.sort(rows), np.sort(cols)
        cases.append(rrlu_case((trial // 2) % 2, rows, cols, gens, (1e-17, 1e-13, 1e-8)[trial % 3]))
    col_cases = [c for c in cases if c[0] == 1]
    fix, ref0 = R.run([("rrlu", dict(cases=cases, fixed=True)), ("rrlu", dict(cases=col_cases))])
    for i, c in enumerate(cases):
        a, b = fix[i], oracle_rrlu(c)
        assert a[0] == b[0], (i, "rank")
        assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2]) and np.array_equal(a[3], b[3]), i
        assert a[4] == b[4], (i, "swaps")
    assert len(col_cases) == 120
    for i, c in enumerate(col_cases):  # the column search has no stale read
        a, b = ref0[i], oracle_rrlu(c)
        assert a[0] == b[0] and np.array_equal(a[1], b[1]) and a[4] == b[4], i


def lowrank_plus_diag(n, r, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((n, r)) @ rng.standard_normal((r, n)) + np.diag(np.arange(n, dtype=float))


@pytest.mark.parametrize("name,n,seed,ni,tol", [("fmm", 2000, 1, 64, 1e-11), ("fmm", 1000, 2, 50, 1e-8),
                                                ("lowrank", 600, 3, 40, 1e-10)])
def test_mat2hss_matches_reference(name, n, seed, ni, tol):
    """The reference compresses node by node in postorder, each block row against the
    columns its earlier nodes already compressed (ConstructHssd.f90:475-700); the oracle
    and the device compress a tree level at a time against that level's bases.  The
    matrices decomposed differ by O(tol), so a singular value within that of tol can be
    counted differently (the testfmm matrix with seed 1 at n = 1000, tol 1e-8: one node's
    rank 26 against 27); on these inputs every rank agrees."""
    a = fmm_matrix(n, seed) if name == "fmm" else lowrank_plus_diag(n, 8, seed)
    b = np.random.default_rng(5).standard_normal((n, 200))  # columns a multiple of NCOL
    ref, mm = R.run([("mat2hss", dict(a=a, ni=ni, tol=tol)), ("dhssmm", dict(a=a, b=b, tol=tol, ni=ni, ncol=100))])
    H = O.mat2hss(a, ni, tol)
    try:
        od = oracle_dims(H)
        hd = O.hss_dense(H)
    finally:
        O.hss_free(H)
    assert same_dims(ref, od)
    assert mm["info"] == 0
    # two truncations of A at tol: their products differ by O(tol |A| |B|)
    ref_prod, orc_prod = mm["b"], hd @ b
    assert np.abs(ref_prod - orc_prod).max() <= 100 * max(tol, 1e-14) * np.abs(a).max() * np.abs(b).sum(0).max()


@pytest.mark.parametrize("gen,n,ni,tol", [(interlaced, 1000, 120, 1e-17), (separated, 800, 64, 1e-12),
                                          (interlaced, 2000, 64, 1e-11)])
def test_cauchy2hss_matches_reference(gen, n, ni, tol):
    d, f, u, v = gen(n, 1)
    b = np.random.default_rng(5).standard_normal((n, n))  # CauchyHssMM takes B as N x N
    ref, mm = R.run([("cauchy2hss", dict(d=d, f=f, u=u, v=v, ni=ni, tol=tol)),
                     ("cauchyhssmm", dict(d=d, f=f, u=u, v=v, ni=ni, tol=tol, ncol=128, b=b))])
    H = O.cauchy2hss(d, f, u, v, ni, tol)
    try:
        od = oracle_dims(H)
        hd = O.hss_dense(H)
    finally:
        O.hss_free(H)
    assert same_dims(ref, od)
    assert np.abs(mm["b"] - hd @ b).max() <= 1e-12 * np.abs(hd @ b).max()
