"""Pins the CPU oracle of the bidiagonal SVD path (MDBDSDC / MDLASD0-2 and the LAPACK
routines they call) against LAPACK, closed forms and invariants.  CPU only."""
import numpy as np
import pytest

import lapack_ref as L
import oracle_py as O
from hsspack_amd.workloads import bidiag_cfg

EPS = 2.0 ** -53


def bidiag(d, e, sqre=0):
    n = len(d)
    B = np.zeros((n, n + sqre))
    B[np.arange(n), np.arange(n)] = d
    B[np.arange(n - 1 + sqre), np.arange(1, n + sqre)] = e[: n - 1 + sqre]
    return B


def check_svd(B, s, U, VT, tol=1e-12):
    n, m = B.shape
    S = np.zeros((n, m))
    S[np.arange(n), np.arange(n)] = s
    nrm = np.abs(B).max()
    assert np.abs(U.T @ U - np.eye(n)).max() <= tol
    assert np.abs(VT @ VT.T - np.eye(m)).max() <= tol
    assert np.abs(U @ S @ VT - B).max() / nrm <= tol


# ---------------------------------------------------------------- DLASDT
@pytest.mark.parametrize("n", [26, 27, 51, 100, 257, 1000, 4097, 16384])
@pytest.mark.parametrize("msub", [3, 25, 40])
def test_lasdt_matches_lapack(n, msub):
    a = O.lasdt(n, msub)
    b = L.dlasdt(n, msub)
    assert a[0] == b[0] and a[1] == b[1]
    for x, y in zip(a[2:], b[2:]):
        assert np.array_equal(x, y)


# ---------------------------------------------------------------- leaves
@pytest.mark.parametrize("sqre", [0, 1])
def test_leaf_matches_dlasdq(sqre):
    """Jacobi leaf against LAPACK DLASDQ: singular values to a few ulps, vectors equal
    up to sign, the null vector last; B = U [S 0] VT to rounding."""
    rng = np.random.default_rng(11 + sqre)
    for m in list(range(1, 26)) + [40, 63 - sqre]:
        d = rng.uniform(0.5, 1.5, m)
        e = rng.uniform(0.0, 1.0, m - 1 + sqre)
        s, U, VT, sw = O.bd_leaf(d, e, sqre)
        assert sw > 0
        sl, Ul, VTl, info = L.dlasdq(d, e, sqre)
        assert info == 0
        assert np.all(np.diff(s) >= 0)
        assert np.abs(s - sl).max() <= 4 * m * EPS * sl.max()
        B = bidiag(d, e, sqre)
        check_svd(B, s, U, VT, tol=4e-14)
        gap = np.min(np.diff(sl)) if m > 1 else 1.0
        if gap > 1e-6:
            assert np.abs(np.abs(np.sum(U * Ul, axis=0)) - 1).max() <= 1e-12
            assert np.abs(np.abs(np.sum(VT * VTl, axis=1)) - 1).max() <= 1e-12
        if sqre:
            assert np.abs(B @ VT[m]).max() <= 1e-13


def test_leaf_graded_and_clustered():
    """Graded and repeated entries: still an SVD to rounding."""
    for d, e in [(np.full(20, 1.0), np.full(20, 1.0)),
                 (np.logspace(0, -12, 16), np.logspace(0, -12, 16)),
                 (np.full(8, 2.0), np.zeros(8))]:
        for sqre in (0, 1):
            s, U, VT, sw = O.bd_leaf(d, e, sqre)
            assert sw > 0
            check_svd(bidiag(d, e, sqre), s, U, VT, tol=1e-13)
            sl = L.dlasdq(d, e, sqre)[0]
            assert np.abs(s - sl).max() <= 4 * len(d) * EPS * sl.max()


# ---------------------------------------------------------------- secular roots
def secular_case(rng, k, rho=None, cluster=False):
    d = np.sort(rng.uniform(0.0, 1.0, k))
    if cluster:
        d[1::3] = d[0::3][: len(d[1::3])] + 1e-9
        d = np.sort(d)
    d[0] = 0.0
    d = np.maximum.accumulate(d + np.arange(k) * 1e-13)
    z = rng.uniform(0.1, 1.0, k) * rng.choice([-1, 1], k)
    z /= np.linalg.norm(z)
    return d, z, (rng.uniform(0.1, 2.0) if rho is None else rho)


@pytest.mark.parametrize("k", [2, 3, 5, 17, 64, 300])
def test_lasd4_matches_lapack(k):
    rng = np.random.default_rng(100 + k)
    for cl in (False, True):
        d, z, rho = secular_case(rng, k, cluster=cl and k > 3)
        sig = np.zeros(k)
        for j in range(1, k + 1):
            r = O.lasd4(d, z, rho, j)
            assert r["info"] == 0
            s, delta, work, info = L.dlasd4(d, z, rho, j)
            assert info == 0
            assert abs(r["sigma"] - s) <= 8 * EPS * s
            # d_i - sigma with high relative accuracy near the root
            dorg = d[r["org"] - 1]
            mine = (d - dorg) - r["tau"]
            near = slice(max(0, j - 3), min(k, j + 2))
            rel = np.abs(mine[near] - delta[near]) / np.abs(delta[near])
            assert rel.max() <= 1e-12
            sig[j - 1] = r["sigma"]
        # interlacing and the trace identity: sum s^2 = sum d^2 + rho |z|^2
        assert np.all(sig[:-1] > d[:-1]) and np.all(sig[:-1] < d[1:]) and sig[-1] > d[-1]
        assert abs(np.sum(sig ** 2) - (np.sum(d ** 2) + rho)) <= 1e-13 * (np.sum(d ** 2) + rho)


def test_lasd4_two_poles_closed_form():
    """k = 2, d = (0, d2): s^2 are the roots of x^2 - (d2^2 + rho) x + rho z1^2 d2^2."""
    rng = np.random.default_rng(5)
    for _ in range(50):
        d2 = rng.uniform(0.01, 2.0)
        z = rng.uniform(0.05, 1.0, 2)
        z /= np.linalg.norm(z)
        rho = rng.uniform(0.01, 3.0)
        b = d2 * d2 + rho
        c = rho * z[0] ** 2 * d2 * d2
        lo = 2 * c / (b + np.sqrt(b * b - 4 * c))
        hi = c / lo
        for j, x in ((1, lo), (2, hi)):
            r = O.lasd4(np.array([0.0, d2]), z, rho, j)
            assert abs(r["sigma"] - np.sqrt(x)) <= 8 * EPS * np.sqrt(x)


def test_lasd4_against_high_precision_roots():
    """Roots to a few ulps of 40-digit roots of the same secular function."""
    mp = pytest.importorskip("mpmath")
    mp.mp.dps = 40
    rng = np.random.default_rng(13)
    d, z, rho = secular_case(rng, 17)
    dm = [mp.mpf(x) for x in d]
    zm = [mp.mpf(x) for x in z]
    f = lambda s: 1 / mp.mpf(rho) + sum(zi * zi / ((di - s) * (di + s)) for di, zi in zip(dm, zm))
    for j in range(1, 18):
        lo = dm[j - 1]
        hi = dm[j] if j < 17 else mp.sqrt(dm[-1] ** 2 + mp.mpf(rho))
        a, b = lo + (hi - lo) * mp.mpf(10) ** -30, hi - (hi - lo) * mp.mpf(10) ** -30
        for _ in range(200):
            c = (a + b) / 2
            if f(c) < 0:
                a = c
            else:
                b = c
        s = float((a + b) / 2)
        r = O.lasd4(d, z, rho, j)
        assert abs(r["sigma"] - s) <= 4 * EPS * s


def test_lasd4_matches_eigvals_of_rank_one_update():
    """s_j^2 are the eigenvalues of diag(d^2) + rho z z^T."""
    rng = np.random.default_rng(9)
    d, z, rho = secular_case(rng, 40)
    ev = np.linalg.eigvalsh(np.diag(d * d) + rho * np.outer(z, z))
    sig = np.array([O.lasd4(d, z, rho, j)["sigma"] for j in range(1, 41)])
    assert np.abs(sig ** 2 - ev).max() <= 1e-14 * ev.max()


# ---------------------------------------------------------------- merge inputs
def merge_inputs(nl, nr, sqre, d, e, alpha, beta):
    """Children solved by LAPACK DLASDQ, assembled as DLASD0 hands them to DLASD1:
    (D, U, VT, IDXQ) of the n x m block."""
    n = nl + nr + 1
    m = n + sqre
    D = np.zeros(n)
    U = np.zeros((n, n), order="F")
    VT = np.zeros((m, m), order="F")
    U[nl, nl] = 1.0
    sl, ul, vtl, _ = L.dlasdq(d[:nl], e[:nl], 1)
    sr, ur, vtr, _ = L.dlasdq(d[nl + 1:], e[nl + 1:nl + nr + sqre], sqre)
    D[:nl] = sl
    D[nl + 1:] = sr
    D[nl] = alpha
    U[:nl, :nl] = ul
    U[nl + 1:, nl + 1:] = ur
    VT[:nl + 1, :nl + 1] = vtl
    VT[nl + 1:, nl + 1:] = vtr
    idxq = np.concatenate([np.arange(1, nl + 1), [0], np.arange(1, nr + 1)]).astype(np.int32)
    return D, U, VT, idxq


def merge_cases():
    rng = np.random.default_rng(21)
    out = []
    for nl, nr, sqre in [(10, 12, 1), (12, 12, 0), (25, 24, 1), (40, 39, 0)]:
        n = nl + nr + 1
        d = rng.uniform(0.5, 1.5, n)
        e = rng.uniform(0.0, 1.0, n)
        out.append((nl, nr, sqre, d, e, d[nl], e[nl], "random"))
        # mirrored children: equal singular values on both sides -> rotation deflation
        dm = d.copy()
        em = e.copy()
        if nl == nr:
            dm[nl + 1:] = d[:nl]
            em[nl + 1:nl + nr] = e[:nl - 1]
            em[nl + nr - 1 + sqre] = e[nl - 1] if sqre else em[nl + nr - 1 + sqre]
            out.append((nl, nr, sqre, dm, em, d[nl], e[nl], "mirrored"))
        # tiny coupling -> small-z deflation
        et = e.copy()
        et[nl - 3:nl] = 1e-18
        out.append((nl, nr, sqre, d, et, d[nl], 1e-17, "tiny-z"))
    return out


@pytest.mark.parametrize("case", range(9))
def test_lasd2_matches_lapack(case):
    cases = merge_cases()
    if case >= len(cases):
        pytest.skip("fewer cases")
    nl, nr, sqre, d, e, alpha, beta, _ = cases[case]
    D, U, VT, idxq = merge_inputs(nl, nr, sqre, d, e, alpha, beta)
    a = O.lasd2(nl, nr, sqre, D, alpha, beta, U, VT, idxq)
    b = L.dlasd2(nl, nr, sqre, D, alpha, beta, U, VT, idxq)
    assert b["info"] == 0
    k = a["k"]
    assert k == b["k"]
    n = nl + nr + 1
    for key in ("idxp", "idx", "idxc"):
        assert np.array_equal(a[key][1:], b[key][1:]), key
    assert np.array_equal(a["coltyp"][:4], b["coltyp"][:4])
    assert np.array_equal(a["dsigma"], b["dsigma"])
    assert np.array_equal(a["d"][k:], b["d"][k:])
    assert np.abs(a["z"][:k] - b["z"][:k]).max() <= 4 * EPS * np.abs(b["z"][:k]).max()
    for key in ("u2", "vt2"):
        assert np.abs(a[key][:, :n] - b[key][:, :n]).max() <= 1e-15, key
    assert np.abs(a["u"] - b["u"]).max() <= 1e-15
    assert np.abs(a["vt"] - b["vt"]).max() <= 1e-15


def test_lasd2_deflation_paths_are_exercised():
    kinds = {}
    for nl, nr, sqre, d, e, alpha, beta, kind in merge_cases():
        D, U, VT, idxq = merge_inputs(nl, nr, sqre, d, e, alpha, beta)
        a = O.lasd2(nl, nr, sqre, D, alpha, beta, U, VT, idxq)
        kinds.setdefault(kind, []).append((a["k"], len(a["rot"][0]), nl + nr + 1))
    assert any(nrot > 0 for _, nrot, _ in kinds["mirrored"])
    assert any(k < n for k, _, n in kinds["tiny-z"])


@pytest.mark.parametrize("case", range(9))
def test_lasd1_matches_lapack(case):
    """A whole merge (deflation, roots, vectors, products, DLAMRG) against LAPACK DLASD1."""
    cases = merge_cases()
    if case >= len(cases):
        pytest.skip("fewer cases")
    nl, nr, sqre, d, e, alpha, beta, _ = cases[case]
    D, U, VT, idxq = merge_inputs(nl, nr, sqre, d, e, alpha, beta)
    a = O.lasd1(nl, nr, sqre, D, alpha, beta, U, VT, idxq)
    bd, bu, bvt, bidx, info = L.dlasd1(nl, nr, sqre, D, alpha, beta, U, VT, idxq)
    assert a["info"] == 0 and info == 0
    assert np.array_equal(a["idxq"], bidx)
    assert np.abs(a["d"] - bd).max() <= 8 * EPS * bd.max()
    assert np.abs(a["u"] - bu).max() <= 1e-13
    assert np.abs(a["vt"] - bvt).max() <= 1e-13
    # the merged block is an SVD of the block's bidiagonal
    B = bidiag(d, e, sqre)
    B[nl, nl] = alpha
    B[nl, nl + 1] = beta
    S = np.zeros(B.shape)
    S[np.arange(len(d)), np.arange(len(d))] = a["d"]
    assert np.abs(a["u"] @ S @ a["vt"] - B).max() <= 1e-13


# ---------------------------------------------------------------- driver
@pytest.mark.parametrize("n,uplo", [(20, b"U"), (30, b"U"), (200, b"U"), (1000, b"U"), (301, b"L")])
def test_bdsdc_matches_lapack(n, uplo):
    d, e = bidiag_cfg(n, seed=5)
    r = O.bdsdc(d, e, uplo=uplo)
    assert r["info"] == 0
    s, U, VT, info = L.dbdsdc(d, e, uplo=uplo)
    assert info == 0
    B = bidiag(d, e) if uplo == b"U" else bidiag(d, e).T
    nrm = np.abs(B).max()
    assert np.all(np.diff(r["d"]) <= 0)
    assert np.abs(r["d"] - s).max() <= 4 * n * EPS * nrm
    check_svd(B, r["d"], r["u"], r["vt"])
    gap = np.minimum(np.r_[np.inf, -np.diff(s)], np.r_[-np.diff(s), np.inf])
    ok = gap > 1e-8
    assert np.abs(np.abs(np.sum(r["u"] * U, axis=0))[ok] - 1).max() <= 1e-8
    assert np.abs(np.abs(np.sum(r["vt"] * VT, axis=1))[ok] - 1).max() <= 1e-8


def test_bdsdc_splits_and_small_diagonal():
    """Tiny superdiagonal entries split the problem (mdbdsdc.f90:333-372); a tiny
    diagonal entry is raised to eps."""
    n = 300
    d, e = bidiag_cfg(n, seed=6)
    e[[40, 41, 150, n - 2]] = 1e-20
    d[77] = 1e-30
    r = O.bdsdc(d, e)
    assert r["info"] == 0
    s = L.dbdsdc(d, e)[0]
    assert np.abs(r["d"] - s).max() <= 4 * n * EPS * 1.5
    check_svd(bidiag(d, e), r["d"], r["u"], r["vt"])


def test_bdsdc_trace_k_matches_lapack_merges():
    """Merge sizes K of the oracle's DC tree equal those LAPACK's DLASD1 finds when fed
    the oracle's own children (the deflation decisions agree)."""
    n = 400
    d, e = bidiag_cfg(n, seed=7)
    r = O.bdsdc(d, e, trace_cap=64)
    tr = r["trace"]
    assert tr["level"].size > 0
    assert np.all(tr["k"] >= 1) and np.all(tr["k"] <= tr["n"])
    assert tr["roots"] == int(tr["k"].sum())


# ---------------------------------------------------------------- MDLASD8 and the HSS branch
def lasd8_case(k, seed, toeplitz=False):
    rng = np.random.default_rng(seed)
    if toeplitz:
        d = np.concatenate([[0.0], 1.0 + np.cumsum(rng.uniform(0.5, 1.5, k - 1)) / k])
    else:
        d = np.concatenate([[0.0], np.sort(rng.uniform(0.05, 1.0, k - 1))])
    z = rng.uniform(0.1, 1.0, k) * rng.choice([-1, 1], k)
    return d, z


@pytest.mark.parametrize("k", [3, 17, 64, 300])
def test_lasd8_matches_lapack(k):
    d, z = lasd8_case(k, k)
    o = O.lasd8(d, z)
    r = L.dlasd8(d, z)
    assert o["info"] == 0 and r["info"] == 0
    np.testing.assert_allclose(o["sigma"], r["sigma"], rtol=4 * k * EPS, atol=0)
    # the differences sigma_j - d_j and sigma_j - d_{j+1}, each to a few ulps of itself
    np.testing.assert_allclose(o["difl"], r["difl"], rtol=1e-12, atol=0)
    np.testing.assert_allclose(o["difr"][:-1], r["difr"][:-1], rtol=1e-12, atol=0)
    np.testing.assert_allclose(np.abs(o["z"]), np.abs(r["z"]), rtol=1e-12, atol=0)
    np.testing.assert_allclose(o["aphav"], 1.0 / r["vnorm"], rtol=1e-12, atol=0)


@pytest.mark.parametrize("k,toeplitz", [(40, False), (257, False), (400, True)])
def test_svd_vector_matrices_are_the_merge_singular_vectors(k, toeplitz):
    """From MDLASD8's outputs: V(i,j) = zhat_i APHAV_j / (d_i^2 - sigma_j^2) and U with
    rows d_i zhat_i APHAU_j / (d_i^2 - sigma_j^2), row 1 = -APHAU, are the singular
    vectors of the arrowhead matrix e_1 zhat^T + diag(d): orthonormal, M V = U Sigma."""
    d, z = lasd8_case(k, 3 * k, toeplitz)
    o = O.lasd8(d, z)
    s, zh = o["sigma"], o["z"]
    V = O.cauchy_dense_sq(d, s, zh, o["aphav"], o["difl"], o["difr"])
    U = O.cauchy_dense_sq(d, s, d * zh, o["aphau"], o["difl"], o["difr"])
    U[0, :] = -o["aphau"]
    M = np.diag(d)
    M[0, :] = zh
    assert np.abs(V.T @ V - np.eye(k)).max() <= 1e-13
    assert np.abs(U.T @ U - np.eye(k)).max() <= 1e-13
    assert np.abs(M @ V - U * s[None, :]).max() <= 1e-13 * np.abs(M).max()


@pytest.mark.parametrize("k,ni,toeplitz", [(300, 40, False), (600, 40, True), (1000, 64, False)])
def test_hss_squared_matches_dense(k, ni, toeplitz):
    d, z = lasd8_case(k, k + ni, toeplitz)
    o = O.lasd8(d, z)
    tol = float(np.float32(1e-17))
    for u, v in ((o["z"], o["aphav"]), (d * o["z"], o["aphau"])):
        H = O.hss_build_sq(d, o["sigma"], u, v, o["difl"], o["difr"], ni, tol)
        try:
            Hd = O.hss_dense(H)
            usz, vsz = O.hss_ranks(H)
            x = np.random.default_rng(1).standard_normal((37, k))
            y = O.hss_mmL(H, x)
        finally:
            O.hss_free(H)
        C = O.cauchy_dense_sq(d, o["sigma"], u, v, o["difl"], o["difr"])
        scale = np.abs(C).max()
        assert np.abs(Hd - C).max() <= 64 * EPS * k ** 0.5 * scale
        assert np.abs(y - x @ C).max() <= 1e-13 * np.abs(x).max() * k ** 0.5 * scale
        assert usz[:-1, 1].max() < k // 2   # compressed, not dense


@pytest.mark.parametrize("gen,n,bsmlz", [("toeplitz", 1000, 100), ("uniform", 2000, 64),
                                         ("toeplitz", 2200, 2000)])
def test_bdsdc_hss_branch_matches_lapack(gen, n, bsmlz):
    rng = np.random.default_rng(n)
    if gen == "toeplitz":
        d, e = 1.0 + 1e-3 * rng.standard_normal(n), 1.0 + 1e-3 * rng.standard_normal(n - 1)
    else:
        d, e = 0.5 + rng.random(n), rng.random(n - 1)
    o = O.bdsdc(d, e, prm=O.params(bd_bsmlz=bsmlz), trace_cap=4 * n)
    assert o["info"] == 0
    tr = o["trace"]
    hss = tr["rank_max"] > 0
    assert hss.any() and np.all(tr["k"][hss] >= bsmlz)
    assert np.all(tr["k"][~hss] < bsmlz)
    s_ref = L.dbdsdc(d, e)[0]
    np.testing.assert_allclose(o["d"], s_ref, rtol=0, atol=4 * n * EPS * s_ref[0])
    check_svd(bidiag(d, e), o["d"], o["u"], o["vt"], tol=1e-12)
