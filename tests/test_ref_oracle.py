"""The C oracle against the reference's own Fortran (oracle/_ref, built by
oracle/build_ref.py from /root/reference/SRC with scipy's OpenBLAS as its LAPACK).

* MDSTEDC (mdstedc.f90:139) whole solves: every merge's order N, deflated size K, branch
  (HSS or dense) and deflation rotation count identical, in call order, to the oracle's
  trace; eigenvalues within 4 n eps |T| and vectors within 1e-9 of the oracle's.
* The structured RRLU (rrluCauchyEigrow / rrluCauchyEigcol, CauchyHssEig.f90:7, 215) and
  the HSS construction of DHSSEVC (Cauchy2hssEig_omp, :1945): rank, pivot permutations
  and swap count of every RRLU, and the tree and every node's generator dimensions,
  identical to the oracle's against libhssref_fix.so -- the reference with the row pivot
  test's stale read corrected (:676, :678 read F(PL(k)), a node-local array at a global
  index; the column test reads FF).  Against the unmodified reference the column RRLU is
  identical and the construction's product agrees to rounding; its row pivots may differ
  where the stale value changes the early-exit test.

The reference's automatic arrays exceed the default 8 MB stack at K ~ 4096, so every
reference call runs on a thread with a 1 GB stack.
"""
import ctypes
import os
import threading

import numpy as np
import pytest
import scipy.linalg as sl

import oracle_py as O
from hsspack_amd.workloads import tridiag_cfg

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(os.path.dirname(HERE), "oracle", "_ref")
EPS = 2.0 ** -53

_LIBS = {}


def ref_lib(fixed=False):
    name = "libhssref_fix.so" if fixed else "libhssref.so"
    path = os.path.join(REF_DIR, name)
    if not os.path.exists(path):
        pytest.skip("reference library not built (oracle/build_ref.py needs /root/reference)")
    if name not in _LIBS:
        lib = ctypes.CDLL(path)
        lib.refd_trace.restype = ctypes.c_int
        _LIBS[name] = lib
    return _LIBS[name]


def dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def big_stack(fn, *args):
    out = {}
    old = threading.stack_size(1 << 30)
    try:
        t = threading.Thread(target=lambda: out.setdefault("r", fn(*args)))
        t.start()
        t.join()
    finally:
        threading.stack_size(old)
    return out["r"]


def ref_stedc(lib, d, e):
    def go():
        n = len(d)
        dd = np.array(d, np.float64)
        ee = np.r_[np.asarray(e, np.float64), 0.0]
        z = np.zeros((n, n), order="F")
        info = ctypes.c_int(0)
        lib.refd_stedc(n, dp(dd), dp(ee), dp(z), n, ctypes.byref(info))
        cap = 1 << 16
        bufs = [np.zeros(cap, np.int32) for _ in range(4)]
        m = lib.refd_trace(cap, *[ip(b) for b in bufs])
        return dd, z, info.value, dict(n=bufs[0][:m], k=bufs[1][:m], hss=bufs[2][:m], nrot=bufs[3][:m])
    return big_stack(go)


@pytest.mark.parametrize("cfg,n,seed", [(2, 300, 1), (2, 1000, 2), (2, 4096, 3), (1, 2100, 4), (1, 4096, 1)])
def test_stedc_trace_matches_reference(cfg, n, seed):
    lib = ref_lib()
    d, e = tridiag_cfg(cfg, n, seed)
    lam, z, info, tr = ref_stedc(lib, d, e)
    assert info == 0
    r = O.stedc(d, e, prm=O.params(tol=1e-17), trace_cap=1 << 16)
    assert r["info"] == 0
    to = r["trace"]
    assert len(tr["k"]) == len(to["k"]) > 0
    for key in ("n", "k", "hss", "nrot"):
        assert np.array_equal(tr[key], to[key]), key
    if cfg == 1 and n >= 2100:
        assert tr["hss"].sum() > 0  # the near-Toeplitz solves reach the HSS branch
    T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    nrm = np.abs(sl.eigvalsh_tridiagonal(d, e)).max()
    assert np.abs(lam - r["d"]).max() <= 4 * n * EPS * nrm
    s = np.sign(np.sum(z * r["z"], axis=0))
    assert np.abs(z - r["z"] * s).max() <= 1e-9
    assert np.abs(z.T @ z - np.eye(n)).max() <= 1e-12
    assert np.abs(T @ z - z * lam).max() <= 1e-12 * nrm


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


def ref_hss_build(lib, gens, ni, tol, x):
    def go():
        f, d0, u, v, difl, difr = [np.array(a, np.float64) for a in gens]
        k = len(f)
        cap = 2 * (k // ni + 1)
        tr = np.zeros(cap, np.int32)
        usz, vsz, bsz = [np.zeros((cap, 2), np.int32) for _ in range(3)]
        dsz = np.zeros(cap, np.int32)
        ltr, nf, ns, info = [ctypes.c_int(0) for _ in range(4)]
        xx = np.array(x, np.float64, order="F")
        lib.refd_hss_build(k, ni, dp(f), dp(d0), dp(u), dp(v), dp(difl), dp(difr), ctypes.c_double(tol),
                           xx.shape[0], dp(xx), xx.shape[0], ctypes.byref(ltr), ip(tr), ip(usz), ip(vsz), ip(bsz),
                           ip(dsz), ctypes.byref(nf), ctypes.byref(ns), ctypes.byref(info))
        m = ltr.value
        return dict(tr=tr[:m], usz=usz[:m], vsz=vsz[:m], bsz=bsz[:m], dsz=dsz[:m], x=xx, info=info.value)
    return big_stack(go)


GENS = [("random", 1000, 1), ("random", 3000, 2), ("toeplitz", 2000, 4), ("toeplitz", 4096, 1)]


@pytest.mark.parametrize("kind,k,seed", GENS)
def test_hss_construction_matches_reference(kind, k, seed):
    gens = secular_generators(kind, k, seed)
    ni, tol = 200, 1e-17
    x = np.random.default_rng(3).standard_normal((48, len(gens[0])))
    H = O.hss_build(*gens, ni, tol)
    try:
        us, vs = O.hss_ranks(H)
        ltr = H.ltr
        tro = np.ctypeslib.as_array(H.tr, (ltr,)).copy()
        bso = np.ctypeslib.as_array(H.bsz, (2 * ltr,)).reshape(-1, 2).copy()
        dso = np.ctypeslib.as_array(H.dsz, (ltr,)).copy()
        xo = O.hss_mmL(H, x)
    finally:
        O.hss_free(H)
    # corrected reference: the same tree, every node's generator dimensions identical
    R = ref_hss_build(ref_lib(fixed=True), gens, ni, tol, x)
    assert R["info"] == 0
    assert np.array_equal(R["tr"], tro)
    assert np.array_equal(R["dsz"], dso)
    assert np.array_equal(R["usz"], us) and np.array_equal(R["vsz"], vs) and np.array_equal(R["bsz"], bso)
    assert np.abs(R["x"] - xo).max() <= 1e-13 * np.abs(xo).max()
    # unmodified reference: the same tree and leaves; the product agrees to rounding
    R0 = ref_hss_build(ref_lib(), gens, ni, tol, x)
    assert np.array_equal(R0["tr"], tro) and np.array_equal(R0["dsz"], dso)
    assert np.abs(R0["x"] - xo).max() <= 1e-13 * np.abs(xo).max()


def _rrlu(fn, wide, rowcol, rows_vals, cols_vals, rows_idx, cols_idx, gens, tol):
    """One RRLU on a node: rowcol 0 compresses rows (rrluCauchyEigrow), 1 columns
    (rrluCauchyEigcol, whose 'rows' are the node's columns, as ComprcauchyEig calls it,
    CauchyHssEig.f90:2377, 2409)."""
    FF, D0, Z, AL, difl, difr = gens
    if rowcol == 0:
        F, D, A, V = FF[rows_idx].copy(), D0[cols_idx].copy(), Z[rows_idx].copy(), AL[cols_idx].copy()
        P1, P2 = (rows_idx + 1).astype(np.int32), (cols_idx + 1).astype(np.int32)
        M, N = len(rows_idx), len(cols_idx)
    else:
        F, D, A, V = FF[rows_idx].copy(), D0[cols_idx].copy(), AL[cols_idx].copy(), Z[rows_idx].copy()
        P1, P2 = (cols_idx + 1).astype(np.int32), (rows_idx + 1).astype(np.int32)
        M, N = len(cols_idx), len(rows_idx)
    # generous work arrays: the reference's unmodified row search reads F at global
    # indices, so its F is passed with K entries (the node's values first)
    Fb = np.zeros(len(FF) + 1)
    Fb[:len(F)] = F
    U = np.zeros(M + 1)
    W = np.zeros(min(M, N) + 1)
    PQ = np.zeros(M + 1, np.int32)
    rk = ctypes.c_int(0)
    cnt = ctypes.c_int if wide else ctypes.c_longlong
    nf, ns = cnt(0), cnt(0)
    fn(rowcol, dp(Fb), dp(D), dp(A), dp(V), dp(difl), dp(difr), ctypes.c_double(tol), dp(U), dp(W), dp(FF),
       ip(P1), ip(P2), M, N, ctypes.byref(rk), ip(PQ), ctypes.byref(nf), ctypes.byref(ns))
    return rk.value, PQ[:M].copy(), P1, P2, ns.value


def test_rrlu_matches_corrected_reference():
    gens = secular_generators("random", 800, 5)
    K = len(gens[0])
    ref_fix, ref0 = ref_lib(fixed=True), ref_lib()
    OL = O.lib()
    rng = np.random.default_rng(7)
    col_same_unmodified = 0
    for trial in range(240):
        M = int(rng.integers(5, 120))
        N = int(rng.integers(5, 300))
        rows = rng.choice(K, M, replace=False)
        cols = rng.choice(np.setdiff1d(np.arange(K), rows), N, replace=False)
        if trial % 2 == 0:
            rows, cols = np.sort(rows), np.sort(cols)
        tol = (1e-17, 1e-13, 1e-8)[trial % 3]
        rowcol = (trial // 2) % 2
        a = big_stack(_rrlu, ref_fix.refd_rrlu, True, rowcol, None, None, rows, cols, gens, tol)
        b = _rrlu(OL.orc_rrlu, False, rowcol, None, None, rows, cols, gens, tol)
        assert a[0] == b[0], (trial, "rank")
        assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2]) and np.array_equal(a[3], b[3]), trial
        assert a[4] == b[4], (trial, "swaps")
        if rowcol == 1:
            c = big_stack(_rrlu, ref0.refd_rrlu, True, rowcol, None, None, rows, cols, gens, tol)
            assert c[0] == b[0] and np.array_equal(c[1], b[1]) and c[4] == b[4], (trial, "column RRLU")
            col_same_unmodified += 1
    assert col_same_unmodified == 120
