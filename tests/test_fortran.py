"""The Fortran drop-ins (MDSTEDC, MDLAED0, MDLAED1 -> MDLAED2/MDLAED3) called from a
Fortran program the way HSSPACK's users call them (tests/fortran/fdrive.f90)."""
import os
import subprocess

import numpy as np
import pytest

import oracle_py as O
import lapack_ref as L
from hsspack_amd.workloads import tridiag_cfg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FLIB = os.path.join(ROOT, "hsspack_amd", "libhsdc_f.so")
FDRIVE = os.path.join(ROOT, "tests", "fortran", "fdrive")
EPS = 2.0 ** -53


def run(mode, tmp_path, payload=b""):
    fin, fout = str(tmp_path / (mode + ".in")), str(tmp_path / (mode + ".out"))
    with open(fin, "wb") as f:
        f.write(payload)
    subprocess.run([FDRIVE, mode, fin, fout], check=True, timeout=300)
    return open(fout, "rb").read()


def i32(*v):
    return np.array(v, "<i4").tobytes()


def f64(a):
    return np.ascontiguousarray(a, "<f8").tobytes()


def test_library_exports_the_fortran_names():
    out = subprocess.run(["nm", "-D", "--defined-only", FLIB], capture_output=True, text=True, check=True).stdout
    for name in ("mdstedc_", "mdlaed0_", "mdlaed1_", "mdlaed2_", "mdlaed3_", "mdbdsdc_", "cauchyhssmm_",
                 "mdsbevd_", "dhssmm_", "dhssmml_"):
        assert (" T " + name) in out, name


def test_argument_errors_call_xerbla_like_the_reference(tmp_path):
    """Illegal arguments return INFO = -i and call XERBLA with the LAPACK routine name
    (mdstedc.f90:241, mdlaed0.f90:179, mdlaed1.f90:153, mdlaed3.f90:176, mdbdsdc.f90:174-187,
    mdsbevd.f90:290, whose name the 6-character log of the driver's XERBLA truncates); no
    device needed."""
    raw = run("errors", tmp_path)
    want = [(-1, "DSTEDC"), (-6, "DSTEDC"), (-8, "DSTEDC"), (-1, "DLAED0"), (-9, "DLAED0"), (-7, "DLAED1"),
            (-2, "DLAED3"), (-1, "DBDSDC"), (-2, "DBDSDC"), (-3, "DBDSDC"), (-7, "DBDSDC"), (-9, "DBDSDC"),
            (-1, "MDSBEV"), (-6, "MDSBEV")]
    rec = 14
    assert len(raw) == rec * len(want)
    for i, (info, name) in enumerate(want):
        r = raw[i * rec:(i + 1) * rec]
        a = np.frombuffer(r[:8], "<i4")
        assert (a[0], a[1], r[8:].decode()) == (info, -info, name)


def tri(d, e):
    return np.diag(d) + np.diag(e, 1) + np.diag(e, -1)


def check_eig(d, e, lam, z, tol=1e-12):
    T = tri(d, e)
    nrm = np.abs(T).max()
    orth = np.abs(z.T @ z - np.eye(len(d))).max()
    res = np.abs(T @ z - z * lam).max() / nrm
    assert orth <= tol and res <= tol, (orth, res)


@pytest.mark.gpu
@pytest.mark.parametrize("cfg,n,compz", [(1, 600, "I"), (2, 500, "I"), (1, 301, "V"), (2, 40, "N")])
def test_mdstedc(tmp_path, cfg, n, compz):
    d, e = tridiag_cfg(cfg, n, 77)
    payload = i32(n, ord(compz)) + f64(d) + f64(e)
    zin = None
    if compz == "V":
        zin = np.linalg.qr(np.random.default_rng(3).standard_normal((n, n)))[0]
        payload += f64(zin.T)  # column-major
    raw = run("stedc", tmp_path, payload)
    info, lwmin, liwmin = np.frombuffer(raw[:12], "<i4")
    lam = np.frombuffer(raw[12:12 + 8 * n], "<f8")
    z = np.frombuffer(raw[12 + 8 * n:], "<f8").reshape(n, n).T
    assert info == 0
    if compz == "I":
        assert (lwmin, liwmin) == (1 + 4 * n + n * n, 3 + 5 * n)
    ref = L.dstedc(d, e)[0]
    nrm = np.abs(tri(d, e)).max() * 3
    assert np.abs(lam - ref).max() <= 4 * n * EPS * nrm
    if compz == "I":
        check_eig(d, e, lam, z)
    elif compz == "V":
        # Z := Zin * Q with Q the tridiagonal's eigenvectors
        check_eig(d, e, lam, zin.T @ z)


@pytest.mark.gpu
@pytest.mark.parametrize("icompq", [2, 1])
def test_mdlaed0(tmp_path, icompq):
    n, qsiz = 400, 450
    d, e = tridiag_cfg(1, n, 78)
    payload = i32(icompq, qsiz, n) + f64(d) + f64(e)
    q0 = None
    if icompq == 1:
        q0 = np.random.default_rng(4).standard_normal((qsiz, n))
        payload += f64(q0.T)
    raw = run("laed0", tmp_path, payload)
    info = np.frombuffer(raw[:4], "<i4")[0]
    lam = np.frombuffer(raw[4:4 + 8 * n], "<f8")
    rows = qsiz if icompq == 1 else n
    q = np.frombuffer(raw[4 + 8 * n:], "<f8").reshape(n, rows).T
    assert info == 0
    ref, zref = L.dstedc(d, e)[:2]
    assert np.abs(lam - ref).max() <= 4 * n * EPS * 3 * np.abs(tri(d, e)).max()
    if icompq == 2:
        check_eig(d, e, lam, q)
    else:
        # Q0 * V: compare with Q0 * V_lapack up to the sign of each eigenvector
        qq = q0 @ zref
        sg = np.sign(np.sum(qq * q, axis=0))
        assert np.abs(q - qq * sg).max() <= 1e-10 * np.abs(qq).max()


@pytest.mark.gpu
@pytest.mark.parametrize("cfg,n", [(2, 300), (1, 600), (1, 2000)])
def test_mdlaed1_vs_oracle(tmp_path, cfg, n):
    """One merge through MDLAED1 -> MDLAED2 -> DLAED3 (n <= 600) or MDLAED3 (HSS branch,
    K >= 1500): eigenvalues and INDXQ as the oracle's MDLAED1, vectors to tolerance."""
    from golden.make_golden import merge_inputs
    d, e = tridiag_cfg(cfg, n, 79)
    cut = n // 2
    dd, q, indxq, rho, z = merge_inputs(d, e[: n - 1], cut)
    payload = i32(n, cut) + f64([rho]) + f64(dd) + f64(q.T) + i32(*indxq)
    raw = run("laed1", tmp_path, payload)
    info, wflag = np.frombuffer(raw[:8], "<i4")
    o = 8
    lam = np.frombuffer(raw[o:o + 8 * n], "<f8")
    o += 8 * n
    qq = np.frombuffer(raw[o:o + 8 * n * n], "<f8").reshape(n, n).T
    o += 8 * n * n
    iq = np.frombuffer(raw[o:], "<i4")
    r_info, r_d, r_q, r_iq, k, hss = O.laed1(dd, q, indxq, rho, cut)
    assert info == 0 and r_info == 0
    assert wflag == (1 if hss else 0)
    assert np.array_equal(lam[:k], r_d[:k])
    assert np.array_equal(lam[k:], r_d[k:])
    assert np.array_equal(iq, r_iq)
    assert np.abs(qq - r_q).max() <= 1e-11


@pytest.mark.gpu
@pytest.mark.parametrize("n,uplo", [(300, "U"), (777, "L"), (20, "U")])
def test_mdbdsdc_vs_oracle(tmp_path, n, uplo):
    """MDBDSDC from Fortran: singular values, U and VT bit-exact against the oracle."""
    from hsspack_amd.workloads import bidiag_cfg
    d, e = bidiag_cfg(n, seed=11)
    raw = run("bdsdc", tmp_path, i32(n, ord(uplo)) + f64(d) + f64(e))
    info = np.frombuffer(raw[:4], "<i4")[0]
    s = np.frombuffer(raw[4:4 + 8 * n], "<f8")
    u = np.frombuffer(raw[4 + 8 * n:4 + 8 * n + 8 * n * n], "<f8").reshape(n, n).T
    vt = np.frombuffer(raw[4 + 8 * n + 8 * n * n:], "<f8").reshape(n, n).T
    assert info == 0
    r = O.bdsdc(d, e, uplo=uplo.encode())
    assert np.array_equal(s, r["d"])
    assert np.array_equal(u, r["u"])
    assert np.array_equal(vt, r["vt"])


@pytest.mark.gpu
def test_mdbdsdc_hss_branch_from_fortran(tmp_path):
    """MDBDSDC from Fortran on a bidiagonal whose top merge reaches K >= 2000 (the
    reference's BSMLZ): the HSS branch, with singular values within 4 n eps of the
    oracle's and orthogonality / residual <= 1e-12."""
    from hsspack_amd.workloads import bidiag_toeplitz
    n = 2200
    d, e = bidiag_toeplitz(n, 7)
    raw = run("bdsdc", tmp_path, i32(n, ord("U")) + f64(d) + f64(e))
    info = np.frombuffer(raw[:4], "<i4")[0]
    s = np.frombuffer(raw[4:4 + 8 * n], "<f8")
    u = np.frombuffer(raw[4 + 8 * n:4 + 8 * n + 8 * n * n], "<f8").reshape(n, n).T
    vt = np.frombuffer(raw[4 + 8 * n + 8 * n * n:], "<f8").reshape(n, n).T
    assert info == 0
    r = O.bdsdc(d, e, trace_cap=4 * n)
    assert (r["trace"]["rank_max"] > 0).any()  # the oracle took the HSS branch too
    assert np.abs(s - r["d"]).max() <= 4 * n * 2.0 ** -53 * s[0]
    B = np.diag(d) + np.diag(e, 1)
    assert np.abs(u.T @ u - np.eye(n)).max() <= 1e-12
    assert np.abs(vt @ vt.T - np.eye(n)).max() <= 1e-12
    assert np.abs((u * s) @ vt - B).max() / np.abs(B).max() <= 1e-12


@pytest.mark.gpu
def test_cauchyhssmm_from_fortran(tmp_path):
    """CAUCHYHSSMM(D, F, U, V, N, NI, TOL, NCOL, B, LDB, TIME) as testCauchyHss_serial.f90
    calls it (interlaced points, NI = 120, NCOL = 128, TOL = 1e-17), twice in one program:
    the second call runs on the thread's kept context and gives the same product."""
    from test_oracle_cauchy import interlaced
    n, ni, ncol, tol = 1000, 120, 128, 1e-17
    d, f, u, v = interlaced(n, 21)
    b = np.random.default_rng(4).uniform(0, 1, (n, n))
    payload = i32(n, ni, ncol) + f64([tol]) + f64(d) + f64(f) + f64(u) + f64(v) + f64(b.T)
    raw = run("cauchy", tmp_path, payload)
    y1 = np.frombuffer(raw[:8 * n * n], "<f8").reshape(n, n).T
    y2 = np.frombuffer(raw[8 * n * n:], "<f8").reshape(n, n).T
    A = O.cauchy_dense_general(d, f, u, v)
    ref = A @ b
    assert np.abs(y1 - ref).max() <= 1e-12 * np.abs(ref).max()
    assert np.array_equal(y1, y2)


@pytest.mark.gpu
@pytest.mark.parametrize("uplo", ["U", "L"])
def test_mdsbevd_from_fortran(tmp_path, uplo):
    """MDSBEVD('I', UPLO, N, KD, AB, LDAB, D, Q, LDQ, ...) from Fortran: the same
    eigenpairs as the C entry on the same storage, eigenvalues against LAPACK dsbevd."""
    import scipy.linalg as sl
    from hsspack_amd import hsdc
    from hsspack_amd.workloads import band_cfg
    from test_gpu_band import lower_storage
    n, kd = 600, 8
    ab = band_cfg(n, kd, seed=6, sub=3)
    st = ab if uplo == "U" else lower_storage(ab)
    raw = run("sbevd", tmp_path, i32(n, kd, ord(uplo)) + f64(st.T))
    info = np.frombuffer(raw[:4], "<i4")[0]
    lam = np.frombuffer(raw[4:4 + 8 * n], "<f8")
    q = np.frombuffer(raw[4 + 8 * n:], "<f8").reshape(n, n).T
    assert info == 0
    ctx = hsdc.SbContext()
    w2, q2, info2 = ctx.sbevd(st, kd, uplo=uplo)
    ctx.close()
    assert info2 == 0
    assert np.array_equal(lam, w2) and np.array_equal(q, q2)
    w = sl.eig_banded(ab, lower=False, eigvals_only=True)
    assert np.abs(lam - w).max() <= 4 * n * 2.0 ** -53 * np.abs(w).max()


@pytest.mark.gpu
def test_dhssmm_and_dhssmml_from_fortran(tmp_path):
    """DHSSMM(A, LDA, M, N, B, LDB, K, TOL, PRANK, NI, NCOL, TIME, INFO) and DHSSMML as
    testfmm.f90 / testfmml.f90 call them (Cauchy matrix on interlaced points, PRANK = 45,
    NI = 64, NCOL = 100, TOL = 1e-11): the same products as the C entries."""
    from hsspack_amd import hsdc
    from test_oracle_dhssmm import fmm_matrix
    n, ni, ncol, tol = 1200, 64, 100, 1e-11
    A = fmm_matrix(n, 8)
    b = np.random.default_rng(9).uniform(0, 1, (n, n))
    raw = run("dhssmm", tmp_path, i32(n, ni, ncol) + f64([tol]) + f64(A.T) + f64(b.T))
    rec = 4 + 8 * n * n
    info1 = np.frombuffer(raw[:4], "<i4")[0]
    y1 = np.frombuffer(raw[4:rec], "<f8").reshape(n, n).T
    info2 = np.frombuffer(raw[rec:rec + 4], "<i4")[0]
    y2 = np.frombuffer(raw[rec + 4:], "<f8").reshape(n, n).T
    assert info1 == 0 and info2 == 0
    ctx = hsdc.Context()
    c1, _, _ = ctx.dhssmm(A, b, ni, tol, ncol=ncol)
    c2, _, _ = ctx.dhssmm(A, b, ni, tol, ncol=ncol, right=True)
    ctx.close()
    assert np.array_equal(y1, c1) and np.array_equal(y2, c2)
    assert np.abs(y1 - A @ b).max() <= 1e-8 and np.abs(y2 - b @ A).max() <= 1e-8
