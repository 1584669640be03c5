"""ctypes bindings to the LAPACK routines shipped inside scipy's OpenBLAS.

These are the third-party routines the reference path calls (DLAED4 at
mdlaed3.f90:208, DLAED2/DLAED3 behind mdlaed1.f90:184-199, DLAMRG at
mdlaed1.f90:221, DSTEQR at mdlaed0.f90:263) plus DSTEDC, the LAPACK driver
MDSTEDC replaces.  They serve as golden references for the oracle.  Test-only.
"""
import ctypes
import glob
import os

import numpy as np

_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        import scipy  # noqa: F401  (locates scipy.libs)
        base = os.path.join(os.path.dirname(scipy.__file__), os.pardir, "scipy.libs")
        cand = sorted(glob.glob(os.path.join(base, "libscipy_openblas*.so")))
        if not cand:
            raise RuntimeError("scipy OpenBLAS LAPACK not found")
        _LIB = ctypes.CDLL(cand[0])
    return _LIB


_i = ctypes.c_int
_d = ctypes.c_double


def _ip(x):
    return ctypes.byref(_i(x))


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ipa(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def dlaed4(d, z, rho, j):
    """Root j (1-based); returns (lam, delta, info)."""
    n = len(d)
    d = np.ascontiguousarray(d, np.float64)
    z = np.ascontiguousarray(z, np.float64)
    delta = np.zeros(n)
    lam = _d(0)
    info = _i(0)
    lib().scipy_dlaed4_(_ip(n), _ip(j), _dp(d), _dp(z), _dp(delta), ctypes.byref(_d(rho)),
                        ctypes.byref(lam), ctypes.byref(info))
    return lam.value, delta, info.value


def dlamrg(a, n1, n2, s1, s2):
    a = np.ascontiguousarray(a, np.float64)
    n1, n2, s1, s2 = int(n1), int(n2), int(s1), int(s2)
    idx = np.zeros(n1 + n2, np.int32)
    lib().scipy_dlamrg_(_ip(n1), _ip(n2), _dp(a), _ip(s1), _ip(s2), _ipa(idx))
    return idx


def dsteqr(d, e):
    n = len(d)
    d = np.array(d, np.float64)
    e = np.array(e, np.float64)
    z = np.zeros((n, n), order="F")
    work = np.zeros(max(1, 2 * n - 2))
    info = _i(0)
    lib().scipy_dsteqr_(ctypes.c_char_p(b"I"), _ip(n), _dp(d), _dp(e), _dp(z), _ip(max(1, n)),
                        _dp(work), ctypes.byref(info), ctypes.c_size_t(1))
    return d, z, info.value


def dstedc(d, e, compz=b"I"):
    n = len(d)
    d = np.array(d, np.float64)
    e = np.array(e, np.float64)
    z = np.zeros((n, n), order="F")
    lwork = 1 + 4 * n + n * n
    liwork = 3 + 5 * n
    work = np.zeros(lwork)
    iwork = np.zeros(liwork, np.int32)
    info = _i(0)
    lib().scipy_dstedc_(ctypes.c_char_p(compz), _ip(n), _dp(d), _dp(e), _dp(z), _ip(max(1, n)),
                        _dp(work), _ip(lwork), _ipa(iwork), _ip(liwork), ctypes.byref(info),
                        ctypes.c_size_t(1))
    return d, z, info.value


def dsterf(d, e):
    """LAPACK DSTERF: eigenvalues only (ascending)."""
    n = len(d)
    d = np.array(d, np.float64)
    e = np.array(e, np.float64) if n > 1 else np.zeros(1)
    info = _i(0)
    lib().scipy_dsterf_(_ip(n), _dp(d), _dp(e), ctypes.byref(info))
    return d, info.value


def dlaed2(d, q, indxq, rho, z, n1):
    """LAPACK DLAED2 on copies; returns a dict of every output."""
    n = len(d)
    d = np.array(d, np.float64)
    q = np.array(q, np.float64, order="F")
    indxq = np.array(indxq, np.int32)
    z = np.array(z, np.float64)
    dlamda = np.zeros(n)
    w = np.zeros(n)
    q2 = np.zeros(n * n + n)
    indx = np.zeros(n, np.int32)
    indxc = np.zeros(n, np.int32)
    indxp = np.zeros(n, np.int32)
    coltyp = np.zeros(n, np.int32)
    k = _i(0)
    rho_c = _d(rho)
    info = _i(0)
    lib().scipy_dlaed2_(ctypes.byref(k), _ip(n), _ip(n1), _dp(d), _dp(q), _ip(n), _ipa(indxq),
                        ctypes.byref(rho_c), _dp(z), _dp(dlamda), _dp(w), _dp(q2), _ipa(indx),
                        _ipa(indxc), _ipa(indxp), _ipa(coltyp), ctypes.byref(info))
    return dict(k=k.value, d=d, q=q, indxq=indxq, rho=rho_c.value, z=z, dlamda=dlamda, w=w,
                q2=q2, indx=indx, indxc=indxc, indxp=indxp, coltyp=coltyp, info=info.value)


def dlaed3(k, n, n1, q, rho, dlamda, q2, indx, ctot, w):
    d = np.zeros(n)
    q = np.array(q, np.float64, order="F")
    dlamda = np.array(dlamda, np.float64)
    q2 = np.array(q2, np.float64)
    indx = np.array(indx, np.int32)
    ctot = np.array(ctot, np.int32)
    w = np.array(w, np.float64)
    s = np.zeros(n * n + n)
    info = _i(0)
    lib().scipy_dlaed3_(_ip(k), _ip(n), _ip(n1), _dp(d), _dp(q), _ip(n), ctypes.byref(_d(rho)),
                        _dp(dlamda), _dp(q2), _ipa(indx), _ipa(ctot), _dp(w), _dp(s),
                        ctypes.byref(info))
    return d, q, info.value


# ---------------------------------------------------------------- bidiagonal SVD
# DLASDT, DLASD1-4, DLASDQ, DBDSDC: the LAPACK routines behind MDBDSDC / MDLASD0-2
# (mdlasd0.f90:166-183, mdlasd1.f90:239-244) and the driver MDBDSDC is built from.
def dlasdt(n, msub):
    inode = np.zeros(n, np.int32)
    ndiml = np.zeros(n, np.int32)
    ndimr = np.zeros(n, np.int32)
    lvl, nd = _i(0), _i(0)
    lib().scipy_dlasdt_(_ip(n), ctypes.byref(lvl), ctypes.byref(nd), _ipa(inode), _ipa(ndiml),
                        _ipa(ndimr), _ip(msub))
    return lvl.value, nd.value, inode[:nd.value], ndiml[:nd.value], ndimr[:nd.value]


def dlasd4(d, z, rho, j):
    n = len(d)
    d = np.ascontiguousarray(d, np.float64)
    z = np.ascontiguousarray(z, np.float64)
    delta = np.zeros(n)
    work = np.zeros(n)
    sigma = _d(0)
    info = _i(0)
    lib().scipy_dlasd4_(_ip(n), _ip(j), _dp(d), _dp(z), _dp(delta), ctypes.byref(_d(rho)),
                        ctypes.byref(sigma), _dp(work), ctypes.byref(info))
    return sigma.value, delta, work, info.value


def dlasd8(dsigma, z):
    """DLASD8 (ICOMPQ = 1): roots, updated z, DIFL, DIFR(:,1) and DIFR(:,2) = the 2-norms
    of the right vectors before normalisation (MDLASD8's APHAV holds their reciprocals),
    for a deflated merge with poles dsigma (dsigma[0] = 0)."""
    k = len(dsigma)
    d = np.zeros(k)
    zz = np.array(z, np.float64)
    vf = np.zeros(k)
    vl = np.zeros(k)
    difl = np.zeros(k)
    difr = np.zeros((k, 2), order="F")
    ds = np.array(dsigma, np.float64)
    work = np.zeros(3 * k)
    info = _i(0)
    lib().scipy_dlasd8_(_ip(1), _ip(k), _dp(d), _dp(zz), _dp(vf), _dp(vl), _dp(difl), _dp(difr), _ip(k), _dp(ds),
                        _dp(work), ctypes.byref(info))
    return dict(sigma=d, z=zz, difl=difl, difr=difr[:, 0].copy(), vnorm=difr[:, 1].copy(),
                info=info.value)


def dlasdq(d, e, sqre):
    """SVD of the n x (n+sqre) upper bidiagonal: (s ascending, U, VT, info)."""
    n = len(d)
    m = n + sqre
    d = np.array(d, np.float64)
    e = np.array(e, np.float64) if len(e) else np.zeros(1)
    vt = np.asfortranarray(np.eye(m))
    u = np.asfortranarray(np.eye(n))
    c = np.zeros((1, 1), order="F")
    work = np.zeros(4 * m + 8)
    info = _i(0)
    lib().scipy_dlasdq_(ctypes.c_char_p(b"U"), _ip(sqre), _ip(n), _ip(m), _ip(n), _ip(0), _dp(d),
                        _dp(e), _dp(vt), _ip(m), _dp(u), _ip(n), _dp(c), _ip(1), _dp(work),
                        ctypes.byref(info), ctypes.c_size_t(1))
    return d, u, vt, info.value


def dlasd2(nl, nr, sqre, d, alpha, beta, u, vt, idxq):
    n = nl + nr + 1
    m = n + sqre
    d = np.array(d, np.float64)
    u = np.array(u, np.float64, order="F")
    vt = np.array(vt, np.float64, order="F")
    idxq = np.array(idxq, np.int32)
    z = np.zeros(m)
    dsigma = np.zeros(n)
    u2 = np.zeros((n, n), order="F")
    vt2 = np.zeros((m, m), order="F")
    idxp, idx, idxc, coltyp = (np.zeros(n, np.int32) for _ in range(4))
    k, info = _i(0), _i(0)
    lib().scipy_dlasd2_(_ip(nl), _ip(nr), _ip(sqre), ctypes.byref(k), _dp(d), _dp(z),
                        ctypes.byref(_d(alpha)), ctypes.byref(_d(beta)), _dp(u), _ip(n), _dp(vt),
                        _ip(m), _dp(dsigma), _dp(u2), _ip(n), _dp(vt2), _ip(m), _ipa(idxp),
                        _ipa(idx), _ipa(idxc), _ipa(idxq), _ipa(coltyp), ctypes.byref(info))
    return dict(k=k.value, d=d, z=z, u=u, vt=vt, dsigma=dsigma, u2=u2, vt2=vt2, idxp=idxp,
                idx=idx, idxc=idxc, idxq=idxq, coltyp=coltyp, info=info.value)


def dlasd1(nl, nr, sqre, d, alpha, beta, u, vt, idxq):
    n = nl + nr + 1
    m = n + sqre
    d = np.array(d, np.float64)
    u = np.array(u, np.float64, order="F")
    vt = np.array(vt, np.float64, order="F")
    idxq = np.array(idxq, np.int32)
    iwork = np.zeros(4 * n, np.int32)
    work = np.zeros(3 * m * m + 2 * m)
    a, b, info = _d(alpha), _d(beta), _i(0)
    lib().scipy_dlasd1_(_ip(nl), _ip(nr), _ip(sqre), _dp(d), ctypes.byref(a), ctypes.byref(b),
                        _dp(u), _ip(n), _dp(vt), _ip(m), _ipa(idxq), _ipa(iwork), _dp(work),
                        ctypes.byref(info))
    return d, u, vt, idxq, info.value


def dbdsdc(d, e, uplo=b"U"):
    n = len(d)
    d = np.array(d, np.float64)
    e = np.array(e, np.float64) if n > 1 else np.zeros(1)
    u = np.zeros((n, n), order="F")
    vt = np.zeros((n, n), order="F")
    q = np.zeros(1)
    iq = np.zeros(1, np.int32)
    work = np.zeros(3 * n * n + 4 * n + 8)
    iwork = np.zeros(8 * n, np.int32)
    info = _i(0)
    lib().scipy_dbdsdc_(ctypes.c_char_p(uplo), ctypes.c_char_p(b"I"), _ip(n), _dp(d), _dp(e),
                        _dp(u), _ip(n), _dp(vt), _ip(n), _dp(q), _ipa(iq), _dp(work), _ipa(iwork),
                        ctypes.byref(info), ctypes.c_size_t(1), ctypes.c_size_t(1))
    return d, u, vt, info.value
