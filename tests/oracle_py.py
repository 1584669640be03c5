"""ctypes wrapper around the C oracle (oracle/liborc.so).  Test infrastructure."""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_SO = os.path.join(ROOT, "oracle", "liborc.so")
_LIB = None

_i = ctypes.c_int
_d = ctypes.c_double
_P = ctypes.POINTER


class Params(ctypes.Structure):
    _fields_ = [("smlsiz", _i), ("bsmlz", _i), ("ni", _i), ("tol", _d), ("maxit", _i),
                ("bd_bsmlz", _i), ("bd_ni", _i)]


class Trace(ctypes.Structure):
    _fields_ = [("cap", _i), ("nmerge", _i), ("level", _P(_i)), ("n", _P(_i)), ("k", _P(_i)),
                ("hss", _P(_i)), ("rank_max", _P(_i)), ("roots", ctypes.c_longlong),
                ("iters", ctypes.c_longlong), ("nrot", _P(_i)), ("indxq", _P(_i)), ("indxq_n", _i)]


class Hss(ctypes.Structure):
    _fields_ = [("k", _i), ("ltr", _i), ("lm", _i), ("lvl", _i),
                ("tr", _P(_i)), ("m", _P(_i)), ("ch", _P(_i)), ("lentr", _P(_i)), ("ttr", _P(_i)),
                ("dsz", _P(_i)), ("usz", _P(_i)), ("vsz", _P(_i)), ("bsz", _P(_i)),
                ("pu", _P(ctypes.c_long)), ("pv", _P(ctypes.c_long)), ("pb", _P(ctypes.c_long)),
                ("pd", _P(ctypes.c_long)), ("h", _P(_d)), ("dd", _P(_d)),
                ("nh", ctypes.c_long), ("ndd", ctypes.c_long),
                ("nflops", ctypes.c_longlong), ("nswap", ctypes.c_longlong)]


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(_SO):
            subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
        _LIB = ctypes.CDLL(_SO)
    return _LIB


def dp(a):
    return a.ctypes.data_as(_P(_d))


def ip(a):
    return a.ctypes.data_as(_P(_i))


def params(smlsiz=25, bsmlz=1500, ni=200, tol=None, maxit=100, bd_bsmlz=2000, bd_ni=40):
    p = Params()
    lib().orc_default_params(ctypes.byref(p))
    p.smlsiz, p.bsmlz, p.ni, p.maxit = smlsiz, bsmlz, ni, maxit
    p.bd_bsmlz, p.bd_ni = bd_bsmlz, bd_ni
    if tol is not None:
        p.tol = tol
    return p


def laed4(d, z, rho, j, maxit=100):
    n = len(d)
    d = np.ascontiguousarray(d, np.float64)
    z = np.ascontiguousarray(z, np.float64)
    delta = np.zeros(n)
    lam, tau = _d(0), _d(0)
    org, nit = _i(0), _i(0)
    info = lib().orc_laed4(n, j, dp(d), dp(z), dp(delta), _d(rho), ctypes.byref(lam),
                           ctypes.byref(org), ctypes.byref(tau), ctypes.byref(nit), maxit)
    return dict(lam=lam.value, delta=delta, org=org.value, tau=tau.value, niter=nit.value,
                info=info)


def laed4_mlp(d, z, rho, j, w, maxit=100, gate=0.0):
    """orc_laed4 started from the learned proposal of weights w (float32)."""
    n = len(d)
    d = np.ascontiguousarray(d, np.float64)
    z = np.ascontiguousarray(z, np.float64)
    w = np.ascontiguousarray(w, np.float32)
    delta = np.zeros(n)
    lam, tau = _d(0), _d(0)
    org, nit, used = _i(0), _i(0), _i(0)
    info = lib().orc_laed4_mlp(n, j, dp(d), dp(z), dp(delta), _d(rho), ctypes.byref(lam),
                               ctypes.byref(org), ctypes.byref(tau), ctypes.byref(nit), maxit,
                               w.ctypes.data_as(_P(ctypes.c_float)), ctypes.c_float(gate),
                               ctypes.byref(used))
    return dict(lam=lam.value, org=org.value, tau=tau.value, niter=nit.value, info=info,
                used=used.value)


def mlp_forward(w, x):
    """C restatement of the FP32 forward pass, one sample."""
    w = np.ascontiguousarray(w, np.float32)
    x = np.ascontiguousarray(x, np.float32)
    L = lib()
    L.orc_mlp_forward.restype = ctypes.c_float
    return L.orc_mlp_forward(w.ctypes.data_as(_P(ctypes.c_float)), x.ctypes.data_as(_P(ctypes.c_float)))


def sec_team(n):
    return int(lib().orc_sec_team(int(n)))


def sec_exp(x):
    L = lib()
    L.orc_sec_exp.restype = _d
    return L.orc_sec_exp(_d(x))


def mlp_features(d, z, rho, r, W, xc):
    d = np.ascontiguousarray(d, np.float64)
    z = np.ascontiguousarray(z, np.float64)
    f = np.zeros(34, np.float32)
    lib().orc_mlp_features(len(d), int(r), dp(d), dp(z), _d(rho), _d(W), _d(xc),
                           f.ctypes.data_as(_P(ctypes.c_float)))
    return f


def laed1(d, q, indxq, rho, cutpnt, prm=None):
    """MDLAED1 on copies: returns (info, d, q, indxq, k, hss)."""
    n = len(d)
    d = np.array(d, np.float64)
    q = np.array(q, np.float64, order="F")
    indxq = np.array(indxq, np.int32)
    work = np.zeros(4 * n + n * n + 16)
    iwork = np.zeros(4 * n + 16, np.int32)
    prm = prm or params()
    k, hss, rmax = _i(0), _i(0), _i(0)
    it = ctypes.c_longlong(0)
    info = lib().orc_laed1(n, dp(d), dp(q), n, ip(indxq), _d(rho), int(cutpnt), dp(work), ip(iwork),
                           ctypes.byref(prm), ctypes.byref(k), ctypes.byref(hss), ctypes.byref(rmax),
                           ctypes.byref(it))
    return info, d, q, indxq, k.value, hss.value


def dlamrg(a, n1, n2, s1, s2):
    a = np.ascontiguousarray(a, np.float64)
    idx = np.zeros(n1 + n2, np.int32)
    lib().orc_dlamrg(int(n1), int(n2), dp(a), int(s1), int(s2), ip(idx))
    return idx


def steqr(d, e):
    n = len(d)
    d = np.array(d, np.float64)
    e = np.array(e, np.float64) if n > 1 else np.zeros(1)
    z = np.zeros((n, n), order="F")
    info = lib().orc_steqr(n, dp(d), dp(e), dp(z), max(1, n))
    return d, z, info


def laed2(d, q, indxq, rho, z, n1, bsmlz=1 << 30):
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
    hflag = _i(0)
    nrot = _i(0)
    rot_pj = np.zeros(n, np.int32)
    rot_nj = np.zeros(n, np.int32)
    rot_c = np.zeros(n)
    rot_s = np.zeros(n)
    rho_c = _d(rho)
    info = lib().orc_laed2(ctypes.byref(k), n, n1, dp(d), dp(q), n, ip(indxq), ctypes.byref(rho_c),
                           dp(z), dp(dlamda), dp(w), dp(q2), ip(indx), ip(indxc), ip(indxp),
                           ip(coltyp), bsmlz, ctypes.byref(hflag), ctypes.byref(nrot), ip(rot_pj),
                           ip(rot_nj), dp(rot_c), dp(rot_s))
    r = nrot.value
    return dict(k=k.value, d=d, q=q, indxq=indxq, rho=rho_c.value, z=z, dlamda=dlamda, w=w,
                q2=q2, indx=indx, indxc=indxc, indxp=indxp, coltyp=coltyp, info=info,
                hflag=hflag.value, rot=(rot_pj[:r], rot_nj[:r], rot_c[:r], rot_s[:r]))


def laed3(k, n, n1, q, rho, dlamda, q2, indx, ctot, w, maxit=100):
    d = np.zeros(n)
    q = np.array(q, np.float64, order="F")
    dlamda = np.array(dlamda, np.float64)
    q2 = np.array(q2, np.float64)
    indx = np.array(indx, np.int32)
    ctot = np.array(ctot, np.int32)
    w = np.array(w, np.float64)
    s = np.zeros(n * n + n)
    it = ctypes.c_longlong(0)
    info = lib().orc_laed3(k, n, n1, dp(d), dp(q), n, _d(rho), dp(dlamda), dp(q2), ip(indx),
                           ip(ctot), dp(w), dp(s), maxit, ctypes.byref(it))
    return d, q, info


def stedc(d, e, compz=b"I", prm=None, trace_cap=0, z=None):
    n = len(d)
    d = np.array(d, np.float64)
    e = np.array(e, np.float64) if n > 1 else np.zeros(1)
    if z is None:
        z = np.zeros((n, n), order="F")
    else:
        z = np.array(z, np.float64, order="F")
    prm = prm or params()
    lgn = max(1, int(np.ceil(np.log2(max(n, 2)))) + 1)
    if compz in (b"V", "V"):
        lwork = 1 + 3 * n + 2 * n * lgn + 4 * n * n
        liwork = 6 + 6 * n + 5 * n * lgn
    else:
        lwork = 1 + 4 * n + n * n
        liwork = 3 + 5 * n
    work = np.zeros(max(1, lwork))
    iwork = np.zeros(max(1, liwork), np.int32)
    tr = None
    bufs = None
    if trace_cap:
        tr = Trace()
        bufs = [np.zeros(trace_cap, np.int32) for _ in range(6)]
        qbuf = np.zeros(max(n, 1), np.int32)
        tr.cap = trace_cap
        tr.nmerge = 0
        tr.level, tr.n, tr.k, tr.hss, tr.rank_max, tr.nrot = [ip(b) for b in bufs]
        tr.indxq = ip(qbuf)
        tr.indxq_n = n
        tr.roots = 0
        tr.iters = 0
    c = compz if isinstance(compz, bytes) else compz.encode()
    info = lib().orc_stedc(ctypes.c_char(c), n, dp(d), dp(e), dp(z), max(1, n), dp(work), lwork,
                           ip(iwork), liwork, ctypes.byref(prm),
                           ctypes.byref(tr) if tr is not None else None)
    out = dict(d=d, z=z, info=info, e=e)
    if tr is not None:
        m = min(tr.nmerge, trace_cap)
        out["trace"] = dict(level=bufs[0][:m].copy(), n=bufs[1][:m].copy(), k=bufs[2][:m].copy(),
                            hss=bufs[3][:m].copy(), rank_max=bufs[4][:m].copy(), nrot=bufs[5][:m].copy(),
                            indxq=qbuf[:n].copy(), roots=tr.roots, iters=tr.iters)
    return out


def hss_build(f, d0, u, v, difl, difr, ni, tol):
    k = len(f)
    H = Hss()
    arrs = [np.ascontiguousarray(a, np.float64) for a in (f, d0, u, v, difl, difr)]
    info = lib().orc_hss_build(ctypes.byref(H), k, ni, *[dp(a) for a in arrs], _d(tol))
    if info != 0:
        raise RuntimeError("orc_hss_build failed %d" % info)
    return H


def hss_build_sq(f, d0, u, v, difl, difr, ni, tol):
    """HSS form of u_i v_j / (f_i^2 - d0_j^2) (the SVD's vector matrices)."""
    k = len(f)
    H = Hss()
    arrs = [np.ascontiguousarray(a, np.float64) for a in (f, d0, u, v, difl, difr)]
    info = lib().orc_hss_build_sq(ctypes.byref(H), k, ni, *[dp(a) for a in arrs], _d(tol))
    if info != 0:
        raise RuntimeError("orc_hss_build_sq failed %d" % info)
    return H


def cauchy_dense_sq(f, d0, u, v, difl, difr):
    k = len(f)
    a = np.zeros((k, k), order="F")
    arrs = [np.ascontiguousarray(x, np.float64) for x in (f, d0, u, v, difl, difr)]
    lib().orc_cauchy_dense_sq(k, *[dp(x) for x in arrs], dp(a), k)
    return a


def lasd8(dsigma, z, maxit=100):
    """MDLASD8: roots, updated z, DIFL, DIFR, APHAU, APHAV of one deflated SVD merge."""
    k = len(dsigma)
    ds = np.ascontiguousarray(dsigma, np.float64)
    zz = np.array(z, np.float64)
    out = [np.zeros(k) for _ in range(5)]
    it = ctypes.c_longlong(0)
    info = lib().orc_lasd8(k, dp(out[0]), dp(ds), dp(zz), dp(out[1]), dp(out[2]), dp(out[3]), dp(out[4]),
                           maxit, ctypes.byref(it))
    return dict(sigma=out[0], z=zz, difl=out[1], difr=out[2], aphau=out[3], aphav=out[4], info=info,
                iters=it.value)


def hss_dense(H):
    k = H.k
    a = np.zeros((k, k), order="F")
    lib().orc_hss_dense(ctypes.byref(H), dp(a), k)
    return a


def hss_mmL(H, x):
    x = np.array(x, np.float64, order="F")
    lib().orc_hss_mmL(ctypes.byref(H), dp(x), x.shape[0], x.shape[0])
    return x


def hss_ranks(H):
    usz = np.ctypeslib.as_array(H.usz, (2 * H.ltr,)).reshape(-1, 2)
    vsz = np.ctypeslib.as_array(H.vsz, (2 * H.ltr,)).reshape(-1, 2)
    return usz.copy(), vsz.copy()


def hss_free(H):
    lib().orc_hss_free(ctypes.byref(H))


def cauchy2hss(d, f, u, v, ni, tol):
    """HSS form of the general Cauchy-like u_i v_j / (d_i - f_j) (cauchy2hss, VP2)."""
    n = len(d)
    H = Hss()
    arrs = [np.ascontiguousarray(a, np.float64) for a in (d, f, u, v)]
    info = lib().orc_cauchy2hss(ctypes.byref(H), n, ni, *[dp(a) for a in arrs], _d(tol))
    if info != 0:
        raise RuntimeError("orc_cauchy2hss failed %d" % info)
    return H


def mat2hss(a, ni, tol):
    """HSS form of a dense square matrix by truncated SVDs (mat2hssvd of DHSSMM / DHSSMML)."""
    a = np.array(a, np.float64, order="F")
    n = a.shape[0]
    H = Hss()
    info = lib().orc_mat2hss(ctypes.byref(H), n, ni, dp(a), n, _d(tol))
    if info != 0:
        raise RuntimeError("orc_mat2hss failed %d" % info)
    return H


def mat2hss_rand(a, ni, tol, seed):
    """mat2hss with mat2hssvdr's randomized range finding, Omega from the seed's stream."""
    a = np.array(a, np.float64, order="F")
    n = a.shape[0]
    H = Hss()
    info = lib().orc_mat2hss_rand(ctypes.byref(H), n, ni, dp(a), n, _d(tol), ctypes.c_ulonglong(seed))
    if info != 0:
        raise RuntimeError("orc_mat2hss_rand failed %d" % info)
    return H


def cauchy_dense_general(d, f, u, v):
    n = len(d)
    a = np.zeros((n, n), order="F")
    arrs = [np.ascontiguousarray(x, np.float64) for x in (d, f, u, v)]
    lib().orc_cauchy_dense_general(n, *[dp(x) for x in arrs], dp(a), n)
    return a


class BandTrace(ctypes.Structure):
    _fields_ = [("cap", _i), ("nupd", _i), ("nmerge", _i), ("k", _P(_i)), ("n", _P(_i)), ("nrot", _P(_i)),
                ("iters", ctypes.c_longlong), ("sv", _P(_d))]


def sym_jacobi(a):
    m = a.shape[0]
    a = np.array(a, np.float64, order="F")
    w, v = np.zeros(m), np.zeros((m, m), order="F")
    sw = lib().orc_sym_jacobi(m, dp(a), m, dp(w), dp(v), m)
    return w, v, sw


def jacobi_svd(c):
    m = c.shape[0]
    c = np.array(c, np.float64, order="F")
    s, u, v = np.zeros(m), np.zeros((m, m), order="F"), np.zeros((m, m), order="F")
    sw = lib().orc_jacobi_svd(m, dp(c), m, dp(s), dp(u), m, dp(v), m)
    return s, u, v, sw


def sbevd(ab, kd, smlsiz=None, trace_cap=0):
    """MDSBEVD (COMPZ 'I', UPLO 'U') on the upper band storage ab (kd+1 x n)."""
    ab = np.array(ab, np.float64, order="F")
    n = ab.shape[1]
    smlsiz = smlsiz or max(2 * kd - 1, 25)
    d, q = np.zeros(n), np.zeros((n, n), order="F")
    tr = None
    if trace_cap:
        tr = BandTrace()
        bufs = [np.zeros(trace_cap, np.int32) for _ in range(3)]
        ncut = n // max(1, kd) + 2
        sv = np.zeros(max(1, ncut * kd))
        tr.cap = trace_cap
        tr.k, tr.n, tr.nrot = [ip(b) for b in bufs]
        tr.sv = dp(sv)
    info = lib().orc_sbevd(n, kd, dp(ab), ab.shape[0], dp(d), dp(q), max(n, 1), smlsiz,
                           ctypes.byref(tr) if tr is not None else None)
    out = dict(d=d, q=q, info=info)
    if tr is not None:
        m = min(tr.nupd, trace_cap)
        out["trace"] = dict(k=bufs[0][:m].copy(), n=bufs[1][:m].copy(), nrot=bufs[2][:m].copy(),
                            nmerge=tr.nmerge, iters=tr.iters, sv=sv)
    return out


def cauchy_dense(f, u, v, difl, difr):
    k = len(f)
    a = np.zeros((k, k), order="F")
    arrs = [np.ascontiguousarray(x, np.float64) for x in (f, u, v, difl, difr)]
    lib().orc_cauchy_dense(k, *[dp(x) for x in arrs], dp(a), k)
    return a


def _setup():
    L = lib()
    for f in ("orc_bd_leaf", "orc_lasd4", "orc_lasd2", "orc_lasd3", "orc_lasd1", "orc_lasd0",
              "orc_bdsdc"):
        getattr(L, f).restype = _i
    L.orc_laed4.restype = _i
    L.orc_laed4_mlp.restype = _i
    L.orc_stedc.restype = _i
    L.orc_laed2.restype = _i
    L.orc_laed3.restype = _i
    L.orc_steqr.restype = _i
    L.orc_hss_build.restype = _i
    L.orc_cauchy2hss.restype = _i
    L.orc_mat2hss.restype = _i
    L.orc_mat2hss_rand.restype = _i
    L.orc_sbevd.restype = _i
    L.orc_sym_jacobi.restype = _i
    L.orc_jacobi_svd.restype = _i
    L.orc_hss_build_sq.restype = _i
    L.orc_lasd8.restype = _i
    L.orc_hss_tree_ok.restype = _i


def mdlaed3_vectors(dlamda, z, rho, maxit=100):
    """Roots, updated z, DIFL, DIFR, ALPHA of MDLAED3 (no eigenvector update)."""
    k = len(dlamda)
    dl = np.array(dlamda, np.float64)
    zz = np.array(z, np.float64)
    d = np.zeros(k)
    difl = np.zeros(k)
    difr = np.zeros(k)
    alpha = np.zeros(k)
    it = ctypes.c_longlong(0)
    info = lib().orc_mdlaed3_vectors(k, dp(d), _d(rho), dp(dl), dp(zz), dp(difl), dp(difr),
                                     dp(alpha), maxit, ctypes.byref(it))
    if info != 0:
        raise RuntimeError("mdlaed3_vectors info=%d" % info)
    return dict(d=d, dlamda=dl, z=zz, difl=difl, difr=difr, alpha=alpha, iters=it.value)


# ---------------------------------------------------------------- bidiagonal SVD
def lasdt(n, msub):
    inode = np.zeros(max(n, 1), np.int32)
    ndiml = np.zeros(max(n, 1), np.int32)
    ndimr = np.zeros(max(n, 1), np.int32)
    lvl, nd = _i(0), _i(0)
    lib().orc_lasdt(n, ctypes.byref(lvl), ctypes.byref(nd), ip(inode), ip(ndiml), ip(ndimr), msub)
    return lvl.value, nd.value, inode[:nd.value], ndiml[:nd.value], ndimr[:nd.value]


def bd_leaf(d, e, sqre):
    """(s ascending, U, VT, sweeps) of the m x (m+sqre) upper bidiagonal leaf."""
    m = len(d)
    p = m + sqre
    d = np.array(d, np.float64)
    e = np.array(e, np.float64) if len(e) else np.zeros(1)
    u = np.zeros((m, m), order="F")
    vt = np.zeros((p, p), order="F")
    sw = lib().orc_bd_leaf(m, sqre, dp(d), dp(e), dp(u), m, dp(vt), p)
    return d, u, vt, sw


def lasd4(d, z, rho, j, maxit=100):
    k = len(d)
    d = np.ascontiguousarray(d, np.float64)
    z = np.ascontiguousarray(z, np.float64)
    sg, tau = _d(0), _d(0)
    org, nit = _i(0), _i(0)
    info = lib().orc_lasd4(k, j, dp(d), dp(z), _d(rho), ctypes.byref(sg), ctypes.byref(org),
                           ctypes.byref(tau), ctypes.byref(nit), maxit)
    return dict(sigma=sg.value, org=org.value, tau=tau.value, niter=nit.value, info=info)


def lasd2(nl, nr, sqre, d, alpha, beta, u, vt, idxq):
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
    k, nrot = _i(0), _i(0)
    rp, rq = np.zeros(n, np.int32), np.zeros(n, np.int32)
    rc, rs = np.zeros(n), np.zeros(n)
    lib().orc_lasd2(nl, nr, sqre, ctypes.byref(k), dp(d), dp(z), _d(alpha), _d(beta), dp(u), n,
                    dp(vt), m, dp(dsigma), dp(u2), n, dp(vt2), m, ip(idxp), ip(idx), ip(idxc),
                    ip(idxq), ip(coltyp), ctypes.byref(nrot), ip(rp), ip(rq), dp(rc), dp(rs))
    r = nrot.value
    return dict(k=k.value, d=d, z=z, u=u, vt=vt, dsigma=dsigma, u2=u2, vt2=vt2, idxp=idxp, idx=idx,
                idxc=idxc, idxq=idxq, coltyp=coltyp, rot=(rp[:r], rq[:r], rc[:r], rs[:r]))


def lasd1(nl, nr, sqre, d, alpha, beta, u, vt, idxq, prm=None):
    n = nl + nr + 1
    m = n + sqre
    d = np.array(d, np.float64)
    u = np.array(u, np.float64, order="F")
    vt = np.array(vt, np.float64, order="F")
    idxq = np.array(idxq, np.int32)
    a, b = _d(alpha), _d(beta)
    k, nrot, rk = _i(0), _i(0), _i(0)
    it = ctypes.c_longlong(0)
    prm = prm or params()
    info = lib().orc_lasd1(nl, nr, sqre, dp(d), ctypes.byref(a), ctypes.byref(b), dp(u), n, dp(vt),
                           m, ip(idxq), ctypes.byref(prm), ctypes.byref(k), ctypes.byref(it),
                           ctypes.byref(nrot), ctypes.byref(rk))
    return dict(d=d, u=u, vt=vt, idxq=idxq, k=k.value, iters=it.value, nrot=nrot.value,
                hss_rank=rk.value, info=info)


def bdsdc(d, e, uplo=b"U", compq=b"I", prm=None, trace_cap=0):
    n = len(d)
    d = np.array(d, np.float64)
    e = np.array(e, np.float64) if n > 1 else np.zeros(1)
    u = np.zeros((n, n), order="F")
    vt = np.zeros((n, n), order="F")
    prm = prm or params()
    tr = None
    bufs = None
    if trace_cap:
        tr = Trace()
        bufs = [np.zeros(trace_cap, np.int32) for _ in range(5)]
        tr.cap = trace_cap
        tr.nmerge = 0
        tr.level, tr.n, tr.k, tr.hss, tr.rank_max = [ip(b) for b in bufs]
        tr.roots = 0
        tr.iters = 0
    info = lib().orc_bdsdc(ctypes.c_char(uplo), ctypes.c_char(compq), n, dp(d), dp(e), dp(u),
                           max(1, n), dp(vt), max(1, n), ctypes.byref(prm),
                           ctypes.byref(tr) if tr is not None else None)
    out = dict(d=d, u=u, vt=vt, info=info)
    if tr is not None:
        m = min(tr.nmerge, trace_cap)
        out["trace"] = dict(level=bufs[0][:m].copy(), n=bufs[1][:m].copy(), k=bufs[2][:m].copy(),
                            nrot=bufs[3][:m].copy(), rank_max=bufs[4][:m].copy(), roots=tr.roots,
                            iters=tr.iters)
    return out
