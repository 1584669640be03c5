"""Calls into the reference's own Fortran (oracle/_ref, built by oracle/build_ref.py;
test infrastructure only).

The reference keeps large automatic arrays on the stack (work arrays of order K x 200
and more in its HSS branches), so its calls run in a child process whose stack limit is
raised before it starts: `run(jobs)` pickles a list of (name, kwargs) jobs to a child
running this file, which loads the libraries, runs each job and pickles the results
back.  A child per batch also keeps the reference's own memory errors away from the
test process.

Jobs (numpy arrays in and out):
  stedc(d, e)                       MDSTEDC('I'): eigenvalues, vectors, info, merge trace
  bdsdc(d, e)                       MDBDSDC('U','I'): singular values, U, VT, info, trace
  sbevd(ab, kd)                     MDSBEVD('I','U'): eigenvalues, vectors, info, trace
  hss_build(gens, ni, tol, x)       DHSSEVC's construction: tree, node dims, X * H
  rrlu(cases)                       rrluCauchyEigrow / rrluCauchyEigcol on each case
  mat2hss(a, ni, tol)               mat2hssvd: tree and node dims
  cauchy2hss(d, f, u, v, ni, tol)   cauchy2hss (VP2): tree and node dims
  dhssmm(a, b, tol, ni, ncol)       DHSSMM: b := H b
  cauchyhssmm(d, f, u, v, ni, tol, ncol, b)   CauchyHssMM: b := H b
Every job takes fixed=True to use libhssref_fix.so (the row pivot test corrected).
"""
import ctypes
import os
import pickle
import resource
import subprocess
import sys
import tempfile

import numpy as np

REF_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref")
LIBS = {False: os.path.join(REF_DIR, "libhssref.so"), True: os.path.join(REF_DIR, "libhssref_fix.so")}


def available():
    return all(os.path.exists(p) for p in LIBS.values())


def _raise_stack():
    soft, hard = resource.getrlimit(resource.RLIMIT_STACK)
    resource.setrlimit(resource.RLIMIT_STACK, (hard, hard))


def run(jobs, timeout=600):
    with tempfile.TemporaryDirectory() as td:
        pin, pout = os.path.join(td, "in.pkl"), os.path.join(td, "out.pkl")
        with open(pin, "wb") as fh:
            pickle.dump(jobs, fh)
        subprocess.run([sys.executable, os.path.abspath(__file__), pin, pout], preexec_fn=_raise_stack, check=True,
                       timeout=timeout)
        with open(pout, "rb") as fh:
            return pickle.load(fh)


# ---------------------------------------------------------------- child side
_L = {}


def _lib(fixed):
    if fixed not in _L:
        lib = ctypes.CDLL(LIBS[fixed])
        lib.refd_trace.restype = ctypes.c_int
        _L[fixed] = lib
    return _L[fixed]


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def _trace(lib):
    cap = 1 << 16
    b = [np.zeros(cap, np.int32) for _ in range(4)]
    m = lib.refd_trace(cap, *[_ip(x) for x in b])
    return dict(n=b[0][:m].copy(), k=b[1][:m].copy(), hss=b[2][:m].copy(), nrot=b[3][:m].copy())


def j_stedc(d, e, fixed=False):
    lib = _lib(fixed)
    n = len(d)
    dd = np.array(d, np.float64)
    ee = np.r_[np.asarray(e, np.float64), 0.0]
    z = np.zeros((n, n), order="F")
    info = ctypes.c_int(0)
    lib.refd_stedc(n, _dp(dd), _dp(ee), _dp(z), n, ctypes.byref(info))
    return dict(d=dd, z=z, info=info.value, trace=_trace(lib))


def j_bdsdc(d, e, fixed=False):
    lib = _lib(fixed)
    n = len(d)
    dd = np.array(d, np.float64)
    ee = np.r_[np.asarray(e, np.float64), 0.0]
    u = np.zeros((n, n), order="F")
    vt = np.zeros((n, n), order="F")
    info = ctypes.c_int(0)
    lib.refd_bdsdc(n, _dp(dd), _dp(ee), _dp(u), _dp(vt), ctypes.byref(info))
    return dict(d=dd, u=u, vt=vt, info=info.value, trace=_trace(lib))


def j_sbevd(ab, kd, fixed=False):
    lib = _lib(fixed)
    a = np.array(ab, np.float64, order="F")
    n = a.shape[1]
    w = np.zeros(n)
    q = np.zeros((n, n), order="F")
    info = ctypes.c_int(0)
    lib.refd_sbevd(n, kd, _dp(a), a.shape[0], _dp(w), _dp(q), n, ctypes.byref(info))
    return dict(d=w, q=q, info=info.value, trace=_trace(lib))


def _dims(cap):
    return (np.zeros(cap, np.int32), np.zeros((cap, 2), np.int32), np.zeros((cap, 2), np.int32),
            np.zeros((cap, 2), np.int32), np.zeros(cap, np.int32))


def _dims_out(tr, u, v, b, d, m):
    return dict(tr=tr[:m].copy(), usz=u[:m].copy(), vsz=v[:m].copy(), bsz=b[:m].copy(), dsz=d[:m].copy())


def j_hss_build(gens, ni, tol, x, fixed=False):
    lib = _lib(fixed)
    f, d0, u, v, difl, difr = [np.array(a, np.float64) for a in gens]
    k = len(f)
    tr, us, vs, bs, ds = _dims(2 * (k // ni + 1))
    ltr, nf, ns, info = [ctypes.c_int(0) for _ in range(4)]
    xx = np.array(x, np.float64, order="F")
    lib.refd_hss_build(k, ni, _dp(f), _dp(d0), _dp(u), _dp(v), _dp(difl), _dp(difr), ctypes.c_double(tol),
                       xx.shape[0], _dp(xx), xx.shape[0], ctypes.byref(ltr), _ip(tr), _ip(us), _ip(vs), _ip(bs),
                       _ip(ds), ctypes.byref(nf), ctypes.byref(ns), ctypes.byref(info))
    out = _dims_out(tr, us, vs, bs, ds, ltr.value)
    out.update(x=xx, info=info.value)
    return out


def j_rrlu(cases, fixed=False):
    """cases: (rowcol, F, D, A, V, difl, difr, tol, FF, P1, P2, M, N); the arrays are
    passed as ComprcauchyEig passes them (CauchyHssEig.f90:2377, 2409)."""
    lib = _lib(fixed)
    out = []
    for rowcol, F, D, A, V, difl, difr, tol, FF, P1, P2, M, N in cases:
        F, D, A, V = [np.array(a, np.float64) for a in (F, D, A, V)]
        P1, P2 = np.array(P1, np.int32), np.array(P2, np.int32)
        FF, difl, difr = [np.array(a, np.float64) for a in (FF, difl, difr)]
        U, W, PQ = np.zeros(M + 1), np.zeros(min(M, N) + 1), np.zeros(M + 1, np.int32)
        rk, nf, ns = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
        lib.refd_rrlu(rowcol, _dp(F), _dp(D), _dp(A), _dp(V), _dp(difl), _dp(difr), ctypes.c_double(tol), _dp(U),
                      _dp(W), _dp(FF), _ip(P1), _ip(P2), M, N, ctypes.byref(rk), _ip(PQ), ctypes.byref(nf),
                      ctypes.byref(ns))
        out.append((rk.value, PQ[:M].copy(), P1, P2, ns.value))
    return out


def j_mat2hss(a, ni, tol, fixed=False):
    lib = _lib(fixed)
    a = np.array(a, np.float64, order="F")
    n = a.shape[0]
    tr, us, vs, bs, ds = _dims(2 * (n // ni + 1))
    ltr, prank, info = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
    lib.refd_mat2hss(n, _dp(a), n, ni, ctypes.c_double(tol), ctypes.byref(ltr), _ip(tr), _ip(us), _ip(vs), _ip(bs),
                     _ip(ds), ctypes.byref(prank), ctypes.byref(info))
    return _dims_out(tr, us, vs, bs, ds, ltr.value)


def j_cauchy2hss(d, f, u, v, ni, tol, fixed=False):
    lib = _lib(fixed)
    arr = [np.array(x, np.float64) for x in (d, f, u, v)]
    n = len(arr[0])
    tr, us, vs, bs, ds = _dims(2 * (n // ni + 1))
    ltr, info = ctypes.c_int(0), ctypes.c_int(0)
    lib.refd_cauchy2hss(n, ni, *[_dp(x) for x in arr], ctypes.c_double(tol), ctypes.byref(ltr), _ip(tr), _ip(us),
                        _ip(vs), _ip(bs), _ip(ds), ctypes.byref(info))
    return _dims_out(tr, us, vs, bs, ds, ltr.value)


def j_dhssmm(a, b, tol, ni, ncol, fixed=False):
    lib = _lib(fixed)
    a = np.array(a, np.float64, order="F")
    b = np.array(b, np.float64, order="F")
    info = ctypes.c_int(0)
    lib.refd_dhssmm(a.shape[0], _dp(a), a.shape[0], _dp(b), b.shape[0], b.shape[1], ctypes.c_double(tol), ni, ncol,
                    ctypes.byref(info))
    return dict(b=b, info=info.value)


def j_cauchyhssmm(d, f, u, v, ni, tol, ncol, b, fixed=False):
    lib = _lib(fixed)
    arr = [np.array(x, np.float64) for x in (d, f, u, v)]
    b = np.array(b, np.float64, order="F")
    lib.refd_cauchyhssmm(len(arr[0]), *[_dp(x) for x in arr], ni, ctypes.c_double(tol), ncol, _dp(b), b.shape[0])
    return dict(b=b)


if __name__ == "__main__":
    with open(sys.argv[1], "rb") as fh:
        jobs = pickle.load(fh)
    res = [globals()["j_" + name](**kw) for name, kw in jobs]
    with open(sys.argv[2], "wb") as fh:
        pickle.dump(res, fh)
