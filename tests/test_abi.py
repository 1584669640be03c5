"""CPU checks of the C ABI: the library loads, exports every declared symbol, and
validates arguments exactly as MDSTEDC does (no device needed)."""
import ctypes
import os
import re

import numpy as np
import pytest

from hsspack_amd import hsdc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "hsdc.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|hsdc_\w*ctx\s*\*)\s*(hsdc_\w+)\s*\(", txt, re.M)))


def test_library_exports_every_declared_symbol():
    L = hsdc.lib()
    syms = declared_symbols()
    assert len(syms) >= 10
    for s in syms:
        assert hasattr(L, s), s
    assert set(syms) == set(hsdc.EXPORTS)


def test_stedc_argument_errors_and_query():
    ctx = hsdc.Context()
    n = 50
    d = np.zeros(n)
    e = np.zeros(n)
    L = hsdc.lib()
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    ip = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    z = np.zeros((n, n), order="F")
    work = np.zeros(2)
    iwork = np.zeros(2, np.int32)
    h = ctx._h
    assert L.hsdc_stedc(h, b"X", n, dp(d), dp(e), dp(z), n, dp(work), 1, ip(iwork), 1) == -1
    assert L.hsdc_stedc(h, b"I", -1, dp(d), dp(e), dp(z), n, dp(work), 1, ip(iwork), 1) == -2
    assert L.hsdc_stedc(h, b"I", n, dp(d), dp(e), dp(z), n - 1, dp(work), 1, ip(iwork), 1) == -6
    assert L.hsdc_stedc(h, b"I", n, dp(d), dp(e), dp(z), n, dp(work), 1, ip(iwork), 1) == -8
    assert L.hsdc_stedc(h, b"I", n, dp(d), dp(e), dp(z), n, dp(work), 1 + 4 * n + n * n, ip(iwork), 1) == -10
    assert L.hsdc_stedc(h, b"I", n, dp(d), dp(e), dp(z), n, dp(work), -1, ip(iwork), 1) == 0
    assert work[0] == 1 + 4 * n + n * n and iwork[0] == 3 + 5 * n
    assert L.hsdc_stedc(h, b"V", n, dp(d), dp(e), dp(z), n, dp(work), -1, ip(iwork), 1) == 0
    lgn = int(np.ceil(np.log2(n)))
    assert work[0] == 1 + 3 * n + 2 * n * lgn + 4 * n * n
    ctx.close()


def test_params_validation():
    ctx = hsdc.Context()
    with pytest.raises(ValueError):
        ctx.set_params(smlsiz=100)
    ctx.set_params(smlsiz=25, hss_min_k=1024, hss_leaf=128, hss_tol=1e-13)
    ctx.close()


def test_bdsdc_argument_errors_without_device():
    """MDBDSDC's checks (mdbdsdc.f90:174-187) answer before any device work."""
    L = hsdc.lib()
    ctx = hsdc.BdContext()
    n = 8
    d = np.ones(n)
    e = np.ones(n)
    u = np.zeros((n, n), order="F")
    vt = np.zeros((n, n), order="F")
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    c = ctypes.c_char
    calls = [((b"X", b"I", n, n, n), -1), ((b"U", b"X", n, n, n), -2), ((b"U", b"P", n, n, n), -2),
             ((b"U", b"I", -1, n, n), -3), ((b"U", b"I", n, n - 1, n), -7), ((b"L", b"I", n, n, 0), -9),
             ((b"U", b"N", n, 1, 1), None)]
    for (uplo, compq, nn, ldu, ldvt), want in calls:
        if want is None:
            continue
        rc = L.hsdc_bdsdc(ctx._h, c(uplo), c(compq), nn, dp(d), dp(e), dp(u), ldu, dp(vt), ldvt,
                          None, None, None, None)
        assert rc == want, (uplo, compq, nn, ldu, ldvt, rc)
    assert L.hsdc_bdsdc(ctx._h, c(b"U"), c(b"I"), 0, dp(d), dp(e), dp(u), 1, dp(vt), 1,
                        None, None, None, None) == 0
    p = hsdc.BdParams()
    L.hsdc_bd_default_params(ctypes.byref(p))
    assert (p.smlsiz, p.maxit) == (25, 100)
    # the HSS branch: BSMLZ 2000 (mdlasd2.f90:172), Ni 40 (mdlasd8.f90:277), TOL 1.0E-17
    assert (p.hss_min_k, p.hss_leaf) == (2000, 40) and p.hss_tol == float(np.float32(1e-17))
    for field, bad, want in (("hss_min_k", 0, -4), ("hss_leaf", 4, -5), ("hss_tol", -1.0, -6)):
        q = hsdc.BdParams()
        L.hsdc_bd_default_params(ctypes.byref(q))
        setattr(q, field, bad)
        assert L.hsdc_bd_set_params(ctx._h, ctypes.byref(q)) == want, field
    p.smlsiz = 2
    assert L.hsdc_bd_set_params(ctx._h, ctypes.byref(p)) == -2
    ctx.close()


def test_hss_cmm_argument_errors_without_device():
    """hsdc_hss_cmm rejects a bad side, mode or column count before any device work."""
    L = hsdc.lib()
    ctx = hsdc.Context()
    k = 4
    a = np.ones(k)
    x = np.zeros((k, 2), order="F")
    dp = lambda v: v.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    args = lambda side, sq, ncol: (ctx._h, ctypes.c_char(side), sq, k, 8, 1e-17, dp(a), dp(a), dp(a), dp(a), dp(a),
                                   dp(a), dp(x), k, ncol, None)
    assert L.hsdc_hss_cmm(*args(b"X", 0, 2)) == -2
    assert L.hsdc_hss_cmm(*args(b"L", 2, 2)) == -3
    assert L.hsdc_hss_cmm(*args(b"L", 0, -1)) == -15
    assert L.hsdc_hss_cmm(*args(b"R", 1, 0)) == 0
    ctx.close()


def test_cauchyhssmm_argument_errors():
    """CauchyHssMM's arguments are checked before any device work (LAPACK order:
    every argument, then the quick return for n = 0)."""
    ctx = hsdc.Context()
    L = hsdc.lib()
    n = 40
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    gap = 0.01
    f = 0.1 + (2 * np.arange(1, n + 1) - 1) * gap
    d = f + gap
    u, v = np.ones(n), np.ones(n)
    b = np.zeros((n, n), order="F")
    t = np.zeros(2)
    h = ctx._h
    call = lambda nn, ni, tol, ncol, ldb, dd=d: L.hsdc_cauchyhssmm(h, dp(dd), dp(f), dp(u), dp(v), nn, ni, tol, ncol,
                                                                   dp(b), ldb, dp(t), None)
    assert call(-1, 16, 1e-12, 8, n) == -5
    assert call(n, 4, 1e-12, 8, n) == -6
    assert call(n, 16, -1.0, 8, n) == -7
    assert call(n, 16, 1e-12, 0, n) == -8
    assert call(n, 16, 1e-12, 8, n - 1) == -10
    assert call(0, 16, 1e-12, 8, 1) == 0
    dd = d.copy()
    dd[7] = f[3]  # A(8, 4) would divide by zero
    assert call(n, 16, 1e-12, 8, n, dd) == -1
    dd[7] = np.nan
    assert call(n, 16, 1e-12, 8, n, dd) == -1
    ctx.close()


def test_hss_cmm_checks_arguments_before_the_quick_return():
    ctx = hsdc.Context()
    L = hsdc.lib()
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    a = np.ones(8)
    x = np.zeros((8, 8), order="F")
    h = ctx._h
    # ncol = 0 is a quick return only once k, ni and ldx are valid
    assert L.hsdc_hss_cmm(h, b"R", 0, 0, 16, 1e-12, *[dp(a)] * 6, dp(x), 8, 0, None) == -4
    assert L.hsdc_hss_cmm(h, b"R", 0, 8, 4, 1e-12, *[dp(a)] * 6, dp(x), 8, 0, None) == -5
    assert L.hsdc_hss_cmm(h, b"L", 0, 8, 16, 1e-12, *[dp(a)] * 6, dp(x), 4, 0, None) == -14
    assert L.hsdc_hss_cmm(h, b"R", 0, 8, 16, 1e-12, *[dp(a)] * 6, dp(x), 8, 0, None) == 0
    assert L.hsdc_hss_cmm(h, b"R", 0, 8, 16, 1e-12, *[dp(a)] * 6, dp(x), 8, -1, None) == -15
    ctx.close()


def test_dhssmm_argument_errors():
    """DHSSMM / DHSSMML argument checks run before any device work."""
    ctx = hsdc.Context()
    L = hsdc.lib()
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    a = np.zeros((10, 10), order="F")
    b = np.zeros((10, 10), order="F")
    h = ctx._h
    for fn in (L.hsdc_dhssmm, L.hsdc_dhssmml):
        assert fn(h, None, 10, 10, 10, dp(b), 10, 10, 1e-11, 0, 4, 2, None, None) == -1
        assert fn(h, dp(a), 9, 10, 10, dp(b), 10, 10, 1e-11, 0, 4, 2, None, None) == -2
        assert fn(h, dp(a), 10, 9, 10, dp(b), 10, 10, 1e-11, 0, 4, 2, None, None) == -3
        assert fn(h, dp(a), 10, -1, -1, dp(b), 10, 10, 1e-11, 0, 4, 2, None, None) == -4
        assert fn(h, dp(a), 10, 10, 10, None, 10, 10, 1e-11, 0, 4, 2, None, None) == -5
        assert fn(h, dp(a), 10, 10, 10, dp(b), 9, 10, 1e-11, 0, 4, 2, None, None) == -6
        assert fn(h, dp(a), 10, 10, 10, dp(b), 10, -1, 1e-11, 0, 4, 2, None, None) == -7
        assert fn(h, dp(a), 10, 10, 10, dp(b), 10, 10, -1.0, 0, 4, 2, None, None) == -8
        assert fn(h, dp(a), 10, 10, 10, dp(b), 10, 10, 1e-11, 0, 0, 2, None, None) == -10
        assert fn(h, dp(a), 10, 10, 10, dp(b), 10, 10, 1e-11, 0, 4, 0, None, None) == -11
    ctx.close()
