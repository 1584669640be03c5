"""Python host mirror of the hsdc C ABI (include/hsdc.h).

Mirrors HSSPACK's LAPACK-style interface for the tridiagonal DC path:
``stedc`` is MDSTEDC (SRC/mdstedc.f90:139) with the same argument meaning and
INFO convention; ``deflate`` / ``secular`` expose the MDLAED2 deflation and the
DLAED4 secular solve of one merge.  Device memory is owned by torch tensors or by
the library's context.  The library must be built (hsspack_amd/build.py); there is
no fallback: a missing libhsdc.so raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhsdc.so")

_i = ctypes.c_int
_d = ctypes.c_double
_P = ctypes.POINTER
_vp = ctypes.c_void_p

HSDC_ERR_HIP = -1000
HSDC_ERR_NOMEM = -1001
HSDC_ERR_NODEVICE = -1002

# every symbol include/hsdc.h declares
EXPORTS = ["hsdc_default_params", "hsdc_ctx_create", "hsdc_ctx_destroy", "hsdc_set_params",
           "hsdc_get_stats", "hsdc_set_stream", "hsdc_stedc", "hsdc_stedc_dev",
           "hsdc_stedc_batched_dev", "hsdc_autotune", "hsdc_laed0", "hsdc_deflate", "hsdc_secular", "hsdc_laed3", "hsdc_loewner",
           "hsdc_hss_evc", "hsdc_hss_svc", "hsdc_hss_cmm", "hsdc_hss_dense", "hsdc_mlp_set_weights", "hsdc_mlp_get_weights",
           "hsdc_mlp_load", "hsdc_mlp_harvest", "hsdc_mlp_samples", "hsdc_mlp_grad", "hsdc_mlp_train",
           "hsdc_bd_default_params", "hsdc_bd_ctx_create", "hsdc_bd_ctx_destroy", "hsdc_bd_set_params",
           "hsdc_bd_get_stats", "hsdc_bd_set_stream", "hsdc_bdsdc", "hsdc_bdsdc_dev", "hsdc_bd_trace",
           "hsdc_trace", "hsdc_trace_indxq", "hsdc_loewner_noalloc", "hsdc_cauchyhssmm", "hsdc_thread_ctx",
           "hsdc_bd_thread_ctx", "hsdc_thread_ctx_release", "hsdc_sb_default_params", "hsdc_sb_ctx_create",
           "hsdc_sb_ctx_destroy", "hsdc_sb_set_params", "hsdc_sb_get_stats", "hsdc_sb_set_stream",
           "hsdc_sb_thread_ctx", "hsdc_sbevd", "hsdc_sbevd_dev", "hsdc_sb_trace", "hsdc_dhssmm", "hsdc_dhssmml", "hsdc_dhssmm_rand"]

# learned secular start: 34 inputs -> 64 -> 64 -> 1
MLP_IN, MLP_H = 34, 64
MLP_NPARAM = MLP_H * MLP_IN + MLP_H + MLP_H * MLP_H + MLP_H + MLP_H + 1


class Params(ctypes.Structure):
    _fields_ = [("smlsiz", _i), ("hss_min_k", _i), ("hss_leaf", _i), ("hss_tol", _d),
                ("maxit", _i), ("batched", _i), ("learned", _i), ("profile", _i),
                ("learned_min_k", _i), ("learned_gate", _d),
                ("level_sync", _i), ("graph", _i)]


class Stats(ctypes.Structure):
    _fields_ = [("levels", _i), ("merges", _i), ("hss_merges", _i), ("roots", ctypes.c_longlong),
                ("secular_iters", ctypes.c_longlong), ("deflated", ctypes.c_longlong),
                ("k_top", _i), ("rank_max", _i), ("ms_total", _d), ("ms_final", _d),
                ("final_bytes", ctypes.c_longlong), ("learned_roots", ctypes.c_longlong),
                ("secular_wave_iters", ctypes.c_longlong), ("secular_crit_iters", ctypes.c_longlong),
                ("hss_mm_flops", ctypes.c_longlong), ("ms_hss_mm", _d), ("ms_zfill", _d),
                ("zfill_bytes", ctypes.c_longlong), ("secular_evals", ctypes.c_longlong),
                ("ms_secular", _d), ("ms_loewner", _d), ("loewner_pairs", ctypes.c_longlong),
                ("hss_nflops", ctypes.c_longlong), ("ms_hss_build", _d)]


class BdParams(ctypes.Structure):
    _fields_ = [("smlsiz", _i), ("maxit", _i), ("hss_min_k", _i), ("hss_leaf", _i), ("hss_tol", _d)]


class BdStats(ctypes.Structure):
    _fields_ = [("levels", _i), ("merges", _i), ("leaves", _i), ("k_top", _i),
                ("roots", ctypes.c_longlong), ("deflated", ctypes.c_longlong),
                ("rotations", ctypes.c_longlong), ("secular_iters", ctypes.c_longlong),
                ("ms_total", _d), ("hss_merges", _i), ("hss_rank_max", _i)]


class SbParams(ctypes.Structure):
    _fields_ = [("smlsiz", _i), ("maxit", _i), ("hss_min_k", _i), ("hss_leaf", _i), ("hss_tol", _d),
                ("profile", _i)]


class SbStats(ctypes.Structure):
    _fields_ = [("levels", _i), ("merges", _i), ("leaves", _i), ("updates", _i), ("k_top", _i),
                ("roots", ctypes.c_longlong), ("deflated", ctypes.c_longlong),
                ("rotations", ctypes.c_longlong), ("secular_iters", ctypes.c_longlong),
                ("hss_updates", _i), ("hss_rank_max", _i), ("ms_total", _d),
                ("hss_nflops", ctypes.c_longlong), ("hss_mm_flops", ctypes.c_longlong),
                ("dense_flops", ctypes.c_longlong), ("ms_hss_build", _d), ("ms_hss_mm", _d)]


_LIB = None


def lib():
    """Load libhsdc.so; raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("hsspack_amd/libhsdc.so is not built (run hsspack_amd/build.py)")
        # A process must hold a single HIP runtime.  torch ships its own
        # libamdhip64.so.7; loading torch first makes this library bind to that copy
        # (same soname) instead of bringing up a second runtime beside it.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        L.hsdc_ctx_create.argtypes = [_P(_vp)]
        L.hsdc_ctx_destroy.argtypes = [_vp]
        L.hsdc_set_params.argtypes = [_vp, _P(Params)]
        L.hsdc_get_stats.argtypes = [_vp, _P(Stats)]
        L.hsdc_set_stream.argtypes = [_vp, _vp]
        L.hsdc_stedc.argtypes = [_vp, ctypes.c_char, _i, _P(_d), _P(_d), _P(_d), _i, _P(_d), _i,
                                 _P(_i), _i]
        L.hsdc_stedc_dev.argtypes = [_vp, _i, _vp, _vp, _vp, _i]
        L.hsdc_laed0.argtypes = [_vp, _i, _i, _i, _P(_d), _P(_d), _P(_d), _i]
        L.hsdc_autotune.argtypes = [_vp, _i, _P(_i), _P(ctypes.c_long), _vp, _vp, _vp, _P(ctypes.c_long),
                                    _P(_i), _P(_i), _P(_d)]
        L.hsdc_stedc_batched_dev.argtypes = [_vp, _i, _P(_i), _P(ctypes.c_long), _vp, _vp, _vp,
                                             _P(ctypes.c_long)]
        L.hsdc_deflate.argtypes = [_vp, _i, _i, _P(_d), _P(_d), _P(_i), _d, _i, _P(_i), _P(_d),
                                   _P(_d), _P(_d), _P(_i), _P(_i), _P(_i), _P(_i), _P(_i), _P(_i),
                                   _P(_i), _P(_d), _P(_d)]
        L.hsdc_secular.argtypes = [_vp, _i, _P(_d), _P(_d), _d, _P(_d), _P(_i), _P(_d), _P(_i)]
        L.hsdc_laed3.argtypes = [_vp, _i, _i, _i, _P(_d), _P(_d), _i, _d, _P(_d), _P(_d), _P(_i),
                                 _P(_i), _P(_d)]
        L.hsdc_loewner.argtypes = [_vp, _i, _d, _P(_d), _P(_d), _P(_d), _P(_d), _P(_d), _P(_d)]
        L.hsdc_hss_evc.argtypes = [_vp, _i, _i, _d, _P(_d), _P(_d), _P(_d), _P(_d), _P(_d), _P(_d),
                                   _P(_d), _i, _i, _P(_i)]
        L.hsdc_hss_svc.argtypes = [_vp, _i, _i, _d, _P(_d), _P(_d), _P(_d), _P(_d), _P(_d), _P(_d),
                                   _P(_d), _i, _i, _P(_i)]
        L.hsdc_hss_cmm.argtypes = [_vp, ctypes.c_char, _i, _i, _i, _d, _P(_d), _P(_d), _P(_d), _P(_d), _P(_d),
                                   _P(_d), _P(_d), _i, _i, _P(_i)]
        L.hsdc_hss_dense.argtypes = [_vp, _P(_d), _i]
        _f = _P(ctypes.c_float)
        _ll = ctypes.c_longlong
        L.hsdc_mlp_set_weights.argtypes = [_vp, _f, _i]
        L.hsdc_mlp_get_weights.argtypes = [_vp, _f, _i]
        L.hsdc_mlp_load.argtypes = [_vp, ctypes.c_char_p]
        L.hsdc_mlp_harvest.argtypes = [_vp, _ll]
        L.hsdc_mlp_samples.argtypes = [_vp, _f, _f, _ll, _P(_ll)]
        L.hsdc_mlp_grad.argtypes = [_vp, _f, _f, _i, _f, _f]
        L.hsdc_mlp_train.argtypes = [_vp, _f, _f, _ll, _i, _i, ctypes.c_float, ctypes.c_ulonglong, _f]
        L.hsdc_bd_ctx_create.argtypes = [_P(_vp)]
        L.hsdc_bd_ctx_destroy.argtypes = [_vp]
        L.hsdc_bd_set_params.argtypes = [_vp, _P(BdParams)]
        L.hsdc_bd_get_stats.argtypes = [_vp, _P(BdStats)]
        L.hsdc_bd_set_stream.argtypes = [_vp, _vp]
        L.hsdc_bdsdc.argtypes = [_vp, ctypes.c_char, ctypes.c_char, _i, _P(_d), _P(_d), _P(_d), _i,
                                 _P(_d), _i, _P(_d), _P(_i), _P(_d), _P(_i)]
        L.hsdc_bdsdc_dev.argtypes = [_vp, ctypes.c_char, _i, _vp, _vp, _vp, _i, _vp, _i]
        L.hsdc_bd_trace.argtypes = [_vp, _P(_i), _P(_i), _P(_i), _P(_i), _P(_i), _i]
        L.hsdc_trace.argtypes = [_vp, _P(_i), _P(_i), _P(_i), _P(_i), _P(_i), _P(_i), _i]
        L.hsdc_trace_indxq.argtypes = [_vp, _P(_i), _i]
        L.hsdc_loewner_noalloc.argtypes = [_i, _d, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]
        L.hsdc_cauchyhssmm.argtypes = [_vp, _P(_d), _P(_d), _P(_d), _P(_d), _i, _i, _d, _i, _P(_d), _i, _P(_d),
                                       _P(_i)]
        L.hsdc_dhssmm.argtypes = [_vp, _P(_d), _i, _i, _i, _P(_d), _i, _i, _d, _i, _i, _i, _P(_d), _P(_i)]
        L.hsdc_dhssmml.argtypes = [_vp, _P(_d), _i, _i, _i, _P(_d), _i, _i, _d, _i, _i, _i, _P(_d), _P(_i)]
        L.hsdc_dhssmm_rand.argtypes = [_vp, _P(_d), _i, _i, _i, _P(_d), _i, _i, _d, _i, _i, _i, ctypes.c_ulonglong,
                                       _P(_d), _P(_i)]
        L.hsdc_thread_ctx.restype = _vp
        L.hsdc_bd_thread_ctx.restype = _vp
        L.hsdc_sb_ctx_create.argtypes = [_P(_vp)]
        L.hsdc_sb_ctx_destroy.argtypes = [_vp]
        L.hsdc_sb_set_params.argtypes = [_vp, _P(SbParams)]
        L.hsdc_sb_get_stats.argtypes = [_vp, _P(SbStats)]
        L.hsdc_sb_set_stream.argtypes = [_vp, _vp]
        L.hsdc_sb_thread_ctx.restype = _vp
        L.hsdc_sbevd.argtypes = [_vp, ctypes.c_char, ctypes.c_char, _i, _i, _P(_d), _i, _P(_d), _P(_d), _i,
                                 _P(_d), _i, _P(_i), _i]
        L.hsdc_sbevd_dev.argtypes = [_vp, ctypes.c_char, _i, _i, _vp, _i, _vp, _vp, _i]
        L.hsdc_sb_trace.argtypes = [_vp, _P(_i), _P(_i), _P(_i), _P(_i), _P(_i), _i]
        _LIB = L
    return _LIB


def _dp(a):
    return a.ctypes.data_as(_P(_d))


def _fp(a):
    return a.ctypes.data_as(_P(ctypes.c_float))


def _fnv1a(b):
    h = 2166136261
    for x in b:
        h = ((h ^ x) * 16777619) & 0xFFFFFFFF
    return h


def save_mlp_weights(path, w):
    """Writes the weights file read by hsdc_mlp_load: b"HSDCMLP1", int32 inputs,
    hidden, nparam, uint32 FNV-1a of the weight bytes, float32 weights (little-endian)."""
    w = np.ascontiguousarray(w, "<f4")
    if w.size != MLP_NPARAM:
        raise ValueError("expected %d weights" % MLP_NPARAM)
    b = w.tobytes()
    hdr = np.array([MLP_IN, MLP_H, MLP_NPARAM], "<i4").tobytes()
    with open(path, "wb") as f:
        f.write(b"HSDCMLP1" + hdr + np.array([_fnv1a(b)], "<u4").tobytes() + b)


def load_mlp_weights(path):
    """Reads and verifies a weights file; returns float32 weights."""
    raw = open(path, "rb").read()
    if raw[:8] != b"HSDCMLP1":
        raise ValueError("not a weights file: %s" % path)
    hdr = np.frombuffer(raw[8:20], "<i4")
    if tuple(hdr) != (MLP_IN, MLP_H, MLP_NPARAM):
        raise ValueError("shape mismatch %s" % (tuple(hdr),))
    sum_ = int(np.frombuffer(raw[20:24], "<u4")[0])
    b = raw[24:24 + 4 * MLP_NPARAM]
    if len(b) != 4 * MLP_NPARAM or _fnv1a(b) != sum_:
        raise ValueError("checksum mismatch: %s" % path)
    return np.frombuffer(b, "<f4").copy()


def _ip(a):
    return a.ctypes.data_as(_P(_i))


def default_params():
    p = Params()
    lib().hsdc_default_params(ctypes.byref(p))
    return p


NOALLOC_KMAX = 65536


def loewner_noalloc(k, rho, dlamda, z, d, znew, difl, difr, alpha, info, stream=None):
    """hsdc_loewner_noalloc on torch device tensors (float64 of k, info int32 of 1):
    allocates nothing, queues on `stream` (torch stream, raw handle or None) and
    returns the status without waiting."""
    h = getattr(stream, "cuda_stream", stream)
    return lib().hsdc_loewner_noalloc(int(k), float(rho), dlamda.data_ptr(), z.data_ptr(), d.data_ptr(),
                                      znew.data_ptr(), difl.data_ptr(), difr.data_ptr(), alpha.data_ptr(),
                                      info.data_ptr(), ctypes.c_void_p(h or 0))


class Context:
    """One library context (device buffers + parameters); not thread safe."""

    def __init__(self, **params):
        self._h = _vp()
        rc = lib().hsdc_ctx_create(ctypes.byref(self._h))
        if rc != 0:
            raise RuntimeError("hsdc_ctx_create failed: %d" % rc)
        self.params = default_params()
        self.set_params(**params)

    def set_params(self, **kw):
        for k, v in kw.items():
            setattr(self.params, k, v)
        rc = lib().hsdc_set_params(self._h, ctypes.byref(self.params))
        if rc != 0:
            raise ValueError("hsdc_set_params rejected %r (%d)" % (kw, rc))

    def set_stream(self, stream):
        """HIP stream of the context's launches: a torch stream, a raw handle, or None
        for the null stream."""
        h = getattr(stream, "cuda_stream", stream)
        return lib().hsdc_set_stream(self._h, ctypes.c_void_p(h or 0))

    def stats(self):
        s = Stats()
        lib().hsdc_get_stats(self._h, ctypes.byref(s))
        return {f: getattr(s, f) for f, _ in Stats._fields_}

    def close(self):
        if self._h:
            lib().hsdc_ctx_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def trace(self):
        """Merges of the last solve (level by level, MDLAED0's order): dict of arrays s, n,
        k, nrot, hss (1: HSS branch), hss_rank (largest generator rank, 0 dense)."""
        cnt = lib().hsdc_trace(self._h, None, None, None, None, None, None, 0)
        if cnt < 0:
            raise RuntimeError("hsdc_trace: %d" % cnt)
        a = [np.zeros(max(cnt, 1), np.int32) for _ in range(6)]
        lib().hsdc_trace(self._h, *[_ip(x) for x in a], cnt)
        return dict(zip(("s", "n", "k", "nrot", "hss", "hss_rank"), (x[:cnt] for x in a)))

    def trace_indxq(self, n):
        """INDXQ (1-based) after the top merge of the last single-problem solve of order n."""
        q = np.zeros(n, np.int32)
        rc = lib().hsdc_trace_indxq(self._h, _ip(q), n)
        if rc != 0:
            raise RuntimeError("hsdc_trace_indxq: %d" % rc)
        return q

    # ---- MDSTEDC (host arrays) ----
    def stedc(self, d, e, compz="I", z=None, lwork=None, liwork=None, ldz=None):
        """Returns (d, z, info) like MDSTEDC; d and e are copied.  ldz (>= n): leading
        dimension of the Z array the library writes (z is returned as its n x n part)."""
        n = len(d)
        d = np.array(d, np.float64)
        e = np.array(e, np.float64) if n > 1 else np.zeros(1)
        ldz = max(1, n) if ldz is None else ldz
        zb = np.zeros((ldz, n), order="F")
        if z is not None:
            zb[:n, :] = np.asarray(z, np.float64)
        if lwork is None:
            lwork = 1 + 4 * n + n * n if compz != "V" else 1 + 3 * n + 2 * n * 17 + 4 * n * n
        if liwork is None:
            liwork = 3 + 5 * n if compz != "V" else 6 + 6 * n + 5 * n * 17
        work = np.zeros(max(1, min(lwork, 4)) if lwork > 0 else 1)
        iwork = np.zeros(1, np.int32)
        info = lib().hsdc_stedc(self._h, ctypes.c_char(compz.encode()), n, _dp(d), _dp(e),
                                _dp(zb) if n else None, ldz, _dp(work), lwork, _ip(iwork), liwork)
        return d, np.asfortranarray(zb[:n, :]), info

    # ---- device-resident MDSTEDC on torch tensors ----
    def stedc_dev(self, d, e, z):
        """d (n), e (n-1 or n), z (n x n, column-major i.e. z.T contiguous) on cuda."""
        n = d.shape[0]
        rc = lib().hsdc_stedc_dev(self._h, n, d.data_ptr(), e.data_ptr(), z.data_ptr(), n)
        return rc

    def stedc_batched_dev(self, ns, offs, d, e, z, zoffs):
        ns = np.ascontiguousarray(ns, np.int32)
        offs = np.ascontiguousarray(offs, np.int64)
        zoffs = np.ascontiguousarray(zoffs, np.int64)
        return lib().hsdc_stedc_batched_dev(self._h, len(ns), _ip(ns),
                                            offs.ctypes.data_as(_P(ctypes.c_long)), d.data_ptr(),
                                            e.data_ptr(), z.data_ptr(),
                                            zoffs.ctypes.data_as(_P(ctypes.c_long)))

    def autotune(self, ns, offs, d, e, z, zoffs):
        """Picks the HSS leaf size and DC leaf bound by timing batched solves; returns
        (hss_leaf, smlsiz, times of the 6 candidates)."""
        ns = np.ascontiguousarray(ns, np.int32)
        offs = np.ascontiguousarray(offs, np.int64)
        zoffs = np.ascontiguousarray(zoffs, np.int64)
        bl, bs = _i(0), _i(0)
        times = np.zeros(6)
        rc = lib().hsdc_autotune(self._h, len(ns), _ip(ns), offs.ctypes.data_as(_P(ctypes.c_long)),
                                 d.data_ptr(), e.data_ptr(), z.data_ptr(),
                                 zoffs.ctypes.data_as(_P(ctypes.c_long)), ctypes.byref(bl), ctypes.byref(bs),
                                 _dp(times))
        if rc != 0:
            raise RuntimeError("hsdc_autotune: %d" % rc)
        self.params.hss_leaf, self.params.smlsiz = bl.value, bs.value
        return bl.value, bs.value, times

    # ---- one merge ----
    def deflate(self, d, z, indxq, rho, n1, hss_min_k=1 << 30):
        n = len(d)
        d = np.array(d, np.float64)
        z = np.array(z, np.float64)
        indxq = np.array(indxq, np.int32)
        out = dict(dlamda=np.zeros(n), w=np.zeros(n), indxp=np.zeros(n, np.int32),
                   indx=np.zeros(n, np.int32), indxc=np.zeros(n, np.int32),
                   coltyp=np.zeros(max(n, 4), np.int32), rot_pj=np.zeros(n, np.int32),
                   rot_nj=np.zeros(n, np.int32), rot_c=np.zeros(n), rot_s=np.zeros(n))
        k, nrot = _i(0), _i(0)
        rho_o = _d(0)
        info = lib().hsdc_deflate(self._h, n, n1, _dp(d), _dp(z), _ip(indxq), rho, hss_min_k,
                                  ctypes.byref(k), ctypes.byref(rho_o), _dp(out["dlamda"]),
                                  _dp(out["w"]), _ip(out["indxp"]), _ip(out["indx"]),
                                  _ip(out["indxc"]), _ip(out["coltyp"]), ctypes.byref(nrot),
                                  _ip(out["rot_pj"]), _ip(out["rot_nj"]), _dp(out["rot_c"]),
                                  _dp(out["rot_s"]))
        r = nrot.value
        out.update(info=info, k=k.value, rho=rho_o.value, d=d, indxq=indxq,
                   rot=(out["rot_pj"][:r], out["rot_nj"][:r], out["rot_c"][:r], out["rot_s"][:r]))
        return out

    def secular(self, dlamda, w, rho):
        k = len(dlamda)
        dl = np.ascontiguousarray(dlamda, np.float64)
        ww = np.ascontiguousarray(w, np.float64)
        lam, tau = np.zeros(k), np.zeros(k)
        org, nit = np.zeros(k, np.int32), np.zeros(k, np.int32)
        info = lib().hsdc_secular(self._h, k, _dp(dl), _dp(ww), rho, _dp(lam), _ip(org), _dp(tau),
                                  _ip(nit))
        return dict(lam=lam, org=org, tau=tau, niter=nit, info=info)

    def laed3(self, k, n, n1, rho, dlamda, q2, indx, ctot, w):
        """LAPACK DLAED3 semantics on the device: returns (d, q, w_updated, info)."""
        d = np.zeros(n)
        q = np.zeros((n, max(k, 1)), order="F")
        dl = np.array(dlamda, np.float64)
        q2 = np.ascontiguousarray(q2, np.float64)
        ix = np.ascontiguousarray(indx, np.int32)
        ct = np.ascontiguousarray(ctot[:4], np.int32)
        ww = np.array(w, np.float64)
        info = lib().hsdc_laed3(self._h, k, n, n1, _dp(d), _dp(q), n, rho, _dp(dl), _dp(q2), _ip(ix),
                                _ip(ct), _dp(ww))
        return d, q, ww, info

    def loewner(self, dlamda, z, rho):
        k = len(dlamda)
        dl = np.array(dlamda, np.float64)
        zz = np.array(z, np.float64)
        d, difl, difr, alpha = np.zeros(k), np.zeros(k), np.zeros(k), np.zeros(k)
        info = lib().hsdc_loewner(self._h, k, rho, _dp(dl), _dp(zz), _dp(d), _dp(difl), _dp(difr),
                                  _dp(alpha))
        return dict(d=d, z=zz, difl=difl, difr=difr, alpha=alpha, info=info)

    def hss_evc(self, f, d, u, v, difl, difr, q, ni, tol):
        """q (nrow x k) := q * H; returns (q, ranks per node [U cols, V rows])."""
        return self._hss(lib().hsdc_hss_evc, "hsdc_hss_evc", f, d, u, v, difl, difr, q, ni, tol)

    def hss_svc(self, f, d, u, v, difl, difr, q, ni, tol):
        """As hss_evc for H(i,j) = u_i v_j / (f_i^2 - d_j^2) (the SVD's vector matrices)."""
        return self._hss(lib().hsdc_hss_svc, "hsdc_hss_svc", f, d, u, v, difl, difr, q, ni, tol)

    def hss_cmm(self, side, squared, f, d, u, v, difl, difr, x, ni, tol):
        """CauchyHssMM: side 'R' x (ncol x k) := x * H, side 'L' x (k x ncol) := H * x for the
        HSS form of the Cauchy-like matrix (squared 1: the SVD's u_i v_j / (f_i^2 - d_j^2))."""
        k = len(f)
        side = side.upper()
        arrs = [np.ascontiguousarray(a, np.float64) for a in (f, d, u, v, difl, difr)]
        x = np.array(x, np.float64, order="F")
        if x.ndim != 2 or side not in ("R", "L"):
            raise ValueError("x must be 2-D and side 'R' or 'L'")
        if (side == "R" and x.shape[1] != k) or (side == "L" and x.shape[0] != k):
            raise ValueError("x has shape %s; side %s needs %s k = %d" % (x.shape, side,
                             "columns" if side == "R" else "rows", k))
        ncol = x.shape[0] if side == "R" else x.shape[1]
        ranks = np.zeros(2 * (2 * (k // max(ni, 1)) + 4), np.int32)
        info = lib().hsdc_hss_cmm(self._h, ctypes.c_char(side.encode()), int(squared), k, ni, tol,
                                  *[_dp(a) for a in arrs], _dp(x), max(x.shape[0], 1), ncol, _ip(ranks))
        if info != 0:
            raise RuntimeError("hsdc_hss_cmm: %d" % info)
        return x, ranks

    def _hss(self, fn, name, f, d, u, v, difl, difr, q, ni, tol):
        k = len(f)
        arrs = [np.ascontiguousarray(a, np.float64) for a in (f, d, u, v, difl, difr)]
        q = np.array(q, np.float64, order="F")
        if q.ndim != 2 or q.shape[1] != k:
            raise ValueError("q has shape %s; it needs k = %d columns" % (q.shape, k))
        nrow = q.shape[0]
        ranks = np.zeros(2 * (2 * (k // max(ni, 1)) + 4), np.int32)
        info = fn(self._h, k, ni, tol, *[_dp(a) for a in arrs], _dp(q), max(nrow, 1), nrow, _ip(ranks))
        if info != 0:
            raise RuntimeError("%s: %d" % (name, info))
        return q, ranks

    def cauchyhssmm(self, d, f, u, v, b, ni, tol, ncol=128):
        """CauchyHssMM: b (n x n) := A b for A(i,j) = u_i v_j / (d_i - f_j); returns (b,
        [construction s, multiply s], ranks per node (U, V) or None when applied densely)."""
        n = len(d)
        arrs = [np.ascontiguousarray(a, np.float64) for a in (d, f, u, v)]
        if any(a.shape != (n,) for a in arrs):
            raise ValueError("d, f, u, v must have the same length")
        b = np.array(b, np.float64, order="F")
        if b.shape != (n, n):
            raise ValueError("b must be %d x %d" % (n, n))
        t = np.zeros(2)
        ranks = np.zeros(2 * max(2 * (n // max(ni, 1)) + 4, 1), np.int32)
        info = lib().hsdc_cauchyhssmm(self._h, *[_dp(a) for a in arrs], n, ni, tol, ncol, _dp(b), max(n, 1),
                                      _dp(t), _ip(ranks))
        if info != 0:
            raise RuntimeError("hsdc_cauchyhssmm: %d" % info)
        return b, t, (None if ranks[0] < 0 else ranks)

    def dhssmm(self, a, b, ni, tol, ncol=100, prank=0, right=False, seed=None):
        """DHSSMM (A b, b n x k) or, right=True, DHSSMML (b A, b k x n) with A put in HSS
        form by truncated SVDs, or with seed given by randomized range finding
        (mat2hssvdr, hsdc_dhssmm_rand).  Returns (product, (construction s, multiply s),
        ranks (2 per tree node) or None when A was applied densely)."""
        a = np.asfortranarray(np.asarray(a, np.float64))
        b = np.array(b, np.float64, order="F")
        n = a.shape[0]
        k = b.shape[0] if right else b.shape[1]
        t = np.zeros(2)
        ranks = np.zeros(2 * max(2 * (n // max(ni, 1)) + 2, 2), np.int32)
        if seed is not None:
            name = "hsdc_dhssmm_rand"
            rc = lib().hsdc_dhssmm_rand(self._h, _dp(a), max(1, n), a.shape[1], n, _dp(b), max(1, b.shape[0]), k,
                                        tol, ni, ncol, 1 if right else 0, ctypes.c_ulonglong(seed), _dp(t),
                                        _ip(ranks))
        else:
            name = "hsdc_dhssmml" if right else "hsdc_dhssmm"
            fn = lib().hsdc_dhssmml if right else lib().hsdc_dhssmm
            rc = fn(self._h, _dp(a), max(1, n), a.shape[1], n, _dp(b), max(1, b.shape[0]), k, tol, prank, ni, ncol,
                    _dp(t), _ip(ranks))
        if rc != 0:
            raise RuntimeError("%s: %d" % (name, rc))
        return b, t, (None if ranks[0] < 0 else ranks)

    def hss_dense(self, k):
        a = np.zeros((k, k), order="F")
        info = lib().hsdc_hss_dense(self._h, _dp(a), k)
        if info != 0:
            raise RuntimeError("hsdc_hss_dense: %d" % info)
        return a

    # ---- learned secular start ----
    def mlp_set_weights(self, w):
        if w is None:
            return lib().hsdc_mlp_set_weights(self._h, None, 0)
        w = np.ascontiguousarray(w, np.float32)
        rc = lib().hsdc_mlp_set_weights(self._h, _fp(w), w.size)
        if rc != 0:
            raise ValueError("hsdc_mlp_set_weights: %d" % rc)
        return rc

    def mlp_get_weights(self):
        w = np.zeros(MLP_NPARAM, np.float32)
        rc = lib().hsdc_mlp_get_weights(self._h, _fp(w), MLP_NPARAM)
        return w if rc == 0 else None

    def mlp_load(self, path):
        rc = lib().hsdc_mlp_load(self._h, os.fsencode(path))
        if rc != 0:
            raise ValueError("hsdc_mlp_load(%s): %d" % (path, rc))

    def mlp_harvest(self, cap):
        rc = lib().hsdc_mlp_harvest(self._h, int(cap))
        if rc != 0:
            raise RuntimeError("hsdc_mlp_harvest: %d" % rc)

    def mlp_samples(self, maxn):
        x = np.zeros((int(maxn), MLP_IN), np.float32)
        t = np.zeros(int(maxn), np.float32)
        n = ctypes.c_longlong(0)
        rc = lib().hsdc_mlp_samples(self._h, _fp(x), _fp(t), int(maxn), ctypes.byref(n))
        if rc != 0:
            raise RuntimeError("hsdc_mlp_samples: %d" % rc)
        return x[:n.value], t[:n.value]

    def mlp_count(self):
        n = ctypes.c_longlong(0)
        rc = lib().hsdc_mlp_samples(self._h, None, None, 0, ctypes.byref(n))
        if rc != 0:
            raise RuntimeError("hsdc_mlp_samples: %d" % rc)
        return n.value

    def mlp_grad(self, x, t):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        g = np.zeros(MLP_NPARAM, np.float32)
        loss = np.zeros(1, np.float32)
        rc = lib().hsdc_mlp_grad(self._h, _fp(x), _fp(t), len(t), _fp(g), _fp(loss))
        if rc != 0:
            raise RuntimeError("hsdc_mlp_grad: %d" % rc)
        return g, float(loss[0])

    def mlp_train(self, x, t, epochs, batch, lr, seed=0):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        hist = np.zeros(epochs, np.float32)
        rc = lib().hsdc_mlp_train(self._h, _fp(x), _fp(t), len(t), epochs, batch, lr, seed, _fp(hist))
        if rc != 0:
            raise RuntimeError("hsdc_mlp_train: %d" % rc)
        return hist


class BdContext:
    """Context of the bidiagonal DC SVD (MDBDSDC, SRC/mdbdsdc.f90); not thread safe.

    Parameters (hsdc_bd_params): smlsiz (leaf bound, 25), maxit (secular iteration cap),
    hss_min_k (a merge with K >= hss_min_k takes the HSS branch, MDLASD8 + DHSSVCS;
    2000 as the reference's BSMLZ), hss_leaf (HSS leaf block, 40 as the reference's Ni),
    hss_tol (HSS compression tolerance, 1e-17)."""

    def __init__(self, **params):
        self._h = _vp()
        rc = lib().hsdc_bd_ctx_create(ctypes.byref(self._h))
        if rc != 0:
            raise RuntimeError("hsdc_bd_ctx_create failed: %d" % rc)
        self.params = BdParams()
        lib().hsdc_bd_default_params(ctypes.byref(self.params))
        self.set_params(**params)

    def set_params(self, **kw):
        for k, v in kw.items():
            setattr(self.params, k, v)
        rc = lib().hsdc_bd_set_params(self._h, ctypes.byref(self.params))
        if rc != 0:
            raise ValueError("hsdc_bd_set_params rejected %r (%d)" % (kw, rc))

    def set_stream(self, stream):
        h = getattr(stream, "cuda_stream", stream)
        return lib().hsdc_bd_set_stream(self._h, ctypes.c_void_p(h or 0))

    def stats(self):
        s = BdStats()
        lib().hsdc_bd_get_stats(self._h, ctypes.byref(s))
        return {f: getattr(s, f) for f, _ in BdStats._fields_}

    def trace(self):
        cnt = lib().hsdc_bd_trace(self._h, None, None, None, None, None, 0)
        a = [np.zeros(max(cnt, 1), np.int32) for _ in range(5)]
        lib().hsdc_bd_trace(self._h, *[_ip(x) for x in a], cnt)
        return dict(s=a[0][:cnt], n=a[1][:cnt], k=a[2][:cnt], nrot=a[3][:cnt], hss_rank=a[4][:cnt])

    def close(self):
        if self._h:
            lib().hsdc_bd_ctx_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def bdsdc(self, d, e, uplo="U", compq="I"):
        """MDBDSDC on host arrays: returns (s descending, U, VT, info); d, e are copied."""
        n = len(d)
        d = np.array(d, np.float64)
        e = np.array(e, np.float64) if n > 1 else np.zeros(1)
        u = np.zeros((n, n), order="F")
        vt = np.zeros((n, n), order="F")
        ld = max(1, n)
        info = lib().hsdc_bdsdc(self._h, ctypes.c_char(uplo.encode()), ctypes.c_char(compq.encode()), n,
                                _dp(d), _dp(e), _dp(u) if n else None, ld, _dp(vt) if n else None, ld,
                                None, None, None, None)
        return d, u, vt, info

    def bdsdc_dev(self, d, e, u, vt, uplo="U"):
        """Device-resident MDBDSDC on torch tensors: d (n, overwritten by the singular
        values), e (n-1), u and vt (n x n column-major, i.e. .T contiguous) or None."""
        n = d.shape[0]
        return lib().hsdc_bdsdc_dev(self._h, ctypes.c_char(uplo.encode()), n, d.data_ptr(), e.data_ptr(),
                                    u.data_ptr() if u is not None else None, n,
                                    vt.data_ptr() if vt is not None else None, n)


class SbContext:
    """Context of the banded DC eigensolver (MDSBEVD, SRC/mdsbevd.f90); not thread safe.

    Parameters (hsdc_sb_params): smlsiz (leaf bound, 2 KD - 1 <= smlsiz <= 128, default
    128), maxit (secular iteration cap), hss_min_k (an update with K >= hss_min_k takes the
    HSS branch, MDSBED5 + DHSSEVC; 1500 as the reference's BSMLZ; 0 = never), hss_leaf
    (HSS leaf block, 40 as the reference's Ni), hss_tol (HSS compression tolerance),
    profile (1: event-timed HSS construction and multiply in the stats)."""

    def __init__(self, **params):
        self._h = _vp()
        rc = lib().hsdc_sb_ctx_create(ctypes.byref(self._h))
        if rc != 0:
            raise RuntimeError("hsdc_sb_ctx_create failed: %d" % rc)
        self.params = SbParams()
        lib().hsdc_sb_default_params(ctypes.byref(self.params))
        self.set_params(**params)

    def set_params(self, **kw):
        for k, v in kw.items():
            setattr(self.params, k, v)
        rc = lib().hsdc_sb_set_params(self._h, ctypes.byref(self.params))
        if rc != 0:
            raise ValueError("hsdc_sb_set_params rejected %r (%d)" % (kw, rc))

    def set_stream(self, stream):
        h = getattr(stream, "cuda_stream", stream)
        return lib().hsdc_sb_set_stream(self._h, ctypes.c_void_p(h or 0))

    def stats(self):
        s = SbStats()
        lib().hsdc_sb_get_stats(self._h, ctypes.byref(s))
        return {f: getattr(s, f) for f, _ in SbStats._fields_}

    def trace(self):
        """Per rank-one update (level by level, update by update, merge by merge): block
        start, order N, K, rotations, largest HSS rank (0: dense)."""
        cnt = lib().hsdc_sb_trace(self._h, None, None, None, None, None, 0)
        a = [np.zeros(max(cnt, 1), np.int32) for _ in range(5)]
        lib().hsdc_sb_trace(self._h, *[_ip(x) for x in a], cnt)
        return dict(s=a[0][:cnt], n=a[1][:cnt], k=a[2][:cnt], nrot=a[3][:cnt], hss_rank=a[4][:cnt])

    def close(self):
        if self._h:
            lib().hsdc_sb_ctx_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sbevd(self, ab, kd, uplo="U", compz="I"):
        """MDSBEVD on host arrays: ab (kd+1 x n band storage, any layout; read
        column-major).  Returns (w ascending, Q (n x n) or None, info)."""
        ab = np.asfortranarray(np.asarray(ab, np.float64))
        ldab, n = ab.shape
        w = np.zeros(max(n, 1))
        vec = compz.upper() == "I"
        q = np.zeros((n, n), order="F") if vec else None
        info = lib().hsdc_sbevd(self._h, ctypes.c_char(compz.encode()), ctypes.c_char(uplo.encode()), n, kd,
                                _dp(ab) if n else None, ldab, _dp(w), _dp(q) if (vec and n) else None, max(1, n),
                                None, 0, None, 0)
        return w[:n], q, info

    def sbevd_dev(self, ab, kd, w, q, uplo="U"):
        """Device-resident MDSBEVD on torch tensors: ab (n, ldab) -- band storage
        column-major, i.e. row j of the tensor is column j of AB --, w (n) receives the
        eigenvalues, q (n x n column-major, i.e. .T contiguous) the vectors or None."""
        n, ldab = ab.shape
        return lib().hsdc_sbevd_dev(self._h, ctypes.c_char(uplo.encode()), n, kd, ab.data_ptr(), ldab,
                                    w.data_ptr(), q.data_ptr() if q is not None else None, max(1, n))
