"""Seeded synthetic tridiagonal workloads.

One generator is used everywhere (tests, fixtures, the benchmark) so that every
consumer builds bit-identical matrices: splitmix64 seeded with
``seed * 0x100000001 + sub``; u = (x >> 11) * 2**-53; U(a, b) = a + (b - a) * u.
All diagonal entries are drawn before all off-diagonal entries.

Workloads (sizes are the published benchmark configurations of the HSSPACK
tridiagonal divide-and-conquer path, TEST/Testedc.f90:18-37 for the near-Toeplitz
family):
  cfg 1  d = 2 + U(-1e-3, 1e-3),  e = 1 + U(-1e-3, 1e-3)      (low deflation)
  cfg 2  d, e ~ U(-1, 1)                                       (heavy deflation)
Upper bidiagonal (the SVD path):
  cfg 5  d ~ U(0.5, 1.5),  e ~ U(0, 1)
"""
import numpy as np

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix_uniform(seed, count, sub=0, start=0):
    """count uniforms in [0,1) from stream (seed, sub), skipping `start` draws."""
    with np.errstate(over="ignore"):
        state = np.uint64((int(seed) * 0x100000001 + int(sub)) & 0xFFFFFFFFFFFFFFFF)
        k = np.arange(start + 1, start + count + 1, dtype=np.uint64)
        x = state + k * _GAMMA
        x = (x ^ (x >> np.uint64(30))) * _M1
        x = (x ^ (x >> np.uint64(27))) * _M2
        x = x ^ (x >> np.uint64(31))
    return (x >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def tridiag_cfg(cfg, n, seed, sub=0):
    """(d, e) of order n for workload family cfg (1 or 2)."""
    u = splitmix_uniform(seed, 2 * n - 1, sub)
    ud, ue = u[:n], u[n:]
    if cfg == 1:
        d = 2.0 + (-1e-3 + 2e-3 * ud)
        e = 1.0 + (-1e-3 + 2e-3 * ue)
    elif cfg == 2:
        d = -1.0 + 2.0 * ud
        e = -1.0 + 2.0 * ue
    else:
        raise ValueError("unknown workload family %r" % cfg)
    return d, e


def bidiag_cfg(n, seed=5, sub=0):
    """(d, e) of the order-n upper bidiagonal of workload family 5."""
    u = splitmix_uniform(seed, 2 * n - 1, sub)
    return 0.5 + u[:n], u[n:].copy()


def bidiag_toeplitz(n, seed=11):
    """(d, e) of an order-n near-Toeplitz upper bidiagonal: 1 + 1e-3 N(0,1) entries.
    Its singular vectors are delocalised, so the DC merges barely deflate (K ~ N) and
    the large merges take the HSS branch."""
    rng = np.random.default_rng(seed)
    return 1.0 + 1e-3 * rng.standard_normal(n), 1.0 + 1e-3 * rng.standard_normal(n - 1)


def batch_sizes(count=512, lo=256, hi=8192, seed=3):
    """Problem orders drawn log-uniform in [lo, hi] from stream (seed, 0)."""
    u = splitmix_uniform(seed, count, 0)
    return np.rint(np.exp(np.log(lo) + (np.log(hi) - np.log(lo)) * u)).astype(np.int64)


def batch_problems(count=512, lo=256, hi=8192, seed=3, cfg=1):
    sizes = batch_sizes(count, lo, hi, seed)
    return [tridiag_cfg(cfg, int(n), seed, sub=k + 1) for k, n in enumerate(sizes)]


def band_cfg(n, kd, seed=6, sub=0):
    """(kd + 1) x n upper band storage of the symmetric band matrix of workload family 6:
    entries N(0,1) by Box-Muller on stream (seed, sub), AB filled column-major, the
    entries above the matrix (AB(i, j) with i + j <= kd, 0-based) left zero."""
    cnt = (kd + 1) * n
    u = splitmix_uniform(seed, 2 * ((cnt + 1) // 2), sub)
    u1, u2 = u[0::2], u[1::2]
    r = np.sqrt(-2.0 * np.log1p(-u1))  # 1 - u1 lies in (0, 1]
    g = np.empty(2 * len(u1))
    g[0::2] = r * np.cos(2.0 * np.pi * u2)
    g[1::2] = r * np.sin(2.0 * np.pi * u2)
    ab = g[:cnt].reshape(n, kd + 1).T.copy(order="F")
    for j in range(min(kd, n)):
        ab[:kd - j, j] = 0.0
    return ab


def band_dense(ab):
    """The dense symmetric matrix of an upper band storage."""
    kd = ab.shape[0] - 1
    n = ab.shape[1]
    a = np.zeros((n, n))
    for j in range(n):
        for i in range(max(0, j - kd), j + 1):
            a[i, j] = a[j, i] = ab[kd + i - j, j]
    return a
