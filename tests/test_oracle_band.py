"""The oracle's banded divide and conquer (MDSBEVD, mdsbevd.f90:183-498, with the
MDSBED1 merges) and its two dense kernels pinned against LAPACK (scipy): the Jacobi
eigensolver of the leaves against dsyevd, the Jacobi SVD of the coupling blocks against
dgesvd, the whole driver against dsbevd (eigenvalues <= 4 n eps |A|, orthogonality and
residual <= 1e-12).  The algebra of the split is checked directly: A equals the block
diagonal of the corrected leaves plus s_k [u_k; v_k][u_k; v_k]^T per cut."""
import numpy as np
import pytest
import scipy.linalg as sl

import oracle_py as O
from hsspack_amd.workloads import band_cfg, band_dense

EPS = 2.0 ** -53


@pytest.mark.parametrize("m", [1, 2, 7, 64, 127])
def test_sym_jacobi_vs_dsyevd(m):
    rng = np.random.default_rng(m)
    a = rng.standard_normal((m, m))
    a = a + a.T
    w, v, sw = O.sym_jacobi(a)
    assert sw >= 0
    ref = sl.eigh(a, eigvals_only=True, driver="evd")
    nrm = np.abs(a).max()
    nrm2 = np.abs(ref).max()  # |A|_2, the norm of the eigenvalue perturbation bound
    assert np.abs(w - ref).max() <= 4 * m * EPS * nrm2
    assert np.abs(v.T @ v - np.eye(m)).max() <= 1e-13
    assert np.abs(a @ v - v * w).max() <= 1e-13 * nrm * m
    assert np.all(np.diff(w) >= 0)


@pytest.mark.parametrize("m", [1, 3, 16, 64])
def test_jacobi_svd_vs_dgesvd(m):
    rng = np.random.default_rng(10 + m)
    c = np.tril(rng.standard_normal((m, m)))
    s, u, v, sw = O.jacobi_svd(c)
    assert sw >= 0
    ref = sl.svd(c, compute_uv=False, lapack_driver="gesvd")
    assert np.abs(s - ref).max() <= 8 * m * EPS * ref[0]
    assert np.abs(u @ np.diag(s) @ v.T - c).max() <= 1e-13 * ref[0] * m
    assert np.abs(v.T @ v - np.eye(m)).max() <= 1e-13
    assert np.all(np.diff(s) <= 0)


@pytest.mark.parametrize("n,kd", [(40, 0), (60, 1), (200, 3), (300, 8), (257, 16), (512, 32), (700, 64)])
def test_sbevd_vs_dsbevd(n, kd):
    ab = band_cfg(n, kd, seed=6, sub=n)
    A = band_dense(ab)
    r = O.sbevd(ab, kd, trace_cap=1 << 16)
    assert r["info"] == 0
    w = sl.eig_banded(ab, lower=False, eigvals_only=True)
    nrm = np.abs(A).max()
    nrm2 = np.abs(w).max()
    lam, q = r["d"], r["q"]
    assert np.abs(lam - w).max() <= 4 * n * EPS * nrm2
    assert np.abs(q.T @ q - np.eye(n)).max() <= 1e-12
    assert np.abs(A @ q - q * lam).max() / nrm <= 1e-12
    tr = r["trace"]
    if kd > 0 and n > max(2 * kd - 1, 25):
        assert tr["nmerge"] > 0 and len(tr["k"]) == tr["nmerge"] * kd
        assert np.all(tr["k"] <= tr["n"])


def test_band_storage_round_trip():
    ab = band_cfg(30, 4, seed=1)
    A = band_dense(ab)
    assert np.allclose(A, A.T)
    assert np.all(np.triu(A, 5) == 0)
    w = sl.eig_banded(ab, lower=False, eigvals_only=True)
    assert np.allclose(np.sort(np.linalg.eigvalsh(A)), w, atol=1e-12)
