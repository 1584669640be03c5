"""DHSSMM and DHSSMML on the device (DHSSMM.f90:1-125, DHSSMML.f90:2-122: a dense matrix
put in HSS form by truncated SVDs, mat2hssvd ConstructHssd.f90:371-696, then multiplied)
against the oracle's restatement and the dense product.

* The reference's own case (testfmm.f90:15-64): Cauchy matrix on interlaced points,
  n = 2000, ni = 64, tol = 1e-11, ncol = 100: the rank of every node's U and V equal to
  the oracle's, the product within 1e-12 (relative) of the oracle's HSS product and
  within tol sqrt(2 n lvl) |b_j| of A b; the same for b A (DHSSMML).
* tol = 0 reproduces A b to rounding; diagonal plus rank one gives ranks <= 1; a tree of
  one leaf falls back to the dense product.
* The randomized construction (mat2hssvdr, hsdc_dhssmm_rand): the same case with a
  seed, every node's ranks within one of the oracle's randomized form (at most 1 node
  in 20 differs) and the product within 1e-11 of it (both sides); tol = 0 reproduces
  A b; a side other than 0 or 1 is -12.
"""
import numpy as np
import pytest

import oracle_py as O
from test_oracle_dhssmm import fmm_matrix, node_ranks
from test_gpu_trace import record

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from hsspack_amd import hsdc
    c = hsdc.Context()
    yield c
    c.close()


def dev_ranks(r, ltr):
    r = np.asarray(r)[:2 * ltr].reshape(-1, 2)
    return r[:-1, 0], r[:-1, 1]


@pytest.mark.parametrize("right", [False, True])
def test_reference_case_vs_oracle(ctx, right):
    n, ni, tol = 2000, 64, 1e-11
    A = fmm_matrix(n, 3)
    B = np.random.default_rng(4).uniform(0, 1, (n, n))
    Y, t, ranks = ctx.dhssmm(A, B, ni, tol, ncol=100, right=right)
    assert ranks is not None
    H = O.mat2hss(A, ni, tol)
    lvl, ltr = H.lvl, H.ltr
    ou, ov = node_ranks(H)
    Hd = O.hss_dense(H)
    O.hss_free(H)
    du, dv = dev_ranks(ranks, ltr)
    ref = B @ Hd if right else Hd @ B
    exact = B @ A if right else A @ B
    err_orc = np.abs(Y - ref).max() / np.abs(ref).max()
    err = np.abs(Y - exact).max()
    bcol = np.linalg.norm(B, axis=1 if right else 0).max()
    bound = tol * np.sqrt(2 * n * lvl) * bcol
    record("dhssmm%s n=%d ni=%d tol %g" % ("l" if right else "", n, ni, tol), rank_max=int(max(du.max(), dv.max())),
           ranks_equal_oracle=bool(np.array_equal(du, ou) and np.array_equal(dv, ov)), rel_err_vs_oracle=err_orc,
           max_err_vs_dense=err, bound=bound, construction_s=t[0], multiply_s=t[1])
    assert np.array_equal(du, ou) and np.array_equal(dv, ov)
    assert err_orc <= 1e-12
    assert err <= bound


def test_tol_zero_and_rank_one(ctx):
    n = 300
    rng = np.random.default_rng(1)
    A = rng.standard_normal((n, n))
    B = rng.standard_normal((n, 40))
    Y, _, ranks = ctx.dhssmm(A, B, 32, 0.0)
    assert ranks is not None
    assert np.abs(Y - A @ B).max() <= 1e-12 * np.abs(A @ B).max()
    u, v = rng.standard_normal(400), rng.standard_normal(400)
    A2 = np.diag(rng.standard_normal(400)) + np.outer(u, v)
    B2 = rng.standard_normal((400, 30))
    Y2, _, r2 = ctx.dhssmm(A2, B2, 40, 1e-10)
    assert r2 is not None and np.asarray(r2).max() <= 1
    assert np.abs(Y2 - A2 @ B2).max() <= 1e-12 * np.abs(A2 @ B2).max()


def test_one_leaf_is_dense(ctx):
    A = np.random.default_rng(2).standard_normal((50, 50))
    B = np.random.default_rng(3).standard_normal((50, 7))
    Y, _, ranks = ctx.dhssmm(A, B, 64, 1e-12)
    assert ranks is None
    assert np.abs(Y - A @ B).max() <= 1e-13 * np.abs(A @ B).max()
    Bl = np.random.default_rng(4).standard_normal((9, 50))
    Yl, _, rl = ctx.dhssmm(A, Bl, 64, 1e-12, right=True)
    assert rl is None
    assert np.abs(Yl - Bl @ A).max() <= 1e-13 * np.abs(Bl @ A).max()


# ---- randomized range finding (mat2hssvdr, ConstructHssd.f90:8-368) ----
@pytest.mark.parametrize("right", [False, True])
def test_rand_reference_case_vs_oracle(ctx, right):
    """hsdc_dhssmm_rand on testfmm.f90's matrix: the device forms the same sketches as
    the oracle (one uniform stream per seed), so the product agrees with the oracle's
    randomized HSS form to rounding.  Ranks: Omega's entries are uniform on [0, 1), and
    their mean (1/2) makes sigma_1 of the sketch 17 to 97 times that of the block on this
    matrix (about 1e4), so eps sigma_1(S) is 0.15 to 0.37 tol: the sketch's singular
    values near tol carry rounding of that size, and two computations that round
    differently can count one of them on either side.  The ranks agree to within one, at
    no more than 1 node in 20."""
    n, ni, tol, seed = 2000, 64, 1e-11, 2024
    A = fmm_matrix(n, 3)
    B = np.random.default_rng(5).uniform(0, 1, (100, n) if right else (n, 100))
    Y, t, ranks = ctx.dhssmm(A, B, ni, tol, ncol=64, right=right, seed=seed)
    assert ranks is not None
    H = O.mat2hss_rand(A, ni, tol, seed)
    lvl, ltr = H.lvl, H.ltr
    ou, ov = node_ranks(H)
    Hd = O.hss_dense(H)
    O.hss_free(H)
    du, dv = dev_ranks(ranks, ltr)
    ref = B @ Hd if right else Hd @ B
    exact = B @ A if right else A @ B
    err_orc = np.abs(Y - ref).max() / np.abs(ref).max()
    err = np.abs(Y - exact).max() / np.abs(exact).max()
    record("dhssmm_rand%s n=%d ni=%d tol %g seed %d" % ("l" if right else "", n, ni, tol, seed),
           rank_max=int(max(du.max(), dv.max())), rel_err_vs_oracle=err_orc, rel_err_vs_dense=err,
           construction_s=t[0], multiply_s=t[1], nodes_rank_differs=int(np.sum(du != ou) + np.sum(dv != ov)))
    d = np.r_[du - ou, dv - ov]
    assert np.abs(d).max() <= 1 and np.count_nonzero(d) * 20 <= len(d)
    assert err_orc <= 1e-11
    assert err <= 1e-7


def test_rand_tol_zero_and_bad_side(ctx):
    n = 300
    rng = np.random.default_rng(6)
    A = rng.standard_normal((n, n))
    B = rng.standard_normal((n, 20))
    Y, _, ranks = ctx.dhssmm(A, B, 32, 0.0, seed=7)
    assert ranks is not None
    assert np.abs(Y - A @ B).max() <= 1e-11 * np.abs(A @ B).max()
    import ctypes
    from hsspack_amd import hsdc
    a = np.asfortranarray(A)
    b = np.asfortranarray(B)
    dp = lambda x: x.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    rc = hsdc.lib().hsdc_dhssmm_rand(ctx._h, dp(a), n, n, n, dp(b), n, 20, 1e-10, 32, 100, 2,
                                     ctypes.c_ulonglong(1), None, None)
    assert rc == -12
