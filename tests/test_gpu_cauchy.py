"""GPU parity of CauchyHssMM (cauchyhssmm.f90:1-123) for general Cauchy-like matrices
A(i,j) = u_i v_j / (d_i - f_j): the device walks cauchy2hss's post-order traversal
(CauchyHss_VP2.f90:387-613) with the same structured RRLU as the oracle, so every
node's U and V ranks must be identical to the oracle's; the product B := A B must equal
the oracle's HSS form times B to rounding, and the dense product to the compression
tolerance.  Shapes of the reference's own tests: n = 1000, ni = 120, ncol = 128, tol
1e-17 on interlaced points (testCauchyHss_serial.f90:27-88) and n = 2000, ni = 64,
ncol = 100, tol 1e-11 (testfmm.f90:17-43's points and sizes)."""
import numpy as np
import pytest

import oracle_py as O
from test_oracle_cauchy import interlaced, separated

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


@pytest.mark.parametrize("gen,n,ni,ncol,tol", [(interlaced, 1000, 120, 128, 1e-17), (interlaced, 2000, 64, 100, 1e-11),
                                               (separated, 1500, 100, 256, 1e-12), (interlaced, 3001, 200, 128, 1e-13)])
def test_cauchyhssmm_vs_oracle(ctx, gen, n, ni, ncol, tol):
    d, f, u, v = gen(n, n + 7)
    rng = np.random.default_rng(n)
    B = rng.uniform(0, 1, (n, n))
    Y, t, ranks = ctx.cauchyhssmm(d, f, u, v, B, ni, tol, ncol)
    H = O.cauchy2hss(d, f, u, v, ni, tol)
    try:
        usz, vsz = O.hss_ranks(H)
        Hd = O.hss_dense(H)
    finally:
        O.hss_free(H)
    assert ranks is not None
    got = ranks.reshape(-1, 2)[:len(usz)]
    assert np.array_equal(got[:, 0], usz[:, 1])
    assert np.array_equal(got[:, 1], vsz[:, 0])
    ref_h = Hd @ B
    scale = np.abs(ref_h).max()
    assert np.abs(Y - ref_h).max() <= 1e-12 * scale
    A = O.cauchy_dense_general(d, f, u, v)
    err = np.abs(Y - A @ B).max() / scale
    assert err <= max(100 * tol, 1e-13) * np.sqrt(n), err
    assert t[0] > 0 and t[1] > 0


def test_cauchyhssmm_single_leaf_is_dense(ctx):
    # npart (aux_hss.f90:803-827) folds a remainder below ni / 2 into the last leaf:
    # 90 = 64 + 26 is one leaf of 90, a tree with no off-diagonal blocks
    d, f, u, v = interlaced(90, 5)
    B = np.random.default_rng(1).standard_normal((90, 90))
    Y, t, ranks = ctx.cauchyhssmm(d, f, u, v, B, 64, 1e-12, 32)
    assert ranks is None
    A = O.cauchy_dense_general(d, f, u, v)
    assert np.abs(Y - A @ B).max() <= 1e-13 * np.abs(A @ B).max()


def test_cauchyhssmm_repeated_calls_reuse_the_context(ctx):
    d, f, u, v = separated(900, 3)
    B = np.random.default_rng(2).standard_normal((900, 900))
    Y1, _, r1 = ctx.cauchyhssmm(d, f, u, v, B, 64, 1e-12, 128)
    Y2, _, r2 = ctx.cauchyhssmm(d, f, u, v, B, 64, 1e-12, 900)  # another column blocking
    assert np.array_equal(r1, r2)
    assert np.abs(Y1 - Y2).max() <= 1e-14 * np.abs(Y1).max()
