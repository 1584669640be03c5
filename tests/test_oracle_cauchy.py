"""The oracle's general Cauchy-like HSS construction (cauchy2hss and rrluCauchy,
CauchyHss_VP2.f90:7-613) pinned by what the mathematics fixes: the HSS form expands
to the dense matrix A(i,j) = u_i v_j / (d_i - f_j) within the compression tolerance
(and to rounding at tol 1e-17); ranks do not grow as the tolerance loosens; every
row generator is interpolative (an identity row at each skeleton row); a matrix of
rank one off its diagonal blocks compresses to rank <= 2."""
import numpy as np
import pytest

import oracle_py as O


def interlaced(n, seed):
    """The points of the reference's HSS multiply tests (testCauchyHss_serial.f90:46-55):
    f_i = 0.1 + (2i - 1) gap, d_i = f_i + gap, gap = 8.9 / (2n); u, v uniform."""
    gap = (9.0 - 0.1) / (2 * n)
    f = 0.1 + (2 * np.arange(1, n + 1) - 1) * gap
    d = f + gap
    rng = np.random.default_rng(seed)
    return d, f, rng.uniform(0, 1, n), rng.uniform(0, 1, n)


def separated(n, seed):
    rng = np.random.default_rng(seed)
    return (np.sort(rng.uniform(0, 1, n)), np.sort(rng.uniform(2, 3, n)), rng.standard_normal(n),
            rng.standard_normal(n))


@pytest.mark.parametrize("gen,n,ni,tol", [(interlaced, 1000, 120, 1e-17), (interlaced, 700, 64, 1e-10),
                                          (separated, 600, 50, 1e-12), (interlaced, 130, 40, 1e-17)])
def test_expansion_matches_dense(gen, n, ni, tol):
    d, f, u, v = gen(n, n)
    A = O.cauchy_dense_general(d, f, u, v)
    H = O.cauchy2hss(d, f, u, v, ni, tol)
    try:
        Hd = O.hss_dense(H)
        usz, vsz = O.hss_ranks(H)
    finally:
        O.hss_free(H)
    err = np.abs(Hd - A).max() / np.abs(A).max()
    assert err <= max(100 * tol, 1e-13) * np.sqrt(n), err
    assert usz[:-1, 1].max() < ni  # the off-diagonal blocks are compressed


def test_direct_formula_on_a_tiny_case():
    # brute force: the dense fill is u_i v_j / (d_i - f_j) exactly
    d, f, u, v = interlaced(5, 3)
    A = O.cauchy_dense_general(d, f, u, v)
    assert np.array_equal(A, np.outer(u, v) / (d[:, None] - f[None, :]))


def test_ranks_monotone_in_tolerance():
    d, f, u, v = interlaced(800, 8)
    last = None
    for tol in (1e-4, 1e-8, 1e-12, 1e-17):
        H = O.cauchy2hss(d, f, u, v, 64, tol)
        try:
            usz, vsz = O.hss_ranks(H)
        finally:
            O.hss_free(H)
        r = usz[:-1, 1].sum() + vsz[:-1, 0].sum()
        if last is not None:
            assert r >= last
        last = r


def test_row_generators_are_interpolative():
    d, f, u, v = interlaced(400, 11)
    H = O.cauchy2hss(d, f, u, v, 50, 1e-12)
    try:
        usz, vsz = O.hss_ranks(H)
        pu = np.ctypeslib.as_array(H.pu, (H.ltr,))
        h = np.ctypeslib.as_array(H.h, (H.nh,))
        for i in range(H.ltr - 1):
            M, r = usz[i]
            if r == 0 or r >= M:
                continue
            Ui = h[pu[i]:pu[i] + M * r].reshape(r, M).T
            # every column is a unit vector at some row: the skeleton rows
            eye_rows = [k for k in range(M) if np.count_nonzero(Ui[k]) == 1 and np.max(np.abs(Ui[k])) == 1.0]
            assert len(eye_rows) >= r
    finally:
        O.hss_free(H)


def test_rank_one_off_diagonal_blocks():
    # d and f far apart and nearly constant: every off-diagonal block is numerically rank one
    n = 512
    d = 10.0 + 1e-9 * np.arange(n)
    f = -10.0 - 1e-9 * np.arange(n)
    rng = np.random.default_rng(2)
    u, v = rng.uniform(1, 2, n), rng.uniform(1, 2, n)
    H = O.cauchy2hss(d, f, u, v, 64, 1e-12)
    try:
        usz, vsz = O.hss_ranks(H)
    finally:
        O.hss_free(H)
    assert usz[:-1, 1].max() <= 2 and vsz[:-1, 0].max() <= 2
