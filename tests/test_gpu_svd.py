"""GPU parity of the bidiagonal divide-and-conquer SVD (MDBDSDC / MDLASD0-2) against
the CPU oracle: bit-exact singular values, vectors and per-merge deflation (K and
rotation counts), LAPACK DBDSDC to its error bound, and orthogonality / residual at
the full workload size."""
import numpy as np
import pytest

import lapack_ref as L
import oracle_py as O
from hsspack_amd import hsdc
from hsspack_amd.workloads import bidiag_cfg

pytestmark = pytest.mark.gpu
EPS = 2.0 ** -53


@pytest.fixture(scope="module")
def ctx():
    c = hsdc.BdContext()
    yield c
    c.close()


def bidiag(d, e):
    n = len(d)
    return np.diag(d) + np.diag(e[: n - 1], 1)


def oracle_trace(tr):
    return tr["n"], tr["k"], tr["nrot"]


def cases():
    out = []
    for n in (1, 2, 20, 25, 26, 60, 301, 1000, 2500):
        d, e = bidiag_cfg(n, seed=5)
        out.append(("cfg5-%d" % n, d, e, "U"))
    d, e = bidiag_cfg(700, seed=8)
    out.append(("lower-700", d, e, "L"))
    n = 512
    out.append(("const-512", np.ones(n), np.ones(n - 1), "U"))  # equal child spectra: rotations
    d, e = bidiag_cfg(400, seed=9)
    e[[50, 51, 200, 398]] = 1e-20  # splits, a trailing 1 x 1
    d[123] = 1e-30                 # raised to eps
    out.append(("split-400", d, e, "U"))
    d, e = bidiag_cfg(300, seed=10)
    e[:] *= 1e-9                   # heavy small-z deflation
    out.append(("tiny-e-300", d, e, "U"))
    return out


@pytest.mark.parametrize("case", range(13))
def test_bdsdc_bitexact_vs_oracle(ctx, case):
    name, d, e, uplo = cases()[case]
    print("case", name, flush=True)
    s, u, vt, info = ctx.bdsdc(d, e, uplo=uplo)
    assert info == 0
    r = O.bdsdc(d, e, uplo=uplo.encode(), trace_cap=4096)
    assert r["info"] == 0
    assert np.array_equal(s, r["d"])
    assert np.array_equal(u, r["u"])
    assert np.array_equal(vt, r["vt"])
    tr = ctx.trace()
    ot = r["trace"]
    # merge order: the device runs every problem's level t together, the oracle one
    # problem after another; compare as multisets of (order, K, rotations)
    mine = sorted(zip(tr["n"].tolist(), tr["k"].tolist(), tr["nrot"].tolist()))
    ref = sorted(zip(ot["n"].tolist(), ot["k"].tolist(), ot["nrot"].tolist()))
    assert mine == ref
    st = ctx.stats()
    assert st["roots"] == ot["roots"]
    assert st["secular_iters"] == ot["iters"]


def test_rotation_and_deflation_paths_exercised(ctx):
    for name, d, e, uplo in cases():
        if name in ("const-512", "tiny-e-300"):
            ctx.bdsdc(d, e, uplo=uplo)
            tr = ctx.trace()
            if name == "const-512":
                assert tr["nrot"].sum() > 0
            else:
                assert np.any(tr["k"] < tr["n"])


@pytest.mark.parametrize("n", [1000, 4096])
def test_bdsdc_vs_lapack(ctx, n):
    d, e = bidiag_cfg(n, seed=5)
    s, u, vt, info = ctx.bdsdc(d, e)
    assert info == 0
    sl = L.dbdsdc(d, e)[0]
    B = bidiag(d, e)
    nrm = np.abs(B).max()
    assert np.abs(s - sl).max() <= 4 * n * EPS * nrm
    assert np.abs(u.T @ u - np.eye(n)).max() <= 1e-12
    assert np.abs(vt @ vt.T - np.eye(n)).max() <= 1e-12
    assert np.abs((u * s) @ vt - B).max() / nrm <= 1e-12


def test_values_only_matches_vectors_run(ctx):
    d, e = bidiag_cfg(777, seed=5)
    s1 = ctx.bdsdc(d, e, compq="I")[0]
    s2, _, _, info = ctx.bdsdc(d, e, compq="N")
    assert info == 0
    assert np.array_equal(s1, s2)


def test_bdsdc_device_resident_full_size(ctx):
    """The headline bidiagonal workload (n = 16384) through the device entry: sorted
    singular values, orthogonality and residual of the product through torch."""
    torch = pytest.importorskip("torch")
    n = 16384
    d, e = bidiag_cfg(n, seed=5)
    dd = torch.tensor(d, device="cuda")
    de = torch.tensor(e, device="cuda")
    U = torch.empty((n, n), dtype=torch.float64, device="cuda")
    VT = torch.empty((n, n), dtype=torch.float64, device="cuda")
    rc = ctx.bdsdc_dev(dd, de, U, VT)
    torch.cuda.synchronize()
    assert rc == 0
    s = dd.cpu().numpy()
    assert np.all(np.diff(s) <= 0)
    Uc = U.t()  # column-major storage: U.t() is the matrix
    Vc = VT.t()
    I = torch.eye(n, dtype=torch.float64, device="cuda")
    orth_u = (Uc.t() @ Uc - I).abs().max().item()
    orth_v = (Vc @ Vc.t() - I).abs().max().item()
    B = torch.zeros((n, n), dtype=torch.float64, device="cuda")
    idx = torch.arange(n, device="cuda")
    B[idx, idx] = torch.tensor(d, device="cuda")
    B[idx[:-1], idx[1:]] = torch.tensor(e, device="cuda")
    res = ((Uc * dd) @ Vc - B).abs().max().item() / B.abs().max().item()
    print("orth_u %.2e orth_v %.2e res %.2e" % (orth_u, orth_v, res), flush=True)
    assert orth_u <= 1e-12 and orth_v <= 1e-12 and res <= 1e-12
    st = ctx.stats()
    assert st["levels"] >= 9 and st["merges"] > 0


@pytest.mark.parametrize("n,uplo", [(1, "U"), (1, "L"), (2, "L"), (3, "U"), (25, "L")])
def test_small_orders_bitexact(ctx, n, uplo):
    d, e = bidiag_cfg(n, seed=12)
    d[0] = -d[0]  # a negative diagonal entry: U carries the sign
    s, u, vt, info = ctx.bdsdc(d, e, uplo=uplo)
    assert info == 0
    r = O.bdsdc(d, e, uplo=uplo.encode())
    assert np.array_equal(s, r["d"]) and np.array_equal(u, r["u"]) and np.array_equal(vt, r["vt"])
    B = bidiag(d, e) if uplo == "U" else bidiag(d, e).T
    assert np.abs((u * s) @ vt - B).max() <= 4 * n * EPS * max(1.0, np.abs(B).max()) + 1e-14


def test_device_entry_values_only_and_smlsiz(ctx):
    """hsdc_bdsdc_dev with U and VT omitted gives the same singular values; a larger leaf
    bound (40) stays bit-exact with the oracle run at the same bound."""
    torch = pytest.importorskip("torch")
    n = 900
    d, e = bidiag_cfg(n, seed=13)
    s_ref = ctx.bdsdc(d, e)[0]
    dd = torch.tensor(d, device="cuda")
    de = torch.tensor(e, device="cuda")
    assert ctx.bdsdc_dev(dd, de, None, None) == 0
    torch.cuda.synchronize()
    assert np.array_equal(dd.cpu().numpy(), s_ref)
    c40 = hsdc.BdContext(smlsiz=40)
    s, u, vt, info = c40.bdsdc(d, e)
    assert info == 0
    r = O.bdsdc(d, e, prm=O.params(smlsiz=40))
    assert np.array_equal(s, r["d"]) and np.array_equal(u, r["u"]) and np.array_equal(vt, r["vt"])
    c40.close()


# ---------------------------------------------------------------- HSS branch (K >= hss_min_k)
def near_toeplitz(n, seed):
    from hsspack_amd.workloads import bidiag_toeplitz
    return bidiag_toeplitz(n, seed)


@pytest.mark.parametrize("gen,n,min_k,uplo", [("toeplitz", 1000, 100, "U"), ("uniform", 2000, 64, "U"),
                                              ("toeplitz", 2200, 2000, "U"), ("toeplitz", 900, 100, "L")])
def test_bdsdc_hss_branch_vs_oracle(gen, n, min_k, uplo):
    """Merges with K >= hss_min_k take the HSS branch on both sides: the same merges,
    K and generator ranks as the oracle, singular values within 4 n eps, vectors
    orthonormal with residual <= 1e-12 and within 1e-11 of the oracle's."""
    d, e = near_toeplitz(n, 7) if gen == "toeplitz" else bidiag_cfg(n, seed=7)
    print("case", gen, n, min_k, uplo, flush=True)
    c = hsdc.BdContext(hss_min_k=min_k)
    try:
        s, u, vt, info = c.bdsdc(d, e, uplo=uplo)
        tr = c.trace()
        st = c.stats()
    finally:
        c.close()
    assert info == 0
    r = O.bdsdc(d, e, uplo=uplo.encode(), prm=O.params(bd_bsmlz=min_k), trace_cap=4 * n)
    assert r["info"] == 0
    ot = r["trace"]
    assert np.array_equal(tr["k"], ot["k"])
    hs = tr["hss_rank"] > 0
    assert hs.any() and np.array_equal(hs, ot["rank_max"] > 0)
    assert st["hss_merges"] == hs.sum()
    # ranks of the first HSS level match exactly (same inputs); later levels see
    # vectors that differ in rounding, so their ranks may move by a little
    assert np.all(np.abs(tr["hss_rank"] - ot["rank_max"]) <= 4)
    tol_s = 4 * n * EPS * max(s[0], 1.0)
    assert np.abs(s - r["d"]).max() <= tol_s
    B = bidiag(d, e) if uplo == "U" else bidiag(d, e).T
    assert np.abs(u.T @ u - np.eye(n)).max() <= 1e-12
    assert np.abs(vt @ vt.T - np.eye(n)).max() <= 1e-12
    assert np.abs(u @ np.diag(s) @ vt - B).max() / np.abs(B).max() <= 1e-12
    assert np.abs(u - r["u"]).max() <= 1e-11
    assert np.abs(vt - r["vt"]).max() <= 1e-11


def test_bdsdc_hss_branch_small_trees():
    """HSS branch with small leaves: threshold 16 and leaf 8 put the merges of this
    matrix with K from 47 to 106 on the HSS branch (trees of 6 to 13 leaves of 8); both
    sides take the same merges, and the vectors meet the orthogonality and residual
    bounds."""
    n = 400
    d, e = bidiag_cfg(n, seed=21)
    print("case small trees", n, flush=True)
    c = hsdc.BdContext(hss_min_k=16, hss_leaf=8)
    try:
        s, u, vt, info = c.bdsdc(d, e)
        tr = c.trace()
    finally:
        c.close()
    assert info == 0
    r = O.bdsdc(d, e, prm=O.params(bd_bsmlz=16, bd_ni=8), trace_cap=4 * n)
    assert r["info"] == 0
    ot = r["trace"]
    assert np.array_equal(tr["k"], ot["k"])
    hs = tr["hss_rank"] > 0
    assert hs.sum() >= 4 and np.array_equal(hs, ot["rank_max"] > 0)
    assert np.abs(s - r["d"]).max() <= 4 * n * EPS * s[0]
    B = bidiag(d, e)
    assert np.abs(u.T @ u - np.eye(n)).max() <= 1e-12
    assert np.abs(vt @ vt.T - np.eye(n)).max() <= 1e-12
    assert np.abs(u @ np.diag(s) @ vt - B).max() / np.abs(B).max() <= 1e-12


def test_bdsdc_hss_branch_full_size():
    """The reference threshold (BSMLZ 2000) on a bidiagonal that barely deflates:
    n = 8192, the top merges take the HSS branch; device resident."""
    torch = pytest.importorskip("torch")
    n = 8192
    d, e = near_toeplitz(n, 11)
    print("case toeplitz", n, flush=True)
    c = hsdc.BdContext()
    try:
        dd = torch.tensor(d, device="cuda")
        de = torch.tensor(e, device="cuda")
        U = torch.empty((n, n), dtype=torch.float64, device="cuda")
        VT = torch.empty((n, n), dtype=torch.float64, device="cuda")
        rc = c.bdsdc_dev(dd, de, U, VT)
        torch.cuda.synchronize()
        st = c.stats()
    finally:
        c.close()
    assert rc == 0
    assert st["hss_merges"] >= 3 and st["hss_rank_max"] > 0
    s = dd.cpu().numpy()
    assert np.all(np.diff(s) <= 0)
    Uc = U.t()  # column-major storage: U.t() is the matrix
    Vc = VT.t()
    I = torch.eye(n, dtype=torch.float64, device="cuda")
    orth_u = (Uc.t() @ Uc - I).abs().max().item()
    orth_v = (Vc @ Vc.t() - I).abs().max().item()
    B = torch.zeros((n, n), dtype=torch.float64, device="cuda")
    idx = torch.arange(n, device="cuda")
    B[idx, idx] = torch.tensor(d, device="cuda")
    B[idx[:-1], idx[1:]] = torch.tensor(e, device="cuda")
    res = ((Uc * dd) @ Vc - B).abs().max().item() / B.abs().max().item()
    print("hss merges %d rank %d orth_u %.2e orth_v %.2e res %.2e" % (st["hss_merges"], st["hss_rank_max"],
                                                                  orth_u, orth_v, res), flush=True)
    assert orth_u <= 1e-12 and orth_v <= 1e-12 and res <= 1e-12
