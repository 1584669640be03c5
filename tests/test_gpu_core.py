"""GPU parity of the HIP divide-and-conquer path against the CPU oracle."""
import numpy as np
import pytest

import oracle_py as O
import lapack_ref as L
from hsspack_amd.workloads import tridiag_cfg

pytestmark = pytest.mark.gpu
EPS = 2.0 ** -53


@pytest.fixture(scope="module")
def torch_dev():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


@pytest.fixture(scope="module")
def ctx(torch_dev):
    from hsspack_amd import hsdc
    c = hsdc.Context()
    yield c
    c.close()


def tri(d, e):
    return np.diag(d) + np.diag(e, 1) + np.diag(e, -1)


def solve_dev(torch, ctx, d, e, **params):
    if params:
        ctx.set_params(**params)
    n = len(d)
    dev = torch.device("cuda:0")
    dd = torch.tensor(d, device=dev)
    ee = torch.tensor(np.append(e, 0.0), device=dev)
    zz = torch.zeros((n, n), dtype=torch.float64, device=dev)
    rc = ctx.stedc_dev(dd, ee, zz)
    torch.cuda.synchronize()
    return rc, dd.cpu().numpy(), zz.cpu().numpy().T.copy()


def check_eig(d, e, lam, z, tol=1e-12):
    T = tri(d, e)
    nrm = np.abs(T).max()
    n = len(d)
    orth = np.abs(z.T @ z - np.eye(n)).max()
    res = np.abs(T @ z - z * lam).max() / nrm
    assert orth <= tol, orth
    assert res <= tol, res


def test_secular_bitexact_vs_oracle(ctx):
    rng = np.random.default_rng(11)
    for k in (1, 2, 3, 5, 17, 64, 65, 300):
        d = np.sort(rng.uniform(-1, 1, k))
        if k > 4:
            d[2] = d[1] + 1e-12
        z = rng.standard_normal(k)
        z /= np.linalg.norm(z)
        rho = 0.7
        g = ctx.secular(d, z, rho)
        assert g["info"] == 0
        for j in range(1, k + 1):
            r = O.laed4(d, z, rho, j)
            assert g["lam"][j - 1] == r["lam"], (k, j)
            assert g["tau"][j - 1] == r["tau"], (k, j)
            assert g["org"][j - 1] == r["org"]
            if k > 2:
                assert g["niter"][j - 1] == r["niter"]


def _merge_cases():
    from golden.make_golden import merge_inputs
    out = []
    for name, (d, e) in [("rand", tridiag_cfg(2, 96, 11)), ("toep", tridiag_cfg(1, 80, 12)),
                         ("exact", (np.full(64, 2.0), np.full(63, 1.0))),
                         ("neg", (np.linspace(-1, 1, 50), -np.abs(tridiag_cfg(2, 49, 13)[0]) - 0.1)),
                         ("big", tridiag_cfg(2, 3000, 14)), ("bigexact", (np.full(2500, 2.0), np.full(2499, 1.0)))]:
        n = len(d)
        out.append((name, merge_inputs(d, e[: n - 1], n // 2), n // 2))
    # runs out of order (INDXQ not sorting the halves): the sequential DLAMRG path and
    # the plain deflation scan
    for name, (d, e) in [("unsorted", tridiag_cfg(2, 96, 15)), ("unsorted_toep", tridiag_cfg(1, 300, 16))]:
        n = len(d)
        dd, q, indxq, rho, z = merge_inputs(d, e[: n - 1], n // 2)
        indxq = indxq.copy()
        indxq[[3, 10]] = indxq[[10, 3]]
        indxq[[n // 2 + 2, n // 2 + 7]] = indxq[[n // 2 + 7, n // 2 + 2]]
        out.append((name, (dd, q, indxq, rho, z), n // 2))
    return out


@pytest.mark.parametrize("hss_min_k", [1 << 30, 16])
def test_deflate_bitexact_vs_oracle(ctx, hss_min_k):
    for name, (dd, q, indxq, rho, z), n1 in _merge_cases():
        r = O.laed2(dd, q, indxq, rho, z, n1, bsmlz=hss_min_k)
        g = ctx.deflate(dd, z, indxq, rho, n1, hss_min_k=hss_min_k)
        k = r["k"]
        assert g["info"] == 0 and g["k"] == k, name
        assert g["rho"] == r["rho"]
        assert np.array_equal(g["indxp"], r["indxp"]), name
        assert np.array_equal(g["indx"], r["indx"]), name
        if r["hflag"]:
            assert np.array_equal(g["indxc"], r["indxc"]), name
        assert np.array_equal(g["coltyp"][:4], r["coltyp"][:4]), name
        assert np.array_equal(g["dlamda"][:k], r["dlamda"][:k]), name
        assert np.array_equal(g["w"][:k], r["w"][:k]), name
        for a, b in zip(g["rot"], r["rot"]):
            assert np.array_equal(a, b), name
        assert np.array_equal(g["d"][k:], r["d"][k:]), name


@pytest.mark.parametrize("cfg,n", [(1, 26), (2, 100), (1, 257), (2, 513), (1, 1200), (2, 2000)])
def test_stedc_dev_dense_vs_oracle(torch_dev, ctx, cfg, n):
    d, e = tridiag_cfg(cfg, n, 100 + n)
    rc, lam, Z = solve_dev(torch_dev, ctx, d, e, hss_min_k=1 << 30, batched=1)
    assert rc == 0
    ref = O.stedc(d, e, prm=O.params(bsmlz=1 << 30))
    w, _, _ = L.dstedc(d, e)
    nrm = np.abs(tri(d, e)).max()
    assert np.abs(lam - ref["d"]).max() <= 4 * n * EPS * nrm
    assert np.abs(lam - w).max() <= 4 * n * EPS * nrm
    check_eig(d, e, lam, Z)


def test_plain_schedule_bitexact_with_batched(torch_dev, ctx):
    d, e = tridiag_cfg(2, 700, 5)
    rc1, l1, z1 = solve_dev(torch_dev, ctx, d, e, hss_min_k=1 << 30, batched=1)
    rc2, l2, z2 = solve_dev(torch_dev, ctx, d, e, hss_min_k=1 << 30, batched=0)
    ctx.set_params(batched=1)
    assert rc1 == 0 and rc2 == 0
    assert np.array_equal(l1, l2)
    assert np.array_equal(z1, z2)


STAT_KEYS = ("levels", "merges", "hss_merges", "roots", "deflated", "k_top", "secular_iters", "rank_max")


@pytest.mark.parametrize("cfg,n,hss_min_k,ni", [(2, 3000, 1 << 30, 200), (1, 2100, 128, 64),
                                               (1, 700, 64, 32), (2, 20, 1 << 30, 200)])
def test_device_schedule_bitexact_with_level_sync(torch_dev, ctx, cfg, n, hss_min_k, ni):
    """The queued schedule (launch sizes and tables derived on the device, the host
    called back only for levels holding HSS merges) and the per-level host round
    trip compute the same bits: same kernels, same per-root arithmetic."""
    d, e = tridiag_cfg(cfg, n, 60 + n)
    kw = dict(hss_min_k=hss_min_k, hss_leaf=ni, hss_tol=1e-13, batched=1)
    rc1, l1, z1 = solve_dev(torch_dev, ctx, d, e, level_sync=1, **kw)
    s1 = ctx.stats()
    rc0, l0, z0 = solve_dev(torch_dev, ctx, d, e, level_sync=0, **kw)
    s0 = ctx.stats()
    ctx.set_params(hss_min_k=1 << 30, hss_leaf=200, level_sync=0)
    assert rc0 == 0 and rc1 == 0
    assert np.array_equal(l0, l1)
    assert np.array_equal(z0, z1)
    for key in STAT_KEYS:
        assert s0[key] == s1[key], key
    if hss_min_k < n and cfg == 1:
        assert s0["hss_merges"] > 1  # the host was called back at several levels
    check_eig(d, e, l0, z0)


def test_device_schedule_batched_problems_with_hss(torch_dev, ctx):
    torch = torch_dev
    sizes = [700, 30, 1300, 257, 5, 900]
    ds, es = [], []
    for i, n in enumerate(sizes):
        d, e = tridiag_cfg(1 + (i % 2), n, 700 + i)
        ds.append(d)
        es.append(np.append(e, 0.0))
    offs = np.cumsum([0] + sizes[:-1])
    zoffs = np.cumsum([0] + [n * n for n in sizes[:-1]])
    dev = torch.device("cuda:0")
    out = []
    for sync in (1, 0):
        D = torch.tensor(np.concatenate(ds), device=dev)
        E = torch.tensor(np.concatenate(es), device=dev)
        Z = torch.zeros(int(sum(n * n for n in sizes)), dtype=torch.float64, device=dev)
        ctx.set_params(hss_min_k=200, hss_leaf=32, hss_tol=1e-13, level_sync=sync)
        rc = ctx.stedc_batched_dev(sizes, offs, D, E, Z, zoffs)
        torch.cuda.synchronize()
        assert rc == 0
        out.append((D.cpu().numpy(), Z.cpu().numpy(), ctx.stats()))
    ctx.set_params(hss_min_k=1 << 30, hss_leaf=200, level_sync=0)
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
    for key in STAT_KEYS:
        assert out[0][2][key] == out[1][2][key], key
    assert out[1][2]["hss_merges"] > 0
    Dh, Zh = out[1][0], out[1][1]
    for i, n in enumerate(sizes):
        lam = Dh[offs[i]:offs[i] + n]
        z = Zh[zoffs[i]:zoffs[i] + n * n].reshape(n, n).T
        check_eig(ds[i], es[i][:n - 1], lam, z)


def test_split_matrix(torch_dev, ctx):
    rng = np.random.default_rng(9)
    n = 150
    d = rng.uniform(-1, 1, n)
    e = rng.uniform(-1, 1, n - 1)
    e[[20, 57, 100]] = 0.0
    rc, lam, Z = solve_dev(torch_dev, ctx, d, e, hss_min_k=1 << 30)
    assert rc == 0
    assert np.all(np.diff(lam) >= 0)
    w, _, _ = L.dstedc(d, e)
    assert np.abs(lam - w).max() <= 4 * n * EPS * np.abs(tri(d, e)).max()
    check_eig(d, e, lam, Z)


def test_batched_problems(torch_dev, ctx):
    torch = torch_dev
    sizes = [30, 257, 64, 1000, 5, 400]
    ds, es = [], []
    for i, n in enumerate(sizes):
        d, e = tridiag_cfg(1 + (i % 2), n, 300 + i)
        ds.append(d)
        es.append(np.append(e, 0.0))
    offs = np.cumsum([0] + sizes[:-1])
    zoffs = np.cumsum([0] + [n * n for n in sizes[:-1]])
    dev = torch.device("cuda:0")
    D = torch.tensor(np.concatenate(ds), device=dev)
    E = torch.tensor(np.concatenate(es), device=dev)
    Z = torch.zeros(int(sum(n * n for n in sizes)), dtype=torch.float64, device=dev)
    ctx.set_params(hss_min_k=1 << 30)
    rc = ctx.stedc_batched_dev(sizes, offs, D, E, Z, zoffs)
    torch.cuda.synchronize()
    assert rc == 0
    Dh, Zh = D.cpu().numpy(), Z.cpu().numpy()
    for i, n in enumerate(sizes):
        lam = Dh[offs[i]:offs[i] + n]
        z = Zh[zoffs[i]:zoffs[i] + n * n].reshape(n, n).T
        w, _, _ = L.dstedc(ds[i], es[i][: n - 1])
        assert np.abs(lam - w).max() <= 4 * n * EPS * np.abs(tri(ds[i], es[i][:n - 1])).max()
        check_eig(ds[i], es[i][:n - 1], lam, z)


def test_host_api_and_info(ctx):
    n = 300
    d, e = tridiag_cfg(2, n, 3)
    ctx.set_params(hss_min_k=1 << 30)
    lam, z, info = ctx.stedc(d, e)
    assert info == 0
    w, _, _ = L.dstedc(d, e)
    assert np.abs(lam - w).max() <= 4 * n * EPS * np.abs(tri(d, e)).max()
    check_eig(d, e, lam, z)
    # COMPZ = 'V' applies the input Z
    q0, _ = np.linalg.qr(np.random.default_rng(2).standard_normal((n, n)))
    lam2, z2, info = ctx.stedc(d, e, compz="V", z=q0)
    assert info == 0
    np.testing.assert_allclose(np.abs(z2), np.abs(q0 @ z), atol=1e-12)


@pytest.mark.parametrize("n,pad", [(1500, 0), (1500, 3), (4000, 1)])
def test_host_api_staged_copy(ctx, n, pad):
    """Host-array COMPZ='I' at orders whose Z goes back through the staged copy with
    several host threads (4000: two staged blocks), with ldz = n and ldz > n."""
    d, e = tridiag_cfg(2, n, 5 + pad)
    ctx.set_params(hss_min_k=1 << 30)
    lam, z, info = ctx.stedc(d, e, ldz=n + pad)
    assert info == 0
    check_eig(d, e, lam, z)


@pytest.mark.parametrize("cfg,n,ni", [(1, 700, 32), (2, 700, 32), (1, 1100, 64), (1, 2100, 128)])
def test_stedc_hss_branch_vs_oracle(torch_dev, ctx, cfg, n, ni):
    d, e = tridiag_cfg(cfg, n, 40 + n)
    rc, lam, Z = solve_dev(torch_dev, ctx, d, e, hss_min_k=128, hss_leaf=ni, hss_tol=1e-13)
    st = ctx.stats()
    ctx.set_params(hss_min_k=1 << 30, hss_leaf=200)
    assert rc == 0
    ref = O.stedc(d, e, prm=O.params(bsmlz=128, ni=ni, tol=1e-13), trace_cap=4096)
    w, _, _ = L.dstedc(d, e)
    nrm = np.abs(tri(d, e)).max()
    assert np.abs(lam - w).max() <= 4 * n * EPS * nrm
    assert np.abs(lam - ref["d"]).max() <= 4 * n * EPS * nrm
    check_eig(d, e, lam, Z)
    if cfg == 1:
        assert st["hss_merges"] > 0
        assert st["rank_max"] > 0


def test_autotune_picks_a_candidate_and_restores_inputs(torch_dev, ctx):
    torch = torch_dev
    sizes = [1500, 700, 1200]
    ds, es = [], []
    for i, n in enumerate(sizes):
        d, e = tridiag_cfg(1, n, 500 + i)
        ds.append(d)
        es.append(np.append(e, 0.0))
    offs = np.cumsum([0] + sizes[:-1])
    zoffs = np.cumsum([0] + [n * n for n in sizes[:-1]])
    dev = torch.device("cuda:0")
    D = torch.tensor(np.concatenate(ds), device=dev)
    E = torch.tensor(np.concatenate(es), device=dev)
    D0, E0 = D.clone(), E.clone()
    Z = torch.zeros(int(sum(n * n for n in sizes)), dtype=torch.float64, device=dev)
    ctx.set_params(hss_min_k=1024, hss_tol=1e-15)
    try:
        leaf, smls, times = ctx.autotune(sizes, offs, D, E, Z, zoffs)
        assert leaf in (64, 128, 256) and smls in (25, 40)
        assert np.all(times > 0) and times[[25, 40].index(smls) * 3 + [64, 128, 256].index(leaf)] == times.min()
        assert torch.equal(D, D0) and torch.equal(E, E0)
        st = ctx.stats()
        ctx.set_params()  # the tuned parameters are the context's now
        assert ctx.params.hss_leaf == leaf and ctx.params.smlsiz == smls
        assert ctx.stedc_batched_dev(sizes, offs, D, E, Z, zoffs) == 0
        torch.cuda.synchronize()
        Dh, Zh = D.cpu().numpy(), Z.cpu().numpy()
        for i, n in enumerate(sizes):
            lam = Dh[offs[i]:offs[i] + n]
            z = Zh[zoffs[i]:zoffs[i] + n * n].reshape(n, n).T
            check_eig(ds[i], es[i][:n - 1], lam, z)
        del st
    finally:
        ctx.set_params(hss_min_k=1500, hss_tol=float(np.float32(1e-17)), hss_leaf=200, smlsiz=25)


def _solve_pair(torch, ctx, d, e, **kw):
    rc, lam, z = solve_dev(torch, ctx, d, e, **kw)
    assert rc == 0
    return lam, z, ctx.stats()


@pytest.mark.parametrize("cfg,n,hss_min_k,ni", [(2, 3000, 1 << 30, 200), (1, 2100, 128, 64)])
def test_graph_replays_match_direct_launches(torch_dev, ctx, cfg, n, hss_min_k, ni):
    """A whole-solve HIP graph is captured on the first call of a shape and replayed:
    every replay, with other work and synchronisations between them and with new
    inputs of the same shape, equals the directly launched solve bit for bit.  With
    HSS merges the replay stops at the first such level and the host continues."""
    torch = torch_dev
    kw = dict(hss_min_k=hss_min_k, hss_leaf=ni, hss_tol=1e-13, batched=1, level_sync=0)
    inputs = [tridiag_cfg(cfg, n, 80 + n), tridiag_cfg(cfg, n, 81 + n)]
    eager = [_solve_pair(torch, ctx, d, e, graph=0, **kw) for d, e in inputs]
    junk = torch.randn(1 << 20, device="cuda:0", dtype=torch.float64)
    for rep, i in enumerate([0, 0, 1, 0, 1]):
        lam, z, st = _solve_pair(torch, ctx, *inputs[i], graph=1, **kw)
        assert np.array_equal(lam, eager[i][0]), (rep, i)
        assert np.array_equal(z, eager[i][1]), (rep, i)
        for key in STAT_KEYS:
            assert st[key] == eager[i][2][key], (rep, key)
        junk.mul_(1.0001).add_(0.5)  # other work on the device between replays
        torch.cuda.synchronize()
    # the same on a caller stream
    strm = torch.cuda.Stream()
    ctx.set_stream(strm)
    try:
        with torch.cuda.stream(strm):
            lam, z, st = _solve_pair(torch, ctx, *inputs[1], graph=1, **kw)
            lam2, z2, _ = _solve_pair(torch, ctx, *inputs[0], graph=1, **kw)
    finally:
        ctx.set_stream(None)
    assert np.array_equal(lam, eager[1][0]) and np.array_equal(z, eager[1][1])
    assert np.array_equal(lam2, eager[0][0]) and np.array_equal(z2, eager[0][1])
    ctx.set_params(hss_min_k=1 << 30, hss_leaf=200, graph=1)



def test_device_input_path_matches_host_input_path(torch_dev, ctx):
    """hsdc_stedc_dev scans, scales and tears a device-resident input on the device
    (plan from the order alone, tables kept between calls); with the per-level host
    round trip (level_sync = 1) the host prepares the input instead.  Both give the
    same bits, also after other entry points have reused the context's tables."""
    for cfg, n, seed in [(2, 3000, 31), (1, 2500, 32), (2, 3000, 33)]:
        d, e = tridiag_cfg(cfg, n, seed)
        rc1, lam_h, z_h = solve_dev(torch_dev, ctx, d, e, hss_min_k=1 << 30, batched=1, level_sync=1)
        rc0, lam_d, z_d = solve_dev(torch_dev, ctx, d, e, level_sync=0)
        assert rc0 == 0 and rc1 == 0
        assert np.array_equal(lam_h, lam_d)
        assert np.array_equal(z_h, z_d)
        ctx.secular(np.array([0.1, 0.5, 0.9]), np.array([0.6, 0.6, 0.52915]), 0.3)  # reuses the tables


@pytest.mark.parametrize("case", ["zero", "split", "near_split", "scaled_zero_e"])
def test_device_input_handed_to_host(torch_dev, ctx, case):
    """Inputs the device screen sends back to the host: zero, an exact split, an
    off-diagonal within twice MDSTEDC's split bound, zero off-diagonals."""
    n = 400
    d, e = tridiag_cfg(2, n, 44)
    d, e = d.copy(), e.copy()
    if case == "zero":
        d[:] = 0.0
        e[:] = 0.0
    elif case == "split":
        e[123] = 0.0
    elif case == "near_split":
        e[200] = 1.5 * EPS * np.sqrt(abs(d[200])) * np.sqrt(abs(d[201]))
    else:
        e[::7] = 0.0
    rc1, lam_h, z_h = solve_dev(torch_dev, ctx, d, e, hss_min_k=1 << 30, batched=1, level_sync=1)
    rc0, lam, z = solve_dev(torch_dev, ctx, d, e, level_sync=0)
    assert rc0 == 0 and rc1 == 0
    assert np.array_equal(lam, lam_h)
    assert np.array_equal(z, z_h)
    if case == "zero":
        assert np.array_equal(z, np.eye(n)) and not lam.any()
    else:
        check_eig(d, e, lam, z)
