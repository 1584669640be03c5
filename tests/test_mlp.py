"""Learned starting iterate of the secular solver: the oracle (CPU) and the
device forward / training kernels against it (GPU)."""
import os
import sys

import numpy as np
import pytest

import oracle_py as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import mlp_ref as M  # noqa: E402

from hsspack_amd import hsdc  # noqa: E402
from hsspack_amd.workloads import tridiag_cfg  # noqa: E402

EPS = 2.0 ** -53
WEIGHTS = os.path.join(ROOT, "hsspack_amd", "data", "secular_mlp.bin")


def secular_case(k, seed):
    rng = np.random.default_rng(seed)
    d = np.sort(rng.uniform(-1, 1, k))
    z = rng.standard_normal(k)
    z /= np.linalg.norm(z)
    return d, z, 0.9


def random_weights(seed, scale=0.3):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(M.NPARAM) * scale).astype(np.float32)


def identity_on_classical():
    """A network whose output is 0: the proposal is the classical start itself."""
    return np.zeros(M.NPARAM, np.float32)


def random_batch(ns, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, (ns, M.NIN)).astype(np.float32)
    x[:, 16:32] = np.abs(x[:, 16:32])
    t = rng.normal(0, 0.3, ns).astype(np.float32)
    return x, t


# ------------------------------------------------------------------ CPU: oracle
def test_gradient_oracle_matches_finite_differences():
    x, t = random_batch(40, 1)
    w = random_weights(2, 0.25).astype(np.float64)
    L0, g = M.loss_grad(w, x, t)
    assert abs(L0 - M.loss(w, x, t)) <= 1e-12 * abs(L0)
    rng = np.random.default_rng(3)
    # every block of the layout: W1, b1, W2, b2, w3, b3
    picks = list(rng.choice(M.B1, 12, replace=False)) + [M.B1 + 5, M.B1 + 40]
    picks += [M.W2 + i for i in rng.choice(M.NH * M.NH, 12, replace=False)]
    picks += [M.B2 + 3, M.B2 + 60, M.W3 + 7, M.W3 + 63, M.B3]
    h = 1e-6
    for p in picks:
        wp, wm = w.copy(), w.copy()
        wp[p] += h
        wm[p] -= h
        fd = (M.loss(wp, x, t) - M.loss(wm, x, t)) / (2 * h)
        assert abs(fd - g[p]) <= 1e-6 * (abs(g[p]) + 1e-3), (p, fd, g[p])


def test_c_forward_matches_fp64_forward():
    w = random_weights(4)
    x, _ = random_batch(50, 5)
    y64 = M.forward(w, x)
    for s in range(50):
        y = O.mlp_forward(w, x[s])
        assert abs(y - y64[s]) <= 2e-5 * (1 + abs(y64[s])), s


def test_features_invariants():
    d, z, rho = secular_case(40, 6)
    for r in (0, 3, 20, 38, 39):
        W = d[r + 1] - d[r] if r < 39 else rho * np.sum(z * z)
        f = O.mlp_features(d, z, rho, r, W, 0.25)
        # pole gaps: left ones negative, right ones positive, |u| < 1, nearest largest
        for q, tt in enumerate(list(range(-8, 0)) + list(range(2, 10))):
            i = r + tt
            if 0 <= i < 40:
                u = W / (d[i] - d[r])
                assert np.sign(f[q]) == np.sign(u) and abs(f[q]) < 1
                assert abs(f[q] - u / (1 + abs(u))) <= 1e-7
            else:
                assert f[q] == 0
        # z^2 window normalised to one
        assert abs(float(np.sum(f[16:32], dtype=np.float64)) - 1.0) <= 1e-6
        assert f[33] == np.float32(0.25)
        # rho feature: W / (rho * window) through q/(1+q)
        lo, hi = max(0, r - 7), min(40, r + 9)
        qv = W / (rho * np.sum(z[lo:hi] ** 2))
        assert abs(f[32] - qv / (1 + qv)) <= 1e-7


def test_learned_start_keeps_the_root():
    """Any proposal (in or out of the bracket) converges to the root: the learned
    start meets the accuracy the classical start meets against 40-digit roots, and
    both stop inside the band |f| <= eps * erretm around the same root."""
    mp = pytest.importorskip("mpmath")
    mp.mp.dps = 40
    errs = []
    for seed, k in ((7, 70), (8, 64)):
        d, z, rho = secular_case(k, seed)
        dm = [mp.mpf(x) for x in d]
        zm = [mp.mpf(x) for x in z]
        rm = mp.mpf(rho)
        for w in (random_weights(seed), identity_on_classical()):
            nused = 0
            for j in range(1, k + 1):
                a = O.laed4(d, z, rho, j)
                b = O.laed4_mlp(d, z, rho, j, w, gate=-1.0)
                assert b["info"] == 0 and b["org"] == a["org"]
                assert abs(b["lam"] - a["lam"]) <= 64 * EPS * abs(d).max()
                nused += b["used"]
                org = b["org"]

                def f(t):
                    return 1 / rm + sum(zz ** 2 / ((dd - dm[org - 1]) - t) for dd, zz in zip(dm, zm))
                if j == k:
                    lo, hi = mp.mpf(0), rm * sum(zz ** 2 for zz in zm)
                elif org == j:
                    lo, hi = mp.mpf(0), dm[j] - dm[j - 1]
                else:
                    lo, hi = -(dm[j] - dm[j - 1]), mp.mpf(0)
                wd = hi - lo
                te = mp.findroot(f, (lo + wd * mp.mpf(10) ** -30, hi - wd * mp.mpf(10) ** -30),
                                 solver="anderson")
                errs.append(float(abs((mp.mpf(b["tau"]) - te) / te)) / EPS)
            assert nused > 0
    errs = np.array(errs)
    assert np.median(errs) <= 2.0
    assert np.percentile(errs, 99) <= 64.0


def test_zero_network_reproduces_classical_solve():
    """y = 0 proposes exp(0) times the classical start: the same iterates bit for bit,
    whether the gate takes the proposal (gate < 0) or keeps the classical start."""
    d, z, rho = secular_case(120, 9)
    w = identity_on_classical()
    for gate, used in ((-1.0, 1), (0.0, 0), (0.5, 0)):
        for j in range(1, 121):
            a = O.laed4(d, z, rho, j)
            b = O.laed4_mlp(d, z, rho, j, w, gate=gate)
            assert b["used"] == used
            assert (b["lam"], b["tau"], b["niter"]) == (a["lam"], a["tau"], a["niter"])


def test_gate_selects_large_corrections_only():
    d, z, rho = secular_case(200, 19)
    w = random_weights(20)
    n0 = sum(O.laed4_mlp(d, z, rho, j, w, gate=0.0)["used"] for j in range(1, 201))
    n1 = sum(O.laed4_mlp(d, z, rho, j, w, gate=2.0)["used"] for j in range(1, 201))
    n2 = sum(O.laed4_mlp(d, z, rho, j, w, gate=100.0)["used"] for j in range(1, 201))
    assert n0 >= n1 >= n2 == 0 and n0 > n1


def test_series_exp():
    xs = np.linspace(-16, 16, 4001)
    for x in xs:
        v = O.sec_exp(x)
        assert abs(v - np.exp(x)) <= 8 * EPS * np.exp(x), x
    assert O.sec_exp(0.0) == 1.0


def test_weights_file_round_trip(tmp_path):
    w = random_weights(10)
    p = str(tmp_path / "w.bin")
    hsdc.save_mlp_weights(p, w)
    assert np.array_equal(hsdc.load_mlp_weights(p), w)
    raw = bytearray(open(p, "rb").read())
    raw[100] ^= 1
    open(p, "wb").write(bytes(raw))
    with pytest.raises(ValueError):
        hsdc.load_mlp_weights(p)


def test_shipped_weights_file_is_valid():
    w = hsdc.load_mlp_weights(WEIGHTS)
    assert w.shape == (M.NPARAM,) and np.all(np.isfinite(w))


# ------------------------------------------------------------------ GPU
@pytest.fixture(scope="module")
def torch_dev():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


@pytest.fixture(scope="module")
def ctx(torch_dev):
    c = hsdc.Context()
    yield c
    c.close()


@pytest.mark.gpu
def test_device_gradient_vs_oracle(ctx):
    w = random_weights(11, 0.2)
    ctx.mlp_set_weights(w)
    for ns, seed in ((3001, 12), (17, 13)):
        x, t = random_batch(ns, seed)
        g, loss = ctx.mlp_grad(x, t)
        L64, g64 = M.loss_grad(w.astype(np.float64), x, t)
        assert abs(loss - L64) <= 1e-4 * L64
        err = np.linalg.norm(g - g64) / np.linalg.norm(g64)
        assert err <= 1e-4, err
        assert np.all(np.abs(g - g64) <= 1e-3 * np.abs(g64).max())
        g2, loss2 = ctx.mlp_grad(x, t)
        assert np.array_equal(g, g2) and loss == loss2


@pytest.mark.gpu
@pytest.mark.parametrize("which,gate", [("random", 0.0), ("random", 0.5), ("identity", 0.0),
                                        ("shipped", None)])
def test_device_learned_secular_bitexact_vs_oracle(ctx, which, gate):
    w = {"random": lambda: random_weights(14),
         "identity": identity_on_classical,
         "shipped": lambda: hsdc.load_mlp_weights(WEIGHTS)}[which]()
    gate = float(hsdc.default_params().learned_gate) if gate is None else gate
    ctx.mlp_set_weights(w)
    ctx.set_params(learned=1, learned_min_k=3, learned_gate=gate)
    try:
        for k, seed in ((3, 15), (64, 16), (301, 17), (1000, 18)):
            d, z, rho = secular_case(k, seed)
            g = ctx.secular(d, z, rho)
            st = ctx.stats()
            assert g["info"] == 0
            used = 0
            for j in range(1, k + 1):
                r = O.laed4_mlp(d, z, rho, j, w, gate=gate)
                assert g["lam"][j - 1] == r["lam"], (k, j)
                assert g["tau"][j - 1] == r["tau"], (k, j)
                assert g["org"][j - 1] == r["org"]
                assert g["niter"][j - 1] == r["niter"]
                used += r["used"]
            assert st["learned_roots"] == used
            assert st["secular_iters"] == int(np.sum(g["niter"]))
            nit = np.asarray(g["niter"])
            per_wave = 64 // O.sec_team(k)  # roots per wavefront
            assert st["secular_wave_iters"] == sum(int(nit[b:b + per_wave].max()) for b in range(0, k, per_wave))
    finally:
        ctx.set_params(learned=0, learned_min_k=256, learned_gate=hsdc.default_params().learned_gate)


@pytest.mark.gpu
def test_training_on_harvested_roots_reduces_loss(torch_dev, ctx):
    torch = torch_dev
    ctx.mlp_set_weights(None)
    ctx.set_params(learned=0, learned_min_k=64, hss_min_k=1 << 30)
    ctx.mlp_harvest(1 << 16)
    dev = torch.device("cuda:0")
    for seed in (1000, 1001):
        d, e = tridiag_cfg(1, 2048, seed)
        dd = torch.tensor(d, device=dev)
        ee = torch.tensor(np.append(e, 0.0), device=dev)
        zz = torch.zeros((2048, 2048), dtype=torch.float64, device=dev)
        assert ctx.stedc_dev(dd, ee, zz) == 0
    x, t = ctx.mlp_samples(1 << 16)
    ctx.mlp_harvest(0)
    assert len(t) > 5000
    assert np.all(np.isfinite(t)) and np.all(np.isfinite(x))
    # feature 33, the classical start, is a fraction of the bracket
    assert np.all((x[:, 33] >= 0) & (x[:, 33] <= 1))
    hist = ctx.mlp_train(x, t, epochs=6, batch=1024, lr=2e-3, seed=5)
    assert np.all(np.isfinite(hist))
    assert hist[-1] < hist[0], hist
    w = ctx.mlp_get_weights()
    L_after = M.loss(w.astype(np.float64), x, t)
    L_zero = float(np.mean(t.astype(np.float64) ** 2))  # y = 0: the classical start
    assert L_after < L_zero, (L_after, L_zero)
    ctx.set_params(learned_min_k=256, hss_min_k=1500)
    ctx.mlp_set_weights(None)


@pytest.mark.gpu
def test_shipped_weights_save_iterations_and_keep_accuracy(torch_dev, ctx):
    """The committed network on a held-out config-1 matrix: fewer secular iterations
    than the classical start, same eigenpair accuracy."""
    torch = torch_dev
    n = 8192
    d, e = tridiag_cfg(1, n, 4242)
    dev = torch.device("cuda:0")
    res = {}
    ctx.mlp_load(WEIGHTS)
    try:
        for learned in (0, 1):
            ctx.set_params(learned=learned, learned_min_k=256, hss_min_k=1024, hss_leaf=128, hss_tol=1e-15)
            dd = torch.tensor(d, device=dev)
            ee = torch.tensor(np.append(e, 0.0), device=dev)
            zz = torch.zeros((n, n), dtype=torch.float64, device=dev)
            assert ctx.stedc_dev(dd, ee, zz) == 0
            res[learned] = (dd.cpu().numpy(), zz, ctx.stats())
    finally:
        ctx.set_params(learned=0, hss_min_k=1500, hss_leaf=200, hss_tol=float(np.float32(1e-17)))
    lam0, _, s0 = res[0]
    lam1, z1, s1 = res[1]
    assert s1["learned_roots"] > 0
    assert s1["secular_iters"] < s0["secular_iters"], (s0["secular_iters"], s1["secular_iters"])
    nrm = max(np.abs(d).max(), np.abs(e).max()) * 3
    assert np.abs(lam1 - lam0).max() <= 4 * n * EPS * nrm
    # orthogonality and residual of the learned run, on the device
    Z = z1.T
    T = torch.diag(torch.tensor(d, device=dev)) + torch.diag(torch.tensor(e, device=dev), 1) + \
        torch.diag(torch.tensor(e, device=dev), -1)
    L = torch.tensor(lam1, device=dev)
    orth = (Z.T @ Z - torch.eye(n, dtype=torch.float64, device=dev)).abs().max().item()
    resid = (T @ Z - Z * L).abs().max().item() / T.abs().max().item()
    assert orth <= 1e-12 and resid <= 1e-12, (orth, resid)
