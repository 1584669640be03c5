"""Trains the learned secular starting iterate on the GPU and writes the weights file.

Samples are harvested from the solver's own merges: tridiagonals of order n from the
generators --gens (1: d = 2 + U(-1e-3, 1e-3), e = 1 + U(-1e-3, 1e-3); 2: d, e ~
U(-1, 1)), taken in turn, seeds from --seed0 on, solved with the classical start;
every root of a merge with K >= --min-k gives (34 features, log(converged tau /
classical tau)).  Training is Adam on the device (hsdc_mlp_train); the held-out 5 %
gives the validation loss.  The result is evaluated on held-out matrices
(generator 1 seed 4 and generator 2 seed 2 by default: iterations per root, per-wave
and per-level maxima, solve time, with and without the proposal for each gate) and
written with its provenance:

    python -m hsspack_amd.train_mlp --out hsspack_amd/data/secular_mlp.bin
"""
import argparse
import json
import os
import time

import numpy as np

from hsspack_amd import hsdc
from hsspack_amd.workloads import tridiag_cfg

HERE = os.path.dirname(os.path.abspath(__file__))


def _solve(torch, ctx, d, e, zz):
    dev = zz.device
    dd = torch.tensor(d, device=dev)
    ee = torch.tensor(np.append(e, 0.0), device=dev)
    rc = ctx.stedc_dev(dd, ee, zz)
    torch.cuda.synchronize()
    if rc != 0:
        raise RuntimeError("stedc_dev: %d" % rc)
    return ctx.stats()


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=32768)
    ap.add_argument("--samples", type=int, default=1 << 20)
    ap.add_argument("--seed0", type=int, default=1000)
    ap.add_argument("--min-k", type=int, default=256)
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--gens", type=int, nargs="+", default=[1])
    ap.add_argument("--evals", type=lambda v: tuple(int(x) for x in v.split(":")), nargs="+",
                    default=[(1, 4), (2, 2)], help="generator:seed pairs")
    ap.add_argument("--gates", type=float, nargs="*", default=[0.0, 0.1, 0.25, 0.5, 1.0, 2.0, 4.0])
    ap.add_argument("--out", default=os.path.join(HERE, "data", "secular_mlp.bin"))
    a = ap.parse_args(argv)
    import torch
    if not torch.cuda.is_available():
        raise SystemExit("train_mlp needs a GPU")
    dev = torch.device("cuda:0")
    ctx = hsdc.Context(hss_min_k=1024, hss_leaf=128, hss_tol=1e-13, learned=0, learned_min_k=a.min_k)
    zz = torch.empty((a.n, a.n), dtype=torch.float64, device=dev)
    t0 = time.time()
    ctx.mlp_harvest(a.samples)
    seeds = []
    while ctx.mlp_count() < a.samples:
        seed = a.seed0 + len(seeds)
        gen = a.gens[len(seeds) % len(a.gens)]
        d, e = tridiag_cfg(gen, a.n, seed)
        _solve(torch, ctx, d, e, zz)
        seeds.append((gen, seed))
        print("harvest generator %d seed %d: %d samples" % (gen, seed, ctx.mlp_count()), flush=True)
    x, t = ctx.mlp_samples(a.samples)
    ctx.mlp_harvest(0)
    th = time.time() - t0
    rng = np.random.default_rng(a.seed0)
    perm = rng.permutation(len(t))
    nval = len(t) // 20
    xv, tv = x[perm[:nval]], t[perm[:nval]]
    xt, tt = x[perm[nval:]], t[perm[nval:]]
    ctx.mlp_set_weights(None)
    t1 = time.time()
    hist = ctx.mlp_train(xt, tt, a.epochs, a.batch, a.lr, seed=a.seed0)
    ttrain = time.time() - t1
    _, val = ctx.mlp_grad(xv, tv)
    val0 = float(np.mean(tv.astype(np.float64) ** 2))
    print("train loss %s -> %s, val %.4g (zero predictor %.4g), %.1f s" % (hist[0], hist[-1], val, val0, ttrain),
          flush=True)
    w = ctx.mlp_get_weights()
    hsdc.save_mlp_weights(a.out, w)
    # held-out evaluation: classical start and the proposal behind each gate (a
    # wavefront costs its slowest root's iterations, a level its slowest wavefront's)
    ev = []
    for gen, seed in a.evals:
        d, e = tridiag_cfg(gen, a.n, seed)
        ctx.set_params(learned=0)
        _solve(torch, ctx, d, e, zz)
        for learned, gate in [(0, 0.0)] + [(1, g) for g in a.gates] + [(0, 0.0)]:
            ctx.set_params(learned=learned, learned_gate=gate)
            s = _solve(torch, ctx, d, e, zz)
            r = dict(generator=gen, seed=seed, learned=learned, gate=gate, secular_iters=s["secular_iters"],
                     wave_iters=s["secular_wave_iters"], crit_iters=s["secular_crit_iters"], roots=s["roots"],
                     learned_roots=s["learned_roots"], ms_total=s["ms_total"])
            ev.append(r)
            print("eval gen %d seed %d learned=%d gate=%.2f: iters/root %.4f, wave iters %d, critical path %d, "
                  "learned roots %d, %.1f ms" % (gen, seed, learned, gate, r["secular_iters"] / r["roots"],
                                                 r["wave_iters"], r["crit_iters"], r["learned_roots"],
                                                 r["ms_total"]), flush=True)
    prov = dict(format="HSDCMLP1", n=a.n, harvest=seeds, min_k=a.min_k,
                samples=int(len(t)), validation=int(nval), epochs=a.epochs, batch=a.batch, lr=a.lr,
                loss_hist=[float(v) for v in hist], val_loss=float(val), val_loss_zero_predictor=val0,
                harvest_s=th, train_s=ttrain, eval=ev)
    with open(os.path.splitext(a.out)[0] + ".json", "w") as f:
        json.dump(prov, f, indent=1)
    ctx.close()


if __name__ == "__main__":
    main()
