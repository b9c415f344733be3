"""Per-case scale calibration of the warm-start surrogate.

For each case of an unseen set the network predicts node (rho, u); the case
is solved cold to steady state, and the least-squares factors that best
map the predicted rho - 1 and u onto the converged node fields are its
ideal scales. A ridge regression of log(factor) on ``surrogate.case_features``
(standardised; the regularisation picked by k-fold cross-validation) is
fitted for each of the two, and its coefficients are stored in the weight
file as the ``calib_*`` arrays that ``MeshNet.predict`` applies.

    python -m fvdbm-jax_amd.calibrate_surrogate --cases 512 --seed0 6000000

meshes ``--cases`` cases of the steady-state distribution (seeds from
``--seed0``, disjoint from training and evaluation seeds) ``--group`` at a
time in a process pool forked before the GPU is touched, solves each group
cold in one batched handle (R < 1e-6, every 1000 steps), and writes the
weights with the calibration to ``--out`` (default: the input file).
"""
from __future__ import annotations

import argparse
import importlib
import json
import multiprocessing as mp
import os
import time

import numpy as np

ALPHAS = tuple(float(a) for a in np.logspace(-3, 3, 13))


def ideal_scales(rho_pred, vel_pred, rho_true, vel_true):
    """Least-squares factors (a_rho, a_u) with a_rho (rho_pred - 1) closest
    to rho_true - 1 and a_u u_pred closest to u_true (over the nodes)."""
    p = np.asarray(rho_pred, np.float64) - 1.0
    t = np.asarray(rho_true, np.float64) - 1.0
    q = np.asarray(vel_pred, np.float64).reshape(-1)
    v = np.asarray(vel_true, np.float64).reshape(-1)
    return float(p @ t / (p @ p)), float(q @ v / (q @ q))


def _ridge(z, y, alpha):
    w = np.linalg.solve(z.T @ z + alpha * np.eye(z.shape[1]), z.T @ (y - y.mean()))
    return w, float(y.mean())


def fit_ridge(x, y, alphas=ALPHAS, folds=5, seed=0):
    """Ridge regression of y on standardised x, the penalty chosen by
    ``folds``-fold cross-validation. Returns dict(mu, sd, w, b, alpha, r2)
    with r2 the cross-validated coefficient of determination."""
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    mu = x.mean(0)
    sd = np.where(x.std(0) > 0, x.std(0), 1.0)
    z = (x - mu) / sd
    order = np.random.default_rng(seed).permutation(len(y))
    parts = np.array_split(order, folds)
    best = None
    for a in alphas:
        pred = np.empty_like(y)
        for k in range(folds):
            test = parts[k]
            train = np.concatenate([parts[j] for j in range(folds) if j != k])
            zt = z[train]
            # standardise within the fold's training rows as the final fit does
            w, b = _ridge(zt - zt.mean(0), y[train], a)
            pred[test] = (z[test] - zt.mean(0)) @ w + b
        sse = float(((y - pred) ** 2).sum())
        if best is None or sse < best[1]:
            best = (a, sse)
    a = best[0]
    w, b = _ridge(z, y, a)
    ss = float(((y - y.mean()) ** 2).sum())
    r2 = 1.0 - best[1] / ss if ss > 0 else 0.0  # a constant target: nothing to explain
    return dict(mu=mu, sd=sd, w=w, b=b, alpha=a, r2=r2)


MIN_CASES = 20


def calibration_arrays(features, a_rho, a_u):
    """The ``calib_*`` arrays (and a record) from per-case features and
    ideal factors (at least MIN_CASES cases, every factor positive)."""
    a_rho = np.asarray(a_rho, np.float64)
    a_u = np.asarray(a_u, np.float64)
    if len(a_rho) < MIN_CASES:
        raise ValueError(f"calibration needs at least {MIN_CASES} converged cases, got {len(a_rho)}")
    if not ((a_rho > 0).all() and (a_u > 0).all() and np.isfinite(features).all()):
        raise ValueError("calibration needs positive ideal factors and finite features")
    fr = fit_ridge(features, np.log(a_rho))
    fu = fit_ridge(features, np.log(a_u))
    arrays = dict(calib_mu=fr["mu"], calib_sd=fr["sd"], calib_w_rho=fr["w"],
                  calib_b_rho=np.float64(fr["b"]), calib_w_u=fu["w"], calib_b_u=np.float64(fu["b"]))
    # the u model was standardised with its own (identical) mu and sd
    assert np.array_equal(fr["mu"], fu["mu"]) and np.array_equal(fr["sd"], fu["sd"])
    record = dict(cases=int(len(a_rho)), alpha_rho=fr["alpha"], alpha_u=fu["alpha"],
                  cv_r2_log_rho=fr["r2"], cv_r2_log_u=fu["r2"],
                  ideal_rho_median=float(np.median(a_rho)),
                  ideal_rho_p10_p90=[float(v) for v in np.percentile(a_rho, [10, 90])],
                  ideal_u_median=float(np.median(a_u)))
    return arrays, record


def collect(pool, seeds, weights, group=64, log=print):
    """Per-case features and ideal factors of the cases ``seeds`` (meshed by
    ``pool``, solved cold on the GPU): arrays [n, F], [n], [n] over the
    cases that converged."""
    seeds = list(seeds)
    chunks = [seeds[i:i + group] for i in range(0, len(seeds), group)]
    make = _make_case
    nxt = pool.map_async(make, chunks[0])
    batch = importlib.import_module("fvdbm-jax_amd.batch")
    ws = importlib.import_module("fvdbm-jax_amd.warmstart")
    sg = importlib.import_module("fvdbm-jax_amd.surrogate")
    net = sg.MeshNet.load(weights)
    feats, ar, au = [], [], []
    for ci in range(len(chunks)):
        cases = nxt.get()
        if ci + 1 < len(chunks):
            nxt = pool.map_async(make, chunks[ci + 1])
        b = batch.BatchSolver([c.env_tuple() for c in cases])
        try:
            st = ws.solve(b, None, tol=1e-6, check_every=1000, max_steps=800_000)
            nr, nv = b.node_fields()
        finally:
            b.close()
        pred = net.predict(cases, calibrate=False)
        for k, (c, (pr, pv)) in enumerate(zip(cases, pred)):
            if st["status"][k] != ws.CONVERGED:
                continue
            off0, off1 = b.node_off[k], b.node_off[k + 1]
            a_r, a_u = ideal_scales(pr, pv, nr[off0:off1], nv[off0:off1])
            feats.append(sg.case_features(c, pr, pv))
            ar.append(a_r)
            au.append(a_u)
        log(f"calibration: group {ci + 1}/{len(chunks)}, {len(ar)} cases")
    return np.array(feats), np.array(ar), np.array(au)


def _make_case(seed):
    porous = importlib.import_module("fvdbm-jax_amd.porous")
    return porous.make_case(seed, dist=porous.STEADY)


def main(argv=None):
    here = os.path.dirname(os.path.abspath(__file__))
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--weights", default=os.path.join(here, "data", "surrogate_porous.npz"))
    ap.add_argument("--out", default=None, help="output weights (default: --weights)")
    ap.add_argument("--cases", type=int, default=512)
    ap.add_argument("--seed0", type=int, default=6_000_000)
    ap.add_argument("--group", type=int, default=64)
    ap.add_argument("--workers", type=int, default=16)
    args = ap.parse_args(argv)
    t0 = time.time()

    def log(m):
        print(f"[calibrate] {m} ({time.time() - t0:.0f}s)", flush=True)

    with mp.get_context("fork").Pool(args.workers) as pool:  # before anything touches the GPU
        feats, ar, au = collect(pool, range(args.seed0, args.seed0 + args.cases), args.weights,
                                group=args.group, log=log)
    arrays, record = calibration_arrays(feats, ar, au)
    record.update(seed0=args.seed0, features=list(importlib.import_module(
        "fvdbm-jax_amd.surrogate").CALIB_FEATURES))
    z = np.load(args.weights)
    out = {k: z[k] for k in z.files}
    out.update(arrays)
    out["calib_record"] = np.array(json.dumps(record))
    path = args.out or args.weights
    np.savez(path, **out)
    log(f"wrote {path}: {json.dumps(record)}")


if __name__ == "__main__":
    main()
