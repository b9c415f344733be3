"""Train the warm-start surrogate on steady states the solver generates.

    python -m fvdbm-jax_amd.train_surrogate --cases 2000 --out weights.npz

1. meshes ``--cases`` porous cases of the study distribution (porous.STUDY,
   seeds from ``--seed0``) in a process pool, before the GPU is touched;
2. solves them from the uniform state to a residual below ``--tol`` with
   the batched solver, in groups of at most ``--group-cells`` cells;
3. takes rho and u at every node (fvdbm_node_fields) of each converged case
   as the targets;
4. trains MeshNet (Adam from ``--lr`` to ``--lr-final`` on a cosine,
   ``--batch`` meshes per step, ``--epochs``)
   on all but the last ``--val`` cases and reports the relative L2 error of
   the prediction on those;
5. writes the float32 weights, the target scaling and the run's record.
"""
from __future__ import annotations

import argparse
import json
import multiprocessing as mp
import sys
import time

import numpy as np

from . import porous, surrogate, warmstart
from .batch import BatchSolver


def _build(seed):
    return porous.make_case(seed, dist=porous.STUDY)


def build_cases(seeds, workers=16):
    """porous.Case for each seed, meshed in a fork pool (call before any GPU use)."""
    seeds = list(seeds)
    if workers <= 1:
        return [_build(s) for s in seeds]
    with mp.get_context("fork").Pool(workers) as pool:
        return pool.map(_build, seeds, chunksize=4)


def groups(cases, max_cells):
    out, cur, n = [], [], 0
    for i, c in enumerate(cases):
        if cur and n + c.n_cells > max_cells:
            out.append(cur)
            cur, n = [], 0
        cur.append(i)
        n += c.n_cells
    if cur:
        out.append(cur)
    return out


def steady_states(cases, tol=1e-6, check_every=500, max_steps=400_000, group_cells=4_000_000,
                  log=print):
    """Cold-start batched solves. Returns per case dict(rho [N], vel [N,2],
    status, steps) with node fields from fvdbm_node_fields."""
    res = [None] * len(cases)
    for gi, idx in enumerate(groups(cases, group_cells)):
        t0 = time.perf_counter()
        b = BatchSolver([cases[i].env_tuple() for i in idx])
        out = warmstart.solve(b, None, tol=tol, check_every=check_every, max_steps=max_steps)
        nr, nv = b.node_fields()
        for j, i in enumerate(idx):
            res[i] = dict(rho=b.case_nodes(j, nr).copy(), vel=b.case_nodes(j, nv).copy(),
                          status=int(out["status"][j]), steps=int(out["steps"][j]))
        b.close()
        if log:
            conv = int((out["status"] == warmstart.CONVERGED).sum())
            log(f"group {gi}: {len(idx)} cases, {conv} converged, max steps "
                f"{int(out['steps'].max())}, {time.perf_counter() - t0:.1f} s")
    return res


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--cases", type=int, default=2000)
    ap.add_argument("--seed0", type=int, default=0)
    ap.add_argument("--val", type=int, default=64)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--check-every", type=int, default=500)
    ap.add_argument("--max-steps", type=int, default=400_000)
    ap.add_argument("--group-cells", type=int, default=4_000_000)
    ap.add_argument("--epochs", type=int, default=50)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--lr-final", type=float, default=1e-5)
    ap.add_argument("--workers", type=int, default=16)
    ap.add_argument("--grid", default="32x16", help="virtual grid GXxGY (GY = 1: bands of x)")
    ap.add_argument("--out", required=True)
    a = ap.parse_args(argv)

    def log(msg):
        print(msg, flush=True)

    t0 = time.perf_counter()
    cases = build_cases(range(a.seed0, a.seed0 + a.cases), a.workers)
    log(f"meshed {len(cases)} cases ({sum(c.n_cells for c in cases)} cells) in "
        f"{time.perf_counter() - t0:.1f} s")
    t1 = time.perf_counter()
    fields = steady_states(cases, a.tol, a.check_every, a.max_steps, a.group_cells, log)
    t_solve = time.perf_counter() - t1
    ok = [i for i, f in enumerate(fields) if f["status"] == warmstart.CONVERGED]
    log(f"solved in {t_solve:.1f} s: {len(ok)}/{len(cases)} converged")
    val_idx, train_idx = ok[-a.val:], ok[:-a.val]
    grid = tuple(int(v) for v in a.grid.split("x"))
    t_all = np.concatenate([surrogate.targets(fields[i]["rho"], fields[i]["vel"], cases[i].u_in, (1, 1, 1))
                            for i in train_idx])
    scale = t_all.astype(np.float64).std(axis=0)
    log(f"target scale {scale}")

    def samples(idx):
        out = []
        for i in idx:
            x, g, nm = surrogate.case_inputs(cases[i], grid=grid)
            y = surrogate.targets(fields[i]["rho"], fields[i]["vel"], cases[i].u_in, scale, g["n"] - nm)
            out.append(surrogate.device_sample(x, g, y, nm))
        return out

    tr, va = samples(train_idx), samples(val_idx)
    net = surrogate.MeshNet(seed=0, target_scale=scale, grid=grid)
    t2 = time.perf_counter()
    hist = surrogate.train(net, tr, epochs=a.epochs, batch=a.batch, lr=a.lr, lr_final=a.lr_final, log=log)
    t_train = time.perf_counter() - t2
    err = surrogate.evaluate(net, va)
    log(f"trained in {t_train:.1f} s; validation relative L2 (rho-1, ux, uy): "
        f"median {np.median(err, axis=0)}, max {err.max(axis=0)}")
    state = net.state()
    record = dict(cases=a.cases, seed0=a.seed0, converged=len(ok), train=len(train_idx), val=len(val_idx),
                  tol=a.tol, epochs=a.epochs, batch=a.batch, lr=a.lr, lr_final=a.lr_final, loss=hist,
                  grid=list(grid),
                  val_rel_l2_median=np.median(err, axis=0).tolist(), val_rel_l2_max=err.max(axis=0).tolist(),
                  solve_s=t_solve, train_s=t_train,
                  solve_steps_median=float(np.median([fields[i]["steps"] for i in ok])))
    state["record"] = np.array(json.dumps(record))
    np.savez(a.out, **state)
    log(json.dumps(record))
    return 0


if __name__ == "__main__":
    sys.exit(main())
