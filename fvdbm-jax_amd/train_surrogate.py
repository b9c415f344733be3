"""Train the warm-start surrogate on steady states the solver generates.

    python -m fvdbm-jax_amd.train_surrogate --cases 2000 --out weights.npz

1. meshes ``--cases`` porous cases of ``--dist`` (default ``steady``: 3-8
   discs at 20k-200k cells; seeds from ``--seed0``) in a process pool that
   is forked before the GPU is touched, one group of ``--group`` cases
   ahead of the solver, so meshing overlaps the solves and only two groups
   are ever held in host memory;
2. solves each group to a residual below ``--tol`` with the batched solver,
   from the uniform state or, with ``--init-weights``, from an earlier
   network's prediction (cold fallback per case; the converged states are
   steady states of the solver either way);
3. takes rho and u at every node (fvdbm_node_fields) of each converged case
   as the targets and keeps only the network's inputs and targets, on the
   GPU; the meshes are dropped;
4. trains MeshNet (Adam from ``--lr`` to ``--lr-final`` on a cosine,
   ``--batch`` meshes per step, ``--epochs``) on all but the last ``--val``
   cases and reports the relative L2 error of the prediction on those;
5. writes the float32 weights, the target scaling and the run's record.
"""
from __future__ import annotations

import argparse
import json
import multiprocessing as mp
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import porous, surrogate, warmstart
from .batch import BatchSolver

_DIST = porous.STUDY


def _build(seed):
    return porous.make_case(seed, dist=_DIST)


def build_cases(seeds, workers=16, dist=None):
    """porous.Case for each seed, meshed in a fork pool (call before any GPU use)."""
    global _DIST
    _DIST = dist or porous.STUDY
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
                  log=print, net=None):
    """Batched solves from the uniform state (or from ``net``'s prediction,
    cold fallback per case). Returns per case dict(rho [N], vel [N,2],
    status, steps) with node fields from fvdbm_node_fields."""
    res = [None] * len(cases)
    for gi, idx in enumerate(groups(cases, group_cells)):
        t0 = time.perf_counter()
        b = BatchSolver([cases[i].env_tuple() for i in idx])
        init = None
        if net is not None:
            pred = net.predict([cases[i] for i in idx])
            init = (np.concatenate([p[0] for p in pred]), np.concatenate([p[1] for p in pred]))
        out = warmstart.solve(b, init, tol=tol, check_every=check_every, max_steps=max_steps)
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
    ap.add_argument("--dist", default="steady", choices=sorted(porous.DISTRIBUTIONS))
    ap.add_argument("--seed0", type=int, default=0)
    ap.add_argument("--val", type=int, default=64)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--check-every", type=int, default=1000)
    ap.add_argument("--max-steps", type=int, default=600_000)
    ap.add_argument("--group", type=int, default=100, help="cases per batched solve")
    ap.add_argument("--init-weights", default=None,
                    help="warm-start the data solves from this network (cold fallback)")
    ap.add_argument("--epochs", type=int, default=50)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--lr-final", type=float, default=1e-5)
    ap.add_argument("--workers", type=int, default=16)
    ap.add_argument("--grid", default="32x16", help="virtual grid GXxGY (GY = 1: bands of x)")
    ap.add_argument("--no-resolution", action="store_true",
                    help="leave out the mesh-resolution input column")
    ap.add_argument("--data-seconds", type=float, default=None,
                    help="stop generating data after the group that passes this many seconds")
    ap.add_argument("--out", required=True)
    a = ap.parse_args(argv)

    def log(msg):
        print(msg, flush=True)

    global _DIST
    _DIST = porous.DISTRIBUTIONS[a.dist]
    grid = tuple(int(v) for v in a.grid.split("x"))
    resolution = not a.no_resolution
    seeds = list(range(a.seed0, a.seed0 + a.cases))
    chunks = [seeds[i:i + a.group] for i in range(0, len(seeds), a.group)]
    t0 = time.perf_counter()
    # the workers are forked here, before this process touches the GPU
    pool = mp.get_context("fork").Pool(a.workers)
    pending = pool.map_async(_build, chunks[0], chunksize=2)
    init_net = surrogate.MeshNet.load(a.init_weights) if a.init_weights else None
    samples, info = [], []
    attempted = 0
    t_solve = t_mesh_wait = 0.0
    threads = ThreadPoolExecutor(max_workers=a.workers)
    for gi in range(len(chunks)):
        tw = time.perf_counter()
        cases = pending.get()
        t_mesh_wait += time.perf_counter() - tw
        if gi + 1 < len(chunks):
            pending = pool.map_async(_build, chunks[gi + 1], chunksize=2)
        ts = time.perf_counter()
        fields = steady_states(cases, a.tol, a.check_every, a.max_steps, group_cells=1 << 62,
                               log=None, net=init_net)
        t_solve += time.perf_counter() - ts
        attempted += len(cases)
        ok = [i for i, f in enumerate(fields) if f["status"] == warmstart.CONVERGED]
        ins = list(threads.map(lambda i: surrogate.case_inputs(cases[i], grid, resolution), ok))
        for i, (x, g, nm) in zip(ok, ins):
            c, f = cases[i], fields[i]
            y = surrogate.targets(f["rho"], f["vel"], c.u_in, (1.0, 1.0, 1.0), g["n"] - nm)
            samples.append(surrogate.device_sample(x, g, y, nm))
            info.append(dict(seed=c.seed, cells=c.n_cells, steps=f["steps"]))
        log(f"group {gi}: {len(cases)} cases ({sum(c.n_cells for c in cases)} cells), {len(ok)} converged, "
            f"median steps {np.median([f['steps'] for f in fields]):.0f}, "
            f"{time.perf_counter() - t0:.1f} s elapsed")
        del cases, fields, ins
        if a.data_seconds and time.perf_counter() - t0 > a.data_seconds and gi + 1 < len(chunks):
            log(f"data time limit reached after {gi + 1} of {len(chunks)} groups")
            pending.wait()
            break
    pool.close()
    threads.shutdown()
    log(f"data: {len(samples)} steady states in {time.perf_counter() - t0:.1f} s "
        f"(solves {t_solve:.1f} s, waiting for meshes {t_mesh_wait:.1f} s)")
    torch = surrogate._torch()
    tr, va = samples[:-a.val], samples[-a.val:]
    # per-component scale of the targets over the training nodes
    with torch.no_grad():
        ys = torch.cat([s["y"][s["mask"][:, 0] > 0] for s in tr]).double()
        scale = ys.std(dim=0).cpu().numpy()
        sc = torch.tensor(scale, dtype=torch.float32, device=ys.device)
        for s in samples:
            s["y"] /= sc
    del ys
    log(f"target scale {scale}")
    net = surrogate.MeshNet(seed=0, target_scale=scale, grid=grid, resolution=resolution)
    t2 = time.perf_counter()
    hist = surrogate.train(net, tr, epochs=a.epochs, batch=a.batch, lr=a.lr, lr_final=a.lr_final, log=log)
    t_train = time.perf_counter() - t2
    err = surrogate.evaluate(net, va)
    log(f"trained in {t_train:.1f} s; validation relative L2 (rho-1, ux, uy): "
        f"median {np.median(err, axis=0)}, max {err.max(axis=0)}")
    state = net.state()
    record = dict(cases=attempted, dist=a.dist, seed0=a.seed0, converged=len(samples), train=len(tr),
                  val=len(va), tol=a.tol, epochs=a.epochs, batch=a.batch, lr=a.lr, lr_final=a.lr_final,
                  loss=hist, grid=list(grid), resolution=resolution,
                  init_weights=os.path.basename(a.init_weights) if a.init_weights else None,
                  cells_median=float(np.median([i["cells"] for i in info])),
                  val_rel_l2_median=np.median(err, axis=0).tolist(), val_rel_l2_max=err.max(axis=0).tolist(),
                  solve_s=t_solve, train_s=t_train,
                  solve_steps_median=float(np.median([i["steps"] for i in info])))
    state["record"] = np.array(json.dumps(record))
    np.savez(a.out, **state)
    log(json.dumps(record))
    return 0


if __name__ == "__main__":
    sys.exit(main())
