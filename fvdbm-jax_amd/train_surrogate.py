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
   ``--batch`` meshes per step, ``--epochs``; the squared errors of the
   three targets weighted by ``--loss-weights``) on all but the last
   ``--val`` cases and reports the relative L2 error of the prediction on
   those;
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


class _InProcess:
    """The part of a process pool's interface ``_solve_groups`` uses, meshing
    each group when it is asked for."""

    class _Result:
        def __init__(self, fn, items):
            self.fn, self.items = fn, items

        def get(self):
            return [self.fn(i) for i in self.items]

        def wait(self):
            pass

    def map_async(self, fn, items, chunksize=1):
        return self._Result(fn, items)

    def close(self):
        pass


def _solve_groups(seeds, dist, consume, group=100, tol=1e-6, check_every=1000, max_steps=600_000,
                  workers=16, init_weights=None, data_seconds=None, log=print):
    """Mesh the cases of ``seeds`` (distribution ``dist``) in a process pool,
    ``group`` at a time, one group ahead of the solver; solve each group to
    a residual below ``tol`` (warm-started from ``init_weights`` when given,
    cold fallback per case) and hand ``consume(cases, fields, converged
    indices)`` each group, then drop its meshes. The pool is forked here:
    call before the process touches the GPU; ``workers`` 0 meshes in this
    process instead (for profilers whose handlers a forked child inherits).
    Stops after the group that passes ``data_seconds``. Returns timing and
    count stats."""
    global _DIST
    _DIST = dist
    seeds = list(seeds)
    chunks = [seeds[i:i + group] for i in range(0, len(seeds), group)]
    t0 = time.perf_counter()
    if workers > 0:
        pool = mp.get_context("fork").Pool(workers)  # forked before any GPU use
    else:
        pool = _InProcess()
    pending = pool.map_async(_build, chunks[0], chunksize=2)
    init_net = surrogate.MeshNet.load(init_weights) if init_weights else None
    attempted = 0
    t_solve = t_mesh_wait = 0.0
    steps = []
    for gi in range(len(chunks)):
        tw = time.perf_counter()
        cases = pending.get()
        t_mesh_wait += time.perf_counter() - tw
        if gi + 1 < len(chunks):
            pending = pool.map_async(_build, chunks[gi + 1], chunksize=2)
        ts = time.perf_counter()
        fields = steady_states(cases, tol, check_every, max_steps, group_cells=1 << 62, log=None,
                               net=init_net)
        t_solve += time.perf_counter() - ts
        attempted += len(cases)
        ok = [i for i, f in enumerate(fields) if f["status"] == warmstart.CONVERGED]
        steps += [fields[i]["steps"] for i in ok]
        consume(cases, fields, ok)
        log(f"group {gi}: {len(cases)} cases ({sum(c.n_cells for c in cases)} cells), {len(ok)} converged, "
            f"median steps {np.median([f['steps'] for f in fields]):.0f}, "
            f"{time.perf_counter() - t0:.1f} s elapsed")
        del cases, fields
        if data_seconds and time.perf_counter() - t0 > data_seconds and gi + 1 < len(chunks):
            log(f"data time limit reached after {gi + 1} of {len(chunks)} groups")
            pending.wait()
            break
    pool.close()
    return dict(attempted=attempted, solve_s=t_solve, mesh_wait_s=t_mesh_wait, elapsed_s=time.perf_counter() - t0,
                steps_median=float(np.median(steps)) if steps else None)


def sample_builder(grids, resolution=True, workers=16):
    """A ``consume`` for ``_solve_groups`` that keeps, per virtual-grid
    size in ``grids``, each converged case's network inputs and unscaled
    targets on the GPU; returns (consume, samples {grid: list}, info list)."""
    samples = {g: [] for g in grids}
    info = []
    threads = ThreadPoolExecutor(max_workers=max(1, workers))

    def consume(cases, fields, ok):
        for grid in grids:
            ins = list(threads.map(lambda i: surrogate.case_inputs(cases[i], grid, resolution), ok))
            for i, (x, g, nm) in zip(ok, ins):
                c, f = cases[i], fields[i]
                y = surrogate.targets(f["rho"], f["vel"], c.u_in, (1.0, 1.0, 1.0), g["n"] - nm)
                samples[grid].append(surrogate.device_sample(x, g, y, nm))
        for i in ok:
            info.append(dict(seed=cases[i].seed, cells=cases[i].n_cells, steps=fields[i]["steps"]))

    return consume, samples, info


def fit(samples, val, grid, resolution, epochs, batch, lr, lr_final, weights=(1.0, 1.0, 1.0), log=print,
        layers=surrogate.LAYERS):
    """Scale the targets by their training-set standard deviation, train a
    MeshNet on all but the last ``val`` samples, evaluate on those. The
    samples keep their unscaled targets (the scaled ones are copies), so
    they can train another net. Returns (net, record dict)."""
    if not 1 <= val < len(samples):
        raise ValueError(f"val must leave training samples: val={val}, {len(samples)} samples")
    torch = surrogate._torch()
    with torch.no_grad():
        ys = torch.cat([s["y"][s["mask"][:, 0] > 0] for s in samples[:-val]]).double()
        scale = ys.std(dim=0).cpu().numpy()
        sc = torch.tensor(scale, dtype=torch.float32, device=ys.device)
        scaled = [dict(s, y=s["y"] / sc) for s in samples]
    del ys
    tr, va = scaled[:-val], scaled[-val:]
    log(f"target scale {scale}")
    net = surrogate.MeshNet(seed=0, target_scale=scale, grid=grid, resolution=resolution, layers=layers)
    t2 = time.perf_counter()
    hist = surrogate.train(net, tr, epochs=epochs, batch=batch, lr=lr, lr_final=lr_final, log=log,
                           weights=weights)
    t_train = time.perf_counter() - t2
    err = surrogate.evaluate(net, va)
    ntr, nva = len(tr), len(va)
    del scaled, tr, va
    log(f"trained in {t_train:.1f} s; validation relative L2 (rho-1, ux, uy): "
        f"median {np.median(err, axis=0)}, max {err.max(axis=0)}")
    record = dict(train=ntr, val=nva, epochs=epochs, batch=batch, lr=lr, lr_final=lr_final,
                  loss=hist, grid=list(grid), resolution=resolution, loss_weights=list(weights), layers=layers,
                  val_rel_l2_median=np.median(err, axis=0).tolist(), val_rel_l2_max=err.max(axis=0).tolist(),
                  train_s=t_train)
    return net, record


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
    ap.add_argument("--loss-weights", default="1,1,1",
                    help="weights of the (rho - 1, ux, uy) terms of the loss")
    ap.add_argument("--workers", type=int, default=16,
                    help="meshing processes (0: mesh in this process, for profiling)")
    ap.add_argument("--grid", default="32x16", help="virtual grid GXxGY (GY = 1: bands of x)")
    ap.add_argument("--layers", type=int, default=surrogate.LAYERS, help="message-passing layers")
    ap.add_argument("--no-resolution", action="store_true",
                    help="leave out the mesh-resolution input column")
    ap.add_argument("--data-seconds", type=float, default=None,
                    help="stop generating data after the group that passes this many seconds")
    ap.add_argument("--out", required=True)
    a = ap.parse_args(argv)

    def log(msg):
        print(msg, flush=True)

    grid = tuple(int(v) for v in a.grid.split("x"))
    resolution = not a.no_resolution
    consume, samples, info = sample_builder([grid], resolution, a.workers)
    stats = _solve_groups(range(a.seed0, a.seed0 + a.cases), porous.DISTRIBUTIONS[a.dist], consume,
                          group=a.group, tol=a.tol, check_every=a.check_every, max_steps=a.max_steps,
                          workers=a.workers, init_weights=a.init_weights, data_seconds=a.data_seconds,
                          log=log)
    log(f"data: {len(info)} steady states in {stats['elapsed_s']:.1f} s (solves {stats['solve_s']:.1f} s, "
        f"waiting for meshes {stats['mesh_wait_s']:.1f} s)")
    weights = tuple(float(v) for v in a.loss_weights.split(","))
    net, rec = fit(samples[grid], a.val, grid, resolution, a.epochs, a.batch, a.lr, a.lr_final, weights, log,
                   layers=a.layers)
    state = net.state()
    record = dict(cases=stats["attempted"], dist=a.dist, seed0=a.seed0, converged=len(info), tol=a.tol,
                  init_weights=os.path.basename(a.init_weights) if a.init_weights else None,
                  cells_median=float(np.median([i["cells"] for i in info])), solve_s=stats["solve_s"],
                  solve_steps_median=stats["steps_median"], **rec)
    state["record"] = np.array(json.dumps(record))
    np.savez(a.out, **state)
    log(json.dumps(record))
    return 0


if __name__ == "__main__":
    sys.exit(main())
