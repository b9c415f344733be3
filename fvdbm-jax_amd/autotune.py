"""Autotuner of the step launch for a given mesh (or batch of meshes).

Candidates combine
  * the tile size: cells per thread block (one lane per cell, so this is
    also the wavefronts per tile: block / 64),
  * the kernel variant bits: 1 = no XCD-aware tile order, 2 = boundary
    nodes in their own kernel instead of the fused node phase, 4 = no LDS
    staging of the tile's cells and halo, 8 = one step per launch instead
    of two-step tiles,
  * the device cell order: 4 = Hilbert curve, 3 = Hilbert with boundary
    tiles first, 1 = reverse Cuthill-McKee.
Every candidate must give the same bits: after ``check_steps`` steps its
cell pdf is compared with the first candidate's and a mismatch raises. The
fastest by HIP-event timing of ``steps`` steps wins. Results are cached in
a JSON file keyed by a hash of the static arrays, the precision, scheme and
candidate list, and each decision is logged with every timing.
"""
from __future__ import annotations

import hashlib
import json
import os

import numpy as np

from . import _lib
from .environment import Environment

DEFAULT_CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_build", "autotune.json")
BLOCKS = (128, 256, 512)
VARIANTS = (0, 8, 1, 4)
REORDERS = (4, 3, 1)


def candidates(blocks=BLOCKS, variants=VARIANTS, reorders=REORDERS):
    return [dict(block=b, variant=v, reorder=r) for r in reorders for b in blocks for v in variants]


def _hash_containers(triples, extra):
    h = hashlib.sha256()
    for cells, faces, nodes in triples:
        for a in (cells.face_indices, cells.face_normals, faces.stencil_cells_index,
                  faces.stencil_dists, faces.nodes_index, faces.n, faces.L, nodes.cells_index,
                  nodes.cell_dists, nodes.type):
            h.update(np.ascontiguousarray(a).tobytes())
    h.update(json.dumps(extra, sort_keys=True).encode())
    return h.hexdigest()


def _load(path):
    if path and os.path.exists(path):
        with open(path) as fh:
            return json.load(fh)
    return {}


def _store(path, key, rec):
    if not path:
        return
    db = _load(path)
    db[key] = rec
    os.makedirs(os.path.dirname(path), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as fh:
        json.dump(db, fh, indent=1)
    os.replace(tmp, path)


def _run(make, cands, steps, check_steps, log):
    ref = None
    results = []
    for c in cands:
        obj = make(c)
        try:
            obj.step(check_steps)
            pdf = obj.get_pdf() if hasattr(obj, "get_pdf") else obj.cells.pdf
            if ref is None:
                ref = pdf
            elif not np.array_equal(pdf, ref):
                raise RuntimeError(f"autotune candidate {c} changes the result")
            obj.step(20)
            ms = _time(obj, steps)
        finally:
            obj.close()
        results.append(dict(c, us_per_step=round(1e3 * ms / steps, 3)))
        if log:
            log(f"autotune {c}: {1e3 * ms / steps:.2f} us/step")
    best = min(results, key=lambda r: r["us_per_step"])
    return best, results


def _time(obj, steps):
    if isinstance(obj, Environment):
        return obj.time_steps(steps)
    import ctypes as ct
    ms = ct.c_double()
    _lib.check(_lib.lib().fvdbm_step_timed(obj._handle(), int(steps), ct.byref(ms)), "fvdbm_step_timed")
    return float(ms.value)


def tune(cells, faces, nodes, precision="fp32", cands=None, steps=200, check_steps=20,
         cache=DEFAULT_CACHE, log=None):
    """Best launch options for one mesh. Returns dict(best, results, cached)."""
    cands = cands or candidates()
    key = _hash_containers([(cells, faces, nodes)], dict(precision=precision, scheme=faces.flux_scheme,
                                                         cands=cands, kind="single"))
    db = _load(cache)
    if key in db:
        if log:
            log(f"autotune (cached): {db[key]['best']}")
        return dict(db[key], cached=True)

    def make(c):
        env = Environment(cells, faces, nodes, precision=precision, **c)
        env.init()
        return env

    best, results = _run(make, cands, steps, check_steps, log)
    rec = dict(best=best, results=results)
    _store(cache, key, rec)
    if log:
        log(f"autotune: chose {best}")
    return dict(rec, cached=False)


def tuned_environment(cells, faces, nodes, precision="fp32", **kw):
    """An initialised Environment with the tuned launch options."""
    best = tune(cells, faces, nodes, precision, **kw)["best"]
    env = Environment(cells, faces, nodes, precision=precision, block=best["block"],
                      variant=best["variant"], reorder=best["reorder"])
    env.init()
    return env


def tune_batch(cases, precision="fp32", cands=None, steps=200, check_steps=20, cache=DEFAULT_CACHE,
               log=None):
    """Best launch options for a batch (list of (cells, faces, nodes))."""
    from .batch import BatchSolver
    cands = cands or candidates()
    key = _hash_containers(cases, dict(precision=precision, cands=cands, kind="batch"))
    db = _load(cache)
    if key in db:
        return dict(db[key], cached=True)

    def make(c):
        return BatchSolver(cases, precision=precision, **c)

    best, results = _run(make, cands, steps, check_steps, log)
    rec = dict(best=best, results=results)
    _store(cache, key, rec)
    return dict(rec, cached=False)


def confirm_batch(cases, results, precision="fp32", top=2, steps=100, log=None):
    """Time the best candidate of each of the ``top`` fastest tile sizes in
    ``results`` (``tune_batch`` on a sample of ``cases``) on the whole batch
    and keep the fastest. A two-step launch sizes its LDS classes from the
    largest tables, rings and ghost lists of all its tiles, so a sample can
    rank tile sizes differently from the full batch. Returns (best options,
    full-batch timings, the winning BatchSolver, left open and stepped)."""
    from .batch import BatchSolver
    finalists = []
    for r in sorted(results, key=lambda r: r["us_per_step"]):
        if all(f["block"] != r["block"] for f in finalists):
            finalists.append(dict(block=r["block"], variant=r["variant"], reorder=r["reorder"]))
        if len(finalists) == top:
            break
    if not finalists:
        raise ValueError("no tuning results to confirm")
    keep, timings = None, []
    for c in finalists:
        b = BatchSolver(cases, precision=precision, **c)
        try:
            b.step(20)
            us = 1e3 * _time(b, steps) / steps
        except Exception:
            b.close()
            raise
        timings.append(dict(c, us_per_step=round(us, 3)))
        if log:
            log(f"autotune on the full batch {c}: {us:.2f} us/step")
        if keep is None or us < keep[0]:
            if keep is not None:
                keep[2].close()
            keep = (us, c, b)
        else:
            b.close()
    return keep[1], timings, keep[2]
