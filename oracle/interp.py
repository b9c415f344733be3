"""NumPy restatement of the interpolations around the step: all-node
inverse-distance weighting (cell -> node), its node -> cell counterpart, and
the equilibrium initialisation.

TEST INFRASTRUCTURE ONLY: the checker for fvdbm_node_fields,
fvdbm_set_equilibrium_nodes and fvdbm_set_equilibrium. Only tests/ may
import it.

The arithmetic follows utils.weighted_avg (utils/utils.py:34-45): weights
w = 1/d, set to 0 where negative (the -1 padding), then sum(x*w)/sum(w),
each sum accumulated from 0 in list order, in the working precision (every
operation rounded to float32 for "f32"). NumPy float32 elementwise
operations round each product and sum exactly as the device does with
contraction off, so results compare bitwise.
"""
from __future__ import annotations

import numpy as np


def idw_weights(d, dtype=np.float32):
    d = np.asarray(d, dtype)
    with np.errstate(divide="ignore"):
        w = (dtype(1) / d).astype(dtype)
    return np.where(w < 0, dtype(0), w).astype(dtype)


def node_idw(node_cells, node_dists, x, dtype=np.float32):
    """x: [C] or [C, m] cell values -> [N] or [N, m] node values, every node
    (utils/utils.py:34-45 applied to all nodes)."""
    nc = np.asarray(node_cells)
    w = idw_weights(node_dists, dtype)
    x = np.asarray(x, dtype)
    vec = x.ndim == 2
    X = x if vec else x[:, None]
    N, K = nc.shape
    ws = np.zeros(N, dtype)
    num = np.zeros((N, X.shape[1]), dtype)
    for k in range(K):
        ws = (ws + w[:, k]).astype(dtype)
        c = nc[:, k]
        ok = c >= 0
        term = np.zeros_like(num)
        term[ok] = (X[c[ok]] * w[ok, k][:, None]).astype(dtype)
        # a padding entry adds +0, the same as skipping it
        num = (num + term).astype(dtype)
    out = (num / ws[:, None]).astype(dtype)
    return out if vec else out[:, 0]


def cell_vertex_lists(node_cells, node_dists, n_cells):
    """For each cell, its (node, d) entries in ascending node id."""
    nc = np.asarray(node_cells)
    nd = np.asarray(node_dists)
    lists = [[] for _ in range(n_cells)]
    N, K = nc.shape
    for n in range(N):
        for k in range(K):
            c = nc[n, k]
            if c >= 0:
                lists[c].append((n, nd[n, k]))
    return lists


def cell_from_nodes(node_cells, node_dists, n_cells, x, dtype=np.float32):
    """x: [N, m] node values -> [C, m] cell values, weights 1/d over each
    cell's vertices in ascending node id."""
    x = np.asarray(x, dtype)
    X = x if x.ndim == 2 else x[:, None]
    lists = cell_vertex_lists(node_cells, node_dists, n_cells)
    deg = max(len(l) for l in lists)
    idx = np.zeros((n_cells, deg), np.int64)
    dd = np.ones((n_cells, deg), dtype)
    valid = np.zeros((n_cells, deg), bool)
    for c, l in enumerate(lists):
        for j, (n, d) in enumerate(l):
            idx[c, j], dd[c, j], valid[c, j] = n, d, True
    w = (dtype(1) / dd).astype(dtype)
    ws = np.zeros(n_cells, dtype)
    num = np.zeros((n_cells, X.shape[1]), dtype)
    for j in range(deg):
        v = valid[:, j]
        ws[v] = (ws[v] + w[v, j]).astype(dtype)
        num[v] = (num[v] + (X[idx[v, j]] * w[v, j][:, None]).astype(dtype)).astype(dtype)
    out = (num / ws[:, None]).astype(dtype)
    return out if x.ndim == 2 else out[:, 0]
