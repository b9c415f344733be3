"""Seeded mesh generators for the benchmark workloads.

The reference builds meshes with meshpy/Triangle (src/mesher.py:48-61), which
is absent here; these generators produce the same kind of input
(``points``, ``cells``, ``faces``, ``point_markers``) for ``Mesher``:

* ``ldc_structured``: the 32x32 lid-driven cavity, squares split along the
  (i,j)->(i+1,j+1) diagonal (2048 cells, 3136 faces, 1089 nodes).
* ``jittered_delaunay``: a rectangle with circular holes, meshed by Delaunay
  triangulation (scipy) of a jittered grid plus boundary points; used for
  the cylinder channel and the porous channels.

Marker convention (shared by all generators):
  0 interior, 1 no-slip wall, 2 density outlet, 3 moving lid,
  4 velocity inlet, 5 obstacle surface.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

INTERIOR, WALL, OUTLET, LID, INLET, OBSTACLE = 0, 1, 2, 3, 4, 5


@dataclass
class RawMesh:
    points: np.ndarray          # [N,2] float64
    cells: np.ndarray           # [C,3] int32
    faces: np.ndarray           # [F,2] int32
    point_markers: np.ndarray   # [N]   int32
    meta: dict = field(default_factory=dict)


def _unique_edges_in_order(cells):
    """Edges (min,max) in order of first appearance over (cell, local edge)."""
    e = np.stack([cells[:, [0, 1]], cells[:, [1, 2]], cells[:, [2, 0]]], axis=1).reshape(-1, 2)
    e = np.sort(e, axis=1)
    n = int(cells.max()) + 1
    key = e[:, 0].astype(np.int64) * n + e[:, 1]
    _, first = np.unique(key, return_index=True)
    return e[np.sort(first)].astype(np.int32)


def ldc_structured(n=32):
    """Unit-square cavity on an n x n grid, point id j*(n+1)+i at (i/n, j/n).

    Markers: top-row nodes strictly between the corners are LID; every other
    boundary node, the lid corners included, is WALL (the corner convention
    of the reference's cavity example, tests/ldcFVDBM.ipynb cell 8)."""
    ii, jj = np.meshgrid(np.arange(n + 1), np.arange(n + 1))
    pts = np.stack([ii.ravel() / n, jj.ravel() / n], axis=1).astype(np.float64)
    cells = []
    for j in range(n):
        for i in range(n):
            a = j * (n + 1) + i
            b, c, d = a + 1, a + (n + 1) + 1, a + (n + 1)
            cells.append([a, b, c])
            cells.append([a, c, d])
    cells = np.asarray(cells, dtype=np.int32)
    faces = _unique_edges_in_order(cells)
    i_, j_ = ii.ravel(), jj.ravel()
    markers = np.zeros(pts.shape[0], np.int32)
    boundary = (i_ == 0) | (i_ == n) | (j_ == 0) | (j_ == n)
    markers[boundary] = WALL
    markers[(j_ == n) & (i_ > 0) & (i_ < n)] = LID
    return RawMesh(pts, cells, faces, markers, dict(kind="ldc", n=n))


def jittered_delaunay(width, height, discs, h, seed, jitter=0.3):
    """Delaunay mesh of [0,width]x[0,height] minus the discs (cx, cy, r).

    Boundary points sit at spacing ~h on the rectangle and on each circle;
    interior points are a grid of spacing h jittered by U(-jitter*h,
    jitter*h) per coordinate (numpy default_rng(seed)). Interior points
    closer than 0.75 of the local boundary spacing to a circle, or than
    0.2h to the rectangle, are dropped, so every boundary chord is a
    Delaunay (Gabriel) edge and the triangles inside each disc consist of
    circle points only; those are removed by their centroid."""
    from scipy.spatial import Delaunay

    rng = np.random.default_rng(seed)
    nx = max(2, int(round(width / h)))
    ny = max(2, int(round(height / h)))
    xs = np.linspace(0.0, width, nx + 1)
    ys = np.linspace(0.0, height, ny + 1)
    bottom = np.stack([xs, np.zeros_like(xs)], 1)
    top = np.stack([xs, np.full_like(xs, height)], 1)
    left = np.stack([np.zeros(ny - 1), ys[1:-1]], 1)
    right = np.stack([np.full(ny - 1, width), ys[1:-1]], 1)
    rect = np.concatenate([bottom, top, left, right])
    rect_marker = np.concatenate([
        np.full(len(bottom), WALL), np.full(len(top), WALL),
        np.full(len(left), INLET), np.full(len(right), OUTLET)]).astype(np.int32)

    circ_pts, circ_spacing = [], []
    for (cx, cy, r) in discs:
        m = max(12, int(math.ceil(2 * math.pi * r / h)))
        th = 2 * math.pi * np.arange(m) / m
        circ_pts.append(np.stack([cx + r * np.cos(th), cy + r * np.sin(th)], 1))
        circ_spacing.append(2 * r * math.sin(math.pi / m))

    gx = np.arange(h / 2, width, h)
    gy = np.arange(h / 2, height, h)
    GX, GY = np.meshgrid(gx, gy)
    P = np.stack([GX.ravel(), GY.ravel()], 1)
    P = P + rng.uniform(-jitter * h, jitter * h, size=P.shape)
    keep = ((P[:, 0] > 0.2 * h) & (P[:, 0] < width - 0.2 * h)
            & (P[:, 1] > 0.2 * h) & (P[:, 1] < height - 0.2 * h))
    for (cx, cy, r), s in zip(discs, circ_spacing):
        d = np.hypot(P[:, 0] - cx, P[:, 1] - cy)
        keep &= d > r + 0.75 * max(s, h)
    P = P[keep]

    pts = np.concatenate([rect] + circ_pts + [P]).astype(np.float64)
    markers = np.concatenate(
        [rect_marker] + [np.full(len(c), OBSTACLE, np.int32) for c in circ_pts]
        + [np.zeros(len(P), np.int32)]).astype(np.int32)

    tri = Delaunay(pts).simplices.astype(np.int64)
    cen = pts[tri].mean(axis=1)
    inside = np.zeros(len(tri), bool)
    for (cx, cy, r) in discs:
        inside |= np.hypot(cen[:, 0] - cx, cen[:, 1] - cy) < r
    tri = tri[~inside]
    p = pts[tri]
    area = 0.5 * ((p[:, 1, 0] - p[:, 0, 0]) * (p[:, 2, 1] - p[:, 0, 1])
                  - (p[:, 2, 0] - p[:, 0, 0]) * (p[:, 1, 1] - p[:, 0, 1]))
    tri = tri[np.abs(area) > 1e-6 * h * h]
    used = np.zeros(len(pts), bool)
    used[tri.ravel()] = True
    if not used.all():
        remap = -np.ones(len(pts), np.int64)
        remap[used] = np.arange(used.sum())
        pts, markers, tri = pts[used], markers[used], remap[tri]
    cells = tri.astype(np.int32)
    faces = _unique_edges_in_order(cells)
    return RawMesh(pts, cells, faces, markers,
                   dict(kind="jittered_delaunay", width=width, height=height,
                        discs=[tuple(map(float, d)) for d in discs], h=h, seed=seed))


def cylinder_channel(target_cells=1_000_000, seed=0):
    """Channel [0,2.2]x[0,0.41] with a cylinder r=0.05 at (0.2,0.2); h is
    chosen so that the triangle count is about ``target_cells``."""
    area = 2.2 * 0.41 - math.pi * 0.05 ** 2
    h = math.sqrt(2.0 * area / target_cells)
    mesh = jittered_delaunay(2.2, 0.41, [(0.2, 0.2, 0.05)], h, seed)
    mesh.meta.update(kind="cylinder_channel", target_cells=target_cells)
    return mesh


def porous_channel(seed, target_cells=None, n_obstacles=None, r_range=(0.02, 0.08), gap=0.01,
                   count_range=(5, 40), cells_range=(20_000, 200_000)):
    """Porous channel [0,1]x[0,0.5] with non-overlapping discs, radius
    U[r_range], centres uniform over x in [0.1+r, 0.9-r] (rejection sampling
    with ``gap`` between discs and from the walls). The disc count defaults
    to U{count_range} and ``target_cells`` to U[cells_range]; the defaults
    are the 256-case batched workload (5-40 discs, r in [0.02,0.08],
    20k-200k cells)."""
    rng = np.random.default_rng(seed)
    if n_obstacles is None:
        n_obstacles = int(rng.integers(count_range[0], count_range[1] + 1))
    if target_cells is None:
        target_cells = int(rng.integers(cells_range[0], cells_range[1] + 1))
    discs = []
    tries = 0
    while len(discs) < n_obstacles and tries < 20000:
        tries += 1
        r = rng.uniform(*r_range)
        cx = rng.uniform(0.1 + r, 0.9 - r)
        cy = rng.uniform(r + gap, 0.5 - r - gap)
        if all(math.hypot(cx - x, cy - y) > r + s + gap for x, y, s in discs):
            discs.append((cx, cy, r))
    area = 0.5 - sum(math.pi * r * r for _, _, r in discs)
    h = math.sqrt(2.0 * area / target_cells)
    h = min(h, 0.5 * gap + 0.25 * min(r for _, _, r in discs))
    mesh = jittered_delaunay(1.0, 0.5, discs, h, int(rng.integers(2**31)))
    mesh.meta.update(kind="porous_channel", case_seed=seed,
                     target_cells=target_cells, n_obstacles=len(discs))
    return mesh


def parabolic_inlet(y, height, u_max):
    """u(y) = 4 u_max y (H - y) / H^2, v = 0."""
    u = 4.0 * u_max * y * (height - y) / (height * height)
    return np.stack([u, np.zeros_like(u)], axis=-1)
