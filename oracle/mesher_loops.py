"""Loop-by-loop restatement of the reference mesh geometry and ``to_env``.

TEST INFRASTRUCTURE ONLY: only tests/, ``__graft_entry__.smoke()`` and
bench.py's cpu_baseline leg use this module, as the checker of the product's
vectorised Mesher. Pure-Python loops, so use it on small meshes only.

Each function follows one routine of the reference ``src/mesher.py``
(MADALERAU/FVDBM-JAX) and keeps its loop order, because list order and slot
order are observable (node-cell lists, face slots):

  enforce_ccw                          src/mesher.py:63-78
  calc_cell_centers                    src/mesher.py:113-120
  enforce_boundary_face_normals_outward src/mesher.py:80-110
  calc_face_centers/normals/lengths    src/mesher.py:172-195
  calc_cell_face_indices_and_normals   src/mesher.py:122-158
  calc_cell_face_normal_signs          src/mesher.py:160-169
  calc_face_cell_indices               src/mesher.py:197-266
  calc_face_ghost_distances            src/mesher.py:268-283
  calc_point_cell_indices_and_distances src/mesher.py:286-316
  to_env ("upwind")                    src/mesher.py:610-645, 686-693
  calc_normal / calc_dist              utils/utils.py:159-173

``np.dot``/``np.linalg.norm`` on 2-vectors are written out as
``a0*b0 + a1*b1`` and ``sqrt(x0*x0 + x1*x1)``; the reference's BLAS calls may
differ from these by an ulp in FP64, which the FP32 rounding in ``to_env``
almost always absorbs.
"""
from __future__ import annotations

import math

import numpy as np


def _normal(p0, p1):
    """utils.calc_normal: left normal of p0->p1, unit length."""
    tx, ty = p1[0] - p0[0], p1[1] - p0[1]
    nx, ny = -ty, tx
    nrm = math.sqrt(nx * nx + ny * ny)
    return np.array([nx / nrm, ny / nrm])


def _dot(a, b):
    return a[0] * b[0] + a[1] * b[1]


def reference_geometry(points, cells, faces):
    """Return every static array of ``Mesher.calc_mesh_properties`` that the
    upwind path reads, computed with the reference's loops."""
    points = np.asarray(points, dtype=np.float64)
    cells = np.array(cells, dtype=np.int64, copy=True)
    faces = np.array(faces, dtype=np.int64, copy=True)

    # enforce_ccw
    for c in range(cells.shape[0]):
        t = cells[c]
        p = points[t]
        area = 0.5 * ((p[1, 0] - p[0, 0]) * (p[2, 1] - p[0, 1])
                      - (p[2, 0] - p[0, 0]) * (p[1, 1] - p[0, 1]))
        if area < 0:
            cells[c] = [t[0], t[2], t[1]]

    # calc_cell_centers
    centers = np.array([(points[a] + points[b] + points[c]) / 3.0
                        for a, b, c in cells])

    # enforce_boundary_face_normals_outward
    face_to_cell = {}
    for ci, cell in enumerate(cells):
        for i in range(3):
            key = tuple(sorted((int(cell[i]), int(cell[(i + 1) % 3]))))
            face_to_cell.setdefault(key, []).append(ci)
    for i in range(faces.shape[0]):
        face = faces[i]
        lst = face_to_cell.get(tuple(sorted((int(face[0]), int(face[1])))), [])
        if len(lst) == 1:
            p0, p1 = points[face[0]], points[face[1]]
            nrm = _normal(p0, p1)
            mid = (p0 + p1) / 2.0
            if _dot(nrm, centers[lst[0]] - mid) < 0:
                continue
            faces[i] = [face[1], face[0]]

    # face centers / normals / lengths
    face_normals = np.array([_normal(points[a], points[b]) for a, b in faces])
    face_lengths = []
    for a, b in faces:
        tx, ty = points[b][0] - points[a][0], points[b][1] - points[a][1]
        face_lengths.append(math.sqrt(tx * tx + ty * ty))
    face_lengths = np.array(face_lengths)

    # calc_cell_face_indices_and_normals
    key_to_index = {}
    for i, f in enumerate(faces):
        key_to_index[tuple(sorted((int(f[0]), int(f[1]))))] = i
    cell_face_indices = np.zeros((cells.shape[0], 3), dtype=np.int64)
    cell_face_normals = np.zeros((cells.shape[0], 3, 2))
    for ci, cell in enumerate(cells):
        edges = [(cell[0], cell[1]), (cell[1], cell[2]), (cell[2], cell[0])]
        for k, (a, b) in enumerate(edges):
            cell_face_indices[ci, k] = key_to_index[tuple(sorted((int(a), int(b))))]
            p0, p1 = points[a], points[b]
            nrm = _normal(p0, p1)
            mid = (p0 + p1) / 2.0
            cell_face_normals[ci, k] = nrm if _dot(nrm, centers[ci] - mid) < 0 else -nrm

    # calc_cell_face_normal_signs
    signs = np.zeros((cells.shape[0], 3), dtype=np.int32)
    for ci in range(cells.shape[0]):
        for k in range(3):
            d = _dot(cell_face_normals[ci, k], face_normals[cell_face_indices[ci, k]])
            signs[ci, k] = int(np.sign(d))

    # calc_face_cell_indices
    key_to_cells = {}
    for ci, cell in enumerate(cells):
        for i in range(3):
            key = tuple(sorted((int(cell[i]), int(cell[(i + 1) % 3]))))
            key_to_cells.setdefault(key, []).append(ci)
    face_cells = -np.ones((faces.shape[0], 2), dtype=np.int64)
    face_dists = np.zeros((faces.shape[0], 2))
    for i, f in enumerate(faces):
        lst = key_to_cells.get(tuple(sorted((int(f[0]), int(f[1])))), [])
        c0 = lst[0] if len(lst) > 0 else -1
        c1 = lst[1] if len(lst) > 1 else -1
        mid = (points[f[0]] + points[f[1]]) / 2.0
        n = face_normals[i]
        dists = np.array([-1.0, -1.0])
        if c0 == -1 or c1 == -1:
            c = c0 if c0 != -1 else c1
            d = _dot(n, centers[c] - mid)
            if c0 != -1:
                dists[0] = d
            else:
                dists[1] = d
            idx = [c0, c1]
        else:
            d0 = _dot(n, centers[c0] - mid)
            d1 = _dot(n, centers[c1] - mid)
            if d0 < d1:
                idx, dists = [c0, c1], np.array([d0, d1])
            else:
                idx, dists = [c1, c0], np.array([d1, d0])
        face_cells[i] = idx
        face_dists[i] = np.abs(dists)

    # calc_face_ghost_distances
    m0 = face_cells[:, 0] == -1
    m1 = face_cells[:, 1] == -1
    face_dists[m0, 0] = face_dists[m0, 1]
    face_dists[m1, 1] = face_dists[m1, 0]

    # calc_point_cell_indices_and_distances
    npts = points.shape[0]
    lists = [[] for _ in range(npts)]
    for ci, cell in enumerate(cells):
        for pt in cell:
            lists[pt].append(ci)
    kmax = max(len(lst) for lst in lists) if npts else 0
    node_cells = -np.ones((npts, kmax), dtype=np.int64)
    node_dists = -np.ones((npts, kmax))
    for i, lst in enumerate(lists):
        for j, ci in enumerate(lst):
            node_cells[i, j] = ci
            dx = points[i][0] - centers[ci][0]
            dy = points[i][1] - centers[ci][1]
            node_dists[i, j] = math.sqrt(dx * dx + dy * dy)

    return dict(cells=cells, faces=faces, cell_centers=centers,
                face_normals=face_normals, face_lengths=face_lengths,
                cell_face_indices=cell_face_indices,
                cell_face_normal_signs=signs, face_cell_indices=face_cells,
                face_cell_center_distances=face_dists,
                point_cell_indices=node_cells,
                point_cell_center_distances=node_dists)


def to_env_arrays(geom, dim_multiplier=1, precision="f32"):
    """The arrays ``Mesher.to_env(dynamics, "upwind", dim_multiplier)`` hands
    to the containers (src/mesher.py:632-645, 686-691).

    FP32 (the reference path, x64 disabled): every float array is rounded to
    f32 first and then multiplied by the f32-rounded multiplier. FP64 (the
    twin): the same products in f64.
    """
    ft = np.float32 if precision == "f32" else np.float64
    m = ft(dim_multiplier)
    node_d = geom["point_cell_center_distances"].astype(ft)
    pos = node_d > 0
    node_d = node_d.copy()
    node_d[pos] = node_d[pos] * m
    return dict(
        cell_faces=geom["cell_face_indices"].astype(np.int32),
        cell_signs=geom["cell_face_normal_signs"].astype(np.int32),
        face_cells=geom["face_cell_indices"].astype(np.int32),
        face_dists=(geom["face_cell_center_distances"].astype(ft) * m).astype(ft),
        face_nodes=geom["faces"].astype(np.int32),
        face_n=geom["face_normals"].astype(ft),
        face_L=(geom["face_lengths"].astype(ft) * m).astype(ft),
        node_cells=geom["point_cell_indices"].astype(np.int32),
        node_dists=node_d.astype(ft),
    )
