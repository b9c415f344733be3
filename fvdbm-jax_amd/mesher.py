"""Vectorised Mesher with the reference's API (src/mesher.py, class Mesher).

Same attributes (``points``, ``cells``, ``faces``, ``point_markers`` and the
derived arrays), same methods (``enforce_ccw``, ``calc_mesh_properties``,
``verify_stencil_geometry``, ``to_env``, ``set_vel_node``, ``set_rho_node``,
``to_vtk``, ``to_pickle``/``from_pickle``), same orientation, slot order,
padding and list order; the O(C) Python loops of the reference become NumPy
array passes so that 1M-cell meshes take seconds. Integer outputs are equal
to the reference's; float outputs use the same formulas element by element.
"""
from __future__ import annotations

import time

import numpy as np

from .containers import Cells, Faces, Nodes


def _left_normals(p0, p1):
    """utils.calc_normal (utils/utils.py:162-173), row-wise."""
    tx = p1[:, 0] - p0[:, 0]
    ty = p1[:, 1] - p0[:, 1]
    nx, ny = -ty, tx
    nrm = np.sqrt(nx * nx + ny * ny)
    return np.stack([nx / nrm, ny / nrm], axis=1)


def _rowdot(a, b):
    return a[:, 0] * b[:, 0] + a[:, 1] * b[:, 1]


def _edge_table(cells, npts):
    """The (cell, local edge) list the reference builds its dicts from:
    edge i of cell c joins cell[i] and cell[(i+1)%3]; key = sorted pair."""
    a = cells[:, [0, 1, 2]].reshape(-1)
    b = cells[:, [1, 2, 0]].reshape(-1)
    lo, hi = np.minimum(a, b), np.maximum(a, b)
    key = lo.astype(np.int64) * npts + hi
    owner = np.repeat(np.arange(cells.shape[0]), 3)
    return key, owner


def _first_two_cells(face_key, edge_key, edge_owner):
    """For each face key, the first and second cell in (cell, local edge)
    order among the edges with that key (dict.setdefault(...).append order
    in src/mesher.py:207-220), -1 when absent; plus the count."""
    order = np.argsort(edge_key, kind="stable")
    sk = edge_key[order]
    so = edge_owner[order]
    lo = np.searchsorted(sk, face_key, side="left")
    hi = np.searchsorted(sk, face_key, side="right")
    cnt = hi - lo
    n = len(sk)
    c0 = np.where(cnt > 0, so[np.minimum(lo, n - 1)], -1)
    c1 = np.where(cnt > 1, so[np.minimum(lo + 1, n - 1)], -1)
    return c0, c1, cnt


class Mesher:
    """Constructs the static geometry of a 2D triangular mesh
    (reference: src/mesher.py:20-45)."""

    def _face_cells_raw(self):
        """(c0, c1, count) per face from the (cell, local edge) table; the
        sorted-pair key ignores face orientation, so one evaluation serves
        every pass of calc_mesh_properties."""
        cache = self.__dict__.get("_fc_cache")
        if cache is not None and cache[0] is self.cells and cache[1] == len(self.faces):
            return cache[2]
        npts = self.points.shape[0]
        f = self.faces
        fkey = np.minimum(f[:, 0], f[:, 1]).astype(np.int64) * npts + np.maximum(f[:, 0], f[:, 1])
        ekey, eown = _edge_table(self.cells, npts)
        res = _first_two_cells(fkey, ekey, eown)
        self.__dict__["_fc_cache"] = (self.cells, len(self.faces), res)
        return res

    def __init__(self):
        self.points = None
        self.cells = None
        self.neighbors = None
        self.faces = None
        self.point_markers = None
        self.verbose = False

    # --- import and orientation -------------------------------------------
    @classmethod
    def from_raw(cls, raw):
        """Mesher from a ``meshgen.RawMesh`` (points/cells/faces/markers)."""
        m = cls()
        m.points = np.asarray(raw.points, np.float64)
        m.cells = np.asarray(raw.cells, np.int32).copy()
        m.faces = np.asarray(raw.faces, np.int32).copy()
        m.point_markers = np.asarray(raw.point_markers, np.int32).copy()
        return m

    def import_meshpy(self, mesh):
        """src/mesher.py:48-61 (meshpy itself is not needed: any object with
        ``points``, ``elements``, ``faces`` and ``point_markers``)."""
        self.points = np.array(mesh.points, dtype=np.float64)
        self.cells = np.array(mesh.elements, dtype=np.int32)
        self.faces = np.array(mesh.faces, dtype=np.int32)
        self.point_markers = np.array(mesh.point_markers, dtype=np.int32)
        self.enforce_ccw()

    def enforce_ccw(self):
        """Swap to [t0,t2,t1] where the signed area is negative
        (src/mesher.py:63-78)."""
        p = self.points[self.cells]
        area = 0.5 * ((p[:, 1, 0] - p[:, 0, 0]) * (p[:, 2, 1] - p[:, 0, 1])
                      - (p[:, 2, 0] - p[:, 0, 0]) * (p[:, 1, 1] - p[:, 0, 1]))
        neg = area < 0
        cells = np.array(self.cells, copy=True)
        cells[neg] = cells[neg][:, [0, 2, 1]]
        self.cells = cells.astype(self.cells.dtype)

    # --- geometry -----------------------------------------------------------
    def calc_cell_centers(self):
        """(p1 + p2 + p3) / 3 (src/mesher.py:113-120)."""
        P = self.points
        c = self.cells
        self.cell_centers = (P[c[:, 0]] + P[c[:, 1]] + P[c[:, 2]]) / 3.0

    def enforce_boundary_face_normals_outward(self):
        """Flip boundary faces whose left normal points into their cell
        (src/mesher.py:80-110)."""
        f = self.faces
        c0, _, cnt = self._face_cells_raw()
        bnd = np.nonzero(cnt == 1)[0]
        p0 = self.points[f[bnd, 0]]
        p1 = self.points[f[bnd, 1]]
        nrm = _left_normals(p0, p1)
        mid = (p0 + p1) / 2.0
        to_center = self.cell_centers[c0[bnd]] - mid
        flip = ~(_rowdot(nrm, to_center) < 0)
        faces = np.array(self.faces, copy=True)
        idx = bnd[flip]
        faces[idx] = faces[idx][:, ::-1]
        self.faces = faces

    def calc_face_centers(self):
        f = self.faces
        self.face_centers = (self.points[f[:, 0]] + self.points[f[:, 1]]) / 2.0

    def calc_face_normals(self):
        f = self.faces
        self.face_normals = _left_normals(self.points[f[:, 0]], self.points[f[:, 1]])

    def calc_face_lengths(self):
        """calc_dist = norm(p2 - p1) (utils/utils.py:159-160)."""
        f = self.faces
        t = self.points[f[:, 1]] - self.points[f[:, 0]]
        self.face_lengths = np.sqrt(t[:, 0] * t[:, 0] + t[:, 1] * t[:, 1])

    def calc_cell_face_indices_and_normals(self):
        """Faces per cell in order (v0v1, v1v2, v2v0) and the outward normal of
        each (src/mesher.py:122-158). A duplicated face key maps to its last
        index, as the reference's dict comprehension does."""
        npts = self.points.shape[0]
        f = self.faces
        fkey = np.minimum(f[:, 0], f[:, 1]).astype(np.int64) * npts + np.maximum(f[:, 0], f[:, 1])
        # last index per key: sort by (key, index) and take each group's end
        order = np.lexsort((np.arange(len(fkey)), fkey))
        sk = fkey[order]
        ekey, _ = _edge_table(self.cells, npts)
        pos = np.searchsorted(sk, ekey, side="right") - 1
        if np.any(pos < 0) or np.any(sk[np.maximum(pos, 0)] != ekey):
            raise KeyError("a cell edge is missing from Mesher.faces")
        self.cell_face_indices = order[pos].reshape(-1, 3).astype(int)

        c = self.cells
        a = c[:, [0, 1, 2]].reshape(-1)
        b = c[:, [1, 2, 0]].reshape(-1)
        p0, p1 = self.points[a], self.points[b]
        nrm = _left_normals(p0, p1)
        mid = (p0 + p1) / 2.0
        to_center = np.repeat(self.cell_centers, 3, axis=0) - mid
        out = np.where((_rowdot(nrm, to_center) < 0)[:, None], nrm, -nrm)
        self.cell_face_normals = out.reshape(-1, 3, 2)

    def calc_cell_face_normal_signs(self):
        """sign(cell outward normal . face normal) (src/mesher.py:160-169)."""
        fn = self.face_normals[self.cell_face_indices.reshape(-1)]
        d = _rowdot(self.cell_face_normals.reshape(-1, 2), fn)
        self.cell_face_normal_signs = np.sign(d).reshape(-1, 3).astype(np.int32)

    def calc_stencil_norms(self):
        """Cell-to-cell (interior) or cell-to-face (boundary) unit vectors,
        flipped to align with the face normal (src/mesher.py:506-544). Only
        the cell-centred stencil variants read them."""
        c0, c1, cnt = self._face_cells_raw()
        v = np.zeros_like(self.face_normals)
        two = cnt == 2
        v[two] = self.cell_centers[c1[two]] - self.cell_centers[c0[two]]
        one = cnt == 1
        v[one] = self.face_centers[one] - self.cell_centers[c0[one]]
        nrm = np.sqrt(v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1])
        sn = np.where((nrm > 0)[:, None], v / np.where(nrm > 0, nrm, 1.0)[:, None], v)
        fnorm = self.face_normals / np.sqrt(_rowdot(self.face_normals, self.face_normals))[:, None]
        snu = sn / (np.sqrt(_rowdot(sn, sn)) + 1e-14)[:, None]
        flip = _rowdot(fnorm, snu) < 0
        sn[flip] *= -1
        self.stencil_norms = sn

    def calc_face_cell_indices(self):
        """Slot 0 = the cell behind the face normal (smaller n.(centroid -
        midpoint)); boundary faces are [cell, -1]; distances are |n.(x_c -
        x_f)| (src/mesher.py:197-266)."""
        f = self.faces
        c0, c1, _ = self._face_cells_raw()
        mid = (self.points[f[:, 0]] + self.points[f[:, 1]]) / 2.0
        n = self.face_normals
        cen = self.cell_centers
        d0 = _rowdot(n, cen[c0] - mid)      # c0 == -1 reads the last centre, as the reference
        d1 = _rowdot(n, cen[c1] - mid)
        one = (c0 == -1) | (c1 == -1)
        idx = np.stack([c0, c1], axis=1)
        dists = np.full((len(f), 2), -1.0)
        # boundary: the distance goes to the slot of the existing cell
        cb = np.where(c0 != -1, c0, c1)
        db = _rowdot(n, cen[cb] - mid)
        bm0 = one & (c0 != -1)
        bm1 = one & (c0 == -1)
        dists[bm0, 0] = db[bm0]
        dists[bm1, 1] = db[bm1]
        inner = ~one
        keep = inner & (d0 < d1)
        swap = inner & ~(d0 < d1)
        dists[keep] = np.stack([d0[keep], d1[keep]], 1)
        dists[swap] = np.stack([d1[swap], d0[swap]], 1)
        idx[swap] = idx[swap][:, ::-1]
        self.face_cell_indices = idx.astype(int)
        self.face_cell_center_distances = np.abs(dists)
        if hasattr(self, "stencil_norms"):
            sn = self.stencil_norms
            s0 = _rowdot(sn, cen[c0] - mid)
            s1 = _rowdot(sn, cen[c1] - mid)
            sb = _rowdot(sn, cen[cb] - mid)
            cc = np.full((len(f), 2), -1.0)
            cc[bm0, 0] = sb[bm0]
            cc[bm1, 1] = sb[bm1]
            cc[keep] = np.stack([s0[keep], s1[keep]], 1)
            cc[swap] = np.stack([s1[swap], s0[swap]], 1)
            self.cc_stencil_dist = np.abs(cc)

    def calc_face_ghost_distances(self):
        """Ghost distance := known distance (src/mesher.py:268-283)."""
        d = self.face_cell_center_distances.copy()
        idx = self.face_cell_indices
        m0 = idx[:, 0] == -1
        m1 = idx[:, 1] == -1
        d[m0, 0] = d[m0, 1]
        d[m1, 1] = d[m1, 0]
        self.face_cell_center_distances = d
        if hasattr(self, "cc_stencil_dist"):
            cc = self.cc_stencil_dist.copy()
            cc[m0, 0] = cc[m0, 1]
            cc[m1, 1] = cc[m1, 0]
            self.cc_stencil_dist = cc

    def calc_point_cell_indices_and_distances(self):
        """Per point: touching cells in ascending (cell, vertex) order, -1
        padded to the largest degree; Euclidean point-to-centroid distances
        (src/mesher.py:286-316)."""
        npts = self.points.shape[0]
        flat = self.cells.reshape(-1).astype(np.int64)
        owner = np.repeat(np.arange(self.cells.shape[0]), 3)
        order = np.argsort(flat, kind="stable")
        node = flat[order]
        cell = owner[order]
        deg = np.bincount(flat, minlength=npts)
        kmax = int(deg.max()) if npts else 0
        start = np.concatenate([[0], np.cumsum(deg)[:-1]])
        rank = np.arange(len(node)) - start[node]
        idx = -np.ones((npts, kmax), dtype=int)
        idx[node, rank] = cell
        dist = -np.ones((npts, kmax))
        dx = self.points[node, 0] - self.cell_centers[cell, 0]
        dy = self.points[node, 1] - self.cell_centers[cell, 1]
        dist[node, rank] = np.sqrt(dx * dx + dy * dy)
        self.point_cell_indices = idx
        self.point_cell_center_distances = dist

    def calc_face_stencil_angles(self):
        """Angle between face normal and stencil norm (src/mesher.py:546-558)."""
        fn = self.face_normals / np.sqrt(_rowdot(self.face_normals, self.face_normals))[:, None]
        sn = self.stencil_norms / np.sqrt(_rowdot(self.stencil_norms, self.stencil_norms))[:, None]
        self.face_stencil_angles = np.arccos(np.clip(_rowdot(fn, sn), -1.0, 1.0))

    def calc_mesh_properties(self):
        """Every derived array, in the reference's order (src/mesher.py:319-383)."""
        steps = [
            ("cell centers", self.calc_cell_centers),
            ("boundary face normals outward", self.enforce_boundary_face_normals_outward),
            ("face centers", self.calc_face_centers),
            ("face normals", self.calc_face_normals),
            ("face lengths", self.calc_face_lengths),
            ("cell face indices and normals", self.calc_cell_face_indices_and_normals),
            ("cell face normal signs", self.calc_cell_face_normal_signs),
            ("stencil norms", self.calc_stencil_norms),
            ("face cell indices", self.calc_face_cell_indices),
            ("face ghost distances", self.calc_face_ghost_distances),
            ("point cell indices and distances", self.calc_point_cell_indices_and_distances),
            ("face-stencil angles", self.calc_face_stencil_angles),
        ]
        for name, fn in steps:
            t0 = time.time()
            fn()
            if self.verbose:
                print(f"{name}: {time.time() - t0:.3f}s")

    def verify_stencil_geometry(self, quiet=False):
        """Face-normal alignment, projected-distance and face-centre offset
        statistics over interior faces (src/mesher.py:386-503); returns them
        as a dict (and prints them unless ``quiet``)."""
        idx = self.face_cell_indices
        inner = (idx[:, 0] != -1) & (idx[:, 1] != -1)
        c0, c1 = idx[inner, 0], idx[inner, 1]
        cen0, cen1 = self.cell_centers[c0], self.cell_centers[c1]
        n = self.face_normals[inner]
        n = n / np.sqrt(_rowdot(n, n))[:, None]
        v = cen1 - cen0
        vn = v / np.sqrt(_rowdot(v, v))[:, None]
        ang = np.degrees(np.arccos(np.clip(_rowdot(vn, n), -1, 1)))
        L = self.face_lengths[inner]
        d = self.face_cell_center_distances[inner]
        derr = np.abs(_rowdot(v, n) - (d[:, 0] + d[:, 1]))
        mid = (cen0 + cen1) / 2.0
        off = np.sqrt(_rowdot(self.face_centers[inner] - mid, self.face_centers[inner] - mid))
        stats = dict(
            mean_angle_deg=float(ang.mean()), max_angle_deg=float(ang.max()),
            n_bad_angle=int((np.abs(ang) > 30).sum()),
            max_distance_error=float(derr.max()),
            n_bad_distance=int((100.0 * derr / L > 1.0).sum()),
            max_offset_pct=float((100.0 * off / L).max()),
            n_bad_offset=int((100.0 * off / L > 10.0).sum()),
            n_interior_faces=int(inner.sum()),
            n_boundary_faces=int((~inner).sum()),
        )
        if not quiet:
            for k, val in stats.items():
                print(f"{k}: {val}")
        return stats

    # --- containers -----------------------------------------------------------
    def to_env(self, dynamics, flux_method="upwind", dim_multiplier=1):
        """(Cells, Faces, Nodes) for the solver (src/mesher.py:610-693).

        Float arrays are handed over as the reference's FP32 path sees them:
        rounded to f32, then multiplied by the f32-rounded ``dim_multiplier``.
        The FP64 values are kept beside them (``*_f64``) for the FP64 twin."""
        if flux_method not in ("upwind", "lax_wendroff"):
            raise ValueError(f"Unsupported flux method: {flux_method}")
        m32 = np.float32(dim_multiplier)
        m64 = np.float64(dim_multiplier)
        C, F, N = self.cells.shape[0], self.faces.shape[0], self.points.shape[0]

        cells = Cells(C, dynamics)
        cells.face_indices = np.asarray(self.cell_face_indices, np.int32)
        cells.face_normals = np.asarray(self.cell_face_normal_signs, np.int32)
        cells.centers = np.asarray(self.cell_centers, np.float64)  # for the device cell order

        faces = Faces(F, dynamics, flux_scheme=flux_method)
        faces.n = self.face_normals.astype(np.float32)
        faces.L = (self.face_lengths.astype(np.float32)[:, None] * m32).astype(np.float32)
        faces.nodes_index = np.asarray(self.faces, np.int32)
        faces.stencil_cells_index = np.asarray(self.face_cell_indices, np.int32)
        faces.stencil_dists = (self.face_cell_center_distances.astype(np.float32) * m32).astype(np.float32)
        faces.n_f64 = self.face_normals.astype(np.float64)
        faces.L_f64 = self.face_lengths.astype(np.float64)[:, None] * m64
        faces.stencil_dists_f64 = self.face_cell_center_distances.astype(np.float64) * m64

        nodes = Nodes(N, dynamics)
        nodes.cells_index = np.asarray(self.point_cell_indices, np.int32)
        d32 = self.point_cell_center_distances.astype(np.float32)
        pos = d32 > 0
        d32[pos] = d32[pos] * m32
        nodes.cell_dists = d32
        d64 = self.point_cell_center_distances.astype(np.float64).copy()
        pos64 = d64 > 0
        d64[pos64] = d64[pos64] * m64
        nodes.cell_dists_f64 = d64
        nodes.type = np.zeros((N, 1), np.int32)
        return cells, faces, nodes

    # --- boundary conditions -------------------------------------------------
    def set_vel_node(self, nodes, marker, velocity):
        """type <- 1 and vel <- velocity on nodes with ``marker``
        (src/mesher.py:697-719)."""
        velocity = np.asarray(velocity)
        assert velocity.shape == (nodes.dynamics.DIM,)
        sel = self.point_markers == marker
        nodes.type[sel] = 1
        nodes.vel[sel] = velocity.astype(np.float32)
        nodes.vel_f64[sel] = velocity.astype(np.float64)
        return nodes

    def set_vel_profile(self, nodes, marker, profile):
        """Per-node velocity boundary: ``profile`` is an array [k,2] for the k
        nodes with ``marker`` (in node order) or a callable points[k,2] ->
        velocities[k,2]. Extends set_vel_node, which takes one constant."""
        sel = np.nonzero(self.point_markers == marker)[0]
        vals = profile(self.points[sel]) if callable(profile) else profile
        vals = np.asarray(vals, dtype=np.float64)
        assert vals.shape == (len(sel), nodes.dynamics.DIM)
        nodes.type[sel] = 1
        nodes.vel[sel] = vals.astype(np.float32)
        nodes.vel_f64[sel] = vals
        return nodes

    def set_rho_node(self, nodes, marker, rho):
        """type <- 2 and rho <- rho on nodes with ``marker``
        (src/mesher.py:721-742)."""
        assert rho > 0
        sel = self.point_markers == marker
        nodes.type[sel] = 2
        nodes.rho[sel] = np.float32(rho)
        nodes.rho_f64[sel] = np.float64(rho)
        return nodes

    # --- export ----------------------------------------------------------------
    def to_vtk(self, env, filename, save_f=False, save_feq=False):
        """Legacy-ASCII VTK unstructured grid with cell data ``Velocity``
        (3 components, z = 0), ``Density`` and optionally ``pdf``/``feq``
        (the fields of src/mesher.py:562-598, written without pyvista)."""
        from .io import write_vtk
        return write_vtk(self, env, filename, save_f=save_f, save_feq=save_feq)

    def to_pickle(self, env, filename):
        """Checkpoint (env state + mesh) to ``filename.npz`` (replaces the
        reference's pickle, src/mesher.py:600-603)."""
        from .io import save_checkpoint
        return save_checkpoint(self, env, filename)

    @classmethod
    def from_pickle(cls, filename):
        from .io import load_checkpoint
        return load_checkpoint(filename)
