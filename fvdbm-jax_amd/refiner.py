"""MeshRefiner: mesh-quality metrics and in-place improvement of 2D triangle
meshes (reference: src/refiner.py:18-748).

The reference works on a meshpy ``MeshInfo`` and, after each operation,
re-triangulates the whole point set with ``meshpy.triangle.build``
(constrained Delaunay with a 30 degree minimum angle). meshpy is not
available here, so this class applies the same operations in place on the
connectivity and keeps the mesh valid itself: no operation may invert a
triangle, change the covered area, touch a boundary node or a boundary
segment (facet), or change the node markers. What needs a constrained
triangulator (the global re-meshing of ``improve`` and ``coarsen_mesh``)
is not built.

The mesh is any object with ``points`` [N,2], ``elements`` [C,3] and
``facets`` [B,2] (the boundary segments), optionally ``point_markers`` and
``facet_markers`` (meshpy's attribute names); ``faces`` (every edge) is
kept up to date so the result can go to ``Mesher.import_meshpy``.
``from_raw`` wraps a ``meshgen.RawMesh``.

Differences from the reference, by design:
  * ``laplacian_smooth`` takes each interior node's neighbours from the
    mesh edges (the reference lists them from the boundary facets only, so
    there its interior nodes, which lie on no facet, never move) and
    reverts a move that would invert a triangle;
  * ``flip_edges`` flips only convex quadrilaterals and at most one edge
    per triangle per pass (the reference's flips are undone by its
    re-triangulation, which restores a Delaunay mesh);
  * ``patch_remesh`` keeps a Delaunay re-triangulation of the patch only
    when it exactly tiles the removed patch.
"""
from __future__ import annotations

import time
from types import SimpleNamespace

import numpy as np


def _edges_of(elements):
    """[3C,2] edges (sorted) of each triangle, (v0v1, v1v2, v2v0) per cell."""
    e = np.stack([elements[:, [0, 1]], elements[:, [1, 2]], elements[:, [2, 0]]], axis=1).reshape(-1, 2)
    return np.sort(e, axis=1)


def _unique_edges(elements):
    e = _edges_of(elements)
    u, first = np.unique(e, axis=0, return_index=True)
    return e[np.sort(first)].astype(np.int32)


def _signed_areas(points, elements):
    p = points[elements]
    return 0.5 * ((p[:, 1, 0] - p[:, 0, 0]) * (p[:, 2, 1] - p[:, 0, 1])
                  - (p[:, 2, 0] - p[:, 0, 0]) * (p[:, 1, 1] - p[:, 0, 1]))


def _angles(cell_points):
    """Angles (degrees) at the vertices of triangles [..., 3, 2], the
    reference's construction (src/refiner.py:165-181)."""
    v = np.roll(cell_points, -1, axis=-2) - cell_points
    a = np.roll(v, 2, axis=-2)
    b = -v
    dot = np.sum(a * b, axis=-1)
    cos_t = dot / (np.linalg.norm(a, axis=-1) * np.linalg.norm(b, axis=-1))
    return np.arccos(np.clip(cos_t, -1, 1)) * 180 / np.pi


def min_angles(points, elements):
    """Smallest angle (degrees) of each triangle."""
    return _angles(np.asarray(points)[np.asarray(elements)]).min(axis=-1)


class MeshRefiner:
    """Mesh-quality analysis and improvement (src/refiner.py:18)."""

    def __init__(self, mesh, holes=None):
        self.holes = holes
        self.mesh = SimpleNamespace(
            points=np.array(mesh.points, np.float64),
            elements=np.array(mesh.elements, np.int32),
            facets=np.array(mesh.facets, np.int32).reshape(-1, 2),
            point_markers=(np.array(mesh.point_markers, np.int32)
                           if getattr(mesh, "point_markers", None) is not None else None),
            facet_markers=(np.array(mesh.facet_markers, np.int32)
                           if getattr(mesh, "facet_markers", None) is not None else None))
        self.point_markers = None if self.mesh.point_markers is None else self.mesh.point_markers.copy()
        self.facet_markers = None if self.mesh.facet_markers is None else self.mesh.facet_markers.copy()
        self._orient()
        self._refresh()

    @classmethod
    def from_raw(cls, raw, holes=None):
        """From a meshgen.RawMesh (its boundary segments are the edges used
        by one triangle)."""
        cells = np.asarray(raw.cells, np.int32)
        e = _edges_of(cells)
        u, cnt = np.unique(e, axis=0, return_counts=True)
        facets = u[cnt == 1]
        return cls(SimpleNamespace(points=raw.points, elements=cells, facets=facets,
                                   point_markers=raw.point_markers, facet_markers=None), holes)

    # --- bookkeeping ----------------------------------------------------------
    def _orient(self):
        m = self.mesh
        neg = _signed_areas(m.points, m.elements) < 0
        m.elements[neg] = m.elements[neg][:, [0, 2, 1]]

    def _refresh(self):
        """faces (every edge) and the boundary-node mask after a change."""
        m = self.mesh
        m.faces = _unique_edges(m.elements)
        self._boundary = np.zeros(len(m.points), bool)
        self._boundary[m.facets.reshape(-1)] = True

    def euler_characteristic(self):
        """V - E + F of the triangulation (1 - holes for a planar domain)."""
        m = self.mesh
        used = np.unique(m.elements)
        return int(len(used) - len(m.faces) + len(m.elements))

    # --- quality metrics (src/refiner.py:44-181) -------------------------------
    @staticmethod
    def _aspect_ratio_single(cell_points):
        """max edge / min edge of one triangle (src/refiner.py:44-60)."""
        edges = np.linalg.norm(cell_points - np.roll(cell_points, -1, axis=0), axis=1)
        return np.max(edges) / np.min(edges)

    @staticmethod
    def _skewness_single(cell_points):
        """Largest |angle - 60 deg| of one triangle (src/refiner.py:62-84)."""
        return float(np.max(np.abs(_angles(np.asarray(cell_points, np.float64)) - 60)))

    def compute_aspect_ratios(self, points, cells):
        """src/refiner.py:86-105."""
        cp = np.asarray(points)[np.asarray(cells)]
        edges = np.linalg.norm(cp - np.roll(cp, -1, axis=1), axis=2)
        return np.max(edges, axis=1) / np.min(edges, axis=1)

    def compute_skewness(self, points, cells):
        """src/refiner.py:107-131."""
        return np.max(np.abs(_angles(np.asarray(points)[np.asarray(cells)]) - 60), axis=1)

    def show_mesh_quality(self, aspect_thresh=3.0, skew_thresh=30.0, plot=True):
        """Histograms (matplotlib) and a summary of aspect ratio and skewness
        (src/refiner.py:133-163); also returns the summary."""
        m = self.mesh
        ar = self.compute_aspect_ratios(m.points, m.elements)
        sk = self.compute_skewness(m.points, m.elements)
        if plot:
            import matplotlib.pyplot as plt
            fig, axs = plt.subplots(1, 2, figsize=(12, 4))
            axs[0].hist(ar, bins=30, color="skyblue", edgecolor="k")
            axs[0].set_title("Aspect Ratio Histogram")
            axs[0].set_xlabel("Aspect Ratio")
            axs[0].set_ylabel("Count")
            axs[1].hist(sk, bins=30, color="salmon", edgecolor="k")
            axs[1].set_title("Skewness Histogram")
            axs[1].set_xlabel("Max Angle Deviation (deg)")
            axs[1].set_ylabel("Count")
            plt.tight_layout()
            plt.show()
        out = dict(bad_aspect=int(np.sum(ar > aspect_thresh)), bad_skew=int(np.sum(sk > skew_thresh)),
                   max_skewness=float(np.max(sk)), min_angle=float(min_angles(m.points, m.elements).min()))
        print(f"Number of cells with aspect ratio > {aspect_thresh}: {out['bad_aspect']}")
        print(f"Number of cells with skewness > {skew_thresh}°: {out['bad_skew']}")
        print(f"Maximum skewness: {out['max_skewness']:.2f}")
        return out

    def _triangle_angles(self, p):
        """(angles, min, max) of one triangle (src/refiner.py:165-181)."""
        a = _angles(np.asarray(p, np.float64))
        return a, float(np.min(a)), float(np.max(a))

    def _grading_ok(self, points, cell, grading_ratio=2.0):
        """Longest / shortest edge within grading_ratio (src/refiner.py:183-202)."""
        v = np.asarray(points)[np.asarray(cell)]
        e = np.linalg.norm(v - np.roll(v, -1, axis=0), axis=1)
        return np.max(e) / np.min(e) <= grading_ratio

    # --- operations ------------------------------------------------------------
    def flip_edges(self, verbose=False):
        """Flip an interior edge when the two triangles on its other diagonal
        have a larger smallest angle (src/refiner.py:205-281). Only convex
        quadrilaterals are flipped and each triangle takes part in at most
        one flip per call, so the smallest angle of the mesh never
        decreases. Returns the number of flips."""
        m = self.mesh
        P, T = m.points, m.elements
        e = _edges_of(T)
        key = e[:, 0].astype(np.int64) * len(P) + e[:, 1]
        order = np.argsort(key, kind="stable")
        ks = key[order]
        pair = np.nonzero(ks[1:] == ks[:-1])[0]
        c1, c2 = order[pair] // 3, order[pair + 1] // 3  # the two cells of each interior edge
        ed = e[order[pair]]
        # each interior edge in order of first appearance (the reference's dict order)
        first = np.minimum(order[pair], order[pair + 1])
        srt = np.argsort(first, kind="stable")
        c1, c2, ed = c1[srt], c2[srt], ed[srt]
        touched = np.zeros(len(T), bool)
        flipped = 0
        for k in range(len(ed)):
            a, b = int(c1[k]), int(c2[k])
            if touched[a] or touched[b]:
                continue
            e0, e1 = int(ed[k, 0]), int(ed[k, 1])
            o1 = int(T[a][(T[a] != e0) & (T[a] != e1)][0])
            o2 = int(T[b][(T[b] != e0) & (T[b] != e1)][0])
            old = min(self._triangle_angles(P[T[a]])[1], self._triangle_angles(P[T[b]])[1])
            n1 = np.array([o1, o2, e0])
            n2 = np.array([o1, o2, e1])
            ar1 = _signed_areas(P, n1[None])[0]
            ar2 = _signed_areas(P, n2[None])[0]
            if ar1 == 0 or ar2 == 0 or np.sign(ar1) == np.sign(ar2):
                continue  # e0 and e1 not on opposite sides of o1-o2: not convex
            new = min(self._triangle_angles(P[n1])[1], self._triangle_angles(P[n2])[1])
            if new > old + 1e-8:
                T[a] = n1 if ar1 > 0 else n1[[0, 2, 1]]
                T[b] = n2 if ar2 > 0 else n2[[0, 2, 1]]
                touched[a] = touched[b] = True
                flipped += 1
        if flipped:
            self._refresh()
        if verbose:
            print(f"  Edge flips performed: {flipped}")
        return flipped

    def _guard(self, old_points, new_points):
        """new_points with the moves reverted at every vertex of a triangle
        the moves would invert or flatten (repeated until none does)."""
        T = self.mesh.elements
        pts = np.array(new_points, np.float64)
        for _ in range(100):
            bad = _signed_areas(pts, T) <= 0
            if not bad.any():
                return pts
            v = np.unique(T[bad])
            pts[v] = old_points[v]
        return np.array(old_points, np.float64)

    def laplacian_smooth(self, points, facets, iterations=1):
        """Each interior node moves to the mean of its edge neighbours,
        boundary nodes (on a facet) stay (src/refiner.py:283-324; see the
        module notes). Returns the points as a list of tuples."""
        pts = np.array(points, np.float64)
        n = len(pts)
        e = _unique_edges(self.mesh.elements)
        boundary = np.zeros(n, bool)
        boundary[np.asarray(facets, np.int64).reshape(-1)] = True
        deg = np.bincount(e.reshape(-1), minlength=n).astype(np.float64)
        for _ in range(iterations):
            s = np.zeros_like(pts)
            np.add.at(s, e[:, 0], pts[e[:, 1]])
            np.add.at(s, e[:, 1], pts[e[:, 0]])
            moved = ~boundary & (deg > 0)
            new = pts.copy()
            new[moved] = s[moved] / deg[moved, None]
            pts = self._guard(pts, new)
        return [tuple(p) for p in pts]

    def angle_based_smooth(self, iterations=1):
        """Move each interior node to the mean of its neighbours when that
        raises the smallest angle of its triangles, node by node
        (src/refiner.py:326-384). The smallest angle of the mesh never
        decreases."""
        m = self.mesh
        P, T = m.points, m.elements
        order = np.argsort(T.reshape(-1), kind="stable")
        tri_of = order // 3
        start = np.searchsorted(T.reshape(-1)[order], np.arange(len(P) + 1))
        for _ in range(iterations):
            for i in range(len(P)):
                if self._boundary[i] or start[i] == start[i + 1]:
                    continue
                tris = T[tri_of[start[i]:start[i + 1]]]
                best = _angles(P[tris]).min()
                nb = np.unique(tris[tris != i])
                cand = P[nb].mean(axis=0)
                old = P[i].copy()
                P[i] = cand
                if np.all(_signed_areas(P, tris) > 0) and _angles(P[tris]).min() > best:
                    continue
                P[i] = old

    def selective_refine(self, points, facets, cells, aspect_thresh, skew_thresh, min_area, max_refine=10):
        """The points plus the centroids of the worst triangles (by
        skewness, then aspect ratio) of area at least min_area
        (src/refiner.py:386-433). Adding them to the mesh needs a
        re-triangulation, which is not built here."""
        pa = np.asarray(points, np.float64)
        ca = np.asarray(cells)
        area = np.abs(_signed_areas(pa, ca))
        skew = self.compute_skewness(pa, ca)
        aspect = self.compute_aspect_ratios(pa, ca)
        idx = np.nonzero(area >= min_area)[0]
        if idx.size == 0:
            return list(points)
        srt = np.lexsort((aspect[idx], skew[idx]))[::-1]
        worst = ca[idx][srt][:max_refine]
        cen = pa[worst].mean(axis=1)
        have = set(map(tuple, pa))
        return list(map(tuple, pa)) + [tuple(c) for c in cen if tuple(c) not in have]

    def patch_remesh(self, worst_cell_idx, patch_radius=1, verbose=True):
        """Replace the triangles within patch_radius edge-neighbour layers of
        worst_cell_idx by a Delaunay triangulation of their nodes
        (src/refiner.py:435-493), kept only when it exactly tiles the
        removed patch (same area, every patch-outline edge present, no new
        boundary node). Returns the number of triangles replaced (0 when the
        patch is kept)."""
        from scipy.spatial import Delaunay
        m = self.mesh
        P, T = m.points, m.elements
        e = _edges_of(T)
        key = e[:, 0].astype(np.int64) * len(P) + e[:, 1]
        patch = {int(worst_cell_idx)}
        for _ in range(patch_radius):
            keys = set(key.reshape(-1, 3)[list(patch)].reshape(-1).tolist())
            patch |= set((np.nonzero(np.isin(key, list(keys)))[0] // 3).tolist())
        patch = np.array(sorted(patch))
        pts_idx = np.unique(T[patch])
        # outline of the patch: its edges used by one patch triangle
        pk = key.reshape(-1, 3)[patch].reshape(-1)
        u, cnt = np.unique(pk, return_counts=True)
        outline = set(u[cnt == 1].tolist())
        replaced = 0
        if len(pts_idx) >= 3:
            try:
                tri = Delaunay(P[pts_idx])
            except Exception:
                tri = None
            if tri is not None:
                cand = pts_idx[tri.simplices].astype(np.int32)
                ar = _signed_areas(P, cand)
                cand[ar < 0] = cand[ar < 0][:, [0, 2, 1]]
                ar = np.abs(ar)
                # keep the triangles whose centroid lies inside the patch
                cen = P[cand].mean(axis=1)
                inside = np.zeros(len(cand), bool)
                for t in T[patch]:
                    a, b, c = P[t]
                    d = (b[1] - c[1]) * (a[0] - c[0]) + (c[0] - b[0]) * (a[1] - c[1])
                    l1 = ((b[1] - c[1]) * (cen[:, 0] - c[0]) + (c[0] - b[0]) * (cen[:, 1] - c[1])) / d
                    l2 = ((c[1] - a[1]) * (cen[:, 0] - c[0]) + (a[0] - c[0]) * (cen[:, 1] - c[1])) / d
                    inside |= (l1 > 0) & (l2 > 0) & (1 - l1 - l2 > 0)
                cand, ar = cand[inside & (ar > 0)], ar[inside & (ar > 0)]
                old_area = np.abs(_signed_areas(P, T[patch])).sum()
                ce = _edges_of(cand)
                ck = set((ce[:, 0].astype(np.int64) * len(P) + ce[:, 1]).tolist())
                if len(cand) and abs(ar.sum() - old_area) <= 1e-12 * max(old_area, 1e-300) * len(cand) \
                        and outline <= ck:
                    keep = np.ones(len(T), bool)
                    keep[patch] = False
                    m.elements = np.concatenate([T[keep], cand]).astype(np.int32)
                    self._refresh()
                    replaced = len(patch)
        if verbose:
            print(f"Patch remeshed: {replaced} triangles replaced.")
        return replaced

    def relocate_bad_points(self, points, cells, bad_cells, move_fraction=0.5):
        """Move the interior nodes of bad triangles toward their centroid
        (src/refiner.py:495-528); a move that would invert a triangle is
        reverted. Returns the points as a list of tuples."""
        old = np.array(points, np.float64)
        pts = old.copy()
        for cell in bad_cells:
            cell = np.asarray(cell)
            c = pts[cell].mean(axis=0)
            for idx in cell:
                if not self._boundary[idx]:
                    pts[idx] = (1 - move_fraction) * pts[idx] + move_fraction * c
        return [tuple(p) for p in self._guard(old, pts)]

    def coarsen_mesh(self, min_edge_length=0.05, max_iterations=5, verbose=True):
        """src/refiner.py:530-606 collapses short edges and re-triangulates
        with meshpy.triangle, which is not available here."""
        raise NotImplementedError("coarsen_mesh re-triangulates with meshpy.triangle, which is not available")

    # --- the improvement loop (src/refiner.py:608-748) --------------------------
    def improve(self, aspect_thresh=3.0, skew_thresh=30.0, min_area=1e-12, max_iter=10,
                smoothing=True, edge_flipping=True, angle_smoothing=True, selective_refinement=True,
                max_refine=10, patch_remesh=True, patch_radius=1, verbose=True,
                global_remesh_interval=5, grading_ratio=2.0, max_volume=None, move_fraction=0.5):
        """Iterate: find bad triangles (skewness above skew_thresh or aspect
        ratio above aspect_thresh, with grading within grading_ratio);
        remesh the worst patch when the worst skewness is twice the
        threshold, else move the bad triangles' nodes toward their
        centroids; then Laplacian smoothing, edge flips and angle-based
        smoothing, until no bad triangle is left or max_iter. The
        reference's re-triangulations (each iteration and every
        global_remesh_interval-th) and the centroid insertion of
        selective_refinement=False need a constrained triangulator and are
        left out. Returns the mesh."""
        m = self.mesh
        counts = []
        for it in range(max_iter):
            t0 = time.time()
            P, T = m.points, m.elements
            area = np.abs(_signed_areas(P, T))
            ok = area >= min_area
            skew = self.compute_skewness(P, T)
            aspect = self.compute_aspect_ratios(P, T)
            e = np.linalg.norm(P[T] - np.roll(P[T], -1, axis=1), axis=2)
            graded = e.max(axis=1) / e.min(axis=1) <= grading_ratio
            bad = ok & ((skew > skew_thresh) | (aspect > aspect_thresh)) & graded
            max_skew = float(skew[ok].max()) if ok.any() else 0.0
            counts.append(len(T))
            if not bad.any():
                if verbose:
                    print(f"Iteration {it}: 0 bad triangles, max skewness: {max_skew:.2f}, "
                          f"elements: {len(T)}, elapsed: {time.time() - t0:.3f}s")
                    print("Mesh quality acceptable.")
                break
            if patch_remesh and max_skew > skew_thresh * 2:
                worst = int(np.argmax(np.where(ok, skew, -np.inf)))
                self.patch_remesh(worst, patch_radius=patch_radius, verbose=verbose)
            elif selective_refinement:
                m.points = np.array(self.relocate_bad_points(m.points, m.elements, m.elements[bad],
                                                             move_fraction), np.float64)
            if smoothing:
                m.points = np.array(self.laplacian_smooth(m.points, m.facets, iterations=2), np.float64)
            if edge_flipping:
                self.flip_edges(verbose=verbose)
            if angle_smoothing:
                self.angle_based_smooth(iterations=1)
            if verbose:
                print(f"Iteration {it + 1}: {int(bad.sum())} bad triangles, max skewness: {max_skew:.2f}, "
                      f"elements: {len(m.elements)}, elapsed: {time.time() - t0:.3f}s")
        if verbose:
            print(f"Element counts per iteration: {counts}")
        return self.mesh
