"""Mesher parity: the vectorised product Mesher against the loop restatement
of the reference geometry (oracle/mesher_loops.py), and the loop restatement
against closed forms on the structured cavity mesh."""
import math

import numpy as np
import pytest

from cases import build_mesher, fv, meshgen
from oracle import mesher_loops

INT_KEYS = ["cells", "faces", "cell_face_indices", "cell_face_normal_signs",
            "face_cell_indices", "point_cell_indices"]
FLOAT_KEYS = ["cell_centers", "face_normals", "face_lengths",
              "face_cell_center_distances", "point_cell_center_distances"]


def _small_meshes():
    yield "ldc8", meshgen.ldc_structured(8)
    yield "ldc32", meshgen.ldc_structured(32)
    yield "disc", meshgen.jittered_delaunay(1.0, 0.5, [(0.4, 0.25, 0.1)], 0.05, seed=3)
    yield "two_discs", meshgen.jittered_delaunay(1.0, 0.5, [(0.3, 0.2, 0.08), (0.7, 0.3, 0.1)],
                                                 0.04, seed=11)


def _scrambled(raw, seed):
    """Same mesh with random vertex order per cell (half of them clockwise)
    and random face node order: exercises enforce_ccw and the boundary flip."""
    rng = np.random.default_rng(seed)
    cells = raw.cells.copy()
    for i in range(len(cells)):
        cells[i] = cells[i][rng.permutation(3)]
    faces = raw.faces.copy()
    flip = rng.random(len(faces)) < 0.5
    faces[flip] = faces[flip][:, ::-1]
    perm = rng.permutation(len(faces))
    return meshgen.RawMesh(raw.points, cells, faces[perm], raw.point_markers, raw.meta)


@pytest.mark.parametrize("name,raw", list(_small_meshes()))
@pytest.mark.parametrize("scramble", [False, True])
def test_vectorised_matches_loops(name, raw, scramble):
    if scramble:
        raw = _scrambled(raw, 7)
    m = build_mesher(raw)
    cells = raw.cells.copy()
    # the loop restatement runs enforce_ccw itself
    ref = mesher_loops.reference_geometry(raw.points, cells, raw.faces)
    for k in INT_KEYS:
        got = {"cells": m.cells, "faces": m.faces}.get(k, getattr(m, k, None))
        assert np.array_equal(np.asarray(got), ref[k]), k
    for k in FLOAT_KEYS:
        got = getattr(m, k)
        np.testing.assert_array_equal(got, ref[k], err_msg=k)


def test_to_env_rounding_and_multiplier():
    raw = meshgen.jittered_delaunay(1.0, 0.5, [(0.4, 0.25, 0.1)], 0.05, seed=3)
    m = build_mesher(raw)
    dyn = fv.D2Q9(0.6, 0.01)
    mult = 3.7
    cells, faces, nodes = m.to_env(dyn, "upwind", dim_multiplier=mult)
    ref = mesher_loops.to_env_arrays(mesher_loops.reference_geometry(raw.points, raw.cells, raw.faces),
                                     mult, "f32")
    assert faces.L.dtype == np.float32 and faces.L.shape == (len(m.faces), 1)
    np.testing.assert_array_equal(faces.L.reshape(-1), ref["face_L"])
    np.testing.assert_array_equal(faces.stencil_dists, ref["face_dists"])
    np.testing.assert_array_equal(nodes.cell_dists, ref["node_dists"])
    np.testing.assert_array_equal(faces.n, ref["face_n"])
    # padding stays -1 (only positive distances are scaled, src/mesher.py:690)
    assert np.all(nodes.cell_dists[nodes.cells_index == -1] == -1)
    # f32(f32(L) * f32(m)), not f32(L * m)
    L64 = m.face_lengths
    assert np.array_equal(faces.L.reshape(-1),
                          (L64.astype(np.float32) * np.float32(mult)).astype(np.float32))
    assert nodes.type.shape == (len(m.points), 1) and not nodes.type.any()


def test_ldc_counts_and_closed_forms():
    n = 32
    m = build_mesher(meshgen.ldc_structured(n))
    h = 1.0 / n
    C, F, N = m.cells.shape[0], m.faces.shape[0], m.points.shape[0]
    assert (C, F, N) == (2048, 3136, 1089)
    assert m.point_cell_indices.shape == (1089, 6)
    # normals: +-x, +-y or diagonal; lengths h or sqrt(2) h
    nrm = m.face_normals
    ok_axis = np.isclose(np.abs(nrm).max(axis=1), 1.0)
    ok_diag = np.allclose(np.abs(nrm[~ok_axis]), 1 / math.sqrt(2))
    assert ok_diag
    L = m.face_lengths
    assert np.allclose(L[ok_axis], h) and np.allclose(L[~ok_axis], math.sqrt(2) * h)
    # slot 0 is behind the normal, slot 1 in front; boundary faces are [c, -1]
    fc = m.face_cell_indices
    mid = (m.points[m.faces[:, 0]] + m.points[m.faces[:, 1]]) / 2
    d0 = np.einsum("ij,ij->i", nrm, m.cell_centers[fc[:, 0]] - mid)
    assert np.all(d0 < 0)
    inner = fc[:, 1] >= 0
    d1 = np.einsum("ij,ij->i", nrm[inner], m.cell_centers[fc[inner, 1]] - mid[inner])
    assert np.all(d1 > 0)
    assert (~inner).sum() == 4 * n
    # boundary normals point out of the unit square
    out = mid[~inner] + 1e-3 * nrm[~inner]
    assert np.all(np.any((out < 0) | (out > 1), axis=1)), "boundary normal points inward"
    # sign +1 exactly where the cell is in slot 0
    cf, sg = m.cell_face_indices, m.cell_face_normal_signs
    slot0 = fc[cf, 0] == np.arange(C)[:, None]
    assert np.array_equal(sg == 1, slot0)
    # node-cell lists ascending, -1 padded at the end; IDW distances exact
    idx = m.point_cell_indices
    for i in range(N):
        row = idx[i][idx[i] >= 0]
        assert np.all(np.diff(row) > 0)
        assert np.all(idx[i][len(row):] == -1)
    valid = idx >= 0
    d = np.linalg.norm(m.points[np.nonzero(valid)[0]] - m.cell_centers[idx[valid]], axis=1)
    np.testing.assert_allclose(m.point_cell_center_distances[valid], d, rtol=1e-15)
    # ghost distances equal the known distance
    dist = m.face_cell_center_distances
    assert np.array_equal(dist[~inner, 1], dist[~inner, 0])
    # distance of a cell centre to an edge of a right triangle with legs h
    np.testing.assert_allclose(dist[ok_axis], h / 3, rtol=1e-12)


def test_ldc_markers_and_bcs():
    from cases import ldc_case
    m, dyn, cells, faces, nodes = ldc_case()
    t = nodes.type.reshape(-1)
    pts = m.points
    boundary = (pts[:, 0] == 0) | (pts[:, 0] == 1) | (pts[:, 1] == 0) | (pts[:, 1] == 1)
    assert np.array_equal(t != 0, boundary)
    lid = (pts[:, 1] == 1) & (pts[:, 0] > 0) & (pts[:, 0] < 1)
    assert np.all(nodes.vel[lid] == np.float32([0.1, 0.0]))
    assert np.all(nodes.vel[boundary & ~lid] == 0)
    assert lid.sum() == 31


def test_bc_setter_assertions():
    m = build_mesher(meshgen.ldc_structured(4))
    dyn = fv.D2Q9(0.6, 0.01)
    _, _, nodes = m.to_env(dyn)
    with pytest.raises(AssertionError):
        m.set_vel_node(nodes, 1, np.array([0.1, 0.0, 0.0]))
    with pytest.raises(AssertionError):
        m.set_rho_node(nodes, 1, 0.0)
    with pytest.raises(ValueError):
        m.to_env(dyn, "cc_stencil_nonsense")
    nodes = m.set_rho_node(nodes, meshgen.WALL, 1.05)
    assert np.all(nodes.type[m.point_markers == meshgen.WALL] == 2)
    nodes = m.set_vel_node(nodes, meshgen.WALL, np.array([0.0, 0.0]))  # last call wins
    assert np.all(nodes.type[m.point_markers == meshgen.WALL] == 1)


def test_velocity_profile_extension():
    raw = meshgen.jittered_delaunay(2.2, 0.41, [(0.2, 0.2, 0.05)], 0.03, seed=1)
    m = build_mesher(raw)
    _, _, nodes = m.to_env(fv.D2Q9(0.55, 1e-3))
    nodes = m.set_vel_profile(nodes, meshgen.INLET,
                              lambda p: meshgen.parabolic_inlet(p[:, 1], 0.41, 0.1))
    sel = m.point_markers == meshgen.INLET
    y = m.points[sel, 1]
    np.testing.assert_allclose(nodes.vel[sel, 0], 4 * 0.1 * y * (0.41 - y) / 0.41 ** 2, rtol=1e-6)
    assert np.all(nodes.type[sel] == 1)


def test_generated_meshes_are_valid():
    raw = meshgen.porous_channel(seed=5, target_cells=20_000)
    m = build_mesher(raw)
    fc = m.face_cell_indices
    # every face has a cell in slot 0; interior faces have two distinct cells
    assert np.all(fc[:, 0] >= 0)
    assert np.all(fc[fc[:, 1] >= 0, 0] != fc[fc[:, 1] >= 0, 1])
    # each cell references each face once; each interior face by exactly 2 cells
    cnt = np.bincount(m.cell_face_indices.reshape(-1), minlength=len(m.faces))
    assert np.array_equal(cnt, np.where(fc[:, 1] >= 0, 2, 1))
    # Euler characteristic: V - E + F = 1 - holes
    V, E, Fc = len(m.points), len(m.faces), len(m.cells)
    assert V - E + Fc == 1 - raw.meta["n_obstacles"]
    # boundary faces lie on the rectangle or on an obstacle circle
    bmid = ((m.points[m.faces[:, 0]] + m.points[m.faces[:, 1]]) / 2)[fc[:, 1] < 0]
    on_rect = (np.isclose(bmid[:, 0], 0) | np.isclose(bmid[:, 0], 1.0)
               | np.isclose(bmid[:, 1], 0) | np.isclose(bmid[:, 1], 0.5))
    on_disc = np.zeros(len(bmid), bool)
    for cx, cy, r in raw.meta["discs"]:
        on_disc |= np.abs(np.hypot(bmid[:, 0] - cx, bmid[:, 1] - cy) - r) < 0.05 * r
    assert np.all(on_rect | on_disc)
    st = m.verify_stencil_geometry(quiet=True)
    assert st["max_distance_error"] < 1e-9


def test_import_meshpy_equals_from_raw():
    """Mesher.import_meshpy (src/mesher.py:48-61) takes the attributes of a
    meshpy.triangle.MeshInfo (points, elements, faces, point_markers; meshpy
    itself is not needed) and orients the cells CCW: the result equals the
    from_raw path on the same mesh, also when half the cells come clockwise."""
    from types import SimpleNamespace
    raw = _scrambled(meshgen.jittered_delaunay(1.0, 0.5, [(0.4, 0.25, 0.1)], 0.05, seed=3), 7)
    info = SimpleNamespace(points=raw.points.tolist(), elements=raw.cells.tolist(),
                           faces=raw.faces.tolist(), point_markers=raw.point_markers.tolist())
    m = fv.Mesher()
    m.import_meshpy(info)
    assert m.points.dtype == np.float64 and m.cells.dtype == np.int32
    m.calc_mesh_properties()
    ref = build_mesher(raw)
    for key in INT_KEYS:
        assert np.array_equal(getattr(m, key), getattr(ref, key)), key
    for key in FLOAT_KEYS:
        assert np.array_equal(getattr(m, key), getattr(ref, key)), key
