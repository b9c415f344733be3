"""MeshRefiner (fvdbm-jax_amd/refiner.py; reference src/refiner.py) on small
meshes, against closed forms and brute-force checks: the quality metrics
against the one-triangle formulas and textbook triangles; every operation
keeps the mesh a valid triangulation of the same domain (positive areas,
same total area, Euler characteristic, boundary nodes, segments and
markers unchanged); edge flips and angle-based smoothing never lower the
smallest angle; the improved mesh goes through Mesher and the oracle."""
import importlib

import numpy as np
import pytest

from cases import build_mesher, fv, meshgen, oracle_env, oracle_state
from oracle import oracle

rf = importlib.import_module("fvdbm-jax_amd.refiner")


def _jittered(n=12, amp=0.35, seed=0):
    """A jittered structured mesh of the unit square (interior nodes moved
    by up to amp*h), boundary marker 1."""
    raw = meshgen.ldc_structured(n)
    rng = np.random.default_rng(seed)
    pts = raw.points.copy()
    inner = (pts[:, 0] > 0) & (pts[:, 0] < 1) & (pts[:, 1] > 0) & (pts[:, 1] < 1)
    pts[inner] += amp / n * rng.uniform(-1, 1, (inner.sum(), 2))
    markers = np.where(inner, 0, 1).astype(np.int32)
    a = rf._signed_areas(pts, raw.cells)
    assert (np.sign(a) == np.sign(rf._signed_areas(raw.points, raw.cells))).all()  # nothing inverted
    return meshgen.RawMesh(pts, raw.cells, raw.faces, markers)


def _valid(r, area0, euler0, pts0, facets0, markers0):
    m = r.mesh
    a = rf._signed_areas(m.points, m.elements)
    assert (a > 0).all()
    assert abs(a.sum() - area0) < 1e-12
    assert r.euler_characteristic() == euler0
    b = np.unique(facets0)
    assert np.array_equal(m.points[b], pts0[b])           # boundary nodes stay
    assert np.array_equal(m.facets, facets0)               # segments stay
    assert np.array_equal(m.point_markers, markers0)      # markers stay
    # every boundary segment is still an edge of exactly one triangle
    e = rf._edges_of(m.elements)
    u, cnt = np.unique(e, axis=0, return_counts=True)
    assert set(map(tuple, u[cnt == 1])) == set(map(tuple, np.sort(facets0, axis=1)))
    assert (cnt <= 2).all()


def _snapshot(r):
    m = r.mesh
    return (rf._signed_areas(m.points, m.elements).sum(), r.euler_characteristic(), m.points.copy(),
            m.facets.copy(), m.point_markers.copy())


def test_metrics_closed_forms():
    r = rf.MeshRefiner.from_raw(_jittered())
    eq = np.array([[0, 0], [1, 0], [0.5, np.sqrt(3) / 2]])
    rt = np.array([[0, 0], [1, 0], [0, 1]])
    assert r._aspect_ratio_single(eq) == pytest.approx(1.0)
    assert r._skewness_single(eq) == pytest.approx(0.0, abs=1e-6)
    assert r._aspect_ratio_single(rt) == pytest.approx(np.sqrt(2))
    assert r._skewness_single(rt) == pytest.approx(30.0)
    ang, lo, hi = r._triangle_angles(rt)
    assert sorted(ang) == pytest.approx([45, 45, 90]) and lo == pytest.approx(45) and hi == pytest.approx(90)
    assert r._grading_ok(rt, [0, 1, 2], 1.5) and not r._grading_ok(rt, [0, 1, 2], 1.4)
    # the vectorised forms against the one-triangle forms on the whole mesh
    m = r.mesh
    ar = r.compute_aspect_ratios(m.points, m.elements)
    sk = r.compute_skewness(m.points, m.elements)
    for k in range(0, len(m.elements), 7):
        p = m.points[m.elements[k]]
        assert ar[k] == pytest.approx(r._aspect_ratio_single(p))
        assert sk[k] == pytest.approx(r._skewness_single(p))
    out = r.show_mesh_quality(plot=False)
    assert out["bad_skew"] == int(np.sum(sk > 30)) and out["max_skewness"] == pytest.approx(sk.max())


def test_flip_fixes_a_bad_diagonal():
    """A quad split along its long diagonal: one flip, to the short one."""
    pts = np.array([[0.0, 0.0], [1.0, -0.2], [2.0, 0.0], [1.0, 0.2]])
    m = type("M", (), {})()
    m.points, m.elements = pts, np.array([[0, 1, 2], [0, 2, 3]])
    m.facets = np.array([[0, 1], [1, 2], [2, 3], [3, 0]])
    m.point_markers = np.ones(4, np.int32)
    r = rf.MeshRefiner(m)
    snap = _snapshot(r)
    before = rf.min_angles(r.mesh.points, r.mesh.elements).min()
    assert r.flip_edges() == 1
    assert set(map(tuple, r.mesh.faces)) >= {(1, 3)}
    assert rf.min_angles(r.mesh.points, r.mesh.elements).min() > before
    _valid(r, *snap)
    assert r.flip_edges() == 0  # now optimal


def test_flip_edges_and_angle_smoothing_never_lower_the_min_angle():
    r = rf.MeshRefiner.from_raw(_jittered(16, 0.3, 1))
    snap = _snapshot(r)
    lo = rf.min_angles(r.mesh.points, r.mesh.elements).min()
    assert r.flip_edges() > 0
    lo2 = rf.min_angles(r.mesh.points, r.mesh.elements).min()
    assert lo2 >= lo
    _valid(r, *snap)
    r.angle_based_smooth(iterations=2)
    lo3 = rf.min_angles(r.mesh.points, r.mesh.elements).min()
    assert lo3 >= lo2 and lo3 > lo
    _valid_moved(r, snap)


def _valid_moved(r, snap):
    """Validity after node moves: the same checks, boundary nodes fixed."""
    area0, euler0, pts0, facets0, markers0 = snap
    b = np.unique(facets0)
    assert np.array_equal(r.mesh.points[b], pts0[b])
    _valid(r, area0, euler0, np.where(np.isin(np.arange(len(pts0)), b)[:, None], pts0, r.mesh.points),
           facets0, markers0)


def test_laplacian_smooth_moves_interior_nodes_to_neighbour_means():
    r = rf.MeshRefiner.from_raw(_jittered(10, 0.3, 2))
    snap = _snapshot(r)
    m = r.mesh
    new = np.array(r.laplacian_smooth(m.points, m.facets, iterations=1))
    e = rf._unique_edges(m.elements)
    i = int(np.nonzero(~r._boundary)[0][5])
    nb = np.concatenate([e[e[:, 0] == i, 1], e[e[:, 1] == i, 0]])
    assert np.allclose(new[i], m.points[nb].mean(axis=0))
    m.points = new
    _valid_moved(r, snap)
    sk0 = np.mean(r.compute_skewness(snap[2], m.elements))
    assert np.mean(r.compute_skewness(m.points, m.elements)) < sk0  # smoother on average


def test_patch_remesh_keeps_a_valid_tiling():
    r = rf.MeshRefiner.from_raw(_jittered(12, 0.3, 3))
    snap = _snapshot(r)
    sk = r.compute_skewness(r.mesh.points, r.mesh.elements)
    replaced = 0
    for worst in np.argsort(sk)[::-1][:5]:
        replaced += r.patch_remesh(int(worst), patch_radius=1, verbose=False)
        _valid(r, *snap)
    assert replaced > 0
    assert r.compute_skewness(r.mesh.points, r.mesh.elements).max() <= sk.max() + 1e-9


def test_selective_refine_adds_worst_centroids():
    r = rf.MeshRefiner.from_raw(_jittered(8, 0.3, 4))
    m = r.mesh
    pts = r.selective_refine(m.points, m.facets, m.elements, 3.0, 30.0, 0.0, max_refine=4)
    assert len(pts) == len(m.points) + 4
    worst = np.argsort(r.compute_skewness(m.points, m.elements))[::-1][0]
    assert np.allclose(pts[len(m.points)], m.points[m.elements[worst]].mean(axis=0))


def test_improve_then_step_with_the_oracle():
    raw = _jittered(14, 0.3, 5)
    r = rf.MeshRefiner.from_raw(raw)
    snap = _snapshot(r)
    sk0 = r.compute_skewness(r.mesh.points, r.mesh.elements)
    mesh = r.improve(max_iter=6, verbose=False)
    _valid_moved(r, snap)
    sk1 = r.compute_skewness(mesh.points, mesh.elements)
    assert np.sum(sk1 > 30) < np.sum(sk0 > 30)
    with pytest.raises(NotImplementedError):
        r.coarsen_mesh()
    # the result is a mesh the solver takes
    m = fv.Mesher()
    m.import_meshpy(mesh)
    m.calc_mesh_properties()
    dyn = fv.D2Q9(0.6, 0.01)
    cells, faces, nodes = m.to_env(dyn)
    nodes = m.set_vel_node(nodes, 1, np.array([0.0, 0.0]))
    env = oracle_env(cells, faces, nodes, "f64")
    st = oracle_state(env, nodes, "f64")
    f0 = st["cell_pdf"].copy()
    oracle.run(env, st, dyn.tau, dyn.delta_t, 20, "f64")
    assert np.max(np.abs(st["cell_pdf"] - f0)) < 1e-13  # the rest state stays at rest
