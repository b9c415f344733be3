"""The C-ABI library: it loads, exports every function include/fvdbm.h
declares, and its host-side layout (reverse Cuthill-McKee order, ghost-face
cells first) is a valid permutation. No compute call needs a GPU here."""
import ctypes as ct
import importlib
import os
import re

import numpy as np
import pytest

from cases import build_mesher, fv, ldc_case, meshgen

_lib = importlib.import_module("fvdbm-jax_amd._lib")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "fvdbm.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fvdbm_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_function():
    L = _lib.lib()
    names = header_functions()
    assert len(names) >= 15
    for n in names:
        assert hasattr(L, n), n
    assert sorted(_lib.declared_symbols()) == names
    assert L.fvdbm_abi_version() == 1


def test_device_count_never_fails():
    assert _lib.device_count() >= 0


def _layout(cells, faces, nodes, reorder):
    arrays = dict(cell_faces=cells.face_indices, cell_signs=cells.face_normals,
                  face_cells=faces.stencil_cells_index, face_dists=faces.stencil_dists,
                  face_nodes=faces.nodes_index, face_n=faces.n, face_L=faces.L.reshape(-1),
                  node_cells=nodes.cells_index, node_dists=nodes.cell_dists,
                  node_type=nodes.type.reshape(-1), node_rho=nodes.rho.reshape(-1),
                  node_vel=nodes.vel)
    desc = _lib.mesh_desc(arrays, np.float32)
    perm = np.empty(cells.face_indices.shape[0], np.int32)
    stats = np.empty(4, np.int32)
    rc = _lib.lib().fvdbm_layout_order(ct.byref(desc), _lib.FP32, reorder, _lib.ptr(perm), _lib.ptr(stats))
    return rc, perm, stats


def _bandwidth(perm, fc):
    inv = np.empty_like(perm)
    inv[perm] = np.arange(len(perm))
    inner = fc[:, 1] >= 0
    return int(np.max(np.abs(inv[fc[inner, 0]] - inv[fc[inner, 1]])))


def test_rcm_layout():
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import reverse_cuthill_mckee
    raw = meshgen.porous_channel(seed=2, target_cells=20_000)
    m = build_mesher(raw)
    cells, faces, nodes = m.to_env(fv.D2Q9(0.6, 0.01))
    fc = faces.stencil_cells_index
    C = cells.face_indices.shape[0]
    rc, ident, _ = _layout(cells, faces, nodes, 0)
    assert rc == 0 and np.array_equal(ident, np.arange(C))
    rc, rcm, stats = _layout(cells, faces, nodes, 2)
    assert rc == 0 and np.array_equal(np.sort(rcm), np.arange(C))
    inner = fc[:, 1] >= 0
    A = coo_matrix((np.ones(inner.sum()), (fc[inner, 0], fc[inner, 1])), shape=(C, C))
    sp = np.asarray(reverse_cuthill_mckee((A + A.T).tocsr(), symmetric_mode=True))
    bw_ident, bw_ours, bw_scipy = _bandwidth(ident, fc), _bandwidth(rcm, fc), _bandwidth(sp, fc)
    assert stats[3] == bw_ours
    assert bw_ours < bw_ident / 10
    assert bw_ours <= 1.1 * bw_scipy, (bw_ours, bw_scipy)
    # reorder 1: same RCM order, cells with a ghost face moved to the front
    rc, part, stats = _layout(cells, faces, nodes, 1)
    has_ghost = np.zeros(C, bool)
    has_ghost[fc[fc[:, 1] < 0, 0]] = True
    nb = int(has_ghost.sum())
    assert stats[0] == nb and stats[1] == int((fc[:, 1] < 0).sum())
    assert has_ghost[part[:nb]].all() and not has_ghost[part[nb:]].any()
    pos = np.empty(C, int)
    pos[rcm] = np.arange(C)
    assert np.all(np.diff(pos[part[:nb]]) > 0) and np.all(np.diff(pos[part[nb:]]) > 0)


def test_invalid_mesh_is_rejected():
    m, dyn, cells, faces, nodes = ldc_case(n=4)
    bad = cells.face_indices.copy()
    bad[0, 0] = 10_000
    cells.face_indices = bad
    rc, _, _ = _layout(cells, faces, nodes, 1)
    assert rc == _lib.ERR_INVALID
    m, dyn, cells, faces, nodes = ldc_case(n=4)
    sc = faces.stencil_cells_index.copy()
    sc[cells.face_indices[0, 0]] = [5, 6]  # face no longer holds cell 0
    faces.stencil_cells_index = sc
    rc, _, _ = _layout(cells, faces, nodes, 1)
    assert rc == _lib.ERR_INVALID


def test_missing_library_fails_loudly(monkeypatch):
    monkeypatch.setattr(_lib, "LIB_PATH", "/nonexistent/libfvdbm.so")
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(ImportError, match="no CPU fallback"):
        _lib.lib()


def test_dense_kernels_reject_misaligned_rows():
    """The dense surrogate kernels load four floats per lane: each entry
    point rejects a row array that is not 16-byte aligned before it launches
    anything (checked on the host, so no GPU is needed)."""
    L = _lib.lib()
    ok, bad = 0x10000, 0x10004  # never dereferenced: the check comes first
    rc = L.fvdbm_nn_mlp_forward(bad, 32, 64, ok, ok, ok, ok, 64, ok, None)
    assert rc != 0 and b"16-byte aligned" in L.fvdbm_nn_last_error()
    rc = L.fvdbm_nn_mlp_backward(ok, 32, 64, ok, ok, ok, 64, bad, None, ok, ok, None)
    assert rc != 0 and b"16-byte aligned" in L.fvdbm_nn_last_error()
    rc = L.fvdbm_nn_wgrad(ok, bad, 32, 64, 64, ok, ok, None)
    assert rc != 0 and b"16-byte aligned" in L.fvdbm_nn_last_error()
    rc = L.fvdbm_nn_wgrad_noalloc(bad, ok, 32, 64, 64, ok, None)
    assert rc != 0 and b"16-byte aligned" in L.fvdbm_nn_last_error()
