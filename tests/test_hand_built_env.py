"""Hand-built meshes (src/environment.py:20-36, utils/utils.py:176-229):
``Environment.create`` gives empty containers whose ragged index arrays are
``CustomArray``s filled row by row with ``add_items`` (-1 padded, a column
added whenever a row is full); after ``init()`` they are the dense arrays a
``Mesher.to_env`` run produces, and the environment steps the same.

The sign arrays (``face_normals``, values +-1) are assigned whole, as the
reference's own ``to_env`` does: with ``default_value=-1`` a stored -1 reads
as a free slot to ``add_item`` (a quirk of the reference class, kept)."""
import numpy as np
import pytest

from cases import fv, ldc_case, oracle_env, oracle_state
from oracle import oracle


def _hand_built(cells, faces, nodes, dyn):
    class Env(fv.Environment):
        dynamics = dyn

    C, F, N = cells.size, faces.size, nodes.size
    e = Env.create(C, F, N)
    fi = np.asarray(cells.face_indices)
    for c in range(C):
        e.cells.face_indices.add_items(c, fi[c])
    e.cells.face_normals = np.asarray(cells.face_normals).copy()
    sc, sd, ni = (np.asarray(faces.stencil_cells_index), np.asarray(faces.stencil_dists),
                  np.asarray(faces.nodes_index))
    for j in range(F):
        e.faces.stencil_cells_index.add_items(j, sc[j])
        e.faces.stencil_dists.add_items(j, sd[j])
        e.faces.nodes_index.add_items(j, ni[j])
    e.faces.n = np.asarray(faces.n).copy()
    e.faces.L = np.asarray(faces.L).copy()
    ci, cd = np.asarray(nodes.cells_index), np.asarray(nodes.cell_dists)
    for i in range(N):
        keep = ci[i] != -1
        e.nodes.cells_index.add_items(i, ci[i][keep])
        e.nodes.cell_dists.add_items(i, cd[i][keep])
    e.nodes.type = np.asarray(nodes.type).copy()
    e.nodes.rho = np.asarray(nodes.__dict__["_host_rho"]).copy()
    e.nodes.vel = np.asarray(nodes.__dict__["_host_vel"]).copy()
    return e


def _containers_init(e):
    e.cells.init()
    e.faces.init()
    e.nodes.init()


def test_custom_array_grows_and_pads():
    a = fv.CustomArray(3, dtype=np.int32)
    a.add_items(0, [5, 6, 7])
    a.add_item(2, 9)
    assert np.array_equal(np.asarray(a), [[5, 6, 7], [-1, -1, -1], [9, -1, -1]])


def test_hand_built_arrays_equal_to_env():
    m, dyn, cells, faces, nodes = ldc_case(n=8)
    e = _hand_built(cells, faces, nodes, dyn)
    _containers_init(e)
    for a, b in ((e.cells.face_indices, cells.face_indices), (e.faces.stencil_cells_index, faces.stencil_cells_index),
                 (e.faces.stencil_dists, faces.stencil_dists), (e.faces.nodes_index, faces.nodes_index),
                 (e.nodes.cells_index, nodes.cells_index), (e.nodes.cell_dists, nodes.cell_dists)):
        assert np.array_equal(np.asarray(a), np.asarray(b))
    # and the oracle steps both alike
    out = []
    for c, f, n in ((e.cells, e.faces, e.nodes), (cells, faces, nodes)):
        env = oracle_env(c, f, n)
        st = oracle_state(env, n)
        oracle.run(env, st, dyn.tau, dyn.delta_t, 20, "f32")
        out.append(st["cell_pdf"])
    assert np.array_equal(out[0], out[1])


@pytest.mark.gpu
def test_hand_built_env_steps_like_to_env(gpu):
    m, dyn, cells, faces, nodes = ldc_case(n=8)
    e = _hand_built(cells, faces, nodes, dyn)
    e.init()
    ref = fv.Environment(cells, faces, nodes)
    ref.init()
    e = e.step(37)
    ref = ref.step(37)
    assert np.array_equal(e.cells.pdf, ref.cells.pdf)
    assert np.array_equal(e.faces.pdf, ref.faces.pdf)
