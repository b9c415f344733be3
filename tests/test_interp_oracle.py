"""The interpolation oracle (oracle/interp.py), pinned on the CPU.

* cell -> node: on every node it must equal what the C oracle's node phase
  computes for boundary nodes (Nodes.calc_density / calc_velocity,
  src/containers.py:364-374) — checked bitwise by making every node a
  velocity node (rho interpolated) or a density node (u interpolated);
* constant fields come back as that constant (to rounding), a node with one
  cell returns that cell's value, and padding entries contribute nothing;
* node -> cell: constants reproduce, equidistant vertices give the plain
  mean, and a vertex closer than the others pulls the value towards itself
  (weights 1/d, not d).
"""
import numpy as np
import pytest

from cases import ldc_case, oracle_env, oracle_state
from oracle import interp, oracle


def _stepped_cavity(nsteps=25):
    m, dyn, cells, faces, nodes = ldc_case()
    env = oracle_env(cells, faces, nodes)
    st = oracle_state(env, nodes)
    oracle.run(env, st, dyn.tau, dyn.delta_t, nsteps)
    return m, dyn, env, st


@pytest.mark.parametrize("node_type", [1, 2])
def test_all_node_idw_equals_the_reference_node_phase(node_type):
    m, dyn, env, st = _stepped_cavity()
    N = env["node_cells"].shape[0]
    st["node_type"] = np.full(N, node_type, np.int32)
    st["node_vel"] = np.zeros((N, 2), np.float32)
    st["node_rho"] = np.ones(N, np.float32)
    oracle.run(env, st, dyn.tau, dyn.delta_t, 1)
    # after the step, cell_rho / cell_vel and the node fields describe its start
    if node_type == 1:
        want = interp.node_idw(env["node_cells"], env["node_dists"], st["cell_rho"])
        assert np.array_equal(want, st["node_rho"])
    else:
        want = interp.node_idw(env["node_cells"], env["node_dists"], st["cell_vel"])
        assert np.array_equal(want, st["node_vel"])


def test_idw_constant_single_cell_and_padding():
    m, dyn, env, st = _stepped_cavity(0)
    nc, nd = env["node_cells"], env["node_dists"]
    C = env["cell_faces"].shape[0]
    out = interp.node_idw(nc, nd, np.full(C, 1.25, np.float32))
    assert np.allclose(out, 1.25, rtol=3e-7, atol=0)
    # a node with one adjacent cell (cavity corners) takes that cell's value
    x = np.random.default_rng(0).uniform(0.5, 1.5, C).astype(np.float32)
    out = interp.node_idw(nc, nd, x)
    one = (nc >= 0).sum(axis=1) == 1
    assert one.any()
    assert np.allclose(out[one], x[nc[one, 0]], rtol=2e-7, atol=0)
    # padding entries (index -1, d = -1) contribute nothing: poisoning the
    # value a padding index would wrap to leaves the result unchanged
    y = x.copy()
    y[-1] = np.float32(1e30)
    has_pad = (nc < 0).any(axis=1) & ~(nc == C - 1).any(axis=1)
    assert np.array_equal(interp.node_idw(nc, nd, y)[has_pad], out[has_pad])


def test_node_to_cell_weights():
    # three nodes around one cell (no padding)
    nc = np.array([[0], [0], [0]], np.int32)
    eq = interp.cell_from_nodes(nc, np.full((3, 1), 0.5, np.float32), 1,
                                np.array([1.0, 2.0, 4.0], np.float32))
    assert eq[0] == pytest.approx(7.0 / 3.0, rel=1e-6)
    const = interp.cell_from_nodes(nc, np.array([[0.2], [0.7], [1.3]], np.float32), 1,
                                   np.full(3, 0.75, np.float32))
    assert const[0] == pytest.approx(0.75, rel=1e-6)
    near = interp.cell_from_nodes(nc, np.array([[0.01], [1.0], [1.0]], np.float32), 1,
                                  np.array([1.0, 0.0, 0.0], np.float32))
    assert near[0] > 0.97  # 1/d weighting: the near vertex dominates


def test_node_to_cell_constant_on_mesh():
    m, dyn, env, st = _stepped_cavity(0)
    C = env["cell_faces"].shape[0]
    N = env["node_cells"].shape[0]
    out = interp.cell_from_nodes(env["node_cells"], env["node_dists"], C,
                                 np.full((N, 2), [1.0, -0.5], np.float32))
    assert np.allclose(out, [1.0, -0.5], rtol=3e-7, atol=0)
