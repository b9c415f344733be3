"""All-node interpolation and equilibrium initialisation on the GPU against
the oracle: cell moments (calc_macro), all-node IDW, node -> cell IDW and
calc_eq, bitwise; batched handles against single handles."""
import importlib

import numpy as np
import pytest

from cases import fv, ldc_case, porous_case
from oracle import interp, oracle

pytestmark = pytest.mark.gpu
batch = importlib.import_module("fvdbm-jax_amd.batch")


def _env(nsteps=25, precision="fp32"):
    m, dyn, cells, faces, nodes = ldc_case()
    env = fv.Environment(cells, faces, nodes, precision=precision)
    env.init()
    env.step(nsteps)
    return env, nodes


def test_cell_and_node_fields(gpu):
    env, nodes = _env()
    rho, vel = env.cell_fields()
    r0, v0 = oracle.calc_macro(env.cells.pdf)
    assert np.array_equal(rho, r0) and np.array_equal(vel, v0)
    nr, nv = env.node_fields()
    nc, nd = np.asarray(nodes.cells_index), np.asarray(nodes.cell_dists)
    assert np.array_equal(nr, interp.node_idw(nc, nd, rho))
    assert np.array_equal(nv, interp.node_idw(nc, nd, vel))
    env.close()


def test_init_equilibrium_is_calc_eq(gpu):
    env, nodes = _env(0)
    C, N = env.cells.size, env.nodes.size
    rng = np.random.default_rng(3)
    rho = rng.uniform(0.9, 1.1, C).astype(np.float32)
    vel = rng.uniform(-0.1, 0.1, (C, 2)).astype(np.float32)
    env.init_equilibrium(rho, vel)
    assert np.array_equal(env.cells.pdf, oracle.calc_eq(rho, vel))
    nrho = rng.uniform(0.9, 1.1, N).astype(np.float32)
    nvel = rng.uniform(-0.1, 0.1, (N, 2)).astype(np.float32)
    env.init_equilibrium_nodes(nrho, nvel)
    nc, nd = np.asarray(nodes.cells_index), np.asarray(nodes.cell_dists)
    cr = interp.cell_from_nodes(nc, nd, C, nrho)
    cv = interp.cell_from_nodes(nc, nd, C, nvel)
    assert np.array_equal(env.cells.pdf, oracle.calc_eq(cr, cv))
    # the rest state is W exactly, and stepping from it matches stepping from init
    env.init_equilibrium(np.ones(C, np.float32), np.zeros((C, 2), np.float32))
    assert np.array_equal(env.cells.pdf, oracle.uniform_w(C))
    env.close()


def test_fp64_fields(gpu):
    env, nodes = _env(10, "fp64")
    rho, vel = env.cell_fields()
    nr, nv = env.node_fields()
    nc, nd = np.asarray(nodes.cells_index), np.asarray(nodes.cell_dists_f64)
    assert np.array_equal(nr, interp.node_idw(nc, nd, rho, np.float64))
    C = env.cells.size
    env.init_equilibrium(rho, vel)
    assert np.array_equal(env.cells.pdf, oracle.calc_eq(rho, vel, "f64"))
    env.close()


def test_batch_fields_equal_single(gpu):
    cs = [porous_case(s, target_cells=4000) for s in range(3)]
    b = batch.BatchSolver([(c, f, n) for _, _, c, f, n in cs])
    b.step(40)
    nr, nv = b.node_fields()
    cr, cv = b.cell_fields()
    rng = np.random.default_rng(5)
    wr = rng.uniform(0.95, 1.05, nr.shape).astype(np.float32)
    wv = rng.uniform(-0.05, 0.05, nv.shape).astype(np.float32)
    b.init_equilibrium_nodes(wr, wv)
    warm = b.get_pdf()
    for k, (m, dyn, c, f, n) in enumerate(cs):
        env = fv.Environment(c, f, n)
        env.init()
        env.step(40)
        r1, v1 = env.node_fields()
        assert np.array_equal(b.case_nodes(k, nr), r1) and np.array_equal(b.case_nodes(k, nv), v1)
        r2, v2 = env.cell_fields()
        assert np.array_equal(b.case_cells(k, cr), r2) and np.array_equal(b.case_cells(k, cv), v2)
        env.init_equilibrium_nodes(b.case_nodes(k, wr), b.case_nodes(k, wv))
        assert np.array_equal(b.case_cells(k, warm), env.cells.pdf)
        env.close()
    b.close()
