"""Warm start on the GPU batch: the equilibrium of a converged field's node
interpolation restarts a case in a fraction of the cold steps and stops at a
field within the float32 fixed-point band of the cold one; corrupted
predictions fall back to the cold start, which then reproduces the cold
solve bitwise."""
import importlib

import numpy as np
import pytest

from cases import fv

pytestmark = pytest.mark.gpu
batch = importlib.import_module("fvdbm-jax_amd.batch")
ws = importlib.import_module("fvdbm-jax_amd.warmstart")
TOL, EVERY = 1e-6, 250


def _solver(cases):
    return batch.BatchSolver([c.env_tuple() for c in cases])


def test_warm_and_fallback(gpu):
    cases = [fv.porous.make_case(s, target_cells=3000) for s in (31, 32, 33)]
    b = _solver(cases)
    cold = ws.solve(b, None, tol=TOL, check_every=EVERY, max_steps=300_000)
    assert (cold["status"] == ws.CONVERGED).all(), cold
    pdf_cold = b.get_pdf()
    nr, nv = b.node_fields()
    # case 1 gets a NaN, case 2 an over-fast field: both start cold
    nr2, nv2 = nr.copy(), nv.copy()
    nv2[b.node_off[1] + 5] = np.nan
    nv2[b.node_off[2]:b.node_off[3]] *= 10.0
    warm = ws.solve(b, (nr2, nv2), tol=TOL, check_every=EVERY, max_steps=300_000)
    assert warm["fallback"].tolist() == [False, True, True]
    assert (warm["status"] == ws.CONVERGED).all()
    assert warm["steps"][0] * 3 <= cold["steps"][0]
    pdf_warm = b.get_pdf()
    assert ws.relative_l2(b.case_cells(0, pdf_warm), b.case_cells(0, pdf_cold)) < 1e-3
    for k in (1, 2):
        assert warm["steps"][k] == cold["steps"][k]
        assert np.array_equal(b.case_cells(k, pdf_warm), b.case_cells(k, pdf_cold))
    b.close()
