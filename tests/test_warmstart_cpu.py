"""The warm-start controller on the CPU, driving the oracle through the
BatchSolver interface (tests/oracle_batch.py): a good warm start converges
in fewer steps to the same field as the cold start; corrupted warm starts
(NaN, velocity x10) trigger the cold fallback and the case still converges.
One porous case of about 2k cells, residual tolerance 1e-6."""
import importlib

import numpy as np
import pytest

from cases import fv
from oracle_batch import CONVERGED, OracleBatch

ws = importlib.import_module("fvdbm-jax_amd.warmstart")
TOL, EVERY = 1e-6, 250


@pytest.fixture(scope="module")
def cold():
    case = fv.porous.make_case(21, target_cells=1500)
    b = OracleBatch([case])
    res = ws.solve(b, None, tol=TOL, check_every=EVERY, max_steps=200_000)
    assert res["status"].tolist() == [CONVERGED]
    return case, res, b.get_pdf(), b.node_fields()


def test_warm_start_from_a_converged_field(cold):
    case, res, pdf, (nr, nv) = cold
    b = OracleBatch([case])
    warm = ws.solve(b, (nr, nv), tol=TOL, check_every=EVERY, max_steps=200_000)
    assert warm["status"].tolist() == [CONVERGED] and not warm["fallback"].any()
    assert warm["steps"][0] * 3 <= res["steps"][0]
    # both runs stop at float32 fixed points of the update (f + dt * (...)
    # rounds to f); for this case those lie within 1e-3 of each other
    assert ws.relative_l2(b.get_pdf(), pdf) < 1e-3


@pytest.mark.parametrize("corrupt,gate", [("nan", True), ("x10", True), ("nan", False)])
def test_corrupted_warm_start_falls_back(cold, corrupt, gate):
    """The plausibility gate sends a non-finite or over-fast prediction
    straight to the cold start; without the gate, the NaN run stops as
    diverged at its first check and restarts cold."""
    case, res, pdf, (nr, nv) = cold
    nr2, nv2 = nr.copy(), nv.copy()
    if corrupt == "nan":
        nv2[len(nv2) // 2] = np.nan
    else:
        nv2 *= 10.0
    assert not ws.plausible(nr2, nv2)
    b = OracleBatch([case])
    out = ws.solve(b, (nr2, nv2), tol=TOL, check_every=EVERY, max_steps=200_000, gate=gate)
    assert out["fallback"].tolist() == [True]
    assert out["status"].tolist() == [CONVERGED]
    # the cold restart reproduces the cold solve (bitwise when gated: the
    # same initial state and steps)
    if gate:
        assert out["steps"][0] == res["steps"][0]
        assert np.array_equal(b.get_pdf(), pdf)
    else:
        assert out["steps"][0] == res["steps"][0] + EVERY


def test_forced_residual_growth_falls_back(cold):
    """growth = 1e-12 makes the second check of any run count as growing
    (forced from the host): the warm case restarts cold after two intervals."""
    case, res, pdf, (nr, nv) = cold
    b = OracleBatch([case])
    out = ws.solve(b, (nr, nv), tol=TOL, check_every=EVERY, max_steps=200_000, growth=1e-12)
    assert out["fallback"].tolist() == [True]
    assert out["status"].tolist() == [CONVERGED]
    assert out["steps"][0] == res["steps"][0] + 2 * EVERY
    assert np.array_equal(b.get_pdf(), pdf)


def test_fallback_keeps_within_the_step_budget(cold):
    """A warm case that grows after two intervals restarts cold with what is
    left of max_steps: three intervals in all, never more."""
    case, res, pdf, (nr, nv) = cold
    b = OracleBatch([case])
    out = ws.solve(b, (nr, nv), tol=TOL, check_every=EVERY, max_steps=3 * EVERY + 17,
                   growth=1e-12)
    assert out["fallback"].tolist() == [True]
    assert out["steps"][0] == 3 * EVERY + 17
