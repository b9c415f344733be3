"""The batched workload's choked cases, on the CPU oracle.

A velocity inlet fixes the mass flux G = rho_in u_in into the channel. In
a steady state the density must fall along the obstacles by what their
flow resistance needs to carry that flux; with a kinematic viscosity fixed
by tau (dynamic viscosity rho nu), that drop grows with G, hence with
rho_in itself: rho_in (1 - K) = rho_out with K proportional to u_in times
the resistance. For K >= 1 (a choked channel) no steady state exists: mass
accumulates, the density grows without bound (exponentially), while the
velocity field stays bounded and its per-step residual levels off, so a
residual criterion alone never stops such a case. Most cases of the
workload's distribution (5-40 discs) are choked at 20k-200k cells (the
bench's batched run counts them); the steady-state distribution (3-8
discs) is not. Here: one workload case (37 discs) meshed small enough for
the CPU, and the controller's runaway stop (warmstart.run_until)."""
import importlib

import numpy as np
import pytest

from cases import fv, oracle_env, oracle_state
from oracle import oracle
from oracle_batch import CONVERGED, OracleBatch

ws = importlib.import_module("fvdbm-jax_amd.warmstart")


@pytest.fixture(scope="module")
def choked():
    oracle.set_threads(8)
    case = fv.porous.make_case(13, dist=fv.porous.WORKLOAD, target_cells=3000)
    assert len(case.discs) >= 30
    return case


def _vel(f):
    f = np.asarray(f, np.float64)
    return (f @ oracle.KSI.astype(np.float64)) / f.sum(1)[:, None]


def test_choked_case_has_no_steady_state(choked):
    """Mean density rises at an accelerating rate over 20k steps, the inlet
    density far more than the outlet's fixed 1, while the velocity stays
    bounded and the residual stays above 1e-5."""
    c = choked
    env = oracle_env(c.cells, c.faces, c.nodes)
    st = oracle_state(env, c.nodes)
    inlet = (c.mesher.point_markers == fv.meshgen.INLET) & (c.nodes.type.reshape(-1) == 1)
    rho_mean, rho_in, resid, umax = [], [], [], []
    for _ in range(4):
        oracle.run(env, st, c.dynamics.tau, c.dynamics.delta_t, 4999, "f32")
        u0 = _vel(st["cell_pdf"])
        oracle.run(env, st, c.dynamics.tau, c.dynamics.delta_t, 1, "f32")
        u1 = _vel(st["cell_pdf"])
        rho_mean.append(float(st["cell_pdf"].astype(np.float64).sum(1).mean()))
        rho_in.append(float(st["node_rho"][inlet].mean()))
        resid.append(float(np.linalg.norm(u1 - u0) / np.linalg.norm(u1)))
        umax.append(float(np.abs(u1).max()))
    gain = np.diff([1.0] + rho_mean)
    assert np.all(gain > 0) and np.all(np.diff(gain[1:]) > 0), rho_mean  # accelerating growth
    assert rho_in[-1] > 2.0 and rho_mean[-1] > 1.5, (rho_in, rho_mean)
    assert min(resid) > 1e-5, resid
    assert max(umax) < 2 * c.u_in, umax


def test_run_until_stops_the_choked_case(choked):
    """In one batch with a steady-state case, the choked case is stopped as
    runaway (mean density past the bound) and the other converges."""
    good = fv.porous.make_case(21, target_cells=1500)
    b = OracleBatch([choked, good])
    b.set_active(np.ones(2, np.int32), reset_steps=True)
    out = ws.run_until(b, 200_000, check_every=250, tol=1e-6, rho_max=1.3, chunk=5000)
    assert out["runaway"].tolist() == [True, False]
    assert out["status"].tolist() == [ws.STOPPED, CONVERGED]
    assert ws.mean_density(b)[0] > 1.3 and out["steps"][0] < 30_000
