"""Steady-state solves of a batch of cases from a cold start (uniform
equilibrium, as Environment.init does) or a warm start (equilibrium of
predicted node fields), with the cold start as the fallback of every case
whose warm start diverges or whose residual grows.

The controller only needs the solver interface BatchSolver provides
(reset_uniform, init_equilibrium_nodes, run, status, set_active, get_pdf,
set_pdf, cell_off, n), so the same logic drives the GPU batch and, in the
tests, a CPU stand-in built on the oracle.
"""
from __future__ import annotations

import time

import numpy as np

RUNNING, CONVERGED, DIVERGED, GROWING, STOPPED = range(5)
RUNAWAY_RHO = 2.0  # run_until stops a case whose mean density passes this
W = np.array([4 / 9, 1 / 9, 1 / 9, 1 / 9, 1 / 9, 1 / 36, 1 / 36, 1 / 36, 1 / 36])

WARM_GROWTH = 100.0  # a warm case stops as growing past 100x its best residual
# plausibility gate of a predicted field: finite, |u| below the weakly
# compressible limit of the D2Q9 model (Mach 0.5 at c_s = 1/sqrt(3)), rho near 1
U_MAX = 0.5 / np.sqrt(3.0)
RHO_RANGE = (0.5, 2.0)


def plausible(rho, vel):
    """True when a predicted node field may seed a warm start."""
    rho = np.asarray(rho, np.float64)
    vel = np.asarray(vel, np.float64)
    if not (np.isfinite(rho).all() and np.isfinite(vel).all()):
        return False
    speed = np.hypot(vel[:, 0], vel[:, 1])
    return bool(speed.max() <= U_MAX and rho.min() >= RHO_RANGE[0] and rho.max() <= RHO_RANGE[1])


def _set_uniform(solver, cases):
    pdf = solver.get_pdf()
    for k in cases:
        pdf[solver.cell_off[k]:solver.cell_off[k + 1]] = W.astype(pdf.dtype)
    solver.set_pdf(pdf)


def solve(solver, init=None, tol=1e-6, check_every=500, max_steps=400_000, growth=WARM_GROWTH,
          gate=True, synchronize=None):
    """Run every case of ``solver`` to a residual below ``tol``.

    init: None for a cold start, else (node_rho [N], node_vel [N,2]) over the
    concatenated nodes. With ``gate``, a case whose predicted field is not
    ``plausible`` starts cold at once. A warm case that diverges or grows
    (``growth`` times its smallest residual) is restarted cold from W; its
    step count keeps the warm steps it spent, and the cold restarts share
    what is left of ``max_steps`` after the longest of those warm attempts,
    so no case runs more than ``max_steps`` in all. Returns dict: steps, status,
    residual (per case), fallback (bool per case: started or restarted
    cold), wall_s (seconds, including the initialisation)."""
    t0 = time.perf_counter()
    n = solver.n
    fallback = np.zeros(n, bool)
    if init is None:
        solver.reset_uniform()
    else:
        rho, vel = init
        solver.init_equilibrium_nodes(rho, vel)
        if gate:
            for k in range(n):
                sl = slice(solver.node_off[k], solver.node_off[k + 1])
                fallback[k] = not plausible(rho[sl], vel[sl])
            if fallback.any():
                _set_uniform(solver, np.nonzero(fallback)[0])
    solver.set_active(np.ones(n, np.int32), reset_steps=True)
    st = solver.run(max_steps, check_every, tol, growth if init is not None else 0.0)
    if init is not None:
        bad = np.isin(st["status"], (DIVERGED, GROWING))
        if bad.any():
            fallback |= bad
            _set_uniform(solver, np.nonzero(bad)[0])
            solver.set_active(bad.astype(np.int32), reset_steps=False)
            left = int(max_steps - st["steps"][bad].max())
            if left > 0:
                st = solver.run(left, check_every, tol, 0.0)
    if synchronize is not None:
        synchronize()
    return dict(steps=st["steps"].copy(), status=st["status"].copy(), residual=st["residual"].copy(),
                fallback=fallback, wall_s=time.perf_counter() - t0)


def mean_density(solver):
    """Mean cell density of each case of ``solver`` (its current pdf)."""
    rho, _ = solver.cell_fields()
    return np.array([np.mean(solver.case_cells(k, rho), dtype=np.float64) for k in range(solver.n)])


def run_until(solver, max_steps, check_every=2000, tol=1e-6, rho_max=RUNAWAY_RHO, chunk=10_000,
              log=None):
    """Step every active case until it converges (residual below ``tol`` at
    a check), runs away or ``max_steps`` have run. A case runs away when its
    mean density passes ``rho_max``: a velocity inlet into a channel whose
    flow resistance would need a density drop larger than the inlet density
    has no steady state, its density grows without bound while the velocity
    residual levels off, so the residual alone never stops it. The density
    is read every ``chunk`` steps. Returns the status dict of
    ``solver.status()`` plus ``runaway`` (bool per case), ``wall_s``, and
    ``rho_last``/``rho_prev`` (mean density per case at the last two reads,
    ``chunk`` steps apart: the trend of a case still running)."""
    t0 = time.perf_counter()
    runaway = np.zeros(solver.n, bool)
    done = 0
    st = solver.status()
    rho_prev = rho_last = mean_density(solver)
    while done < max_steps and st["active"].any():
        n = min(chunk, max_steps - done)
        st = solver.run(n, check_every, tol, 0.0)
        done += n
        rho_prev, rho_last = rho_last, mean_density(solver)
        high = (rho_last > rho_max) & st["active"]
        if high.any():
            runaway |= high
            solver.set_active((st["active"] & ~high).astype(np.int32))
            st = solver.status()
        if log:
            log(f"{done} steps: {int(st['active'].sum())} running, "
                f"{int((st['status'] == CONVERGED).sum())} converged, {int(runaway.sum())} runaway")
    out = dict(st)
    out["runaway"] = runaway
    out["wall_s"] = time.perf_counter() - t0
    out["rho_last"], out["rho_prev"] = rho_last, rho_prev
    return out


def relative_l2(a, b):
    """||a - b||_2 / ||b||_2 over all entries (float64)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))
