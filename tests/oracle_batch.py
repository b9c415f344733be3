"""A CPU stand-in for BatchSolver built on the oracle, for testing the
warm-start controller without a GPU: the same interface (reset_uniform,
init_equilibrium_nodes, run, status, set_active, get_pdf, set_pdf,
node_fields, cell_fields, case_cells) and the same per-case semantics (residual of the last step of
each check interval, converged / diverged / growing / stopped)."""
import numpy as np

from cases import oracle_env, oracle_state
from oracle import interp, oracle

RUNNING, CONVERGED, DIVERGED, GROWING, STOPPED = range(5)


def _vel(f):
    f = np.asarray(f, np.float64)
    return (f @ oracle.KSI.astype(np.float64)) / f.sum(1)[:, None]


class OracleBatch:
    def __init__(self, cases):
        """cases: list of porous.Case (or anything with cells/faces/nodes/dynamics)."""
        self.cases = cases
        self.envs = [oracle_env(c.cells, c.faces, c.nodes) for c in cases]
        self.n = len(cases)
        sizes = [e["cell_faces"].shape[0] for e in self.envs]
        nsz = [e["node_cells"].shape[0] for e in self.envs]
        self.cell_off = np.concatenate([[0], np.cumsum(sizes)])
        self.node_off = np.concatenate([[0], np.cumsum(nsz)])
        self.steps = np.zeros(self.n, np.int64)
        self.resid = np.zeros(self.n)
        self.best = np.full(self.n, np.inf)
        self.st = np.zeros(self.n, np.int32)
        self.active = np.ones(self.n, bool)
        self.reset_uniform()

    def reset_uniform(self):
        self.states = [oracle_state(e, c.nodes) for e, c in zip(self.envs, self.cases)]

    def init_equilibrium_nodes(self, rho, vel):
        for k, (e, s) in enumerate(zip(self.envs, self.states)):
            sl = slice(self.node_off[k], self.node_off[k + 1])
            C = e["cell_faces"].shape[0]
            cr = interp.cell_from_nodes(e["node_cells"], e["node_dists"], C, rho[sl])
            cv = interp.cell_from_nodes(e["node_cells"], e["node_dists"], C, vel[sl])
            s["cell_pdf"] = oracle.calc_eq(cr, cv)

    def get_pdf(self):
        return np.concatenate([s["cell_pdf"] for s in self.states]).copy()

    def set_pdf(self, pdf):
        for k, s in enumerate(self.states):
            s["cell_pdf"] = np.array(pdf[self.cell_off[k]:self.cell_off[k + 1]], np.float32)

    def node_fields(self):
        rr, vv = [], []
        for e, s in zip(self.envs, self.states):
            r, v = oracle.calc_macro(s["cell_pdf"])
            rr.append(interp.node_idw(e["node_cells"], e["node_dists"], r))
            vv.append(interp.node_idw(e["node_cells"], e["node_dists"], v))
        return np.concatenate(rr), np.concatenate(vv)

    def cell_fields(self):
        rr, vv = [], []
        for s in self.states:
            r, v = oracle.calc_macro(s["cell_pdf"])
            rr.append(r)
            vv.append(v)
        return np.concatenate(rr), np.concatenate(vv)

    def case_cells(self, k, arr):
        return arr[self.cell_off[k]:self.cell_off[k + 1]]

    def set_active(self, mask, reset_steps=False):
        for k, a in enumerate(np.asarray(mask).astype(bool)):
            if a:
                self.st[k], self.best[k] = RUNNING, np.inf
                if reset_steps:
                    self.steps[k] = 0
            elif self.st[k] == RUNNING:
                self.st[k] = STOPPED
            self.active[k] = a

    def run(self, max_steps, check_every=100, tol=1e-6, growth=0.0):
        done = 0
        while done < max_steps and self.active.any():
            every = min(check_every, max_steps - done)  # the last interval may be shorter
            for k in np.nonzero(self.active)[0]:
                c, e, s = self.cases[k], self.envs[k], self.states[k]
                tau, dt = c.dynamics.tau, c.dynamics.delta_t
                oracle.run(e, s, tau, dt, every - 1)
                u0 = _vel(s["cell_pdf"])
                oracle.run(e, s, tau, dt, 1)
                u1 = _vel(s["cell_pdf"])
                self.steps[k] += every
                with np.errstate(invalid="ignore"):
                    r = np.linalg.norm(u1 - u0) / np.linalg.norm(u1)
                self.resid[k] = r
                if not np.isfinite(r):
                    self.st[k], self.active[k] = DIVERGED, False
                elif r < tol:
                    self.st[k], self.active[k] = CONVERGED, False
                elif growth > 0 and r > growth * self.best[k]:
                    self.st[k], self.active[k] = GROWING, False
                self.best[k] = min(self.best[k], r)
            done += every
        return self.status()

    def status(self):
        return dict(steps=self.steps.copy(), residual=self.resid.copy(), status=self.st.copy(),
                    active=self.active.copy())
