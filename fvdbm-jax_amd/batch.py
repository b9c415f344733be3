"""Batched solves: many independent meshes stepped by one kernel launch.

The reference runs one mesh per Environment and has no convergence test;
``BatchSolver`` packs the containers of several cases (each from
``Mesher.to_env`` plus its boundary conditions, with its own ``D2Q9(tau,
delta_t)``) into one device handle (``fvdbm_create_batch``). Every case
starts on a tile boundary, so a thread block never mixes cases; each case
keeps its own buffer parity, parameters, step count and convergence mask, and
a converged case costs nothing. ``run`` replays one captured graph per check
interval until every case has stopped.
"""
from __future__ import annotations

import ctypes as ct

import numpy as np

from . import _lib

STATUS = {0: "running", 1: "converged", 2: "diverged", 3: "growing", 4: "stopped"}


def _concat(cases, precision):
    """Concatenated to_env arrays of several (cells, faces, nodes) cases."""
    f64 = precision == "fp64"
    K = max(int(np.asarray(n.cells_index).shape[1]) for _, _, n in cases)
    out = {k: [] for k in ("cell_faces", "cell_signs", "face_cells", "face_dists", "face_nodes",
                           "face_n", "face_L", "node_cells", "node_dists", "node_type",
                           "node_rho", "node_vel", "cell_xy", "cell_case")}
    coff = foff = noff = 0
    cell_off, node_off = [0], [0]
    for k, (c, f, n) in enumerate(cases):
        C, F, N = c.face_indices.shape[0], f.nodes_index.shape[0], n.type.shape[0]
        out["cell_faces"].append(np.asarray(c.face_indices, np.int64) + foff)
        out["cell_signs"].append(np.asarray(c.face_normals))
        fc = np.asarray(f.stencil_cells_index, np.int64)
        out["face_cells"].append(np.where(fc >= 0, fc + coff, -1))
        out["face_nodes"].append(np.asarray(f.nodes_index, np.int64) + noff)
        nc = np.asarray(n.cells_index, np.int64)
        nd = np.asarray(n.cell_dists_f64 if f64 else n.cell_dists)
        pad = K - nc.shape[1]
        nc = np.pad(np.where(nc >= 0, nc + coff, -1), ((0, 0), (0, pad)), constant_values=-1)
        nd = np.pad(nd, ((0, 0), (0, pad)), constant_values=-1)
        out["node_cells"].append(nc)
        out["node_dists"].append(nd)
        out["face_dists"].append(np.asarray(f.stencil_dists_f64 if f64 else f.stencil_dists))
        out["face_n"].append(np.asarray(f.n_f64 if f64 else f.n))
        out["face_L"].append(np.asarray(f.L_f64 if f64 else f.L).reshape(-1))
        out["node_type"].append(np.asarray(n.type).reshape(-1))
        out["node_rho"].append(np.asarray(n.rho_f64 if f64 else n.__dict__["_host_rho"]).reshape(-1))
        out["node_vel"].append(np.asarray(n.vel_f64 if f64 else n.__dict__["_host_vel"]).reshape(-1, 2))
        out["cell_xy"].append(np.asarray(c.centers, np.float64))
        out["cell_case"].append(np.full(C, k, np.int32))
        coff += C
        foff += F
        noff += N
        cell_off.append(coff)
        node_off.append(noff)
    arrays = {k: np.concatenate(v) for k, v in out.items()}
    return arrays, np.array(cell_off), np.array(node_off)


class BatchSolver(_lib.FieldsMixin):
    """Several independent FVDBM cases on one device handle."""

    def __init__(self, cases, precision="fp32", reorder=4, block=0, variant=0):
        """cases: list of (cells, faces, nodes) as returned by Mesher.to_env
        and the BC setters; each case's dynamics gives its tau and delta_t."""
        if precision not in ("fp32", "fp64"):
            raise ValueError("precision must be 'fp32' or 'fp64'")
        schemes = {f.flux_scheme for _, f, _ in cases}
        if len(schemes) != 1:
            raise ValueError("all cases of a batch use one flux scheme")
        lattices = {int(c.dynamics.NUM_QUIVERS) for c, _, _ in cases}
        if lattices != {9}:
            raise ValueError(f"batched solves run the D2Q9 lattice; these cases have {sorted(lattices)} "
                             "velocities (a D2Q13 case runs as its own Environment)")
        self.cases = cases
        self.precision = precision
        self.dtype = np.float64 if precision == "fp64" else np.float32
        arrays, self.cell_off, self.node_off = _concat(cases, precision)
        cell_case = arrays.pop("cell_case")
        self._arrays = arrays
        self._cell_case = cell_case
        desc = _lib.mesh_desc(arrays, self.dtype)
        self.tau = np.array([float(c.dynamics.tau) for c, _, _ in cases])
        self.dt = np.array([float(c.dynamics.delta_t) for c, _, _ in cases])
        bd = _lib.BatchDesc(len(cases), _lib.ptr(cell_case), _lib.ptr(self.tau), _lib.ptr(self.dt))
        scheme = schemes.pop()
        p = _lib.Params(1.0, 1.0, _lib.FP64 if precision == "fp64" else _lib.FP32,
                        _lib.LAX_WENDROFF if scheme == "lax_wendroff" else _lib.UPWIND,
                        int(reorder), int(variant), int(block), -1)
        h = ct.c_void_p()
        _lib.check(_lib.lib().fvdbm_create_batch(ct.byref(desc), ct.byref(p), ct.byref(bd), ct.byref(h)),
                   "fvdbm_create_batch")
        self._h = h
        self.n = len(cases)

    def _handle(self):
        if not self._h:
            raise RuntimeError("the batch handle is closed")
        return self._h

    def _counts(self):
        return int(self.cell_off[-1]), int(self.node_off[-1])

    def info(self):
        """Layout and launch facts of the handle (``variant`` has bit 8 set
        when the steps run one per launch)."""
        inf = _lib.Info()
        _lib.check(_lib.lib().fvdbm_info_get(self._handle(), ct.byref(inf)), "fvdbm_info_get")
        return {k: getattr(inf, k) for k, _ in inf._fields_}

    def case_nodes(self, k, arr):
        """Rows of node array ``arr`` that belong to case k."""
        return arr[self.node_off[k]:self.node_off[k + 1]]

    def case_cells(self, k, arr):
        """Rows of cell array ``arr`` that belong to case k."""
        return arr[self.cell_off[k]:self.cell_off[k + 1]]

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().fvdbm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # state ------------------------------------------------------------------
    def get_pdf(self):
        out = np.empty((int(self.cell_off[-1]), 9), self.dtype)
        _lib.check(_lib.lib().fvdbm_get_field(self._h, _lib.CELL_PDF, _lib.ptr(out)), "fvdbm_get_field")
        return out

    def set_pdf(self, pdf):
        a = np.ascontiguousarray(np.asarray(pdf, self.dtype).reshape(-1, 9))
        _lib.check(_lib.lib().fvdbm_set_field(self._h, _lib.CELL_PDF, _lib.ptr(a)), "fvdbm_set_field")

    def case_pdf(self, k, pdf=None):
        """Cell pdfs of case k (reference order of that case)."""
        pdf = self.get_pdf() if pdf is None else pdf
        return pdf[self.cell_off[k]:self.cell_off[k + 1]]

    def reset_uniform(self):
        _lib.check(_lib.lib().fvdbm_init_uniform(self._h), "fvdbm_init_uniform")

    # running ----------------------------------------------------------------
    def run(self, max_steps, check_every=100, tol=1e-6, growth=0.0):
        """Step active cases until each has converged/stopped or max_steps
        have run. Returns the status dict of ``status()``."""
        na = ct.c_int32()
        done = ct.c_int64()
        _lib.check(_lib.lib().fvdbm_batch_run(self._h, int(max_steps), int(check_every), float(tol),
                                              float(growth), ct.byref(na), ct.byref(done)),
                   "fvdbm_batch_run")
        self.steps_run = int(done.value)
        return self.status()

    def status(self):
        steps = np.empty(self.n, np.int64)
        resid = np.empty(self.n, np.float64)
        st = np.empty(self.n, np.int32)
        act = np.empty(self.n, np.int32)
        _lib.check(_lib.lib().fvdbm_batch_status(self._h, _lib.ptr(steps), _lib.ptr(resid),
                                                 _lib.ptr(st), _lib.ptr(act)), "fvdbm_batch_status")
        return dict(steps=steps, residual=resid, status=st, active=act.astype(bool))

    def set_active(self, mask, reset_steps=False):
        m = np.ascontiguousarray(np.asarray(mask, np.int32).reshape(self.n))
        _lib.check(_lib.lib().fvdbm_batch_set_active(self._h, _lib.ptr(m), int(bool(reset_steps))),
                   "fvdbm_batch_set_active")

    def step(self, nsteps=1):
        """Step every active case nsteps times (no convergence checks)."""
        _lib.check(_lib.lib().fvdbm_step(self._h, int(nsteps)), "fvdbm_step")
        return self
