"""Environment: the drop-in for the reference's ``Environment``
(src/environment.py:10-67).

``Environment(cells, faces, nodes)`` then ``.init()`` uploads the static
geometry once to a device handle (libfvdbm); ``env = env.step()`` advances
one step on the GPU and returns ``self``, matching the reference idiom
``for i in range(n): env = env.step()`` (README.md:99-106). After a step,
``cells.pdf`` is the new pdf while ``cells.rho``/``vel``/``pdf_eq``,
``faces.pdf`` and the node fields describe the start of that step, exactly
as the reference leaves them (src/environment.py:60-64).

Differences from the reference, by design:
  * state lives on the device; attribute reads copy to the host lazily;
  * ``step(nsteps)`` accepts a step count (default 1) and runs the steps as
    captured hipGraphs; ``step_residual`` and ``run_to_steady`` add the
    convergence residual the reference lacks;
  * ``precision="fp64"`` runs the same kernels in double precision (the
    FP64 twin used to state the FP32 error);
  * an Environment on ``D2Q13`` dynamics (or any, with
    ``lattice_kernel=True``) runs on the library's any-lattice handle
    (fvdbm_lattice_*: plain kernels, one step per launch); the residual,
    warm-start and field-interpolation methods need the tuned D2Q9 handle.
"""
from __future__ import annotations

import ctypes as ct

import numpy as np

from . import _lib
from .containers import Cells, Faces, Nodes

_Q = None  # the lattice's pdf width
_FIELDS = {
    ("cells", "pdf"): (_lib.CELL_PDF, _Q),
    ("cells", "rho"): (_lib.CELL_RHO, 1),
    ("cells", "vel"): (_lib.CELL_VEL, 2),
    ("cells", "pdf_eq"): (_lib.CELL_PDF_EQ, _Q),
    ("faces", "pdf"): (_lib.FACE_FLUX, _Q),
    ("nodes", "pdf"): (_lib.NODE_PDF, _Q),
    ("nodes", "rho"): (_lib.NODE_RHO, 1),
    ("nodes", "vel"): (_lib.NODE_VEL, 2),
}
# entry points of the two handle kinds
_TUNED_API = dict(destroy="fvdbm_destroy", set="fvdbm_set_field", get="fvdbm_get_field",
                  step="fvdbm_step", timed="fvdbm_step_timed", sync="fvdbm_synchronize",
                  info="fvdbm_info_get", perm="fvdbm_get_permutation", prepare="fvdbm_prepare")
_LATTICE_API = {k: v.replace("fvdbm_", "fvdbm_lattice_", 1) for k, v in _TUNED_API.items()}
_SETTABLE = {("cells", "pdf"), ("nodes", "pdf"), ("nodes", "rho"), ("nodes", "vel")}


def _f64_or(obj, name, fallback):
    v = getattr(obj, name, None)
    return np.asarray(fallback, np.float64) if v is None else np.asarray(v, np.float64)


class Environment(_lib.FieldsMixin):
    """Cells, faces and nodes of one mesh plus the device state that steps them."""

    dynamics = None  # class attribute used by create(), as in the reference

    def __init__(self, cells: Cells, faces: Faces, nodes: Nodes, precision="fp32",
                 reorder=None, variant=0, block=0, graph_steps=0, lattice_kernel=False):
        self.cells = cells
        self.faces = faces
        self.nodes = nodes
        if precision not in ("fp32", "fp64"):
            raise ValueError("precision must be 'fp32' or 'fp64'")
        self.precision = precision
        self.options = dict(reorder=reorder, variant=variant, block=block,
                            graph_steps=graph_steps, lattice_kernel=bool(lattice_kernel))
        self._h = None
        self._api = _TUNED_API
        self._lattice = False

    @classmethod
    def create(cls, num_cells, num_faces, num_nodes):
        """Empty containers to fill by hand (src/environment.py:20-27)."""
        temp = cls.__new__(cls)
        Environment.__init__(temp, Cells(num_cells, cls.dynamics),
                             Faces(num_faces, cls.dynamics), Nodes(num_nodes, cls.dynamics))
        return temp

    @classmethod
    def define(cls, cells, faces, nodes):
        temp = cls.__new__(cls)
        Environment.__init__(temp, cells, faces, nodes)
        return temp

    # ------------------------------------------------------------------
    @property
    def dtype(self):
        return np.float64 if self.precision == "fp64" else np.float32

    def _mesh_arrays(self):
        c, f, n = self.cells, self.faces, self.nodes
        N = n.size
        if self.precision == "fp64":
            face_dists = _f64_or(f, "stencil_dists_f64", f.stencil_dists)
            face_n = _f64_or(f, "n_f64", f.n)
            face_L = _f64_or(f, "L_f64", f.L)
            node_dists = _f64_or(n, "cell_dists_f64", n.cell_dists)
            node_rho = _f64_or(n, "rho_f64", n.__dict__["_host_rho"])
            node_vel = _f64_or(n, "vel_f64", n.__dict__["_host_vel"])
        else:
            face_dists, face_n, face_L = f.stencil_dists, f.n, f.L
            node_dists = n.cell_dists
            node_rho = n.__dict__["_host_rho"]
            node_vel = n.__dict__["_host_vel"]
        return dict(
            cell_faces=c.face_indices, cell_signs=c.face_normals,
            face_cells=f.stencil_cells_index, face_dists=face_dists,
            face_nodes=f.nodes_index, face_n=face_n,
            face_L=np.asarray(face_L).reshape(-1), node_cells=n.cells_index,
            node_dists=node_dists, node_type=np.asarray(n.type).reshape(N),
            node_rho=np.asarray(node_rho).reshape(N), node_vel=np.asarray(node_vel).reshape(N, 2),
            cell_xy=getattr(c, "centers", None))

    def init(self):
        """Convert the ragged arrays, upload the static geometry and the
        initial state (src/environment.py:38-42)."""
        if self.faces.flux_scheme not in ("upwind", "lax_wendroff"):
            raise ValueError(f"Unknown flux scheme: {self.faces.flux_scheme}")
        self.cells.init()
        self.faces.init()
        self.nodes.init()
        self.close()
        dt = self.dtype
        arrays = self._mesh_arrays()
        reorder = self.options["reorder"]
        if reorder is None:  # Hilbert order when centroids are known, else RCM
            reorder = 4 if arrays.get("cell_xy") is not None else 1
        desc = _lib.mesh_desc(arrays, dt)
        dyn = self.cells.dynamics
        p = _lib.Params(float(dyn.tau), float(dyn.delta_t),
                        _lib.FP64 if self.precision == "fp64" else _lib.FP32,
                        _lib.LAX_WENDROFF if self.faces.flux_scheme == "lax_wendroff" else _lib.UPWIND,
                        int(reorder), int(self.options["variant"]),
                        int(self.options["block"]), int(self.options["graph_steps"]))
        h = ct.c_void_p()
        quivers = int(dyn.NUM_QUIVERS)
        self._lattice = quivers != 9 or self.options["lattice_kernel"]
        if self._lattice:
            _lib.check(_lib.lib().fvdbm_lattice_create(ct.byref(desc), ct.byref(p), quivers, ct.byref(h)),
                       "fvdbm_lattice_create")
            self._api = _LATTICE_API
        else:
            _lib.check(_lib.lib().fvdbm_create(ct.byref(desc), ct.byref(p), ct.byref(h)), "fvdbm_create")
            self._api = _TUNED_API
        self._h = h
        self._arrays = arrays  # keep host copies alive for the handle's lifetime
        # initial state: fvdbm_create set every pdf to W in the working
        # precision; upload only pdfs the caller changed before init()
        W32 = np.asarray(dyn.W, np.float32)[None, :]
        for field, obj in ((_lib.CELL_PDF, self.cells), (_lib.NODE_PDF, self.nodes)):
            host = np.asarray(obj.__dict__["_host_pdf"])
            if not (host.dtype == np.float32 and np.array_equal(host, np.broadcast_to(W32, host.shape))):
                self._upload(field, host, quivers)
        for obj in (self.cells, self.faces, self.nodes):
            obj.__dict__["_env"] = self

    def close(self):
        if getattr(self, "_h", None):
            getattr(_lib.lib(), self._api["destroy"])(self._h)
            self._h = None
            for obj in (self.cells, self.faces, self.nodes):
                obj.__dict__["_env"] = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _counts(self):
        return self.cells.size, self.nodes.size

    def _core(self):
        if not self._h:
            raise RuntimeError("Environment.init() has not been called")
        return self._h

    def _handle(self):
        """The tuned D2Q9 handle (residual, warm start and field interpolation)."""
        h = self._core()
        if self._lattice:
            raise RuntimeError("this operation needs the tuned D2Q9 handle; this Environment runs "
                               "on the any-lattice handle")
        return h

    def _call(self, op, *args):
        name = self._api[op]
        _lib.check(getattr(_lib.lib(), name)(self._core(), *args), name)

    def _comps(self, field_comps):
        return int(self.cells.dynamics.NUM_QUIVERS) if field_comps is _Q else field_comps

    def _upload(self, field, value, comps):
        arr = np.ascontiguousarray(np.asarray(value, dtype=self.dtype).reshape(-1, comps))
        self._call("set", field, _lib.ptr(arr))

    # container attribute plumbing ---------------------------------------
    def _read(self, kind, name):
        field, comps = _FIELDS[(kind, name)]
        comps = self._comps(comps)
        n = {"cells": self.cells, "faces": self.faces, "nodes": self.nodes}[kind].size
        out = np.empty((n, comps), dtype=self.dtype)
        self._call("get", field, _lib.ptr(out))
        return out

    def _write(self, kind, name, value):
        if (kind, name) not in _SETTABLE:
            raise ValueError(f"{kind}.{name} is derived from the pdf and cannot be set")
        field, comps = _FIELDS[(kind, name)]
        self._upload(field, value, self._comps(comps))

    # stepping --------------------------------------------------------------
    def step(self, nsteps=1):
        """One FVDBM step (or ``nsteps``): macros -> equilibrium -> node pdfs
        -> face fluxes -> cell update (src/environment.py:55-65). Returns self."""
        self._call("step", int(nsteps))
        return self

    def step_residual(self, nsteps=1):
        """Advance ``nsteps`` and return R = ||u^{n+1}-u^n|| / ||u^{n+1}|| of
        the last step (over all cells, FP64 accumulation)."""
        r = ct.c_double()
        _lib.check(_lib.lib().fvdbm_step_residual(self._handle(), int(nsteps), ct.byref(r)),
                   "fvdbm_step_residual")
        return float(r.value)

    def run_to_steady(self, tol=1e-6, check_every=100, max_steps=1_000_000):
        """Step until the residual of a checked step is below ``tol``.
        Returns (steps taken, residual history)."""
        hist = []
        steps = 0
        while steps < max_steps:
            n = min(check_every, max_steps - steps)
            r = self.step_residual(n)
            steps += n
            hist.append((steps, r))
            if not np.isfinite(r) or r < tol:
                break
        return steps, hist

    def prepare(self, nsteps):
        """Capture, on the host, the graphs a later run of ``nsteps`` steps
        replays (nothing runs), so that run starts without capturing."""
        self._call("prepare", int(nsteps))

    def time_steps(self, nsteps):
        """Run ``nsteps`` between HIP events; returns elapsed milliseconds.
        Steps already queued (``step`` is asynchronous) run before the first
        event, so a warm-up queued just before keeps the device busy while
        this run is submitted."""
        ms = ct.c_double()
        self._call("timed", int(nsteps), ct.byref(ms))
        return float(ms.value)

    def synchronize(self):
        self._call("sync")

    def info(self):
        inf = _lib.Info()
        self._call("info", ct.byref(inf))
        return {k: getattr(inf, k) for k, _ in inf._fields_}

    def permutation(self):
        out = np.empty(self.cells.size, np.int32)
        self._call("perm", _lib.ptr(out))
        return out

    def __repr__(self):
        return f"Environment(cells={self.cells}, faces={self.faces}, nodes={self.nodes})"
