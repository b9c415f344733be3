"""ctypes binding of libfvdbm.so (the C ABI declared in include/fvdbm.h).

The library is loaded from fvdbm-jax_amd/_build/ and nothing else: there is
no CPU fallback. A missing library raises ``ImportError`` naming the build
step. Before loading it, the HIP runtime bundled with PyTorch-ROCm is loaded
by path (when PyTorch is installed), so the solver and any PyTorch code in
the process share one HIP runtime whichever is imported first.
"""
from __future__ import annotations

import ctypes as ct
import importlib.util
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libfvdbm.so")

FP32, FP64 = 0, 1
UPWIND, LAX_WENDROFF = 0, 1
CELL_PDF, CELL_RHO, CELL_VEL, CELL_PDF_EQ, FACE_FLUX, NODE_PDF, NODE_RHO, NODE_VEL = range(8)
ERR_INVALID, ERR_HIP, ERR_ALLOC, ERR_STATE = 1, 2, 3, 4


class MeshDesc(ct.Structure):
    _fields_ = [("n_cells", ct.c_int32), ("n_faces", ct.c_int32),
                ("n_nodes", ct.c_int32), ("k_max", ct.c_int32),
                ("cell_faces", ct.c_void_p), ("cell_signs", ct.c_void_p),
                ("face_cells", ct.c_void_p), ("face_dists", ct.c_void_p),
                ("face_nodes", ct.c_void_p), ("face_n", ct.c_void_p),
                ("face_L", ct.c_void_p), ("node_cells", ct.c_void_p),
                ("node_dists", ct.c_void_p), ("node_type", ct.c_void_p),
                ("node_rho", ct.c_void_p), ("node_vel", ct.c_void_p),
                ("cell_xy", ct.c_void_p)]


class Params(ct.Structure):
    _fields_ = [("tau", ct.c_double), ("delta_t", ct.c_double),
                ("precision", ct.c_int32), ("scheme", ct.c_int32),
                ("reorder", ct.c_int32), ("variant", ct.c_int32),
                ("block", ct.c_int32), ("graph_steps", ct.c_int32)]


class Info(ct.Structure):
    _fields_ = [(n, ct.c_int32) for n in (
        "n_cells", "n_faces", "n_nodes", "k_max", "n_cells_padded",
        "n_boundary_cells", "n_boundary_faces", "n_active_nodes", "bandwidth",
        "precision", "scheme", "variant", "block", "graph_steps")] + [
        ("steps_done", ct.c_int64), ("device_bytes", ct.c_int64), ("n_cases", ct.c_int32),
        ("launch_classes", ct.c_int32)]


class BatchDesc(ct.Structure):
    _fields_ = [("n_cases", ct.c_int32), ("cell_case", ct.c_void_p),
                ("case_tau", ct.c_void_p), ("case_dt", ct.c_void_p)]


RUNNING, CONVERGED, DIVERGED, GROWING, STOPPED = range(5)


# every function the header declares, with its ctypes signature
_SIGNATURES = {
    "fvdbm_last_error": (ct.c_char_p, []),
    "fvdbm_abi_version": (ct.c_int, []),
    "fvdbm_device_count": (ct.c_int, [ct.POINTER(ct.c_int32)]),
    "fvdbm_variant_count": (ct.c_int, [ct.POINTER(ct.c_int32)]),
    "fvdbm_create": (ct.c_int, [ct.POINTER(MeshDesc), ct.POINTER(Params), ct.POINTER(ct.c_void_p)]),
    "fvdbm_destroy": (ct.c_int, [ct.c_void_p]),
    "fvdbm_info_get": (ct.c_int, [ct.c_void_p, ct.POINTER(Info)]),
    "fvdbm_init_uniform": (ct.c_int, [ct.c_void_p]),
    "fvdbm_set_field": (ct.c_int, [ct.c_void_p, ct.c_int32, ct.c_void_p]),
    "fvdbm_get_field": (ct.c_int, [ct.c_void_p, ct.c_int32, ct.c_void_p]),
    "fvdbm_step": (ct.c_int, [ct.c_void_p, ct.c_int64]),
    "fvdbm_step_residual": (ct.c_int, [ct.c_void_p, ct.c_int64, ct.POINTER(ct.c_double)]),
    "fvdbm_prepare": (ct.c_int, [ct.c_void_p, ct.c_int64]),
    "fvdbm_step_timed": (ct.c_int, [ct.c_void_p, ct.c_int64, ct.POINTER(ct.c_double)]),
    "fvdbm_synchronize": (ct.c_int, [ct.c_void_p]),
    "fvdbm_get_permutation": (ct.c_int, [ct.c_void_p, ct.c_void_p]),
    "fvdbm_copy_probe": (ct.c_int, [ct.c_int64, ct.c_int32, ct.POINTER(ct.c_double)]),
    "fvdbm_lattice_create": (ct.c_int, [ct.POINTER(MeshDesc), ct.POINTER(Params), ct.c_int32,
                                        ct.POINTER(ct.c_void_p)]),
    "fvdbm_lattice_destroy": (ct.c_int, [ct.c_void_p]),
    "fvdbm_lattice_init_uniform": (ct.c_int, [ct.c_void_p]),
    "fvdbm_lattice_step": (ct.c_int, [ct.c_void_p, ct.c_int64]),
    "fvdbm_lattice_prepare": (ct.c_int, [ct.c_void_p, ct.c_int64]),
    "fvdbm_lattice_step_timed": (ct.c_int, [ct.c_void_p, ct.c_int64, ct.POINTER(ct.c_double)]),
    "fvdbm_lattice_synchronize": (ct.c_int, [ct.c_void_p]),
    "fvdbm_lattice_get_field": (ct.c_int, [ct.c_void_p, ct.c_int32, ct.c_void_p]),
    "fvdbm_lattice_set_field": (ct.c_int, [ct.c_void_p, ct.c_int32, ct.c_void_p]),
    "fvdbm_lattice_info_get": (ct.c_int, [ct.c_void_p, ct.POINTER(Info)]),
    "fvdbm_lattice_get_permutation": (ct.c_int, [ct.c_void_p, ct.c_void_p]),
    "fvdbm_create_batch": (ct.c_int, [ct.POINTER(MeshDesc), ct.POINTER(Params), ct.POINTER(BatchDesc),
                                      ct.POINTER(ct.c_void_p)]),
    "fvdbm_batch_run": (ct.c_int, [ct.c_void_p, ct.c_int64, ct.c_int32, ct.c_double, ct.c_double,
                                   ct.POINTER(ct.c_int32), ct.POINTER(ct.c_int64)]),
    "fvdbm_batch_status": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p]),
    "fvdbm_batch_set_active": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.c_int32]),
    "fvdbm_node_fields": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.c_void_p]),
    "fvdbm_cell_fields": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.c_void_p]),
    "fvdbm_set_equilibrium": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.c_void_p]),
    "fvdbm_set_equilibrium_nodes": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.c_void_p]),
    "fvdbm_nn_last_error": (ct.c_char_p, []),
    "fvdbm_nn_chunk_rows": (ct.c_int, []),
    "fvdbm_nn_mlp_forward": (ct.c_int, [ct.c_void_p, ct.c_int64, ct.c_int32, ct.c_void_p, ct.c_void_p,
                                        ct.c_void_p, ct.c_void_p, ct.c_int32, ct.c_void_p, ct.c_void_p]),
    "fvdbm_nn_mlp_backward": (ct.c_int, [ct.c_void_p, ct.c_int64, ct.c_int32, ct.c_void_p, ct.c_void_p,
                                         ct.c_void_p, ct.c_int32, ct.c_void_p, ct.c_void_p, ct.c_void_p,
                                         ct.c_void_p, ct.c_void_p]),
    "fvdbm_nn_wgrad": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.c_int64, ct.c_int32, ct.c_int32,
                                  ct.c_void_p, ct.c_void_p, ct.c_void_p]),
    "fvdbm_nn_wgrad_noalloc": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.c_int64, ct.c_int32,
                                          ct.c_int32, ct.c_void_p, ct.c_void_p]),
    "fvdbm_nn_colsum": (ct.c_int, [ct.c_void_p, ct.c_int64, ct.c_int32, ct.c_void_p, ct.c_void_p,
                                   ct.c_void_p]),
    "fvdbm_nn_aggregate": (ct.c_int, [ct.c_void_p, ct.c_int64, ct.c_void_p, ct.c_void_p, ct.c_void_p,
                                      ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p]),
    "fvdbm_nn_aggregate_backward": (ct.c_int, [ct.c_void_p, ct.c_int64, ct.c_void_p, ct.c_void_p,
                                               ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p,
                                               ct.c_void_p]),
    "fvdbm_nn_graph_edges": (ct.c_int64, [ct.c_int64, ct.c_int64, ct.c_int32, ct.c_int32]),
    "fvdbm_nn_graph_error": (ct.c_char_p, []),
    "fvdbm_nn_case_features": (ct.c_int, [ct.c_void_p, ct.c_int64, ct.c_void_p, ct.c_void_p, ct.c_void_p,
                                          ct.c_int32, ct.c_double, ct.c_double, ct.c_int32, ct.c_int32,
                                          ct.c_double, ct.c_int32, ct.c_void_p, ct.c_void_p, ct.c_void_p,
                                          ct.c_void_p, ct.c_int32, ct.c_void_p]),
    "fvdbm_nn_case_graph": (ct.c_int, [ct.c_void_p, ct.c_int64, ct.c_void_p, ct.c_int64, ct.c_int32,
                                       ct.c_int32, ct.c_double] + [ct.c_void_p] * 8),
    "fvdbm_layout_order": (ct.c_int, [ct.POINTER(MeshDesc), ct.c_int32, ct.c_int32,
                                      ct.c_void_p, ct.c_void_p]),
}

_lib = None


def _preload_torch_hip_runtime():
    spec = importlib.util.find_spec("torch")
    if spec is None or not spec.submodule_search_locations:
        return None
    path = os.path.join(list(spec.submodule_search_locations)[0], "lib", "libamdhip64.so")
    if os.path.exists(path):
        return ct.CDLL(path, mode=ct.RTLD_GLOBAL)
    return None


def lib():
    """The loaded library (raises ImportError when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with "
                "`python -c 'import __graft_entry__ as g; g.build()'` "
                "(the solver has no CPU fallback)")
        _preload_torch_hip_runtime()
        L = ct.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def declared_symbols():
    return list(_SIGNATURES)


class FvdbmError(RuntimeError):
    pass


def check(rc, what=""):
    """Raise for a non-zero status: ValueError for invalid input (the
    reference's convention), RuntimeError otherwise."""
    if rc == 0:
        return
    msg = lib().fvdbm_last_error().decode(errors="replace")
    text = f"{what}: {msg}" if what else msg
    if rc == ERR_INVALID:
        raise ValueError(text)
    raise FvdbmError(f"{text} (status {rc})")


def device_count():
    n = ct.c_int32(0)
    check(lib().fvdbm_device_count(ct.byref(n)), "fvdbm_device_count")
    return int(n.value)


def ptr(a):
    return ct.c_void_p(a.ctypes.data) if a is not None else ct.c_void_p(0)


def mesh_desc(arrays, dtype):
    """MeshDesc over a dict of contiguous arrays (kept alive by the caller).
    Keys: cell_faces, cell_signs, face_cells, face_dists, face_nodes, face_n,
    face_L, node_cells, node_dists, node_type, node_rho, node_vel, and the
    optional cell_xy (centroids, for the Hilbert cell order)."""
    a = arrays
    for k in ("cell_faces", "cell_signs", "face_cells", "face_nodes", "node_cells", "node_type"):
        a[k] = np.ascontiguousarray(a[k], dtype=np.int32)
    for k in ("face_dists", "face_n", "face_L", "node_dists", "node_rho", "node_vel"):
        a[k] = np.ascontiguousarray(a[k], dtype=dtype)
    C = a["cell_faces"].shape[0]
    F = a["face_nodes"].shape[0]
    N = a["node_type"].reshape(-1).shape[0]
    K = a["node_cells"].shape[1] if a["node_cells"].ndim == 2 else 0
    xy = a.get("cell_xy")
    if xy is not None:
        a["cell_xy"] = xy = np.ascontiguousarray(xy, dtype=np.float64).reshape(C, 2)
    return MeshDesc(C, F, N, K, *[ptr(a[k]) for k in (
        "cell_faces", "cell_signs", "face_cells", "face_dists", "face_nodes",
        "face_n", "face_L", "node_cells", "node_dists", "node_type",
        "node_rho", "node_vel")], ptr(xy))


def copy_probe(nbytes, reps=20):
    """Milliseconds per device copy of ``nbytes`` (k_copy_probe)."""
    ms = ct.c_double()
    check(lib().fvdbm_copy_probe(int(nbytes), int(reps), ct.byref(ms)), "fvdbm_copy_probe")
    return float(ms.value)


class FieldsMixin:
    """All-node interpolation and equilibrium initialisation, shared by the
    single-mesh Environment and BatchSolver. Subclasses provide
    ``_handle()``, ``dtype`` and ``_counts()`` -> (cells, nodes)."""

    def node_fields(self):
        """(rho [N], u [N,2]) at every node: inverse-distance weighting of the
        moments of the current cell pdf."""
        C, N = self._counts()
        rho = np.empty(N, self.dtype)
        vel = np.empty((N, 2), self.dtype)
        check(lib().fvdbm_node_fields(self._handle(), ptr(rho), ptr(vel)), "fvdbm_node_fields")
        return rho, vel

    def cell_fields(self):
        """(rho [C], u [C,2]) of the current cell pdf."""
        C, N = self._counts()
        rho = np.empty(C, self.dtype)
        vel = np.empty((C, 2), self.dtype)
        check(lib().fvdbm_cell_fields(self._handle(), ptr(rho), ptr(vel)), "fvdbm_cell_fields")
        return rho, vel

    def init_equilibrium(self, rho, vel):
        """Cell pdf <- feq(rho, u) from cell fields rho [C], u [C,2]."""
        C, N = self._counts()
        r = np.ascontiguousarray(np.asarray(rho, self.dtype).reshape(C))
        v = np.ascontiguousarray(np.asarray(vel, self.dtype).reshape(C, 2))
        check(lib().fvdbm_set_equilibrium(self._handle(), ptr(r), ptr(v)), "fvdbm_set_equilibrium")

    def init_equilibrium_nodes(self, rho, vel):
        """Cell pdf <- feq of the node fields rho [N], u [N,2] interpolated to
        the cells (inverse-distance weighting of each cell's vertices)."""
        C, N = self._counts()
        r = np.ascontiguousarray(np.asarray(rho, self.dtype).reshape(N))
        v = np.ascontiguousarray(np.asarray(vel, self.dtype).reshape(N, 2))
        check(lib().fvdbm_set_equilibrium_nodes(self._handle(), ptr(r), ptr(v)),
              "fvdbm_set_equilibrium_nodes")
