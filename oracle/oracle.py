"""ctypes front end of the C oracle (fvdbm_oracle.c).

TEST INFRASTRUCTURE ONLY: the checker for the HIP step. Only tests/,
``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg may import it.

``run(env_arrays, tau, dt, nsteps, ...)`` advances a state in the reference's
container layout ([C,9] cell pdfs, [N,9] node pdfs ...) with the reference's
phase order (src/environment.py:55-65) and returns every observable array.
"""
from __future__ import annotations

import ctypes as ct
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "_build", "libfvdbm_oracle.so")

MUTATIONS = {
    "none": 0, "downwind": 1, "no_ghost": 2, "no_sign": 3, "transpose_ksi": 4,
    "drop_quadratic": 5, "idw_uniform": 6, "idw_unnormalised": 7,
    "no_collision": 8, "lw_d1_for_d0": 9, "lw_dt_sign": 10, "lw_swap_ud": 11,
}

# src/dynamics.py:54-63
KSI = np.array([[0, 0], [1, 0], [0, 1], [-1, 0], [0, -1],
                [1, 1], [-1, 1], [-1, -1], [1, -1]], dtype=np.int32)
W64 = np.array([4 / 9, 1 / 9, 1 / 9, 1 / 9, 1 / 9, 1 / 36, 1 / 36, 1 / 36, 1 / 36])
# D2Q13, src/dynamics.py:80-96
KSI13 = np.array([[0, 0], [1, 0], [0, 1], [-1, 0], [0, -1], [1, 1], [-1, 1], [-1, -1], [1, -1],
                  [2, 0], [0, 2], [-2, 0], [0, -2]], dtype=np.int32)
W13_64 = np.array([3 / 8] + [1 / 12] * 4 + [1 / 16] * 4 + [1 / 96] * 4)


def _sfx(precision, quivers):
    if quivers not in (9, 13):
        raise ValueError("quivers must be 9 (D2Q9) or 13 (D2Q13)")
    return ("q13_" if quivers == 13 else "") + precision


class _Mesh(ct.Structure):
    _fields_ = [("n_cells", ct.c_int32), ("n_faces", ct.c_int32),
                ("n_nodes", ct.c_int32), ("k_max", ct.c_int32),
                ("cell_faces", ct.c_void_p), ("cell_signs", ct.c_void_p),
                ("face_cells", ct.c_void_p), ("face_dists", ct.c_void_p),
                ("face_nodes", ct.c_void_p), ("face_n", ct.c_void_p),
                ("face_L", ct.c_void_p), ("node_cells", ct.c_void_p),
                ("node_dists", ct.c_void_p), ("node_type", ct.c_void_p)]


class _State(ct.Structure):
    _fields_ = [(n, ct.c_void_p) for n in (
        "cell_pdf", "cell_rho", "cell_vel", "cell_feq", "face_pdf",
        "node_pdf", "node_rho", "node_vel")]


_lib = None


def build():
    """Compile the oracle with its Makefile (gcc, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        _lib = ct.CDLL(_SO)
        for name in ("oracle_step_f32", "oracle_step_f64", "oracle_step_q13_f32",
                     "oracle_step_q13_f64"):
            fn = getattr(_lib, name)
            fn.restype = ct.c_int
            fn.argtypes = [ct.POINTER(_Mesh), ct.c_double, ct.c_double,
                           ct.POINTER(_State), ct.c_int, ct.c_int, ct.c_int]
        for name in ("oracle_calc_eq_f32", "oracle_calc_eq_f64", "oracle_calc_macro_f32",
                     "oracle_calc_eq_q13_f32", "oracle_calc_eq_q13_f64",
                     "oracle_calc_macro_q13_f32"):
            fn = getattr(_lib, name)
            fn.restype = ct.c_int
            fn.argtypes = [ct.c_int, ct.c_void_p, ct.c_void_p, ct.c_void_p]
    return _lib


def set_threads(n):
    """Threads of the oracle's element loops (OpenMP); returns the count in
    use. The results do not depend on it."""
    fn = lib().oracle_set_threads
    fn.restype = ct.c_int
    fn.argtypes = [ct.c_int]
    return int(fn(int(n)))


def _ptr(a):
    return a.ctypes.data_as(ct.c_void_p)


def uniform_w(n, precision="f32", quivers=9):
    """Container.__init__ pdf: calc_eq(1, [0,0]) repeated (src/containers.py:29)."""
    ft = np.float32 if precision == "f32" else np.float64
    w = W13_64 if quivers == 13 else W64
    return np.repeat(w.astype(ft)[None, :], n, axis=0)


def initial_state(env, node_type, node_rho=None, node_vel=None, precision="f32", quivers=9):
    """State right after ``Environment.init()``: every pdf is W, moments and
    pdf_eq are zero, node rho/vel are the boundary-condition values."""
    ft = np.float32 if precision == "f32" else np.float64
    C = env["cell_faces"].shape[0]
    F = env["face_nodes"].shape[0]
    N = env["node_cells"].shape[0]
    return dict(
        cell_pdf=uniform_w(C, precision, quivers), cell_rho=np.zeros(C, ft),
        cell_vel=np.zeros((C, 2), ft), cell_feq=np.zeros((C, quivers), ft),
        face_pdf=uniform_w(F, precision, quivers), node_pdf=uniform_w(N, precision, quivers),
        node_rho=np.zeros(N, ft) if node_rho is None else np.asarray(node_rho, ft).reshape(N).copy(),
        node_vel=np.zeros((N, 2), ft) if node_vel is None else np.asarray(node_vel, ft).reshape(N, 2).copy(),
        node_type=np.asarray(node_type, np.int32).reshape(N).copy(),
    )


def run(env, state, tau, dt, nsteps, precision="f32", mutation="none", scheme="upwind",
        quivers=9):
    """Advance ``state`` (dict from ``initial_state``) by ``nsteps`` in place
    and return it. ``env`` holds the to_env arrays in the working precision;
    ``scheme`` is "upwind" or "lax_wendroff"; ``quivers`` 9 (D2Q9) or 13
    (D2Q13), matching the state's pdf width."""
    ft = np.float32 if precision == "f32" else np.float64
    arrs = {}
    for k in ("cell_faces", "cell_signs", "face_cells", "face_nodes", "node_cells"):
        arrs[k] = np.ascontiguousarray(env[k], dtype=np.int32)
    for k in ("face_dists", "face_n", "face_L", "node_dists"):
        arrs[k] = np.ascontiguousarray(env[k], dtype=ft)
    arrs["node_type"] = np.ascontiguousarray(state["node_type"], dtype=np.int32)
    C, F, N = arrs["cell_faces"].shape[0], arrs["face_nodes"].shape[0], arrs["node_cells"].shape[0]
    K = arrs["node_cells"].shape[1] if arrs["node_cells"].ndim == 2 else 0
    m = _Mesh(C, F, N, K, *[_ptr(arrs[k]) for k in (
        "cell_faces", "cell_signs", "face_cells", "face_dists", "face_nodes",
        "face_n", "face_L", "node_cells", "node_dists", "node_type")])
    for k in ("cell_pdf", "cell_rho", "cell_vel", "cell_feq", "face_pdf",
              "node_pdf", "node_rho", "node_vel"):
        state[k] = np.ascontiguousarray(state[k], dtype=ft)
    s = _State(*[_ptr(state[k]) for k in (
        "cell_pdf", "cell_rho", "cell_vel", "cell_feq", "face_pdf",
        "node_pdf", "node_rho", "node_vel")])
    if state["cell_pdf"].shape[1] != quivers:
        raise ValueError("state pdf width does not match quivers")
    fn = getattr(lib(), "oracle_step_" + _sfx(precision, quivers))
    rc = fn(ct.byref(m), float(tau), float(dt), ct.byref(s), int(nsteps),
            {"upwind": 0, "lax_wendroff": 1}[scheme], MUTATIONS[mutation])
    if rc != 0:
        raise RuntimeError(f"oracle step failed ({rc})")
    return state


def calc_eq(rho, vel, precision="f32", quivers=9):
    """D2Q9.calc_eq (src/dynamics.py:70-74) or D2Q13.calc_eq (:101-102) for
    arrays rho[n], vel[n,2]."""
    ft = np.float32 if precision == "f32" else np.float64
    rho = np.ascontiguousarray(rho, ft).reshape(-1)
    vel = np.ascontiguousarray(vel, ft).reshape(-1, 2)
    out = np.empty((rho.shape[0], quivers), ft)
    fn = getattr(lib(), "oracle_calc_eq_" + _sfx(precision, quivers))
    fn(rho.shape[0], _ptr(rho), _ptr(vel), _ptr(out))
    return out


def calc_macro(f, quivers=9):
    """Dynamics.calc_macro for f[n,Q] in FP32 (src/dynamics.py:35-47)."""
    f = np.ascontiguousarray(f, np.float32).reshape(-1, quivers)
    rho = np.empty(f.shape[0], np.float32)
    vel = np.empty((f.shape[0], 2), np.float32)
    getattr(lib(), "oracle_calc_macro_" + _sfx("f32", quivers))(f.shape[0], _ptr(f), _ptr(rho), _ptr(vel))
    return rho, vel
