"""Export and checkpoint (reference: Mesher.to_vtk / to_pickle / from_pickle,
src/mesher.py:562-608).

* ``write_vtk``: legacy-ASCII VTK unstructured grid of triangles with cell
  data ``Velocity`` (3 components, z = 0) and ``Density``, and optionally
  ``pdf`` and ``feq`` — the fields the reference's pyvista export writes —
  without pyvista.
* ``save_checkpoint`` / ``load_checkpoint``: ``.npz`` holding the mesh input
  arrays, the containers' static arrays, the dynamics parameters, the cell
  and node pdfs (reference order) and the step count, plus a hash of the
  static arrays. Nothing is unpickled on load (``allow_pickle=False``).
"""
from __future__ import annotations

import hashlib

import numpy as np


def write_vtk(mesher, env, filename, save_f=False, save_feq=False):
    pts = np.asarray(mesher.points, np.float64)
    tri = np.asarray(mesher.cells, np.int64)
    vel = np.asarray(env.cells.vel, np.float64).reshape(-1, 2)
    rho = np.asarray(env.cells.rho, np.float64).reshape(-1)
    path = f"{filename}.vtk"
    with open(path, "w") as fh:
        fh.write("# vtk DataFile Version 3.0\nfvdbm\nASCII\nDATASET UNSTRUCTURED_GRID\n")
        fh.write(f"POINTS {len(pts)} double\n")
        np.savetxt(fh, np.column_stack([pts, np.zeros(len(pts))]), fmt="%.17g")
        fh.write(f"CELLS {len(tri)} {4 * len(tri)}\n")
        np.savetxt(fh, np.column_stack([np.full(len(tri), 3), tri]), fmt="%d")
        fh.write(f"CELL_TYPES {len(tri)}\n")
        np.savetxt(fh, np.full(len(tri), 5), fmt="%d")  # VTK_TRIANGLE
        fh.write(f"CELL_DATA {len(tri)}\n")
        fh.write("VECTORS Velocity double\n")
        np.savetxt(fh, np.column_stack([vel, np.zeros(len(vel))]), fmt="%.9g")
        fh.write("SCALARS Density double 1\nLOOKUP_TABLE default\n")
        np.savetxt(fh, rho, fmt="%.9g")
        for name, on, arr in (("pdf", save_f, lambda: env.cells.pdf),
                              ("feq", save_feq, lambda: env.cells.pdf_eq)):
            if on:
                a = np.asarray(arr(), np.float64)
                fh.write(f"FIELD FieldData 1\n{name} {a.shape[1]} {a.shape[0]} double\n")
                np.savetxt(fh, a, fmt="%.9g")
    return path


_STATIC = {
    "cells": ("face_indices", "face_normals"),
    "faces": ("nodes_index", "stencil_cells_index", "stencil_dists", "n", "L"),
    "nodes": ("cells_index", "cell_dists", "type"),
}


def static_hash(env):
    h = hashlib.sha256()
    for kind, names in _STATIC.items():
        obj = getattr(env, kind)
        for n in names:
            h.update(np.ascontiguousarray(getattr(obj, n)).tobytes())
    return h.hexdigest()


def save_checkpoint(mesher, env, filename):
    out = dict(points=mesher.points, mesh_cells=mesher.cells, mesh_faces=mesher.faces,
               point_markers=mesher.point_markers,
               lattice=np.array(type(env.cells.dynamics).__name__),
               tau=np.float64(env.cells.dynamics.tau), delta_t=np.float64(env.cells.dynamics.delta_t),
               flux_scheme=np.array(env.faces.flux_scheme), precision=np.array(env.precision),
               cell_pdf=env.cells.pdf, node_pdf=env.nodes.pdf,
               node_rho=env.nodes.rho, node_vel=env.nodes.vel,
               static_hash=np.array(static_hash(env)),
               steps=np.int64(env.info()["steps_done"]))
    for kind, names in _STATIC.items():
        for n in names:
            out[f"{kind}__{n}"] = np.asarray(getattr(getattr(env, kind), n))
    for n in ("n_f64", "L_f64", "stencil_dists_f64"):
        if getattr(env.faces, n, None) is not None:
            out[f"faces__{n}"] = getattr(env.faces, n)
    for n in ("cell_dists_f64", "rho_f64", "vel_f64"):
        if getattr(env.nodes, n, None) is not None:
            out[f"nodes__{n}"] = getattr(env.nodes, n)
    if getattr(env.cells, "centers", None) is not None:
        out["cells__centers"] = env.cells.centers
    path = f"{filename}.npz"
    np.savez(path, **out)
    return path


def load_checkpoint(filename):
    """Returns (env, mesher): a new initialised Environment holding the saved
    pdfs, and the Mesher with the saved input arrays."""
    from .containers import Cells, Faces, Nodes
    from .dynamics import D2Q9, D2Q13
    from .environment import Environment
    from .mesher import Mesher

    z = np.load(f"{filename}.npz" if not str(filename).endswith(".npz") else filename)
    m = Mesher()
    m.points, m.cells, m.faces, m.point_markers = (z["points"], z["mesh_cells"], z["mesh_faces"],
                                                   z["point_markers"])
    lattices = {"D2Q9": D2Q9, "D2Q13": D2Q13}
    name = str(z["lattice"]) if "lattice" in z.files else "D2Q9"  # files written before it was stored
    if name not in lattices:
        raise ValueError(f"checkpoint lattice {name!r} is not one of {sorted(lattices)}")
    dyn = lattices[name](float(z["tau"]), float(z["delta_t"]))
    C, F, N = z["cells__face_indices"].shape[0], z["faces__nodes_index"].shape[0], z["nodes__type"].shape[0]
    cells, faces, nodes = Cells(C, dyn), Faces(F, dyn, str(z["flux_scheme"])), Nodes(N, dyn)
    for kind, obj in (("cells", cells), ("faces", faces), ("nodes", nodes)):
        for key in z.files:
            if key.startswith(kind + "__"):
                setattr(obj, key.split("__", 1)[1], z[key])
    nodes.rho = z["node_rho"]
    nodes.vel = z["node_vel"]
    cells.pdf = z["cell_pdf"]
    nodes.pdf = z["node_pdf"]
    env = Environment(cells, faces, nodes, precision=str(z["precision"]))
    if static_hash(env) != str(z["static_hash"]):
        raise ValueError("checkpoint static arrays do not match their hash")
    env.init()
    return env, m
