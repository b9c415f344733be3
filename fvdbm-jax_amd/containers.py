"""Host mirrors of the reference containers (src/containers.py:18-408).

``Cells``, ``Faces`` and ``Nodes`` keep the reference's constructor
signatures and attribute names. Static geometry lives here as NumPy arrays
until ``Environment.init()`` uploads it; after that the dynamic attributes
(``pdf``, ``rho``, ``vel``, ``pdf_eq``) are read from, and written to, the
device state owned by the environment's library handle. All per-element
physics runs in the HIP library, not here.
"""
from __future__ import annotations

import numpy as np


class CustomArray:
    """-1 padded ragged rows (utils/utils.py:176-229): ``add_item`` fills the
    first free slot of a row and grows every row by one column when none is
    free."""

    def __init__(self, size, dtype=np.float32, default_value=-1):
        self.data = default_value * np.ones((size, 1), dtype=dtype)
        self.default_value = default_value

    def add_item(self, index, item):
        free = np.nonzero(self.data[index] == self.default_value)[0]
        if free.size != 0:
            self.data[index, free[0]] = item
        else:
            col = -np.ones_like(self.data[:, 0:1])
            col[index] = item
            self.data = np.concatenate((self.data, col), axis=-1)

    def add_items(self, index, items):
        for item in items:
            self.add_item(index, item)

    def __array__(self, dtype=None, copy=None):
        return np.asarray(self.data, dtype=dtype)

    def __getitem__(self, idx):
        return self.data[idx]

    def shape(self):
        return self.data.shape

    def dtype(self):
        return self.data.dtype

    def __repr__(self):
        return f"CustomArray(size={self.data.shape[0]}, data={self.data})"


class _DeviceField:
    """Attribute that lives on the host until the container is bound to an
    initialised Environment, then reads/writes the device state."""

    def __init__(self, name):
        self.name = name

    def __get__(self, obj, objtype=None):
        if obj is None:
            return self
        env = obj.__dict__.get("_env")
        if env is not None:
            return env._read(obj._kind, self.name)
        return obj.__dict__["_host_" + self.name]

    def __set__(self, obj, value):
        env = obj.__dict__.get("_env")
        if env is not None:
            env._write(obj._kind, self.name, value)
        else:
            obj.__dict__["_host_" + self.name] = value


def _uniform_pdf(size, dynamics, dtype=np.float32):
    """calc_eq(1, [0,0]) repeated (src/containers.py:29): exactly W."""
    return np.repeat(np.asarray(dynamics.W, dtype)[None, :], size, axis=0)


class Container:
    """src/containers.py:18-45."""
    _kind = "container"
    pdf = _DeviceField("pdf")

    def __init__(self, size, dynamics):
        self.__dict__["_env"] = None
        self.dynamics = dynamics
        self.pdf = _uniform_pdf(size, dynamics)

    def __repr__(self):
        return f"{type(self).__name__}(size={self.size})"

    @property
    def size(self):
        return int(np.asarray(self.__dict__["_host_pdf"]).shape[0])


class Cells(Container):
    """src/containers.py:47-134."""
    _kind = "cells"
    rho = _DeviceField("rho")
    vel = _DeviceField("vel")
    pdf_eq = _DeviceField("pdf_eq")

    def __init__(self, size, dynamics):
        super().__init__(size, dynamics)
        self.rho = np.zeros((size, 1), np.float32)
        self.vel = np.zeros((size, dynamics.DIM), np.float32)
        self.pdf_eq = np.zeros((size, dynamics.NUM_QUIVERS), np.float32)
        self.face_indices = CustomArray(size, dtype=np.int32, default_value=-1)
        self.face_normals = CustomArray(size, dtype=np.int32, default_value=-1)

    def init(self):
        self.face_indices = np.asarray(self.face_indices, np.int32)
        self.face_normals = np.asarray(self.face_normals, np.int32)


class Faces(Container):
    """src/containers.py:136-291. ``pdf`` holds the face flux."""
    _kind = "faces"

    def __init__(self, size, dynamics, flux_scheme="upwind"):
        super().__init__(size, dynamics)
        self.nodes_index = CustomArray(size, dtype=np.int32, default_value=-1)
        self.stencil_cells_index = CustomArray(size, dtype=np.int32, default_value=-1)
        self.stencil_dists = CustomArray(size, dtype=np.float32, default_value=-1)
        self.n = np.zeros((size, dynamics.DIM), np.float32)
        self.L = np.zeros((size, 1), np.float32)
        self.flux_scheme = flux_scheme

    def init(self):
        if self.flux_scheme not in ("upwind", "lax_wendroff"):
            raise ValueError(f"Unknown flux scheme: {self.flux_scheme}")
        self.nodes_index = np.asarray(self.nodes_index, np.int32)
        self.stencil_cells_index = np.asarray(self.stencil_cells_index, np.int32)
        self.stencil_dists = np.asarray(self.stencil_dists, np.float32)


class Nodes(Container):
    """src/containers.py:293-408. ``type``: 0 none, 1 velocity, 2 density."""
    _kind = "nodes"
    rho = _DeviceField("rho")
    vel = _DeviceField("vel")

    def __init__(self, size, dynamics):
        super().__init__(size, dynamics)
        self.rho = np.zeros((size, 1), np.float32)
        self.vel = np.zeros((size, dynamics.DIM), np.float32)
        self.rho_f64 = np.zeros((size, 1), np.float64)
        self.vel_f64 = np.zeros((size, dynamics.DIM), np.float64)
        self.type = np.zeros((size, 1), np.int32)
        self.cells_index = CustomArray(size, dtype=np.int32, default_value=-1)
        self.cell_dists = CustomArray(size, dtype=np.float32, default_value=-1)

    def init(self):
        self.cells_index = np.asarray(self.cells_index, np.int32)
        self.cell_dists = np.asarray(self.cell_dists, np.float32)
