"""Porous-channel cases: the batched workloads and the data of the
warm-start surrogate.

Each case is a channel [0,1]x[0,0.5] with circular obstacles
(``meshgen.porous_channel``): parabolic velocity inlet on x = 0 (maximum
``u_in``), density outlet rho = 1 on x = 1, no-slip walls and obstacles, its
own ``D2Q9(tau, dt)``. The reference's update has no 1/A (the face lengths
are only scaled by ``dim_multiplier``, src/mesher.py:642), so a case fixes
its scales explicitly:

* ``dim_multiplier = 1 / (SCALE * mean cell area)``: the update is the
  area-normalised finite-volume scheme on the channel stretched to
  ``SCALE`` lattice lengths (unit lattice speed, viscosity tau/3);
* ``dt = CFL / (1/tau + max_cell sqrt(2) * sum_faces L')``, the explicit
  Euler positivity bound of the upwind update, rounded to float32.

Three distributions:

* ``WORKLOAD`` — 5-40 discs, r in [0.02,0.08], 20k-200k cells, u_in in
  [0.02,0.1], tau in [0.55,0.9]: the batched throughput workload. Most of
  these cases have no steady state (below); ``warmstart.run_until`` stops
  them as runaway once their mean density passes a bound.
* ``STEADY`` — the obstacle and flow ranges of ``STUDY`` (3-8 discs, r in
  [0.02,0.06], gap 0.03, u_in in [0.02,0.05], tau in [0.55,0.9]) at the
  batched workload's 20k-200k cells: cases that reach a 1e-6 residual
  (median ~110k steps); the surrogate is trained and warm starts are
  measured on these.
* ``STUDY`` — the same at 4k-16k cells (small cases for tests).

The batched workload's velocity inlet pushes mass into a channel whose
flow resistance grows with the obstacle count: in steady state the density
drop across the obstacles must carry that mass flux, and when it would
have to exceed the inlet density itself (a choked channel) no steady state
exists and the mean density grows without bound while the velocity
residual levels off (``warmstart.mean_density`` is the quantity a runaway
check watches).

``node_features`` gives the surrogate's per-node inputs: x, y, signed
distance to the nearest obstacle, a one-hot of the boundary marker (wall,
inlet, outlet, obstacle; all zero inside), u_in and tau.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import meshgen
from .dynamics import D2Q9
from .mesher import Mesher

SCALE = 100.0
CFL = 0.5
HEIGHT = 0.5

WORKLOAD = dict(count_range=(5, 40), r_range=(0.02, 0.08), gap=0.01,
                cells_range=(20_000, 200_000), u_range=(0.02, 0.1), tau_range=(0.55, 0.9))
STUDY = dict(count_range=(3, 8), r_range=(0.02, 0.06), gap=0.03,
             cells_range=(4_000, 16_000), u_range=(0.02, 0.05), tau_range=(0.55, 0.9))
STEADY = dict(STUDY, cells_range=(20_000, 200_000))
DISTRIBUTIONS = dict(workload=WORKLOAD, steady=STEADY, study=STUDY)

N_FEATURES = 9
MARKER_COLUMNS = (meshgen.WALL, meshgen.INLET, meshgen.OUTLET, meshgen.OBSTACLE)


@dataclass
class Case:
    seed: int
    mesher: Mesher
    dynamics: D2Q9
    cells: object
    faces: object
    nodes: object
    u_in: float
    discs: list = field(default_factory=list)

    @property
    def n_cells(self):
        return int(self.mesher.cells.shape[0])

    @property
    def n_nodes(self):
        return int(self.mesher.points.shape[0])

    def env_tuple(self):
        return self.cells, self.faces, self.nodes


def mean_area(mesher):
    p = mesher.points[mesher.cells]
    a = 0.5 * np.abs((p[:, 1, 0] - p[:, 0, 0]) * (p[:, 2, 1] - p[:, 0, 1])
                     - (p[:, 2, 0] - p[:, 0, 0]) * (p[:, 1, 1] - p[:, 0, 1]))
    return float(a.mean())


def stable_dt(mesher, faces_L, tau, cfl=CFL):
    """cfl / (1/tau + max_cell sqrt(2) * sum of the cell's scaled face lengths)."""
    L = np.asarray(faces_L, np.float64).reshape(-1)
    per_cell = L[mesher.cell_face_indices].sum(axis=1) * np.sqrt(2.0)
    return cfl / (1.0 / tau + per_cell.max())


def make_case(seed, dist=STUDY, target_cells=None, scheme="upwind", scale=SCALE, cfl=CFL,
              tau=None, u_in=None):
    """Mesh, dynamics and containers (BCs set) of porous case ``seed``."""
    rng = np.random.default_rng(10_000 + seed)
    tau = float(rng.uniform(*dist["tau_range"])) if tau is None else float(tau)
    u_in = float(rng.uniform(*dist["u_range"])) if u_in is None else float(u_in)
    raw = meshgen.porous_channel(seed, target_cells=target_cells, r_range=dist["r_range"],
                                 gap=dist["gap"], count_range=dist["count_range"],
                                 cells_range=dist["cells_range"])
    m = Mesher.from_raw(raw)
    m.enforce_ccw()
    m.calc_mesh_properties()
    mult = 1.0 / (scale * mean_area(m))
    dyn = D2Q9(tau=tau, delta_t=1.0)
    cells, faces, nodes = m.to_env(dyn, scheme, dim_multiplier=mult)
    dyn.delta_t = float(np.float32(stable_dt(m, faces.L, tau, cfl)))
    nodes = m.set_vel_profile(nodes, meshgen.INLET,
                              lambda p: meshgen.parabolic_inlet(p[:, 1], HEIGHT, u_in))
    nodes = m.set_rho_node(nodes, meshgen.OUTLET, 1.0)
    nodes = m.set_vel_node(nodes, meshgen.WALL, np.array([0.0, 0.0]))
    nodes = m.set_vel_node(nodes, meshgen.OBSTACLE, np.array([0.0, 0.0]))
    return Case(seed, m, dyn, cells, faces, nodes, u_in, list(raw.meta["discs"]))


def signed_distance(points, discs):
    """min over discs of |p - c| - r (positive in the fluid)."""
    pts = np.asarray(points, np.float64)
    d = np.full(len(pts), np.inf)
    for cx, cy, r in discs:
        d = np.minimum(d, np.hypot(pts[:, 0] - cx, pts[:, 1] - cy) - r)
    return d


def node_features(case):
    """[N, 9] float32: x, y, sdf, one-hot(wall, inlet, outlet, obstacle), u_in, tau."""
    m = case.mesher
    N = m.points.shape[0]
    out = np.zeros((N, N_FEATURES), np.float32)
    out[:, 0:2] = m.points
    out[:, 2] = signed_distance(m.points, case.discs)
    for j, mk in enumerate(MARKER_COLUMNS):
        out[:, 3 + j] = m.point_markers == mk
    out[:, 7] = case.u_in
    out[:, 8] = case.dynamics.tau
    return out


def section_features(points, discs, grid=2048):
    """Per-node cross-section descriptors of the channel at the node's x:
    open fraction a(x) = 1 - (sum of disc chords)/HEIGHT, 1/a(x), the share
    of the total hydraulic resistance int 1/a^3 dx lying downstream of x,
    and log of that total. [N, 4] float32."""
    xs = (np.arange(grid) + 0.5) / grid
    blocked = np.zeros(grid)
    for cx, cy, r in discs:
        d2 = r * r - (xs - cx) ** 2
        blocked += np.where(d2 > 0, 2.0 * np.sqrt(np.maximum(d2, 0.0)), 0.0)
    a = np.clip(1.0 - blocked / HEIGHT, 0.05, 1.0)
    res = 1.0 / a ** 3
    total = res.sum() / grid
    down = np.cumsum(res[::-1])[::-1] / grid / total
    px = np.asarray(points, np.float64)[:, 0]
    i = np.clip((px * grid).astype(np.int64), 0, grid - 1)
    return np.stack([a[i], 1.0 / a[i], down[i], np.full(len(px), np.log(total))], 1).astype(np.float32)


def mesh_edges(case):
    """Undirected node pairs of the mesh (its faces), [F, 2] int32."""
    return np.asarray(case.mesher.faces, np.int32)
