"""Warm-start surrogate: a mesh message-passing network that predicts the
steady (rho - 1, u) of a porous-channel case at every mesh node.

The network (``MeshNet``) works on the node graph of the mesh (its faces as
undirected edges):

  h0   = mlp_enc(x)                      x: 9 node features (+2 flags) padded to 16
  z_l  = [h | mean_j h_j | mean_j ex_ij h_j | mean_j ey_ij h_j]     (256)
  h_l+1 = h_l + mlp_l(z_l)               l = 0..5
  y    = mlp_dec(h_6)[:, :3]             (rho - 1, ux, uy), normalised

with mlp(x) = silu(x W1 + b1) W2 + b2, hidden width 64, and (ex, ey) the
unit vector along edge i -> j. Besides the mesh nodes the graph holds
virtual nodes that give every layer a long reach: SLABS "slab" nodes, each
linked both ways to the mesh nodes in its band of x and to its neighbouring
slabs, and one "global" node linked both ways to every slab. Virtual nodes
carry flags in input columns 9 (slab) and 10 (global) and no targets.
Targets are (rho - 1, ux, uy) divided by the case's inlet speed and by a
per-component scale fixed from the training set. Every dense layer, its backward pass, the
weight gradients and the aggregation are this library's HIP kernels
(fvdbm_nn_*: bf16 MFMA operands, float32 accumulation, float32 master
weights) behind ``torch.autograd.Function``; PyTorch only holds the tensors,
runs Adam and the loss. There is no CPU path: without a GPU and the built
library the functions here raise.

Per-case calibration: the network's fields are close in shape, but their
overall level varies from case to case (the least-squares factor between a
prediction's rho - 1 and the converged one spreads over 0.7-1.3). A ridge
model of log(factor) on case-level features (``case_features``: obstacle
count and sizes, tau, inlet speed, cell count and summaries of the
prediction itself), fitted on cold-solved cases by ``calibrate_surrogate``,
rescales rho - 1 and u of each prediction (``apply_calibration``). Weight
files without the ``calib_*`` arrays predict uncalibrated.
"""
from __future__ import annotations

import ctypes as ct
import math
import time

import numpy as np

from . import _lib, porous

HID = 64
IN_PAD = 16
OUT = 3
OUT_PAD = 16
LAYERS = 6

# case-level features of the per-case calibration (``case_features``)
CALIB_FEATURES = ("obstacles", "solid_fraction", "r_mean", "r_max", "tau", "u_in", "cells",
                  "pred_inlet_rho", "pred_mean_rho", "pred_umax", "log_cells", "pred_inlet_rho_tau")
CALIB_KEYS = ("calib_mu", "calib_sd", "calib_w_rho", "calib_b_rho", "calib_w_u", "calib_b_u")
CALIB_CLIP = (0.5, 2.0)  # bounds of an applied factor

# fixed input scaling: (x - shift) / scale per feature column
FEATURE_SHIFT = np.array([0.5, 0.25, 0.126, 0, 0, 0, 0, 0.035, 0.725], np.float32)
FEATURE_SCALE = np.array([0.29, 0.144, 0.095, 1, 1, 1, 1, 0.009, 0.1], np.float32)


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("the surrogate kernels need a GPU (no CPU path)")
    return torch


def _check(rc, what):
    if rc:
        msg = _lib.lib().fvdbm_nn_last_error().decode(errors="replace")
        raise RuntimeError(f"{what}: {msg} (status {rc})")


def _p(t):
    return ct.c_void_p(t.data_ptr())


def _stream():
    torch = _torch()
    return ct.c_void_p(torch.cuda.current_stream().cuda_stream)


# ---------------------------------------------------------------------------
# graphs
# ---------------------------------------------------------------------------
def graph_arrays(points, edges):
    """CSR node graph of a mesh alone: both directions of every edge,
    sorted by (source, target), w = 1/deg(source), (ex, ey) =
    unit(p_target - p_source), and for the backward pass the reverse edge's
    weight wt = 1/deg(target) and direction (ext, eyt) = -(ex, ey)."""
    pts = np.asarray(points, np.float64)
    e = np.asarray(edges, np.int64)
    src = np.concatenate([e[:, 0], e[:, 1]])
    dst = np.concatenate([e[:, 1], e[:, 0]])
    d = pts[dst] - pts[src]
    L = np.hypot(d[:, 0], d[:, 1])
    return graph_from_edges(len(pts), src, dst, d[:, 0] / L, d[:, 1] / L)


def concat_graphs(graphs, edge_keys=("w", "ex", "ey", "wt", "ext", "eyt")):
    """Block-diagonal union of several graph_arrays dicts (with the per-edge
    arrays ``edge_keys``)."""
    out = {k: [] for k in ("off", "col") + tuple(edge_keys)}
    n0 = e0 = 0
    for g in graphs:
        out["off"].append(g["off"][:-1] + e0)
        out["col"].append(g["col"] + n0)
        for k in edge_keys:
            out[k].append(g[k])
        n0 += g["n"]
        e0 += len(g["col"])
    res = {k: np.concatenate(v) for k, v in out.items()}
    res["off"] = np.concatenate([res["off"], [e0]]).astype(np.int32)
    res["n"] = n0
    return res


def graph_to_device(g, device="cuda"):
    torch = _torch()
    return {k: (torch.from_numpy(np.ascontiguousarray(v)).to(device) if isinstance(v, np.ndarray) else v)
            for k, v in g.items()}


SLABS = 32


def graph_from_edges(n, src, dst, ex, ey, paired=False):
    """CSR graph from directed edges (any order): rows sorted by (source,
    target); weights 1/deg(source); transpose arrays for the backward pass
    from the reverse edge (which must exist). With ``paired``, edges 2k and
    2k + 1 are each other's reverse, which spares the search for it."""
    src = np.asarray(src, np.int64)
    dst = np.asarray(dst, np.int64)
    ex = np.asarray(ex, np.float64)
    ey = np.asarray(ey, np.float64)
    key = src * n + dst
    order = np.argsort(key, kind="stable")
    key = key[order]
    if paired:
        pos = np.empty_like(order)
        pos[order] = np.arange(len(order))
        rev = pos[order ^ 1]
    else:
        rkey = dst[order] * n + src[order]
        rev = np.searchsorted(key, rkey)
        if not np.array_equal(key[np.minimum(rev, len(key) - 1)], rkey):
            raise ValueError("graph edges must come in both directions")
    src, dst, ex, ey = src[order], dst[order], ex[order], ey[order]
    deg = np.bincount(src, minlength=n)
    off = np.concatenate([[0], np.cumsum(deg)]).astype(np.int32)
    return dict(off=off, col=dst.astype(np.int32), w=(1.0 / deg[src]).astype(np.float32),
                ex=ex.astype(np.float32), ey=ey.astype(np.float32),
                wt=(1.0 / deg[dst]).astype(np.float32), ext=ex[rev].astype(np.float32),
                eyt=ey[rev].astype(np.float32), n=int(n))


def _interleave(u, v, ex, ey):
    """Directed pairs (u -> v, v -> u) of undirected edges, as arrays in
    the order 2k, 2k + 1."""
    m = len(u)
    src = np.empty(2 * m, np.int64)
    dst = np.empty(2 * m, np.int64)
    dx = np.empty(2 * m)
    dy = np.empty(2 * m)
    src[0::2], src[1::2] = u, v
    dst[0::2], dst[1::2] = v, u
    dx[0::2], dx[1::2] = ex, -np.asarray(ex, np.float64)
    dy[0::2], dy[1::2] = ey, -np.asarray(ey, np.float64)
    return src, dst, dx, dy


SECTION_SHIFT = np.array([0.95, 1.06, 0.5, 0.26], np.float32)
SECTION_SCALE = np.array([0.085, 0.116, 0.31, 0.106], np.float32)
# optional column 15: the mesh resolution, (log(cells) - shift) / scale, the
# same on every node of a case. First-order upwind fluxes add a numerical
# viscosity that grows with the cell size, so the steady state of one
# geometry depends on the resolution it is meshed at.
RESOLUTION_COLUMN = 15
RESOLUTION_SHIFT = 11.0   # log(60k cells)
RESOLUTION_SCALE = 0.7


SECTION_GRID = 2048  # x resolution of porous.section_features


def case_inputs(case, grid=(SLABS, 1), resolution=False):
    """Scaled and padded node inputs [N + GX*GY + 1, 16] of a porous.Case
    (mesh nodes first, then the virtual grid nodes in row-major order, then
    the global node), its graph, and the number of mesh nodes.

    The virtual grid covers the channel with GX x GY cells: each mesh node
    is linked both ways to the node of its grid cell, grid nodes to their
    8 neighbours (with the unit direction between centres), and the global
    node to every grid node. GY = 1 gives bands of x ("slabs"). Columns
    0-8: the node features, 9-10: virtual-node flags (grid, global), 11-14:
    porous.section_features; with ``resolution``, column 15: the mesh
    resolution (RESOLUTION_COLUMN). Built by the library's host code
    (fvdbm_nn_case_features, fvdbm_nn_case_graph), which releases the GIL,
    so threads prepare several cases at once; ``case_inputs_numpy`` is the
    same construction in NumPy."""
    L = _lib.lib()
    pts = np.ascontiguousarray(case.mesher.points, np.float64)
    markers = np.ascontiguousarray(case.mesher.point_markers, np.int32)
    codes = np.asarray(porous.MARKER_COLUMNS, np.int32)
    discs = np.ascontiguousarray(np.asarray(case.discs, np.float64).reshape(-1, 3))
    GX, GY = int(grid[0]), int(grid[1])
    N = len(pts)
    n = N + GX * GY + 1
    x = np.empty((n, IN_PAD), np.float32)
    rc = L.fvdbm_nn_case_features(_lib.ptr(pts), N, _lib.ptr(markers), _lib.ptr(codes), _lib.ptr(discs),
                                  len(discs), float(case.u_in), float(case.dynamics.tau), GX, GY,
                                  porous.HEIGHT, SECTION_GRID, _lib.ptr(FEATURE_SHIFT),
                                  _lib.ptr(FEATURE_SCALE), _lib.ptr(SECTION_SHIFT),
                                  _lib.ptr(SECTION_SCALE), IN_PAD, _lib.ptr(x))
    if rc:
        raise ValueError(L.fvdbm_nn_graph_error().decode())
    if resolution:
        x[:, RESOLUTION_COLUMN] = (np.log(case.n_cells) - RESOLUTION_SHIFT) / RESOLUTION_SCALE
    g = case_graph(pts, porous.mesh_edges(case), grid)
    return x, g, N


def case_inputs_numpy(case, grid=(SLABS, 1)):
    """``case_inputs`` built with NumPy (the check of the native builder;
    features equal to float32 rounding, graph bitwise)."""
    f = porous.node_features(case)
    pts = np.asarray(case.mesher.points, np.float64)
    N = f.shape[0]
    GX, GY = grid
    S = GX * GY
    n = N + S + 1
    x = np.zeros((n, IN_PAD), np.float32)
    x[:N, :f.shape[1]] = (f - FEATURE_SHIFT) / FEATURE_SCALE
    gi, gj = np.meshgrid(np.arange(GX), np.arange(GY), indexing="xy")
    gi, gj = gi.ravel(), gj.ravel()  # row-major: index = j * GX + i
    cx = (gi + 0.5) / GX
    cy = (gj + 0.5) / GY * porous.HEIGHT
    vx = np.concatenate([pts[:, 0], cx, [0.5]])
    sf = porous.section_features(np.stack([vx, np.zeros_like(vx)], 1), case.discs, grid=SECTION_GRID)
    x[:, 11:15] = (sf - SECTION_SHIFT) / SECTION_SCALE
    virt = np.zeros((S + 1, f.shape[1]), np.float32)
    virt[:S, 0] = cx
    virt[S, 0] = 0.5
    virt[:, 1] = 0.25
    virt[:, 2] = FEATURE_SHIFT[2]
    if GY > 1:
        virt[:S, 1] = cy
        virt[:S, 2] = porous.signed_distance(np.stack([cx, cy], 1), case.discs)
    virt[:, 7] = case.u_in
    virt[:, 8] = case.dynamics.tau
    x[N:, :f.shape[1]] = (virt - FEATURE_SHIFT) / FEATURE_SCALE
    x[N:N + S, 9] = 1.0
    x[N + S, 10] = 1.0
    g = case_graph_numpy(pts, porous.mesh_edges(case), grid)
    return x, g, N


def case_graph_numpy(points, edges, grid):
    """The graph of ``case_graph`` built with NumPy (sort of the directed
    edges by (source, target)); kept as the check of the native builder."""
    pts = np.asarray(points, np.float64)
    N = len(pts)
    GX, GY = grid
    S = GX * GY
    n = N + S + 1
    gi, gj = np.meshgrid(np.arange(GX), np.arange(GY), indexing="xy")
    gi, gj = gi.ravel(), gj.ravel()
    # undirected edges: mesh faces, mesh node - its grid cell, grid cell -
    # 4 of its 8 neighbours (the other 4 are the reverses), grid - global
    us, vs, exs, eys = [], [], [], []
    e = np.asarray(edges).astype(np.int64)
    d = pts[e[:, 1]] - pts[e[:, 0]]
    L = np.hypot(d[:, 0], d[:, 1])
    us.append(e[:, 0]); vs.append(e[:, 1]); exs.append(d[:, 0] / L); eys.append(d[:, 1] / L)
    ci = np.minimum((pts[:, 0] * GX).astype(np.int64), GX - 1)
    cj = np.minimum((pts[:, 1] / porous.HEIGHT * GY).astype(np.int64), GY - 1)
    us.append(np.arange(N)); vs.append(cj * GX + ci + N); exs.append(np.zeros(N)); eys.append(np.zeros(N))
    for di, dj in ((1, 0), (1, 1), (0, 1), (-1, 1)):
        ok = (gi + di >= 0) & (gi + di < GX) & (gj + dj >= 0) & (gj + dj < GY)
        a = (gj * GX + gi)[ok] + N
        b = ((gj + dj) * GX + gi + di)[ok] + N
        dx, dy = di / GX, dj / GY * porous.HEIGHT
        L2 = np.hypot(dx, dy)
        us.append(a); vs.append(b); exs.append(np.full(len(a), dx / L2)); eys.append(np.full(len(a), dy / L2))
    us.append(np.arange(S) + N); vs.append(np.full(S, N + S)); exs.append(np.zeros(S)); eys.append(np.zeros(S))
    src, dst, dx, dy = _interleave(np.concatenate(us), np.concatenate(vs), np.concatenate(exs),
                                   np.concatenate(eys))
    return graph_from_edges(n, src, dst, dx, dy, paired=True)


def case_graph(points, edges, grid):
    """Input graph of one case (mesh nodes ``points`` [N, 2], mesh edges
    [F, 2], virtual grid ``grid`` = (GX, GY) and the global node) as CSR
    arrays, built by the library's host code (fvdbm_nn_case_graph; it
    releases the GIL, so threads build several cases at once)."""
    L = _lib.lib()
    pts = np.ascontiguousarray(points, np.float64)
    e = np.ascontiguousarray(edges, np.int32)
    GX, GY = int(grid[0]), int(grid[1])
    N = len(pts)
    E = int(L.fvdbm_nn_graph_edges(N, len(e), GX, GY))
    if E < 0:
        raise ValueError("bad graph sizes")
    n = N + GX * GY + 1
    g = dict(off=np.empty(n + 1, np.int32), col=np.empty(E, np.int32))
    for k in ("w", "ex", "ey", "wt", "ext", "eyt"):
        g[k] = np.empty(E, np.float32)
    rc = L.fvdbm_nn_case_graph(_lib.ptr(pts), N, _lib.ptr(e), len(e), GX, GY, porous.HEIGHT,
                               *(_lib.ptr(g[k]) for k in ("off", "col", "w", "ex", "ey", "wt", "ext",
                                                          "eyt")))
    if rc:
        raise ValueError(L.fvdbm_nn_graph_error().decode())
    g["n"] = n
    return g


# ---------------------------------------------------------------------------
# autograd functions over the HIP kernels
# ---------------------------------------------------------------------------
def _aligned(t):
    """Contiguous float32 with a 16-byte aligned start (the kernels read and
    write rows as float4)."""
    t = t.contiguous()
    return t if t.data_ptr() % 16 == 0 else t.clone()


def _workspace(torch, m, n, like):
    chunks = max(1, -(-m // _lib.lib().fvdbm_nn_chunk_rows()))
    return torch.empty(chunks * n, dtype=torch.float32, device=like.device)


def _make_functions():
    import torch

    class MLP2(torch.autograd.Function):
        """y = silu(x W1 + b1) W2 + b2 (fvdbm_nn_mlp_forward / _backward,
        fvdbm_nn_wgrad, fvdbm_nn_colsum)."""

        @staticmethod
        def forward(ctx, x, w1, b1, w2, b2):
            x = _aligned(x)
            M, K = x.shape
            NO = w2.shape[1]
            y = torch.empty((M, NO), dtype=torch.float32, device=x.device)
            _check(_lib.lib().fvdbm_nn_mlp_forward(_p(x), M, K, _p(w1), _p(b1), _p(w2), _p(b2), NO,
                                                   _p(y), _stream()), "fvdbm_nn_mlp_forward")
            ctx.save_for_backward(x, w1, b1, w2)
            return y

        @staticmethod
        def backward(ctx, dy):
            x, w1, b1, w2 = ctx.saved_tensors
            dy = _aligned(dy)
            M, K = x.shape
            NO = w2.shape[1]
            dev = x.device
            dx = torch.empty_like(x) if ctx.needs_input_grad[0] else None
            dz = torch.empty((M, HID), dtype=torch.float32, device=dev)
            h = torch.empty((M, HID), dtype=torch.float32, device=dev)
            lib = _lib.lib()
            _check(lib.fvdbm_nn_mlp_backward(_p(x), M, K, _p(w1), _p(b1), _p(w2), NO, _p(dy),
                                             _p(dx) if dx is not None else None, _p(dz), _p(h),
                                             _stream()), "fvdbm_nn_mlp_backward")
            ws = _workspace(torch, M, max(K * HID, HID * NO), x)
            dw1 = torch.empty((K, HID), dtype=torch.float32, device=dev)
            dw2 = torch.empty((HID, NO), dtype=torch.float32, device=dev)
            db1 = torch.empty(HID, dtype=torch.float32, device=dev)
            db2 = torch.empty(NO, dtype=torch.float32, device=dev)
            s = _stream()
            _check(lib.fvdbm_nn_wgrad(_p(x), _p(dz), M, K, HID, _p(ws), _p(dw1), s), "fvdbm_nn_wgrad")
            _check(lib.fvdbm_nn_wgrad(_p(h), _p(dy), M, HID, NO, _p(ws), _p(dw2), s), "fvdbm_nn_wgrad")
            _check(lib.fvdbm_nn_colsum(_p(dz), M, HID, _p(ws), _p(db1), s), "fvdbm_nn_colsum")
            _check(lib.fvdbm_nn_colsum(_p(dy), M, NO, _p(ws), _p(db2), s), "fvdbm_nn_colsum")
            return dx, dw1, db1, dw2, db2

    class Aggregate(torch.autograd.Function):
        """z = [h | mean h_j | mean ex h_j | mean ey h_j] (fvdbm_nn_aggregate
        and its transpose)."""

        @staticmethod
        def forward(ctx, h, g):
            h = h.contiguous()
            n = h.shape[0]
            z = torch.empty((n, 4 * HID), dtype=torch.float32, device=h.device)
            _check(_lib.lib().fvdbm_nn_aggregate(_p(h), n, _p(g["off"]), _p(g["col"]), _p(g["w"]),
                                                 _p(g["ex"]), _p(g["ey"]), _p(z), _stream()),
                   "fvdbm_nn_aggregate")
            ctx.g = g
            return z

        @staticmethod
        def backward(ctx, dz):
            g = ctx.g
            dz = dz.contiguous()
            n = dz.shape[0]
            dh = torch.empty((n, HID), dtype=torch.float32, device=dz.device)
            _check(_lib.lib().fvdbm_nn_aggregate_backward(_p(dz), n, _p(g["off"]), _p(g["col"]),
                                                          _p(g["wt"]), _p(g["ext"]), _p(g["eyt"]),
                                                          _p(dh), _stream()),
                   "fvdbm_nn_aggregate_backward")
            return dh, None

    return MLP2, Aggregate


_FN = None


def functions():
    """(MLP2, Aggregate) autograd functions."""
    global _FN
    if _FN is None:
        _FN = _make_functions()
    return _FN


def mlp2(x, w1, b1, w2, b2):
    return functions()[0].apply(x, w1, b1, w2, b2)


def aggregate(h, g):
    return functions()[1].apply(h, g)


# ---------------------------------------------------------------------------
# the network
# ---------------------------------------------------------------------------
def init_params(seed=0, layers=LAYERS):
    """Float32 master weights as a flat dict of NumPy arrays: enc (16->64->64),
    proc.l (256->64->64, l < layers), dec (64->64->16). W ~ N(0, 1/fan_in);
    the last layer of each residual block starts 10x smaller."""
    rng = np.random.default_rng(seed)
    out = {}

    def block(name, k, n, last_scale=1.0):
        out[f"{name}.w1"] = (rng.standard_normal((k, HID)) / math.sqrt(k)).astype(np.float32)
        out[f"{name}.b1"] = np.zeros(HID, np.float32)
        out[f"{name}.w2"] = (last_scale * rng.standard_normal((HID, n)) / math.sqrt(HID)).astype(np.float32)
        out[f"{name}.b2"] = np.zeros(n, np.float32)

    block("enc", IN_PAD, HID)
    for l in range(layers):
        block(f"proc{l}", 4 * HID, HID, 0.1)
    block("dec", HID, OUT_PAD)
    return out


def _block_names(layers):
    return ["enc"] + [f"proc{l}" for l in range(layers)] + ["dec"]


class MeshNet:
    """The surrogate network; parameters are torch tensors on the GPU."""

    def __init__(self, params=None, layers=LAYERS, seed=0, target_scale=(1.0, 1.0, 1.0),
                 grid=(SLABS, 1), resolution=False):
        torch = _torch()
        self.layers = layers
        self.grid = tuple(int(v) for v in (params["grid"] if params is not None and "grid" in params
                                           else grid))
        # whether the network reads the mesh-resolution input column
        self.resolution = bool(params["resolution"]) if params is not None and "resolution" in params \
            else bool(resolution)
        p = init_params(seed, layers) if params is None else params
        self.params = {k: torch.tensor(np.asarray(v, np.float32), device="cuda", requires_grad=True)
                       for k, v in p.items() if "." in k}
        self.target_scale = np.asarray(
            params.get("target_scale", target_scale) if params is not None else target_scale, np.float32)
        self.calib = ({k: np.asarray(params[k], np.float64) for k in CALIB_KEYS}
                      if params is not None and all(k in params for k in CALIB_KEYS) else None)

    def parameters(self):
        return list(self.params.values())

    def block(self, name):
        p = self.params
        return p[f"{name}.w1"], p[f"{name}.b1"], p[f"{name}.w2"], p[f"{name}.b2"]

    def forward(self, x, g):
        """x [N,16] float32 on the GPU, g a device graph -> [N,3] (scaled targets)."""
        h = mlp2(x, *self.block("enc"))
        for l in range(self.layers):
            h = h + mlp2(aggregate(h, g), *self.block(f"proc{l}"))
        return mlp2(h, *self.block("dec"))[:, :OUT]

    __call__ = forward

    def state(self):
        out = {k: v.detach().cpu().numpy() for k, v in self.params.items()}
        out["target_scale"] = self.target_scale
        out["layers"] = np.int32(self.layers)
        out["grid"] = np.array(self.grid, np.int32)
        out["resolution"] = np.int32(self.resolution)
        if self.calib is not None:
            out.update(self.calib)
        return out

    def save(self, path):
        np.savez(path, **self.state())

    @classmethod
    def load(cls, path):
        z = np.load(path)
        params = {k: z[k] for k in z.files if k != "record"}
        return cls(params, layers=int(z["layers"]))

    def predict(self, cases, project=False, chunk=64, calibrate=True):
        """Node (rho, u) of each porous.Case: list of (rho [N], u [N,2]);
        with ``project``, u_x passes through ``project_mass_flux``; with
        ``calibrate`` and calibration arrays in the weights, each case's
        rho - 1 and u are rescaled by ``apply_calibration``. Cases go
        through the network ``chunk`` at a time; their inputs are prepared
        by a thread pool (the native builders release the GIL) while the
        previous chunk runs on the GPU."""
        torch = _torch()
        import os
        from concurrent.futures import ThreadPoolExecutor
        out = []
        with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as pool:
            futs = [pool.submit(case_inputs, c, self.grid, self.resolution) for c in cases]
            for c0 in range(0, len(cases), chunk):
                ins = [f.result() for f in futs[c0:c0 + chunk]]
                g = graph_to_device(concat_graphs([i[1] for i in ins], edge_keys=("w", "ex", "ey")))
                x = torch.from_numpy(np.concatenate([i[0] for i in ins])).cuda()
                with torch.no_grad():
                    y = self.forward(x, g).cpu().numpy().astype(np.float64) * self.target_scale
                o = 0
                for c, (_, gi, N) in zip(cases[c0:c0 + chunk], ins):
                    yy = y[o:o + N] * c.u_in
                    o += gi["n"]
                    r = (1.0 + yy[:, 0]).astype(np.float32)
                    v = yy[:, 1:3].astype(np.float32)
                    if calibrate and self.calib is not None:
                        r, v = apply_calibration(self.calib, c, r, v)
                    out.append((r, project_mass_flux(c, r, v) if project else v))
        return out


def case_features(case, rho, vel):
    """The CALIB_FEATURES of a porous.Case and its predicted node (rho, u)
    (speeds and rho - 1 relative to the inlet speed)."""
    pts = np.asarray(case.mesher.points, np.float64)
    width = pts[:, 0].max() - pts[:, 0].min()
    height = pts[:, 1].max() - pts[:, 1].min()
    d = np.asarray(case.discs, np.float64).reshape(-1, 3)
    p = np.asarray(rho, np.float64) - 1.0
    q = np.asarray(vel, np.float64).reshape(-1, 2)
    inlet = case.mesher.point_markers == porous.meshgen.INLET
    u_in = float(case.u_in)
    tau = float(case.dynamics.tau)
    cells = float(case.n_cells)
    inlet_rho = float(p[inlet].mean() / u_in) if inlet.any() else 0.0
    r = d[:, 2] if len(d) else np.zeros(1)
    return np.array([len(d), np.pi * (d[:, 2] ** 2).sum() / (width * height), r.mean(), r.max(), tau,
                     u_in, cells, inlet_rho, p.mean() / u_in, np.abs(q).max() / u_in, np.log(cells),
                     inlet_rho * tau])


def calibration_factors(calib, features):
    """(rho - 1 factor, u factor) of a feature vector: exp of the ridge
    models on the standardised features, clipped to CALIB_CLIP; (1, 1) for
    a feature vector that is not finite."""
    z = (np.asarray(features, np.float64) - calib["calib_mu"]) / calib["calib_sd"]
    if not np.isfinite(z).all():
        return 1.0, 1.0  # a non-finite prediction is left as it is (the warm start rejects it)
    fr = math.exp(float(z @ calib["calib_w_rho"] + calib["calib_b_rho"]))
    fu = math.exp(float(z @ calib["calib_w_u"] + calib["calib_b_u"]))
    return (min(max(fr, CALIB_CLIP[0]), CALIB_CLIP[1]), min(max(fu, CALIB_CLIP[0]), CALIB_CLIP[1]))


def apply_calibration(calib, case, rho, vel):
    """rho and u of one case's prediction rescaled by its calibration
    factors: rho = 1 + f_rho (rho - 1), u = f_u u (float64, stored float32)."""
    fr, fu = calibration_factors(calib, case_features(case, rho, vel))
    r = (1.0 + fr * (np.asarray(rho, np.float64) - 1.0)).astype(np.float32)
    v = (fu * np.asarray(vel, np.float64)).astype(np.float32)
    return r, v


def project_mass_flux(case, rho, vel, bins=64):
    """Rescale a predicted u_x band by band along x so that each band
    carries the inlet's mass flux: for the parabolic inlet the flux is
    (2/3) u_in HEIGHT times the mean inlet density, and a band's flux is
    estimated as the mean of rho u_x over its nodes times its open height
    a(x) HEIGHT. The per-band factors (clipped to [0.5, 2]) are interpolated
    linearly in x between band centres. Returns the corrected velocity."""
    pts = np.asarray(case.mesher.points, np.float64)
    rho = np.asarray(rho, np.float64)
    vel = np.array(vel, np.float64)
    inlet = case.mesher.point_markers == porous.meshgen.INLET
    q = (2.0 / 3.0) * case.u_in * porous.HEIGHT * (rho[inlet].mean() if inlet.any() else 1.0)
    k = np.minimum((pts[:, 0] * bins).astype(np.int64), bins - 1)
    centres = (np.arange(bins) + 0.5) / bins
    a = porous.section_features(np.stack([centres, np.zeros(bins)], 1), case.discs)[:, 0]
    num = np.bincount(k, weights=rho * vel[:, 0], minlength=bins)
    cnt = np.bincount(k, minlength=bins)
    flux = np.where(cnt > 0, num / np.maximum(cnt, 1), 0.0) * a * porous.HEIGHT
    fac = np.where(flux > 0, q / np.where(flux > 0, flux, 1.0), 1.0)
    fac = np.clip(fac, 0.5, 2.0)
    vel[:, 0] *= np.interp(pts[:, 0], centres, fac)
    return vel.astype(np.float32)


def targets(rho, vel, u_in, scale, n_virtual=0):
    """Training targets [(rho - 1), ux, uy] / u_in / scale per component,
    with zero rows for the virtual nodes."""
    t = np.stack([np.asarray(rho, np.float64) - 1.0, vel[:, 0], vel[:, 1]], axis=1)
    t = t / u_in / np.asarray(scale, np.float64)
    return np.concatenate([t, np.zeros((n_virtual, 3))]).astype(np.float32)


def device_sample(x, g, y, n_mesh=None):
    """Upload one (x, graph, y) sample to the GPU once; rows from n_mesh on
    (virtual nodes) are masked out of the loss."""
    torch = _torch()
    d = graph_to_device(g)
    d["x"] = torch.from_numpy(np.ascontiguousarray(x)).cuda()
    d["y"] = torch.from_numpy(np.ascontiguousarray(y)).cuda()
    m = np.zeros((g["n"], 1), np.float32)
    m[:g["n"] if n_mesh is None else n_mesh] = 1.0
    d["mask"] = torch.from_numpy(m).cuda()
    return d


def concat_device(samples):
    """Block-diagonal batch of device samples: (x, graph, y)."""
    torch = _torch()
    offs, cols = [], []
    n0 = e0 = 0
    for s in samples:
        offs.append(s["off"][:-1] + e0)
        cols.append(s["col"] + n0)
        n0 += s["n"]
        e0 += int(s["col"].shape[0])
    dev = samples[0]["x"].device
    g = {"off": torch.cat(offs + [torch.tensor([e0], dtype=torch.int32, device=dev)]),
         "col": torch.cat(cols), "n": n0}
    for k in ("w", "ex", "ey", "wt", "ext", "eyt"):
        g[k] = torch.cat([s[k] for s in samples])
    return (torch.cat([s["x"] for s in samples]), g, torch.cat([s["y"] for s in samples]),
            torch.cat([s["mask"] for s in samples]))


def train(net, samples, epochs=50, batch=16, lr=1e-3, seed=0, log=None, lr_final=1e-5,
          weights=(1.0, 1.0, 1.0)):
    """Adam on the mean squared error over mesh nodes (per-component
    ``weights`` on the three scaled targets, normalised to a mean of 1),
    learning rate from ``lr`` to ``lr_final`` on a cosine over the steps.
    samples: list of (x, graph, y[, n_mesh]) with y already scaled, or of
    ``device_sample`` dicts. Returns the per-epoch mean loss."""
    torch = _torch()
    samples = [s if isinstance(s, dict) else device_sample(*s) for s in samples]
    wq = np.asarray(weights, np.float64)
    wq = torch.tensor(wq / wq.mean(), dtype=torch.float32, device="cuda")
    opt = torch.optim.Adam(net.parameters(), lr=lr)
    per_epoch = -(-len(samples) // batch)
    total = max(1, epochs * per_epoch)
    rng = np.random.default_rng(seed)
    hist = []
    it = 0
    for ep in range(epochs):
        order = rng.permutation(len(samples))
        losses = []
        t_ep = time.perf_counter()
        for b0 in range(0, len(order), batch):
            for grp in opt.param_groups:
                grp["lr"] = lr_final + 0.5 * (lr - lr_final) * (1.0 + math.cos(math.pi * it / total))
            x, g, y, m = concat_device([samples[i] for i in order[b0:b0 + batch]])
            opt.zero_grad(set_to_none=True)
            loss = torch.sum(m * wq * (net(x, g) - y) ** 2) / (3.0 * torch.sum(m))
            loss.backward()
            opt.step()
            losses.append(loss.detach())
            it += 1
        hist.append(float(torch.stack(losses).mean()))
        if log:
            log(f"epoch {ep + 1}/{epochs} loss {hist[-1]:.4e} "
                f"({time.perf_counter() - t_ep:.1f} s)")
    return hist


def evaluate(net, samples):
    """Relative L2 error of the predicted (rho - 1, ux, uy) on the mesh nodes
    of each sample, as an array [len(samples), 3]."""
    torch = _torch()
    samples = [s if isinstance(s, dict) else device_sample(*s) for s in samples]
    out = []
    with torch.no_grad():
        for s in samples:
            m = s["mask"][:, 0] > 0
            p = net(s["x"], s)[m]
            t = s["y"][m]
            out.append((torch.linalg.norm(p - t, dim=0) / torch.linalg.norm(t, dim=0)).cpu().numpy())
    return np.array(out)
