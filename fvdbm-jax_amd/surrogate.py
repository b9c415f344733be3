"""Warm-start surrogate: a mesh message-passing network that predicts the
steady (rho - 1, u) of a porous-channel case at every mesh node.

The network (``MeshNet``) works on the node graph of the mesh (its faces as
undirected edges):

  h0   = mlp_enc(x)                      x: 9 node features padded to 16
  z_l  = [h | mean_j h_j | mean_j ex_ij h_j | mean_j ey_ij h_j]     (256)
  h_l+1 = h_l + mlp_l(z_l)               l = 0..5
  y    = mlp_dec(h_6)[:, :3]             (rho - 1, ux, uy), normalised

with mlp(x) = silu(x W1 + b1) W2 + b2, hidden width 64, and (ex, ey) the
unit vector along edge i -> j. Every dense layer, its backward pass, the
weight gradients and the aggregation are this library's HIP kernels
(fvdbm_nn_*: bf16 MFMA operands, float32 accumulation, float32 master
weights) behind ``torch.autograd.Function``; PyTorch only holds the tensors,
runs Adam and the loss. There is no CPU path: without a GPU and the built
library the functions here raise.
"""
from __future__ import annotations

import ctypes as ct
import math

import numpy as np

from . import _lib, porous

HID = 64
IN_PAD = 16
OUT = 3
OUT_PAD = 16
LAYERS = 6

# fixed input scaling: (x - shift) / scale per feature column
FEATURE_SHIFT = np.array([0.5, 0.25, 0.05, 0, 0, 0, 0, 0.035, 0.725], np.float32)
FEATURE_SCALE = np.array([0.29, 0.144, 0.05, 1, 1, 1, 1, 0.01, 0.1], np.float32)


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
    """CSR node graph of a mesh: both directions of every edge, sorted by
    (source, target). Returns dict of NumPy arrays: off [N+1], col [E],
    w = 1/deg(source), (ex, ey) = unit(p_target - p_source), and for the
    backward pass the reverse edge's weight wt = 1/deg(target) and
    direction (ext, eyt) = -(ex, ey)."""
    pts = np.asarray(points, np.float64)
    e = np.asarray(edges, np.int64)
    e = np.concatenate([e, e[:, ::-1]])
    e = e[np.lexsort((e[:, 1], e[:, 0]))]
    src, dst = e[:, 0], e[:, 1]
    N = len(pts)
    deg = np.bincount(src, minlength=N)
    off = np.concatenate([[0], np.cumsum(deg)]).astype(np.int32)
    d = pts[dst] - pts[src]
    L = np.hypot(d[:, 0], d[:, 1])
    ex, ey = d[:, 0] / L, d[:, 1] / L
    return dict(off=off, col=dst.astype(np.int32), w=(1.0 / deg[src]).astype(np.float32),
                ex=ex.astype(np.float32), ey=ey.astype(np.float32),
                wt=(1.0 / deg[dst]).astype(np.float32), ext=(-ex).astype(np.float32),
                eyt=(-ey).astype(np.float32), n=N)


def concat_graphs(graphs):
    """Block-diagonal union of several graph_arrays dicts."""
    out = {k: [] for k in ("off", "col", "w", "ex", "ey", "wt", "ext", "eyt")}
    n0 = e0 = 0
    for g in graphs:
        out["off"].append(g["off"][:-1] + e0)
        out["col"].append(g["col"] + n0)
        for k in ("w", "ex", "ey", "wt", "ext", "eyt"):
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


def case_inputs(case):
    """Scaled and padded node inputs [N, 16] of a porous.Case, and its graph."""
    f = porous.node_features(case)
    x = np.zeros((f.shape[0], IN_PAD), np.float32)
    x[:, :f.shape[1]] = (f - FEATURE_SHIFT) / FEATURE_SCALE
    return x, graph_arrays(case.mesher.points, porous.mesh_edges(case))


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

    def __init__(self, params=None, layers=LAYERS, seed=0, target_scale=(1.0, 1.0, 1.0)):
        torch = _torch()
        self.layers = layers
        p = init_params(seed, layers) if params is None else params
        self.params = {k: torch.tensor(np.asarray(v, np.float32), device="cuda", requires_grad=True)
                       for k, v in p.items() if "." in k}
        self.target_scale = np.asarray(
            params.get("target_scale", target_scale) if params is not None else target_scale, np.float32)

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
        return out

    def save(self, path):
        np.savez(path, **self.state())

    @classmethod
    def load(cls, path):
        z = np.load(path)
        params = {k: z[k] for k in z.files}
        return cls(params, layers=int(z["layers"]))

    def predict(self, cases):
        """Node (rho, u) of each porous.Case: list of (rho [N], u [N,2])."""
        torch = _torch()
        xs, gs = zip(*[case_inputs(c) for c in cases])
        g = graph_to_device(concat_graphs(gs))
        x = torch.from_numpy(np.concatenate(xs)).cuda()
        with torch.no_grad():
            y = self.forward(x, g).cpu().numpy().astype(np.float64) * self.target_scale
        out, o = [], 0
        for c in cases:
            n = c.n_nodes
            yy = y[o:o + n]
            o += n
            out.append(((1.0 + yy[:, 0]).astype(np.float32), yy[:, 1:3].astype(np.float32)))
        return out


def targets(rho, vel, scale):
    """Scaled training targets [(rho - 1)/s0, ux/s1, uy/s2]."""
    t = np.stack([np.asarray(rho, np.float64) - 1.0, vel[:, 0], vel[:, 1]], axis=1)
    return (t / np.asarray(scale, np.float64)).astype(np.float32)


def train(net, samples, epochs=50, batch=16, lr=1e-3, seed=0, log=None):
    """Adam on the mean squared error over nodes. samples: list of
    (x [N,16], graph_arrays, y [N,3]) with y already scaled. Returns the
    per-epoch mean loss."""
    torch = _torch()
    opt = torch.optim.Adam(net.parameters(), lr=lr)
    rng = np.random.default_rng(seed)
    hist = []
    for ep in range(epochs):
        order = rng.permutation(len(samples))
        tot, cnt = 0.0, 0
        for b0 in range(0, len(order), batch):
            sel = [samples[i] for i in order[b0:b0 + batch]]
            x = torch.from_numpy(np.concatenate([s[0] for s in sel])).cuda()
            y = torch.from_numpy(np.concatenate([s[2] for s in sel])).cuda()
            g = graph_to_device(concat_graphs([s[1] for s in sel]))
            opt.zero_grad(set_to_none=True)
            loss = torch.mean((net(x, g) - y) ** 2)
            loss.backward()
            opt.step()
            tot += float(loss.detach()) * len(sel)
            cnt += len(sel)
        hist.append(tot / cnt)
        if log:
            log(f"epoch {ep + 1}/{epochs} loss {hist[-1]:.4e}")
    return hist
