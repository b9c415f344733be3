"""CPU twin of the surrogate network (PyTorch on the CPU, float32 or
float64).

TEST INFRASTRUCTURE ONLY: the checker for the fvdbm_nn_* kernels and
surrogate.MeshNet. Only tests/ may import it.

The same architecture as the product network: mlp(x) = silu(x W1 + b1) W2
+ b2, aggregation z = [h | sum w h_j | sum w ex h_j | sum w ey h_j] over the
CSR node graph. With ``bf16=True`` the operands of every product are rounded
to bfloat16 (round to nearest even) as the MFMA kernels do, while sums stay
in the working precision.
"""
from __future__ import annotations

import numpy as np
import torch


def _round_bf16(t):
    return t.to(torch.bfloat16).to(t.dtype)


def mlp2(x, w1, b1, w2, b2, bf16=False):
    r = _round_bf16 if bf16 else (lambda t: t)
    z = r(x) @ r(w1) + b1
    h = z * torch.sigmoid(z)
    return r(h) @ r(w2) + b2


def aggregate(h, g):
    """g: graph_arrays dict (NumPy)."""
    src = torch.from_numpy(np.repeat(np.arange(g["n"]), np.diff(g["off"])))
    col = torch.from_numpy(g["col"].astype(np.int64))
    w = torch.from_numpy(g["w"]).to(h.dtype)[:, None]
    ex = torch.from_numpy(g["ex"]).to(h.dtype)[:, None]
    ey = torch.from_numpy(g["ey"]).to(h.dtype)[:, None]
    v = h[col] * w
    out = []
    for term in (v, v * ex, v * ey):
        acc = torch.zeros_like(h)
        out.append(acc.index_add(0, src, term))
    return torch.cat([h] + out, dim=1)


def forward(params, x, g, layers, bf16=False):
    """params: dict name -> tensor (enc.*, procL.*, dec.*); returns [N,3]."""
    def blk(n):
        return params[f"{n}.w1"], params[f"{n}.b1"], params[f"{n}.w2"], params[f"{n}.b2"]
    h = mlp2(x, *blk("enc"), bf16=bf16)
    for l in range(layers):
        h = h + mlp2(aggregate(h, g), *blk(f"proc{l}"), bf16=bf16)
    return mlp2(h, *blk("dec"), bf16=bf16)[:, :3]


def params_from_numpy(p, dtype=torch.float64, requires_grad=True):
    return {k: torch.tensor(np.asarray(v), dtype=dtype, requires_grad=requires_grad)
            for k, v in p.items() if "." in k}
