"""Surrogate kernels on the GPU against the CPU twin (oracle/surrogate_twin.py):

* each dense layer (all four width pairs) and the aggregation, forward;
* their backward passes against the twin's autograd (float32) and the whole
  network's gradient against central finite differences of the float64
  twin along each parameter block's gradient direction;
* determinism (no atomics: two backward passes are bitwise equal).

Tolerances: the kernels round every product operand to bfloat16 (8-bit
significand, relative rounding 2^-9), so a dense layer agrees with the
bf16-emulating twin to 2e-3 (relative to the largest output) and with plain
float32 to 2e-2; the aggregation has no bf16 step and agrees to 1e-5.
"""
import importlib

import numpy as np
import pytest

from cases import fv

pytestmark = pytest.mark.gpu
sg = importlib.import_module("fvdbm-jax_amd.surrogate")


def _torch():
    import torch
    return torch


def _twin():
    from oracle import surrogate_twin
    return surrogate_twin


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


@pytest.mark.parametrize("k,no", [(16, 64), (64, 64), (256, 64), (64, 16)])
@pytest.mark.parametrize("m", [1, 37, 5000])
def test_mlp_forward_backward(gpu, k, no, m):
    torch = _torch()
    tw = _twin()
    rng = np.random.default_rng(k * 7 + no + m)
    x = rng.standard_normal((m, k)).astype(np.float32)
    w1 = (rng.standard_normal((k, 64)) / np.sqrt(k)).astype(np.float32)
    b1 = (0.1 * rng.standard_normal(64)).astype(np.float32)
    w2 = (rng.standard_normal((64, no)) / 8).astype(np.float32)
    b2 = (0.1 * rng.standard_normal(no)).astype(np.float32)
    dy = rng.standard_normal((m, no)).astype(np.float32)
    dev = [torch.tensor(a, device="cuda", requires_grad=True) for a in (x, w1, b1, w2, b2)]
    y = sg.mlp2(*dev)
    y.backward(torch.tensor(dy, device="cuda"))
    got = [y.detach().cpu().numpy()] + [t.grad.cpu().numpy() for t in dev]

    cpu = [torch.tensor(a, requires_grad=True) for a in (x, w1, b1, w2, b2)]
    yb = tw.mlp2(*cpu, bf16=True)
    assert _rel(got[0], yb.detach().numpy()) < 2e-3
    y32 = tw.mlp2(*cpu)
    y32.backward(torch.tensor(dy))
    assert _rel(got[0], y32.detach().numpy()) < 2e-2
    for g, t in zip(got[1:], cpu):
        assert _rel(g, t.grad.numpy()) < 2e-2

    # deterministic: a second backward gives the same bits
    for t in dev:
        t.grad = None
    sg.mlp2(*dev).backward(torch.tensor(dy, device="cuda"))
    again = [t.grad.cpu().numpy() for t in dev]
    for a, b in zip(again, got[1:]):
        assert np.array_equal(a, b)


def _small_graph(seed=0):
    c = fv.porous.make_case(seed, target_cells=1500)
    x, g, nm = sg.case_inputs(c)
    return c, x, g


def test_aggregate_forward_backward(gpu):
    torch = _torch()
    tw = _twin()
    c, x, g = _small_graph()
    n = g["n"]
    rng = np.random.default_rng(1)
    h = rng.standard_normal((n, 64)).astype(np.float32)
    dz = rng.standard_normal((n, 256)).astype(np.float32)
    hd = torch.tensor(h, device="cuda", requires_grad=True)
    z = sg.aggregate(hd, sg.graph_to_device(g))
    z.backward(torch.tensor(dz, device="cuda"))
    hc = torch.tensor(h, requires_grad=True)
    zc = tw.aggregate(hc, g)
    zc.backward(torch.tensor(dz))
    assert _rel(z.detach().cpu().numpy(), zc.detach().numpy()) < 1e-5
    assert _rel(hd.grad.cpu().numpy(), hc.grad.numpy()) < 1e-5


def test_network_gradient_finite_differences(gpu):
    """d loss / d theta (loss = sum(y * r)): GPU kernels against the float32
    twin's autograd, block by block, and against central differences of
    the float64 twin along each block's gradient direction."""
    torch = _torch()
    tw = _twin()
    c, x, g = _small_graph(3)
    params = sg.init_params(seed=4, layers=2)
    net = sg.MeshNet(params, layers=2)
    r = np.random.default_rng(2).standard_normal((g["n"], 3))
    y = net(torch.tensor(x, device="cuda"), sg.graph_to_device(g))
    (y * torch.tensor(r, device="cuda", dtype=torch.float32)).sum().backward()
    grads = {k: v.grad.cpu().numpy().astype(np.float64) for k, v in net.params.items()}
    # the GPU gradient against the float32 twin's autograd, block by block
    p32 = tw.params_from_numpy(params, torch.float32)
    (tw.forward(p32, torch.tensor(x), g, 2) * torch.tensor(r, dtype=torch.float32)).sum().backward()
    for k, t in p32.items():
        a, b = grads[k], t.grad.numpy()
        assert np.linalg.norm(a - b) <= 3e-2 * np.linalg.norm(b) + 1e-6, k

    p64 = tw.params_from_numpy(params, torch.float64, requires_grad=False)
    xd = torch.tensor(x, dtype=torch.float64)
    rd = torch.tensor(r)

    def loss(pp):
        return float((tw.forward(pp, xd, g, 2) * rd).sum())

    # along each block's own gradient direction (unit norm), the float64
    # central difference must reproduce the GPU gradient's norm: a wrong
    # direction or scale in any block shows up as a smaller or larger slope
    for k in grads:
        nrm = np.linalg.norm(grads[k])
        if nrm == 0.0:
            continue
        d = grads[k] / nrm
        eps = 1e-3
        plus = dict(p64)
        minus = dict(p64)
        plus[k] = p64[k] + eps * torch.tensor(d)
        minus[k] = p64[k] - eps * torch.tensor(d)
        fd = (loss(plus) - loss(minus)) / (2 * eps)
        assert fd == pytest.approx(nrm, rel=3e-2), (k, nrm, fd)


def test_training_reduces_loss(gpu):
    """A few Adam steps on two small cases with a smooth synthetic target."""
    samples = []
    for s in range(2):
        c, x, g = _small_graph(s)
        p = c.mesher.points
        y = np.stack([np.sin(3 * p[:, 0]), p[:, 1] * 2, np.cos(2 * p[:, 0] + p[:, 1])], 1)
        y = np.concatenate([y, np.zeros((g["n"] - len(p), 3))])
        samples.append((x, g, y.astype(np.float32), len(p)))
    net = sg.MeshNet(sg.init_params(seed=0, layers=2), layers=2)
    hist = sg.train(net, samples, epochs=30, batch=2, lr=3e-3)
    assert hist[-1] < 0.2 * hist[0], hist


def test_loss_weights_train_only_the_weighted_targets(gpu):
    """train_surrogate.fit with loss weights (1, 0, 0): the first target's
    error falls far more than the two unweighted ones', and fit leaves the
    samples' unscaled targets untouched (it trains on scaled copies)."""
    torch = _torch()
    ts = importlib.import_module("fvdbm-jax_amd.train_surrogate")
    samples = []
    for s in range(4):
        c, x, g = _small_graph(s)
        p = c.mesher.points
        y = np.stack([np.sin(3 * p[:, 0]), p[:, 1] * 2, np.cos(2 * p[:, 0] + p[:, 1])], 1)
        y = np.concatenate([y, np.zeros((g["n"] - len(p), 3))]).astype(np.float32)
        samples.append(sg.device_sample(x, g, y, len(p)))
    before = [s["y"].clone() for s in samples]
    net0 = sg.MeshNet(seed=0)
    sc = torch.tensor([float(s) for s in np.std(np.concatenate(
        [b[s["mask"][:, 0] > 0].cpu().numpy() for b, s in zip(before[:-1], samples[:-1])]), axis=0)],
        device="cuda")
    e0 = sg.evaluate(net0, [dict(s, y=s["y"] / sc) for s in samples[:-1]]).mean(axis=0)
    net, rec = ts.fit(samples, 1, (32, 16), False, epochs=80, batch=3, lr=3e-3, lr_final=1e-3,
                      weights=(1.0, 0.0, 0.0), log=lambda m: None)
    for b, s in zip(before, samples):
        assert torch.equal(b, s["y"])
    e1 = sg.evaluate(net, [dict(s, y=s["y"] / sc) for s in samples[:-1]]).mean(axis=0)
    gain = e1 / e0
    print("relative L2 before", e0, "after", e1, flush=True)
    assert gain[0] < 0.5, (e0, e1)
    assert gain[0] < 0.7 * min(gain[1], gain[2]), (e0, e1)
    assert rec["loss_weights"] == [1.0, 0.0, 0.0]
