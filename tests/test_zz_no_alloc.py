"""fvdbm_nn_wgrad_noalloc (the allocation-free entry point) against
fvdbm_nn_wgrad with a caller workspace, through the C ABI: bitwise equal
for every supported width pair, on two calls in a row (the second smaller,
so stale partials of the first call would show), plus the size limit and
the empty case."""
import ctypes as ct
import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

_lib = importlib.import_module("fvdbm-jax_amd._lib")

WIDTHS = [(16, 64), (64, 64), (256, 64), (64, 16)]
SCRATCH_FLOATS = 4 << 20


def _p(t):
    return ct.c_void_p(t.data_ptr())


def _stream(torch):
    return ct.c_void_p(torch.cuda.current_stream().cuda_stream)


def _nn_error(L):
    return L.fvdbm_nn_last_error().decode(errors="replace")


def _both(torch, L, a, b, ka, nb):
    m = a.shape[0]
    chunk = L.fvdbm_nn_chunk_rows()
    ws = torch.empty(max(1, -(-m // chunk)) * ka * nb, device="cuda", dtype=torch.float32)
    ref = torch.full((ka, nb), float("nan"), device="cuda", dtype=torch.float32)
    got = torch.full((ka, nb), float("nan"), device="cuda", dtype=torch.float32)
    s = _stream(torch)
    rc = L.fvdbm_nn_wgrad(_p(a), _p(b), m, ka, nb, _p(ws), _p(ref), s)
    assert rc == 0, _nn_error(L)
    rc = L.fvdbm_nn_wgrad_noalloc(_p(a), _p(b), m, ka, nb, _p(got), s)
    assert rc == 0, _nn_error(L)
    torch.cuda.synchronize()
    return ref, got


@pytest.mark.parametrize("ka,nb", WIDTHS)
def test_noalloc_wgrad_equals_workspace_wgrad(gpu, ka, nb):
    import torch
    L = _lib.lib()
    chunk = L.fvdbm_nn_chunk_rows()
    max_rows = (SCRATCH_FLOATS // (ka * nb)) * chunk
    g = torch.Generator(device="cuda").manual_seed(ka * 1000 + nb)
    # a large call (up to the limit, capped for time), then a smaller one
    for m in (min(max_rows, 200_003), 1_537):
        a = torch.randn(m, ka, device="cuda", generator=g)
        b = torch.randn(m, nb, device="cuda", generator=g)
        print(f"wgrad_noalloc ka={ka} nb={nb} m={m}", flush=True)
        ref, got = _both(torch, L, a, b, ka, nb)
        assert torch.isfinite(ref).all()
        assert torch.equal(ref, got), f"m={m}: max |diff| {(ref - got).abs().max().item()}"
        # the product itself, against float64 on the CPU (bf16 operands)
        exact = a.double().cpu().T @ b.double().cpu()
        err = (got.double().cpu() - exact).abs().max().item()
        assert err <= 4e-2 * np.sqrt(m), err


def test_noalloc_wgrad_limits(gpu):
    import torch
    L = _lib.lib()
    chunk = L.fvdbm_nn_chunk_rows()
    ka, nb = 256, 64
    max_rows = (SCRATCH_FLOATS // (ka * nb)) * chunk
    a = torch.randn(max_rows + 1, ka, device="cuda")
    b = torch.randn(max_rows + 1, nb, device="cuda")
    out = torch.full((ka, nb), 7.0, device="cuda")
    s = _stream(torch)
    rc = L.fvdbm_nn_wgrad_noalloc(_p(a), _p(b), max_rows + 1, ka, nb, _p(out), s)
    assert rc == _lib.ERR_INVALID
    assert f"at most {max_rows} rows" in _nn_error(L)
    torch.cuda.synchronize()
    assert (out == 7.0).all()  # nothing launched
    # exactly at the limit is accepted and equals the workspace path
    ref, got = _both(torch, L, a[:max_rows], b[:max_rows], ka, nb)
    assert torch.equal(ref, got)
    # unsupported widths and the empty case
    assert L.fvdbm_nn_wgrad_noalloc(_p(a), _p(b), 10, 32, 64, _p(out), s) == _lib.ERR_INVALID
    rc = L.fvdbm_nn_wgrad_noalloc(_p(a), _p(b), 0, ka, nb, _p(out), s)
    assert rc == 0, _nn_error(L)
    torch.cuda.synchronize()
    assert (out == 0).all()
