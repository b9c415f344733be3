"""Host parts of the surrogate, on the CPU: the node graph built for the
aggregation kernels, the CPU twin's aggregation against a dense adjacency
construction, and the twin's bfloat16 rounding against the kernels'
round-to-nearest-even bit formula."""
import importlib

import numpy as np
import torch

from cases import fv
from oracle import surrogate_twin as tw

sg = importlib.import_module("fvdbm-jax_amd.surrogate")


def _case():
    c = fv.porous.make_case(2, target_cells=1200)
    return c, sg.graph_arrays(c.mesher.points, c.mesher.faces)


def test_graph_arrays():
    c, g = _case()
    N = c.n_nodes
    off, col = g["off"], g["col"]
    src = np.repeat(np.arange(N), np.diff(off))
    pairs = set(zip(src.tolist(), col.tolist()))
    assert all((j, i) in pairs for i, j in pairs)            # symmetric
    assert len(pairs) == 2 * len(c.mesher.faces)             # every face, both ways
    assert np.all(np.diff(col.astype(np.int64) + src.astype(np.int64) * N) > 0)  # sorted, unique
    wsum = np.add.reduceat(g["w"].astype(np.float64), off[:-1])
    assert np.allclose(wsum, 1.0, atol=1e-6)                 # mean weights
    assert np.allclose(g["ex"] ** 2 + g["ey"] ** 2, 1.0, atol=1e-6)
    # the transpose arrays describe the reverse edge
    deg = np.diff(off)
    assert np.allclose(g["wt"], 1.0 / deg[col]) and np.array_equal(g["ext"], -g["ex"])


def test_virtual_nodes():
    c, _ = _case()
    x, g, N = sg.case_inputs(c)
    S = sg.SLABS
    assert N == c.n_nodes and g["n"] == N + S + 1 and x.shape == (N + S + 1, 16)
    assert np.all(x[:N, 9:11] == 0) and np.all(x[N:N + S, 9] == 1) and x[N + S, 10] == 1
    assert np.all(x[:, 15] == 0)  # the last input column is padding
    off, col = g["off"], g["col"]
    deg = np.diff(off)
    # each mesh node: its mesh neighbours plus its slab
    mesh_deg = np.bincount(np.asarray(c.mesher.faces).ravel(), minlength=N)
    assert np.array_equal(deg[:N], mesh_deg + 1)
    # slab k holds the nodes with floor(x * S) == k, plus neighbour slabs and the global node
    k = np.minimum((c.mesher.points[:, 0] * S).astype(int), S - 1)
    cnt = np.bincount(k, minlength=S)
    nb = np.full(S, 2)
    nb[0] = nb[-1] = 1
    assert np.array_equal(deg[N:N + S], cnt + nb + 1)
    assert deg[N + S] == S
    # transpose arrays: the weight and direction of each reverse edge
    src = np.repeat(np.arange(g["n"]), deg)
    key = {(a, b): i for i, (a, b) in enumerate(zip(src.tolist(), col.tolist()))}
    for i in range(0, len(col), 97):
        r = key[(col[i], src[i])]
        assert g["wt"][i] == g["w"][r] and g["ext"][i] == g["ex"][r] and g["eyt"][i] == g["ey"][r]


def test_native_inputs_match_numpy():
    """The library's host builders of the surrogate inputs
    (fvdbm_nn_case_features, fvdbm_nn_case_graph) against the NumPy
    construction: graph arrays bitwise, features to float32 rounding (the
    NumPy total resistance is a pairwise sum, the native one sequential)."""
    for seed, cells in ((2, 1200), (5, 3000)):
        c = fv.porous.make_case(seed, target_cells=cells)
        for grid in ((32, 16), (32, 1), (5, 3)):
            x, g, N = sg.case_inputs(c, grid)
            x2, g2, N2 = sg.case_inputs_numpy(c, grid)
            assert N == N2 and g["n"] == g2["n"]
            np.testing.assert_allclose(x, x2, rtol=1e-6, atol=1e-6)
            for k in ("off", "col", "w", "ex", "ey", "wt", "ext", "eyt"):
                assert g[k].dtype == g2[k].dtype and np.array_equal(g[k], g2[k]), (grid, k)
    bad = np.array(c.mesher.faces, np.int32)
    bad[3, 1] = c.n_nodes  # out of range
    try:
        sg.case_graph(c.mesher.points, bad, (4, 2))
    except ValueError as e:
        assert "out of range" in str(e)
    else:
        raise AssertionError("an out-of-range mesh edge was accepted")


def test_twin_aggregate_matches_dense_adjacency():
    c, g = _case()
    N = c.n_nodes
    p = c.mesher.points
    A = np.zeros((N, N))
    DX = np.zeros((N, N))
    DY = np.zeros((N, N))
    for a, b in np.asarray(c.mesher.faces):
        for i, j in ((a, b), (b, a)):
            d = p[j] - p[i]
            L = np.hypot(*d)
            A[i, j], DX[i, j], DY[i, j] = 1.0, d[0] / L, d[1] / L
    deg = A.sum(1, keepdims=True)
    h = np.random.default_rng(0).standard_normal((N, 64))
    want = np.concatenate([h, (A / deg) @ h, (A * DX / deg) @ h, (A * DY / deg) @ h], 1)
    got = tw.aggregate(torch.tensor(h), g).numpy()
    assert np.allclose(got, want, rtol=1e-6, atol=1e-6)


def test_bf16_rounding_matches_kernel_formula():
    rng = np.random.default_rng(0)
    f = np.concatenate([rng.standard_normal(100000).astype(np.float32),
                        np.array([1.00390625, 1.01171875, -2.0039062, 3.0e-39], np.float32)])
    u = f.view(np.uint32).astype(np.uint64)
    bits = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint32) << 16
    want = bits.astype(np.uint32).view(np.float32)
    got = tw._round_bf16(torch.tensor(f)).numpy()
    assert np.array_equal(got, want)
