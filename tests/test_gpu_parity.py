"""GPU parity: the HIP step (through the C ABI) against the CPU oracle.

Integer work is bit-exact by construction and the floating-point step keeps
the oracle's operation order with no contraction, so the bar here is bitwise
equality (np.array_equal) of every observable array, in FP32 and in FP64,
for both flux schemes, every cell order, kernel variant and block size.
"""
import numpy as np
import pytest

from cases import channel_case, fv, ldc_case, mean_area, meshgen, oracle_env, oracle_state, stable_dt
from oracle import oracle

pytestmark = pytest.mark.gpu

FIELDS = [("cells", "pdf", "cell_pdf"), ("cells", "rho", "cell_rho"), ("cells", "vel", "cell_vel"),
          ("cells", "pdf_eq", "cell_feq"), ("faces", "pdf", "face_pdf"),
          ("nodes", "pdf", "node_pdf"), ("nodes", "rho", "node_rho"), ("nodes", "vel", "node_vel")]


def _gpu_env(cells, faces, nodes, precision="fp32", **opts):
    env = fv.Environment(cells, faces, nodes, precision=precision, **opts)
    env.init()
    return env


def _assert_fields_equal(env, st, fields=FIELDS, what=""):
    for kind, name, key in fields:
        got = getattr(getattr(env, kind), name)
        exp = np.asarray(st[key]).reshape(got.shape)
        if not np.array_equal(got, exp):
            bad = np.argwhere(got != exp)
            d = np.max(np.abs(got.astype(np.float64) - exp))
            raise AssertionError(f"{what} {kind}.{name}: {len(bad)} mismatches, max |diff| {d:.3e}, "
                                 f"first at {bad[0].tolist()}")


def test_initial_observables(gpu):
    m, dyn, cells, faces, nodes = ldc_case()
    env = _gpu_env(cells, faces, nodes)
    W = np.float32(dyn.W)
    assert np.array_equal(env.cells.pdf, np.repeat(W[None], 2048, 0))
    assert not env.cells.rho.any() and not env.cells.vel.any() and not env.cells.pdf_eq.any()
    assert np.array_equal(env.faces.pdf, np.repeat(W[None], 3136, 0))
    assert np.array_equal(env.nodes.pdf, np.repeat(W[None], 1089, 0))
    assert np.array_equal(env.nodes.vel, nodes.__dict__["_host_vel"])


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_cavity_matches_oracle(gpu, precision):
    """32x32 cavity, 1000 steps (checked after 1, 10, 100 and 1000)."""
    m, dyn, cells, faces, nodes = ldc_case()
    p = "f32" if precision == "fp32" else "f64"
    oenv = oracle_env(cells, faces, nodes, p)
    st = oracle_state(oenv, nodes, p)
    env = _gpu_env(cells, faces, nodes, precision)
    done = 0
    for target in (1, 10, 100, 1000):
        oracle.run(oenv, st, dyn.tau, dyn.delta_t, target - done, p)
        env = env.step(target - done)
        done = target
        _assert_fields_equal(env, st, FIELDS if target in (1, 1000) else FIELDS[:1], f"step {target}")


def test_cavity_matches_committed_fixtures(gpu):
    """The GPU path against the committed oracle snapshots (tests/golden)."""
    import os
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "ldc32_oracle.npz"))
    m, dyn, cells, faces, nodes = ldc_case()
    env = _gpu_env(cells, faces, nodes)
    done = 0
    for n in (1, 10, 100, 1000):
        env.step(n - done)
        done = n
        assert np.array_equal(env.cells.pdf, z[f"pdf_f32_{n}"]), n
    env64 = _gpu_env(cells, faces, nodes, "fp64")
    env64.step(1000)
    assert np.array_equal(env64.cells.pdf, z["pdf_f64_1000"])


def test_cavity_lax_wendroff(gpu):
    m, dyn, cells, faces, nodes = ldc_case(scheme="lax_wendroff")
    oenv = oracle_env(cells, faces, nodes)
    st = oracle_state(oenv, nodes)
    oracle.run(oenv, st, dyn.tau, dyn.delta_t, 300, "f32", scheme="lax_wendroff")
    env = _gpu_env(cells, faces, nodes)
    env.step(300)
    _assert_fields_equal(env, st)


def _channel(target_cells=20_000, seed=0, scheme="upwind"):
    """dim_multiplier 1/mean(A) for upwind; 1/h for Lax-Wendroff, whose
    correction term (dt / (d0+d1)) is only stable with unit-order spacing."""
    raw = meshgen.cylinder_channel(target_cells, seed=seed)
    from cases import build_mesher
    mm = build_mesher(raw)
    A = mean_area(mm)
    mult = 1.0 / A if scheme == "upwind" else 1.0 / np.sqrt(2 * A)
    m, dyn, cells, faces, nodes = channel_case(raw, tau=0.55, dt=1.0, u_max=0.1,
                                               dim_multiplier=mult, scheme=scheme)
    dyn.delta_t = float(np.float32(stable_dt(m, faces.L, dyn.tau)))
    return m, dyn, cells, faces, nodes


@pytest.mark.parametrize("scheme", ["upwind", "lax_wendroff"])
def test_channel_matches_oracle(gpu, scheme):
    """Cylinder channel: parabolic velocity inlet (per-node values), density
    outlet (type-2 nodes), no-slip walls and cylinder."""
    m, dyn, cells, faces, nodes = _channel(scheme=scheme)
    assert (nodes.type == 2).any() and (nodes.type == 1).any()
    oenv = oracle_env(cells, faces, nodes)
    st = oracle_state(oenv, nodes)
    oracle.run(oenv, st, dyn.tau, dyn.delta_t, 200, "f32", scheme=scheme)
    assert np.isfinite(st["cell_pdf"]).all()
    env = _gpu_env(cells, faces, nodes)
    env.step(200)
    _assert_fields_equal(env, st)


def test_full_size_channel_matches_oracle(gpu):
    """The benchmarked workload itself: the 1M-cell cylinder channel (seed 0,
    tau 0.55, u_max 0.1, dim_multiplier 1/mean(A), dt at the explicit bound;
    the geometry, tile count, XCD-aware tile order and block-256 LDS sizing
    of the bench), 50 steps, every observable bitwise equal to the OpenMP
    oracle."""
    import os
    oracle.set_threads(min(16, os.cpu_count() or 1))
    m, dyn, cells, faces, nodes = _channel(target_cells=1_000_000, seed=0)
    assert 990_000 <= len(m.cells) <= 1_010_000
    print(f"full-size channel: {len(m.cells)} cells", flush=True)
    oenv = oracle_env(cells, faces, nodes)
    st = oracle_state(oenv, nodes)
    oracle.run(oenv, st, dyn.tau, dyn.delta_t, 50, "f32")
    env = _gpu_env(cells, faces, nodes)
    info = env.info()
    assert not (info["variant"] & 8) and info["block"] == 256  # two-step tiles, bench tile size
    env.step(50)
    _assert_fields_equal(env, st, what="1M channel, 50 steps")


@pytest.mark.parametrize("opts", [
    dict(reorder=0), dict(reorder=1), dict(reorder=2), dict(reorder=3), dict(variant=1), dict(variant=2),
    dict(variant=3), dict(variant=4), dict(variant=7), dict(reorder=3, block=1024),
    dict(block=64), dict(block=512), dict(block=1024), dict(graph_steps=-1), dict(graph_steps=6),
    dict(variant=8), dict(variant=9), dict(block=128), dict(graph_steps=10),
])
def test_layout_and_variants_bitwise_equal(gpu, opts):
    """Cell order (n1), kernel variant / block size (autotuner candidates),
    one or two steps per launch and graph capture change nothing in the
    results."""
    m, dyn, cells, faces, nodes = _channel(target_cells=8_000, seed=1)
    base = _gpu_env(cells, faces, nodes)
    base.step(77)
    ref = base.cells.pdf
    ref_flux = base.faces.pdf
    base.close()
    env = _gpu_env(cells, faces, nodes, **opts)
    env.step(77)
    assert np.array_equal(env.cells.pdf, ref)
    assert np.array_equal(env.faces.pdf, ref_flux)


def test_set_state_and_residual(gpu):
    m, dyn, cells, faces, nodes = ldc_case()
    rng = np.random.default_rng(0)
    f0 = (np.float32(dyn.W) * (1 + 0.05 * rng.standard_normal((2048, 9)))).astype(np.float32)
    env = _gpu_env(cells, faces, nodes)
    env.cells.pdf = f0
    assert np.array_equal(env.cells.pdf, f0)
    oenv = oracle_env(cells, faces, nodes)
    st = oracle_state(oenv, nodes)
    st["cell_pdf"] = f0.copy()
    oracle.run(oenv, st, dyn.tau, dyn.delta_t, 9, "f32")
    before = st["cell_pdf"].copy()
    oracle.run(oenv, st, dyn.tau, dyn.delta_t, 1, "f32")
    r = env.step_residual(10)
    _assert_fields_equal(env, st)
    _, u0 = oracle.calc_macro(before)
    _, u1 = oracle.calc_macro(st["cell_pdf"])
    d = (u1 - u0).astype(np.float32)
    num = np.sum((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]).astype(np.float32), dtype=np.float64)
    den = np.sum((u1[:, 0] * u1[:, 0] + u1[:, 1] * u1[:, 1]).astype(np.float32), dtype=np.float64)
    assert abs(r - np.sqrt(num) / np.sqrt(den)) <= 1e-9 * r


def test_moments_describe_start_of_last_step(gpu):
    """cells.rho/vel/pdf_eq after step() are those of the pdf the step read
    (src/environment.py:60-64)."""
    m, dyn, cells, faces, nodes = ldc_case()
    env = _gpu_env(cells, faces, nodes)
    env.step(5)
    prev = env.cells.pdf
    env.step()
    rho, vel = oracle.calc_macro(prev)
    assert np.array_equal(env.cells.rho.reshape(-1), rho)
    assert np.array_equal(env.cells.vel, vel)
    assert np.array_equal(env.cells.pdf_eq, oracle.calc_eq(rho, vel))


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_observables_after_runs_of_every_length(gpu, precision):
    """Runs of two or more steps end in a two-step launch that keeps its
    intermediate state for the start-of-step observables; odd runs start
    with one single step. Every observable equals the oracle's after runs
    of 1, 2, 3, 4, 7, 2 and 66 steps (66 crosses a captured graph)."""
    m, dyn, cells, faces, nodes = _channel(target_cells=8_000, seed=4)
    p = "f32" if precision == "fp32" else "f64"
    oenv = oracle_env(cells, faces, nodes, p)
    st = oracle_state(oenv, nodes, p)
    env = _gpu_env(cells, faces, nodes, precision)
    for n in (1, 2, 3, 4, 7, 2, 66):
        oracle.run(oenv, st, dyn.tau, dyn.delta_t, n, p)
        env = env.step(n)
        _assert_fields_equal(env, st, what=f"run of {n}")


def test_boundary_values_set_after_a_two_step_run(gpu):
    """Velocity-node values changed after a run that ended in a two-step
    launch: the cell observables still describe the start of the last step
    as that run computed it (with the old values), and the following steps
    use the new values."""
    m, dyn, cells, faces, nodes = _channel(target_cells=8_000, seed=5)
    oenv = oracle_env(cells, faces, nodes)
    st = oracle_state(oenv, nodes)
    env = _gpu_env(cells, faces, nodes)
    env.step(4)
    oracle.run(oenv, st, dyn.tau, dyn.delta_t, 4, "f32")
    vel = np.asarray(st["node_vel"], np.float32).copy()
    vnode = nodes.type.reshape(-1) == 1
    vel[vnode] *= np.float32(1.5)
    env.nodes.vel = vel
    _assert_fields_equal(env, st, FIELDS[:4], "after setting node velocities")
    st["node_vel"][vnode] = vel[vnode]
    oracle.run(oenv, st, dyn.tau, dyn.delta_t, 3, "f32")
    env.step(3)
    _assert_fields_equal(env, st, what="3 steps with the new values")


def test_prepare_changes_nothing(gpu):
    """Capturing a run's graphs ahead (prepare) leaves the state as it was,
    and the run then gives the oracle's result, from either parity."""
    m, dyn, cells, faces, nodes = _channel(target_cells=8_000, seed=6)
    oenv = oracle_env(cells, faces, nodes)
    st = oracle_state(oenv, nodes)
    env = _gpu_env(cells, faces, nodes)
    env.step(3)
    before = env.cells.pdf
    env.prepare(20)
    assert np.array_equal(env.cells.pdf, before) and env.info()["steps_done"] == 3
    env.step(5)  # odd: the timed run below starts from the other parity
    ms = env.time_steps(20)
    assert ms > 0
    oracle.run(oenv, st, dyn.tau, dyn.delta_t, 28, "f32")
    _assert_fields_equal(env, st, what="prepared run")


def test_launch_mode_default_by_mesh_size(gpu):
    """By default a handle below 131072 padded cells replays captured graphs
    (32 steps each) and a larger one launches directly; either mode, set
    explicitly, gives the same bits as the other after runs that end on
    pair launches and on a single step."""
    small = _gpu_env(*_channel(target_cells=8_000, seed=6)[2:])
    assert small.info()["graph_steps"] == 32
    small.close()
    m, dyn, cells, faces, nodes = _channel(target_cells=150_000, seed=2)
    direct = _gpu_env(cells, faces, nodes)
    graphs = _gpu_env(cells, faces, nodes, graph_steps=32)
    assert direct.info()["n_cells_padded"] >= 131072
    assert direct.info()["graph_steps"] == -1 and graphs.info()["graph_steps"] == 32
    for k in (40, 33):
        print(f"launch modes, 150k cells: {k} more steps", flush=True)
        direct.step(k)
        graphs.step(k)
        assert np.array_equal(direct.cells.pdf, graphs.cells.pdf), k
    assert np.array_equal(direct.nodes.pdf, graphs.nodes.pdf)
    direct.close()
    graphs.close()


def test_every_cell_written_and_deterministic(gpu):
    """Two runs agree bitwise and every cell holds a finite pdf."""
    m, dyn, cells, faces, nodes = _channel(target_cells=8_000, seed=2)
    outs = []
    for _ in range(2):
        env = _gpu_env(cells, faces, nodes)
        env.step(50)
        outs.append(env.cells.pdf)
        env.close()
    assert np.isfinite(outs[0]).all()
    assert np.array_equal(outs[0], outs[1])


def test_checkpoint_resume_bitwise(gpu, tmp_path):
    """k steps, save, load, N - k steps equal the oracle's N steps run
    straight through (every observable, bitwise)."""
    m, dyn, cells, faces, nodes = _channel(target_cells=8_000, seed=3)
    oenv = oracle_env(cells, faces, nodes)
    st = oracle_state(oenv, nodes)
    oracle.run(oenv, st, dyn.tau, dyn.delta_t, 60, "f32")
    env = _gpu_env(cells, faces, nodes)
    env.step(25)
    m.to_pickle(env, str(tmp_path / "ck"))
    env.close()
    env2, m2 = fv.Mesher.from_pickle(str(tmp_path / "ck"))
    assert np.array_equal(m2.cells, m.cells)
    env2.step(35)
    _assert_fields_equal(env2, st, what="resumed run")


def test_vtk_export_of_a_gpu_state(gpu, tmp_path):
    """Mesher.to_vtk of a GPU state after 30 channel steps: the exported
    Velocity, Density, pdf and feq cell data read back to the oracle's
    float32 values bit for bit (9 significant digits round-trip float32)."""
    m, dyn, cells, faces, nodes = _channel(target_cells=8_000, seed=4)
    oenv = oracle_env(cells, faces, nodes)
    st = oracle_state(oenv, nodes)
    oracle.run(oenv, st, dyn.tau, dyn.delta_t, 30, "f32")
    env = _gpu_env(cells, faces, nodes)
    env.step(30)
    path = m.to_vtk(env, str(tmp_path / "chan"), save_f=True, save_feq=True)
    env.close()
    lines = open(path).read().split("\n")
    C = len(m.cells)

    def block(header, ncols):
        i = lines.index(header) + 1
        if header.startswith("SCALARS"):
            i += 1  # LOOKUP_TABLE default
        rows = np.array([[float(v) for v in ln.split()] for ln in lines[i:i + C]])
        assert rows.shape == (C, ncols), header
        return rows.astype(np.float32)

    vel = block("VECTORS Velocity double", 3)
    assert np.array_equal(vel[:, :2], np.asarray(st["cell_vel"]).reshape(C, 2))
    assert not vel[:, 2].any()
    assert np.array_equal(block("SCALARS Density double 1", 1).reshape(C),
                          np.asarray(st["cell_rho"]).reshape(C))
    assert np.array_equal(block(f"pdf 9 {C} double", 9), st["cell_pdf"])
    assert np.array_equal(block(f"feq 9 {C} double", 9), np.asarray(st["cell_feq"]).reshape(C, 9))
    assert "FIELD FieldData 1" in lines


def test_close_with_steps_still_queued(gpu):
    """Steps are queued asynchronously; closing a handle right after a long
    queued run lets it finish before its graphs and buffers go. A new handle
    afterwards steps as the oracle does."""
    m, dyn, cells, faces, nodes = ldc_case()
    env = fv.Environment(cells, faces, nodes)
    env.init()
    env.step(3001)  # a graph replay of two-step launches plus one single step, not awaited
    env.close()
    oenv = oracle_env(cells, faces, nodes)
    st = oracle_state(oenv, nodes)
    oracle.run(oenv, st, dyn.tau, dyn.delta_t, 12, "f32")
    env2 = fv.Environment(cells, faces, nodes)
    env2.init()
    env2.step(12)
    assert np.array_equal(env2.cells.pdf, st["cell_pdf"])
    env2.close()
