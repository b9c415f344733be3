"""The any-lattice handle (fvdbm_lattice_*, csrc/fvdbm_lattice.hip) on the GPU:

* D2Q13 against the oracle's D2Q13 step, every observable bitwise, FP32 and
  FP64, upwind and Lax-Wendroff, cavity and cylinder channel (velocity and
  density nodes);
* the same kernels on D2Q9 against the tuned D2Q9 handle, bitwise (a check
  of the plain kernels independent of the D2Q13 oracle);
* what the handle does not offer fails loudly.
"""
import numpy as np
import pytest

from cases import build_mesher, channel_case, fv, ldc_case, mean_area, meshgen, oracle_env, oracle_state, \
    stable_dt
from oracle import oracle

pytestmark = pytest.mark.gpu

FIELDS = [("cells", "pdf", "cell_pdf"), ("cells", "rho", "cell_rho"), ("cells", "vel", "cell_vel"),
          ("cells", "pdf_eq", "cell_feq"), ("faces", "pdf", "face_pdf"),
          ("nodes", "pdf", "node_pdf"), ("nodes", "rho", "node_rho"), ("nodes", "vel", "node_vel")]


def _equal(env, st, what=""):
    for kind, name, key in FIELDS:
        got = getattr(getattr(env, kind), name)
        exp = np.asarray(st[key]).reshape(got.shape)
        if not np.array_equal(got, exp):
            bad = np.argwhere(got != exp)
            d = np.max(np.abs(got.astype(np.float64) - exp))
            raise AssertionError(f"{what} {kind}.{name}: {len(bad)} mismatches, max |diff| {d:.3e}, "
                                 f"first at {bad[0].tolist()}")


def _channel13(target_cells=12_000, scheme="upwind"):
    raw = meshgen.cylinder_channel(target_cells, seed=2)
    A = mean_area(build_mesher(raw))
    mult = 1.0 / A if scheme == "upwind" else 1.0 / np.sqrt(2 * A)
    m, dyn, cells, faces, nodes = channel_case(raw, tau=0.6, dt=1.0, u_max=0.08, dim_multiplier=mult,
                                               scheme=scheme, lattice=fv.D2Q13)
    # speeds up to 2: half the D2Q9 explicit bound
    dyn.delta_t = float(np.float32(0.5 * stable_dt(m, faces.L, dyn.tau)))
    return m, dyn, cells, faces, nodes


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_d2q13_cavity_matches_oracle(gpu, precision):
    m, dyn, cells, faces, nodes = ldc_case(lattice=fv.D2Q13)
    p = "f32" if precision == "fp32" else "f64"
    oenv = oracle_env(cells, faces, nodes, p)
    st = oracle_state(oenv, nodes, p, 13)
    env = fv.Environment(cells, faces, nodes, precision=precision)
    env.init()
    assert env.info()["variant"] == 4  # two-step tiles (upwind), one step per launch otherwise
    W = np.asarray(oracle.W13_64, env.dtype)  # W in the working precision
    assert np.array_equal(env.cells.pdf, np.repeat(W[None], cells.size, 0))
    assert np.array_equal(env.faces.pdf, np.repeat(W[None], faces.size, 0))
    done = 0
    for target in (1, 33, 300):  # single launches, a graph replay plus remainder
        oracle.run(oenv, st, dyn.tau, dyn.delta_t, target - done, p, quivers=13)
        env.step(target - done)
        done = target
        _equal(env, st, f"step {target}")
    assert np.isfinite(env.cells.pdf).all()


def test_d2q13_observables_after_runs_of_every_length(gpu):
    """Runs of two or more steps end in a two-step launch that keeps its
    intermediate state for the start-of-step observables; every observable
    equals the oracle's after runs of 2, 3, 1, 4 and 5 steps."""
    m, dyn, cells, faces, nodes = _channel13(target_cells=6_000)
    oenv = oracle_env(cells, faces, nodes)
    st = oracle_state(oenv, nodes, "f32", 13)
    env = fv.Environment(cells, faces, nodes)
    env.init()
    for n in (2, 3, 1, 4, 5):
        oracle.run(oenv, st, dyn.tau, dyn.delta_t, n, "f32", quivers=13)
        env.step(n)
        _equal(env, st, f"run of {n}")


@pytest.mark.parametrize("scheme", ["upwind", "lax_wendroff"])
def test_d2q13_channel_matches_oracle(gpu, scheme):
    m, dyn, cells, faces, nodes = _channel13(scheme=scheme)
    assert (nodes.type == 2).any() and (nodes.type == 1).any()
    oenv = oracle_env(cells, faces, nodes)
    st = oracle_state(oenv, nodes, "f32", 13)
    oracle.run(oenv, st, dyn.tau, dyn.delta_t, 120, "f32", scheme=scheme, quivers=13)
    assert np.isfinite(st["cell_pdf"]).all()
    env = fv.Environment(cells, faces, nodes)
    env.init()
    env.step(120)
    _equal(env, st, scheme)


def test_d2q13_set_fields_then_step(gpu):
    """A perturbed start state (cell and node pdfs set through the fields)
    steps exactly as the oracle does from the same state."""
    m, dyn, cells, faces, nodes = ldc_case(lattice=fv.D2Q13)
    rng = np.random.default_rng(5)
    f = (np.asarray(fv.D2Q13.W, np.float32)[None] * (1 + 0.05 * rng.standard_normal((cells.size, 13)))
         ).astype(np.float32)
    oenv = oracle_env(cells, faces, nodes)
    st = oracle_state(oenv, nodes, "f32", 13)
    st["cell_pdf"] = f.copy()
    env = fv.Environment(cells, faces, nodes)
    env.init()
    env.cells.pdf = f
    oracle.run(oenv, st, dyn.tau, dyn.delta_t, 40, "f32", quivers=13)
    env.step(40)
    _equal(env, st, "perturbed")


@pytest.mark.parametrize("scheme", ["upwind", "lax_wendroff"])
def test_plain_kernels_equal_tuned_d2q9(gpu, scheme):
    """D2Q9 through the any-lattice handle equals the tuned handle bitwise."""
    raw = meshgen.cylinder_channel(15_000, seed=4)
    A = mean_area(build_mesher(raw))
    mult = 1.0 / A if scheme == "upwind" else 1.0 / np.sqrt(2 * A)
    m, dyn, cells, faces, nodes = channel_case(raw, tau=0.55, dt=1.0, u_max=0.1, dim_multiplier=mult,
                                               scheme=scheme)
    dyn.delta_t = float(np.float32(stable_dt(m, faces.L, dyn.tau)))
    tuned = fv.Environment(cells, faces, nodes)
    tuned.init()
    tuned.step(101)
    ref = {(k, n): getattr(getattr(tuned, k), n) for k, n, _ in FIELDS}
    tuned.close()
    plain = fv.Environment(cells, faces, nodes, lattice_kernel=True)
    plain.init()
    plain.step(101)
    for (k, n), want in ref.items():
        assert np.array_equal(getattr(getattr(plain, k), n), want), (k, n)


def test_lattice_handle_limits(gpu):
    m, dyn, cells, faces, nodes = ldc_case(lattice=fv.D2Q13)
    env = fv.Environment(cells, faces, nodes)
    env.init()
    with pytest.raises(RuntimeError, match="tuned D2Q9 handle"):
        env.step_residual(1)
    with pytest.raises(RuntimeError, match="tuned D2Q9 handle"):
        env.node_fields()
    with pytest.raises(ValueError, match="derived"):
        env.cells.rho = np.zeros((cells.size, 1), np.float32)
    ms = env.time_steps(64)
    assert ms > 0


def test_d2q13_checkpoint_resume_bitwise(gpu, tmp_path):
    """to_pickle / from_pickle keep the lattice: 25 D2Q13 steps, a
    checkpoint, a new environment from it and 35 more steps equal the D2Q13
    oracle's 60 straight steps, every observable bitwise."""
    m, dyn, cells, faces, nodes = _channel13(target_cells=6_000)
    oenv = oracle_env(cells, faces, nodes)
    st = oracle_state(oenv, nodes, "f32", 13)
    oracle.run(oenv, st, dyn.tau, dyn.delta_t, 60, "f32", quivers=13)
    env = fv.Environment(cells, faces, nodes)
    env.init()
    env.step(25)
    m.to_pickle(env, str(tmp_path / "ck"))
    env.close()
    env2, m2 = fv.Mesher.from_pickle(str(tmp_path / "ck"))
    assert type(env2.cells.dynamics).__name__ == "D2Q13"
    env2.step(35)
    _equal(env2, st, "resumed D2Q13 run")
    env2.close()
