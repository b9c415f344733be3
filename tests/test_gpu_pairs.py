"""Two-step tiles (the default for single upwind handles; variant bit 8
forces one step per launch): every observable equals the one-step kernel
bitwise, for step counts that use pair launches, pair graphs and the
single-step remainder, on meshes with velocity and density nodes."""
import numpy as np
import pytest

from cases import channel_case, fv, ldc_case, meshgen, porous_case

pytestmark = pytest.mark.gpu


def _cases():
    yield "cavity", ldc_case()[2:]
    m, dyn, c, f, n = channel_case(meshgen.cylinder_channel(20_000, seed=5), tau=0.6, dt=1.0,
                                   u_max=0.05, dim_multiplier=1.0)
    dyn.delta_t = 1e-4
    yield "channel", (c, f, n)
    yield "porous", porous_case(4, target_cells=6000)[2:]


@pytest.mark.parametrize("name,case", list(_cases()), ids=lambda v: v if isinstance(v, str) else "")
def test_pairs_bitwise(gpu, name, case):
    c, f, n = case
    ref = fv.Environment(c, f, n, variant=8)
    two = fv.Environment(c, f, n)
    ref.init()
    two.init()
    assert two.info()["variant"] == 0 and ref.info()["variant"] == 8
    done = 0
    for k in (1, 2, 3, 7, 64, 65, 100):
        ref.step(k)
        two.step(k)
        done += k
        assert np.array_equal(two.cells.pdf, ref.cells.pdf), (name, done)
    for attr in ("rho", "vel", "pdf_eq"):
        assert np.array_equal(getattr(two.cells, attr), getattr(ref.cells, attr)), attr
    assert np.array_equal(two.nodes.pdf, ref.nodes.pdf)
    assert two.step_residual(33) == ref.step_residual(33)
    assert np.array_equal(two.cells.pdf, ref.cells.pdf)
    ms = two.time_steps(64)
    ref.time_steps(64)
    assert ms > 0 and np.array_equal(two.cells.pdf, ref.cells.pdf)
    ref.close()
    two.close()


@pytest.mark.parametrize("block", [64, 128])
def test_ghost_records_on_a_strip_one_cell_thick(gpu, block):
    """A channel one triangle thick: half its cells own a wall face and the
    inlet/outlet cells up to two, so most waves of every tile read ghost
    records (one per ghost face and step, staged after the tile's ring).
    Two-step tiles of 64 and 128 cells against the oracle, bitwise, after
    runs that end on pair launches and on a single step."""
    from cases import oracle_env, oracle_state
    from oracle import oracle
    raw = meshgen.jittered_delaunay(2.0, 0.01, [], 0.01, seed=5)
    m, dyn, c, f, n = channel_case(raw, tau=0.6, dt=1.0, u_max=0.05, dim_multiplier=1.0)
    dyn.delta_t = 1e-3
    env = fv.Environment(c, f, n, block=block)
    env.init()
    assert env.info()["variant"] == 0  # two-step tiles
    oenv = oracle_env(c, f, n)
    st = oracle_state(oenv, n)
    done = 0
    for k in (40, 41):
        print(f"strip, block {block}: {k} more steps", flush=True)
        env.step(k)
        oracle.run(oenv, st, dyn.tau, dyn.delta_t, k, "f32")
        done += k
        assert np.array_equal(env.cells.pdf, st["cell_pdf"]), (block, done)
    assert np.isfinite(st["cell_pdf"]).all()
    env.close()
