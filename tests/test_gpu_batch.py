"""Batched solves on the GPU: each case of a batch is bitwise equal to its
single solve; a batch of one is the single path; a stopped case is frozen;
per-case residuals are the single-solve residuals."""
import importlib

import numpy as np
import pytest

from cases import fv, porous_case

pytestmark = pytest.mark.gpu
batch = importlib.import_module("fvdbm-jax_amd.batch")


def _cases(n=4, target=6000):
    return [porous_case(seed, target_cells=target + 1500 * seed) for seed in range(n)]


def _single(m, dyn, cells, faces, nodes, nsteps, resid=False):
    env = fv.Environment(cells, faces, nodes)
    env.init()
    r = None
    if resid:
        r = env.step_residual(nsteps)
    else:
        env.step(nsteps)
    out = env.cells.pdf
    env.close()
    return out, r


def test_batch_cases_equal_single_solves(gpu):
    cs = _cases()
    assert len({round(c[1].tau, 6) for c in cs}) == len(cs)  # different tau per case
    b = batch.BatchSolver([(c, f, n) for _, _, c, f, n in cs])
    b.step(123)
    pdf = b.get_pdf()
    for k, (m, dyn, c, f, n) in enumerate(cs):
        single, _ = _single(m, dyn, c, f, n, 123)
        assert np.array_equal(b.case_pdf(k, pdf), single), k


def test_batch_of_one_is_single(gpu):
    m, dyn, c, f, n = porous_case(7, target_cells=5000)
    b = batch.BatchSolver([(c, f, n)])
    b.step(50)
    single, _ = _single(m, dyn, c, f, n, 50)
    assert np.array_equal(b.get_pdf(), single)


def test_lds_classes_equal_one_launch(gpu):
    """Where the largest tile's LDS, not the registers, limits the resident
    blocks per CU, a two-step launch is split into two launches by tile LDS
    size. The result equals one launch over all tiles (variant bit 1: plain
    tile order, never split), at every block size; small blocks on these
    boundary-heavy meshes take the split."""
    cs = [(c, f, n) for _, _, c, f, n in _cases(3)]
    classes = []
    for block in (64, 128, 256):
        b = batch.BatchSolver(cs, block=block)
        ref = batch.BatchSolver(cs, block=block, variant=1)
        classes.append(b.info()["launch_classes"])
        assert ref.info()["launch_classes"] == 1
        b.step(37)
        ref.step(37)
        assert np.array_equal(b.get_pdf(), ref.get_pdf()), block
        b.close()
        ref.close()
    assert 2 in classes, classes


def test_workload_batch_matches_oracle(gpu):
    """Three cases of the batched throughput workload (5-40 discs, 20k-47k
    cells, velocity inlet, density outlet, each its own tau and dt) in one
    batch, at three tile sizes: after 100 steps every case's pdf is bitwise
    equal to the oracle's single solve of it. At one tile size at least,
    these boundary-heavy meshes take the two-launch split by tile LDS size,
    so both launch classes are checked against the oracle."""
    from cases import oracle_env, oracle_state
    from oracle import oracle
    wl = fv.porous.WORKLOAD
    cases = [fv.porous.make_case(seed, dist=wl, target_cells=t)
             for seed, t in ((11, 20_000), (12, 31_000), (13, 47_000))]
    for c in cases:
        assert wl["count_range"][0] <= len(c.discs) <= wl["count_range"][1]
        assert c.n_cells >= 19_000
    expect = []
    for c in cases:
        oenv = oracle_env(c.cells, c.faces, c.nodes)
        st = oracle_state(oenv, c.nodes)
        oracle.run(oenv, st, c.dynamics.tau, c.dynamics.delta_t, 100, "f32")
        assert np.isfinite(st["cell_pdf"]).all()
        expect.append(st["cell_pdf"])
    classes = []
    for block in (64, 128, 256):
        print(f"workload batch, block {block}", flush=True)
        b = batch.BatchSolver([c.env_tuple() for c in cases], block=block)
        classes.append(b.info()["launch_classes"])
        b.step(100)
        pdf = b.get_pdf()
        for k, want in enumerate(expect):
            assert np.array_equal(b.case_pdf(k, pdf), want), (block, k)
        b.close()
    assert 2 in classes, classes


@pytest.mark.parametrize("variant", [0, 2, 8])
def test_run_residuals_and_freezing(gpu, variant):
    """variant 0 runs two steps per launch (two-step tiles), 8 one step per
    launch, 2 boundary nodes in their own kernel; each must read each case's
    own buffer once the cases' parities differ."""
    cs = _cases(3)
    b = batch.BatchSolver([(c, f, n) for _, _, c, f, n in cs], variant=variant)
    assert bool(b.info()["variant"] & 8) == (variant != 0)
    st = b.run(max_steps=200, check_every=50, tol=0.0)  # tol 0: nobody converges
    assert b.steps_run == 200 and np.all(st["steps"] == 200) and st["active"].all()
    for k, (m, dyn, c, f, n) in enumerate(cs):
        single, r = _single(m, dyn, c, f, n, 200, resid=True)
        assert np.array_equal(b.case_pdf(k), single)
        assert st["residual"][k] == pytest.approx(r, rel=1e-12)
    # stop case 1; the others go on an odd number of steps (their buffer
    # parity now differs from case 1's), case 1 stays bitwise frozen
    frozen = b.case_pdf(1).copy()
    b.set_active([1, 0, 1])
    st = b.run(max_steps=75, check_every=25, tol=0.0)
    assert np.array_equal(b.case_pdf(1), frozen)
    assert st["steps"].tolist() == [275, 200, 275]
    assert st["status"][1] == 4  # stopped
    single, _ = _single(*cs[0], 275)
    assert np.array_equal(b.case_pdf(0), single)
    b.step(33)  # plain steps (graph replays plus a remainder)
    single, _ = _single(*cs[2], 308)
    assert np.array_equal(b.case_pdf(2), single)
    assert np.array_equal(b.case_pdf(1), frozen)
    assert b.status()["steps"].tolist() == [308, 200, 308]


def test_run_stops_at_max_steps(gpu):
    """A step cap that is not a multiple of the check interval ends with a
    shorter interval: every case runs exactly max_steps, its pdf and last
    residual are those of a single solve of that many steps."""
    cs = _cases(2)
    b = batch.BatchSolver([(c, f, n) for _, _, c, f, n in cs])
    st = b.run(max_steps=130, check_every=50, tol=0.0)
    assert b.steps_run == 130 and st["steps"].tolist() == [130, 130]
    for k, (m, dyn, c, f, n) in enumerate(cs):
        single, r = _single(m, dyn, c, f, n, 130, resid=True)
        assert np.array_equal(b.case_pdf(k), single)
        assert st["residual"][k] == pytest.approx(r, rel=1e-12)
    st = b.run(max_steps=20, check_every=50, tol=0.0)  # a cap below one interval
    assert b.steps_run == 20 and st["steps"].tolist() == [150, 150]


def test_run_until_converged(gpu):
    cs = _cases(2, target=4000)
    b = batch.BatchSolver([(c, f, n) for _, _, c, f, n in cs])
    st = b.run(max_steps=300_000, check_every=500, tol=1e-6)
    assert (st["status"] == 1).all(), st
    assert (st["residual"] < 1e-6).all()
    assert np.isfinite(b.get_pdf()).all()
    # a converged case no longer changes
    before = b.get_pdf()
    b.run(max_steps=1000, check_every=500, tol=1e-6)
    assert np.array_equal(b.get_pdf(), before)


def test_nonfinite_case_stops_as_diverged(gpu):
    cs = _cases(2, target=4000)
    b = batch.BatchSolver([(c, f, n) for _, _, c, f, n in cs])
    pdf = b.get_pdf()
    pdf[b.cell_off[1] + 7, 3] = np.nan  # one poisoned cell in case 1
    b.set_pdf(pdf)
    st = b.run(max_steps=400, check_every=100, tol=1e-12)
    assert st["status"].tolist() == [0, 2] and st["active"].tolist() == [True, False]
    assert np.isnan(st["residual"][1])


def test_fp64_batch_equals_single_solves_and_oracle(gpu):
    """The batched handle in FP64 (the precision the steady-state reference
    runs in): each case after an odd number of steps is bitwise equal to its
    FP64 single solve, and the first case to the FP64 oracle."""
    from cases import oracle_env, oracle_state
    from oracle import oracle
    cs = _cases(3)
    b = batch.BatchSolver([(c, f, n) for _, _, c, f, n in cs], precision="fp64")
    b.step(61)
    pdf = b.get_pdf()
    assert pdf.dtype == np.float64
    for k, (m, dyn, c, f, n) in enumerate(cs):
        env = fv.Environment(c, f, n, precision="fp64")
        env.init()
        env.step(61)
        assert np.array_equal(b.case_pdf(k, pdf), env.cells.pdf), k
        env.close()
    m, dyn, c, f, n = cs[0]
    oenv = oracle_env(c, f, n, "f64")
    st = oracle_state(oenv, n, "f64")
    oracle.run(oenv, st, dyn.tau, dyn.delta_t, 61, "f64")
    assert np.array_equal(b.case_pdf(0, pdf), st["cell_pdf"])
    b.close()


def test_lax_wendroff_batch_equals_single_solves(gpu):
    """A batch of Lax-Wendroff cases (one step per launch, the scheme's
    kernel) equals each case's single Lax-Wendroff solve bitwise."""
    cs = [porous_case(seed, target_cells=5000 + 1000 * seed, scheme="lax_wendroff") for seed in range(3)]
    b = batch.BatchSolver([(c, f, n) for _, _, c, f, n in cs])
    assert b.info()["variant"] & 8  # Lax-Wendroff runs one step per launch
    b.step(45)
    pdf = b.get_pdf()
    assert np.isfinite(pdf).all()
    for k, (m, dyn, c, f, n) in enumerate(cs):
        single, _ = _single(m, dyn, c, f, n, 45)
        assert np.array_equal(b.case_pdf(k, pdf), single), k
    b.close()
