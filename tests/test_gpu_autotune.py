"""Autotuner: every candidate launch gives the same bits (checked inside
tune), the choice is cached and reused, and a tuned environment steps
exactly like the default one."""
import importlib

import numpy as np
import pytest

from cases import channel_case, fv, ldc_case, meshgen

pytestmark = pytest.mark.gpu
at = importlib.import_module("fvdbm-jax_amd.autotune")


def test_tune_single_and_cache(gpu, tmp_path):
    m, dyn, cells, faces, nodes = channel_case(meshgen.cylinder_channel(20_000, seed=3), tau=0.6,
                                               dt=1.0, u_max=0.05, dim_multiplier=1.0)
    dyn.delta_t = 1e-4
    cache = str(tmp_path / "tune.json")
    cands = at.candidates(blocks=(128, 256), variants=(0, 4), reorders=(4, 1))
    logs = []
    r = at.tune(cells, faces, nodes, cands=cands, steps=50, cache=cache, log=logs.append)
    assert not r["cached"] and len(r["results"]) == len(cands)
    assert r["best"] in r["results"]
    r2 = at.tune(cells, faces, nodes, cands=cands, steps=50, cache=cache)
    assert r2["cached"] and r2["best"] == r["best"]
    env = at.tuned_environment(cells, faces, nodes, cands=cands, cache=cache)
    env.step(30)
    ref = fv.Environment(cells, faces, nodes)
    ref.init()
    ref.step(30)
    assert np.array_equal(env.cells.pdf, ref.cells.pdf)
    env.close()
    ref.close()


def test_tune_batch(gpu, tmp_path):
    cases = [fv.porous.make_case(s, target_cells=4000).env_tuple() for s in (41, 42)]
    cands = at.candidates(blocks=(128, 256), variants=(0, 2), reorders=(4,))
    r = at.tune_batch(cases, cands=cands, steps=50, cache=str(tmp_path / "t.json"))
    assert len(r["results"]) == 4 and r["best"]["block"] in (128, 256)


def test_confirm_batch_on_the_whole_batch(gpu):
    cases = [fv.porous.make_case(s, target_cells=4000).env_tuple() for s in (41, 42, 43)]
    cands = at.candidates(blocks=(128, 256), variants=(0,), reorders=(4,))
    r = at.tune_batch(cases[:2], cands=cands, steps=50, cache=None)
    best, timings, b = at.confirm_batch(cases, r["results"], steps=40)
    try:
        assert sorted(t["block"] for t in timings) == [128, 256]
        assert best["block"] == min(timings, key=lambda t: t["us_per_step"])["block"]
        # the kept handle has run 20 + 40 steps; a fresh one of the default
        # options stepped as far gives the same bits
        ref = importlib.import_module("fvdbm-jax_amd.batch").BatchSolver(cases)
        ref.step(60)
        assert np.array_equal(b.get_pdf(), ref.get_pdf())
        ref.close()
    finally:
        b.close()
