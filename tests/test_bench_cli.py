"""bench.py measures one GPU: a run asked for several exits at once with a
message that says so (before any GPU or mesh work), and the help text
lists the options the measurement contract uses."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True,
                          text=True, timeout=120, cwd=ROOT)


def test_more_than_one_gpu_is_refused():
    r = _bench("--gpus", "2", "--steps", "5", "--warmup", "1")
    assert r.returncode == 2
    assert "one GPU only" in r.stderr and "--gpus 2" in r.stderr
    assert r.stdout == ""


def test_help_lists_the_contract_options():
    r = _bench("--help")
    assert r.returncode == 0
    for opt in ("--gpus", "--steps", "--warmup", "--lattice", "--precision", "--batch-cases",
                "--steady-cases"):
        assert opt in r.stdout, opt


def test_grouped_steady_study_aggregates_every_case():
    """steady_study solves groups one after the other and reports over all
    cases: medians and totals of the concatenated per-case arrays, summed
    wall times, the FP64 reference of the first group only (the solves are
    replaced by a stand-in that records what it was asked)."""
    import importlib

    import numpy as np

    sys.path.insert(0, ROOT)
    bench = importlib.import_module("bench")
    ts = importlib.import_module("fvdbm-jax_amd.train_surrogate")
    calls = []

    def fake_steady_state(cases, weights, ref_cases=16, raw=None, **kw):
        n = len(cases)
        cold = np.array([100 + 10 * s for s in cases])
        warm = np.full(n, 50)
        raw.update(cells=np.full(n, 1000), cold_steps=cold, warm_steps=warm, ideal_steps=np.full(n, 10),
                   ok=np.ones(n, bool), du=np.full(n, 0.01), dr=np.full(n, 0.02),
                   converged_cold=np.ones(n, bool), converged_warm=np.ones(n, bool),
                   fallback=np.zeros(n, bool), wall_cold=2.0, wall_warm=1.0, inference=0.1,
                   ref=dict(cases=ref_cases) if ref_cases else None)
        calls.append((list(cases), ref_cases))

    orig = bench.steady_state, bench._build_steady
    bench.steady_state, bench._build_steady = fake_steady_state, (lambda seed: seed)
    try:
        out = bench.steady_study(range(10), "weights", 4, ts._InProcess(), ref_cases=16)
    finally:
        bench.steady_state, bench._build_steady = orig
    assert calls == [([0, 1, 2, 3], 16), ([4, 5, 6, 7], 0), ([8, 9], 0)]
    assert out["cases"] == 10 and out["groups"] == 3
    cold = 100 + 10 * np.arange(10)
    assert out["steps_cold_total"] == cold.sum() and out["steps_warm_total"] == 500
    assert out["step_ratio_median"] == np.median(cold / 50)
    assert out["wall_cold_s"] == 6.0 and out["wall_warm_s"] == 3.0 and out["wall_ratio"] == 2.0
    assert out["fp64_reference"] == dict(cases=16)
