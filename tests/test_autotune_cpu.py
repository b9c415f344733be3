"""Autotuner host logic: the full-batch confirmation times the best
candidate of each of the two fastest tile sizes, keeps the faster handle
open and closes the other (solver and timer replaced by stand-ins)."""
import importlib

at = importlib.import_module("fvdbm-jax_amd.autotune")
batch = importlib.import_module("fvdbm-jax_amd.batch")


class _FakeSolver:
    made = []

    def __init__(self, cases, precision="fp32", block=0, variant=0, reorder=4):
        self.block, self.variant, self.reorder = block, variant, reorder
        self.closed = False
        self.steps = 0
        _FakeSolver.made.append(self)

    def step(self, n):
        self.steps += n

    def close(self):
        self.closed = True


def test_confirm_batch_keeps_the_faster_tile_size(monkeypatch):
    _FakeSolver.made = []
    monkeypatch.setattr(batch, "BatchSolver", _FakeSolver)
    full_ms = {128: 46.0, 256: 52.0, 512: 60.0}  # per 100 steps on the whole batch
    monkeypatch.setattr(at, "_time", lambda obj, steps: full_ms[obj.block] * steps / 100)
    sample = [dict(block=256, variant=0, reorder=4, us_per_step=92.7),
              dict(block=256, variant=8, reorder=4, us_per_step=93.0),
              dict(block=128, variant=0, reorder=4, us_per_step=94.3),
              dict(block=512, variant=0, reorder=4, us_per_step=105.2)]
    logs = []
    best, timings, solver = at.confirm_batch(["case"], sample, log=logs.append)
    assert best == dict(block=128, variant=0, reorder=4)
    assert [(t["block"], t["variant"]) for t in timings] == [(256, 0), (128, 0)]
    assert [t["us_per_step"] for t in timings] == [520.0, 460.0]
    assert len(_FakeSolver.made) == 2 and solver is _FakeSolver.made[1]
    assert _FakeSolver.made[0].closed and not solver.closed and solver.steps == 20
    assert len(logs) == 2


def test_confirm_batch_keeps_the_first_when_it_stays_fastest(monkeypatch):
    _FakeSolver.made = []
    monkeypatch.setattr(batch, "BatchSolver", _FakeSolver)
    monkeypatch.setattr(at, "_time", lambda obj, steps: {128: 40.0, 256: 50.0}[obj.block])
    sample = [dict(block=128, variant=0, reorder=4, us_per_step=90.0),
              dict(block=256, variant=0, reorder=4, us_per_step=95.0)]
    best, timings, solver = at.confirm_batch(["case"], sample, top=2)
    assert best["block"] == 128 and solver is _FakeSolver.made[0]
    assert _FakeSolver.made[1].closed and not solver.closed
