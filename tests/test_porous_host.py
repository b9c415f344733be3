"""Host logic of the porous workloads and the warm start, on the CPU:
cross-section features against closed forms, the explicit time-step bound,
case grouping for batched solves, the plausibility gate and the batched
solver's input checks."""
import importlib
import math

import numpy as np
import pytest

from cases import fv

porous = fv.porous
ws = importlib.import_module("fvdbm-jax_amd.warmstart")
ts = importlib.import_module("fvdbm-jax_amd.train_surrogate")


def test_section_features_single_disc():
    cx, cy, r = 0.4, 0.25, 0.1
    xs = np.array([0.05, cx, cx + 0.05, 0.95])
    f = porous.section_features(np.stack([xs, np.zeros(4)], 1), [(cx, cy, r)], grid=100_000)
    a = 1.0 - 2.0 * np.sqrt(np.maximum(r * r - (xs - cx) ** 2, 0.0)) / porous.HEIGHT
    assert np.allclose(f[:, 0], a, atol=1e-3)
    assert np.allclose(f[:, 1], 1.0 / a, rtol=2e-3)
    # share of the resistance downstream: 1 at the inlet, decreasing, 0 at the outlet
    assert f[0, 2] > 0.9 and f[3, 2] < 0.1 and np.all(np.diff(f[:, 2]) < 0)
    # total resistance int 1/a^3 dx over [0,1]: 1 outside the disc plus the bump
    x = np.linspace(0, 1, 400_001)
    aa = 1.0 - 2.0 * np.sqrt(np.maximum(r * r - (x - cx) ** 2, 0.0)) / porous.HEIGHT
    total = np.trapezoid(1.0 / aa ** 3, x)
    assert math.isclose(float(np.exp(f[0, 3])), total, rel_tol=1e-3)


def test_case_time_step_bound():
    c = porous.make_case(3, target_cells=2000)
    L = np.asarray(c.faces.L, np.float64).reshape(-1)
    worst = (L[c.mesher.cell_face_indices].sum(axis=1) * np.sqrt(2.0)).max()
    cfl = c.dynamics.delta_t * (1.0 / c.dynamics.tau + worst)
    assert abs(cfl - porous.CFL) < 1e-6
    # dim_multiplier makes the mean cell 1/SCALE^2 of the unit domain's area in the update
    p = c.mesher.points[c.mesher.cells]
    A = 0.5 * np.abs((p[:, 1, 0] - p[:, 0, 0]) * (p[:, 2, 1] - p[:, 0, 1])
                     - (p[:, 2, 0] - p[:, 0, 0]) * (p[:, 1, 1] - p[:, 0, 1]))
    Lraw = np.hypot(*(c.mesher.points[c.mesher.faces[:, 1]] - c.mesher.points[c.mesher.faces[:, 0]]).T)
    assert np.allclose(L / Lraw, 1.0 / (porous.SCALE * A.mean()), rtol=1e-6)


def test_groups_respect_the_cell_budget():
    class C:
        def __init__(self, n):
            self.n_cells = n
    cases = [C(n) for n in (5, 7, 3, 9, 2, 8)]
    g = ts.groups(cases, 12)
    assert [i for grp in g for i in grp] == list(range(6))
    assert all(sum(cases[i].n_cells for i in grp) <= 12 for grp in g)
    assert ts.groups([C(50)], 12) == [[0]]  # a case above the budget gets a group of its own


def test_plausibility_gate():
    rho = np.ones(10)
    vel = np.zeros((10, 2))
    assert ws.plausible(rho, vel)
    v = vel.copy()
    v[3] = [0.3, 0.0]  # Mach 0.52
    assert not ws.plausible(rho, v)
    r = rho.copy()
    r[0] = 2.5
    assert not ws.plausible(r, vel)
    v = vel.copy()
    v[1, 1] = np.inf
    assert not ws.plausible(rho, v)


def test_batch_rejects_other_lattices_and_mixed_schemes():
    """The batched kernels are D2Q9: a D2Q13 case is refused (not stepped with
    D2Q9 arithmetic), as is a batch mixing flux schemes. Both checks come
    before anything is uploaded, so no GPU is needed."""
    from cases import ldc_case
    batch = importlib.import_module("fvdbm-jax_amd.batch")
    _, _, c13, f13, n13 = ldc_case(lattice=fv.D2Q13)
    with pytest.raises(ValueError, match="D2Q9"):
        batch.BatchSolver([(c13, f13, n13)])
    _, _, c, f, n = ldc_case()
    _, _, cl, fl, nl = ldc_case(scheme="lax_wendroff")
    with pytest.raises(ValueError, match="one flux scheme"):
        batch.BatchSolver([(c, f, n), (cl, fl, nl)])
