"""Start states that exercise both divisions of the FP32 equilibrium.

The kernels divide by the lattice constants with an FMA-corrected reciprocal
when u.u < 1/2 and with the IEEE division otherwise, and rely on quotients of
dividends below 2^-102 being absorbed in the equilibrium's brackets. The
cells of the cavity get pdfs whose velocities are exactly zero, tiny (a huge
rest population makes u = m / rho as small as a subnormal), with ux = -uy or
ux = uy (ksi.u of q5 or q6 exactly zero), moderate, on either side of
u.u = 1/2 (D2Q9), and beyond it. One and two steps from such a state (a
single launch, a two-step launch), which stay finite, must equal the oracle
bit for bit, cells.pdf_eq (the equilibrium of the start state after one
step) included: the tuned D2Q9 kernels, the any-lattice kernels on D2Q9 and
on D2Q13.
"""
import numpy as np
import pytest

from cases import fv, ldc_case, oracle_env, oracle_state
from oracle import oracle

pytestmark = pytest.mark.gpu

FIELDS = [("cells", "pdf", "cell_pdf"), ("cells", "rho", "cell_rho"), ("cells", "vel", "cell_vel"),
          ("cells", "pdf_eq", "cell_feq"), ("faces", "pdf", "face_pdf"),
          ("nodes", "pdf", "node_pdf"), ("nodes", "rho", "node_rho"), ("nodes", "vel", "node_vel")]

# (rest population f0 or None for W0, {q: added amount}); q = 1..4 are the
# axis velocities (1: +x, 2: +y, 3: -x, 4: -y), so m = (f1 - f3, f2 - f4)
RECIPES = [
    (None, {}),                                   # u = 0 exactly
    (1e25, {1: 2.0 ** -26}),                      # |ux| ~ 1e-33: below 2^-102
    (3e31, {2: 2.0 ** -26}),                      # |uy| subnormal
    (1e20, {1: 2.0 ** -20, 4: 2.0 ** -20}),      # ux = -uy ~ 1e-26: ksi.u of q5 = 0
    (1e12, {1: 2.0 ** -24, 2: 2.0 ** -24}),      # ux = uy ~ 6e-20: ksi.u of q6 = 0
    (1e8, {1: 0.01, 2: 0.003}),                   # |u| ~ 1e-10
    (None, {1: 0.05, 2: 0.02}),                   # moderate
    (None, {1: 0.5, 4: 0.2}),                     # u.u ~ 0.1
    (None, {1: 2.414213}),                        # u.u just below 1/2 (0.49999997)
    (None, {1: 2.4142135}),                       # u.u just above 1/2 (0.50000006)
    (None, {1: 3.0, 2: 1.0}),                     # u.u > 1/2 (D2Q9): IEEE division
    (None, {1: 6.0}),                             # u.u > 1/2 (D2Q13; with D2Q9 the
]                                                 # second step turns non-finite)


def _recipes(quivers):
    return RECIPES if quivers == 13 else RECIPES[:-1]


def _state(cells_n, quivers):
    W = np.asarray(oracle.W13_64 if quivers == 13 else oracle.W64, np.float32)
    f = np.empty((cells_n, quivers), np.float32)
    recipes = _recipes(quivers)
    for c in range(cells_n):
        rest, add = recipes[c % len(recipes)]
        row = W.copy()
        if rest is not None:
            row[0] = np.float32(rest)
        for q, v in add.items():
            row[q] = np.float32(row[q] + np.float32(v))
        f[c] = row
    return f


def _equal(env, st, what):
    for kind, name, key in FIELDS:
        got = getattr(getattr(env, kind), name)
        exp = np.asarray(st[key]).reshape(got.shape)
        if not np.array_equal(got, exp, equal_nan=True):
            bad = np.argwhere(~((got == exp) | (np.isnan(got) & np.isnan(exp))))
            raise AssertionError(f"{what} {kind}.{name}: {len(bad)} mismatches, first at {bad[0].tolist()}")


@pytest.mark.parametrize("kind", ["tuned_d2q9", "lattice_d2q9", "lattice_d2q13"])
def test_division_paths_match_oracle(gpu, kind):
    quivers = 13 if kind == "lattice_d2q13" else 9
    m, dyn, cells, faces, nodes = ldc_case(n=8, lattice=fv.D2Q13 if quivers == 13 else None)
    f0 = _state(cells.size, quivers)
    rho, vel = oracle.calc_macro(f0, quivers)
    uu = vel[:, 0].astype(np.float64) ** 2 + vel[:, 1].astype(np.float64) ** 2
    assert (uu == 0).any() and ((uu > 0) & (uu < 2.0 ** -200)).any() and (uu >= 0.5).any()
    assert ((np.abs(vel) < 2.0 ** -126) & (vel != 0)).any() and (vel[:, 0] == -vel[:, 1]).any()
    oenv = oracle_env(cells, faces, nodes)
    env = fv.Environment(cells, faces, nodes, lattice_kernel=(kind == "lattice_d2q9"))
    env.init()
    for n in (1, 2):  # one k_step launch; one two-step launch
        st = oracle_state(oenv, nodes, "f32", quivers)
        st["cell_pdf"] = f0.copy()
        oracle.run(oenv, st, dyn.tau, dyn.delta_t, n, "f32", quivers=quivers)
        assert all(np.isfinite(np.asarray(st[key])).all() for _, _, key in FIELDS)
        env.cells.pdf = f0
        env.step(n)
        _equal(env, st, f"{kind} after {n} steps")
