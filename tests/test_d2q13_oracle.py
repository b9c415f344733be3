"""Pins for the oracle's D2Q13 step (src/dynamics.py:77-102), on the CPU.

As for D2Q9, the mathematics fixes most of the lattice: the weights and
velocities by the lattice moments (sum W = 1, sum W xi xi = c^2 I,
isotropic fourth moment), the equilibrium by its moments (density,
momentum, momentum flux), and the step by the uniform fixed point, mass
balance, a brute-force two-triangle step (upwind and Lax-Wendroff) and the
mirror symmetry. The two
cubic equilibrium terms are pinned in ratio: sum xi feq = rho u holds only
when the cu^3 and cu uu coefficients are in the reference's 1 : 3 ratio.
Their common factor (the reference's 1/2 where a third-order Hermite
expansion has 1/6) no identity fixes: that part is the reference's
expression as written (parity unpinned). ``test_planted_mistakes_q13``
shows each planted mistake fails a check.
"""
import numpy as np
import pytest

from cases import build_mesher, fv, ldc_case, meshgen, oracle_env, oracle_state
from oracle import oracle

Q = 13
KSI = oracle.KSI13.astype(np.float64)
W = oracle.W13_64
CS2 = 0.5
MIRROR_Q = [0, 3, 2, 1, 4, 6, 5, 8, 7, 11, 10, 9, 12]


def feq_ref(rho, u):
    """The reference's D2Q13.calc_eq in float64 (its expression as written)."""
    cu = u @ KSI.T
    uu = np.sum(u * u, axis=-1, keepdims=True)
    c2, c4, c6 = CS2, CS2 ** 2, CS2 ** 3
    br = 1 + cu / c2 + cu ** 2 / (2 * c4) - uu / (2 * c2) + cu ** 3 / (2 * c6) - 3 * cu * uu / (2 * c4)
    return W * np.asarray(rho)[..., None] * br


def random_state(C, seed, amp=0.02):
    rng = np.random.default_rng(seed)
    rho = 1 + 0.05 * rng.standard_normal(C)
    u = 0.04 * rng.standard_normal((C, 2))
    return feq_ref(rho, u) * (1 + amp * rng.standard_normal((C, Q)))


def test_lattice_moments():
    """Weights and velocities: normalised, isotropic to fourth order, c^2 = 1/2."""
    assert W.sum() == pytest.approx(1.0, abs=1e-15)
    assert np.allclose(W @ KSI, 0.0)
    assert np.allclose(np.einsum("q,qa,qb->ab", W, KSI, KSI), CS2 * np.eye(2), atol=1e-15)
    m4 = np.einsum("q,qa,qb,qc,qd->abcd", W, KSI, KSI, KSI, KSI)
    d = np.eye(2)
    iso = CS2 ** 2 * (np.einsum("ab,cd->abcd", d, d) + np.einsum("ac,bd->abcd", d, d)
                      + np.einsum("ad,bc->abcd", d, d))
    assert np.allclose(m4, iso, atol=1e-15)
    assert np.allclose(np.einsum("q,qa,qb,qc->abc", W, KSI, KSI, KSI), 0.0)
    # the constants the oracle and the dynamics class use are these, in f32
    assert np.array_equal(fv.D2Q13.W, W.astype(np.float32))
    assert np.array_equal(fv.D2Q13.KSI, oracle.KSI13)
    assert fv.D2Q13.C == np.float32(1) / np.sqrt(np.float32(2))


def test_equilibrium_moments():
    rng = np.random.default_rng(0)
    rho = rng.uniform(0.8, 1.2, 300)
    u = rng.uniform(-0.12, 0.12, (300, 2))
    f = oracle.calc_eq(rho, u, "f64", 13)
    assert np.allclose(f.sum(1), rho, rtol=1e-14)
    assert np.allclose(f @ KSI, rho[:, None] * u, atol=1e-15)
    P = np.einsum("nq,qa,qb->nab", f, KSI, KSI)
    assert np.allclose(P, rho[:, None, None] * (CS2 * np.eye(2) + np.einsum("na,nb->nab", u, u)),
                       atol=1e-14)
    # the f32 oracle and the host class agree with the f64 oracle to f32 rounding
    f32 = oracle.calc_eq(rho, u, "f32", 13)
    assert np.allclose(f32, f, rtol=2e-6, atol=1e-7)
    host = fv.D2Q13(0.6, 0.01).calc_eq(rho.astype(np.float32)[:, None], u.astype(np.float32))
    assert np.allclose(host, f32, rtol=2e-6, atol=1e-7)
    # the rest state is W
    assert np.allclose(oracle.calc_eq(np.ones(1), np.zeros((1, 2)), "f64", 13)[0], W, atol=1e-16)


def test_cubic_ratio_is_what_keeps_momentum():
    """With the cu uu coefficient changed from 3 to 1, sum xi feq != rho u:
    the momentum check pins the ratio of the two cubic terms."""
    rng = np.random.default_rng(1)
    rho = rng.uniform(0.8, 1.2, 50)
    u = rng.uniform(-0.12, 0.12, (50, 2))
    cu = u @ KSI.T
    uu = np.sum(u * u, -1, keepdims=True)
    good = feq_ref(rho, u)
    bad = good + W * rho[:, None] * (2 * cu * uu / (2 * CS2 ** 2))  # -3 -> -1
    assert np.allclose(good @ KSI, rho[:, None] * u, atol=1e-15)
    assert not np.allclose(bad @ KSI, rho[:, None] * u, atol=1e-6)


# ------------------------------------------------------------- the step ----
def _ldc13(mutation, nsteps=1, state=None, u_lid=0.1):
    m, dyn, cells, faces, nodes = ldc_case(u_lid=u_lid, lattice=fv.D2Q13)
    env = oracle_env(cells, faces, nodes, "f64")
    st = oracle_state(env, nodes, "f64", Q)
    if state is not None:
        st["cell_pdf"] = state.copy()
    f0 = st["cell_pdf"].copy()
    oracle.run(env, st, dyn.tau, dyn.delta_t, nsteps, "f64", mutation, quivers=Q)
    return m, dyn, cells, faces, nodes, env, f0, st


def check_moments(mutation):
    f = random_state(2048, 1)
    *_, f0, st = _ldc13(mutation, 1, f)
    rho = f0.sum(1)
    u = (f0 @ KSI) / rho[:, None]
    feq = st["cell_feq"]
    ok = np.allclose(st["cell_rho"], rho, rtol=1e-13) and np.allclose(st["cell_vel"], u, atol=1e-13)
    ok &= np.allclose(feq.sum(1), rho, rtol=1e-13)
    ok &= np.allclose(feq @ KSI, rho[:, None] * u, atol=1e-13)
    P = np.einsum("cq,qi,qj->cij", feq, KSI, KSI)
    ok &= np.allclose(P, rho[:, None, None] * (CS2 * np.eye(2) + np.einsum("ci,cj->cij", u, u)),
                      atol=1e-12)
    return bool(ok)


def check_node_constant(mutation):
    g = random_state(1, 2)[0]
    f = np.repeat(g[None], 2048, axis=0)
    m, dyn, cells, faces, nodes, env, f0, st = _ldc13(mutation, 1, f)
    rho_g = g.sum()
    u_g = (g @ KSI) / rho_g
    vn = np.nonzero(nodes.type.reshape(-1) == 1)[0]
    un = np.asarray(nodes.vel_f64)[vn]
    expect = feq_ref(np.full(len(vn), rho_g), un) + (g - feq_ref(np.array([rho_g]), u_g[None])[0])
    return bool(np.allclose(st["node_rho"][vn], rho_g, rtol=1e-13)
                and np.allclose(st["node_pdf"][vn], expect, atol=1e-13))


def check_fixed_point(mutation):
    *_, f0, st = _ldc13(mutation, 200, None, u_lid=0.0)
    return bool(np.max(np.abs(st["cell_pdf"] - f0)) < 1e-13)


def check_mass_balance(mutation):
    f = random_state(2048, 3)
    m, dyn, cells, faces, nodes, env, f0, st = _ldc13(mutation, 1, f)
    bnd = faces.stencil_cells_index[:, 1] < 0
    out = st["face_pdf"][bnd].sum()
    lhs = st["cell_pdf"].sum() - f0.sum()
    return bool(abs(lhs + dyn.delta_t * out) < 1e-12 * np.abs(f0).sum())


def brute_force_step(points, tris, f, node_type, node_rho, node_vel, tau, dt, scheme="upwind"):
    """One step from the method's definition (see the D2Q9 version in
    test_oracle.py, upwind or Lax-Wendroff face values), with the thirteen
    velocities."""
    C = len(tris)
    cen = points[tris].mean(axis=1)
    rho = f.sum(1)
    u = (f @ KSI) / rho[:, None]
    feq = feq_ref(rho, u)
    npdf = np.repeat(W[None], len(points), axis=0)
    for i in range(len(points)):
        if node_type[i] == 0:
            continue
        cs = [c for c in range(C) if i in tris[c]]
        w = np.array([1 / np.linalg.norm(points[i] - cen[c]) for c in cs])
        if node_type[i] == 1:
            r, v = np.average(rho[cs], weights=w), node_vel[i]
        else:
            r, v = node_rho[i], np.average(u[cs], axis=0, weights=w)
        npdf[i] = feq_ref(np.array([r]), v[None])[0] + np.average(f[cs] - feq[cs], axis=0, weights=w)
    owners = {}
    for c, t in enumerate(tris):
        for k in range(3):
            owners.setdefault(frozenset((t[k], t[(k + 1) % 3])), []).append(c)
    new = f.copy()
    for c, t in enumerate(tris):
        total = np.zeros(Q)
        for k in range(3):
            a, b = t[k], t[(k + 1) % 3]
            e = points[b] - points[a]
            L = np.hypot(*e)
            n = np.array([e[1], -e[0]]) / L
            if np.dot(n, cen[c] - (points[a] + points[b]) / 2) > 0:
                n = -n
            other = [o for o in owners[frozenset((a, b))] if o != c]
            fo = f[other[0]] if other else (npdf[a] + npdf[b]) - f[c]
            mid = (points[a] + points[b]) / 2
            dU = abs(np.dot(n, cen[c] - mid))
            dD = abs(np.dot(n, cen[other[0]] - mid)) if other else dU
            cn = KSI @ n
            if scheme == "upwind":
                face = np.where(cn >= 0, f[c], fo)
            else:
                face = f[c] + (fo - f[c]) * (dU / (dU + dD) - cn * dt / (2 * (dU + dD)))
            total += L * cn * face
        new[c] = f[c] + dt * ((feq[c] - f[c]) / tau - total)
    return new, npdf


def check_brute_force(mutation, scheme="upwind"):
    pts = np.array([[0.0, 0.0], [1.0, 0.1], [1.1, 1.0], [-0.1, 0.9]])
    raw = meshgen.RawMesh(pts, np.array([[0, 1, 2], [0, 2, 3]], np.int32),
                          np.array([[0, 1], [1, 2], [2, 0], [2, 3], [3, 0]], np.int32),
                          np.array([1, 1, 2, 2], np.int32))
    m = build_mesher(raw)
    dyn = fv.D2Q13(tau=0.7, delta_t=0.05)
    cells, faces, nodes = m.to_env(dyn, scheme)
    nodes = m.set_vel_node(nodes, 1, np.array([0.03, -0.02]))
    nodes = m.set_rho_node(nodes, 2, 1.02)
    env = oracle_env(cells, faces, nodes, "f64")
    st = oracle_state(env, nodes, "f64", Q)
    rng = np.random.default_rng(4)
    f = feq_ref(np.array([1.0, 1.03]), np.array([[0.04, 0.01], [-0.02, 0.03]])) * \
        (1 + 0.05 * rng.standard_normal((2, Q)))
    st["cell_pdf"] = f.copy()
    oracle.run(env, st, dyn.tau, dyn.delta_t, 1, "f64", mutation, scheme=scheme, quivers=Q)
    new, npdf = brute_force_step(m.points, m.cells, f, nodes.type.reshape(-1),
                                 nodes.rho_f64.reshape(-1), nodes.vel_f64, dyn.tau, dyn.delta_t, scheme)
    active = nodes.type.reshape(-1) > 0
    return bool(np.allclose(st["cell_pdf"], new, rtol=0, atol=1e-14)
                and np.allclose(st["node_pdf"][active], npdf[active], rtol=0, atol=1e-14))


def check_mirror(mutation):
    raw = meshgen.ldc_structured(16)
    mir = meshgen.RawMesh(np.stack([1 - raw.points[:, 0], raw.points[:, 1]], 1), raw.cells,
                          raw.faces, raw.point_markers)
    outs = []
    for r, ulid in ((raw, 0.1), (mir, -0.1)):
        m = build_mesher(r)
        dyn = fv.D2Q13(0.6, 0.02)
        cells, faces, nodes = m.to_env(dyn, dim_multiplier=16.0)
        nodes = m.set_vel_node(nodes, meshgen.LID, np.array([ulid, 0.0]))
        nodes = m.set_vel_node(nodes, meshgen.WALL, np.array([0.0, 0.0]))
        env = oracle_env(cells, faces, nodes, "f64")
        st = oracle_state(env, nodes, "f64", Q)
        oracle.run(env, st, dyn.tau, dyn.delta_t, 50, "f64", mutation, quivers=Q)
        outs.append(st["cell_pdf"])
    return bool(np.allclose(outs[1], outs[0][:, MIRROR_Q], rtol=0, atol=1e-12))


CHECKS = {
    "moments": check_moments,
    "node_constant": check_node_constant,
    "fixed_point": check_fixed_point,
    "mass_balance": check_mass_balance,
    "brute_force": check_brute_force,
    "mirror": check_mirror,
    "lw_brute_force": lambda mut: check_brute_force(mut, "lax_wendroff"),
}


@pytest.mark.parametrize("name", list(CHECKS))
def test_faithful_q13_oracle_passes(name):
    assert CHECKS[name]("none")


@pytest.mark.parametrize("mutation", [k for k in oracle.MUTATIONS if k != "none"])
def test_planted_mistakes_q13(mutation):
    failed = [name for name, chk in CHECKS.items() if not chk(mutation)]
    assert failed, f"mutation {mutation} passes every D2Q13 check"


def test_q13_fp32_matches_fp64_on_cavity():
    m, dyn, cells, faces, nodes = ldc_case(lattice=fv.D2Q13)
    out = []
    for prec in ("f32", "f64"):
        env = oracle_env(cells, faces, nodes, prec)
        st = oracle_state(env, nodes, prec, Q)
        oracle.run(env, st, dyn.tau, dyn.delta_t, 300, prec, quivers=Q)
        out.append(st["cell_pdf"].astype(np.float64))
    rel = np.linalg.norm(out[0] - out[1]) / np.linalg.norm(out[1])
    assert rel < 1e-5, rel
