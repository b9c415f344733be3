"""Pins for the CPU oracle (oracle/fvdbm_oracle.c).

The reference cannot run here (jax is absent) and its own tests hold no
golden vectors, so the oracle is pinned by what the mathematics fixes:
moment identities of the equilibrium, IDW of a constant field, the uniform
fixed point, global mass balance, mirror symmetry, and a brute-force step on
a two-triangle mesh and on a fan of irregular triangles, written from the
textbook cell-centred form (upwind: outflow uses the cell's own pdf,
inflow the neighbour's; Lax-Wendroff: the second-order face value from
both sides' centre distances). ``test_planted_mistakes``
shows that each plausible mistake planted in the oracle fails at least one
of these checks.
"""
import numpy as np
import pytest

from cases import build_mesher, fv, ldc_case, meshgen, oracle_env, oracle_state
from oracle import oracle

KSI = np.array(fv.D2Q9.KSI, np.float64)
W = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)
CS2 = 1.0 / 3.0


def feq_textbook(rho, u):
    """w rho (1 + 3 c.u + 4.5 (c.u)^2 - 1.5 u.u)."""
    cu = u @ KSI.T
    uu = np.sum(u * u, axis=-1, keepdims=True)
    return W * rho[..., None] * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * uu)


def random_state(C, seed, amp=0.02):
    rng = np.random.default_rng(seed)
    rho = 1 + 0.05 * rng.standard_normal(C)
    u = 0.05 * rng.standard_normal((C, 2))
    return feq_textbook(rho, u) * (1 + amp * rng.standard_normal((C, 9)))


# ---------------------------------------------------------------- checks ----
def _ldc64(mutation, nsteps=1, state=None, u_lid=0.1):
    m, dyn, cells, faces, nodes = ldc_case(u_lid=u_lid)
    env = oracle_env(cells, faces, nodes, "f64")
    st = oracle_state(env, nodes, "f64")
    if state is not None:
        st["cell_pdf"] = state.copy()
    f0 = st["cell_pdf"].copy()
    oracle.run(env, st, dyn.tau, dyn.delta_t, nsteps, "f64", mutation)
    return m, dyn, cells, faces, nodes, env, f0, st


def check_moments(mutation):
    """cell rho/u are the moments of the input pdf and pdf_eq has the D2Q9
    equilibrium moments: sum feq = rho, sum c feq = rho u,
    sum c c feq = rho (cs^2 I + u u)."""
    f = random_state(2048, 1)
    *_, f0, st = _ldc64(mutation, 1, f)
    rho = f0.sum(1)
    u = (f0 @ KSI) / rho[:, None]
    feq = st["cell_feq"]
    ok = np.allclose(st["cell_rho"], rho, rtol=1e-13) and np.allclose(st["cell_vel"], u, atol=1e-13)
    ok &= np.allclose(feq.sum(1), rho, rtol=1e-13)
    ok &= np.allclose(feq @ KSI, rho[:, None] * u, atol=1e-13)
    P = np.einsum("cq,qi,qj->cij", feq, KSI, KSI)
    Pref = rho[:, None, None] * (CS2 * np.eye(2) + np.einsum("ci,cj->cij", u, u))
    ok &= np.allclose(P, Pref, atol=1e-12)
    return bool(ok)


def check_node_constant(mutation):
    """All cells hold one non-equilibrium pdf g: a velocity node gets
    rho_n = rho(g) and pdf = feq(rho(g), u_n) + (g - feq(g))."""
    g = random_state(1, 2)[0]
    f = np.repeat(g[None], 2048, axis=0)
    m, dyn, cells, faces, nodes, env, f0, st = _ldc64(mutation, 1, f)
    rho_g = g.sum()
    u_g = (g @ KSI) / rho_g
    vnodes = np.nonzero(nodes.type.reshape(-1) == 1)[0]
    un = np.asarray(nodes.vel_f64)[vnodes]
    expect = feq_textbook(np.full(len(vnodes), rho_g), un) + (g - feq_textbook(np.array([rho_g]), u_g[None])[0])
    ok = np.allclose(st["node_rho"][vnodes], rho_g, rtol=1e-13)
    ok &= np.allclose(st["node_pdf"][vnodes], expect, atol=1e-13)
    return bool(ok)


def check_fixed_point(mutation):
    """Uniform W with every boundary node a zero-velocity node stays W."""
    *_, f0, st = _ldc64(mutation, 200, None, u_lid=0.0)
    return bool(np.max(np.abs(st["cell_pdf"] - f0)) < 1e-13)


def check_mass_balance(mutation):
    """sum_c sum_q (f^{n+1} - f^n) = -dt * sum over boundary faces of the
    outward flux (interior fluxes cancel, collision conserves mass)."""
    f = random_state(2048, 3)
    m, dyn, cells, faces, nodes, env, f0, st = _ldc64(mutation, 1, f)
    bnd = faces.stencil_cells_index[:, 1] < 0
    out = st["face_pdf"][bnd].sum()  # boundary normals point out; sign +1
    lhs = st["cell_pdf"].sum() - f0.sum()
    return bool(abs(lhs + dyn.delta_t * out) < 1e-12 * np.abs(f0).sum())


def _two_cells(scheme="upwind"):
    pts = np.array([[0.0, 0.0], [1.0, 0.1], [1.1, 1.0], [-0.1, 0.9]])
    raw = meshgen.RawMesh(pts, np.array([[0, 1, 2], [0, 2, 3]], np.int32),
                          np.array([[0, 1], [1, 2], [2, 0], [2, 3], [3, 0]], np.int32),
                          np.array([1, 1, 2, 2], np.int32))
    m = build_mesher(raw)
    dyn = fv.D2Q9(tau=0.7, delta_t=0.05)
    cells, faces, nodes = m.to_env(dyn, scheme)
    nodes = m.set_vel_node(nodes, 1, np.array([0.03, -0.02]))
    nodes = m.set_rho_node(nodes, 2, 1.02)
    return m, dyn, cells, faces, nodes


def brute_force_step(points, tris, f, node_type, node_rho, node_vel, tau, dt, scheme="upwind"):
    """One step written cell-centred from the method's definition:
    node pdfs from inverse-distance weights, ghost = 2 f_face - f_cell,
    and f += dt ((feq - f)/tau - sum_edges L (c.n_out) f_face) with
    f_face, per direction c, from the cell's own side (U, distance d_U
    from its centroid to the edge line) and the other side (D, d_D; a
    ghost sits at the cell's own distance):
      upwind:       f_U where c.n_out >= 0, else f_D;
      Lax-Wendroff: f_U + (f_D - f_U) (d_U/(d_U+d_D) - (c.n_out) dt/(2 (d_U+d_D))),
    the textbook second-order face value for speed c.n_out over the
    centre distance d_U + d_D. It does not depend on which side is
    called U: exchanging U and D negates c.n and maps the bracket to
    1 minus itself, so it is the same value seen from either cell."""
    C = len(tris)
    cen = points[tris].mean(axis=1)
    rho = f.sum(1)
    u = (f @ KSI) / rho[:, None]
    feq = feq_textbook(rho, u)
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
        npdf[i] = feq_textbook(np.array([r]), v[None])[0] + np.average(f[cs] - feq[cs], axis=0, weights=w)
    # edges -> cells
    owners = {}
    for c, t in enumerate(tris):
        for k in range(3):
            owners.setdefault(frozenset((t[k], t[(k + 1) % 3])), []).append(c)
    new = f.copy()
    for c, t in enumerate(tris):
        total = np.zeros(9)
        for k in range(3):
            a, b = t[k], t[(k + 1) % 3]
            e = points[b] - points[a]
            L = np.hypot(*e)
            n = np.array([e[1], -e[0]]) / L  # outward for a CCW triangle
            if np.dot(n, cen[c] - (points[a] + points[b]) / 2) > 0:
                n = -n
            other = [o for o in owners[frozenset((a, b))] if o != c]
            mid = (points[a] + points[b]) / 2
            dU = abs(np.dot(n, cen[c] - mid))
            if other:
                fo = f[other[0]]
                dD = abs(np.dot(n, cen[other[0]] - mid))
            else:
                fo = 2 * (npdf[a] + npdf[b]) / 2 - f[c]
                dD = dU
            cn = KSI @ n
            if scheme == "upwind":
                face = np.where(cn >= 0, f[c], fo)
            else:
                face = f[c] + (fo - f[c]) * (dU / (dU + dD) - cn * dt / (2 * (dU + dD)))
            total += L * cn * face
        new[c] = f[c] + dt * ((feq[c] - f[c]) / tau - total)
    return new, npdf


def check_brute_force(mutation, scheme="upwind"):
    m, dyn, cells, faces, nodes = _two_cells(scheme)
    env = oracle_env(cells, faces, nodes, "f64")
    st = oracle_state(env, nodes, "f64")
    rng = np.random.default_rng(4)
    f = feq_textbook(np.array([1.0, 1.03]), np.array([[0.04, 0.01], [-0.02, 0.03]])) * \
        (1 + 0.05 * rng.standard_normal((2, 9)))
    st["cell_pdf"] = f.copy()
    oracle.run(env, st, dyn.tau, dyn.delta_t, 1, "f64", mutation, scheme=scheme)
    new, npdf = brute_force_step(m.points, m.cells, f, nodes.type.reshape(-1),
                                 nodes.rho_f64.reshape(-1), nodes.vel_f64, dyn.tau, dyn.delta_t,
                                 scheme)
    active = nodes.type.reshape(-1) > 0
    return bool(np.allclose(st["cell_pdf"], new, rtol=0, atol=1e-14)
                and np.allclose(st["node_pdf"][active], npdf[active], rtol=0, atol=1e-14))


def _fan(n=7, seed=5):
    """n irregular triangles around one interior point, with a velocity and
    a density marker on the rim: interior faces of unequal centre distances
    on both sides."""
    rng = np.random.default_rng(seed)
    ang = np.sort(rng.uniform(0, 2 * np.pi, n))
    rad = rng.uniform(0.7, 1.3, n)
    pts = np.vstack([[[0.05, -0.03]], np.stack([rad * np.cos(ang), rad * np.sin(ang)], 1)])
    tris = np.array([[0, 1 + i, 1 + (i + 1) % n] for i in range(n)], np.int32)
    edges = [[0, 1 + i] for i in range(n)] + [[1 + i, 1 + (i + 1) % n] for i in range(n)]
    markers = np.array([0] + [1 if i < n // 2 else 2 for i in range(n)], np.int32)
    return meshgen.RawMesh(pts, tris, np.array(edges, np.int32), markers)


def check_brute_force_fan(mutation, scheme):
    m = build_mesher(_fan())
    dyn = fv.D2Q9(tau=0.65, delta_t=0.04)
    cells, faces, nodes = m.to_env(dyn, scheme)
    nodes = m.set_vel_node(nodes, 1, np.array([0.02, 0.035]))
    nodes = m.set_rho_node(nodes, 2, 0.98)
    env = oracle_env(cells, faces, nodes, "f64")
    st = oracle_state(env, nodes, "f64")
    C = len(m.cells)
    rng = np.random.default_rng(6)
    f = feq_textbook(1 + 0.02 * rng.standard_normal(C), 0.04 * rng.standard_normal((C, 2))) * \
        (1 + 0.05 * rng.standard_normal((C, 9)))
    st["cell_pdf"] = f.copy()
    oracle.run(env, st, dyn.tau, dyn.delta_t, 1, "f64", mutation, scheme=scheme)
    new, _ = brute_force_step(m.points, m.cells, f, nodes.type.reshape(-1),
                              nodes.rho_f64.reshape(-1), nodes.vel_f64, dyn.tau, dyn.delta_t,
                              scheme)
    return bool(np.allclose(st["cell_pdf"], new, rtol=0, atol=1e-14))


def check_lw_fixed_point(mutation):
    """A uniform equilibrium with zero-velocity walls stays put under the
    Lax-Wendroff flux (constant fields interpolate to themselves)."""
    m, dyn, cells, faces, nodes = ldc_case(u_lid=0.0, scheme="lax_wendroff")
    env = oracle_env(cells, faces, nodes, "f64")
    st = oracle_state(env, nodes, "f64")
    f0 = st["cell_pdf"].copy()
    oracle.run(env, st, dyn.tau, dyn.delta_t, 100, "f64", mutation, scheme="lax_wendroff")
    return bool(np.max(np.abs(st["cell_pdf"] - f0)) < 1e-13)


MIRROR_Q = [0, 3, 2, 1, 4, 6, 5, 8, 7]


def check_mirror(mutation):
    """x -> 1-x mirror of the cavity with the lid velocity reversed: the
    solution is the mirror image with q relabelled (1<->3, 5<->6, 7<->8)."""
    raw = meshgen.ldc_structured(16)
    mir = meshgen.RawMesh(np.stack([1 - raw.points[:, 0], raw.points[:, 1]], 1), raw.cells,
                          raw.faces, raw.point_markers)
    outs = []
    for r, ulid in ((raw, 0.1), (mir, -0.1)):
        m = build_mesher(r)
        dyn = fv.D2Q9(0.6, 0.05)   # stable: dt * sum_faces L' |c.n| < 1
        cells, faces, nodes = m.to_env(dyn, dim_multiplier=16.0)
        nodes = m.set_vel_node(nodes, meshgen.LID, np.array([ulid, 0.0]))
        nodes = m.set_vel_node(nodes, meshgen.WALL, np.array([0.0, 0.0]))
        env = oracle_env(cells, faces, nodes, "f64")
        st = oracle_state(env, nodes, "f64")
        oracle.run(env, st, dyn.tau, dyn.delta_t, 50, "f64", mutation)
        outs.append(st["cell_pdf"])
    # cells keep their index under the mirror (same connectivity)
    return bool(np.allclose(outs[1], outs[0][:, MIRROR_Q], rtol=0, atol=1e-12))


CHECKS = {
    "moments": check_moments,
    "node_constant": check_node_constant,
    "fixed_point": check_fixed_point,
    "mass_balance": check_mass_balance,
    "brute_force": check_brute_force,
    "mirror": check_mirror,
    "brute_force_fan": lambda mut: check_brute_force_fan(mut, "upwind"),
    "lw_brute_force": lambda mut: check_brute_force(mut, "lax_wendroff"),
    "lw_brute_force_fan": lambda mut: check_brute_force_fan(mut, "lax_wendroff"),
    "lw_fixed_point": check_lw_fixed_point,
}

LW_MUTATIONS = [k for k in oracle.MUTATIONS if k.startswith("lw_")]


@pytest.mark.parametrize("name", list(CHECKS))
def test_faithful_oracle_passes(name):
    assert CHECKS[name]("none")


@pytest.mark.parametrize("mutation", [k for k in oracle.MUTATIONS if k != "none"])
def test_planted_mistakes(mutation):
    failed = [name for name, chk in CHECKS.items() if not chk(mutation)]
    assert failed, f"planted mistake {mutation!r} passes every oracle check"


@pytest.mark.parametrize("mutation", LW_MUTATIONS)
def test_lax_wendroff_mistakes_fail_the_brute_force(mutation):
    """Each Lax-Wendroff mistake (reference src/containers.py:266-277: d1
    for d0, the sign of the dt term, U and D exchanged) fails the
    brute-force Lax-Wendroff step on both meshes, while the upwind checks
    (which never take that branch) still pass."""
    assert not check_brute_force(mutation, "lax_wendroff")
    assert not check_brute_force_fan(mutation, "lax_wendroff")
    assert check_brute_force(mutation, "upwind")


def test_fp32_matches_fp64_on_cavity():
    """FP32 oracle against its FP64 twin after 1000 cavity steps."""
    m, dyn, cells, faces, nodes = ldc_case()
    res = {}
    for prec in ("f32", "f64"):
        env = oracle_env(cells, faces, nodes, prec)
        st = oracle_state(env, nodes, prec)
        oracle.run(env, st, dyn.tau, dyn.delta_t, 1000, prec)
        res[prec] = st["cell_pdf"].astype(np.float64)
    # FP32 rounding of the arithmetic and of the geometry (n, L, d) together
    l2 = np.linalg.norm(res["f32"] - res["f64"]) / np.linalg.norm(res["f64"])
    rel = np.max(np.abs(res["f32"] - res["f64"]) / np.abs(res["f64"]))
    assert l2 < 2e-6 and rel < 3e-5, (l2, rel)


def test_eq_of_rest_state_is_w():
    out = oracle.calc_eq(np.ones(3), np.zeros((3, 2)), "f32")
    assert np.array_equal(out, np.repeat(np.float32(W)[None], 3, 0))


def test_oracle_reproduces_committed_fixtures():
    """The oracle still produces the committed cavity snapshots bit for bit."""
    import os
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "ldc32_oracle.npz"))
    m, dyn, cells, faces, nodes = ldc_case()
    env = oracle_env(cells, faces, nodes, "f32")
    st = oracle_state(env, nodes, "f32")
    done = 0
    for n in (1, 10, 100, 1000):
        oracle.run(env, st, dyn.tau, dyn.delta_t, n - done, "f32")
        done = n
        assert np.array_equal(st["cell_pdf"], z[f"pdf_f32_{n}"]), n
    env64 = oracle_env(cells, faces, nodes, "f64")
    st64 = oracle_state(env64, nodes, "f64")
    oracle.run(env64, st64, dyn.tau, dyn.delta_t, 1000, "f64")
    assert np.array_equal(st64["cell_pdf"], z["pdf_f64_1000"])
