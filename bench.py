"""Benchmark: MLUPS of the FVDBM step on the 1M-cell cylinder channel.

Workload: channel [0,2.2]x[0,0.41] with a cylinder r=0.05 at (0.2,0.2),
Delaunay mesh of seeded jittered points (~1.0M cells), parabolic velocity
inlet u_max=0.1, density outlet rho=1, no-slip walls and cylinder, D2Q9,
tau=0.55, first-order upwind, FP32. The update has no 1/A term (as in the
reference), so dim_multiplier = 1/mean(cell area) and dt is the explicit
stability bound at CFL 0.5; both are reported.

Steps run as two-step launches (k_step2: each advances every cell by two
steps, the intermediate state held in LDS; an odd count starts with one
single step, k_step). The timed region is K steps between two HIP events
on the solver's stream (launched as captured hipGraphs where K allows),
with the inputs resident in HBM. value = cells * K / elapsed / 1e6 (MLUPS). roofline.achieved = the
algorithmic 108 B per cell update * cells * 2 per k_step2 launch / the
launch's duration, i.e. 108 B/LUP * cells / (elapsed/K).
cpu_baseline = the C oracle (OpenMP over the element loops, min(16, CPUs)
threads) on a bounded number of steps of the same mesh, on this host.

fp32_vs_fp64: relative L2 of u, rho and rho - 1 against the same kernels in FP64
after a fixed number of steps from the uniform state: the channel after
``--fp64-steps`` (10k) steps, the 32x32 cavity after 1000; the batched and
steady-state lines carry their own.

batched: ``--batch-cases`` (default 256) porous cases of the batched
workload (5-40 discs, 20k-200k cells, seeds 0..) in one handle: the
autotuner picks tile size and kernel variant on the first 32 cases (every
candidate checked bitwise), then MLUPS of the batched step over all their
cells, then a masked run to a 1e-6 residual capped at ``--masked-steps``
(700k: every case of the 256 is decided by 630k) steps, with a runaway stop (mean density past 2: most of these
channels are choked and have no steady state), reported as converged /
runaway / still-running counts; FP32 against FP64 on 16 cases after 20k
steps.

steady_state (the second half of the metric): ``--steady-cases`` (64)
unseen porous channels of the steady-state distribution (3-8 discs,
20k-200k cells; seeds from 1,000,000, the surrogate's training seeds lie
below) solved as one batch to a per-step residual of 1e-6, first from the
uniform state (cold), then from the surrogate's prediction (warm, cold
fallback per case). Reported: steps per case against the 3x target, batch
wall time (the warm time includes input preparation and inference), and the
final fields: warm against cold, and both against an FP64 reference (the
first ``--steady-ref-cases`` cases, the cold solution continued in FP64 to
a residual of 1e-10) against the 1e-4 target.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--parts single,batched,steady]
"""
from __future__ import annotations

import argparse
import importlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

BYTES_PER_LUP_F32 = 108  # pdf 36 read + 36 write, 3 neighbour ids 12, 3 (L n) pairs 24
BYTES_PER_LUP_F64 = 204


def bytes_per_lup(quivers, precision):
    """Algorithmic bytes of one cell update: pdf read and write (Q values
    each), 3 neighbour ids, 3 (L, n) pairs."""
    es = 4 if precision == "fp32" else 8
    return 2 * quivers * es + 12 + 3 * 2 * es
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_case(cells, seed, scheme="upwind", cache_dir=None, lattice="d2q9"):
    fv = importlib.import_module("fvdbm-jax_amd")
    mg = fv.meshgen
    t0 = time.time()
    path = os.path.join(cache_dir, f"cylinder_{cells}_{seed}.npz") if cache_dir else None
    if path and os.path.exists(path):
        z = np.load(path)
        raw = mg.RawMesh(z["points"], z["cells"], z["faces"], z["point_markers"])
    else:
        raw = mg.cylinder_channel(cells, seed=seed)
        if path:
            os.makedirs(cache_dir, exist_ok=True)
            np.savez(path, points=raw.points, cells=raw.cells, faces=raw.faces,
                     point_markers=raw.point_markers)
    m = fv.Mesher.from_raw(raw)
    m.enforce_ccw()
    m.calc_mesh_properties()
    p = m.points[m.cells]
    area = 0.5 * np.abs((p[:, 1, 0] - p[:, 0, 0]) * (p[:, 2, 1] - p[:, 0, 1])
                        - (p[:, 2, 0] - p[:, 0, 0]) * (p[:, 1, 1] - p[:, 0, 1]))
    mult = float(1.0 / area.mean())
    tau = 0.55
    dyn = (fv.D2Q13 if lattice == "d2q13" else fv.D2Q9)(tau=tau, delta_t=1.0)
    c, f, n = m.to_env(dyn, scheme, dim_multiplier=mult)
    L = np.asarray(f.L, np.float64).reshape(-1)
    # explicit-Euler bound with the lattice's largest speed (sqrt 2, or 2 for D2Q13)
    per_cell = L[m.cell_face_indices].sum(axis=1) * (2.0 if lattice == "d2q13" else np.sqrt(2.0))
    dyn.delta_t = float(np.float32(0.5 / (1.0 / tau + per_cell.max())))
    n = m.set_vel_profile(n, mg.INLET, lambda q: mg.parabolic_inlet(q[:, 1], 0.41, 0.1))
    n = m.set_rho_node(n, mg.OUTLET, 1.0)
    n = m.set_vel_node(n, mg.WALL, np.array([0.0, 0.0]))
    n = m.set_vel_node(n, mg.OBSTACLE, np.array([0.0, 0.0]))
    log(f"mesh: {len(m.cells)} cells, {len(m.faces)} faces, {len(m.points)} nodes "
        f"({time.time() - t0:.1f}s)")
    return fv, m, dyn, c, f, n, dict(dim_multiplier=mult, dt=dyn.delta_t, tau=tau)


def _moments(env):
    """Cell rho and u of the current pdf (the library's kernel on the tuned
    handle, the host for the any-lattice handle)."""
    if not getattr(env, "_lattice", False):
        return env.cell_fields()
    pdf = env.cells.pdf.astype(np.float64)
    rho = pdf.sum(1)
    return rho, (pdf @ np.asarray(env.cells.dynamics.KSI, np.float64)) / rho[:, None]


def cpu_baseline(m, dyn, c, f, n, budget_s=12.0, threads=None):
    """The oracle (C, element loops split over ``threads`` OpenMP threads,
    default min(16, CPUs): the host share of one GPU) on a bounded number
    of steps."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle import oracle
    threads = oracle.set_threads(threads or min(16, os.cpu_count() or 1))
    env = dict(cell_faces=c.face_indices, cell_signs=c.face_normals,
               face_cells=f.stencil_cells_index, face_dists=f.stencil_dists,
               face_nodes=f.nodes_index, face_n=f.n, face_L=np.asarray(f.L).reshape(-1),
               node_cells=n.cells_index, node_dists=n.cell_dists)
    Q = int(dyn.NUM_QUIVERS)
    st = oracle.initial_state(env, n.type, n.__dict__["_host_rho"], n.__dict__["_host_vel"], "f32", Q)
    t0 = time.perf_counter()
    oracle.run(env, st, dyn.tau, dyn.delta_t, 1, "f32", quivers=Q)
    one = time.perf_counter() - t0
    steps = max(1, min(200, int(budget_s / max(one, 1e-6))))
    t0 = time.perf_counter()
    oracle.run(env, st, dyn.tau, dyn.delta_t, steps, "f32", quivers=Q)
    el = time.perf_counter() - t0
    C = c.face_indices.shape[0]
    return dict(value=round(C * steps / el / 1e6, 3), unit="MLUPS", cores=threads, kind="port",
                sample=f"{steps} steps of the same {C}-cell mesh, FP32, oracle C restatement "
                       f"(oracle/fvdbm_oracle.c, gcc -O2 -fopenmp), {threads} threads, {el:.1f}s")


WEIGHTS = os.path.join(ROOT, "fvdbm-jax_amd", "data", "surrogate_porous.npz")


def _build_workload(seed):
    """(cells, faces, nodes) of batched-workload case ``seed`` (the mesher
    and its geometry stay in the worker)."""
    fv = importlib.import_module("fvdbm-jax_amd")
    return fv.porous.make_case(seed, dist=fv.porous.WORKLOAD).env_tuple()


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _moment_errors(r32, u32, r64, u64):
    """Relative L2 of u, of rho and of the density variation rho - 1."""
    return dict(rel_l2_u=_rel(u32, u64), rel_l2_rho=_rel(r32, r64),
                rel_l2_rho_minus_1=_rel(np.asarray(r32, np.float64) - 1.0, np.asarray(r64, np.float64) - 1.0))


def fp64_single(fv, c, f, n, steps, **opts):
    """rho and u after ``steps`` steps from the uniform state, FP32 against
    the same kernels in FP64."""
    out = {}
    for prec in ("fp32", "fp64"):
        e = fv.Environment(c, f, n, precision=prec, **opts)
        e.init()
        e.step(steps)
        out[prec] = _moments(e)
        e.close()
    return dict(steps=steps, **_moment_errors(*out["fp32"], *out["fp64"]))


def fp64_cavity(fv):
    """The 32x32 cavity (tau 0.6, dt 0.01, lid u 0.1), 1000 steps."""
    mg = fv.meshgen
    m = fv.Mesher.from_raw(mg.ldc_structured(32))
    m.enforce_ccw()
    m.calc_mesh_properties()
    dyn = fv.D2Q9(tau=0.6, delta_t=0.01)
    c, f, n = m.to_env(dyn, "upwind", dim_multiplier=1)
    n = m.set_vel_node(n, mg.LID, np.array([0.1, 0.0]))
    n = m.set_vel_node(n, mg.WALL, np.array([0.0, 0.0]))
    return fp64_single(fv, c, f, n, 1000)


def batched(triples, steps=1000, masked_steps=700_000, check_every=2000, tune_cases=32,
            fp64_cases=16, fp64_steps=20_000, block=None):
    """Throughput of one batched launch over uneven porous cases (the
    batched workload's distribution), then a masked run: every case steps
    until its residual is below 1e-6 (converged), its mean density passes 2
    (runaway: a choked channel has no steady state) or ``masked_steps`` have
    run. The launch options are tuned on the first ``tune_cases`` cases
    (every candidate checked bitwise) and the best candidates of the two
    fastest tile sizes timed again on the whole batch, or fixed to tile size
    ``block`` with two-step tiles. FP32 against FP64 on the first ``fp64_cases`` cases after
    ``fp64_steps`` steps (per-case medians; 0 cases: skipped)."""
    batch = importlib.import_module("fvdbm-jax_amd.batch")
    ws = importlib.import_module("fvdbm-jax_amd.warmstart")
    _lib = importlib.import_module("fvdbm-jax_amd._lib")
    import ctypes as ct
    at = importlib.import_module("fvdbm-jax_amd.autotune")
    b, confirmed = None, None
    if block:
        best, default = dict(block=int(block), variant=0, reorder=4), []
    else:
        tuned = at.tune_batch(triples[:tune_cases],
                              cands=at.candidates(blocks=(128, 256, 512), variants=(0, 8, 2, 4),
                                                  reorders=(4,)),
                              steps=200, cache=None, log=log)
        default = [r for r in tuned["results"] if (r["block"], r["variant"]) == (256, 0)]
        if len(triples) > tune_cases:
            # the two best tile sizes of the sample, timed on the whole batch
            t0 = time.perf_counter()
            best, confirmed, b = at.confirm_batch(triples, tuned["results"], log=log)
            log(f"batched handles: {len(triples)} cases, two tile sizes in {time.perf_counter() - t0:.1f}s")
        else:
            best = tuned["best"]
    if b is None:
        t0 = time.perf_counter()
        b = batch.BatchSolver(triples, block=best["block"], variant=best["variant"], reorder=best["reorder"])
        log(f"batched handle: {len(triples)} cases uploaded in {time.perf_counter() - t0:.1f}s")
    cells = int(sum(c.face_indices.shape[0] for c, _, _ in triples))
    info = b.info()
    b.step(100)  # queued: the timed run below starts behind it
    ms = ct.c_double()
    _lib.check(_lib.lib().fvdbm_step_timed(b._handle(), int(steps), ct.byref(ms)), "fvdbm_step_timed")
    per = ms.value / 1e3 / steps
    mlups = cells / per / 1e6
    masked = None
    if masked_steps > 0:
        b.reset_uniform()
        b.set_active(np.ones(b.n, np.int32), reset_steps=True)
        st = ws.run_until(b, masked_steps, check_every, 1e-6, rho_max=ws.RUNAWAY_RHO, chunk=10_000,
                          log=lambda m: log("masked: " + m))
        masked = _masked_summary(st, masked_steps, check_every)
    b.close()
    fp64 = _batch_fp64(triples[:fp64_cases], best, fp64_steps) if fp64_cases > 0 else None
    return dict(cases=len(triples), cells=cells, padded_cells=int(info["n_cells_padded"]), steps=steps,
                launch_classes=int(info["launch_classes"]),
                tuned=dict(block=best["block"], variant=best["variant"], reorder=best["reorder"],
                           on_cases=0 if block else min(tune_cases, len(triples)),
                           full_batch=confirmed),
                default_us_per_step=default[0]["us_per_step"] if default else None,
                us_per_step=round(per * 1e6, 3), mlups=round(mlups, 1),
                hbm_frac=round(BYTES_PER_LUP_F32 * cells / per / 1e9 / HBM_PEAK_GBS, 4),
                masked=masked, fp32_vs_fp64=fp64)


def _masked_summary(st, masked_steps, check_every, chunk=10_000):
    """Counts per outcome of a masked run; for the cases still running at the
    cap, how many still gain mass (mean density up over the last ``chunk``
    steps) and their median density, its growth and residual."""
    ws = importlib.import_module("fvdbm-jax_amd.warmstart")
    conv = st["status"] == ws.CONVERGED
    running = st["active"]
    trend = None
    if running.any():
        grow = (st["rho_last"] - st["rho_prev"])[running]
        trend = dict(gaining_mass=int((grow > 0).sum()),
                     mean_rho_median=float(np.median(st["rho_last"][running])),
                     mean_rho_growth_per_10k_steps_median=float(np.median(grow) * 10_000 / chunk),
                     residual_median=float(np.median(st["residual"][running])))
    return dict(running_trend=trend, max_steps=masked_steps, check_every=check_every,
                runaway_mean_rho=ws.RUNAWAY_RHO,
                wall_s=round(st["wall_s"], 2), converged=int(conv.sum()),
                runaway=int(st["runaway"].sum()), running=int(running.sum()),
                diverged=int((st["status"] == ws.DIVERGED).sum()),
                steps_converged_median=float(np.median(st["steps"][conv])) if conv.any() else None,
                steps_runaway_median=float(np.median(st["steps"][st["runaway"]]))
                if st["runaway"].any() else None)


def _batch_fp64(sub, best, fp64_steps):
    """FP32 against FP64 after a fixed number of steps, per case."""
    batch = importlib.import_module("fvdbm-jax_amd.batch")
    res = {}
    for prec in ("fp32", "fp64"):
        bb = batch.BatchSolver(sub, precision=prec, block=best["block"], variant=best["variant"],
                               reorder=best["reorder"])
        bb.step(fp64_steps)
        r, u = bb.cell_fields()
        res[prec] = [(bb.case_cells(k, r), bb.case_cells(k, u)) for k in range(bb.n)]
        bb.close()
    errs = [_moment_errors(*res["fp32"][k], *res["fp64"][k]) for k in range(len(sub))]
    return dict(cases=len(sub), steps=fp64_steps,
                rel_l2_u_median=float(np.median([e["rel_l2_u"] for e in errs])),
                rel_l2_u_max=float(np.max([e["rel_l2_u"] for e in errs])),
                rel_l2_rho_median=float(np.median([e["rel_l2_rho"] for e in errs])),
                rel_l2_rho_minus_1_median=float(np.median([e["rel_l2_rho_minus_1"] for e in errs])))


def steady_state(cases, weights, tol=1e-6, check_every=1000, max_steps=800_000, ref_cases=16,
                 ref_tol=1e-10, raw=None):
    """Cold against warm-started batched solves of the same unseen cases,
    both in FP32 to a per-step residual below ``tol``; then an FP64 run of
    the first ``ref_cases`` cases, continued from the cold solution to a
    residual below ``ref_tol``, judges both final fields. A ``raw`` dict
    receives the per-case arrays (for studies run in groups)."""
    batch = importlib.import_module("fvdbm-jax_amd.batch")
    ws = importlib.import_module("fvdbm-jax_amd.warmstart")
    sg = importlib.import_module("fvdbm-jax_amd.surrogate")
    b = batch.BatchSolver([c.env_tuple() for c in cases])
    # a short run first so graph capture and first-touch are outside both timings
    ws.solve(b, None, tol=tol, check_every=check_every, max_steps=2 * check_every)
    cold = ws.solve(b, None, tol=tol, check_every=check_every, max_steps=max_steps)
    cold_pdf = b.get_pdf()
    cr, cv = b.cell_fields()
    # the best a node-field predictor can do: restart from the converged
    # solution's own node fields
    ideal = ws.solve(b, b.node_fields(), tol=tol, check_every=check_every, max_steps=max_steps)
    net = sg.MeshNet.load(weights)
    t0 = time.perf_counter()
    pred = net.predict(cases)
    t_pred = time.perf_counter() - t0
    init = (np.concatenate([p[0] for p in pred]), np.concatenate([p[1] for p in pred]))
    warm = ws.solve(b, init, tol=tol, check_every=check_every, max_steps=max_steps)
    wr, wv = b.cell_fields()
    du = np.array([ws.relative_l2(b.case_cells(k, wv), b.case_cells(k, cv)) for k in range(b.n)])
    dr = np.array([ws.relative_l2(b.case_cells(k, wr) - 1.0, b.case_cells(k, cr) - 1.0)
                   for k in range(b.n)])
    b.close()
    ok = (cold["status"] == ws.CONVERGED) & (warm["status"] == ws.CONVERGED)
    ratio = cold["steps"][ok] / np.maximum(warm["steps"][ok], 1)
    warm_s = warm["wall_s"] + t_pred
    # FP64 reference: the cold FP32 solution continued in FP64 to ref_tol
    nref = min(ref_cases, len(cases))
    ref = None
    if nref > 0:
        t1 = time.perf_counter()
        rb = batch.BatchSolver([c.env_tuple() for c in cases[:nref]], precision="fp64")
        rb.set_pdf(cold_pdf[:int(rb.cell_off[-1])].astype(np.float64))
        rb.set_active(np.ones(rb.n, np.int32), reset_steps=True)
        rs = rb.run(4_000_000, check_every, ref_tol, 0.0)
        rr, rv = rb.cell_fields()
        e = {}
        for name, (fr, fvv) in (("cold", (cr, cv)), ("warm", (wr, wv))):
            e[name] = np.array([ws.relative_l2(b.case_cells(k, fvv), rb.case_cells(k, rv))
                                for k in range(nref)])
            e[name + "_rho"] = np.array([ws.relative_l2(b.case_cells(k, fr) - 1.0,
                                                        rb.case_cells(k, rr) - 1.0) for k in range(nref)])
        rb.close()
        ref = dict(cases=nref, tol=ref_tol, converged=int((rs["status"] == ws.CONVERGED).sum()),
                   extra_steps_median=float(np.median(rs["steps"])), wall_s=round(time.perf_counter() - t1, 2),
                   rel_l2_u_cold_median=float(np.median(e["cold"])), rel_l2_u_cold_max=float(e["cold"].max()),
                   rel_l2_u_warm_median=float(np.median(e["warm"])), rel_l2_u_warm_max=float(e["warm"].max()),
                   rel_l2_rho_cold_median=float(np.median(e["cold_rho"])),
                   rel_l2_rho_warm_median=float(np.median(e["warm_rho"])))
    cells = np.array([c.n_cells for c in cases])
    if raw is not None:
        raw.update(cells=cells, cold_steps=cold["steps"].copy(), warm_steps=warm["steps"].copy(),
                   ideal_steps=ideal["steps"].copy(), ok=ok, du=du, dr=dr,
                   converged_cold=cold["status"] == ws.CONVERGED, converged_warm=warm["status"] == ws.CONVERGED,
                   fallback=warm["fallback"].copy(), wall_cold=cold["wall_s"], wall_warm=warm_s,
                   inference=t_pred, ref=ref)
    return dict(
        cases=len(cases), cells=int(cells.sum()), cells_per_case=[int(cells.min()), int(cells.max())],
        tol=tol, check_every=check_every,
        converged_cold=int((cold["status"] == ws.CONVERGED).sum()),
        converged_warm=int((warm["status"] == ws.CONVERGED).sum()),
        fallbacks=int(warm["fallback"].sum()),
        steps_cold_median=float(np.median(cold["steps"])), steps_warm_median=float(np.median(warm["steps"])),
        steps_cold_total=int(cold["steps"].sum()), steps_warm_total=int(warm["steps"].sum()),
        steps_exact_fields_median=float(np.median(ideal["steps"])),
        step_ratio_median=float(np.median(ratio)) if ratio.size else None,
        step_ratio_total=float(cold["steps"].sum() / max(warm["steps"].sum(), 1)),
        step_ratio_target=3.0,
        wall_cold_s=round(cold["wall_s"], 3), wall_warm_s=round(warm_s, 3),
        inference_s=round(t_pred, 3), wall_ratio=round(cold["wall_s"] / warm_s, 3),
        final_rel_l2_u_warm_vs_cold_median=float(np.median(du)), final_rel_l2_u_warm_vs_cold_max=float(du.max()),
        final_rel_l2_rho_warm_vs_cold_median=float(np.median(dr)),
        final_rel_l2_target=1e-4,
        fp64_reference=ref)


def _build_steady(seed):
    fv = importlib.import_module("fvdbm-jax_amd")
    return fv.porous.make_case(seed, dist=fv.porous.STEADY)


def steady_study(seeds, weights, group, pool, ref_cases=16):
    """``steady_state`` over many cases in groups of ``group``: ``pool``
    (forked before the GPU was touched) meshes the next group while the
    current one is solved, and one line is logged per group. The FP64
    reference covers the first group's first ``ref_cases`` cases. Returns
    the same record as ``steady_state`` over all cases."""
    seeds = list(seeds)
    chunks = [seeds[i:i + group] for i in range(0, len(seeds), group)]
    pending = pool.map_async(_build_steady, chunks[0], chunksize=2)
    raws = []
    t0 = time.time()
    for gi in range(len(chunks)):
        cases = pending.get()
        if gi + 1 < len(chunks):
            pending = pool.map_async(_build_steady, chunks[gi + 1], chunksize=2)
        raw = {}
        steady_state(cases, weights, ref_cases=ref_cases if gi == 0 else 0, raw=raw)
        raws.append(raw)
        ok = raw["ok"]
        log(f"steady group {gi + 1}/{len(chunks)}: {len(cases)} cases, step ratio median "
            f"{np.median(raw['cold_steps'][ok] / np.maximum(raw['warm_steps'][ok], 1)):.3f}, "
            f"{time.time() - t0:.0f}s")
        del cases
    cat = {k: np.concatenate([r[k] for r in raws]) for k in
           ("cells", "cold_steps", "warm_steps", "ideal_steps", "ok", "du", "dr", "converged_cold",
            "converged_warm", "fallback")}
    wall_cold = sum(r["wall_cold"] for r in raws)
    wall_warm = sum(r["wall_warm"] for r in raws)
    ok = cat["ok"]
    ratio = cat["cold_steps"][ok] / np.maximum(cat["warm_steps"][ok], 1)
    ref = raws[0]["ref"]
    return dict(
        cases=len(seeds), groups=len(chunks), group=group, cells=int(cat["cells"].sum()),
        cells_per_case=[int(cat["cells"].min()), int(cat["cells"].max())], tol=1e-6, check_every=1000,
        converged_cold=int(cat["converged_cold"].sum()), converged_warm=int(cat["converged_warm"].sum()),
        fallbacks=int(cat["fallback"].sum()),
        steps_cold_median=float(np.median(cat["cold_steps"])), steps_warm_median=float(np.median(cat["warm_steps"])),
        steps_cold_total=int(cat["cold_steps"].sum()), steps_warm_total=int(cat["warm_steps"].sum()),
        steps_exact_fields_median=float(np.median(cat["ideal_steps"])),
        step_ratio_median=float(np.median(ratio)) if ratio.size else None,
        step_ratio_p10_p90=[float(np.percentile(ratio, 10)), float(np.percentile(ratio, 90))] if ratio.size else None,
        step_ratio_total=float(cat["cold_steps"].sum() / max(cat["warm_steps"].sum(), 1)),
        step_ratio_target=3.0,
        wall_cold_s=round(wall_cold, 3), wall_warm_s=round(wall_warm, 3),
        inference_s=round(sum(r["inference"] for r in raws), 3), wall_ratio=round(wall_cold / wall_warm, 3),
        final_rel_l2_u_warm_vs_cold_median=float(np.median(cat["du"])),
        final_rel_l2_u_warm_vs_cold_max=float(cat["du"].max()),
        final_rel_l2_rho_warm_vs_cold_median=float(np.median(cat["dr"])),
        final_rel_l2_target=1e-4,
        fp64_reference=ref)


def load_traffic(workload):
    p = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    if os.path.exists(p):
        with open(p) as fh:
            return json.load(fh)
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--cells", type=int, default=1_000_000)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp64"])
    ap.add_argument("--lattice", default="d2q9", choices=["d2q9", "d2q13"],
                    help="d2q13: the any-lattice kernels (no batched or steady-state parts)")
    ap.add_argument("--variant", type=int, default=0)
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--reorder", type=int, default=None)
    ap.add_argument("--parts", default="single,batched,steady",
                    help="comma list of: single (the 1M channel line), batched, steady")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--fp64-steps", type=int, default=10_000,
                    help="steps of the FP32-against-FP64 comparison on the channel")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also time a 256 MiB device copy (counter calibration)")
    ap.add_argument("--mesh-cache", default=None,
                    help="directory to keep the generated mesh between runs")
    ap.add_argument("--steady-cases", type=int, default=64,
                    help="unseen porous cases (20k-200k cells) for the cold/warm steady-state study (0: skip)")
    ap.add_argument("--steady-ref-cases", type=int, default=16,
                    help="of those, cases that get the FP64 reference solution")
    ap.add_argument("--steady-seed0", type=int, default=1_000_000)
    ap.add_argument("--steady-group", type=int, default=0,
                    help="solve the steady-state study in batches of this many cases, meshing the next "
                         "batch meanwhile (for studies too large for one handle, e.g. 1024 cases)")
    ap.add_argument("--weights", default=WEIGHTS)
    ap.add_argument("--batch-cases", type=int, default=256,
                    help="porous cases of the batched workload for the batched throughput line (0: skip)")
    ap.add_argument("--batch-steps", type=int, default=1000, help="timed steps of the batched launch")
    ap.add_argument("--masked-steps", type=int, default=700_000,
                    help="step cap of the batched masked run (0: skip it)")
    ap.add_argument("--mesh-workers", type=int, default=16,
                    help="processes meshing the porous cases (0: in this process, e.g. under a "
                         "profiler that must not see forked children)")
    ap.add_argument("--batch-block", type=int, default=0,
                    help="fix the batched tile size (two-step tiles) instead of tuning it")
    ap.add_argument("--batch-fp64-cases", type=int, default=16,
                    help="batched cases of the FP32-against-FP64 comparison (0: skip it)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (this build targets a single MI355X); "
              f"--gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)
    parts = set(p.strip() for p in args.parts.split(",") if p.strip())
    if parts - {"single", "batched", "steady"}:
        raise SystemExit(f"unknown --parts {sorted(parts - {'single', 'batched', 'steady'})}")
    if args.lattice != "d2q9":  # the batched and steady-state parts run on D2Q9 handles
        parts &= {"single"}

    steady_cases = None
    steady_pool = None
    steady_seeds = range(args.steady_seed0, args.steady_seed0 + args.steady_cases)
    if ("steady" in parts and args.steady_cases > 0 and os.path.exists(args.weights)
            and 0 < args.steady_group < args.steady_cases):
        # groups meshed by a pool forked now, before anything touches the GPU
        import multiprocessing as mp
        ts = importlib.import_module("fvdbm-jax_amd.train_surrogate")
        steady_pool = (mp.get_context("fork").Pool(args.mesh_workers) if args.mesh_workers > 0
                       else ts._InProcess())
    elif "steady" in parts and args.steady_cases > 0 and os.path.exists(args.weights):
        # meshed in a fork pool before anything touches the GPU
        fvp = importlib.import_module("fvdbm-jax_amd")
        ts = importlib.import_module("fvdbm-jax_amd.train_surrogate")
        t0 = time.time()
        steady_cases = ts.build_cases(steady_seeds, workers=max(1, args.mesh_workers), dist=fvp.porous.STEADY)
        log(f"steady-state study: {len(steady_cases)} cases meshed in {time.time() - t0:.1f}s")
    batch_cases = None
    if "batched" in parts and args.batch_cases > 0:
        import multiprocessing as mp
        t0 = time.time()
        if args.mesh_workers > 0:
            with mp.get_context("fork").Pool(args.mesh_workers) as pool:
                batch_cases = pool.map(_build_workload, range(args.batch_cases))
        else:
            batch_cases = [_build_workload(k) for k in range(args.batch_cases)]
        log(f"batched workload: {len(batch_cases)} cases meshed in {time.time() - t0:.1f}s")

    copy = None
    if args.copy_probe:
        _lib = importlib.import_module("fvdbm-jax_amd._lib")
        nbytes = 256 << 20
        cms = _lib.copy_probe(nbytes, 20)
        copy = dict(bytes=nbytes, ms=round(cms, 5), GBps=round(2 * nbytes / (cms / 1e3) / 1e9, 1))
        log(f"copy probe: 256 MiB in {cms:.4f} ms -> {copy['GBps']} GB/s (read + write)")
        for mib in (16, 32, 64, 128):
            t = _lib.copy_probe(mib << 20, 50)
            copy[f"GBps_{mib}MiB"] = round(2 * (mib << 20) / (t / 1e3) / 1e9, 1)
        log(f"copy probe by size: {copy}")
    out = {
        "metric": "MLUPS at 1M cells (%HBM roofline); time-to-steady-state, warm vs cold start",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if args.precision == "fp32" else "f64",
        "data": "synthetic (seeded jittered-Delaunay meshes, uniform equilibrium start)",
    }
    if "single" in parts:
        fv, m, dyn, c, f, n, case = build_case(args.cells, args.seed, cache_dir=args.mesh_cache,
                                               lattice=args.lattice)
        C = len(m.cells)
        opts = dict(variant=args.variant, block=args.block, reorder=args.reorder)
        env = fv.Environment(c, f, n, precision=args.precision, **opts)
        t0 = time.time()
        env.init()
        info = env.info()
        log(f"upload {time.time() - t0:.1f}s; bandwidth {info['bandwidth']}, boundary cells "
            f"{info['n_boundary_cells']}, device bytes {info['device_bytes'] / 1e6:.1f} MB")
        # graphs captured first; the warm-up is queued right before the timed
        # run, so the device is busy when the first event is recorded and the
        # timed region holds the K steps' work, not the host's submission
        env.prepare(args.steps)
        env.step(args.warmup)
        ms = env.time_steps(args.steps)
        if getattr(env, "_lattice", False):  # the any-lattice handle has no residual kernel
            env.step(1)
            r = None
            if not np.isfinite(env.cells.pdf).all():
                raise SystemExit("non-finite pdf: the run diverged")
        else:
            r = env.step_residual(1)
            if not np.isfinite(r):
                raise SystemExit("non-finite residual: the run diverged")
        env.close()
        mlups = C * args.steps / (ms / 1e3) / 1e6
        # two-step tiles unless the handle runs one step per launch (variant bit 8)
        pairs = not (int(info["variant"]) & 8)
        lattice_handle = getattr(env, "_lattice", False)
        kernel = ("k_lat_step2" if pairs else "k_lat_step") if lattice_handle else ("k_step2" if pairs else "k_step")
        updates_per_launch = 2 if pairs else 1
        per_launch_s = ms / 1e3 / args.steps * updates_per_launch
        bpl = bytes_per_lup(int(dyn.NUM_QUIVERS), args.precision)
        achieved = bpl * C * updates_per_launch / per_launch_s / 1e9
        workload = f"cylinder_channel_{C // 1000}k" if C < 999_500 else f"cylinder_channel_{round(C / 1e6)}M"
        tr = load_traffic("cylinder_channel_1M") if abs(C - 1_000_000) < 50_000 else None
        if tr is not None and (tr.get("kernel") != kernel or args.precision != "fp32"):
            tr = None  # the committed counters are for another kernel or precision
        roof = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 4),
                    traffic=(tr["bytes_per_launch"] if tr else None))
        cpu = None if args.no_cpu_baseline else cpu_baseline(m, dyn, c, f, n)
        fp64 = None
        if args.precision == "fp32" and args.fp64_steps > 0:
            # the FP32 error of the method per workload: the same kernels in FP64,
            # after a fixed number of steps from the uniform state
            fp64 = {workload: fp64_single(fv, c, f, n, args.fp64_steps, **opts)}
            if not lattice_handle:
                fp64["cavity_32x32"] = fp64_cavity(fv)
            log(f"fp32 vs fp64: {fp64}")
        out.update({
            "value": round(mlups, 1), "unit": "MLUPS", "ms_per_step": round(ms / args.steps, 5),
            "config": {"workload": workload, "cells": C, "faces": len(m.faces), "nodes": len(m.points),
                       "scheme": "upwind", "tau": case["tau"], "dt": case["dt"],
                       "dim_multiplier": case["dim_multiplier"], "seed": args.seed,
                       "lattice": args.lattice,
                       "kernel": {"k_lat_step2": "k_lat_step2 (any-lattice kernels, two cell steps per launch)",
                                  "k_lat_step": "k_lat_step (any-lattice kernels, one step per launch)",
                                  "k_step2": "k_step2 (two fused cell steps per launch)",
                                  "k_step": "k_step (one step per launch)"}[kernel],
                       "bytes_per_lup": bpl,
                       "residual_after": r},
            "roofline": roof,
            "fp32_vs_fp64": fp64,
            "cpu_baseline": cpu,
        })
    if batch_cases is not None:
        t0 = time.time()
        out["batched"] = batched(batch_cases, steps=args.batch_steps, masked_steps=args.masked_steps,
                                 block=args.batch_block, fp64_cases=args.batch_fp64_cases)
        out["batched"]["workload"] = ("porous channels, 5-40 discs, r in [0.02,0.08], 20k-200k cells, "
                                      "u_in in [0.02,0.1], tau in [0.55,0.9], seeds 0..")
        log(f"batched {time.time() - t0:.1f}s: {out['batched']}")
    if steady_cases is not None or steady_pool is not None:
        t0 = time.time()
        if steady_pool is not None:
            out["steady_state"] = steady_study(steady_seeds, args.weights, args.steady_group, steady_pool,
                                               ref_cases=args.steady_ref_cases)
            steady_pool.close()
        else:
            out["steady_state"] = steady_state(steady_cases, args.weights, ref_cases=args.steady_ref_cases)
        out["steady_state"]["workload"] = (
            "unseen porous channels (seeds from %d), 3-8 discs, r in [0.02,0.06], gap 0.03, "
            "20k-200k cells, u_in in [0.02,0.05], tau in [0.55,0.9]" % args.steady_seed0)
        log(f"steady-state study {time.time() - t0:.1f}s: {out['steady_state']}")
    if copy:
        out["copy_probe"] = copy
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
