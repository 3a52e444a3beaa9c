"""Benchmark: GDoF/s per V-cycle of the matrix-free P1 geometric multigrid on one MI355X, FP32 V-cycles inside
FP64 iterative refinement against plain FP64 V-cycles.

Headline workload: unit cube meshed as 2x2x2 sub-cubes x 6 tetrahedra (48 macro-tets), refinement level 8
(133,432,831 unknowns), 3D Poisson with b = M I_h f + U[-1, 1] * 1e-3 per DoF (f = 3 pi^2 sin(pi x) sin(pi y)
sin(pi z), seeded noise), zero Dirichlet boundary, solved from x = 0 to relative residual 1e-10.  A "step" is
one whole solve: FP32 V(2,2) cycles (multicolour SOR, omega = 1, device CG at level 0), several per FP64
refinement step, all data resident in HBM and the cycles replayed from captured hipGraphs.  `value` = unknowns
x FP32 V-cycles of the solve / solve time (GDoF/s per V-cycle); the same line gives the FP64 V(2,2) solve of
the same system (FP64 GDoF/s per V-cycle) and the ratio of the two solve times.

Prints one JSON line.  `roofline` covers the dominant kernel of the solve (the fused SOR cell phase on the
finest level, weighted over its launches as a smoothing sequence runs them), timed live with hipEvents on the
library's stream; `roofline_apply` the 15-point stencil apply.  `config2_vcycle` is the FP64 V(3,3) cycle at
level 7 (16,581,375 unknowns).  `cpu_baseline` times the CPU oracle (the reference's algorithm restated in C,
lexicographic SOR, OpenMP over macro-primitives) on a bounded sample of the headline workload.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "GDoF/s per V-cycle, mixed vs FP64 at rel. res 1e-10; serving p50/p99 latency"
HBM_PEAK_GBS = 8000.0
SEED = 20251202


def interior_count(n, level):
    m = n * 2 ** level - 1
    return m ** 3


def cell_interior_count(n, level):
    w = 2 ** level + 1
    return 6 * n ** 3 * (w - 4) * (w - 3) * (w - 2) // 6


def time_kernel(H, A, gmg_h, x, b, y, level, which, reps):
    ms = ctypes.c_double()
    H._lib.call("hymi_time_kernel", A.h, gmg_h, x.h, b.h if b is not None else None,
                y.h if y is not None else None, level, which, reps, ctypes.byref(ms))
    return ms.value


KIND_NAMES = {0: "kCellMarch (row windows)", 1: "kCellFlat (flat slab windows)", 2: "kCellTile (LDS slab tiles)",
              3: "kCellColumn (columns marched along z)"}


def traffic_from_profile(name, kernel_name):
    """HBM bytes per launch measured by rocprofv3 --pmc (FETCH_SIZE, WRITE_SIZE, scaled by a plain copy), if a
    summary profiles/traffic_<name>.json exists and names the kernel that runs."""
    path = os.path.join(ROOT, "profiles", f"traffic_{name}.json")
    if os.path.exists(path):
        with open(path) as fh:
            d = json.load(fh)
        if d.get("kernel", "").split("<")[0] == kernel_name.split(" ")[0].split("<")[0]:
            return d.get("bytes_per_launch")
    return None


def sine_modes():
    """f = 3 pi^2 sin(pi x) sin(pi y) sin(pi z) as one Fourier mode (amplitude, k, l, m)."""
    return np.array([[3 * np.pi ** 2, 1, 1, 1]])


def storage(H, n, tune_to=None):
    """The n x n x n x 6 cube; with tune_to = L, launch schedules and the fused cell phase of levels 2..L picked
    by the library's autotuner (hymi_mesh_autotune) for FP64 and FP32."""
    S = H.PrimitiveStorage(H.MeshInfo.meshCuboid((0, 0, 0), (1, 1, 1), n, n, n))
    if tune_to is not None:
        S.tuning = {lv: S.autotune(lv, reps=5) for lv in range(2, tune_to + 1)}
    return S


def tuned_choices(S, L):
    return {lv: dict(S.get_schedule(lv), cell_sor=list(S.get_cell_sor_setting(lv))) for lv in range(2, L + 1)}


def gmg_solver(H, S, L, nu):
    return H.GeometricMultigridSolver(S, H.SORSmoother(1.0), H.CGSolver(S, 0, 0), H.P1toP1LinearRestriction(),
                                      H.P1toP1LinearProlongation(), 0, L, nu, nu)


def noisy_sine_rhs(H, S, L, seed):
    """b = M I_h f + U[-1, 1] * 1e-3 per DoF (the headline's right-hand side)."""
    M = H.P1ConstantMassOperator(S, 0, L)
    b, t = H.P1Function("b", S, 0, L), H.P1Function("t", S, 0, L)
    t.interpolate_modes(sine_modes(), L)
    M.apply(t, b, L, H.ALL)
    bv = b.download(L)
    bv += 1e-3 * np.random.default_rng(seed).uniform(-1.0, 1.0, bv.shape)
    b.upload(L, bv)
    return b, t


def headline(H, n, L, tol, seed, cpo, steps, warmup):
    """Config 3: `steps` timed solves (after `warmup` untimed ones) from x = 0 to relative residual tol, FP32
    V(2,2) cycles (cpo per FP64 refinement step) against FP64 V(2,2) cycles on the same system.  The FP64
    solve replays its cycle count (counted once, untimed) as one graph launch sequence."""
    S = storage(H, n, tune_to=L)
    A = H.P1ConstantLaplaceOperator(S, 0, L)
    b, t = noisy_sine_rhs(H, S, L, seed)
    x = H.P1Function("x", S, 0, L)
    N = interior_count(n, L)

    def rel_residual():
        A.residual(x, b, t, L, H.INNER)
        return float(np.sqrt(t.dotGlobal(t, L, H.INNER)) / r0)

    A.residual(x, b, t, L, H.INNER)
    r0 = float(np.sqrt(t.dotGlobal(t, L, H.INNER)))
    g64 = gmg_solver(H, S, L, 2)
    cycles64 = 0
    while cycles64 < 80:
        g64.solve_graph(A, x, b, L, 1)
        cycles64 += 1
        if rel_residual() <= tol:
            break
    x64 = x.download(L)
    rel64 = rel_residual()

    def fp64_solve():
        x.interpolate(0.0, L, H.ALL)
        g64.solve_graph(A, x, b, L, cycles64)

    R = H.MixedPrecisionRefinementSolver(S, gmg_solver(H, S, L, 2), tolerance=tol, maxOuter=80, cyclesPerOuter=cpo)
    rel = None

    def mixed_solve():
        nonlocal rel
        x.interpolate(0.0, L, H.ALL)
        rel = R.solve(A, x, b, L)

    def timed(fn, k):
        S.ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(k):
            fn()
        S.ctx.synchronize()
        return (time.perf_counter() - t0) / k

    for _ in range(max(1, warmup)):
        mixed_solve()
        fp64_solve()
    t64 = timed(fp64_solve, steps)
    tmix = timed(mixed_solve, steps)
    xm = x.download(L)
    rel_mixed_check = rel_residual()
    cycles32 = R.outerSteps * cpo
    err = float(np.abs(xm - x64).max() / np.abs(x64).max())
    out = {
        "N": N, "S": S, "A": A, "x": x, "b": b, "t": t,
        "fp64": {"cycle": "V(2,2)", "cycles": cycles64, "seconds_per_solve": t64,
                 "GDoF_per_s_per_vcycle": N * cycles64 / t64 / 1e9, "final_rel_residual": rel64},
        "mixed": {"cycle": f"FP32 V(2,2), {cpo} per FP64 refinement step", "outer_steps": R.outerSteps,
                  "fp32_cycles": cycles32, "seconds_per_solve": tmix,
                  "GDoF_per_s_per_vcycle": N * cycles32 / tmix / 1e9,
                  "final_rel_residual": float(np.max(rel)), "final_rel_residual_fp64_check": rel_mixed_check,
                  "history": R.history},
        "speedup_time_to_tol": t64 / tmix,
        "error_vs_fp64_solution": err,
        "tuned": tuned_choices(S, L),
    }
    return out


def cell_phase_roofline(H, S, A, L, prec, Ncell, reps):
    """The SOR cell phase on level L in precision prec as the level runs it, measured as one smoothing sequence
    of a V(2,2) cycle (two sweeps; selector 8): with the fused phase selected, two ping-pong passes (u -> scratch,
    scratch -> u), else the colour passes of two sweeps.  `achieved` = algorithmic bytes of the sequence (2 sweeps
    x 3 s per cell DoF: read x, b, write x) / its measured time.  For the fused phase the in-place form (tile
    kernel alone and with its band copy, which a single sweep uses) is reported beside it."""
    s = 8 if prec == H.F64 else 4
    tn = "double" if s == 8 else "float"
    x, b = H.P1Function("rx", S, L, L, precision=prec), H.P1Function("rb", S, L, L, precision=prec)
    x.interpolate_modes(sine_modes(), L)
    b.interpolate(1e-3, L, H.ALL)
    tile, bits = S.get_cell_sor_setting(L)
    fused = bool(tile) and bool(bits & (1 if s == 8 else 2))
    t_seq = time_kernel(H, A, None, x, b, None, L, 8, reps)
    seq_bytes = 2 * 3 * s * Ncell
    if fused:
        kname = f"kCellSorTile<{tn}>"
        what = f"{kname}: SOR cell phase, all four colours in one launch, two ping-pong passes"
    else:
        kname = KIND_NAMES[S.get_schedule(L)["kind"]].split(" ")[0] + f"<{tn}>"
        what = f"{kname}: SOR cell phase as four colour passes, two sweeps"
    tag = f"cell_sor_{'tile' if fused else 'passes'}_{'f64' if s == 8 else 'f32'}_L{L}"
    r = {"bound": "hbm", "achieved": seq_bytes / (t_seq * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
         "traffic": traffic_from_profile(tag, kname)}
    r["frac"] = r["achieved"] / HBM_PEAK_GBS
    r.update({"kernel": f"{what}, level {L} (one smoothing sequence of V(2,2))",
              "algorithmic_bytes_per_sequence": seq_bytes, "ms_per_sequence": t_seq})
    if fused:
        t_tile = time_kernel(H, A, None, x, b, None, L, 7, reps)
        t_phase = time_kernel(H, A, None, x, b, None, L, 6, reps)
        r.update({"in_place_sweep_ms": t_phase, "in_place_tile_ms": t_tile, "in_place_band_copy_ms": t_phase - t_tile})
    del x, b
    return r


def residual_restriction(H, S, A, L, Ncell, reps):
    """The residual on level L restricted to level L-1 as a V-cycle runs them (selector 9), in one pass
    (kResidualRestrict) and as two passes (the residual, then the restriction), FP32 and FP64.  `fused_GBs` =
    algorithmic bytes of the one pass (read x and b at the fine cell DoFs, write the coarse cell DoFs: 2 s + s / 8 per
    fine cell DoF) / its measured time."""
    out = {}
    default = S.get_residual_restrict(L)
    for prec, name, s in ((H.F32, "fp32", 4), (H.F64, "fp64", 8)):
        x, b = H.P1Function("qx", S, L - 1, L, precision=prec), H.P1Function("qb", S, L - 1, L, precision=prec)
        r = H.P1Function("qr", S, L, L, precision=prec)
        x.interpolate_modes(sine_modes(), L)
        b.interpolate(1e-3, L, H.ALL)
        t = {}
        for mode in (1, 0):
            S.set_residual_restrict(L, mode)
            t[mode] = time_kernel(H, A, None, x, b, r, L, 9, reps)
        out[name] = {"fused_ms": t[1], "two_pass_ms": t[0],
                     "fused_GBs": (2 * s + s / 8) * Ncell / (t[1] * 1e-3) / 1e9}
        del x, b, r
    S.set_residual_restrict(L, -1)
    out["default_fused"] = bool(default)
    return out


def apply_roofline(H, S, A, L, Ncell, reps):
    """The macro-cell 15-point apply kernel (FP64) and a plain copy of the same level, both timed live."""
    x, y = H.P1Function("ax", S, L, L), H.P1Function("ay", S, L, L)
    x.interpolate_modes(sine_modes(), L)
    t_apply = time_kernel(H, A, None, x, None, y, L, 0, reps)
    t_copy = time_kernel(H, A, None, x, None, y, L, 5, reps)
    sched = S.get_schedule(L)
    kname = KIND_NAMES[sched["kind"]]
    apply_bytes = 2 * 8 * Ncell
    r = {"bound": "hbm", "achieved": apply_bytes / (t_apply * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
         "traffic": traffic_from_profile(f"cell_apply_L{L}", kname)}
    r["frac"] = r["achieved"] / HBM_PEAK_GBS
    r.update({"kernel": f"{kname}, 15-point apply, FP64, level {L}", "schedule": sched,
              "algorithmic_bytes_per_launch": apply_bytes, "ms_per_launch": t_apply,
              "plain_copy_ms": t_copy})
    del x, y
    return r


def config2_vcycle(H, n, L, steps, warmup):
    """Config 2: FP64 V(3,3) cycles at level L on the sine problem, replayed from a captured graph."""
    S = storage(H, n, tune_to=L)
    A = H.P1ConstantLaplaceOperator(S, 0, L)
    M = H.P1ConstantMassOperator(S, 0, L)
    x, b, t = (H.P1Function(k, S, 0, L) for k in ("u", "f", "tmp"))
    t.interpolate_modes(sine_modes(), L)
    M.apply(t, b, L, H.ALL)
    gmg = gmg_solver(H, S, L, 3)
    A.residual(x, b, t, L, H.INNER)
    r0 = np.sqrt(t.dotGlobal(t, L, H.INNER))
    gmg.solve_graph(A, x, b, L, max(1, warmup))
    S.ctx.synchronize()
    t0 = time.perf_counter()
    gmg.solve_graph(A, x, b, L, steps)
    S.ctx.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    A.residual(x, b, t, L, H.INNER)
    N = interior_count(n, L)
    Ncell = cell_interior_count(n, L)
    t_vc = time_kernel(H, A, gmg._handle(A, x), x, b, None, L, 4, 3)
    t_seq = time_kernel(H, A, None, x, b, None, L, 8, 5) if S.get_cell_sor(L) else None
    return {"workload": f"unit cube {n}x{n}x{n} x 6 tets, level {L}, {N} unknowns, FP64 V(3,3) multicolour SOR "
                        f"omega=1, CG at level 0, sine RHS, zero Dirichlet",
            "ms_per_vcycle": ms, "GDoF_per_s_per_vcycle": N / (ms * 1e-3) / 1e9, "vcycle_ms_events": t_vc,
            "rel_residual_after_cycles": float(np.sqrt(t.dotGlobal(t, L, H.INNER)) / r0),
            "cycles": 1 + max(1, warmup) + steps, "tuned": tuned_choices(S, L),
            "cell_phase_two_sweeps_ms": t_seq,
            "cell_phase_GBs": (2 * 3 * 8 * Ncell / (t_seq * 1e-3) / 1e9) if t_seq else None}


def batched_rhs(H, n, L, nrhs, tol, seed, cycles_per_outer):
    """Config-4 batch: nrhs right-hand sides, each f_r a sum of 8 seeded modes (amplitude N(0,1), k, l, m in
    1..8), b_r = M I_h f_r, solved together in the interleaved [dof][rhs] layout with FP32 V(2,2) cycles in
    FP64 refinement to tol; against the same solves one right-hand side at a time."""
    S = storage(H, n)
    A = H.P1ConstantLaplaceOperator(S, 0, L)
    M = H.P1ConstantMassOperator(S, 0, L)
    rng = np.random.default_rng(seed)
    modes = np.zeros((nrhs, 8, 4))
    modes[:, :, 0] = rng.standard_normal((nrhs, 8))
    modes[:, :, 1:] = rng.integers(1, 9, (nrhs, 8, 3))
    xb, bb = H.P1Function("xb", S, 0, L, nrhs=nrhs), H.P1Function("bb", S, 0, L, nrhs=nrhs)
    xb.interpolate_modes(modes, L)
    M.apply(xb, bb, L, H.ALL)
    N = interior_count(n, L)

    def refine():
        return H.MixedPrecisionRefinementSolver(S, gmg_solver(H, S, L, 2), tolerance=tol, maxOuter=60,
                                                cyclesPerOuter=cycles_per_outer)

    Rb = refine()
    for timed in (False, True):                      # the first solve captures graphs and allocates
        xb.interpolate(0.0, L, H.ALL)
        S.ctx.synchronize()
        t0 = time.perf_counter()
        rel = Rb.solve(A, xb, bb, L)
        S.ctx.synchronize()
        tb = time.perf_counter() - t0
    cycles_b = Rb.outerSteps * cycles_per_outer
    x1, b1 = H.P1Function("x1", S, 0, L), H.P1Function("b1", S, 0, L)
    R1 = refine()
    t1, cycles_1 = 0.0, 0
    for r in range(nrhs):
        x1.interpolate_modes(modes[r], L)
        M.apply(x1, b1, L, H.ALL)
        x1.interpolate(0.0, L, H.ALL)
        if r == 0:
            R1.solve(A, x1, b1, L)                   # capture
            x1.interpolate(0.0, L, H.ALL)
        S.ctx.synchronize()
        t0 = time.perf_counter()
        R1.solve(A, x1, b1, L)
        S.ctx.synchronize()
        t1 += time.perf_counter() - t0
        cycles_1 += R1.outerSteps * cycles_per_outer
    return {
        "workload": f"unit cube {n}x{n}x{n} x 6 tets, level {L}, {N} unknowns x {nrhs} right-hand sides (8 modes "
                    f"each, seed {seed}), FP32 V(2,2) x {cycles_per_outer} per FP64 refinement step, to {tol:g}",
        "batched": {"seconds": tb, "fp32_cycles": cycles_b, "outer_steps": Rb.outerSteps,
                    "max_rel_residual": float(np.max(rel)),
                    "GDoF_per_s_per_vcycle": N * nrhs * cycles_b / tb / 1e9, "solves_per_s": nrhs / tb},
        "one_at_a_time": {"seconds": t1, "fp32_cycles": cycles_1,
                          "GDoF_per_s_per_vcycle": N * cycles_1 / t1 / 1e9, "solves_per_s": nrhs / t1},
        "speedup": t1 / tb,
    }


def serving_run(count, rate, slots, seed, low_rate_count=120):
    """Config-5 stream: count Poisson requests at rate/s (n in 1..3, level in 4..7, 4 modes, tol 1e-8), with
    every (n, level) hierarchy resident beforehand; small problems are batched into 32-column lanes, large ones
    run on `slots` solve lanes.  The stream is then served again at 70 % of the throughput the first run
    reached (below capacity), so that its latencies describe the service rather than a backlog."""
    from hyteg_amd.serving import PoissonServer, poisson_stream, summarize
    srv = PoissonServer(slots=slots)
    t0 = time.perf_counter()
    srv.warm([(n, L) for n in (1, 2, 3) for L in (4, 5, 6, 7)])
    warm_s = time.perf_counter() - t0
    res = srv.serve(poisson_stream(count, rate=rate, seed=seed), realtime=True)
    out = summarize(res, srv.wall)
    out.update({"rate_per_s": rate, "slots": slots, "seed": seed, "wall_s": srv.wall, "warm_s": warm_s,
                "errors": len(srv.errors), "error_samples": [e for _, e in srv.errors[:3]]})
    low = 0.7 * out["solves_per_s"]
    res = srv.serve(poisson_stream(low_rate_count, rate=low, seed=seed + 1), realtime=True)
    below = summarize(res, srv.wall)
    below.update({"rate_per_s": low, "wall_s": srv.wall, "errors": len(srv.errors)})
    out["below_capacity"] = below
    srv.close()
    return out


def cpu_baseline(n, L, cycles, apply_runs, seed):
    """The CPU oracle (the reference's algorithm, lexicographic SOR, OpenMP over macro-primitives) on a bounded
    sample of the headline workload: the stencil apply (best of apply_runs runs, each at least 1 s of
    back-to-back applies after a warm-up, as ApplyBenchmark.cpp:78-94 times it) and `cycles` FP64 V(2,2)
    cycles from x = 0 on the same right-hand side, the mean of cycles 2.. .  Per cycle it records, as
    MultigridStudies.cpp:228-255 computes them, the discrete L2 residual and the discrete L2 error against
    u = sin(pi x) sin(pi y) sin(pi z) (sums over inner DoFs divided by the number of all DoFs), and the
    residual ratio."""
    import oracle as O
    m = O.Mesh.cuboid(n)
    A = O.Operator(m, 0, L)
    M = O.Operator(m, 0, L, 1)
    x, b, t = O.Function(m, 0, L), O.Function(m, 0, L), O.Function(m, 0, L)
    X = m.owned_coords(L)
    u_exact = np.prod(np.sin(np.pi * X), axis=1)
    t.set(L, 3 * np.pi ** 2 * u_exact)
    M.apply(t, b, L, flag=O.ALL)
    bv = b.get(L)
    bv += 1e-3 * np.random.default_rng(seed).uniform(-1.0, 1.0, bv.shape)
    b.set(L, bv)
    del X, bv
    # apply: y = A t on Inner DoFs
    A.apply(t, x, L, flag=O.INNER)
    best = None
    for _ in range(apply_runs):
        k, t0 = 0, O.wtime()
        while True:
            A.apply(t, x, L, flag=O.INNER)
            k += 1
            dt = O.wtime() - t0
            if dt >= 1.0:
                break
        best = dt / k if best is None else min(best, dt / k)
    x.set(L, np.zeros(len(t.get(L))))

    def resnorm():
        A.apply(x, t, L, flag=O.INNER)
        r = b.get(L) - t.get(L)
        return float(np.sqrt(np.sum(r[owned_inner] ** 2)))

    def l2_error():
        e = u_exact - x.get(L)
        return float(np.sqrt(np.sum(e[owned_inner] ** 2) / n_all))

    owned_inner = m.owned_types(L) == O.INNER
    n_all = len(owned_inner)
    bkeep = b.get(L)
    g = O.GMG(m, 0, L, 2, 2, 1.0, 0)
    res, err = [resnorm()], [l2_error()]
    times = []
    for _ in range(cycles):
        t0 = O.wtime()
        g.solve(A, x, b, L)
        times.append(O.wtime() - t0)
        b.set(L, bkeep)
        res.append(resnorm())
        err.append(l2_error())
    N = interior_count(n, L)
    dt = float(np.mean(times[1:])) if cycles > 1 else times[0]
    Ncell = cell_interior_count(n, L)
    return {"value": N / dt / 1e9, "unit": "GDoF/s", "cores": O.num_threads(), "kind": "port",
            "sample": f"the C oracle (lexicographic SOR) on the headline system: {cycles} FP64 V(2,2) cycles from "
                      f"x = 0, value = unknowns / mean time of cycles 2..{cycles} ({dt:.2f} s per cycle); apply "
                      f"best of {apply_runs} runs of >= 1 s",
            "vcycle_seconds": times, "residual_ratio_per_cycle": [res[k + 1] / res[k] for k in range(cycles)],
            "l2_residual_per_cycle": [r / np.sqrt(n_all) for r in res[1:]], "l2_error_per_cycle": err[1:],
            "apply_ms": best * 1e3, "apply_GDoF_per_s": N / best / 1e9,
            "apply_cell_GBs": 2 * 8 * Ncell / best / 1e9}


def kernels_only(H, n, L, reps):
    """Every measured kernel of the headline workload alone, launched eagerly (no graphs): for rocprofv3 kernel
    traces and counter passes.  One V(2,2) cycle in FP32 and in FP64 (x = 0, b = the headline's right-hand
    side), then `reps` launches each of the SOR cell phase of two sweeps (FP32 and FP64), the residual and
    restriction in one pass and in two (FP32 and FP64), the FP64 macro-cell apply and a plain copy of the level (the
    reference for scaling FETCH_SIZE / WRITE_SIZE)."""
    S = storage(H, n)
    A = H.P1ConstantLaplaceOperator(S, 0, L)
    b, t = noisy_sine_rhs(H, S, L, SEED)
    out = {}
    for prec, name in ((H.F32, "fp32"), (H.F64, "fp64")):
        x, bb = H.P1Function("x", S, 0, L, precision=prec), H.P1Function("bb", S, 0, L, precision=prec)
        H._lib.call("hymi_convert", bb.h, b.h, L, H.ALL)
        g = gmg_solver(H, S, L, 2)
        out[f"vcycle_{name}_ms"] = time_kernel(H, A, g._handle(A, x), x, bb, None, L, 4, 1)
        del x, bb, g
    Ncell = cell_interior_count(n, L)
    out["cell_phase_fp32"] = cell_phase_roofline(H, S, A, L, H.F32, Ncell, reps)["ms_per_sequence"]
    out["cell_phase_fp64"] = cell_phase_roofline(H, S, A, L, H.F64, Ncell, reps)["ms_per_sequence"]
    out["residual_restriction"] = residual_restriction(H, S, A, L, Ncell, reps)
    ra = apply_roofline(H, S, A, L, Ncell, reps)
    out.update({"apply_ms": ra["ms_per_launch"], "copy_ms": ra["plain_copy_ms"],
                "copy_bytes_each_way": 8 * S.owned_count(L)})
    return out


def stage(what):
    """One progress line on stderr per finished part of the run (stdout carries only the JSON line)."""
    print(f"[bench] {what} done", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5, help="timed solves of the headline workload")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--level", type=int, default=8)
    ap.add_argument("--n", type=int, default=2, help="sub-cubes per side")
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--cycles-per-step", type=int, default=4, help="FP32 V-cycles per FP64 refinement step")
    ap.add_argument("--config2-level", type=int, default=7)
    ap.add_argument("--config2-steps", type=int, default=20)
    ap.add_argument("--cpu-cycles", type=int, default=10)
    ap.add_argument("--cpu-apply-runs", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-config2", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--serving-requests", type=int, default=200)
    ap.add_argument("--serving-slots", type=int, default=4)
    ap.add_argument("--kernels-only", action="store_true",
                    help="launch the measured kernels eagerly (for rocprofv3) and print their timings")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only: this build has no multi-GPU path (requested --gpus {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)

    import hyteg_amd as H

    n, L = args.n, args.level
    if args.kernels_only:
        print(json.dumps(kernels_only(H, n, L, max(3, args.steps))))
        return
    hl = headline(H, n, L, args.tol, SEED, args.cycles_per_step, args.steps, args.warmup)
    stage("headline solves")
    N = hl["N"]
    Ncell = cell_interior_count(n, L)
    S, A = hl["S"], hl["A"]
    reps = 10
    roof32 = cell_phase_roofline(H, S, A, L, H.F32, Ncell, reps)
    roof64 = cell_phase_roofline(H, S, A, L, H.F64, Ncell, reps)
    roof_apply = apply_roofline(H, S, A, L, Ncell, reps)
    res_restrict = residual_restriction(H, S, A, L, Ncell, reps)
    stage("kernel rooflines")
    for k in ("x", "b", "t", "S", "A"):
        hl.pop(k)
    del S, A

    mixed, fp64 = hl["mixed"], hl["fp64"]
    out = {
        "metric": METRIC,
        "value": mixed["GDoF_per_s_per_vcycle"],
        "unit": "GDoF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mixed["seconds_per_solve"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 V-cycles in f64 refinement",
        "data": "synthetic",
        "config": {
            "workload": f"unit cube {n}x{n}x{n} sub-cubes x 6 tets ({6 * n ** 3} macro-tets), level {L}, {N} "
                        f"unknowns, b = M I_h f + U[-1,1]*1e-3 (seed {SEED}), zero Dirichlet, from x = 0 to relative "
                        f"residual {args.tol:g}; one step = one solve: FP32 V(2,2) multicolour SOR omega=1 cycles, "
                        f"{args.cycles_per_step} per FP64 refinement step, CG at level 0",
            "unknowns": N, "level": L, "macro_tets": 6 * n ** 3,
            "mixed_GDoF_per_s_per_vcycle": mixed["GDoF_per_s_per_vcycle"],
            "fp64_GDoF_per_s_per_vcycle": fp64["GDoF_per_s_per_vcycle"],
            "mixed_over_fp64_time_to_tol": hl["speedup_time_to_tol"],
            "mixed_ms_per_solve": mixed["seconds_per_solve"] * 1e3,
            "fp64_ms_per_solve": fp64["seconds_per_solve"] * 1e3,
            "fp32_cycles": mixed["fp32_cycles"], "fp64_cycles": fp64["cycles"],
            "error_vs_fp64_solution": hl["error_vs_fp64_solution"],
            "final_rel_residual": mixed["final_rel_residual"],
        },
        "roofline": roof32,
        "roofline_fp64": roof64,
        "roofline_apply": roof_apply,
        "residual_restriction": res_restrict,
        "mixed_precision": {k: v for k, v in hl.items() if k != "N"},
    }
    if not args.no_config2:
        out["config2_vcycle"] = config2_vcycle(H, n, args.config2_level, args.config2_steps, 3)
        stage("level-7 FP64 V(3,3) cycles")
    if not args.no_batched:
        out["batched_rhs"] = batched_rhs(H, n, 6, 32, 1e-8, SEED, args.cycles_per_step)
        stage("32 right-hand sides")
    if not args.no_serving:
        out["serving"] = serving_run(args.serving_requests, 200.0, args.serving_slots, SEED)
        stage("serving")
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(n, L, args.cpu_cycles, args.cpu_apply_runs, SEED)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
