"""Benchmark: GDoF/s per V-cycle of the matrix-free P1 geometric multigrid on one MI355X.

Workload (default): unit cube meshed as 2x2x2 sub-cubes x 6 tetrahedra (48 macro-tets), refinement level 7
(16,581,375 unknowns), 3D Poisson with f = 3 pi^2 sin(pi x) sin(pi y) sin(pi z) and zero Dirichlet
boundary, FP64 V(3,3) cycles with multicolour SOR (omega = 1) and a device CG at level 0.  A "step" is one
V-cycle, replayed from a captured hipGraph, with all data resident in HBM.

Prints one JSON line.  `roofline` covers the dominant kernel of the V-cycle, timed live with hipEvents on
the library's stream; `roofline_apply` the 15-point stencil apply kernel; `cpu_baseline` times the CPU
oracle (the reference's algorithm restated in C, OpenMP over macro-primitives) on a bounded sample.
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


def traffic_from_profile(kernel, kernel_name):
    """HBM bytes per launch measured by rocprofv3 --pmc (FETCH_SIZE, WRITE_SIZE), if a summary for the kernel
    that runs exists (profiles/traffic_<kernel>.json names the kernel it was measured on)."""
    path = os.path.join(ROOT, "profiles", f"traffic_{kernel}.json")
    if os.path.exists(path):
        with open(path) as fh:
            d = json.load(fh)
        if d.get("kernel", "").split("<")[0] == kernel_name.split(" ")[0]:
            return d.get("bytes_per_launch")
    return None


def cpu_baseline(n, level, cycles):
    """The CPU oracle (reference algorithm, lexicographic SOR) on the same workload: `cycles` V(3,3)."""
    import oracle as O
    m = O.Mesh.cuboid(n)
    A = O.Operator(m, 0, level)
    M = O.Operator(m, 0, level, 1)
    x, b, t = O.Function(m, 0, level), O.Function(m, 0, level), O.Function(m, 0, level)
    X = m.owned_coords(level)
    t.set(level, 3 * np.pi ** 2 * np.prod(np.sin(np.pi * X), axis=1))
    M.apply(t, b, level, flag=O.ALL)
    g = O.GMG(m, 0, level, 3, 3, 1.0, 0)
    t0 = O.wtime()
    for _ in range(cycles):
        g.solve(A, x, b, level)
    dt = (O.wtime() - t0) / cycles
    return interior_count(n, level) / dt / 1e9, dt, O.num_threads()


def sine_modes():
    """f = 3 pi^2 sin(pi x) sin(pi y) sin(pi z) as one Fourier mode (amplitude, k, l, m)."""
    return np.array([[3 * np.pi ** 2, 1, 1, 1]])


def kernels_only(H, n, L, reps):
    """Each measured kernel alone, `reps` launches: the macro-cell apply, one SOR colour pass, and a plain
    copy of a whole level (the reference for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE counters)."""
    S = H.PrimitiveStorage(H.MeshInfo.meshCuboid((0, 0, 0), (1, 1, 1), n, n, n))
    S.use_column_schedules(L)
    A = H.P1ConstantLaplaceOperator(S, L, L)
    x, b, y = (H.P1Function(k, S, L, L) for k in "xby")
    x.interpolate_modes(sine_modes(), L)
    b.interpolate(1.0, L, H.ALL)
    t_apply = time_kernel(H, A, None, x, None, y, L, 0, reps)
    t_sor = time_kernel(H, A, None, x, b, None, L, 1, reps)
    t_copy = time_kernel(H, A, None, x, None, y, L, 5, reps)
    out = {"apply_ms": t_apply, "sor_pass_ms": t_sor, "copy_ms": t_copy}
    if S.get_cell_sor(L):
        out["sor_tile_ms"] = time_kernel(H, A, None, x, b, None, L, 7, reps)
        out["sor_cell_phase_ms"] = time_kernel(H, A, None, x, b, None, L, 6, reps)
    copy_bytes = S.ctx.bytes_in_use() // 3       # one level of one function (x, b, y are alike)
    out.update({"copy_bytes_each_way": copy_bytes, "copy_GBs": 2 * copy_bytes / (t_copy * 1e-3) / 1e9})
    return out


def mixed_vs_fp64(H, n, L, tol, seed, cycles_per_outer):
    """FP32 V(2,2) cycles inside FP64 iterative refinement against plain FP64 V(2,2) cycles, both from x = 0
    to relative residual tol, on b = M I_h f + U[-1, 1] * 1e-3 per DoF."""
    S = H.PrimitiveStorage(H.MeshInfo.meshCuboid((0, 0, 0), (1, 1, 1), n, n, n))
    S.use_column_schedules(L)
    A = H.P1ConstantLaplaceOperator(S, 0, L)
    M = H.P1ConstantMassOperator(S, 0, L)
    x, b, t = (H.P1Function(k, S, 0, L) for k in ("x", "b", "t"))
    t.interpolate_modes(sine_modes(), L)
    M.apply(t, b, L, H.ALL)
    bv = b.download(L)
    bv += 1e-3 * np.random.default_rng(seed).uniform(-1.0, 1.0, bv.shape)
    b.upload(L, bv)
    del bv
    N = interior_count(n, L)

    def gmg():
        return H.GeometricMultigridSolver(S, H.SORSmoother(1.0), H.CGSolver(S, 0, 0), H.P1toP1LinearRestriction(),
                                          H.P1toP1LinearProlongation(), 0, L, 2, 2)

    def rel_residual():
        A.residual(x, b, t, L, H.INNER)
        return float(np.sqrt(t.dotGlobal(t, L, H.INNER)) / r0)

    A.residual(x, b, t, L, H.INNER)
    r0 = float(np.sqrt(t.dotGlobal(t, L, H.INNER)))
    # FP64: count the cycles to tol (untimed), then time that many from x = 0 as one graph replay
    g64 = gmg()
    cycles64 = 0
    while cycles64 < 60:
        g64.solve_graph(A, x, b, L, 1)
        cycles64 += 1
        if rel_residual() <= tol:
            break
    x64 = x.download(L)
    x.interpolate(0.0, L, H.ALL)
    S.ctx.synchronize()
    t0 = time.perf_counter()
    g64.solve_graph(A, x, b, L, cycles64)
    S.ctx.synchronize()
    t64 = time.perf_counter() - t0
    # mixed: one untimed solve (captures graphs, allocates refinement work), then a timed one from x = 0
    R = H.MixedPrecisionRefinementSolver(S, gmg(), tolerance=tol, maxOuter=60, cyclesPerOuter=cycles_per_outer)
    x.interpolate(0.0, L, H.ALL)
    R.solve(A, x, b, L)
    x.interpolate(0.0, L, H.ALL)
    S.ctx.synchronize()
    t0 = time.perf_counter()
    rel = R.solve(A, x, b, L)
    S.ctx.synchronize()
    tmix = time.perf_counter() - t0
    cycles32 = R.outerSteps * cycles_per_outer
    xm = x.download(L)
    err = float(np.abs(xm - x64).max() / np.abs(x64).max())
    return {
        "workload": f"unit cube {n}x{n}x{n} x 6 tets, level {L}, {N} unknowns, b = M I_h f + U[-1,1]*1e-3 "
                    f"(seed {seed}), from x = 0 to relative residual {tol:g}",
        "fp64": {"cycle": "V(2,2)", "cycles": cycles64, "seconds": t64, "GDoF_per_s_per_vcycle": N * cycles64 / t64 / 1e9},
        "mixed": {"cycle": f"FP32 V(2,2), {cycles_per_outer} per FP64 refinement step", "outer_steps": R.outerSteps,
                  "fp32_cycles": cycles32, "seconds": tmix, "GDoF_per_s_per_vcycle": N * cycles32 / tmix / 1e9,
                  "final_rel_residual": float(np.max(rel)), "history": R.history},
        "speedup_time_to_tol": t64 / tmix,
        "error_vs_fp64_solution": err,
    }


def batched_rhs(H, n, L, nrhs, tol, seed, cycles_per_outer):
    """Config-4 batch: nrhs right-hand sides, each f_r a sum of 8 seeded modes (amplitude N(0,1), k, l, m in
    1..8), b_r = M I_h f_r, solved together in the interleaved [dof][rhs] layout with FP32 V(2,2) cycles in
    FP64 refinement to tol; against the same solves one right-hand side at a time."""
    S = H.PrimitiveStorage(H.MeshInfo.meshCuboid((0, 0, 0), (1, 1, 1), n, n, n))
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
        g = H.GeometricMultigridSolver(S, H.SORSmoother(1.0), H.CGSolver(S, 0, 0), H.P1toP1LinearRestriction(),
                                       H.P1toP1LinearProlongation(), 0, L, 2, 2)
        return H.MixedPrecisionRefinementSolver(S, g, tolerance=tol, maxOuter=60, cyclesPerOuter=cycles_per_outer)

    Rb = refine()
    for timed in (False, True):                      # the first solve captures graphs and allocates
        xb.interpolate(0.0, L, H.ALL)
        S.ctx.synchronize()
        t0 = time.perf_counter()
        rel = Rb.solve(A, xb, bb, L)
        S.ctx.synchronize()
        tb = time.perf_counter() - t0
    cycles_b = Rb.outerSteps * cycles_per_outer
    # one right-hand side at a time
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


def serving_run(count, rate, slots, seed):
    """Config-5 stream: count Poisson requests at rate/s (n in 1..3, level in 4..7, 4 modes, tol 1e-8),
    served with every (n, level) hierarchy resident on every worker beforehand."""
    from hyteg_amd.serving import PoissonServer, poisson_stream, summarize
    reqs = poisson_stream(count, rate=rate, seed=seed)
    srv = PoissonServer(slots=slots)
    t0 = time.perf_counter()
    srv.warm([(n, L) for n in (1, 2, 3) for L in (4, 5, 6, 7) for _ in range(slots)])
    warm_s = time.perf_counter() - t0
    res = srv.serve(reqs, realtime=True)
    srv.close()
    out = summarize(res, srv.wall)
    out.update({"rate_per_s": rate, "slots": slots, "seed": seed, "wall_s": srv.wall, "warm_s": warm_s,
                "errors": len(srv.errors), "error_samples": [e for _, e in srv.errors[:3]]})
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--level", type=int, default=7)
    ap.add_argument("--n", type=int, default=2, help="sub-cubes per side")
    ap.add_argument("--cpu-cycles", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--mixed-level", type=int, default=8)
    ap.add_argument("--mixed-cycles-per-step", type=int, default=4)
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--serving-requests", type=int, default=200)
    ap.add_argument("--serving-slots", type=int, default=4)
    ap.add_argument("--kernels-only", action="store_true",
                    help="run only the measured kernels (for rocprofv3 counter passes) and print their timings")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only: this build has no multi-GPU path (requested --gpus {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)

    import hyteg_amd as H

    n, L = args.n, args.level
    if args.kernels_only:
        print(json.dumps(kernels_only(H, n, L, max(5, args.steps))))
        return
    S = H.PrimitiveStorage(H.MeshInfo.meshCuboid((0, 0, 0), (1, 1, 1), n, n, n))
    S.use_column_schedules(L)
    A = H.P1ConstantLaplaceOperator(S, 0, L)
    M = H.P1ConstantMassOperator(S, 0, L)
    x = H.P1Function("u", S, 0, L)
    b = H.P1Function("f", S, 0, L)
    t = H.P1Function("tmp", S, 0, L)
    t.interpolate(lambda X: 3 * np.pi ** 2 * np.prod(np.sin(np.pi * X), axis=1), L)
    M.apply(t, b, L, H.ALL)
    gmg = H.GeometricMultigridSolver(S, H.SORSmoother(1.0), H.CGSolver(S, 0, 0), H.P1toP1LinearRestriction(),
                                     H.P1toP1LinearProlongation(), 0, L, 3, 3)
    gh = gmg._handle(A, x)

    def resnorm():
        A.residual(x, b, t, L, H.INNER)
        return np.sqrt(t.dotGlobal(t, L, H.INNER))

    r0 = resnorm()
    gmg.solve_graph(A, x, b, L, args.warmup)
    S.ctx.synchronize()
    t0 = time.perf_counter()
    gmg.solve_graph(A, x, b, L, args.steps)
    S.ctx.synchronize()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / args.steps
    rel_res = resnorm() / r0
    N = interior_count(n, L)
    gdofs = N / (ms / 1e3) / 1e9

    # per-kernel timing (hipEvents on the library stream)
    y = H.P1Function("y", S, 0, L)
    Ncell = cell_interior_count(n, L)
    reps = max(5, min(50, args.steps))
    t_apply = time_kernel(H, A, None, x, None, y, L, 0, reps)
    t_sor = time_kernel(H, A, None, x, b, None, L, 1, reps)
    t_vc = time_kernel(H, A, gh, x, b, None, L, 4, 3)
    sched = S.get_schedule(L)
    kname = KIND_NAMES[sched["kind"]]
    s = 8
    apply_bytes = 2 * s * Ncell                 # read x once, write y once
    pass_bytes = 3 * s * Ncell / 4              # one colour pass = a quarter of a sweep's read x, b / write x
    roof_pass = {"bound": "hbm", "achieved": pass_bytes / (t_sor * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                 "unit": "GB/s", "traffic": traffic_from_profile("cell_sor", kname)}
    roof_pass["frac"] = roof_pass["achieved"] / HBM_PEAK_GBS
    roof_pass = dict(roof_pass, kernel=f"{kname}, SOR, one colour pass", schedule=sched,
                     algorithmic_bytes_per_launch=pass_bytes, ms_per_launch=t_sor)
    if S.get_cell_sor(L):
        # the fused SOR cell phase: all four colours in one tile launch, then the band copy
        t_tile = time_kernel(H, A, None, x, b, None, L, 7, reps)
        t_phase = time_kernel(H, A, None, x, b, None, L, 6, reps)
        sor_bytes = 3 * s * Ncell               # a whole sweep of the cell interiors: read x, b, write x
        roof_sor = {"bound": "hbm", "achieved": sor_bytes / (t_tile * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "traffic": traffic_from_profile("cell_sor_tile", "kCellSorTile")}
        roof_sor["frac"] = roof_sor["achieved"] / HBM_PEAK_GBS
        roof_sor = dict(roof_sor, kernel="kCellSorTile<double>, SOR cell phase, all four colours",
                        tile=S.get_cell_sor(L), algorithmic_bytes_per_launch=sor_bytes, ms_per_launch=t_tile,
                        share_of_vcycle=6 * t_tile / t_vc, band_copy_ms=t_phase - t_tile)
    else:
        roof_sor = dict(roof_pass, share_of_vcycle=6 * 4 * t_sor / t_vc)
    roof_sor["frac"] = roof_sor["achieved"] / HBM_PEAK_GBS
    roof_apply = {"bound": "hbm", "achieved": apply_bytes / (t_apply * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                  "unit": "GB/s", "traffic": traffic_from_profile("cell_apply", kname)}
    roof_apply["frac"] = roof_apply["achieved"] / HBM_PEAK_GBS

    out = {
        "metric": METRIC,
        "value": gdofs,
        "unit": "GDoF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"unit cube {n}x{n}x{n} sub-cubes x 6 tets ({6 * n ** 3} macro-tets), level {L}, "
                        f"{N} unknowns, FP64 V(3,3) multicolour SOR omega=1, CG at level 0, sine RHS, zero Dirichlet",
            "unknowns": N,
            "level": L,
            "macro_tets": 6 * n ** 3,
            "precision": "fp64",
            "cycle": "V(3,3)",
            "rel_residual_after_all_cycles": rel_res,
            "cycles_total": 1 + args.warmup + args.steps + reps * 0,
        },
        "roofline": roof_sor,
        "roofline_colour_pass": roof_pass,
        "roofline_apply": dict(roof_apply, kernel=f"{kname}, 15-point apply", schedule=sched,
                               algorithmic_bytes_per_launch=apply_bytes, ms_per_launch=t_apply),
        "vcycle_ms_events": t_vc,
    }
    if not args.no_mixed:
        out["mixed_precision"] = mixed_vs_fp64(H, n, args.mixed_level, 1e-10, 20251202, args.mixed_cycles_per_step)
    if not args.no_batched:
        out["batched_rhs"] = batched_rhs(H, n, 6, 32, 1e-8, 20251202, args.mixed_cycles_per_step)
    if not args.no_serving:
        out["serving"] = serving_run(args.serving_requests, 200.0, args.serving_slots, 20251202)
    if not args.no_cpu:
        cpu_level = min(L, 7)
        v, dt, cores = cpu_baseline(n, cpu_level, args.cpu_cycles)
        out["cpu_baseline"] = {"value": v, "unit": "GDoF/s", "cores": cores, "kind": "port",
                               "sample": f"{args.cpu_cycles} V(3,3) cycle(s) of the C oracle (lexicographic SOR) on "
                                         f"the same mesh at level {cpu_level}, {dt:.2f} s per cycle"}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
