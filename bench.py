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


def traffic_from_profile(kernel):
    """HBM bytes per launch measured by rocprofv3 --pmc (FETCH_SIZE, WRITE_SIZE), if a summary exists."""
    path = os.path.join(ROOT, "profiles", f"traffic_{kernel}.json")
    if os.path.exists(path):
        with open(path) as fh:
            return json.load(fh).get("bytes_per_launch")
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--level", type=int, default=7)
    ap.add_argument("--n", type=int, default=2, help="sub-cubes per side")
    ap.add_argument("--cpu-cycles", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only: this build has no multi-GPU path (requested --gpus {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)

    import hyteg_amd as H

    n, L = args.n, args.level
    S = H.PrimitiveStorage(H.MeshInfo.meshCuboid((0, 0, 0), (1, 1, 1), n, n, n))
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
    s = 8
    apply_bytes = 2 * s * Ncell                 # read x once, write y once
    sor_bytes = 3 * s * Ncell / 4               # one colour pass = a quarter of a sweep's read x, b / write x
    sor_share = 6 * 4 * t_sor / t_vc            # V(3,3): 6 sweeps x 4 colour passes on the finest level
    roof_sor = {"bound": "hbm", "achieved": sor_bytes / (t_sor * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "traffic": traffic_from_profile("cell_sor")}
    roof_sor["frac"] = roof_sor["achieved"] / HBM_PEAK_GBS
    roof_apply = {"bound": "hbm", "achieved": apply_bytes / (t_apply * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                  "unit": "GB/s", "traffic": traffic_from_profile("cell_apply")}
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
        "roofline": dict(roof_sor, kernel="kCellSor (one colour pass)",
                         algorithmic_bytes_per_launch=sor_bytes, ms_per_launch=t_sor,
                         share_of_vcycle=sor_share),
        "roofline_apply": dict(roof_apply, kernel="kCellApply (15-point stencil)",
                               algorithmic_bytes_per_launch=apply_bytes, ms_per_launch=t_apply),
        "vcycle_ms_events": t_vc,
    }
    if not args.no_cpu:
        cpu_level = min(L, 7)
        v, dt, cores = cpu_baseline(n, cpu_level, args.cpu_cycles)
        out["cpu_baseline"] = {"value": v, "unit": "GDoF/s", "cores": cores, "kind": "port",
                               "sample": f"{args.cpu_cycles} V(3,3) cycle(s) of the C oracle (lexicographic SOR) on "
                                         f"the same mesh at level {cpu_level}, {dt:.2f} s per cycle"}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
