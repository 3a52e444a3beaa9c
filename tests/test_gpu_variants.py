"""GPU tests of the solver variants next to the V-cycle: backward SOR sweeps, symmetric SOR smoothing,
W-cycles and full multigrid (SymmetricSORSmoother.hpp, GeometricMultigridSolver.hpp:201-204,
FullMultigridSolver.hpp), against the CPU oracle's multicolour schedule of the same algorithms.

Bars: a backward sweep matches the oracle's backward multicolour sweep to 1e-12 relative; cycles and FMG
match the oracle to 1e-11 relative (FP64, different summation orders only)."""
import numpy as np
import pytest

from helpers import mesh_pair, rel_err, say, sine_rhs

pytestmark = pytest.mark.gpu

MESHES = [("cube", 1), ("cube", 2), ("gmsh", "regular_octahedron_8el.msh")]


@pytest.mark.parametrize("kind", MESHES)
@pytest.mark.parametrize("level", [2, 3, 5])
def test_backward_sor_matches_oracle(kind, level, gpu_ctx, oracle_lib):
    import hyteg_amd as H
    O = oracle_lib
    say(f"[sor backward] mesh={kind} level={level}")
    om, S = mesh_pair(kind, gpu_ctx)
    rng = np.random.default_rng(31 + level)
    x, b = rng.standard_normal((2, om.owned_count(level)))
    A, OA = H.P1ConstantLaplaceOperator(S, 0, level), O.Operator(om, 0, level)
    u, f = H.P1Function("u", S, 0, level), H.P1Function("f", S, 0, level)
    u.upload(level, x)
    f.upload(level, b)
    ou, of = O.Function(om, 0, level), O.Function(om, 0, level)
    ou.set(level, x)
    of.set(level, b)
    for sweep in range(2):            # backward, then forward after backward (closure copies stay current)
        A.smooth_sor_backwards(u, f, 0.9, level, H.INNER)
        OA.smooth_sor_backward(ou, of, level, omega=0.9, flag=O.INNER, order=1)
        assert rel_err(u.download(level), ou.get(level)) <= 1e-12, ("backward", sweep)
        A.smooth_sor(u, f, 0.9, level, H.INNER)
        OA.smooth_sor(ou, of, level, omega=0.9, flag=O.INNER, order=1)
        assert rel_err(u.download(level), ou.get(level)) <= 1e-12, ("forward", sweep)


def test_backward_sor_column_schedule_level6(gpu_ctx, oracle_lib):
    """Backward colour passes through the macro-cell march kernels (levels above the one-launch cell phase),
    with the flat and the column schedule."""
    import hyteg_amd as H
    O = oracle_lib
    L = 6
    say(f"[sor backward] cube n=1 level={L}, flat and column schedules")
    om, S = mesh_pair(("cube", 1), gpu_ctx)
    rng = np.random.default_rng(7)
    x, b = rng.standard_normal((2, om.owned_count(L)))
    A, OA = H.P1ConstantLaplaceOperator(S, 0, L), O.Operator(om, 0, L)
    ou, of = O.Function(om, 0, L), O.Function(om, 0, L)
    ou.set(L, x)
    of.set(L, b)
    OA.smooth_sor_backward(ou, of, L, omega=1.0, flag=O.INNER, order=1)
    for sched in ((1, 4, 1, 1), (8, 4, 1, 3)):
        S.set_schedule(L, *sched)
        u, f = H.P1Function("u", S, 0, L), H.P1Function("f", S, 0, L)
        u.upload(L, x)
        f.upload(L, b)
        A.smooth_sor_backwards(u, f, 1.0, L, H.INNER)
        assert rel_err(u.download(L), ou.get(L)) <= 1e-12, sched
    S.set_schedule(L)


def _setup(H, O, om, S, L, nu, cycle, symmetric):
    X = om.owned_coords(L)
    A, M = H.P1ConstantLaplaceOperator(S, 0, L), H.P1ConstantMassOperator(S, 0, L)
    OA, OM = O.Operator(om, 0, L), O.Operator(om, 0, L, 1)
    t, b, x = (H.P1Function(n, S, 0, L) for n in "tbx")
    ot, ob, ox = (O.Function(om, 0, L) for _ in range(3))
    for lev in range(0, L + 1):                     # the right-hand side on every level (FMG needs it)
        Xl = om.owned_coords(lev)
        t.upload(lev, sine_rhs(Xl))
        M.apply(t, b, lev, H.ALL)
        ot.set(lev, sine_rhs(Xl))
        OM.apply(ot, ob, lev, flag=O.ALL)
    smoother = H.SymmetricSORSmoother(1.0) if symmetric else H.SORSmoother(1.0)
    gmg = H.GeometricMultigridSolver(S, smoother, H.CGSolver(S, 0, 0), H.P1toP1LinearRestriction(),
                                     H.P1toP1LinearProlongation(), 0, L, nu, nu, 0,
                                     H.CycleType.WCYCLE if cycle == "W" else H.CycleType.VCYCLE)
    og = O.GMG(om, 0, L, nu, nu, 1.0, 1, wcycle=(cycle == "W"), symmetric=symmetric)
    return A, OA, b, x, ob, ox, gmg, og


@pytest.mark.parametrize("kind,L", [(("cube", 2), 4), (("gmsh", "regular_octahedron_8el.msh"), 3), (("cube", 1), 5)])
@pytest.mark.parametrize("cycle,symmetric", [("V", True), ("W", False), ("W", True)])
def test_cycle_variants_match_oracle(kind, L, cycle, symmetric, gpu_ctx, oracle_lib):
    import hyteg_amd as H
    O = oracle_lib
    say(f"[cycle] mesh={kind} level={L} cycle={cycle} symmetric={symmetric}")
    om, S = mesh_pair(kind, gpu_ctx)
    A, OA, b, x, ob, ox, gmg, og = _setup(H, O, om, S, L, 2, cycle, symmetric)
    for cyc in range(3):
        gmg.solve(A, x, b, L)
        og.solve(OA, ox, ob, L)
        assert rel_err(x.download(L), ox.get(L)) <= 1e-11, cyc
    # the same cycles replayed from a captured graph
    xg = H.P1Function("xg", S, 0, L)
    gmg.solve_graph(A, xg, b, L, 3)
    assert rel_err(xg.download(L), ox.get(L)) <= 1e-11


@pytest.mark.parametrize("kind,L", [(("cube", 1), 5), (("cube", 2), 4)])
def test_fmg_matches_oracle(kind, L, gpu_ctx, oracle_lib):
    import hyteg_amd as H
    O = oracle_lib
    say(f"[fmg] mesh={kind} level={L}")
    om, S = mesh_pair(kind, gpu_ctx)
    A, OA, b, x, ob, ox, gmg, og = _setup(H, O, om, S, L, 3, "V", False)
    fmg = H.FullMultigridSolver(S, gmg, H.P1toP1LinearProlongation(), 0, L, 2)
    fmg.solve(A, x, b, L)
    og.fmg(OA, ox, ob, L, 2)
    assert rel_err(x.download(L), ox.get(L)) <= 1e-11
