"""GPU tests of FP32 V-cycles inside FP64 iterative refinement (hymi_solve_refined), through the C-ABI.

Bars: the refined solution reaches the requested relative residual, checked independently with the CPU
oracle's FP64 operator; it agrees with the oracle's own FP64 multigrid solution of the same system to
1e-8 relative (both are solved to <= 1e-10 relative residual, and the condition number at these levels
bounds the difference well below that); the 32-RHS interleaved refinement matches single-RHS FP64
solutions column by column to 1e-6 relative (tolerance 1e-8 on each column's residual).
"""
import numpy as np
import pytest

from helpers import mesh_pair, rel_err, say, sine_rhs

pytestmark = pytest.mark.gpu


def _rhs_with_noise(H, S, A, M, L, rng):
    """b = M I_h f + U[-1,1] * 1e-3 per DoF (the mixed-precision workload's right-hand side)."""
    t, b = H.P1Function("t", S, 0, L), H.P1Function("b", S, 0, L)
    t.interpolate(sine_rhs, L)
    M.apply(t, b, L, H.ALL)
    bv = b.download(L) + 1e-3 * rng.uniform(-1.0, 1.0, S.owned_count(L))
    b.upload(L, bv)
    return b, bv


def _oracle_rel_residual(O, om, L, xv, bv):
    OA = O.Operator(om, 0, L)
    ox, ob, t, r = (O.Function(om, 0, L) for _ in range(4))
    ox.set(L, xv)
    ob.set(L, bv)
    OA.apply(ox, t, L, flag=O.INNER)
    r.assign(L, [1, -1], [ob, t], flag=O.INNER)
    r0 = np.sqrt(ob.dot(ob, L, O.INNER))          # residual of x = 0
    return np.sqrt(r.dot(r, L, O.INNER)) / r0


@pytest.mark.parametrize("kind,L", [(("cube", 2), 4), (("gmsh", "regular_octahedron_8el.msh"), 4), (("cube", 1), 5)])
def test_refined_fp32_reaches_tolerance_and_matches_oracle(kind, L, gpu_ctx, oracle_lib):
    import hyteg_amd as H
    O = oracle_lib
    say(f"[refine] mesh={kind} level={L}")
    om, S = mesh_pair(kind, gpu_ctx)
    rng = np.random.default_rng(20251202)
    A, M = H.P1ConstantLaplaceOperator(S, 0, L), H.P1ConstantMassOperator(S, 0, L)
    b, bv = _rhs_with_noise(H, S, A, M, L, rng)
    x = H.P1Function("x", S, 0, L)
    inner = H.GeometricMultigridSolver(S, H.SORSmoother(1.0), H.CGSolver(S, 0, 0), H.P1toP1LinearRestriction(),
                                       H.P1toP1LinearProlongation(), 0, L, 2, 2)
    R = H.MixedPrecisionRefinementSolver(S, inner, tolerance=1e-10, maxOuter=40, cyclesPerOuter=2)
    rel = R.solve(A, x, b, L)
    say(f"  outer steps {R.outerSteps}, history {['%.1e' % h for h in R.history]}")
    assert rel.max() <= 1e-10
    assert R.outerSteps >= 2
    assert all(R.history[k + 1] < R.history[k] for k in range(len(R.history) - 1))
    xv = x.download(L)
    # independent FP64 residual on the CPU oracle (rounding slack over the device's own check)
    assert _oracle_rel_residual(O, om, L, xv, bv) <= 2e-10
    # the oracle's FP64 multigrid solution of the same system
    OA = O.Operator(om, 0, L)
    og = O.GMG(om, 0, L, 3, 3, 1.0, 1)
    oy, ob = O.Function(om, 0, L), O.Function(om, 0, L)
    ob.set(L, bv)
    for _ in range(40):
        og.solve(OA, oy, ob, L)
        if _oracle_rel_residual(O, om, L, oy.get(L), bv) <= 1e-13:
            break
    err = rel_err(xv, oy.get(L))
    say(f"  error vs FP64 oracle solution {err:.2e}")
    assert err <= 1e-8


def test_refined_repeat_solves_reuse_buffers(gpu_ctx):
    """A second solve from the converged x takes no correction step; a fresh x converges the same way."""
    import hyteg_amd as H
    L = 4
    say(f"[refine repeat] cube n=2 level={L}")
    _, S = mesh_pair(("cube", 2), gpu_ctx)
    rng = np.random.default_rng(7)
    A, M = H.P1ConstantLaplaceOperator(S, 0, L), H.P1ConstantMassOperator(S, 0, L)
    b, _ = _rhs_with_noise(H, S, A, M, L, rng)
    inner = H.GeometricMultigridSolver(S, H.SORSmoother(1.0), H.CGSolver(S, 0, 0), H.P1toP1LinearRestriction(),
                                       H.P1toP1LinearProlongation(), 0, L, 2, 2)
    R = H.MixedPrecisionRefinementSolver(S, inner, tolerance=1e-10, maxOuter=40, cyclesPerOuter=2)
    x1, x2 = H.P1Function("x1", S, 0, L), H.P1Function("x2", S, 0, L)
    R.solve(A, x1, b, L)
    steps = R.outerSteps
    R.solve(A, x2, b, L)
    assert R.outerSteps == steps
    assert np.array_equal(x1.download(L), x2.download(L))      # same graph, same buffers, same order


def test_refined_nrhs32_matches_single(gpu_ctx):
    import hyteg_amd as H
    L = 3
    say(f"[refine nrhs32] cube n=2 level={L}")
    _, S = mesh_pair(("cube", 2), gpu_ctx)
    n = S.owned_count(L)
    X = S.owned_coords(L)
    rng = np.random.default_rng(3)
    B = np.zeros((n, 32))
    for r in range(32):                               # 8 Fourier modes per right-hand side
        for _ in range(8):
            k, l, m = rng.integers(1, 9, 3)
            B[:, r] += rng.standard_normal() * (np.sin(k * np.pi * X[:, 0]) * np.sin(l * np.pi * X[:, 1])
                                                * np.sin(m * np.pi * X[:, 2]))
    A = H.P1ConstantLaplaceOperator(S, 0, L)
    bb, xb = H.P1Function("bb", S, 0, L, nrhs=32), H.P1Function("xb", S, 0, L, nrhs=32)
    bb.upload(L, B)
    inner = H.GeometricMultigridSolver(S, H.SORSmoother(1.0), H.CGSolver(S, 0, 0), H.P1toP1LinearRestriction(),
                                       H.P1toP1LinearProlongation(), 0, L, 2, 2)
    R = H.MixedPrecisionRefinementSolver(S, inner, tolerance=1e-8, maxOuter=40, cyclesPerOuter=2)
    rel = R.solve(A, xb, bb, L)
    assert rel.shape == (32,) and rel.max() <= 1e-8
    Xb = xb.download(L)
    g64 = H.GeometricMultigridSolver(S, H.SORSmoother(1.0), H.CGSolver(S, 0, 0), H.P1toP1LinearRestriction(),
                                     H.P1toP1LinearProlongation(), 0, L, 3, 3)
    b1, x1 = H.P1Function("b1", S, 0, L), H.P1Function("x1", S, 0, L)
    for r in (0, 13, 31):
        b1.upload(L, B[:, r])
        x1.interpolate(0.0, L, H.ALL)
        for _ in range(25):
            g64.solve(A, x1, b1, L)
        assert rel_err(Xb[:, r], x1.download(L)) <= 1e-6, r


def test_refined_split_residual_matches_two_pass(gpu_ctx):
    """With the column schedule the refinement residual is written rounded to FP32 in the cell kernel and its
    norm is summed on the way; the corrections, and so the solution, equal the two-pass form (FP64 residual,
    separate conversion and dot) bit for bit, and the recorded norms agree to rounding."""
    import hyteg_amd as H
    L = 5
    say(f"[refine split] cube n=2 level={L}")
    _, S = mesh_pair(("cube", 2), gpu_ctx)
    rng = np.random.default_rng(17)
    A, M = H.P1ConstantLaplaceOperator(S, 0, L), H.P1ConstantMassOperator(S, 0, L)
    b, _ = _rhs_with_noise(H, S, A, M, L, rng)

    def solve():
        inner = H.GeometricMultigridSolver(S, H.SORSmoother(1.0), H.CGSolver(S, 0, 0), H.P1toP1LinearRestriction(),
                                           H.P1toP1LinearProlongation(), 0, L, 2, 2)
        R = H.MixedPrecisionRefinementSolver(S, inner, tolerance=1e-10, maxOuter=40, cyclesPerOuter=2)
        x = H.P1Function("x", S, 0, L)
        rel = R.solve(A, x, b, L)
        return x.download(L), R.outerSteps, np.array(R.history), float(np.max(rel))

    S.set_schedule(L, 1, 4, 1, 1)                 # flat windows on the finest level: two-pass residual
    x1, n1, h1, r1 = solve()
    S.set_schedule(L, 4, 2, 1, 3)                 # column march on the finest level: split residual
    x2, n2, h2, r2 = solve()
    S.set_schedule(L)
    say(f"  outer steps {n1} / {n2}, final {r1:.2e} / {r2:.2e}")
    assert n1 == n2 and r2 <= 1e-10
    assert np.array_equal(x1, x2)
    assert np.allclose(h1, h2, rtol=1e-12, atol=0.0)
