"""GPU tests of the fused residual and restriction (the macro-cell interiors' residual restricted to the coarse
level in one pass, r written only next to the faces): V-cycles with it switched on at every level of the hierarchy
must follow the oracle's V-cycle to 1e-11 relative, and equal V-cycles that run the two passes bit for bit (same
terms, same order) -- through captured graphs recorded before a switch, FP64 and FP32, levels 6 to 8, on the cube and
on an unstructured mesh."""
import numpy as np
import pytest

from helpers import mesh_pair, rel_err, say, sine_rhs

pytestmark = pytest.mark.gpu


def _select(S, L, mode):
    for k in range(3, L + 1):
        S.set_residual_restrict(k, mode)
        assert S.get_residual_restrict(k) == (int(k >= 8) if mode < 0 else mode)


def _setup(H, S, L, prec, nu=3):
    A = H.P1ConstantLaplaceOperator(S, 0, L)
    M = H.P1ConstantMassOperator(S, 0, L)
    t, b = H.P1Function("t", S, 0, L, precision=prec), H.P1Function("b", S, 0, L, precision=prec)
    t.interpolate(sine_rhs, L)
    M.apply(t, b, L, H.ALL)
    gmg = H.GeometricMultigridSolver(S, H.SORSmoother(1.0), H.CGSolver(S, 0, 0), H.P1toP1LinearRestriction(),
                                     H.P1toP1LinearProlongation(), 0, L, nu, nu)
    return A, b, gmg


@pytest.mark.parametrize("kind,L", [(("cube", 1), 5), (("cube", 2), 4), (("gmsh", "regular_octahedron_8el.msh"), 4)])
def test_fused_vcycles_match_oracle(kind, L, gpu_ctx, oracle_lib):
    import hyteg_amd as H
    O = oracle_lib
    say(f"[residual-restrict] mesh={kind} level={L} fused on levels 3..{L} vs oracle")
    om, S = mesh_pair(kind, gpu_ctx)
    _select(S, L, 1)
    A, b, gmg = _setup(H, S, L, H.F64)
    x = H.P1Function("x", S, 0, L)
    OA, OM = O.Operator(om, 0, L), O.Operator(om, 0, L, 1)
    ot, ob, ox = (O.Function(om, 0, L) for _ in range(3))
    ot.set(L, sine_rhs(om.owned_coords(L)))
    OM.apply(ot, ob, L, flag=O.ALL)
    og = O.GMG(om, 0, L, 3, 3, 1.0, 1)
    assert rel_err(b.download(L), ob.get(L)) <= 1e-13
    for cyc in range(3):
        gmg.solve(A, x, b, L)
        og.solve(OA, ox, ob, L)
        assert rel_err(x.download(L), ox.get(L)) <= 1e-11, cyc
    _select(S, L, -1)


@pytest.mark.parametrize("prec_name,kind,L", [("F64", ("cube", 1), 7), ("F32", ("cube", 1), 7),
                                              ("F64", ("gmsh", "regular_octahedron_8el.msh"), 6),
                                              ("F64", ("cube", 1), 8), ("F32", ("cube", 1), 8)])
def test_fused_equals_two_passes_in_graphs(prec_name, kind, L, gpu_ctx):
    """The same V(2,2) cycles replayed from captured graphs with the fused pass off, on and off again (each switch
    re-captures the graph): all three give the same iterates bit for bit."""
    import hyteg_amd as H
    prec = getattr(H, prec_name)
    om, S = mesh_pair(kind, gpu_ctx)
    A, b, gmg = _setup(H, S, L, prec, nu=2)
    u = H.P1Function("u", S, 0, L, precision=prec)
    runs = []
    for mode in (0, 1, 0):
        say(f"[residual-restrict] mesh={kind} level={L} {prec_name} V(2,2) graphs mode={mode}")
        _select(S, L, mode)
        u.interpolate(0.0, L, H.ALL)
        gmg.solve_graph(A, u, b, L, count=3)
        runs.append(u.download(L))
    _select(S, L, -1)
    assert np.array_equal(runs[0], runs[2])
    say(f"fused vs two passes: rel err {rel_err(runs[1], runs[0]):.3e}")
    assert np.array_equal(runs[1], runs[0])
