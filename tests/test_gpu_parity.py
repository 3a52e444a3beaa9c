"""GPU parity of the HIP path against the CPU oracle, through the C-ABI (hyteg_amd binds include/hymi.h).

Tolerances: FP64 stencil apply / residual <= 1e-13 relative (the target for this path); transfers and
the multicolour SOR, whose operation order the oracle reproduces, <= 1e-12; FP32 storage <= 1e-5.
"""
import numpy as np
import pytest

from helpers import mesh_pair, rel_err, say, sine_exact, sine_rhs

pytestmark = pytest.mark.gpu

MESH_CASES = [("cube", 1), ("cube", 2), ("gmsh", "regular_octahedron_8el.msh"), ("gmsh", "cube_48el.msh")]


def _pair(kind, gpu_ctx):
    return mesh_pair(kind, gpu_ctx)


@pytest.mark.parametrize("kind", MESH_CASES)
@pytest.mark.parametrize("level", [1, 2, 3, 4])
def test_apply_fp64(kind, level, gpu_ctx, oracle_lib):
    import hyteg_amd as H
    O = oracle_lib
    say(f"[apply_fp64] mesh={kind} level={level}")
    om, S = _pair(kind, gpu_ctx)
    rng = np.random.default_rng(1000 + level)
    x = rng.standard_normal(om.owned_count(level))
    A = H.P1ConstantLaplaceOperator(S, 0, level)
    OA = O.Operator(om, 0, level)
    for flag in (H.ALL, H.INNER):
        u = H.P1Function("u", S, 0, level)
        y = H.P1Function("y", S, 0, level)
        u.upload(level, x)
        y0 = rng.standard_normal(len(x))
        y.upload(level, y0)
        A.apply(u, y, level, flag)
        ou, oy = O.Function(om, 0, level), O.Function(om, 0, level)
        ou.set(level, x)
        oy.set(level, y0)
        OA.apply(ou, oy, level, flag=flag)
        got, ref = y.download(level), oy.get(level)
        assert rel_err(got, ref) <= 1e-13, (kind, level, flag, rel_err(got, ref))
        # Add mode
        A.apply(u, y, level, flag, H.ADD)
        OA.apply(ou, oy, level, flag=flag, update=1)
        assert rel_err(y.download(level), oy.get(level)) <= 1e-13


@pytest.mark.parametrize("kind", MESH_CASES[:2])
def test_mass_and_residual_fp64(kind, gpu_ctx, oracle_lib):
    import hyteg_amd as H
    O = oracle_lib
    level = 3
    say(f"[mass_residual] mesh={kind} level={level}")
    om, S = _pair(kind, gpu_ctx)
    rng = np.random.default_rng(7)
    x, b = rng.standard_normal((2, om.owned_count(level)))
    M = H.P1ConstantMassOperator(S, 0, level)
    OM = O.Operator(om, 0, level, 1)
    u, y, r, bb = (H.P1Function(n, S, 0, level) for n in "uyrb")
    u.upload(level, x)
    bb.upload(level, b)
    M.apply(u, y, level, H.ALL)
    ou, oy = O.Function(om, 0, level), O.Function(om, 0, level)
    ou.set(level, x)
    OM.apply(ou, oy, level, flag=O.ALL)
    assert rel_err(y.download(level), oy.get(level)) <= 1e-13
    A = H.P1ConstantLaplaceOperator(S, 0, level)
    OA = O.Operator(om, 0, level)
    A.residual(u, bb, r, level, H.INNER)
    ob, ot, orr = O.Function(om, 0, level), O.Function(om, 0, level), O.Function(om, 0, level)
    ob.set(level, b)
    OA.apply(ou, ot, level, flag=O.INNER)
    orr.assign(level, [1, -1], [ob, ot], flag=O.INNER)
    types = om.owned_types(level)
    sel = types == O.INNER
    assert rel_err(r.download(level)[sel], orr.get(level)[sel]) <= 1e-13


@pytest.mark.parametrize("kind", MESH_CASES)
@pytest.mark.parametrize("level", [1, 2, 3, 5])
def test_sor_multicolour_matches_oracle_schedule(kind, level, gpu_ctx, oracle_lib):
    import hyteg_amd as H
    O = oracle_lib
    say(f"[sor] mesh={kind} level={level}")
    om, S = _pair(kind, gpu_ctx)
    rng = np.random.default_rng(11 + level)
    x, b = rng.standard_normal((2, om.owned_count(level)))
    A = H.P1ConstantLaplaceOperator(S, 0, level)
    OA = O.Operator(om, 0, level)
    u, f = H.P1Function("u", S, 0, level), H.P1Function("f", S, 0, level)
    u.upload(level, x)
    f.upload(level, b)
    ou, of = O.Function(om, 0, level), O.Function(om, 0, level)
    ou.set(level, x)
    of.set(level, b)
    for sweep in range(3):
        A.smooth_sor(u, f, 0.9, level, H.INNER)
        OA.smooth_sor(ou, of, level, omega=0.9, flag=O.INNER, order=1)
        assert rel_err(u.download(level), ou.get(level)) <= 1e-12, sweep


@pytest.mark.parametrize("kind", MESH_CASES)
@pytest.mark.parametrize("level", [1, 2, 3, 4])
def test_transfers(kind, level, gpu_ctx, oracle_lib):
    import hyteg_amd as H
    O = oracle_lib
    say(f"[transfers] mesh={kind} level={level}")
    om, S = _pair(kind, gpu_ctx)
    rng = np.random.default_rng(level)
    f = H.P1Function("f", S, 0, level)
    of = O.Function(om, 0, level)
    fine = rng.standard_normal(om.owned_count(level))
    coarse = rng.standard_normal(om.owned_count(level - 1))
    for L, v in ((level, fine), (level - 1, coarse)):
        f.upload(L, v)
        of.set(L, v)
    H.P1toP1LinearRestriction().restrict(f, level, H.INNER)
    O.restrict(of, level, flag=O.INNER)
    assert rel_err(f.download(level - 1), of.get(level - 1)) <= 1e-13
    H.P1toP1LinearProlongation().prolongateAndAdd(f, level - 1, H.INNER)
    O.prolongate(of, level - 1, flag=O.INNER, add=True)
    assert rel_err(f.download(level), of.get(level)) <= 1e-13


@pytest.mark.parametrize("kind", [("cube", 2), ("gmsh", "regular_octahedron_8el.msh")])
def test_cg(kind, gpu_ctx, oracle_lib):
    import hyteg_amd as H
    O = oracle_lib
    level = 2
    say(f"[cg] mesh={kind} level={level}")
    om, S = _pair(kind, gpu_ctx)
    X = om.owned_coords(level)
    b = sine_rhs(X)
    A = H.P1ConstantLaplaceOperator(S, 0, level)
    OA = O.Operator(om, 0, level)
    x, f = H.P1Function("x", S, 0, level), H.P1Function("f", S, 0, level)
    f.upload(level, b)
    ox, of = O.Function(om, 0, level), O.Function(om, 0, level)
    of.set(level, b)
    H.CGSolver(S, level, level, maxIter=40, tolerance=1e-14).solve(A, x, f, level)
    O.CG(om, level, level, 40, 1e-14).solve(OA, ox, of, level)
    assert rel_err(x.download(level), ox.get(level)) <= 1e-9


def _gmg_setup(H, O, om, S, L, order):
    X = om.owned_coords(L)
    A = H.P1ConstantLaplaceOperator(S, 0, L)
    M = H.P1ConstantMassOperator(S, 0, L)
    OA, OM = O.Operator(om, 0, L), O.Operator(om, 0, L, 1)
    t, b, x = (H.P1Function(n, S, 0, L) for n in "tbx")
    t.interpolate(sine_rhs, L)
    M.apply(t, b, L, H.ALL)
    ot, ob, ox = (O.Function(om, 0, L) for _ in range(3))
    ot.set(L, sine_rhs(X))
    OM.apply(ot, ob, L, flag=O.ALL)
    gmg = H.GeometricMultigridSolver(S, H.SORSmoother(1.0), H.CGSolver(S, 0, 0), H.P1toP1LinearRestriction(),
                                     H.P1toP1LinearProlongation(), 0, L, 3, 3)
    og = O.GMG(om, 0, L, 3, 3, 1.0, order)
    return A, OA, b, x, ob, ox, gmg, og


def _residual_norm(H, A, x, b, L, S):
    r = H.P1Function("r", S, 0, L)
    A.residual(x, b, r, L, H.INNER)
    return np.sqrt(r.dotGlobal(r, L, H.INNER))


def _oracle_residual_norm(O, OA, ox, ob, L, om):
    t, r = O.Function(om, 0, L), O.Function(om, 0, L)
    OA.apply(ox, t, L, flag=O.INNER)
    r.assign(L, [1, -1], [ob, t], flag=O.INNER)
    return np.sqrt(r.dot(r, L, O.INNER))


# cube n=3: 8 interior vertices on level 0, so the one-workgroup coarse CG runs several iterations
@pytest.mark.parametrize("kind,L", [(("cube", 1), 4), (("cube", 2), 3), (("cube", 3), 3),
                                    (("gmsh", "regular_octahedron_8el.msh"), 3)])
def test_gmg_vcycle_matches_oracle_schedule(kind, L, gpu_ctx, oracle_lib):
    import hyteg_amd as H
    O = oracle_lib
    say(f"[gmg_schedule] mesh={kind} level={L}")
    om, S = _pair(kind, gpu_ctx)
    A, OA, b, x, ob, ox, gmg, og = _gmg_setup(H, O, om, S, L, order=1)
    assert rel_err(b.download(L), ob.get(L)) <= 1e-13
    for cyc in range(3):
        gmg.solve(A, x, b, L)
        og.solve(OA, ox, ob, L)
        assert rel_err(x.download(L), ox.get(L)) <= 1e-11, cyc


def test_gmg_config1_against_lexicographic_reference(gpu_ctx, oracle_lib):
    """6-tet unit cube, level 5, V(3,3), SOR omega=1, CG at level 0, 10 cycles: the GPU multicolour SOR's
    per-cycle convergence factor must be within 5 % of the reference's lexicographic SOR, and the final
    discretisation error within 1 %."""
    import hyteg_amd as H
    O = oracle_lib
    L = 5
    say(f"[gmg_config1] cube n=1 level={L}")
    om, S = _pair(("cube", 1), gpu_ctx)
    A, OA, b, x, ob, ox, gmg, og = _gmg_setup(H, O, om, S, L, order=0)
    r_gpu = [_residual_norm(H, A, x, b, L, S)]
    r_ref = [_oracle_residual_norm(O, OA, ox, ob, L, om)]
    for _ in range(10):
        gmg.solve(A, x, b, L)
        og.solve(OA, ox, ob, L)
        r_gpu.append(_residual_norm(H, A, x, b, L, S))
        r_ref.append(_oracle_residual_norm(O, OA, ox, ob, L, om))
    # asymptotic convergence factor: geometric mean over cycles 2..10
    rho_gpu = (r_gpu[10] / r_gpu[1]) ** (1 / 9)
    rho_ref = (r_ref[10] / r_ref[1]) ** (1 / 9)
    say(f"rho gpu={rho_gpu:.4f} ref={rho_ref:.4f}")
    assert abs(rho_gpu - rho_ref) <= 0.05 * rho_ref
    X = om.owned_coords(L)
    inner = om.owned_types(L) == O.INNER
    u = sine_exact(X)
    err_gpu = np.sqrt(np.sum((x.download(L) - u)[inner] ** 2) / len(X))
    err_ref = np.sqrt(np.sum((ox.get(L) - u)[inner] ** 2) / len(X))
    say(f"L2 error gpu={err_gpu:.6e} ref={err_ref:.6e}")
    assert abs(err_gpu - err_ref) <= 0.01 * err_ref


def test_graph_replay_matches_eager(gpu_ctx, oracle_lib):
    import hyteg_amd as H
    O = oracle_lib
    L = 4
    say(f"[graph] cube n=2 level={L}")
    om, S = _pair(("cube", 2), gpu_ctx)
    A, OA, b, x, ob, ox, gmg, og = _gmg_setup(H, O, om, S, L, order=1)
    x2 = H.P1Function("x2", S, 0, L)
    for k in range(3):
        gmg.solve(A, x, b, L)                 # eager
        gmg.solve_graph(A, x2, b, L, 1)       # captured once, replayed
        _ = x.dotGlobal(x, L, H.INNER)        # eager work and a sync between replays
        assert rel_err(x2.download(L), x.download(L)) == 0.0, k


@pytest.mark.parametrize("prec", [0, 1])
def test_march_schedules_bitwise(prec, gpu_ctx, oracle_lib):
    """Every launch schedule of the row-marching kernels computes the same values in the same order."""
    import hyteg_amd as H
    O = oracle_lib
    L = 4
    say(f"[schedules] cube n=2 level={L} precision={prec}")
    om, S = _pair(("cube", 2), gpu_ctx)
    rng = np.random.default_rng(5)
    xh, bh = rng.standard_normal((2, om.owned_count(L)))
    A = H.P1ConstantLaplaceOperator(S, 0, L)
    x, b, y = (H.P1Function(n, S, 0, L, precision=prec) for n in "xby")
    b.upload(L, bh)

    def run():
        x.upload(L, xh)
        A.apply(x, y, L, H.INNER)
        out = [y.download(L)]
        A.smooth_sor(x, b, 0.8, L, H.INNER)
        out.append(x.download(L))
        return out

    S.set_schedule(L)
    ref = run()
    for rows, waves, block, kind in [(1, 1, 1, 0), (3, 2, 2, 0), (4, 4, 8, 0), (8, 4, 4, 0), (64, 2, 8, 0),
                                     (2, 1, 8, 0), (13, 4, 2, 0), (7, 1, 4, 0), (1, 1, 1, 1), (4, 2, 2, 1),
                                     (8, 4, 4, 1), (3, 1, 8, 1), (2, 2, 1, 1), (64, 4, 4, 1),
                                     (4, 4, 1, 2), (3, 4, 2, 2), (8, 2, 5, 2), (12, 4, 64, 2), (2, 1, 3, 2),
                                     (1, 1, 1, 3), (2, 2, 2, 3), (3, 4, 4, 3), (5, 1, 2, 3), (16, 2, 2, 3),
                                     (64, 4, 1, 3), (4, 8, 9, 3), (16, 16, 1, 3), (2, 16, 10, 3)]:
        S.set_schedule(L, rows, waves, block, kind)
        got = run()
        for a, r in zip(got, ref):
            assert np.array_equal(a, r), (rows, waves, block, kind)
    S.set_schedule(L)


@pytest.mark.parametrize("prec", [0, 1])
def test_column_schedules_bitwise_level6(prec, gpu_ctx):
    """The column march at its production parameters (and neighbours of them) on level 6 equals the default
    schedule bit for bit: apply, residual and an SOR sweep."""
    import hyteg_amd as H
    L = 6
    say(f"[column schedules] cube n=2 level={L} precision={prec}")
    S = H.PrimitiveStorage(H.MeshInfo.meshCuboid((0, 0, 0), (1, 1, 1), 2, 2, 2), gpu_ctx)
    rng = np.random.default_rng(11)
    xh, bh = rng.standard_normal((2, S.owned_count(L)))
    A = H.P1ConstantLaplaceOperator(S, L, L)
    x, b, y = (H.P1Function(n, S, L, L, precision=prec) for n in "xby")
    b.upload(L, bh)

    def run():
        x.upload(L, xh)
        A.apply(x, y, L, H.INNER)
        out = [y.download(L)]
        A.residual(x, b, y, L, H.INNER)
        out.append(y.download(L))
        A.smooth_sor(x, b, 1.0, L, H.INNER)
        out.append(x.download(L))
        return out

    S.set_schedule(L, 1, 4, 1, 1)   # flat slab windows: the reference schedule
    ref = run()
    for rows, waves, block in [(8, 4, 1), (16, 4, 1), (4, 2, 1), (8, 1, 2), (32, 2, 4), (1, 4, 1), (16, 16, 9),
                               (8, 8, 1), (16, 8, 12)]:
        S.set_schedule(L, rows, waves, block, 3)
        got = run()
        for a, r in zip(got, ref):
            assert np.array_equal(a, r), (rows, waves, block)
    S.set_schedule(L)


def test_graph_recaptured_after_schedule_change(gpu_ctx, oracle_lib):
    import hyteg_amd as H
    O = oracle_lib
    L = 4
    say(f"[graph schedule] cube n=2 level={L}")
    om, S = _pair(("cube", 2), gpu_ctx)
    A, OA, b, x, ob, ox, gmg, og = _gmg_setup(H, O, om, S, L, order=1)
    x2 = H.P1Function("x2", S, 0, L)
    gmg.solve(A, x, b, L)
    gmg.solve_graph(A, x2, b, L, 1)
    assert rel_err(x2.download(L), x.download(L)) == 0.0
    S.set_schedule(L, 2, 4, 8)      # frees the work-unit table the captured graph used
    gmg.solve(A, x, b, L)
    gmg.solve_graph(A, x2, b, L, 1)
    _ = x.dotGlobal(x, L, H.INNER)
    assert rel_err(x2.download(L), x.download(L)) == 0.0
    S.set_schedule(L)


def test_fp32_apply_and_vcycle(gpu_ctx, oracle_lib):
    import hyteg_amd as H
    O = oracle_lib
    L = 4
    say(f"[fp32] cube n=2 level={L}")
    om, S = _pair(("cube", 2), gpu_ctx)
    rng = np.random.default_rng(3)
    x = rng.standard_normal(om.owned_count(L))
    A = H.P1ConstantLaplaceOperator(S, 0, L)
    u = H.P1Function("u", S, 0, L, precision=H.F32)
    y = H.P1Function("y", S, 0, L, precision=H.F32)
    u.upload(L, x)
    A.apply(u, y, L, H.ALL)
    OA = O.Operator(om, 0, L)
    ou, oy = O.Function(om, 0, L), O.Function(om, 0, L)
    ou.set(L, x.astype(np.float32).astype(np.float64))
    OA.apply(ou, oy, L, flag=O.ALL)
    assert rel_err(y.download(L), oy.get(L)) <= 1e-5


def test_nrhs32_matches_single(gpu_ctx):
    import hyteg_amd as H
    L = 3
    say(f"[nrhs32] cube n=2 level={L}")
    _, S = _pair(("cube", 2), gpu_ctx)
    n = S.owned_count(L)
    rng = np.random.default_rng(5)
    X = rng.standard_normal((n, 32))
    A = H.P1ConstantLaplaceOperator(S, 0, L)
    ub = H.P1Function("ub", S, 0, L, nrhs=32)
    yb = H.P1Function("yb", S, 0, L, nrhs=32)
    ub.upload(L, X)
    A.apply(ub, yb, L, H.ALL)
    Yb = yb.download(L)
    u1, y1 = H.P1Function("u1", S, 0, L), H.P1Function("y1", S, 0, L)
    for r in (0, 7, 31):
        u1.upload(L, X[:, r])
        A.apply(u1, y1, L, H.ALL)
        assert rel_err(Yb[:, r], y1.download(L)) <= 1e-14


@pytest.mark.parametrize("kind", [("cube", 2), ("gmsh", "regular_octahedron_8el.msh")])
def test_reference_primitive_layout(kind, gpu_ctx, oracle_lib):
    """download_primitive returns the reference's FunctionMemory layout (including ghost layers)."""
    import hyteg_amd as H
    O = oracle_lib
    L = 3
    say(f"[layout] mesh={kind} level={L}")
    om, S = _pair(kind, gpu_ctx)
    x = np.random.default_rng(9).standard_normal(om.owned_count(L))
    f = H.P1Function("f", S, 0, L)
    f.upload(L, x)
    of = O.Function(om, 0, L)
    of.set(L, x)
    for d in range(6):
        of.communicate(L, d)
    nV, nE, nF, nC = om.counts()
    for cls, n in ((H.VERTEX, nV), (H.EDGE, nE), (H.FACE, nF), (H.CELL, nC)):
        for idx in range(n):
            got = f.download_primitive(L, cls, idx)
            ref = of.prim(L, cls, idx)
            assert got.shape == ref.shape, (cls, idx)
            assert np.array_equal(got, ref), (cls, idx)


@pytest.mark.parametrize("nrhs,prec", [(1, 0), (32, 0), (1, 1)])
@pytest.mark.parametrize("kind", [("cube", 2), ("gmsh", "regular_octahedron_8el.msh")])
def test_interpolate_modes_matches_host(kind, nrhs, prec, gpu_ctx):
    """Device evaluation of Fourier-mode sums equals host evaluation at the owned coordinates; only flag DoFs
    are written and the closure copies are current (an apply sees the same values as after an upload)."""
    import hyteg_amd as H
    L = 3
    say(f"[modes] mesh={kind} level={L} nrhs={nrhs} precision={prec}")
    _, S = _pair(kind, gpu_ctx)
    rng = np.random.default_rng(17)
    modes = np.zeros((nrhs, 4, 4))
    modes[:, :, 0] = rng.standard_normal((nrhs, 4))
    modes[:, :, 1:] = rng.integers(1, 9, (nrhs, 4, 3))
    X = S.owned_coords(L)
    want = np.zeros((len(X), nrhs))
    for r in range(nrhs):
        for a, k, l, m in modes[r]:
            want[:, r] += a * np.sin(k * np.pi * X[:, 0]) * np.sin(l * np.pi * X[:, 1]) * np.sin(m * np.pi * X[:, 2])
    want = want[:, 0] if nrhs == 1 else want
    tol = 1e-13 if prec == 0 else 1e-6
    f = H.P1Function("f", S, 0, L, precision=prec, nrhs=nrhs)
    f.interpolate_modes(modes, L, H.ALL)
    assert rel_err(f.download(L), want) <= tol
    inner = S.owned_types(L) == H.INNER
    g = H.P1Function("g", S, 0, L, precision=prec, nrhs=nrhs)
    g.interpolate_modes(modes, L, H.INNER)
    got = g.download(L)
    assert np.all(got[~inner] == 0)
    assert rel_err(got[inner], want[inner]) <= tol
    A = H.P1ConstantLaplaceOperator(S, 0, L)
    h, y1, y2 = (H.P1Function(n, S, 0, L, precision=prec, nrhs=nrhs) for n in ("h", "y1", "y2"))
    h.upload(L, f.download(L))
    A.apply(f, y1, L, H.INNER)
    A.apply(h, y2, L, H.INNER)
    assert np.array_equal(y1.download(L), y2.download(L))


def test_graph_replays_survive_other_solvers_and_new_functions(gpu_ctx):
    """A captured V-cycle keeps working after larger solvers are created on the same context (the context's
    reduction buffers are never reallocated) and is never replayed for a new function that happens to reuse a
    destroyed function's address (graphs are keyed on function ids)."""
    import hyteg_amd as H
    L = 4
    say(f"[graph lifetime] cube n=1 level={L}, then cube n=2 level 5 on the same context")
    S = H.PrimitiveStorage(H.MeshInfo.meshCuboid((0, 0, 0), (1, 1, 1), 1, 1, 1), gpu_ctx)
    A = H.P1ConstantLaplaceOperator(S, 0, L)

    def gmg(St, Lmax):
        return H.GeometricMultigridSolver(St, H.SORSmoother(1.0), H.CGSolver(St, 0, 0), H.P1toP1LinearRestriction(),
                                          H.P1toP1LinearProlongation(), 0, Lmax, 2, 2)

    g = gmg(S, L)
    b = H.P1Function("b", S, 0, L)
    b.interpolate(1.0, L, H.INNER)
    x = H.P1Function("x", S, 0, L)
    g.solve_graph(A, x, b, L, 2)                       # captured here
    S2 = H.PrimitiveStorage(H.MeshInfo.meshCuboid((0, 0, 0), (1, 1, 1), 2, 2, 2), gpu_ctx)
    A2 = H.P1ConstantLaplaceOperator(S2, 0, 5)
    g2 = gmg(S2, 5)
    x2, b2 = H.P1Function("x2", S2, 0, 5), H.P1Function("b2", S2, 0, 5)
    b2.interpolate(1.0, 5, H.INNER)
    g2.solve_graph(A2, x2, b2, 5, 1)
    ref = H.P1Function("ref", S, 0, L)
    ref.interpolate(0.0, L, H.ALL)
    for _ in range(5):
        g.solve(A, ref, b, L)                          # eager cycles
    x.interpolate(0.0, L, H.ALL)
    g.solve_graph(A, x, b, L, 5)                       # replay of the graph captured before g2 existed
    assert np.array_equal(x.download(L), ref.download(L))
    for k in range(3):                                 # new functions, possibly at a reused address
        del x
        x = H.P1Function("x", S, 0, L)
        g.solve_graph(A, x, b, L, 5)
        assert np.array_equal(x.download(L), ref.download(L)), k
