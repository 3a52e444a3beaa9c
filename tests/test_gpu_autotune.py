"""The autotuner (hymi_mesh_autotune): it keeps one of its candidates and the fastest measured one, and the
results of the apply, an SOR sweep and a V-cycle do not change, bit for bit."""
import numpy as np
import pytest

from helpers import mesh_pair, say

pytestmark = pytest.mark.gpu

COLUMN = [(rows, waves, block) for rows in (4, 8, 16, 32) for waves in (2, 4, 8) for block in (1, 2)]


def _outputs(H, S, L):
    A = H.P1ConstantLaplaceOperator(S, 0, L)
    M = H.P1ConstantMassOperator(S, 0, L)
    x, y, b, t = (H.P1Function(k, S, 0, L) for k in "xybt")
    rng = np.random.default_rng(3)
    x.upload(L, rng.standard_normal(S.owned_count(L)))
    A.apply(x, y, L, H.INNER)
    t.interpolate_modes(np.array([[3 * np.pi ** 2, 1, 1, 1]]), L)
    M.apply(t, b, L, H.ALL)
    A.smooth_sor(x, b, 1.0, L, H.INNER)
    sweep = x.download(L)
    x.interpolate(0.0, L, H.ALL)
    g = H.GeometricMultigridSolver(S, H.SORSmoother(1.0), H.CGSolver(S, 0, 0), H.P1toP1LinearRestriction(),
                                   H.P1toP1LinearProlongation(), 0, L, 2, 2)
    g.solve_graph(A, x, b, L, 2)
    return y.download(L), sweep, x.download(L)


@pytest.mark.parametrize("L", [4, 6])
def test_autotune_keeps_fastest_candidate_and_results(L, gpu_ctx):
    import hyteg_amd as H
    say(f"[autotune] cube n=2 level {L}")
    _, S = mesh_pair(("cube", 2), gpu_ctx)
    before = _outputs(H, S, L)
    rep = S.autotune(L, reps=3)
    sched = S.get_schedule(L)
    if L >= 5:
        assert len(rep) == len(COLUMN) + 4
        times = rep[:len(COLUMN)]
        assert sched["kind"] == 3
        k = COLUMN.index((sched["rows"], sched["waves"], sched["block"]))
    else:
        assert len(rep) == 3 + 4
        times = rep[:3]
        assert sched["kind"] == 1 and sched["waves"] == 4 and sched["block"] == 1
        k = (1, 2, 4).index(sched["rows"])
    valid = [t for t in times if t > 0]
    assert times[k] == min(valid)
    tile, bits = S.get_cell_sor_setting(L)
    pairs = rep[-4:]
    for i, bit in enumerate((1, 2)):
        fused_faster = pairs[2 * i + 1] < pairs[2 * i]
        assert bool(tile and (bits & bit)) == fused_faster, (pairs, tile, bits)
    say(f"  schedule {sched}, cell phase tile {tile} precisions {bits}, fp64 passes/fused {pairs[:2]}")
    after = _outputs(H, S, L)
    for a, c in zip(before, after):
        assert np.array_equal(a, c)
