"""GPU tests of the serving loop (hyteg_amd.serving): continuous batching of small problems into 32-column batch
lanes, large problems on solve lanes, all sharing the GPU, resident hierarchies, HBM budget.

Bars: every request's solution in the mix equals the same request served alone, bit for bit; every solve
reaches its tolerance; batches hold several requests at once; work of different lanes overlaps in time;
resident hierarchies never exceed the budget.
"""
import numpy as np
import pytest

from helpers import say

pytestmark = pytest.mark.gpu


def test_serving_mix_equals_alone(gpu_ctx):
    from hyteg_amd.serving import PoissonServer, poisson_stream, summarize
    reqs = poisson_stream(24, rate=2000.0, seed=5, ns=(1, 2), levels=(2, 3, 4), tol=1e-8)
    kw = dict(slots=2, keep_solutions=True, batch_max_unknowns=4000)   # level 4 on two sub-cubes: solve lanes
    say("[serving] each request alone")
    alone = PoissonServer(**kw)
    ra = alone.serve(reqs, one_at_a_time=True)
    alone.close()
    say("[serving] the stream")
    mix = PoissonServer(**kw)
    rm = mix.serve(reqs, realtime=True)
    mix.close()
    assert not alone.errors and not mix.errors, (alone.errors, mix.errors)
    for q, a, m in zip(reqs, ra, rm):
        assert a is not None and m is not None
        assert a.rel_residual <= 1e-8 and m.rel_residual <= 1e-8
        assert a.batched == m.batched == mix.batched(q.n, q.level)
        assert a.outer_steps == m.outer_steps
        assert np.array_equal(a.solution, m.solution), a.rid
    assert any(m.batched for m in rm) and any(not m.batched for m in rm)
    # some request joined or shared a batch in flight: its time in the lane overlaps another of the lane's
    same = [(p, q) for p in rm for q in rm if p.rid < q.rid and p.batched and p.slot == q.slot]
    assert any(p.start < q.end and q.start < p.end for p, q in same)
    assert len({r.slot for r in rm}) > 1
    s = summarize(rm, mix.wall)
    say(f"  {s}")
    assert s["completed"] == len(reqs)


def test_serving_respects_hbm_budget(gpu_ctx):
    from hyteg_amd.serving import PoissonServer, hierarchy_bytes, poisson_stream
    reqs = poisson_stream(10, rate=2000.0, seed=9, ns=(1, 2), levels=(3, 4), tol=1e-6)
    need = max(hierarchy_bytes(r.n, r.level) for r in reqs)
    say(f"[serving budget] budget {int(1.5 * need)} B")
    srv = PoissonServer(slots=3, hbm_budget=int(1.5 * need), batch_max_unknowns=0)
    res = srv.serve(reqs, realtime=False)
    srv.close()
    assert not srv.errors, srv.errors
    assert all(r is not None and r.rel_residual <= 1e-6 for r in res)
    assert srv.peak_resident <= srv.budget
