"""GPU tests of the serving loop (hyteg_amd.serving): continuous admission over several workers that share
the GPU, resident hierarchies, HBM budget.

Bars: every request's solution in the mix equals the same request served alone, bit for bit; every solve
reaches its tolerance; solves of different workers overlap in time; resident hierarchies never exceed the
budget.
"""
import numpy as np
import pytest

from helpers import say

pytestmark = pytest.mark.gpu


def test_serving_mix_equals_alone(gpu_ctx):
    from hyteg_amd.serving import PoissonServer, poisson_stream, summarize
    reqs = poisson_stream(12, rate=2000.0, seed=5, ns=(1, 2), levels=(2, 3, 4), tol=1e-8)
    say("[serving] alone")
    alone = PoissonServer(slots=1, keep_solutions=True)
    ra = alone.serve(reqs, realtime=False)
    alone.close()
    say("[serving] mix of 3 workers")
    mix = PoissonServer(slots=3, keep_solutions=True)
    rm = mix.serve(reqs, realtime=True)
    mix.close()
    assert not alone.errors and not mix.errors, (alone.errors, mix.errors)
    for a, m in zip(ra, rm):
        assert a is not None and m is not None
        assert a.rel_residual <= 1e-8 and m.rel_residual <= 1e-8
        assert np.array_equal(a.solution, m.solution), a.rid
    assert len({r.slot for r in rm}) > 1
    overlap = any(p.slot != q.slot and p.start < q.end and q.start < p.end for p in rm for q in rm)
    assert overlap
    s = summarize(rm, mix.wall)
    say(f"  {s}")
    assert s["completed"] == len(reqs)


def test_serving_respects_hbm_budget(gpu_ctx):
    from hyteg_amd.serving import PoissonServer, hierarchy_bytes, poisson_stream
    reqs = poisson_stream(10, rate=2000.0, seed=9, ns=(1, 2), levels=(3, 4), tol=1e-6)
    need = max(hierarchy_bytes(r.n, r.level) for r in reqs)
    say(f"[serving budget] budget {int(1.5 * need)} B")
    srv = PoissonServer(slots=3, hbm_budget=int(1.5 * need))
    res = srv.serve(reqs, realtime=False)
    srv.close()
    assert not srv.errors, srv.errors
    assert all(r is not None and r.rel_residual <= 1e-6 for r in res)
    assert srv.peak_resident <= srv.budget
