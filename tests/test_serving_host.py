"""Host-side tests of the serving loop's scheduling (hyteg_amd.serving.PoissonServer) with stand-in lanes that
compute nothing on a device: every request completes exactly once, small problems go to batch lanes and join
batches in flight, arrival never waits for the device, the HBM budget is respected, and a request's result does
not depend on what else was in flight."""
import threading
import time

import numpy as np

from hyteg_amd import serving as SV


def _steps_needed(req):
    # a deterministic "solve": the number of refinement steps a request needs follows from its data
    return 1 + int(abs(req.modes[0, 0]) * 7) % 5


class FakeCtx:
    def __init__(self, device):
        self.device = device


class FakeHierarchy:
    built = 0

    def __init__(self, ctx, n, level, cycles):
        FakeHierarchy.built += 1
        self.n, self.level, self.bytes, self.last_used = n, level, SV.hierarchy_bytes(n, level), 0.0

    def solve(self, req, keep):
        time.sleep(0.002)
        k = _steps_needed(req)
        return k, 1e-9, float(k), (np.full(4, float(req.rid)) if keep else None)


class FakeLane:
    """The _BatchLane interface: columns, admission, one step for the unconverged columns, retirement."""
    steps_run = []

    def __init__(self, ctx, n, level, cycles, max_modes):
        self.n, self.level = n, level
        self.cols = [None] * SV.BATCH
        self.bytes = SV.batch_hierarchy_bytes(n, level)

    def free_columns(self):
        return [c for c in range(SV.BATCH) if self.cols[c] is None]

    def active(self):
        return [c for c in range(SV.BATCH) if self.cols[c] is not None]

    def admit(self, entries, now):
        for (req, admitted), c in zip(entries, self.free_columns()):
            self.cols[c] = [req, admitted, now(), 1.0, 0]
        return self._norms()

    def _norms(self):
        out = np.zeros(SV.BATCH)
        for c in self.active():
            req, _, _, _, steps = self.cols[c]
            out[c] = 0.0 if steps >= _steps_needed(req) else 1.0
        return out

    def done(self, c, norms):
        req, _, _, r0, steps = self.cols[c]
        return norms[c] <= req.tol * r0 or steps >= SV.MAX_OUTER

    def step(self, norms):
        go = [c for c in self.active() if not self.done(c, norms)]
        if go:
            time.sleep(0.001)
            FakeLane.steps_run.append(len(go))
            for c in go:
                self.cols[c][4] += 1
        return self._norms()

    def retire(self, c, norms, keep):
        req, admitted, start, r0, steps = self.cols[c]
        self.cols[c] = None
        return req, admitted, start, steps, 1e-9, float(steps), (np.full(4, float(req.rid)) if keep else None)


def _server(**kw):
    return SV.PoissonServer(_factories=(FakeCtx, FakeHierarchy, FakeLane), keep_solutions=True, **kw)


def test_every_request_completes_once_and_small_ones_are_batched():
    reqs = SV.poisson_stream(120, rate=3000.0, seed=3, ns=(1, 2, 3), levels=(3, 4, 5, 6, 7), tol=1e-8)
    srv = _server(slots=3)
    FakeLane.steps_run = []
    res = srv.serve(reqs, realtime=True)
    lanes = len(srv._lanes)
    srv.close()
    assert not srv.errors
    assert all(r is not None for r in res)
    for q, r in zip(reqs, res):
        assert r.rid == q.rid and r.outer_steps == _steps_needed(q)
        assert r.batched == srv.batched(q.n, q.level)
        assert r.end >= r.start >= r.admitted >= 0.0
    assert any(k > 1 for k in FakeLane.steps_run)        # several columns stepped together
    assert lanes == len({(q.n, q.level) for q in reqs if srv.batched(q.n, q.level)})


def test_alone_and_mix_give_the_same_results():
    reqs = SV.poisson_stream(40, rate=5000.0, seed=8, ns=(1, 2, 3), levels=(3, 4, 5, 6), tol=1e-8)
    alone = _server(slots=2)
    ra = alone.serve(reqs, one_at_a_time=True)
    alone.close()
    mix = _server(slots=2)
    rm = mix.serve(reqs, realtime=True)
    mix.close()
    for a, m in zip(ra, rm):
        assert a.outer_steps == m.outer_steps and np.array_equal(a.solution, m.solution)


def test_submit_does_not_wait_for_lanes():
    srv = _server(slots=1)
    srv._start()
    block = threading.Event()
    orig = FakeHierarchy.solve

    def slow(self, req, keep):
        block.wait(5.0)
        return orig(self, req, keep)

    FakeHierarchy.solve = slow
    try:
        t0 = time.perf_counter()
        for q in SV.poisson_stream(20, rate=1e6, seed=1, ns=(3,), levels=(7,)):
            srv.submit(q)
        assert time.perf_counter() - t0 < 1.0               # 20 large requests queued behind one busy lane
    finally:
        block.set()
        FakeHierarchy.solve = orig
    srv._drain()
    srv.close()
    assert not srv.errors


def test_budget_holds_and_oversized_requests_fail_alone():
    reqs = SV.poisson_stream(30, rate=1e5, seed=4, ns=(2, 3), levels=(6, 7), tol=1e-8)
    big = max(SV.hierarchy_bytes(q.n, q.level) for q in reqs)
    srv = _server(slots=3, hbm_budget=int(1.2 * big), batch_max_unknowns=0)
    res = srv.serve(reqs, realtime=False)
    srv.close()
    assert not srv.errors and all(r is not None for r in res)
    assert srv.peak_resident <= srv.budget
    tiny = _server(slots=1, hbm_budget=1 << 20, batch_max_unknowns=0)
    tiny.serve(reqs[:2], realtime=False)
    tiny.close()
    assert len(tiny.errors) == 2
