"""Serving loop for Poisson solves on one GPU: continuous admission, resident hierarchies, overlapping slots.

A request asks for the solution of -Laplace u = f with zero Dirichlet boundary on the unit cube meshed as
n x n x n sub-cubes x 6 tetrahedra, refined to `level`, with f a sum of Fourier modes, to a relative
residual `tol`.  The server runs `slots` workers.  Each worker owns a context (its own HIP stream), so the
solves of different workers overlap on the device.  A request is admitted when it has arrived, a worker
is free and the HBM budget holds its hierarchy.  Hierarchies (mesh plans, operator, multigrid work and
the x / b functions) stay resident per worker and (n, level), and are reused by later requests.  When the
budget is short, the least recently used hierarchies of idle workers are released first.  Each solve is
b = M I_h f on the device (mass operator applied to the interpolated modes) followed by FP32 V(2,2)
cycles inside FP64 iterative refinement, with the V-cycles replayed from captured hipGraphs.

A request's result depends only on the request: every kernel reduces in a fixed order and every worker
runs the same launch sequence, so a solve in the mix equals the same solve alone bit for bit.
"""
import queue
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from . import (ALL, F32, F64, INNER, CGSolver, Context, GeometricMultigridSolver, MeshInfo,
               MixedPrecisionRefinementSolver, P1ConstantLaplaceOperator, P1ConstantMassOperator, P1Function,
               P1toP1LinearProlongation, P1toP1LinearRestriction, PrimitiveStorage, SORSmoother)


@dataclass
class PoissonRequest:
    rid: int
    arrival: float                 # seconds after the start of the stream
    n: int                         # sub-cubes per side
    level: int
    modes: np.ndarray              # (nModes, 4) rows of (amplitude, k, l, m)
    tol: float = 1e-8


@dataclass
class PoissonResult:
    rid: int
    n: int
    level: int
    unknowns: int
    arrival: float
    admitted: float
    start: float
    end: float
    outer_steps: int
    rel_residual: float
    solution_norm: float
    slot: int
    solution: np.ndarray = field(default=None, repr=False)

    @property
    def latency(self):
        return self.end - self.arrival


def poisson_stream(count, rate=200.0, seed=20251202, ns=(1, 2, 3), levels=(4, 5, 6, 7), n_modes=4, tol=1e-8):
    """A seeded Poisson-process stream: exponential inter-arrival times with mean 1/rate, n and level drawn
    uniformly, f = sum of n_modes modes with amplitudes N(0, 1) and wave numbers k, l, m in 1..8."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    for rid in range(count):
        t += rng.exponential(1.0 / rate)
        modes = np.zeros((n_modes, 4))
        modes[:, 0] = rng.standard_normal(n_modes)
        modes[:, 1:] = rng.integers(1, 9, (n_modes, 3))
        out.append(PoissonRequest(rid, t, int(rng.choice(ns)), int(rng.choice(levels)), modes, tol))
    return out


def unknowns(n, level):
    return (n * 2 ** level - 1) ** 3


def _round_up(v, m):
    return (v + m - 1) // m * m


def cuboid_counts(n):
    """(vertices, edges, faces, cells) of the n x n x n x 6 cuboid mesh: C = 6n^3, 12n^2 boundary triangles,
    F = (4C + 12n^2) / 2, and E from Euler's formula V - E + F - C = 1."""
    nv, nc = (n + 1) ** 3, 6 * n ** 3
    nf = (4 * nc + 12 * n ** 2) // 2
    return nv, nv + nf - nc - 1, nf, nc


def hierarchy_bytes(n, level):
    """HBM held by one resident hierarchy: FP64 x, b and refinement residual, FP32 refinement right-hand side,
    correction and multigrid scratch, over levels 0..level (padded primitive strides), plus small tables."""
    nv, ne, nf, nc = cuboid_counts(n)
    total = 0
    for L in range(0, level + 1):
        w = 2 ** L + 1
        total += nv + ne * _round_up(w, 16) + nf * _round_up(w * (w + 1) // 2, 32) + nc * _round_up(
            w * (w + 1) * (w + 2) // 6, 64)
    return total * (3 * 8 + 3 * 4) + (8 << 20)


class _Reserved:
    """Budget held for a hierarchy that its worker has not built yet."""

    def __init__(self, nbytes):
        self.bytes, self.last_used = nbytes, 0.0


class _Hierarchy:
    """Everything one (n, level) solve needs, resident on one worker's context."""

    def __init__(self, ctx, n, level, cycles_per_outer):
        self.n, self.level = n, level
        self.S = PrimitiveStorage(MeshInfo.meshCuboid((0, 0, 0), (1, 1, 1), n, n, n), ctx)
        self.A = P1ConstantLaplaceOperator(self.S, 0, level)
        self.M = P1ConstantMassOperator(self.S, 0, level)
        self.x = P1Function("x", self.S, 0, level, precision=F64)
        self.b = P1Function("b", self.S, 0, level, precision=F64)
        inner = GeometricMultigridSolver(self.S, SORSmoother(1.0), CGSolver(self.S, 0, 0), P1toP1LinearRestriction(),
                                         P1toP1LinearProlongation(), 0, level, 2, 2)
        self.R = MixedPrecisionRefinementSolver(self.S, inner, tolerance=1e-8, maxOuter=60,
                                                cyclesPerOuter=cycles_per_outer, innerPrecision=F32)
        self.bytes = hierarchy_bytes(n, level)
        self.last_used = 0.0

    def solve(self, req, keep_solution):
        L = self.level
        self.x.interpolate_modes(req.modes, L, ALL)         # I_h f (x used as scratch)
        self.M.apply(self.x, self.b, L, ALL)                  # b = M I_h f
        self.x.interpolate(0.0, L, ALL)
        self.R.tolerance = req.tol
        rel = self.R.solve(self.A, self.x, self.b, L)
        norm = float(np.sqrt(self.x.dotGlobal(self.x, L, INNER)))
        sol = self.x.download(L) if keep_solution else None
        return self.R.outerSteps, float(np.max(rel)), norm, sol


class PoissonServer:
    """Continuous-admission server over `slots` workers sharing one GPU and an HBM budget (bytes)."""

    def __init__(self, slots=4, hbm_budget=64 << 30, device=0, cycles_per_outer=2, keep_solutions=False):
        self.slots, self.budget, self.device = slots, int(hbm_budget), device
        self.cycles_per_outer, self.keep = cycles_per_outer, keep_solutions
        self._lock = threading.Condition()
        self._resident = [dict() for _ in range(slots)]    # slot -> {(n, level): _Hierarchy}
        self._busy = [False] * slots
        self._queues = [queue.Queue() for _ in range(slots)]
        self._ctx = [None] * slots
        self.peak_resident = 0
        self.errors = []

    # -- memory accounting (called with the lock held) --
    def _resident_bytes(self):
        return sum(h.bytes for r in self._resident for h in r.values())

    def _admit(self, req):
        """With the lock held: pick a free worker for req, release least recently used hierarchies of idle
        workers if the budget needs it, and reserve req's hierarchy bytes.  Returns the worker index, or None
        if req must wait for a busy worker to finish."""
        key = (req.n, req.level)
        need = hierarchy_bytes(req.n, req.level)
        free = [s for s in range(self.slots) if not self._busy[s]]
        if not free:
            return None
        for s in free:                                     # a free worker that already holds the hierarchy
            if key in self._resident[s]:
                return s
        if need > self.budget:
            raise MemoryError(f"request {req.rid} needs {need} B, budget {self.budget} B")
        have = self.budget - self._resident_bytes()
        victims = sorted((h.last_used, s, k) for s in free for k, h in self._resident[s].items())
        drop = []
        for _, s, k in victims:
            if have >= need:
                break
            drop.append((s, k))
            have += self._resident[s][k].bytes
        if have < need:
            return None
        for s, k in drop:                                  # idle workers: nothing of theirs is in flight
            del self._resident[s][k]
        target = min(free, key=lambda s: sum(h.bytes for h in self._resident[s].values()))
        self._resident[target][key] = _Reserved(need)
        self.peak_resident = max(self.peak_resident, self._resident_bytes())
        return target

    # -- workers --
    def _worker(self, s):
        self._ctx[s] = Context(self.device)
        while True:
            item = self._queues[s].get()
            if item is None:
                break
            req, admitted = item
            key = (req.n, req.level)
            try:
                start = time.perf_counter()
                with self._lock:
                    h = self._resident[s].get(key)
                if not isinstance(h, _Hierarchy):
                    h = _Hierarchy(self._ctx[s], req.n, req.level, self.cycles_per_outer)
                    with self._lock:
                        self._resident[s][key] = h
                steps, rel, norm, sol = h.solve(req, self.keep)
                end = time.perf_counter()
                self._results[req.rid] = PoissonResult(req.rid, req.n, req.level, unknowns(req.n, req.level),
                                                       req.arrival, admitted, start, end, steps, rel, norm, s, sol)
                h.last_used = end
            except Exception as exc:                       # reported, the server keeps going
                self.errors.append((req.rid, repr(exc)))
                with self._lock:
                    if isinstance(self._resident[s].get(key), _Reserved):
                        del self._resident[s][key]
            with self._lock:
                self._busy[s] = False
                self._lock.notify_all()

    def warm(self, shapes):
        """Build hierarchies for (n, level) shapes ahead of the stream (round robin over the workers)."""
        self._start_workers()
        self._results = {}
        for i, (n, level) in enumerate(shapes):
            req = PoissonRequest(-1 - i, 0.0, n, level, np.array([[1.0, 1, 1, 1]]), 1e-6)
            with self._lock:
                s = i % self.slots
                while self._busy[s]:
                    self._lock.wait()
                self._busy[s] = True
                if (n, level) not in self._resident[s]:
                    self._resident[s][(n, level)] = _Reserved(hierarchy_bytes(n, level))
            self._queues[s].put((req, 0.0))
        with self._lock:
            while any(self._busy):
                self._lock.wait()

    def _start_workers(self):
        if getattr(self, "_threads", None):
            return
        self._results = {}
        self._threads = [threading.Thread(target=self._worker, args=(s,), daemon=True) for s in range(self.slots)]
        for t in self._threads:
            t.start()

    def serve(self, requests, realtime=True):
        """Run the stream; returns the results in request order, times in seconds from the stream start.
        realtime=False admits every request at once (arrival times are then only labels)."""
        self._start_workers()
        self._results = {}
        t0 = time.perf_counter()
        for req in sorted(requests, key=lambda r: r.arrival):
            if realtime:
                delay = req.arrival - (time.perf_counter() - t0)
                if delay > 0:
                    time.sleep(delay)
            with self._lock:
                while True:
                    s = self._admit(req)
                    if s is not None:
                        break
                    self._lock.wait()
                self._busy[s] = True
            self._queues[s].put((req, time.perf_counter() - t0))
        with self._lock:
            while any(self._busy):
                self._lock.wait()
        self.wall = time.perf_counter() - t0
        out = []
        for req in requests:
            r = self._results.get(req.rid)
            if r is not None:
                r.start -= t0
                r.end -= t0
                if not realtime:
                    r.arrival = 0.0
            out.append(r)
        return out

    def close(self):
        for q in self._queues:
            q.put(None)
        for t in getattr(self, "_threads", []):
            t.join()
        self._threads = []
        with self._lock:
            for r in self._resident:
                r.clear()


def summarize(results, wall):
    """p50 / p99 latency, throughput and solve statistics of a served stream."""
    done = [r for r in results if r is not None]
    lat = np.array([r.latency for r in done])
    dofs = sum(r.unknowns for r in done)
    return {
        "requests": len(results),
        "completed": len(done),
        "p50_latency_ms": float(np.percentile(lat, 50) * 1e3) if len(lat) else None,
        "p99_latency_ms": float(np.percentile(lat, 99) * 1e3) if len(lat) else None,
        "solves_per_s": len(done) / wall if wall > 0 else None,
        "unknowns_per_s": dofs / wall if wall > 0 else None,
        "max_rel_residual": float(max(r.rel_residual for r in done)) if done else None,
        "mean_outer_steps": float(np.mean([r.outer_steps for r in done])) if done else None,
    }
