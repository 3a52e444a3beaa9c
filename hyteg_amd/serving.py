"""Serving loop for Poisson solves on one GPU: continuous batching, resident hierarchies, overlapping lanes.

A request asks for the solution of -Laplace u = f with zero Dirichlet boundary on the unit cube meshed as
n x n x n sub-cubes x 6 tetrahedra, refined to `level`, with f a sum of Fourier modes, to a relative
residual `tol`.  Each solve is b = M I_h f on the device (mass operator applied to the interpolated modes)
followed by FP32 V(2,2) cycles inside FP64 iterative refinement, the V-cycles replayed from captured hipGraphs.

Arrival never waits: `submit` only queues a request.  Two kinds of lanes take work from the queues, each lane a
thread with its own context (its own HIP stream), so the work of different lanes overlaps on the device:

* Batch lanes, one per (n, level) whose problem is small (at most `batch_max_unknowns` unknowns): the lane holds
  a hierarchy of interleaved functions with 32 right-hand sides.  Between two refinement steps it admits
  queued requests of its shape into free columns (zeroing the column of x and setting its right-hand side),
  then runs one refinement step for all columns, with the correction added only to the columns that have not
  converged, and retires the columns that reached their tolerance (or the step limit).  A request joins a
  batch already in flight at the next step boundary and leaves as soon as it has converged.
* Solve lanes (`slots` of them) for the large problems: one request at a time, FIFO, admitted when the HBM
  budget holds its hierarchy (least recently used idle hierarchies are released first).

Every column of an interleaved function is computed independently of the others, and every kernel reduces in
a fixed order, so a request's solution depends only on the request: it is the same, bit for bit, whether it
was served alone or in a mix (a batch lane with other columns active, other lanes running beside it).  The
outer loop restates the reference's cycle-and-check driver (apps/MultigridStudies/MultigridStudies.cpp:616-690)
per column: residual, stop below tolerance, else a correction.
"""
import collections
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from . import (ALL, F32, F64, INNER, CGSolver, Context, GeometricMultigridSolver, MeshInfo,
               MixedPrecisionRefinementSolver, P1ConstantLaplaceOperator, P1ConstantMassOperator, P1Function,
               P1toP1LinearProlongation, P1toP1LinearRestriction, PrimitiveStorage, SORSmoother)

BATCH = 32            # columns of a batch lane
MAX_OUTER = 60        # refinement steps per request at most


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
    slot: int                      # lane index (batch lanes after the solve lanes)
    batched: bool = False
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


def _level_values(n, level):
    nv, ne, nf, nc = cuboid_counts(n)
    total = 0
    for L in range(0, level + 1):
        w = 2 ** L + 1
        total += nv + ne * _round_up(w, 16) + nf * _round_up(w * (w + 1) // 2, 32) + nc * _round_up(
            w * (w + 1) * (w + 2) // 6, 64)
    return total


def hierarchy_bytes(n, level):
    """HBM held by one resident single-solve hierarchy: FP64 x, b and refinement residual, FP32 refinement
    right-hand side, correction and multigrid scratch, over levels 0..level (padded primitive strides), plus
    small tables."""
    return _level_values(n, level) * (3 * 8 + 3 * 4) + (8 << 20)


def batch_hierarchy_bytes(n, level):
    """The same for a batch lane: 32 interleaved columns, and one more FP64 function for I_h f."""
    return _level_values(n, level) * BATCH * (4 * 8 + 3 * 4) + (16 << 20)


def _solver(S, level, cycles_per_outer):
    inner = GeometricMultigridSolver(S, SORSmoother(1.0), CGSolver(S, 0, 0), P1toP1LinearRestriction(),
                                     P1toP1LinearProlongation(), 0, level, 2, 2)
    return MixedPrecisionRefinementSolver(S, inner, tolerance=1e-8, maxOuter=MAX_OUTER,
                                          cyclesPerOuter=cycles_per_outer, innerPrecision=F32)


class _Reserved:
    """Budget held for a hierarchy that its lane has not built yet."""

    def __init__(self, nbytes):
        self.bytes, self.last_used = nbytes, 0.0


class _Hierarchy:
    """Everything one (n, level) single solve needs, resident on one lane's context."""

    def __init__(self, ctx, n, level, cycles_per_outer):
        self.n, self.level = n, level
        self.S = PrimitiveStorage(MeshInfo.meshCuboid((0, 0, 0), (1, 1, 1), n, n, n), ctx)
        self.A = P1ConstantLaplaceOperator(self.S, 0, level)
        self.M = P1ConstantMassOperator(self.S, 0, level)
        self.x = P1Function("x", self.S, 0, level, precision=F64)
        self.b = P1Function("b", self.S, 0, level, precision=F64)
        self.R = _solver(self.S, level, cycles_per_outer)
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


class _BatchLane:
    """A resident 32-column hierarchy of one (n, level) and the state of its columns."""

    def __init__(self, ctx, n, level, cycles_per_outer, max_modes):
        self.n, self.level, self.cycles = n, level, cycles_per_outer
        self.S = PrimitiveStorage(MeshInfo.meshCuboid((0, 0, 0), (1, 1, 1), n, n, n), ctx)
        L = level
        self.A = P1ConstantLaplaceOperator(self.S, 0, L)
        self.M = P1ConstantMassOperator(self.S, 0, L)
        self.x = P1Function("x", self.S, 0, L, precision=F64, nrhs=BATCH)
        self.b = P1Function("b", self.S, 0, L, precision=F64, nrhs=BATCH)
        self.f = P1Function("f", self.S, 0, L, precision=F64, nrhs=BATCH)
        self.R = _solver(self.S, L, cycles_per_outer)
        self.modes = np.zeros((BATCH, max_modes, 4))
        self.modes[:, :, 1:] = 1.0                            # zero amplitude: an empty column holds f = 0
        self.cols = [None] * BATCH                            # column -> [req, admitted, start, r0, steps]
        self.bytes = batch_hierarchy_bytes(n, level)

    def free_columns(self):
        return [c for c in range(BATCH) if self.cols[c] is None]

    def active(self):
        return [c for c in range(BATCH) if self.cols[c] is not None]

    def admit(self, entries, now):
        """Put (req, admitted) entries into free columns: their x columns are zeroed, f and b = M I_h f
        rebuilt for all columns (the other columns' values do not change), and the residual norms of every
        column computed; a new column's initial norm is its r0."""
        L = self.level
        free = self.free_columns()
        keep = self.active()
        placed = []
        for (req, admitted), c in zip(entries, free):
            m = req.modes
            if len(m) > self.modes.shape[1]:
                raise ValueError(f"request {req.rid}: more than {self.modes.shape[1]} modes")
            self.modes[c, :, 0] = 0.0
            self.modes[c, :, 1:] = 1.0
            self.modes[c, :len(m)] = m
            self.cols[c] = [req, admitted, now(), 0.0, 0]
            placed.append(c)
        self.x.select_columns(L, keep)
        self.f.interpolate_modes(self.modes, L, ALL)
        self.M.apply(self.f, self.b, L, ALL)
        norms = self.R.step(self.A, self.x, self.b, L, cycles=0)
        for c in placed:
            self.cols[c][3] = norms[c]
        return norms

    def done(self, c, norms):
        req, _, _, r0, steps = self.cols[c]
        return norms[c] <= req.tol * r0 or steps >= MAX_OUTER

    def step(self, norms):
        """One refinement step on the columns that have not converged; returns the new norms."""
        go = [c for c in self.active() if not self.done(c, norms)]
        if not go:
            return norms
        out = self.R.step(self.A, self.x, self.b, self.level, cycles=self.cycles, columns=go)
        for c in go:
            self.cols[c][4] += 1
        return out

    def retire(self, c, norms, keep_solution):
        """The result of column c (which has converged) and its column freed."""
        req, admitted, start, r0, steps = self.cols[c]
        sol = self.x.download_column(self.level, c)
        self.cols[c] = None
        self.modes[c, :, 0] = 0.0
        rel = float(norms[c] / r0) if r0 > 0 else 0.0
        owned = self.S.owned_types(self.level) == INNER
        norm = float(np.sqrt(np.sum(sol[owned] ** 2)))
        return req, admitted, start, steps, rel, norm, (sol if keep_solution else None)


class PoissonServer:
    """Continuous-batching server on one GPU: batch lanes for small problems (one per (n, level)), `slots`
    solve lanes for large ones, an HBM budget (bytes) for the resident hierarchies."""

    def __init__(self, slots=3, hbm_budget=64 << 30, device=0, cycles_per_outer=2, keep_solutions=False,
                 batch_max_unknowns=2_200_000, max_modes=8, _factories=None):
        self.slots, self.budget, self.device = slots, int(hbm_budget), device
        # (context, single-solve hierarchy, batch lane) constructors; replaced only by host-side tests of the
        # scheduling logic
        self._make_ctx, self._make_hierarchy, self._make_lane = _factories or (Context, _Hierarchy, _BatchLane)
        self.cycles_per_outer, self.keep = cycles_per_outer, keep_solutions
        self.batch_max_unknowns, self.max_modes = batch_max_unknowns, max_modes
        self._lock = threading.Condition()
        self._pending_batch = collections.defaultdict(collections.deque)   # (n, level) -> deque of entries
        self._pending_single = collections.deque()
        self._resident = [dict() for _ in range(slots)]    # solve lane -> {(n, level): _Hierarchy}
        self._busy = [False] * slots
        self._lanes = {}                                   # (n, level) -> thread of its batch lane
        self._batch_bytes = {}                             # (n, level) -> bytes of a built or reserved lane
        self._batch_busy = set()
        self._results = {}
        self._outstanding = 0
        self._stop = False
        self._threads = []
        self.peak_resident = 0
        self.errors = []
        self._t0 = time.perf_counter()

    def batched(self, n, level):
        return unknowns(n, level) <= self.batch_max_unknowns

    # -- memory accounting (called with the lock held) --
    def _resident_bytes(self):
        return sum(h.bytes for r in self._resident for h in r.values()) + sum(self._batch_bytes.values())

    def _now(self):
        return time.perf_counter() - self._t0

    # -- arrival (never blocks on the GPU) --
    def submit(self, req):
        """Queue req (its arrival is now); returns at once."""
        with self._lock:
            entry = (req, self._now())
            self._outstanding += 1
            key = (req.n, req.level)
            if self.batched(req.n, req.level) and key not in self._lanes:
                # a new batch lane if the budget holds it; otherwise the request is solved on its own
                need = batch_hierarchy_bytes(*key)
                if self._resident_bytes() + need <= self.budget:
                    self._batch_bytes[key] = need
                    self.peak_resident = max(self.peak_resident, self._resident_bytes())
                    t = threading.Thread(target=self._batch_lane, args=(key, len(self._lanes)), daemon=True)
                    self._lanes[key] = t
                    t.start()
            if key in self._lanes:
                self._pending_batch[key].append(entry)
            else:
                self._pending_single.append(entry)
            self._lock.notify_all()

    def _record(self, req, admitted, start, end, steps, rel, norm, lane, batched, sol):
        with self._lock:
            self._results[req.rid] = PoissonResult(req.rid, req.n, req.level, unknowns(req.n, req.level),
                                                   req.arrival, admitted, start, end, steps, rel, norm, lane,
                                                   batched, sol)
            self._outstanding -= 1
            self._lock.notify_all()

    def _fail(self, req, exc):
        with self._lock:
            self.errors.append((req.rid, repr(exc)))
            self._outstanding -= 1
            self._lock.notify_all()

    # -- batch lanes --
    def _batch_lane(self, key, index):
        n, level = key
        lane_id = self.slots + index
        ctx = self._make_ctx(self.device)
        lane = None
        while True:
            with self._lock:
                while not self._stop and not self._pending_batch[key] and (lane is None or not lane.active()):
                    self._lock.wait()
                if self._stop:
                    return
                free = BATCH if lane is None else len(lane.free_columns())
                take = [self._pending_batch[key].popleft() for _ in range(min(free, len(self._pending_batch[key])))]
                self._batch_busy.add(key)
            try:
                if lane is None:
                    lane = self._make_lane(ctx, n, level, self.cycles_per_outer, self.max_modes)
                if take:
                    norms = lane.admit(take, self._now)
                norms = lane.step(norms)
                for c in lane.active():
                    if lane.done(c, norms):
                        req, admitted, start, steps, rel, norm, sol = lane.retire(c, norms, self.keep)
                        self._record(req, admitted, start, self._now(), steps, rel, norm, lane_id, True, sol)
            except Exception as exc:                       # reported; the lane drops its columns and goes on
                for entry in take:
                    self._fail(entry[0], exc)
                if lane is not None:
                    for c in lane.active():
                        self._fail(lane.cols[c][0], exc)
                        lane.cols[c] = None
            finally:
                with self._lock:
                    self._batch_busy.discard(key)
                    self._lock.notify_all()

    # -- solve lanes --
    def _admit_single(self, req):
        """With the lock held: a free solve lane for req (releasing least recently used hierarchies of idle
        lanes if the budget needs it, reserving req's bytes), or None if req must wait."""
        key = (req.n, req.level)
        need = hierarchy_bytes(req.n, req.level)
        free = [s for s in range(self.slots) if not self._busy[s]]
        if not free:
            return None
        for s in free:                                     # a free lane that already holds the hierarchy
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
        for s, k in drop:                                  # idle lanes: nothing of theirs is in flight
            del self._resident[s][k]
        target = min(free, key=lambda s: sum(h.bytes for h in self._resident[s].values()))
        self._resident[target][key] = _Reserved(need)
        self.peak_resident = max(self.peak_resident, self._resident_bytes())
        return target

    def _dispatch(self):
        """Hands queued large requests to free solve lanes, in arrival order, as lanes and budget free up."""
        while True:
            with self._lock:
                while True:
                    if self._stop:
                        return
                    if self._pending_single:
                        req, _ = self._pending_single[0]
                        try:
                            s = self._admit_single(req)
                        except MemoryError as exc:
                            self._pending_single.popleft()
                            self.errors.append((req.rid, repr(exc)))
                            self._outstanding -= 1
                            self._lock.notify_all()
                            continue
                        if s is not None:
                            entry = self._pending_single.popleft()
                            self._busy[s] = True
                            self._lane_work[s].append(entry)
                            self._lock.notify_all()
                            continue
                    self._lock.wait()

    def _solve_lane(self, s):
        ctx = self._make_ctx(self.device)
        while True:
            with self._lock:
                while not self._stop and not self._lane_work[s]:
                    self._lock.wait()
                if self._stop:
                    return
                req, admitted = self._lane_work[s].popleft()
            key = (req.n, req.level)
            try:
                start = self._now()
                with self._lock:
                    h = self._resident[s].get(key)
                if h is None or isinstance(h, _Reserved):
                    h = self._make_hierarchy(ctx, req.n, req.level, self.cycles_per_outer)
                    with self._lock:
                        self._resident[s][key] = h
                steps, rel, norm, sol = h.solve(req, self.keep)
                end = self._now()
                h.last_used = end
                self._record(req, admitted, start, end, steps, rel, norm, s, False, sol)
            except Exception as exc:                       # reported, the server keeps going
                with self._lock:
                    if isinstance(self._resident[s].get(key), _Reserved):
                        del self._resident[s][key]
                self._fail(req, exc)
            with self._lock:
                self._busy[s] = False
                self._lock.notify_all()

    def _start(self):
        if self._threads:
            return
        self._lane_work = [collections.deque() for _ in range(self.slots)]
        self._threads = [threading.Thread(target=self._solve_lane, args=(s,), daemon=True) for s in range(self.slots)]
        self._threads.append(threading.Thread(target=self._dispatch, daemon=True))
        for t in self._threads:
            t.start()

    def _drain(self):
        with self._lock:
            while self._outstanding > 0:
                self._lock.wait()

    def warm(self, shapes):
        """Build hierarchies and capture graphs for (n, level) shapes ahead of a stream: one small solve per
        batched shape (its lane), and for a large shape one per solve lane, submitted together so that every
        lane builds one."""
        self._start()
        for i, (n, level) in enumerate(shapes):
            copies = 1 if self.batched(n, level) else self.slots
            for k in range(copies):
                self.submit(PoissonRequest(-1 - i * self.slots - k, 0.0, n, level, np.array([[1.0, 1, 1, 1]]), 1e-6))
            self._drain()
        with self._lock:
            self._results = {}

    def serve(self, requests, realtime=True, one_at_a_time=False):
        """Run the stream; returns the results in request order, times in seconds from the stream start.
        realtime=False submits every request at once (arrival times are then only labels); one_at_a_time
        submits each request after the previous one has finished (every request served alone)."""
        self._start()
        with self._lock:
            self._results = {}
            self._t0 = time.perf_counter()
        for req in sorted(requests, key=lambda r: r.arrival):
            if realtime and not one_at_a_time:
                delay = req.arrival - self._now()
                if delay > 0:
                    time.sleep(delay)
            self.submit(req)
            if one_at_a_time:
                self._drain()
        self._drain()
        self.wall = self._now()
        out = []
        for req in requests:
            r = self._results.get(req.rid)
            if r is not None and not realtime:
                r.arrival = 0.0
            out.append(r)
        return out

    def close(self):
        with self._lock:
            self._stop = True
            self._lock.notify_all()
        for t in self._threads + list(self._lanes.values()):
            t.join()
        self._threads = []
        self._lanes = {}
        with self._lock:
            for r in self._resident:
                r.clear()
            self._batch_bytes.clear()


def summarize(results, wall):
    """p50 / p99 latency, throughput and solve statistics of a served stream."""
    done = [r for r in results if r is not None]
    lat = np.array([r.latency for r in done])
    dofs = sum(r.unknowns for r in done)
    return {
        "requests": len(results),
        "completed": len(done),
        "batched": sum(1 for r in done if r.batched),
        "p50_latency_ms": float(np.percentile(lat, 50) * 1e3) if len(lat) else None,
        "p99_latency_ms": float(np.percentile(lat, 99) * 1e3) if len(lat) else None,
        "solves_per_s": len(done) / wall if wall > 0 else None,
        "unknowns_per_s": dofs / wall if wall > 0 else None,
        "max_rel_residual": float(max(r.rel_residual for r in done)) if done else None,
        "mean_outer_steps": float(np.mean([r.outer_steps for r in done])) if done else None,
    }
