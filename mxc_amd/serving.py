"""Serving: resident factorized problems, HBM-budget admission and
continuous batching of right-hand sides.

A Server keeps H^2 problems (operator + ULV-factorized preconditioner)
resident on one GPU, each on its own device context, so that the work of
different problems runs on different HIP streams and overlaps on the device.
Requests (problem id + one or more right-hand sides) queue per problem; a
fixed number of worker slots each take the problem with the oldest waiting
request that is not already on the GPU, gather every waiting right-hand side
of that problem (first come first served, up to max_rhs) and solve them in
one lockstep GMRES call (mxg_gmres_batch). Requests that arrive while a
problem is on the GPU join its next call, so a call carries whatever has
arrived by the time a slot frees up. Each right-hand side of a call gets bit
for bit the solution it would get alone: the batched block products sum
every row in the same order whatever the batch width, and each right-hand
side keeps its own Krylov basis and stopping test.

Admission: a problem is built only when the resident footprint plus its
estimated footprint fits the HBM budget; otherwise residents that no slot
holds are evicted (idle ones least recently used first, then those whose
waiting requests are all younger than the one being admitted), and the
request waits until memory frees up. Problems are taken up in the order of
their oldest request, but one that cannot be admitted yet does not hold up
resident problems behind it. The estimate is bytes per point of the problems
built so far (a fixed guess before the first; with nothing resident a
problem is built whatever the guess), and the footprint of a built problem is
measured as the drop of free device memory over its construction,
factorization and a first solve of max_rhs right-hand sides (which sizes its
Krylov workspace); builds run one at a time and solves allocate nothing, so
the drop is the problem's own.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import threading
import time
from collections import deque

import numpy as np

from . import _native as N
from .device import Context
from .solver import H2, Problem, generate_points


@dataclasses.dataclass
class ProblemSpec:
    n: int
    seed: int
    theta: float = 2.0
    leaf: int = 256
    kernel: str = "laplace"
    diag: float = 1e3
    alpha: float = 1.0
    geometry: str = "cube"
    op_eps: float = 1e-13       # operator: the residual of every GMRES cycle (1e-12 leaves up to 6e-11
                                # of construction error on cube solutions at N = 65536)
    precond_eps: float = 1e-6   # preconditioner truncation, with a rank cap
    precond_rank: int = 64
    sampling: int = 2000
    factor: int = N.MXG_BF16    # FP32 factorization, bfloat16-stored factors


class Request:
    def __init__(self, pid: int, B: np.ndarray):
        self.pid = pid
        self.B = np.asfortranarray(B.reshape(B.shape[0], -1), dtype=np.float64)
        self.X = None
        self.resid = None
        self.t_submit = time.perf_counter()
        self.t_start = None
        self.t_done = None
        self.error = None
        self.done = threading.Event()

    @property
    def nrhs(self) -> int:
        return self.B.shape[1]

    @property
    def latency(self) -> float:
        return self.t_done - self.t_submit

    def wait(self, timeout=None):
        if not self.done.wait(timeout):
            raise TimeoutError("request not served in time")
        if self.error:
            raise self.error
        return self.X


class Resident:
    """One problem built and factorized on its own context."""

    def __init__(self, pid: int, spec: ProblemSpec, device: int, max_rhs: int):
        self.pid, self.spec = pid, spec
        self.ctx = Context(device)
        free0, _ = self.ctx.mem_info()
        s = spec
        self.problem = Problem(self.ctx, generate_points(s.geometry, s.n, s.seed), s.leaf)
        self.A = H2.operator(self.problem, s.op_eps, s.sampling, s.kernel, s.diag, s.alpha, s.theta)
        self.M = H2.preconditioner(self.problem, s.precond_rank, s.sampling, s.kernel, s.diag, s.alpha, s.theta,
                                   eps=s.precond_eps)
        info = self.M.factorize(s.factor)
        if info.any():
            raise RuntimeError(f"problem {pid}: singular ULV blocks per level {info}")
        self.max_rhs = max_rhs
        self.dB = self.ctx.zeros((s.n, max_rhs))
        self.dX = self.ctx.zeros((s.n, max_rhs))
        self.solve(np.zeros((s.n, max_rhs)), 1e-10)  # sizes the Krylov workspace of the context
        free1, _ = self.ctx.mem_info()
        self.footprint = max(free0 - free1, 0)
        self.last_used = time.perf_counter()

    def solve(self, B: np.ndarray, tol: float, restart: int = 30, max_cycles: int = 20):
        """X for the columns of B (n x k, k <= max_rhs): FP64 GMRES (CGS2)
        with FP32 inner operator applications against the FP32 factors, in
        the problem's resident right-hand-side buffers (no allocation); the
        inner steps after a call's first replay one captured HIP graph."""
        n, k = B.shape
        assert n == self.spec.n and 1 <= k <= self.max_rhs
        B = np.asfortranarray(B, dtype=np.float64)
        G, H = N.gpu(), N.h2()
        N.check_gpu(G.mxg_memcpy_h2d(self.ctx.handle, self.dB.ptr, B.ctypes.data, B.nbytes))
        N.check_gpu(G.mxg_memset(self.ctx.handle, self.dX.ptr, 0, B.nbytes))
        o = N.MxgGmresOpts(tol, restart, max_cycles, N.GMRES_CGS2, N.MXG_FP32, None, 1)  # inner steps as a graph
        resid = np.zeros((k, max_cycles + 1))
        it = np.zeros(k, np.int64)
        inn = np.zeros(k, np.int64)
        N.check_h2(H.mxc_solver_gmres_batch(self.A.handle, self.M.handle, C.byref(o), k, self.dB.ptr, n, self.dX.ptr, n,
                                            resid.ctypes.data, it.ctypes.data, inn.ctypes.data))
        X = np.empty((n, k), order="F")
        N.check_gpu(G.mxg_memcpy_d2h(self.ctx.handle, X.ctypes.data, self.dX.ptr, X.nbytes))
        return X, [float(resid[q, it[q]]) for q in range(k)]

    def close(self):
        self.ctx.close()


class Server:
    def __init__(self, hbm_budget: int, slots: int = 4, max_rhs: int = 32, tol: float = 1e-10, device: int = 0,
                 bytes_per_point_guess: float = 400e3, resident_factory=None):
        self.budget = int(hbm_budget)
        self.max_rhs, self.tol, self.device = max_rhs, tol, device
        # GMRES stops at half the target on the H^2 residual: the operator's
        # construction error on a solution (about 1e-11 at eps 1e-12) takes the
        # rest of the dense-check budget
        self.h2_tol = 0.5 * tol
        self.specs = {}
        self.resident = {}
        self.queues = {}
        self.busy = set()
        self.building = set()
        self.cv = threading.Condition()
        self.stop = False
        self.bpp_guess = bytes_per_point_guess
        self.factory = resident_factory or Resident
        self.bpp_seen = 0.0
        self.evictions = 0
        self.calls = []  # (pid, nrhs, requests, t_start, t_end)
        self.workers = [threading.Thread(target=self._work, daemon=True) for _ in range(slots)]
        for w in self.workers:
            w.start()

    # ---------------------------------------------------------------- API
    def register(self, pid: int, spec: ProblemSpec) -> None:
        with self.cv:
            self.specs[pid] = spec
            self.queues[pid] = deque()

    def submit(self, pid: int, B: np.ndarray) -> Request:
        r = Request(pid, B)
        if r.nrhs > self.max_rhs:
            raise ValueError(f"a request carries at most {self.max_rhs} right-hand sides")
        with self.cv:
            if pid not in self.specs:
                raise KeyError(f"unknown problem {pid}")
            if r.B.shape[0] != self.specs[pid].n:
                raise ValueError("right-hand sides do not match the problem size")
            self.queues[pid].append(r)
            self.cv.notify_all()
        return r

    def resident_bytes(self) -> int:
        with self.cv:
            return sum(p.footprint for p in self.resident.values())

    def state(self) -> str:
        """One line: residents with their footprints, problems being built or
        served, and the waiting requests per problem."""
        with self.cv:
            res = {pid: round(p.footprint / 2**20) for pid, p in self.resident.items()}
            waiting = {pid: len(q) for pid, q in self.queues.items() if q}
            return (f"budget {self.budget / 2**20:.0f} MiB, resident MiB {res}, building {sorted(self.building)}, "
                    f"busy {sorted(self.busy)}, waiting {waiting}, estimate MiB/point "
                    f"{(self.bpp_seen or self.bpp_guess) / 2**20:.4f}, evictions {self.evictions}")

    def close(self):
        with self.cv:
            self.stop = True
            self.cv.notify_all()
        for w in self.workers:
            w.join()
        for p in self.resident.values():
            p.close()
        self.resident.clear()

    # ---------------------------------------------------------- internals
    def _estimate(self, pid: int) -> int:
        return int((self.bpp_seen or self.bpp_guess) * self.specs[pid].n)

    def _next(self):
        """The problem to take up now, and whether it must be built first:
        waiting problems in the order of their oldest request, the first that
        is resident (and no slot holds), or that is not resident and can be
        admitted (no other build running, and room made for it). A problem
        that cannot be admitted yet does not hold up the resident problems
        behind it, whose calls then free up memory."""
        order = sorted((q[0].t_submit, pid) for pid, q in self.queues.items()
                       if q and pid not in self.busy and pid not in self.building)
        for t, pid in order:
            if pid in self.resident:
                return pid, False
            if not self.building and self._make_room(self._estimate(pid), t):
                return pid, True
        return None, False

    def _make_room(self, need: int, t_oldest: float) -> bool:
        """Evict residents until `need` more bytes fit the budget, and say
        whether they fit. Candidates are the residents no slot holds that have
        no waiting request (least recently used first), then those whose
        waiting requests all arrived after t_oldest (latest first); nothing is
        evicted unless enough can be. With nothing resident the problem is
        admitted whatever the estimate, and one estimated above the whole budget
        once every other resident is evicted: its build measures its real
        footprint, and one that exceeds the budget is refused then."""
        if not self.resident:
            return True
        need = min(need, self.budget)  # an estimate above the budget: admitted alone, then measured
        used = sum(p.footprint for p in self.resident.values())
        if used + need <= self.budget:
            return True
        free = [p for pid, p in self.resident.items() if pid not in self.busy]
        idle = sorted((p for p in free if not self.queues[p.pid]), key=lambda p: p.last_used)
        later = sorted((p for p in free if self.queues[p.pid] and self.queues[p.pid][0].t_submit > t_oldest),
                       key=lambda p: -self.queues[p.pid][0].t_submit)
        victims, freed = [], 0
        for p in idle + later:
            if used - freed + need <= self.budget:
                break
            victims.append(p)
            freed += p.footprint
        if used - freed + need > self.budget:
            return False
        for p in victims:
            del self.resident[p.pid]
            p.close()
            self.evictions += 1
        return True

    def _work(self):
        while True:
            with self.cv:
                while True:
                    if self.stop:
                        return
                    pid, build = self._next()
                    if pid is not None:
                        break
                    self.cv.wait(0.05)
                if build:
                    self.building.add(pid)
                else:
                    self.busy.add(pid)
            if build:
                try:
                    res = self.factory(pid, self.specs[pid], self.device, self.max_rhs)
                    err = None
                except Exception as e:  # noqa: BLE001 - the waiting requests receive the error
                    res, err = None, e
                if res is not None and res.footprint > self.budget:
                    fp = res.footprint
                    res.close()
                    res, err = None, MemoryError(f"problem {pid} needs {fp} bytes of device memory, more than the "
                                                 f"budget of {self.budget}")
                with self.cv:
                    self.building.discard(pid)
                    if res is not None:
                        self.resident[pid] = res
                        self.bpp_seen = max(self.bpp_seen, res.footprint / self.specs[pid].n)
                    else:
                        for r in self.queues[pid]:
                            r.error = err
                            r.done.set()
                        self.queues[pid].clear()
                    self.cv.notify_all()
                continue
            with self.cv:
                q = self.queues[pid]
                take, cols = [], 0
                while q and cols + q[0].nrhs <= self.max_rhs:
                    r = q.popleft()
                    take.append(r)
                    cols += r.nrhs
                res = self.resident[pid]
            t0 = time.perf_counter()
            for r in take:
                r.t_start = t0
            try:
                X, rr = res.solve(np.concatenate([r.B for r in take], axis=1), self.h2_tol)
                o = 0
                for r in take:
                    r.X, r.resid = X[:, o:o + r.nrhs], rr[o:o + r.nrhs]
                    o += r.nrhs
            except Exception as e:  # noqa: BLE001
                for r in take:
                    r.error = e
            t1 = time.perf_counter()
            for r in take:
                r.t_done = t1
                r.done.set()
            with self.cv:
                res.last_used = t1
                self.calls.append((pid, cols, len(take), t0, t1))
                self.busy.discard(pid)
                self.cv.notify_all()


def poisson_load(server: Server, pids, sizes, rate: float, count: int, seed: int, max_rhs_per_request: int = 32,
                 timeout: float = 120.0):
    """Submits `count` requests at Poisson arrivals of `rate` per second: each
    names a problem uniformly from pids and carries 1..max_rhs_per_request
    right-hand sides uniform in [-1, 1]. Returns the requests (served); a
    request not served within `timeout` seconds of the last arrival raises
    TimeoutError with the server's state."""
    rng = np.random.default_rng(seed)
    reqs = []
    t = time.perf_counter()
    for _ in range(count):
        t += rng.exponential(1.0 / rate)
        k = int(rng.integers(0, len(pids)))
        pid = pids[k]
        nr = int(rng.integers(1, max_rhs_per_request + 1))
        B = rng.uniform(-1.0, 1.0, size=(sizes[k], nr))
        while time.perf_counter() < t:
            time.sleep(max(0.0, min(t - time.perf_counter(), 0.01)))
        reqs.append(server.submit(pid, B))
    for r in reqs:
        try:
            r.wait(timeout=timeout)
        except TimeoutError as e:
            raise TimeoutError(f"{e}: {server.state()}") from None
    return reqs


def latency_stats(reqs, calls, t0=None):
    lat = np.array([r.latency for r in reqs])
    t_first = min(r.t_submit for r in reqs) if t0 is None else t0
    t_last = max(r.t_done for r in reqs)
    busy = sum(c[4] - c[3] for c in calls)
    return {"requests": len(reqs), "rhs": int(sum(r.nrhs for r in reqs)),
            "p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "mean_ms": float(lat.mean() * 1e3), "wall_s": t_last - t_first,
            "solves_per_s": float(sum(r.nrhs for r in reqs) / (t_last - t_first)),
            "gpu_calls": len(calls), "mean_rhs_per_call": float(np.mean([c[1] for c in calls])) if calls else 0.,
            "overlap": busy / (t_last - t_first)}
