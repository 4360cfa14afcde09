"""Device-resident serving loop: continuous batching of simulation requests.

Requests (3D bluff-body runs, 2D thermal cavities) arrive over time.  Each
admitted request becomes a Domain whose populations stay in HBM until its
last step; requests of one kind (lattice, population storage, collision,
LES, temperature coupling) share a Batch that advances all of them with one
launch per step.  At every chunk boundary the loop completes the chunks in
flight (one per batch, each on its own stream, so the kinds overlap on the
GPU), retires the requests that reached their step count, puts in the
domains built meanwhile and launches the next chunk; while that chunk runs
it builds domains for queued requests while the HBM budget allows (first
fit: a request that does not fit waits; others may still enter).  A request
therefore joins the running mix at the next chunk boundary after memory
frees up, and building its domain does not stop the GPU.

Admission order (`policy`):
  * "least_work" (default): the queued requests with the fewest node-steps
    (nodes x steps, the time a request needs on a full GPU) first, and only
    while the running mix holds fewer than `fill_nodes` nodes -- enough to
    fill the GPU -- so short requests finish early instead of sharing the
    GPU with every long one (shortest-job-first; the mean and median time to
    result fall, the makespan does not grow);
    A request that has waited longer than `max_wait_s` goes ahead of every
    request that has not (oldest first), so a stream of short arrivals
    cannot hold a long request back indefinitely;
  * "arrival": arrival order, everything that fits in the budget at once.

A request whose flow diverges (the step kernels raise a per-domain flag
when a density comes out of the collision non-finite) is retired at the next
chunk boundary with an error result instead of running on with NaNs, and
its slot and memory go to the queue.

Results do not depend on the mix: a domain's arithmetic is the same alone
and batched (tests/test_gpu_serve.py checks that).

`plain_loop` is the reference-style baseline: one request at a time, in
arrival order, FP32 populations, the single-domain step.
"""
import math
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from .analysis import aerodynamic_coefficients
from .batch import Batch
from .grid import LatticeGrid2D, LatticeGrid3D, Storage


@dataclass
class SimRequest:
    rid: int
    kind: str                 # "bluff3d" or "cavity2d"
    n: int                    # grid edge: n^3 (3D) or n^2 (2D)
    steps: int
    re: float = 0.0           # bluff3d: Reynolds number on U and 2 x the largest semi-axis
    ra: float = 0.0           # cavity2d: Rayleigh number (Pr 0.71)
    semi_axes: tuple = (8.0, 8.0, 8.0)
    axes: np.ndarray = None   # rows = body axes (bluff3d)
    tol: float = 1e-5         # allowed rel L2 error vs FP64: >= 1e-3 admits FP16 storage
    arrival: float = 0.0      # seconds after the start of the stream
    # filled in by the server
    admitted: float = None
    finished: float = None
    result: dict = field(default_factory=dict)

    @property
    def storage(self):
        return Storage.F16 if self.tol >= 1e-3 else Storage.F32


U_BLUFF = 0.05
CS_LES = 0.17
PR = 0.71
U_FREEFALL = 0.1


def bluff_geometry(req):
    n = req.n
    c = (n / 3.0, n / 2.0, n / 2.0)
    rmax = max(req.semi_axes)
    nu = U_BLUFF * 2.0 * rmax / req.re
    return c, rmax, 3.0 * nu + 0.5


def make_domain(req, storage=None):
    """Grid for a request (storage: None = the request's own choice)."""
    st = req.storage if storage is None else storage
    if req.kind == "bluff3d":
        c, rmax, tau = bluff_geometry(req)
        n = req.n
        g = LatticeGrid3D(n, n, n, tau, storage=st)
        g.set_periodic_axes(True, True)
        g.initialize(1.0, U_BLUFF, 0.0, 0.0)
        g.set_les_coupled(True)
        g.set_inflow_west(U_BLUFF, 0.0)
        g.set_convective_outflow_east(U_BLUFF)
        g.set_solid_ellipsoid(*c, *req.semi_axes, req.axes)
        return g
    n = req.n
    H = n - 2
    nu = U_FREEFALL * H * math.sqrt(PR / req.ra)
    alpha = nu / PR
    g = LatticeGrid2D(n, n, 3.0 * nu + 0.5, storage=st)
    g.initialize(1.0, 0.0, 0.0)
    g.set_solid_box(0, 0, n - 1, 0)
    g.set_solid_box(0, n - 1, n - 1, n - 1)
    g.set_solid_box(0, 0, 0, n - 1)
    g.set_solid_box(n - 1, 0, n - 1, n - 1)
    g.enable_thermal(3.0 * alpha + 0.5, U_FREEFALL ** 2 / H, 0.5, 0.5)
    g.set_wall_temperature(0, 0, 0, n - 1, 1.0)
    g.set_wall_temperature(n - 1, 0, n - 1, n - 1, 0.0)
    return g


def batch_key(req, storage=None):
    st = int(req.storage if storage is None else storage)
    if req.kind == "bluff3d":
        return (19, st, "MRT", True, False)
    return (9, st, "MRT", False, True)


def step_kwargs(req):
    return {"use_les": True, "smag": CS_LES} if req.kind == "bluff3d" else {}


def nusselt(T, ux, n, alpha):
    T = T.reshape(n, n).astype(np.float64)
    ux = ux.reshape(n, n).astype(np.float64)
    H = n - 2
    rows = slice(1, n - 1)
    q = [(0.5 * (ux[rows, x] * T[rows, x] + ux[rows, x + 1] * T[rows, x + 1])
          - alpha * (T[rows, x + 1] - T[rows, x])).sum() / H for x in range(1, n - 2)]
    return float(np.mean(q)) * H / alpha


def result_of(req, g, force=None):
    """bluff3d: C_D from the momentum-exchange force on the final state
    (equivalent-sphere area); cavity2d: mean Nusselt number."""
    if req.kind == "bluff3d":
        c, rmax, _ = bluff_geometry(req)
        F = force if force is not None else g.momentum_exchange(*c, 1.5 * rmax)
        r_eq = float(np.prod(req.semi_axes)) ** (1.0 / 3.0)
        co = aerodynamic_coefficients(F[0], F[1], F[2], 0.0, 1.0, U_BLUFF, math.pi * r_eq * r_eq)
        return {"C_D": float(co["cd"]), "C_L": float(co["cl"])}
    n = req.n
    nu = U_FREEFALL * (n - 2) * math.sqrt(PR / req.ra)
    return {"Nu": nusselt(g.temperature_data(), g.ux_data(), n, nu / PR)}


def bytes_needed(req, storage=None):
    """Device bytes of the request's domain (wake_domain_bytes_needed)."""
    from . import _lib

    st = int(req.storage if storage is None else storage)
    if req.kind == "bluff3d":
        return _lib.lib().wake_domain_bytes_needed(3, 19, req.n, req.n, req.n, st, -1, 0)
    return _lib.lib().wake_domain_bytes_needed(2, 9, req.n, req.n, 1, st, -1, 1)


def node_count(req):
    return req.n ** 3 if req.kind == "bluff3d" else req.n ** 2


def admission_order(queue, now, policy="least_work", max_wait_s=60.0):
    """The order in which queued requests are offered a place.  least_work:
    requests waiting longer than max_wait_s first (by arrival), then the
    rest by node-steps (then arrival); arrival: arrival order."""
    order = list(queue)
    if policy == "least_work":
        order.sort(key=lambda r: (0, r.arrival, 0) if now - r.arrival > max_wait_s
                   else (1, node_count(r) * r.steps, r.arrival))
    else:
        order.sort(key=lambda r: r.arrival)
    return order


class Server:
    """Continuous-batching request loop under an HBM admission budget."""

    def __init__(self, budget_bytes=200e9, chunk=100, use_graph=True, tune=True, policy="least_work",
                 fill_nodes=4_000_000, max_wait_s=60.0):
        if policy not in ("least_work", "arrival"):
            raise ValueError(f"unknown admission policy {policy!r}")
        self.budget = budget_bytes
        self.chunk = chunk
        self.use_graph = use_graph
        self.tune = tune
        self.policy = policy
        self.fill_nodes = fill_nodes
        self.max_wait_s = max_wait_s
        self.active_nodes = 0
        self.batches = {}
        self.active = {}        # rid -> (req, grid, key, slot)
        self.queue = deque()
        self.used = 0
        self.done = []
        self.rejected = []      # requests that need more HBM than the whole budget
        self.diverged = []      # ids of requests retired because their flow diverged
        self.peak_used = 0
        self.ticks = 0
        self.node_steps = 0
        self.prepared = []      # (request, domain) built while a chunk ran, joining at the next boundary
        self.inflight = []      # batches with a chunk in flight
        # host seconds: waiting for chunks, at chunk boundaries (results,
        # joins, launches) and building domains (overlapping the chunks)
        self.stats = {"wait_s": 0.0, "boundary_s": 0.0, "build_s": 0.0}

    def submit(self, req):
        """Queue a request; one that could never fit the budget is rejected
        (kept in `rejected` with the reason) instead of waiting forever."""
        need = bytes_needed(req)
        if need > self.budget:
            req.result = {"error": f"needs {need / 1e9:.2f} GB, more than the {self.budget / 1e9:.2f} GB budget"}
            self.rejected.append(req)
            return
        self.queue.append(req)

    def _admit(self, now):
        """Build domains for queued requests that fit; they join their batch
        at the next chunk boundary."""
        kept = deque()
        for req in admission_order(self.queue, now, self.policy, self.max_wait_s):
            need = bytes_needed(req)
            busy = self.active or self.prepared
            full = self.policy == "least_work" and busy and self.active_nodes >= self.fill_nodes
            if full or self.used + need > self.budget:
                kept.append(req)
                continue
            t0 = time.perf_counter()
            self.prepared.append((req, make_domain(req)))
            self.stats["build_s"] += time.perf_counter() - t0
            self.used += need
            self.active_nodes += node_count(req)
            self.peak_used = max(self.peak_used, self.used)
            req.admitted = now
        self.queue = kept

    def _join(self, req, g):
        key = batch_key(req)
        b = self.batches.get(key)
        if b is None:
            lat, st, coll, les, thermal = key
            b = Batch(lattice=lat, storage=st, collision=coll, use_les=les, thermal=thermal, chunk=self.chunk)
            # row segments of 64 lanes: no idle lanes for the 64-multiple
            # grid edges of the request mix
            b.set_launch(64, 0)
            self.batches[key] = b
        slot = b.add(g, req.steps, **({"smag": CS_LES} if req.kind == "bluff3d" else {}))
        self.active[req.rid] = (req, g, key, slot)

    def tick(self, now_fn):
        """One chunk boundary.  Complete the chunks in flight; take the
        requests that reached their step count (or diverged) out of their
        batches and compute their results; put in the domains built
        meanwhile; launch the next chunk on every batch; then, while it runs,
        build domains for queued requests that now fit.  Returns whether
        anything is in flight or waiting to join."""
        t0 = time.perf_counter()
        for b in self.inflight:
            b.wait()
        self.inflight = []
        self.stats["wait_s"] += time.perf_counter() - t0
        t0 = time.perf_counter()
        for rid in list(self.active):
            req, g, key, slot = self.active[rid]
            b = self.batches[key]
            # a request whose flow diverged (the step kernels flag a
            # non-finite density) is retired with an error, freeing its slot
            diverged = g.nonfinite_seen()
            if diverged or b.remaining(slot) == 0:
                b.remove(slot)
                del self.active[rid]
                if diverged:
                    req.result = {"error": f"diverged: non-finite density by step {g.current_step()}"}
                    self.diverged.append(req.rid)
                else:
                    req.result = result_of(req, g)
                req.finished = now_fn()
                self.used -= bytes_needed(req)
                self.active_nodes -= node_count(req)
                self.node_steps += g.total_nodes() * req.steps
                self.done.append(req)
                del g
        for req, g in self.prepared:
            self._join(req, g)
        self.prepared = []
        live = [b for b in self.batches.values() if b.info()["members"] > 0]
        for b in live:
            b.launch(self.chunk, self.use_graph)
        self.inflight = live
        self.stats["boundary_s"] += time.perf_counter() - t0
        if live:
            self.ticks += 1
        self._admit(now_fn())
        return bool(live) or bool(self.prepared)

    def serve(self, requests, realtime=True):
        """Replay `requests` (sorted by arrival) against the wall clock."""
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0
        while pending or self.queue or self.active or self.prepared or self.inflight:
            while pending and (not realtime or pending[0].arrival <= now()):
                self.submit(pending.popleft())
            if not self.inflight and not self.prepared:
                self._admit(now())
            if not self.tick(now) and pending:
                wait = pending[0].arrival - now()
                if wait > 0:
                    time.sleep(wait)
        return now()


def plain_loop(requests, realtime=True):
    """One request at a time in arrival order, FP32, single-domain step."""
    t0 = time.perf_counter()
    now = lambda: time.perf_counter() - t0
    out = []
    for req in sorted(requests, key=lambda r: r.arrival):
        if realtime and req.arrival > now():
            time.sleep(req.arrival - now())
        r = SimRequest(**{k: getattr(req, k) for k in ("rid", "kind", "n", "steps", "re", "ra", "semi_axes", "axes",
                                                         "tol", "arrival")})
        r.admitted = now()
        g = make_domain(r, storage=Storage.F32)
        g.run(r.steps, "MRT", **step_kwargs(r))
        if g.nonfinite_seen():
            r.result = {"error": f"diverged: non-finite density by step {g.current_step()}"}
        else:
            r.result = result_of(r, g)
        r.finished = now()
        out.append(r)
        del g
    return out, now()


def random_rotation(rng):
    q, r = np.linalg.qr(rng.standard_normal((3, 3)))
    q = q * np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q.astype(np.float32)


def request_stream(count, rate=20.0, seed=5, step_scale=1.0, n3=(64, 96, 128, 160, 192),
                   n2=(128, 192, 256, 320, 384, 448, 512), p_fp16=0.7):
    """Poisson arrivals (mean `rate` per second): 60% D3Q19 MRT-LES bluff
    bodies (rotated ellipsoid, semi-axes U[8, min(32, n/6)], n^3 grid,
    Re U[100, 1000]), 40% D2Q9 + D2Q5 cavities (n^2, Ra log-U[1e4, 1e6]);
    steps U[2000, 20000] x step_scale; error budget 1e-3 (FP16 storage) with
    probability p_fp16, else 1e-5."""
    rng = np.random.default_rng(seed)
    t = 0.0
    reqs = []
    for i in range(count):
        t += rng.exponential(1.0 / rate)
        steps = max(10, int(rng.integers(2000, 20001) * step_scale))
        tol = 1e-3 if rng.random() < p_fp16 else 1e-5
        if rng.random() < 0.6:
            n = int(rng.choice(n3))
            hi = min(32.0, n / 6.0)
            ax = tuple(float(v) for v in rng.uniform(8.0, hi, 3))
            reqs.append(SimRequest(i, "bluff3d", n, steps, re=float(rng.uniform(100, 1000)), semi_axes=ax,
                                   axes=random_rotation(rng), tol=tol, arrival=t))
        else:
            n = int(rng.choice(n2))
            reqs.append(SimRequest(i, "cavity2d", n, steps, ra=float(10 ** rng.uniform(4, 6)), tol=tol, arrival=t))
    return reqs


def write_results_jsonl(done, path):
    """One JSON object per finished request: its parameters, storage,
    arrival / admission / finish times (s), time to result and result."""
    import json

    with open(path, "w") as f:
        for r in sorted(done, key=lambda r: r.rid):
            rec = {"rid": r.rid, "kind": r.kind, "n": r.n, "steps": r.steps, "tol": r.tol,
                   "storage": Storage(r.storage).name.lower(), "arrival_s": r.arrival, "admitted_s": r.admitted,
                   "finished_s": r.finished,
                   "time_to_result_s": None if r.finished is None else r.finished - r.arrival,
                   "result": r.result}
            if r.kind == "bluff3d":
                rec.update(re=r.re, semi_axes=list(r.semi_axes))
            else:
                rec["ra"] = r.ra
            f.write(json.dumps(rec) + "\n")


def latency_stats(done):
    ttr = np.array([r.finished - r.arrival for r in done])
    return {"p50_s": float(np.percentile(ttr, 50)), "p99_s": float(np.percentile(ttr, 99)),
            "mean_s": float(ttr.mean()), "max_s": float(ttr.max())}
