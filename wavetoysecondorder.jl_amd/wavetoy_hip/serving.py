"""Serving: evolve-to-time-T requests admitted by device-memory footprint,
kept resident on the device, and advanced together by continuous batching.

The reference has no server: each request here is what one evolve3d call does
(src/evolve.jl:1083-1089) -- build the mesh, geometry and Gaussian initial data,
run RK4 to t = T, return the final u and the energy history
(src/evolve.jl:1067-1072).  The loop below admits arrived requests in arrival
order while their footprint (wt2_ctx_footprint plus the caller-owned state and
energy buffer) fits the memory budget, adds each to one wt2_batch, queues a
chunk of steps for all live requests, admits further arrivals while that chunk
runs on the device, then retires the requests whose last step has completed.
"""
import math
import sys
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from .core import (
    Workspace, Batch, adm_bump_background, ctx_footprint, gaussian_pulse, host_coords,
    make_element, make_uniform_hex, make_warped_uniform_hex, recommended_dt_rk4,
    sample_background3d, tau_effective,
)

try:
    import torch
except ImportError:  # pragma: no cover
    torch = None

TAU_LITERAL = 20.0      # the configured SIPG tau; raised to the coercive value per N
EPS_KO = 0.1
WARP_A = 0.05           # x = xi + 0.05 sin(2 pi xi1) sin(2 pi xi2) sin(2 pi xi3)


@dataclass
class Request:
    """One evolution: M^3 elements of N^3 nodes on [0,1]^3, flat or warped with
    the ADM bump background, Gaussian pulse (exp(-r^2/(2 width^2)), amplitude),
    u_t = 0, RK4 with dt = cfl * h_min to t = T."""
    rid: int
    N: int
    M: int
    curved: bool
    centre: tuple
    width: float
    T: float
    arrival: float = 0.0        # seconds after the server starts
    amplitude: float = 1.0
    cfl: float = 0.2

    def mesh(self):
        if self.curved:
            return make_warped_uniform_hex(self.M, self.M, self.M, 0.0, 1.0, WARP_A)
        return make_uniform_hex(self.M, 0.0, 1.0)

    @property
    def npoints(self):
        return self.N ** 3 * self.M ** 3


@dataclass
class Response:
    rid: int
    u: np.ndarray               # final u (host, FP64)
    energy: np.ndarray          # energy at step 0, energy_every, 2 energy_every, ...
    nsteps: int
    dt: float
    arrival: float
    admitted: float
    finished: float
    error: str = ""

    @property
    def latency(self):
        return self.finished - self.arrival


def draw_requests(n, rate, seed, t_range=(0.5, 2.0), N_choices=(4, 6, 8, 10), M_range=(2, 12)):
    """n requests with Poisson arrivals at `rate` per second; per request
    M ~ U{M_range}, N ~ U(N_choices), alternately flat and warped + ADM,
    centre ~ U[0.3,0.7]^3, width ~ U[0.05,0.2], T ~ U[t_range]."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for q in range(n):
        if rate and rate > 0:
            t += float(rng.exponential(1.0 / rate))
        out.append(Request(rid=q, N=int(rng.choice(N_choices)), M=int(rng.integers(M_range[0], M_range[1] + 1)),
                           curved=(q % 2 == 1), centre=tuple(float(c) for c in rng.uniform(0.3, 0.7, size=3)),
                           width=float(rng.uniform(0.05, 0.2)), T=float(rng.uniform(*t_range)), arrival=t))
    return out


def with_arrivals(requests, rate, seed):
    """The same requests with fresh Poisson arrival times at `rate` per second."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for r in requests:
        t += float(rng.exponential(1.0 / rate))
        out.append(Request(**{**r.__dict__, "arrival": t}))
    return out


def build_instance(req, precision):
    """Workspace, device state (u, u_t), dt and step count of a request."""
    mesh = req.mesh()
    elem = make_element(req.N)
    x = host_coords(mesh, elem)
    bg = sample_background3d(adm_bump_background(), 0.0, x[:, 0], x[:, 1], x[:, 2]) if req.curved else None
    ws = Workspace(mesh, elem, tau=tau_effective(TAU_LITERAL, req.N), eps_ko=EPS_KO, precision=precision, bg=bg)
    u0, v0 = gaussian_pulse(x, req.centre, width=req.width, amplitude=req.amplitude)
    dt = recommended_dt_rk4(ws, req.cfl)
    nsteps = max(1, int(math.ceil(req.T / dt - 1e-9)))
    dt = req.T / nsteps
    u = torch.from_numpy(u0).cuda()
    v = torch.from_numpy(v0).cuda()
    return ws, u, v, dt, nsteps


def request_work_estimate(req):
    """Point-updates a request needs: N^3 M^3 points times its RK4 steps, the
    step count from h_min estimated as the smallest GLL gap / M (times 0.82, the
    warp's smallest stretch, on warped meshes)."""
    xs = np.sort(np.asarray(make_element(req.N).xs, dtype=np.float64))
    h = float(np.min(np.diff(xs))) / req.M * (0.82 if req.curved else 1.0)
    return req.npoints * int(math.ceil(req.T / (req.cfl * h)))


def request_footprint(req, precision, energy_every):
    """Device bytes a live request holds: the context's allocations, the FP64
    state u and u_t, and the energy-sample buffer.  The step count behind the
    last is bounded with h_min >= 0.8 x the flat GLL gap (the warp's smallest
    stretch is about 0.82)."""
    mesh = req.mesh()
    elem = make_element(req.N)
    b = ctx_footprint(mesh, elem, precision, with_background=req.curved)
    b += 2 * 8 * req.npoints
    if energy_every > 0:
        xs = np.sort(np.asarray(elem.xs, dtype=np.float64))
        h_lo = 0.8 * float(np.min(np.diff(xs))) / req.M
        nsteps_hi = int(math.ceil(req.T / (req.cfl * h_lo))) + 1
        b += 8 * (nsteps_hi // energy_every + 2)
    return b


def admit_fifo(queue, used, budget, footprint, live_points=0, max_live_points=None):
    """Arrival-order admission: take requests from the head of `queue` while
    their footprint fits in `budget - used` (and the live point count stays
    under max_live_points, if given).  A request larger than the whole budget
    is rejected.  Returns (admitted, rejected, used)."""
    admitted, rejected = [], []
    while queue:
        r = queue[0]
        need = footprint(r)
        if need > budget:
            rejected.append(queue.popleft())
            continue
        if used + need > budget:
            break
        if max_live_points is not None and live_points > 0 and live_points + r.npoints > max_live_points:
            break
        queue.popleft()
        admitted.append((r, need))
        used += need
        live_points += r.npoints
    return admitted, rejected, used


@dataclass
class _Live:
    req: Request
    ws: object
    u: object
    v: object
    dt: float
    nsteps: int
    need: int
    admitted: float


@dataclass
class ServeStats:
    responses: list = field(default_factory=list)
    wall: float = 0.0
    point_updates: int = 0
    max_live: int = 0
    peak_bytes: int = 0
    chunks: int = 0
    host_admit_s: float = 0.0     # building contexts and state of admitted requests
    host_retire_s: float = 0.0    # copying results out, freeing contexts
    wait_s: float = 0.0           # host waiting for queued chunks

    def summary(self):
        lat = np.array([r.latency for r in self.responses if not r.error])
        return {
            "requests": len(self.responses),
            "rejected": sum(1 for r in self.responses if r.error),
            "p50_latency_s": float(np.percentile(lat, 50)) if lat.size else None,
            "p99_latency_s": float(np.percentile(lat, 99)) if lat.size else None,
            "mean_latency_s": float(lat.mean()) if lat.size else None,
            "throughput_point_updates_per_s": self.point_updates / self.wall if self.wall > 0 else None,
            "wall_s": self.wall,
            "max_live_requests": self.max_live,
            "peak_footprint_bytes": self.peak_bytes,
            "chunks": self.chunks,
            "host_admit_s": self.host_admit_s,
            "host_retire_s": self.host_retire_s,
            "host_wait_s": self.wait_s,
        }


class Server:
    """Continuous-batching server over one wt2_batch (mixed precision by
    default).  `run(requests)` replays the requests' arrival times against the
    wall clock and returns ServeStats with one Response per request."""

    def __init__(self, budget_bytes=200e9, precision="mixed", energy_every=20, chunk_ms=20.0,
                 max_live_points=None, clock=time.perf_counter):
        if torch is None or not torch.cuda.is_available():
            raise RuntimeError("the server needs a HIP device")
        self.budget = int(budget_bytes)
        self.precision = precision
        self.energy_every = int(energy_every)
        self.chunk_ms = float(chunk_ms)
        self.max_live_points = max_live_points
        self.clock = clock
        self._fp = {}

    def footprint(self, req):
        key = (req.N, req.M, req.curved)
        if key not in self._fp:
            self._fp[key] = request_footprint(req, self.precision, self.energy_every)
        return self._fp[key]

    def prewarm(self, requests, verbose=False):
        """Tune the pass tiles once per distinct (N, M, flat/warped) shape, as
        a server does when it starts, so that no request pays for it."""
        seen = set()
        for r in requests:
            key = (r.N, r.M, r.curved)
            if key in seen:
                continue
            seen.add(key)
            if verbose:
                print(f"tuning N={r.N} M={r.M} {'warped' if r.curved else 'flat'}", file=sys.stderr, flush=True)
            ws, u, v, dt, _ = build_instance(r, self.precision)
            ws.autotune(u, v, dt)
            ws.close()
            del u, v
        torch.cuda.synchronize()

    def run(self, requests):
        reqs = sorted(requests, key=lambda r: r.arrival)
        pending = deque(reqs)
        queue = deque()
        live = {}                       # slot -> _Live
        stats = ServeStats()
        batch = Batch()
        used = 0
        live_pts = 0
        t_step = None                   # seconds per step of the current live set (EMA)
        clk = self.clock
        t0 = clk()

        def now():
            return clk() - t0

        def admit():
            nonlocal used, live_pts
            ta = clk()
            t = now()
            while pending and pending[0].arrival <= t:
                queue.append(pending.popleft())
            adm, rej, used = admit_fifo(queue, used, self.budget, self.footprint, live_pts, self.max_live_points)
            for r in rej:
                stats.responses.append(Response(r.rid, np.zeros(0), np.zeros(0), 0, 0.0, r.arrival, t, t,
                                                error="footprint exceeds the memory budget"))
            for r, need in adm:
                ws, u, v, dt, ns = build_instance(r, self.precision)
                slot = batch.add(ws, u, v, dt, ns, energy_every=self.energy_every)
                live[slot] = _Live(r, ws, u, v, dt, ns, need, now())
                live_pts += r.npoints
                stats.point_updates += r.npoints * ns
            stats.max_live = max(stats.max_live, len(live))
            stats.peak_bytes = max(stats.peak_bytes, used)
            stats.host_admit_s += clk() - ta

        def retire(slots):
            nonlocal used, live_pts
            tr = clk()
            for slot in slots:
                L = live.pop(slot)
                done, E = batch.slot_info(slot)
                batch.remove(slot)
                u = L.u.cpu().numpy()
                stats.responses.append(Response(L.req.rid, u, E, L.nsteps, L.dt, L.req.arrival, L.admitted, now()))
                L.ws.close()
                used -= L.need
                live_pts -= L.req.npoints
            stats.host_retire_s += clk() - tr

        while pending or queue or live:
            admit()
            if not live:
                if pending and not queue:
                    time.sleep(max(0.0, pending[0].arrival - now()))
                continue
            k = 8 if t_step is None else int(min(512, max(1, round(self.chunk_ms * 1e-3 / t_step))))
            before = {s: batch.steps_done(s) for s in live}
            t_issue = clk()
            batch.step_async(k)
            stats.chunks += 1
            admit()                    # host work for new arrivals overlaps the queued chunk
            tw = clk()
            batch.sync()
            stats.wait_s += clk() - tw
            adv = max(batch.steps_done(s) - before[s] for s in before)
            if adv > 0:
                dt_meas = (clk() - t_issue) / adv
                t_step = dt_meas if t_step is None else 0.5 * (t_step + dt_meas)
            retire([s for s in before if batch.steps_done(s) >= live[s].nsteps])
        stats.wall = now()
        batch.close()
        return stats


def run_one_at_a_time(requests, precision="fp64", energy_every=20):
    """The same requests served one at a time through the single-instance
    path (wt2_rk4_steps), in arrival order.  Each request's service time is
    measured (set-up, steps, copy-out); latencies follow from a FIFO queue on
    the arrival times.  Returns (responses, service_times, point_updates)."""
    reqs = sorted(requests, key=lambda r: r.arrival)
    service = []
    out = []
    work = 0
    for r in reqs:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ws, u, v, dt, ns = build_instance(r, precision)
        E = ws.rk4_steps(u, v, dt, ns, energy_every=energy_every)
        uh = u.cpu().numpy()
        ws.close()
        service.append(time.perf_counter() - t0)
        out.append(Response(r.rid, uh, np.asarray(E), ns, dt, r.arrival, 0.0, 0.0))
        work += r.npoints * ns
    return out, service, work


def fifo_latencies(arrivals, service):
    """Completion-minus-arrival of a single FIFO server given service times."""
    t = 0.0
    lat = []
    for a, s in zip(arrivals, service):
        start = max(a, t)
        t = start + s
        lat.append(t - a)
    return np.array(lat)
