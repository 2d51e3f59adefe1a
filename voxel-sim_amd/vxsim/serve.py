"""Serving loop: continuous batching of solve requests on one tile arena.

Requests (a box scene, a number of fixed steps, step parameters) arrive over
time. Each scheduler iteration admits every queued request that has arrived
and fits in the arena (a request that does not fit yet waits while others
behind it may go), advances the admitted scenes by one tick, and retires
finished scenes: their state is read back, their tiles freed. Matrices,
rotations and warm-start vectors stay resident in HBM for the whole life of a
scene. The arena is sized from an HBM budget (`vx_batch_bytes_per_row`).

Policies:
  * "fifo": admission oldest first, every admitted scene in every tick. A
    tick's launches cover every running scene's tiles, so a small request
    takes about as long as the largest one it shares the arena with.
  * "sjf" (shortest job first, the default): admission smallest first (REVs x
    steps, divided by 1 + waiting time / `aging_s`), and the small running
    scenes (fewer than active_revs / 2 REVs) run first: each tick takes the
    smallest of them up to `active_revs` REVs together (at least one), and
    holds the rest (vx_batch_hold: state kept in HBM, no work; a tick
    launches only its scenes' tiles). The large scenes all run together, as
    many per tick as possible, when no small scene runs, or in every tick once
    one of them has waited `aging_s` since its arrival (a stream of small
    requests does not hold them forever).

Latency of a request = time its result is back on the host - its arrival.
"""
from __future__ import annotations

import dataclasses
import time
from typing import Callable, List, Optional

import numpy as np

from . import batch as _batch
from . import grid as _grid


@dataclasses.dataclass
class Request:
    rid: int
    arrival: float           # seconds from the start of the run
    scene: object            # vxsim.scenes.Scene (a full box)
    steps: int
    params: object           # vxsim._abi.StepParams
    # filled in by the server
    admitted: Optional[float] = None
    finished: Optional[float] = None
    state: int = 0
    status: int = 0
    report: Optional[dict] = None
    result: Optional[dict] = None

    @property
    def latency(self) -> Optional[float]:
        return None if self.finished is None else self.finished - self.arrival


POLICIES = ("sjf", "fifo")


class Server:
    def __init__(self, capacity_rows: int, max_scenes: int = 256, xcd_chunk_tiles: int = 0,
                 keep_results: bool = False, policy: str = "sjf", active_revs: int = 1 << 20,
                 aging_s: float = 10.0):
        if policy not in POLICIES:
            raise ValueError("policy must be one of %s" % (POLICIES,))
        if active_revs <= 0 or aging_s <= 0:
            raise ValueError("active_revs and aging_s must be positive")
        self.batch = _batch.Batch(capacity_rows, max_scenes, xcd_chunk_tiles)
        self.max_scenes = max_scenes
        self.keep_results = keep_results
        self.policy = policy
        self.active_revs = int(active_revs)
        self.aging_s = float(aging_s)
        self.ticks = 0

    def _age(self, req: "Request", now: float) -> float:
        return 1.0 + max(now - req.arrival, 0.0) / self.aging_s

    def _key(self, req: "Request", now: float) -> float:
        """Size of a request for shortest-job-first, reduced by its waiting time."""
        return req.scene.n * req.steps / self._age(req, now)

    def _choose(self, running: dict, held: dict, now: float) -> None:
        """Hold or resume the running scenes for the next tick (policy "sjf")."""
        small = [slot for slot in running if running[slot].scene.n < self.active_revs / 2]
        large = [slot for slot in running if running[slot].scene.n >= self.active_revs / 2]
        run = set()
        total = 0
        for slot in sorted(small, key=lambda q: self._key(running[q], now)):
            n = running[slot].scene.n
            if run and total + n > self.active_revs:
                break
            run.add(slot)
            total += n
        if not small or any(now - running[q].arrival >= self.aging_s for q in large):
            run.update(large)
        for slot in running:
            hold = slot not in run
            if held.get(slot, False) != hold:
                self.batch.hold(slot, hold)
                held[slot] = hold

    @classmethod
    def for_budget(cls, hbm_bytes: int, **kw) -> "Server":
        return cls(_batch.rows_for_bytes(hbm_bytes), **kw)

    def close(self):
        self.batch.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def run(self, requests: List[Request], clock: Callable[[], float] = time.perf_counter,
            on_tick: Optional[Callable[["Server"], None]] = None, max_ticks: int = 10_000_000) -> List[Request]:
        """Serve `requests` (sorted by arrival) to completion; arrival times
        are honoured against `clock` from the start of the call."""
        B = self.batch
        sjf = self.policy == "sjf"
        pending = sorted(requests, key=lambda r: r.arrival)
        queue: List[Request] = []
        running = {}  # slot -> Request
        held = {}     # slot -> whether the scene is held
        t0 = clock()
        nxt = 0
        while nxt < len(pending) or queue or running:
            now = clock() - t0
            while nxt < len(pending) and pending[nxt].arrival <= now:
                queue.append(pending[nxt])
                nxt += 1
            # admission: oldest (fifo) or smallest (sjf) first, skipping
            # requests that do not fit yet
            if queue:
                if sjf:
                    queue.sort(key=lambda r: self._key(r, now))
                free = B.free_rows()[0]
                kept = []
                for req in queue:
                    need = -(-req.scene.n // _batch.TILE) * _batch.TILE
                    slot = None
                    if need <= free and len(running) < self.max_scenes:
                        slot = B.admit(req.scene, req.steps, req.params)
                    if slot is None:
                        kept.append(req)
                        continue
                    req.admitted = clock() - t0
                    running[slot] = req
                    free = B.free_rows()[0]
                queue = kept
            if not running:
                # idle until the next arrival
                if nxt < len(pending):
                    wait = pending[nxt].arrival - (clock() - t0)
                    if wait > 0:
                        time.sleep(min(wait, 0.01))
                continue
            if sjf:
                self._choose(running, held, clock() - t0)
            B.tick()
            self.ticks += 1
            if on_tick:
                on_tick(self)
            if self.ticks >= max_ticks:
                raise RuntimeError("serving loop exceeded %d ticks" % max_ticks)
            states = B.states()
            for slot in list(running):
                if states[slot] in (_batch.DONE, _batch.FAILED):
                    rep = B.report(slot)
                    req = running.pop(slot)
                    if self.keep_results or rep["state"] == _batch.DONE:
                        res = B.download(slot)
                        req.result = res if self.keep_results else None
                    req.finished = clock() - t0
                    req.state = rep["state"]
                    req.status = rep["status"]
                    req.report = rep
                    B.release(slot)
                    held.pop(slot, None)
        return requests


def serve_alone_fp64(requests: List[Request], clock: Callable[[], float] = time.perf_counter) -> List[dict]:
    """The same requests handled one at a time, in arrival order, by the plain
    FP64 structured solve (a `GridScene` per request). Returns per-request
    dicts with latency (including waiting for the previous request) and the
    DOF-iterations done."""
    out = []
    t0 = clock()
    for req in sorted(requests, key=lambda r: r.arrival):
        now = clock() - t0
        if req.arrival > now:
            time.sleep(req.arrival - now)
        p = req.params
        solve = _grid.solve_params(tolerance=p.solve.tolerance, max_inner=max(p.solve.max_inner, 20000),
                                   mode=_grid.F64, warm_start=bool(p.solve.warm_start))
        dof_iters = 0.0
        ok = True
        with _grid.GridScene(req.scene) as gs:
            for _ in range(req.steps):
                try:
                    r = gs.step(fixed_dt=p.fixed_dt, gravity=tuple(p.gravity), substeps=p.substeps,
                                corotation_iterations=p.corotation_iterations,
                                corotation_tolerance=p.corotation_tolerance, solve=solve)
                except Exception:
                    ok = False
                    break
                dof_iters += 3.0 * r["active_rows"] * (r["inner_iterations"] + r["outer_sweeps"])
        out.append({"rid": req.rid, "latency": clock() - t0 - req.arrival, "dof_iters": dof_iters, "ok": ok})
    return out


def percentiles(latencies, qs=(50, 99)):
    a = np.asarray([x for x in latencies if x is not None], dtype=float)
    if a.size == 0:
        return {q: None for q in qs}
    return {q: float(np.percentile(a, q)) for q in qs}


def request_stream(rate: float, seconds: float, rev_lo: float, rev_hi: float, steps: int, seed: int,
                   solve=None, fixed_dt=0.001, gravity=(0.0, 0.0, -9.81), corotation_iterations=2,
                   corotation_tolerance=1e-5, max_requests: Optional[int] = None) -> List[Request]:
    """Poisson arrivals at `rate` per second for `seconds`; each request a
    box whose REV count is log-uniform in [rev_lo, rev_hi] with aspect ratios
    U[0.5, 2], regional materials and one random fixed face, for `steps`
    steps with the warm start kept on the device."""
    from . import scenes
    rng = scenes.SplitMix64(seed)
    reqs: List[Request] = []
    t = 0.0
    solve = solve if solve is not None else _grid.solve_params(tolerance=1e-6, warm_start=True)
    params = _grid.step_params(fixed_dt=fixed_dt, gravity=gravity, substeps=1,
                               corotation_iterations=corotation_iterations,
                               corotation_tolerance=corotation_tolerance, solve=solve)
    while True:
        u = rng.uniform(5)
        t += -np.log(max(1.0 - float(u[0]), 1e-300)) / rate
        if t >= seconds or (max_requests is not None and len(reqs) >= max_requests):
            break
        n = float(np.exp(np.log(rev_lo) + float(u[1]) * (np.log(rev_hi) - np.log(rev_lo))))
        ay, az = 0.5 + 1.5 * float(u[2]), 0.5 + 1.5 * float(u[3])
        nx = max(2, int(round((n / (ay * az)) ** (1.0 / 3.0))))
        dims = (nx, max(2, int(round(ay * nx))), max(2, int(round(az * nx))))
        face = "xyz"[int(float(u[4]) * 6) // 2 % 3] + "-+"[int(float(u[4]) * 6) % 2]
        s = scenes.mixed_box(dims, seed * 1_000_003 + len(reqs), face)
        reqs.append(Request(rid=len(reqs), arrival=t, scene=s, steps=steps, params=params))
    return reqs
