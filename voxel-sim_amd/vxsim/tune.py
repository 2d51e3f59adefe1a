"""Schedule autotuner.

Per scene-size bucket it measures, on the device, the wall time of whole
solve steps under each candidate schedule and keeps the fastest:

* one-scene arenas (`GridScene`): inner iterations per tick (the length of
  the hipGraph-captured chunk between host reads of the scene state);
* multi-scene arenas (`Batch`, the serving loop): the workgroup-to-tile XCD
  interleave (tiles per XCD chunk) and iterations per tick.

The result is a JSON table. A table measured on MI355X ships next to this
module (`schedule_table.json`); `lookup` reads it (or the file named by
VXSIM_SCHEDULE_TABLE) and falls back to the built-in defaults. Run
`python -m vxsim.tune --out FILE` on a GPU to re-measure.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import time

TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "schedule_table.json")
# Upper REV bounds of the size buckets (the last one is open).
BUCKETS = [2e4, 2e5, 2e6, math.inf]
DEFAULT = {"single": {"iters": 16}, "batch": {"chunk": 16, "iters": 16}}
_cache = {}


def bucket_of(revs: float) -> int:
    for i, hi in enumerate(BUCKETS):
        if revs < hi:
            return i
    return len(BUCKETS) - 1


def load_table(path: str = None) -> dict:
    path = path or os.environ.get("VXSIM_SCHEDULE_TABLE") or TABLE
    if path not in _cache:
        try:
            with open(path) as f:
                _cache[path] = json.load(f)
        except (OSError, ValueError):
            _cache[path] = {}
    return _cache[path]


def lookup(revs: float, kind: str = "single", table: dict = None) -> dict:
    """Schedule for a scene of `revs` REVs ("single") or an arena whose
    scenes are typically `revs` REVs ("batch")."""
    t = table if table is not None else load_table()
    entry = t.get(kind, {}).get(str(bucket_of(revs)))
    out = dict(DEFAULT[kind])
    if entry:
        out.update({k: int(v) for k, v in entry.items() if k in out})
    return out


def _step_params(s, grid):
    c = s.config
    return grid.step_params(fixed_dt=c.fixed_dt, gravity=c.gravity, substeps=c.substeps,
                            corotation_iterations=c.corotation_iterations,
                            corotation_tolerance=c.corotation_tolerance,
                            solve=grid.solve_params(tolerance=1e-6))


def _cube(revs: float):
    e = max(4, int(round(revs ** (1.0 / 3.0))))
    return (e, e, e)


def tune_single(revs: float, iters_candidates=(8, 16, 32), steps: int = 2) -> dict:
    """Fastest iterations per tick for one scene of about `revs` REVs."""
    from . import grid, scenes
    s = scenes.large_64cube(dims=_cube(revs))
    p = _step_params(s, grid)
    best = None
    times = {}
    with grid.GridScene(s, tuned=False) as gs:
        gs.step(fixed_dt=p.fixed_dt, gravity=tuple(p.gravity), solve=p.solve)  # warm-up (graph capture)
        for it in iters_candidates:
            gs.set_schedule(it)
            gs.step(fixed_dt=p.fixed_dt, gravity=tuple(p.gravity), solve=p.solve)
            t = time.perf_counter()
            for _ in range(steps):
                gs.step(fixed_dt=p.fixed_dt, gravity=tuple(p.gravity), solve=p.solve)
            times[it] = (time.perf_counter() - t) / steps
            if best is None or times[it] < times[best]:
                best = it
    return {"iters": best, "seconds": times}


def tune_batch(revs: float, chunk_candidates=(4, 8, 16, 32, 64), iters_candidates=(8, 16, 32),
               total_revs: float = 2e6, seed: int = 7) -> dict:
    """Fastest (chunk, iterations per tick) for an arena of boxes of about
    `revs` REVs each (edges drawn around the cube root), one step each."""
    from . import batch, grid, scenes
    e = max(4, int(round(revs ** (1.0 / 3.0))))
    count = int(max(2, min(256, total_revs // max(revs, 1))))
    boxes = scenes.batch_boxes(count=count, seed=seed, lo=max(3, int(0.7 * e)), hi=max(4, int(1.3 * e)))
    tiles = sum(-(-s.n // batch.TILE) for s in boxes)
    params = [_step_params(s, grid) for s in boxes]
    times = {}
    best = None
    with batch.Batch(tiles * batch.TILE, max_scenes=count) as B:
        for chunk in chunk_candidates:
            for it in iters_candidates:
                B.set_schedule(chunk, it)
                for rep in range(2):  # first pass captures the graphs
                    slots = [B.admit(s, 1, p) for s, p in zip(boxes, params)]
                    t = time.perf_counter()
                    B.run()
                    dt = time.perf_counter() - t
                    for x in slots:
                        B.release(x)
                times["%d/%d" % (chunk, it)] = dt
                if best is None or dt < times["%d/%d" % best]:
                    best = (chunk, it)
    return {"chunk": best[0], "iters": best[1], "seconds": times}


def main():
    ap = argparse.ArgumentParser(description="measure the schedule table on this GPU")
    ap.add_argument("--out", default=TABLE)
    ap.add_argument("--quick", action="store_true", help="fewer candidates")
    args = ap.parse_args()
    table = {"single": {}, "batch": {}, "measured": {}}
    reps = {0: 8e3, 1: 6e4, 2: 5e5, 3: 3e6}
    for b, revs in reps.items():
        r = tune_single(revs)
        table["single"][str(b)] = {"iters": r["iters"]}
        table["measured"]["single_%d" % b] = r["seconds"]
        print("single bucket %d (%g REVs): %s" % (b, revs, r), flush=True)
        if b < 3:
            chunks = (8, 16, 64) if args.quick else (4, 8, 16, 32, 64)
            r = tune_batch(revs, chunk_candidates=chunks)
            table["batch"][str(b)] = {"chunk": r["chunk"], "iters": r["iters"]}
            table["measured"]["batch_%d" % b] = r["seconds"]
            print("batch bucket %d (%g REVs/scene): %s" % (b, revs, r), flush=True)
    with open(args.out, "w") as f:
        json.dump(table, f, indent=1)
    print("wrote", args.out)


if __name__ == "__main__":
    main()
