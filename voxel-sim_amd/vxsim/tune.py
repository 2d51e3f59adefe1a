"""Schedule and launch-shape autotuner.

Per scene-size bucket it measures, on the device, candidate schedules and
launch shapes and keeps the fastest:

* one-scene arenas (`GridScene`): the launch shape of the inner kernels --
  the z-column group of the tile order (0, 4, 8, 16) and the XCD deal of the
  SpMV and of the kernels after it, then the update kernel's workgroup per
  512-row tile (512 threads with one row each or 256 with two)
  (vx_scene_set_shape) -- and then the inner iterations per tick (the length
  of the hipGraph-captured chunk between host reads of the scene state),
  timed over whole solves of the bucket's representative scene;
* multi-scene arenas (`Batch`, the serving loop): the workgroup-to-tile XCD
  interleave (tiles per XCD chunk) and iterations per tick, timed over one
  step of an arena of boxes.

A candidate replaces the default only when it is at least 1 % faster. The
tile itself (512 rows, SoA planes, per-tile partial sums) is the arena's
storage unit and does not vary; the shape decides which workgroup runs which
tile and how many threads it has, and the results are identical bit for bit
under every shape.

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
BUCKETS = {"single": [2e4, 2e5, 4e5, 1.5e6, math.inf], "batch": [2e4, 2e5, 2e6, math.inf]}
DEFAULT = {"single": {"iters": 16, "zcols": 8, "deal": 0, "update_threads": 512},
           "batch": {"chunk": 16, "iters": 16}}
UPDATE_THREADS = (512, 256)
ZCOLS = (0, 4, 8, 16)
DEALS = (0, 1, 2, 3)  # VX_DEAL_AUTO, _CONTIGUOUS, _SPLIT, _ROUND_ROBIN
_cache = {}


def bucket_of(revs: float, kind: str = "single") -> int:
    bounds = BUCKETS[kind]
    for i, hi in enumerate(bounds):
        if revs < hi:
            return i
    return len(bounds) - 1


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
    """Schedule (and for "single", launch shape) for a scene of `revs` REVs
    ("single") or an arena whose scenes are typically `revs` REVs ("batch")."""
    t = table if table is not None else load_table()
    entry = t.get(kind, {}).get(str(bucket_of(revs, kind)))
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


def single_scene(bucket: int):
    """The representative scene of a one-scene bucket: mixed-material cubes
    of 20^3, 39^3 and 64^3 REVs, then the cutting scene's wood blocks of
    96x96x64 and 192x192x96."""
    from . import scenes
    if bucket <= 2:
        return scenes.large_64cube(dims=((20, 20, 20), (39, 39, 39), (64, 64, 64))[bucket])
    return scenes.wood_block_blade(dims=(96, 96, 64) if bucket == 3 else (192, 192, 96))


def tune_single(scene, iters_candidates=(8, 16, 32), reps: int = 2, log=None) -> dict:
    """Fastest launch shape, then iterations per tick, for one scene: the
    time of whole solves of its first assembled system (the same iterations
    under every candidate)."""
    from . import grid
    cfg = scene.config
    h = cfg.fixed_dt / cfg.substeps
    sp = grid.solve_params(tolerance=1e-9, inner_tolerance=3e-3, max_inner=40000, max_outer=12)
    d = DEFAULT["single"]
    times = {}
    with grid.GridScene(scene, tuned=False) as gs:
        gs.predict(h, cfg.gravity)
        gs.fit_rotations()
        gs.assemble(h)

        def measure(iters, zcols, deal, upd):
            gs.set_schedule(iters)
            gs.set_shape(zcols, deal, upd)
            gs.solve(sp)  # captures the chunk graphs
            best = math.inf
            for _ in range(reps):
                t = time.perf_counter()
                gs.solve(sp)
                best = min(best, time.perf_counter() - t)
            key = "%d/%d/%d/%d" % (iters, zcols, deal, upd)
            times[key] = best
            if log:
                log("  %s: %.3f ms" % (key, 1e3 * best))
            return best

        shape = (d["zcols"], d["deal"], d["update_threads"])
        t_best = measure(d["iters"], *shape)
        for zc in ZCOLS:
            for dl in DEALS:
                if (zc, dl) == shape[:2]:
                    continue
                t = measure(d["iters"], zc, dl, shape[2])
                if t < 0.99 * t_best:
                    shape, t_best = (zc, dl, shape[2]), t
        for upd in UPDATE_THREADS:
            if upd == shape[2]:
                continue
            t = measure(d["iters"], shape[0], shape[1], upd)
            if t < 0.99 * t_best:
                shape, t_best = (shape[0], shape[1], upd), t
        iters = d["iters"]
        for it in iters_candidates:
            if it == d["iters"]:
                continue
            t = measure(it, *shape)
            if t < 0.99 * t_best:
                iters, t_best = it, t
    return {"iters": iters, "zcols": shape[0], "deal": shape[1], "update_threads": shape[2], "seconds": times}


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
    ap = argparse.ArgumentParser(description="measure the schedule and launch-shape table on this GPU")
    ap.add_argument("--out", default=TABLE)
    ap.add_argument("--quick", action="store_true", help="fewer batch candidates")
    ap.add_argument("--no-batch", action="store_true", help="keep the batch entries of the current table")
    args = ap.parse_args()

    def log(msg):
        print(msg, flush=True)

    old = load_table(args.out if os.path.exists(args.out) else TABLE)
    table = {"buckets": {k: [b if math.isfinite(b) else None for b in v] for k, v in BUCKETS.items()},
             "single": {}, "batch": dict(old.get("batch", {})) if args.no_batch else {}, "measured": {}}
    if args.no_batch:
        table["measured"].update({k: v for k, v in old.get("measured", {}).items() if k.startswith("batch_")})
    for b in range(len(BUCKETS["single"])):
        s = single_scene(b)
        log("single bucket %d (%s, %d REVs)" % (b, s.name, s.n))
        r = tune_single(s, log=log)
        table["single"][str(b)] = {k: r[k] for k in ("iters", "zcols", "deal", "update_threads")}
        table["measured"]["single_%d" % b] = r["seconds"]
        log("single bucket %d: iters %d, zcols %d, deal %d, update threads %d" %
            (b, r["iters"], r["zcols"], r["deal"], r["update_threads"]))
    if not args.no_batch:
        reps = {0: 8e3, 1: 6e4, 2: 5e5}
        for b, revs in reps.items():
            chunks = (8, 16, 64) if args.quick else (4, 8, 16, 32, 64)
            r = tune_batch(revs, chunk_candidates=chunks)
            table["batch"][str(b)] = {"chunk": r["chunk"], "iters": r["iters"]}
            table["measured"]["batch_%d" % b] = r["seconds"]
            log("batch bucket %d (%g REVs/scene): %s" % (b, revs, r))
    with open(args.out, "w") as f:
        json.dump(table, f, indent=1)
    log("wrote " + args.out)


if __name__ == "__main__":
    main()
