#!/usr/bin/env python3
"""Benchmark of the corotational elastic PCG solve on one MI355X.

Workload: the single large mixed-material scene (64^3 = 262,144 REVs, bottom
layer fixed, gravity, +0.2 N on 5% of the top layer; materials per 8^3
region, per-bond stiffness jitter), 1 ms steps with one substep and up to two
corotation passes. One bench step is one physics step of the solve part on
the device: predict, rotation fit, and per pass bond assembly + FP32-inner /
FP64-refined block-Jacobi PCG to the stated relative residual + apply, then
velocities.

value = DOF-PCG-iterations/s = sum over solves of 3 * active rows *
(inner iterations + outer FP64 residual sweeps) / wall time of the timed
steps. Prints one JSON line (rank 0, N = 1).
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "voxel-sim_amd"))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

SOLVE_TOL = 1e-9        # outer FP64 relative true residual
INNER_TOL = 3e-3        # FP32 residual drop between FP64 residual replacements
PEAK_HBM_GBS = 8000.0   # MI355X HBM3E
SPMV_BYTES_PER_ROW = 145  # z 12 + p_old 12 + D 24 + 3 bond blocks 72 + mask 1 + p_new 12 + Ap 12
# fracture_bonds per REV row: reads bond bits 1 + kinematic 1 + material 44 + rotation and assembly rotation 64 +
# position 24 + strength row 48 + per +axis bond (kscale 4, offset 12, plastic 12, damage 4, history 24) x 3 = 350;
# writes plastic 36 + damage 12 + history 72 = 120
FRACTURE_BYTES_PER_ROW = 470
UPDATE_BYTES_PER_ROW = 108  # p, Ap, r, d 48 + Dinv 24 read; d, r, z 36 written
SPMV_KERNEL = "pcg_bond_spmv<float>"
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "pcg_bond_spmv_64cube_traffic.txt")
TRAFFIC_FILE_LARGE = os.path.join(ROOT, "profiles", "pcg_bond_spmv_192x192x96_traffic.txt")
# Relative residual of the batched and served solves: a residual of 1e-6 leaves
# up to 1e-3 relative L2 error against the FP64 oracle on these boxes (condition
# numbers near 1e6); 1e-10 keeps it below 1e-6 (tests/test_gpu_configs.py).
BATCH_TOL = 1e-10
CUT_STEPS = 10          # steps of the cutting scene in the default line
COPY_BYTES = 1 << 30      # calibration copy for the HBM byte counters


def cpu_model() -> str:
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name"):
                return line.split(":", 1)[1].strip()
    except Exception:
        pass
    return platform.processor() or "unknown"


def host_threads() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except Exception:
        return os.cpu_count() or 1


def host_info() -> dict:
    """The host cores this process may use: all CPUs it is affine to, capped by
    OMP_NUM_THREADS where the environment sets it (a host shared by several
    GPUs gives each GPU's processes its share that way)."""
    affinity = host_threads()
    omp = os.environ.get("OMP_NUM_THREADS", "")
    threads = min(affinity, int(omp)) if omp.isdigit() and int(omp) > 0 else affinity
    return {"nproc": os.cpu_count(), "affinity_cpus": affinity, "omp_num_threads": omp or None,
            "threads_used": threads, "cpu_model": cpu_model()}


def oracle_system(O, scene, h, gravity):
    """The oracle's FP64 system of a scene's first solve: predict
    (src/solver.rs:500-523: v += g h, targeted axes at their targets with zero
    velocity, p += v h on free voxels), fit rotations, assemble. Returns the
    oracle scene (positions = predicted) and the packed rows."""
    os_ = O.OracleScene(scene)
    t = scene.targets.astype(np.float64)
    m = ~np.isnan(t)
    kin = scene.kinematic != 0
    v = os_.velocities + np.array(gravity) * h
    v[m] = 0.0
    os_.positions[~kin] += v[~kin] * h
    os_.positions[m] = t[m]
    O.fit_rotations(os_, O.F64)
    sy = O.assemble(os_, O.F64, h, os_.positions.copy(), scene.free_mask(), scene.loads)
    return os_, sy.rows()


def error_vs_oracle_first_solve(scene, sp) -> dict:
    """Relative L2 error of the workload's first device solve (the bench's
    solve parameters) against the oracle's FP64 solution of the same system
    (FP64 block-Jacobi CG to 1e-13), outside the timed region."""
    from oracle import oracle as O
    from vxsim import grid
    cfg = scene.config
    h = cfg.fixed_dt / cfg.substeps
    with grid.GridScene(scene) as gs:
        gs.predict(h, cfg.gravity)
        gs.fit_rotations()
        gs.assemble(h)
        rep = gs.solve(sp)
        x = gs.download()["correction"]
    _, rows = oracle_system(O, scene, h, cfg.gravity)
    t0 = time.perf_counter()
    x_o, it_o, rel_o = O.solve_exact(rows, 1e-13)
    xa = x[rows["active"]]
    return {"relative_l2": float(np.linalg.norm(xa - x_o) / np.linalg.norm(x_o)),
            "device_relative_residual": rep["relative_residual"], "oracle_relative_residual": rel_o,
            "oracle_iterations": it_o, "oracle_seconds": time.perf_counter() - t0,
            "what": "first solve of the workload vs the oracle's FP64 solution of the oracle-assembled system"}


def error_vs_oracle_steps(scene, result_positions, steps: int, step_cfg) -> float:
    """Relative L2 error of the displacement after `steps` device steps
    against the oracle's FP64 steps with exact solves."""
    from oracle import oracle as O
    os_ = O.OracleScene(scene)
    start = os_.positions.copy()
    for _ in range(steps):
        O.step(os_, step_cfg, O.F64, O.BJ, exact=True, exact_tolerance=1e-13, loads=scene.loads)
    d_o = os_.positions - start
    return float(np.linalg.norm((result_positions - start) - d_o) / np.linalg.norm(d_o))


def cpu_baseline(scene, budget_s: float = 12.0):
    """The oracle's restatement of the reference CPU PCG (src/solver.rs:
    1833-1911: f32 vectors, f64 dots, f32 alpha/beta, block-Jacobi in the
    Cholesky form) on the same 64^3 system, OpenMP over all host cores with
    no serial cutoff, capped at a bounded number of iterations."""
    from oracle import oracle as O
    os_ = O.OracleScene(scene)
    h = scene.config.fixed_dt / scene.config.substeps
    g = np.array(scene.config.gravity)
    kin = scene.kinematic != 0
    os_.positions[~kin] += g * h * h
    t0 = time.perf_counter()
    O.fit_rotations(os_, O.F32)
    sy = O.assemble(os_, O.F32, h, os_.positions.copy(), scene.free_mask(), scene.loads)
    assemble_s = time.perf_counter() - t0
    hi = host_info()
    threads = hi["threads_used"]
    results = {}
    for label, nthreads, iters in (("all", threads, 400), ("one", 1, 60)):
        _, it, res, sec = sy.solve_cpu(O.BJ, iters, 1e-30, nthreads, 0)
        results[label] = (it, sec, res)
        if sec > budget_s:
            break
    it, sec, _ = results["all"]
    value = 3.0 * sy.na * it / sec
    out = {"value": value, "unit": "DOF-PCG-iters/s", "cores": threads, "kind": "port",
           "sample": "oracle f32 CPU PCG with block-Jacobi on the 64^3 system (%d active rows), "
                     "%d iterations timed, OpenMP %d threads, no serial cutoff" % (sy.na, it, threads),
           "cpu_model": hi["cpu_model"], "host": hi, "assemble_s": assemble_s}
    if "one" in results:
        it1, sec1, _ = results["one"]
        out["single_thread_value"] = 3.0 * sy.na * it1 / sec1
    # The reference's CPU solve as configured by default (src/scene.rs:16-24,
    # src/solver.rs:1390-1496): block IC(0), the rayon cutoff of 32,768 rows,
    # to its default tolerance 1e-5; and block-Jacobi with the same settings.
    for label, pc in (("ic0", O.IC0), ("block_jacobi", O.BJ)):
        _, it_s, res_s, sec_s = sy.solve_cpu(pc, 4000, 1e-5, threads, 32768)
        out[label + "_to_1e-5"] = {"iterations": it_s, "seconds": sec_s, "relative_residual": res_s,
                                   "dof_iters_per_s": 3.0 * sy.na * it_s / sec_s, "threads": threads,
                                   "parallel_cutoff": 32768}
    return out


def stream_copy(nbytes: int, reps: int = 20) -> dict:
    """A STREAM copy on the device (vx_copy_bandwidth: the library's own
    16-byte-vector copy kernel, median of `reps` launches timed with HIP
    events). GB/s counts the bytes read plus written."""
    import ctypes
    from vxsim import lib
    from vxsim.solver import check
    nbytes = nbytes // 16 * 16
    us = ctypes.c_double()
    check(lib().vx_copy_bandwidth(nbytes, reps, ctypes.byref(us)))
    return {"bytes_read": nbytes, "us": us.value, "gbs": 2.0 * nbytes / (us.value * 1e-6) / 1e9}


def measured_traffic(path=TRAFFIC_FILE):
    """HBM bytes per SpMV launch from a committed rocprofv3 counter summary
    (FETCH_SIZE and WRITE_SIZE scaled by a 1 GiB copy), or None."""
    try:
        with open(path) as f:
            return json.load(f)["traffic_bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def spmv_in_solve(gs, scene, sp) -> dict:
    """Roofline of the FP32 bond SpMV as it runs inside a solve: one solve of
    the scene's assembled system with every pcg_bond_spmv and pcg_update
    dispatch timestamped (vx_scene_solve_timed: begin and end of each
    dispatch, the figures a kernel trace reports). achieved = algorithmic
    bytes per launch / mean dispatch duration."""
    t = gs.solve_timed(sp)
    st, up = t["spmv"], t["update"]
    b = SPMV_BYTES_PER_ROW * scene.n
    bu = UPDATE_BYTES_PER_ROW * scene.n
    achieved = b / (st["mean_us"] * 1e-6) / 1e9
    return {"bound": "hbm", "kernel": SPMV_KERNEL, "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": achieved / PEAK_HBM_GBS, "bytes_per_launch": b, "avg_us": st["mean_us"],
            "median_us": st["median_us"], "min_us": st["min_us"], "max_us": st["max_us"],
            "dispatches": st["dispatches"], "inner_iterations": t["inner_iterations"],
            "update_kernel": {"kernel": "pcg_update<float>", "bytes_per_launch": bu, "avg_us": up["mean_us"],
                              "achieved": bu / (up["mean_us"] * 1e-6) / 1e9,
                              "frac": bu / (up["mean_us"] * 1e-6) / 1e9 / PEAK_HBM_GBS},
            "timing": "mean over every pcg_bond_spmv dispatch of one solve (begin/end timestamps of each dispatch, "
                      "vx_scene_solve_timed), including the dispatches a chunk issues after the inner solve ends"}


def large_roofline() -> dict:
    """The same in-solve SpMV roofline on the 192x192x96 wood block (3.54 M
    rows, 513 MB per launch, far beyond the 256 MB Infinity Cache): the HBM
    figure proper."""
    from vxsim import grid, scenes
    scene = scenes.wood_block_blade()
    cfg = scene.config
    h = cfg.fixed_dt / cfg.substeps
    sp = grid.solve_params(tolerance=1e-9, inner_tolerance=INNER_TOL, max_inner=40000, max_outer=12)
    with grid.GridScene(scene) as gs:
        gs.predict(h, cfg.gravity)
        gs.fit_rotations()
        gs.assemble(h)
        gs.solve(sp)  # warm-up (first touch of the inner planes)
        out = spmv_in_solve(gs, scene, sp)
        # the same system with the multigrid preconditioner, beside block-Jacobi
        t = time.perf_counter()
        rbj = gs.solve(sp)
        tbj = time.perf_counter() - t
        mg_sp = grid.solve_params(tolerance=1e-9, inner_tolerance=INNER_TOL, max_inner=40000, max_outer=12,
                                  preconditioner=grid.MULTIGRID)
        gs.solve(mg_sp)  # builds the hierarchy, captures its graphs
        t = time.perf_counter()
        rmg = gs.solve(mg_sp)
        tmg = time.perf_counter() - t
    out["solve_block_jacobi"] = {"ms": 1000 * tbj, "inner": rbj["inner_iterations"], "sweeps": rbj["outer_sweeps"]}
    out["solve_multigrid"] = {"ms": 1000 * tmg, "inner": rmg["inner_iterations"], "sweeps": rmg["outer_sweeps"],
                              "rel_residual": rmg["relative_residual"], "time_vs_block_jacobi": tbj / tmg}
    out["workload"] = "%s: %d rows per launch, one solve to relative residual 1e-9" % (scene.name, scene.n)
    out["traffic"] = measured_traffic(TRAFFIC_FILE_LARGE)
    out["stream_copy_same_bytes"] = stream_copy(out["bytes_per_launch"] // 2)
    out["frac_of_copy"] = out["stream_copy_same_bytes"]["us"] / out["avg_us"]
    return out


def shard_slabs(depths=(24, 12)) -> dict:
    """One GPU's share of the 192x192x96 block under Z-slab decomposition (4
    and 8 GPUs): the 192x192xd slab solved on its own at its real shape, its
    in-solve SpMV roofline and its time per inner iteration (one solve to
    relative residual 1e-9, after a warm-up solve), beside the whole block's.
    The exchange a decomposition adds per iteration is not part of it."""
    from vxsim import grid, scenes
    out = {}
    for d in (96,) + tuple(depths):
        scene = scenes.wood_block_blade(dims=(192, 192, d))
        cfg = scene.config
        h = cfg.fixed_dt / cfg.substeps
        sp = grid.solve_params(tolerance=1e-9, inner_tolerance=INNER_TOL, max_inner=40000, max_outer=12)
        with grid.GridScene(scene) as gs:
            gs.predict(h, cfg.gravity)
            gs.fit_rotations()
            gs.assemble(h)
            gs.solve(sp)  # warm-up (first touch of the inner planes)
            t0 = time.perf_counter()
            rep = gs.solve(sp)
            wall = time.perf_counter() - t0
            r = spmv_in_solve(gs, scene, sp)
        its = rep["inner_iterations"] + rep["outer_sweeps"]
        out["192x192x%d" % d] = {
            "rows": scene.n, "gpus_sharing_the_block": 96 // d, "solve_ms": 1e3 * wall,
            "inner_iterations": rep["inner_iterations"], "outer_sweeps": rep["outer_sweeps"],
            "us_per_iteration": 1e6 * wall / its, "spmv_avg_us": r["avg_us"], "spmv_frac": r["frac"],
            "update_avg_us": r["update_kernel"]["avg_us"]}
    whole = out["192x192x96"]["us_per_iteration"]
    for d in depths:
        e = out["192x192x%d" % d]
        e["iteration_vs_whole_block_share"] = e["us_per_iteration"] / (whole / e["gpus_sharing_the_block"])
    return out


def batched(steps: int = 1, count: int = 256, preconditioner: int = 0, cpu: bool = False):
    """256 independent boxes (edges U[8, 48], one face fixed, random initial
    velocities) on one tile arena, `steps` steps each, all scenes advanced
    together tick by tick, solves to relative residual BATCH_TOL. Returns
    DOF-PCG-iters/s over the timed run; with `cpu`, the smallest and the
    largest box's step against the oracle's FP64 step."""
    from vxsim import batch, grid, scenes
    boxes = scenes.batch_boxes(count=count, seed=3, lo=8, hi=48)
    sp = grid.solve_params(tolerance=BATCH_TOL, inner_tolerance=INNER_TOL, max_inner=20000, max_outer=12,
                           preconditioner=preconditioner)
    tiles = sum(-(-s.n // batch.TILE) for s in boxes)
    ns = [s.n for s in boxes]
    probe = sorted({int(np.argmin(ns)), int(np.argmax(ns))})
    out = {}
    with batch.Batch(tiles * batch.TILE, max_scenes=count) as B:
        for timed in (False, True):
            slots = []
            t_admit = time.perf_counter()
            for s in boxes:
                c = s.config
                slots.append(B.admit(s, steps, grid.step_params(
                    fixed_dt=c.fixed_dt, gravity=c.gravity, substeps=c.substeps,
                    corotation_iterations=c.corotation_iterations, corotation_tolerance=c.corotation_tolerance,
                    solve=sp)))
            t_admit = time.perf_counter() - t_admit
            t0 = time.perf_counter()
            ticks = B.run()
            wall = time.perf_counter() - t0
            reps = [B.report(x) for x in slots]
            probed = {b: B.download(slots[b])["positions"] for b in probe}
            for x in slots:
                B.release(x)
        bad = [r for r in reps if r["state"] != batch.DONE]
        dof_iters = sum(3.0 * r["active_rows"] * (r["inner_iterations"] + r["outer_sweeps"]) for r in reps)
        out = {"workload": "batch_256_boxes: %d scenes, %d REVs, %d active rows, %d step(s) each, rel. residual %g"
                           % (count, sum(ns), sum(r["active_rows"] for r in reps), steps, BATCH_TOL),
               "dof_iters_per_s": dof_iters / wall, "wall_s": wall, "ticks": ticks, "ms_per_tick": 1000 * wall / ticks,
               "admit_s": t_admit, "failed": len(bad),
               "inner_iterations": sum(r["inner_iterations"] for r in reps),
               "max_rel_residual": max(r["last_relative_residual"] for r in reps)}
    if cpu and steps == 1:
        out["error_vs_fp64_oracle"] = {
            boxes[b].name: error_vs_oracle_steps(boxes[b], probed[b], steps, boxes[b].config) for b in probe}
    return out


def served(count: int, compare: bool = True, cpu: bool = True, preconditioner: int = 0, policy: str = "sjf") -> dict:
    """The served stream at its size mix, bounded: the first `count` arrivals
    of Poisson arrivals at 200/s, REVs log-uniform in [1e3, 2e6], aspect
    ratios U[0.5, 2], ten warm-started steps per request, solves to relative
    residual BATCH_TOL, continuous admission into an arena sized from a 200 GB
    HBM budget, scheduled by `policy` (vxsim.serve: "sjf" holds the larger
    scenes while the smaller ones run, "fifo" runs every admitted scene in
    every tick). p50/p99 latency (arrival to result on the host) and aggregate
    DOF-PCG-iters/s, beside the plain FP64 structured solve handling the same
    requests one at a time in arrival order; with `cpu`, the three smallest
    requests' ten steps against the oracle's FP64 steps."""
    from vxsim import batch, grid, serve
    sp = grid.solve_params(tolerance=BATCH_TOL, inner_tolerance=INNER_TOL, warm_start=True, max_inner=40000,
                           preconditioner=preconditioner)
    t = time.perf_counter()
    reqs = serve.request_stream(200.0, 60.0, 1e3, 2e6, 10, seed=4, max_requests=count, solve=sp)
    gen_s = time.perf_counter() - t
    with serve.Server.for_budget(200 << 30, max_scenes=256, keep_results=cpu, policy=policy) as S:
        t0 = time.perf_counter()
        S.run(reqs)
        wall = time.perf_counter() - t0
        ticks = S.ticks
    done = [r for r in reqs if r.state == batch.DONE]
    dof_iters = sum(3.0 * r.report["active_rows"] * (r.report["inner_iterations"] + r.report["outer_sweeps"])
                    for r in reqs if r.report)
    lat = serve.percentiles([r.latency for r in reqs])
    revs = [r.scene.n for r in reqs]
    out = {"workload": "first %d arrivals of the served stream (Poisson 200/s, REVs log-uniform [1e3, 2e6], "
                       "aspect U[0.5, 2], 10 warm-started steps each, rel. residual %g, 200 GB HBM budget)"
                       % (count, BATCH_TOL),
           "policy": policy, "p50_s": lat[50], "p99_s": lat[99], "requests": len(reqs), "done": len(done),
           "revs": revs, "total_revs": int(sum(revs)), "dof_iters_per_s": dof_iters / wall, "wall_s": wall,
           "ticks": ticks, "scene_generation_s": gen_s,
           "max_rel_residual": max(r.report["last_relative_residual"] for r in reqs if r.report)}
    if compare:
        t0 = time.perf_counter()
        base = serve.serve_alone_fp64(reqs)
        bwall = time.perf_counter() - t0
        bl = serve.percentiles([b["latency"] for b in base])
        bdi = sum(b["dof_iters"] for b in base) / bwall
        out["plain_fp64_one_at_a_time"] = {"p50_s": bl[50], "p99_s": bl[99], "wall_s": bwall, "dof_iters_per_s": bdi,
                                           "failed": sum(1 for b in base if not b["ok"])}
        out["served_vs_plain_fp64_dof_iters"] = out["dof_iters_per_s"] / bdi
        out["served_vs_plain_fp64_wall"] = bwall / wall
    if cpu:
        errs = {}
        for r in sorted(reqs, key=lambda q: q.scene.n)[:3]:
            cfg = r.scene.config
            cfg.corotation_iterations = r.params.corotation_iterations
            cfg.corotation_tolerance = r.params.corotation_tolerance
            cfg.gravity = tuple(r.params.gravity)
            errs["request%d_%drevs" % (r.rid, r.scene.n)] = error_vs_oracle_steps(r.scene, r.result["positions"],
                                                                                 r.steps, cfg)
        out["error_vs_fp64_oracle"] = errs
    return out


def serve_bench(rate: float, seconds: float, rev_lo: float, rev_hi: float, steps: int, budget_gb: float,
                compare: bool = True, preconditioner: int = 0):
    """Served request stream: Poisson arrivals, boxes log-uniform in REVs,
    `steps` warm-started steps each, admission against an HBM budget.
    Prints one JSON line with p50/p99 latency and aggregate DOF-PCG-iters/s,
    beside the plain FP64 solve handling the same requests one at a time."""
    from vxsim import batch, grid, serve
    t = time.perf_counter()
    reqs = serve.request_stream(rate, seconds, rev_lo, rev_hi, steps, seed=4,
                                solve=grid.solve_params(tolerance=BATCH_TOL, inner_tolerance=INNER_TOL, warm_start=True,
                                                        max_inner=40000, preconditioner=preconditioner))
    gen_s = time.perf_counter() - t
    budget = int(budget_gb * (1 << 30))
    last = [time.perf_counter()]

    def progress(server):
        now = time.perf_counter()
        if now - last[0] >= 10.0:
            last[0] = now
            print("serve: %d ticks, %d of %d requests finished" %
                  (server.ticks, sum(1 for r in reqs if r.finished is not None), len(reqs)), file=sys.stderr, flush=True)

    with serve.Server.for_budget(budget, max_scenes=256) as S:
        t0 = time.perf_counter()
        S.run(reqs, on_tick=progress)
        wall = time.perf_counter() - t0
        ticks = S.ticks
    done = [r for r in reqs if r.state == batch.DONE]
    dof_iters = sum(3.0 * r.report["active_rows"] * (r.report["inner_iterations"] + r.report["outer_sweeps"])
                    for r in reqs if r.report)
    lat = serve.percentiles([r.latency for r in reqs])
    revs = [r.scene.n for r in reqs]
    line = {
        "metric": "served p50/p99 solve latency", "unit": "s", "higher_is_better": False,
        "p50": lat[50], "p99": lat[99], "requests": len(reqs), "done": len(done),
        "preconditioner": "red_black" if preconditioner else "block_jacobi",
        "dof_iters_per_s": dof_iters / wall, "wall_s": wall, "ticks": ticks,
        "offered": {"rate_per_s": rate, "seconds": seconds, "revs_log_uniform": [rev_lo, rev_hi],
                    "mean_revs": float(np.mean(revs)), "steps_per_request": steps,
                    "rev_steps_per_s": rate * steps * float(np.mean(revs))},
        "served_rev_steps_per_s": sum(r.scene.n * r.steps for r in done) / wall,
        "hbm_budget_gb": budget_gb, "scene_generation_s": gen_s,
        "data": "synthetic", "dtype": "f32 inner / f64 refinement and reductions",
    }
    if compare:
        t0 = time.perf_counter()
        base = serve.serve_alone_fp64(reqs)
        bwall = time.perf_counter() - t0
        bl = serve.percentiles([b["latency"] for b in base])
        line["plain_fp64_one_at_a_time"] = {"p50": bl[50], "p99": bl[99], "wall_s": bwall,
                                             "dof_iters_per_s": sum(b["dof_iters"] for b in base) / bwall,
                                             "failed": sum(1 for b in base if not b["ok"])}
    print(json.dumps(line))


def red_black_steps(scene, step_kw: dict, steps: int, warmup: int, ms_block_jacobi: float,
                    preconditioner: int = 1) -> dict:
    """The same steps from the same initial scene with another preconditioner
    (red-black by default; its wall time per step beside block-Jacobi's)."""
    from vxsim import grid
    sp = step_kw["solve"]
    kw = dict(step_kw, solve=grid.solve_params(tolerance=sp.tolerance, inner_tolerance=sp.inner_tolerance,
                                               max_inner=sp.max_inner, max_outer=sp.max_outer,
                                               preconditioner=preconditioner))
    with grid.GridScene(scene) as gs:
        for _ in range(warmup):
            gs.step(**kw)
        reps = []
        t0 = time.perf_counter()
        for _ in range(steps):
            reps.append(gs.step(**kw))
        wall = time.perf_counter() - t0
    ms = 1000.0 * wall / steps
    return {"ms_per_step": ms, "inner_iterations": sum(r["inner_iterations"] for r in reps),
            "solves": sum(r["solves"] for r in reps), "time_vs_block_jacobi": ms_block_jacobi / ms}


def large_bench(steps: int, warmup: int, compare: bool = True):
    """The 192x192x96 wood block (3.54 M REVs, out of the 256 MB Infinity
    Cache): whole steps, the SpMV roofline at that size, and the plain FP64
    solve of the same system. Prints one JSON line."""
    from vxsim import grid, scenes
    scene = scenes.wood_block_blade()
    cfg = scene.config
    sp = grid.solve_params(tolerance=1e-6, inner_tolerance=INNER_TOL, max_inner=20000, max_outer=12)
    step_kw = dict(fixed_dt=cfg.fixed_dt, gravity=cfg.gravity, substeps=cfg.substeps,
                   corotation_iterations=cfg.corotation_iterations, corotation_tolerance=cfg.corotation_tolerance,
                   solve=sp)
    active = scene.active_count()
    with grid.GridScene(scene) as gs:
        for _ in range(warmup):
            gs.step(**step_kw)
        reps = []
        t0 = time.perf_counter()
        for _ in range(steps):
            reps.append(gs.step(**step_kw))
        wall = time.perf_counter() - t0
        iters = sum(r["inner_iterations"] + r["outer_sweeps"] for r in reps)
        gs.assemble(cfg.fixed_dt / cfg.substeps)
        spmv_us = gs.time_spmv(50)
        extra = {}
        if compare:
            h = cfg.fixed_dt / cfg.substeps
            gs.predict(h, cfg.gravity)
            gs.fit_rotations()
            gs.assemble(h)
            t = time.perf_counter()
            rm = gs.solve(tolerance=1e-6, inner_tolerance=INNER_TOL, max_inner=20000)
            tm = time.perf_counter() - t
            t = time.perf_counter()
            r64 = gs.solve(tolerance=1e-6, max_inner=40000, mode=grid.F64)
            t64 = time.perf_counter() - t
            rb_sp = grid.solve_params(tolerance=1e-6, inner_tolerance=INNER_TOL, max_inner=20000,
                                      preconditioner=grid.RED_BLACK)
            gs.solve(rb_sp)
            t = time.perf_counter()
            rrb = gs.solve(rb_sp)
            trb = time.perf_counter() - t
            # The drop-in over the reference row packing at this size (host
            # buffers and PCIe upload included), the reference's f32 semantics.
            rows = gs.export_rows()
            inv = np.linalg.inv(rows["diagonal"].reshape(-1, 3, 3)).reshape(-1, 9)
            row_args = (np.ascontiguousarray(rows["rhs"], np.float32),
                        np.ascontiguousarray(rows["neighbors"], np.uint32),
                        np.ascontiguousarray(rows["diagonal"], np.float32),
                        np.ascontiguousarray(rows["blocks"], np.float32), np.ascontiguousarray(inv, np.float32))
            del rows, inv
            from vxsim import solver
            try:
                solver.solve_pcg(*row_args, 20000, 1e-5)
                t = time.perf_counter()
                _, it32, res32 = solver.solve_pcg(*row_args, 20000, 1e-5)
                t32 = time.perf_counter() - t
                drop32 = {"ms": 1000 * t32, "iters": it32, "relative_residual": res32, "tolerance": 1e-5,
                          "dof_iters_per_s": 3.0 * active * it32 / t32}
            except solver.SolverError as e:
                drop32 = {"status": e.status, "error": str(e)}
            del row_args
            extra = {"dropin_rows_f32": drop32,
                     "mixed_solve": {"ms": 1000 * tm, "inner": rm["inner_iterations"], "sweeps": rm["outer_sweeps"],
                                     "dof_iters_per_s": 3.0 * active * (rm["inner_iterations"] + rm["outer_sweeps"]) / tm},
                     "plain_fp64_structured": {"ms": 1000 * t64, "iters": r64["inner_iterations"],
                                               "dof_iters_per_s": 3.0 * active * (r64["inner_iterations"] +
                                                                                   r64["outer_sweeps"]) / t64},
                     "time_to_solution_vs_fp64": t64 / tm,
                     "red_black_solve": {"ms": 1000 * trb, "inner": rrb["inner_iterations"],
                                         "sweeps": rrb["outer_sweeps"], "rel_residual": rrb["relative_residual"],
                                         "time_vs_block_jacobi": tm / trb}}
    b = SPMV_BYTES_PER_ROW * scene.n
    line = {"metric": "DOF-PCG-iters/s", "value": 3.0 * active * iters / wall, "unit": "DOF-PCG-iters/s",
            "workload": "%s: %d REVs, %d active rows, %d steps, rel. residual 1e-6" % (scene.name, scene.n, active, steps),
            "ms_per_step": 1000 * wall / steps, "solves": sum(r["solves"] for r in reps), "iterations": iters,
            "roofline": {"bound": "hbm", "kernel": SPMV_KERNEL, "bytes_per_launch": b, "avg_us": spmv_us,
                         "achieved": b / (spmv_us * 1e-6) / 1e9, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                         "frac": b / (spmv_us * 1e-6) / 1e9 / PEAK_HBM_GBS,
                         "stream_copy": {"1GiB": stream_copy(COPY_BYTES), "same_bytes": stream_copy(b // 2)}},
            "data": "synthetic", "dtype": "f32 inner / f64 refinement and reductions"}
    line.update(extra)
    print(json.dumps(line))


def cut_contact(scene, gs, prev_pos, targets, cpu: bool) -> dict:
    """collision::project on the cut's own state after its last step: the
    device positions (f64, rounded to the reference's f32), the positions at
    the step's start as `previous`, the lattice centres as rest positions, the
    bonds left after fracture and the blade's axis targets."""
    from vxsim import contact
    cur = gs.download()["positions"].astype(np.float32)
    bonds = gs.download_bonds()["neighbor"]
    rest = scene.positions.astype(np.float32)
    held = ~np.isnan(targets)
    args = dict(positions=cur, previous=prev_pos.astype(np.float32), rest_positions=rest,
                mass=scene.mass.astype(np.float32), kinematic=scene.kinematic != 0, bond_neighbor=bonds)
    runs = []
    for k in range(2):
        runs.append(contact.project(**args, targets=held))
        print("cut contact: device call %d: %.1f ms, %d pairs, %d contacts" %
              (k, runs[-1][3]["device_us"] / 1e3, runs[-1][3]["pairs"], runs[-1][3]["contacts"]),
              file=sys.stderr, flush=True)
    p, q, c, rep = runs[-1]
    out = {"voxels": scene.n, "broken_bonds_so_far": int(3 * scene.n - np.count_nonzero(bonds != 0xFFFFFFFF)),
           "candidate_pairs": rep["candidate_pairs"], "pairs": rep["pairs"], "contacts": rep["contacts"],
           "passes": rep["passes"], "relaxed_passes": rep["relaxed_passes"],
           "relaxation_rounds": rep["relaxation_rounds"], "windows": rep["rounds"],
           "pair_evaluations": rep["evaluated"], "device_ms": rep["device_us"] / 1e3,
           "projection_ms": rep["projection_us"] / 1e3}
    # the same call with every pass by the window walker
    contact.set_method(contact.WINDOWS)
    try:
        wp, wq, wc, wrep = contact.project(**args, targets=held)
    finally:
        contact.set_method(contact.AUTO)
    out["windows_method"] = {
        "device_ms": wrep["device_us"] / 1e3, "windows": wrep["rounds"], "pair_evaluations": wrep["evaluated"],
        "same_result": bool(np.array_equal(wp.view(np.uint32), p.view(np.uint32))
                            and np.array_equal(wq.view(np.uint32), q.view(np.uint32))
                            and all(np.array_equal(wc[k], c[k]) for k in ("i", "j", "new_impact")))}
    print("cut contact: windows method %.1f ms, same result %s" %
          (out["windows_method"]["device_ms"], out["windows_method"]["same_result"]), file=sys.stderr, flush=True)
    if cpu:
        from oracle import oracle
        print("cut contact: the oracle on one host thread", file=sys.stderr, flush=True)
        os.environ["VXO_PROGRESS"] = "1"
        t0 = time.perf_counter()
        po, qo, co, _ = oracle.contact_project(**args, targets=held)
        out["cpu_oracle_ms"] = 1000 * (time.perf_counter() - t0)
        print("cut contact: oracle %.1f s" % (out["cpu_oracle_ms"] / 1e3), file=sys.stderr, flush=True)
        out["bit_identical_to_oracle"] = bool(
            np.array_equal(p.view(np.uint32), po.view(np.uint32)) and np.array_equal(q.view(np.uint32), qo.view(np.uint32))
            and all(np.array_equal(c[k], co[k]) for k in ("i", "j", "new_impact")))
    return out


def multigrid_bench(reps: int = 3) -> None:
    """Solves of the 64^3 mixed box's first assembled system with the
    multigrid preconditioner (the hierarchy built by the first), beside one
    block-Jacobi solve; one JSON line (the run the multigrid kernel trace
    profiles)."""
    from vxsim import grid, scenes
    scene = scenes.large_64cube()
    cfg = scene.config
    h = cfg.fixed_dt / cfg.substeps
    with grid.GridScene(scene) as gs:
        gs.predict(h, cfg.gravity)
        gs.fit_rotations()
        gs.assemble(h)
        runs = []
        for prec in [grid.MULTIGRID] * reps + [grid.BLOCK_JACOBI]:
            sp = grid.solve_params(tolerance=SOLVE_TOL, inner_tolerance=INNER_TOL, max_inner=8000, max_outer=12,
                                   preconditioner=prec)
            t = time.perf_counter()
            r = gs.solve(sp)
            runs.append({"preconditioner": "multigrid" if prec == grid.MULTIGRID else "block_jacobi",
                         "ms": 1000 * (time.perf_counter() - t), "inner": r["inner_iterations"],
                         "sweeps": r["outer_sweeps"], "rel_residual": r["relative_residual"]})
    print(json.dumps({"workload": "large_64x64x64_mixed: first assembled system, rel. residual %g" % SOLVE_TOL,
                      "solves": runs}))


def cut_bench(steps: int, dims=(192, 192, 96), preconditioner: int = 0, contact: bool = False, cpu: bool = True,
              speed: float = 5.0):
    print(json.dumps(cut_run(steps, dims, preconditioner, contact, cpu, speed)))


def cut_run(steps: int, dims=(192, 192, 96), preconditioner: int = 0, contact: bool = False, cpu: bool = True,
            speed: float = 5.0, quiet: bool = False) -> dict:
    """The cutting scene with fracture: the 192x192x96 wood block (bottom
    layer fixed, gravity) and the blade as a rigid tool, Tool::new's tapered
    40x2x20 blade (realised with a 5-REV spine) of iron, mounted on a guide
    along -Z and launched at `speed` m/s onto the top face; every step
    advances the tool per substep (and its zero-time pass), runs up to six
    corotation passes and the fracture pass (broken bonds leave the lattice
    before the next assembly). Prints one JSON line with the step time, the
    solve throughput, the bonds broken and the tool's track per step (quiet:
    no line per step)."""
    from vxsim import grid, scenes
    blade = (40, 2, 20)
    scene = scenes.wood_block(dims=dims)
    tool = scenes.blade_tool(scene, blade, speed=speed, mounted=True)
    cfg = scene.config
    sp = grid.solve_params(tolerance=1e-6, inner_tolerance=INNER_TOL, max_inner=40000, max_outer=12,
                           preconditioner=preconditioner)
    reps, walls, track = [], [], []
    contact_out = None
    no_targets = np.full((scene.n, 3), np.nan, np.float32)
    with grid.GridScene(scene) as gs:
        gs.set_tool(tool)
        for k in range(steps):
            if contact and k == steps - 1:
                prev_pos = gs.download()["positions"]
            t = time.perf_counter()
            r = gs.step(fixed_dt=cfg.fixed_dt, gravity=cfg.gravity, substeps=cfg.substeps,
                        corotation_iterations=cfg.corotation_iterations, corotation_tolerance=cfg.corotation_tolerance,
                        solve=sp, fracture=scenes.FractureConfig())
            walls.append(time.perf_counter() - t)
            st = gs.tool_state()
            r["active"] = gs.active
            reps.append(r)
            track.append({"z": round(float(st["position"][2]), 6), "vz": round(float(st["velocity"][2]), 4),
                          "contacts": int(st["contacts"]), "microsteps": int(st["microsteps"]),
                          "candidates": int(st["candidates"])})
            if not quiet:
                print("cut step %d: %.1f ms, %d solves, %d inner iterations, %d bonds broken, fracture pass %.3f ms, "
                      "tool z %.5f vz %.3f, %d contacts" %
                      (k, 1000 * walls[-1], r["solves"], r["inner_iterations"], r["fracture"]["bonds_broken"],
                       1000 * r["fracture"]["seconds"], st["position"][2], st["velocity"][2], st["contacts"]),
                      file=sys.stderr, flush=True)
        if contact:
            contact_out = cut_contact(scene, gs, prev_pos, no_targets, cpu)
    dof_iters = sum(3.0 * r["active"] * (r["inner_iterations"] + r["outer_sweeps"]) for r in reps)
    line = {"metric": "DOF-PCG-iters/s", "value": dof_iters / sum(walls), "unit": "DOF-PCG-iters/s",
            "workload": "%s with fracture: %d REVs, rigid blade tool %dx%dx%d (%d cells, %.2f kg) on a -Z guide "
                        "launched at %g m/s, %d steps of 1 ms, <=%d passes, rel. residual 1e-6" %
                        (scene.name, scene.n, blade[0], blade[1], blade[2], len(tool.cells), tool.mass, speed, steps,
                         cfg.corotation_iterations),
            "preconditioner": {0: "block_jacobi", 1: "red_black", 2: "multigrid"}[preconditioner],
            "mean_ms_per_step": 1000 * sum(walls) / len(walls),
            "ms_per_step": [round(1000 * w, 2) for w in walls], "solves": [r["solves"] for r in reps],
            "inner_iterations": [r["inner_iterations"] for r in reps],
            "bonds_broken": [int(r["fracture"]["bonds_broken"]) for r in reps],
            "tool": track,
            "fracture_pass": {"kernels": "fracture_bonds + reduce_fracture",
                              "ms_per_step": [round(1000 * r["fracture"]["seconds"], 4) for r in reps],
                              "bytes_per_row": FRACTURE_BYTES_PER_ROW,
                              "achieved_gbs": FRACTURE_BYTES_PER_ROW * scene.n / max(
                                  sum(r["fracture"]["seconds"] for r in reps) / len(reps), 1e-12) / 1e9,
                              "peak_gbs": PEAK_HBM_GBS},
            "dissipated_j": {e: sum(r["fracture"][e] for r in reps) for e in
                             ("plastic_heat", "stored_plastic", "surface_work", "damage_heat", "released")},
            "data": "synthetic", "dtype": "f32 inner / f64 refinement, reductions, assembly, fracture and tool"}
    if contact_out is not None:
        line["contact_projection_after_last_step"] = contact_out
    return line


def contact_bench(dims=(64, 64, 64), reps: int = 3, cpu: bool = True) -> dict:
    """Contact projection (collision::project) on the device against the oracle:
    a lattice box with 2% of its bonds broken, a fixed floor, per-voxel jitter,
    fast movers and targets (vxsim.scenes.contact_box). Device time with the
    inputs resident (transfers excluded), median of `reps` calls; the oracle's
    single-thread time for one call on the same input, and whether every
    position, previous position and contact matched bit for bit."""
    from vxsim import contact, scenes
    s = scenes.contact_box(dims, seed=2)
    targets = s.pop("targets")
    contact.set_method(contact.AUTO)
    runs = [contact.project(**s, targets=targets) for _ in range(reps)]
    p, q, c, rep = runs[-1]
    dev_us = sorted(r[3]["device_us"] for r in runs)[reps // 2]
    proj_us = sorted(r[3]["projection_us"] for r in runs)[reps // 2]
    out = {"workload": "contact_box %dx%dx%d seed 2: %d voxels" % (dims + (len(p),)),
           "candidate_pairs": rep["candidate_pairs"], "pairs": rep["pairs"], "contacts": rep["contacts"],
           "passes": rep["passes"], "relaxed_passes": rep["relaxed_passes"],
           "relaxation_rounds": rep["relaxation_rounds"], "windows": rep["rounds"], "device_ms": dev_us / 1e3,
           "projection_ms": proj_us / 1e3, "broad_phase_ms": (dev_us - proj_us) / 1e3,
           "pair_evaluations": rep["evaluated"], "pair_evaluations_per_s": rep["evaluated"] / (proj_us * 1e-6)}
    # the same call with every pass by the window walker (identical results)
    contact.set_method(contact.WINDOWS)
    try:
        wr = [contact.project(**s, targets=targets) for _ in range(reps)]
    finally:
        contact.set_method(contact.AUTO)
    wp, wq, wc, wrep = wr[-1]
    w_us = sorted(r[3]["device_us"] for r in wr)[reps // 2]
    out["windows_method"] = {"device_ms": w_us / 1e3, "windows": wrep["rounds"],
                             "us_per_window": sorted(r[3]["projection_us"] for r in wr)[reps // 2] / max(wrep["rounds"], 1),
                             "pair_evaluations": wrep["evaluated"], "passes_redone": wrep["redone"],
                             "relaxation_speedup": w_us / dev_us,
                             "same_result": bool(np.array_equal(wp.view(np.uint32), p.view(np.uint32))
                                                 and np.array_equal(wq.view(np.uint32), q.view(np.uint32))
                                                 and np.array_equal(wc["i"], c["i"]) and np.array_equal(wc["j"], c["j"]))}
    if cpu:
        from oracle import oracle
        t0 = time.perf_counter()
        po, qo, co, st = oracle.contact_project(**s, targets=targets)
        t_cpu = time.perf_counter() - t0
        same = (np.array_equal(p.view(np.uint32), po.view(np.uint32))
                and np.array_equal(q.view(np.uint32), qo.view(np.uint32))
                and all(np.array_equal(c[k], co[k]) for k in ("i", "j", "new_impact"))
                and np.array_equal(c["normal"].view(np.uint32), co["normal"].view(np.uint32)))
        out["cpu_oracle"] = {"ms": 1000 * t_cpu, "cores": 1, "kind": "port",
                             "device_speedup": 1000 * t_cpu / (dev_us / 1e3)}
        out["bit_identical_to_oracle"] = bool(same)
    return out


def pmc_driver(large: bool = False, red_black: bool = False):
    """Short run for the rocprofv3 counter passes: 40 inner iterations of the
    solve of the assembled 64^3 system (or the 192x192x96 block), block-Jacobi
    (pcg_bond_spmv, pcg_update) or with red_black the red-black reduced
    system (pcg_rb_red, its black SpMV and update), then 3 copies of 1 GiB by
    the library's copy kernel (the calibration of FETCH_SIZE / WRITE_SIZE
    against known bytes)."""
    import ctypes
    from vxsim import grid, lib, scenes
    scene = scenes.wood_block_blade() if large else scenes.large_64cube()
    cfg = scene.config
    with grid.GridScene(scene) as gs:
        h = cfg.fixed_dt / cfg.substeps
        gs.predict(h, cfg.gravity)
        gs.fit_rotations()
        gs.assemble(h)
        rep = gs.solve(grid.solve_params(tolerance=1e-14, inner_tolerance=1e-14, max_inner=40, max_outer=1,
                                         preconditioner=grid.RED_BLACK if red_black else grid.BLOCK_JACOBI),
                       raise_on_error=False)
    out_us = ctypes.c_double()
    lib().vx_copy_bandwidth(COPY_BYTES, 3, ctypes.byref(out_us))
    print(json.dumps({"inner_iterations": rep["inner_iterations"], "copy_bytes": COPY_BYTES, "rows": scene.n,
                      "spmv_bytes": SPMV_BYTES_PER_ROW * scene.n}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-compare", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--pmc-driver", action="store_true", help="short SpMV + copy run for counter passes")
    ap.add_argument("--serve", action="store_true", help="served request stream instead of the step benchmark")
    ap.add_argument("--large", action="store_true", help="192x192x96 block instead of the 64^3 box")
    ap.add_argument("--cut", action="store_true", help="192x192x96 block cut by the blade tool with fracture")
    ap.add_argument("--cut-speed", type=float, default=5.0, help="--cut: the blade's launch speed (m/s)")
    ap.add_argument("--red-black", action="store_true", help="--cut / --serve with the red-black preconditioner")
    ap.add_argument("--multigrid", action="store_true",
                    help="--cut with the aggregation multigrid preconditioner; alone: multigrid solves of the 64^3 system")
    ap.add_argument("--contact", action="store_true", help="contact projection leg only (64^3 contact box)")
    ap.add_argument("--no-contact", action="store_true")
    ap.add_argument("--cut-contact", action="store_true",
                    help="with --cut: the contact projection on the state after the last step, against the oracle")
    ap.add_argument("--serve-rate", type=float, default=20.0)
    ap.add_argument("--serve-seconds", type=float, default=10.0)
    ap.add_argument("--serve-revs", type=str, default="1e3,2e5")
    ap.add_argument("--serve-steps", type=int, default=10)
    ap.add_argument("--serve-budget-gb", type=float, default=200.0)
    ap.add_argument("--serve-count", type=int, default=12,
                    help="requests of the default served leg (the first arrivals of the 200/s stream)")
    ap.add_argument("--no-serve", action="store_true", help="skip the default served leg")
    ap.add_argument("--no-large-roofline", action="store_true", help="skip the 192x192x96 in-solve SpMV roofline")
    ap.add_argument("--no-cut", action="store_true", help="skip the default line's %d-step cut leg" % CUT_STEPS)
    ap.add_argument("--direct-launch", action="store_true",
                    help="launch the PCG chunks kernel by kernel instead of replaying captured hipGraphs "
                         "(identical results; for dispatch tracing)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path (independent scenes shard "
              "without communication; a single scene would need a slab decomposition). Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)
    if args.direct_launch:
        from vxsim import grid
        grid.set_launch_mode(True)
    if args.pmc_driver:
        pmc_driver(args.large, red_black=args.red_black)
        return
    if args.contact:
        print(json.dumps({"contact_projection": contact_bench(cpu=not args.no_cpu_baseline)}))
        return
    if args.multigrid and not args.cut:
        multigrid_bench()
        return
    if args.cut:
        cut_bench(args.steps, preconditioner=2 if args.multigrid else (1 if args.red_black else 0),
                  contact=args.cut_contact, speed=args.cut_speed, cpu=not args.no_cpu_baseline)
        return
    if args.large:
        large_bench(args.steps, args.warmup, compare=not args.no_compare)
        return
    if args.serve:
        lo, hi = (float(x) for x in args.serve_revs.split(","))
        serve_bench(args.serve_rate, args.serve_seconds, lo, hi, args.serve_steps, args.serve_budget_gb,
                    compare=not args.no_compare, preconditioner=1 if args.red_black else 0)
        return

    from vxsim import grid, scenes, solver

    scene = scenes.large_64cube()
    cfg = scene.config
    h = cfg.fixed_dt / cfg.substeps
    sp = grid.solve_params(tolerance=SOLVE_TOL, inner_tolerance=INNER_TOL, max_inner=8000, max_outer=12)
    gs = grid.GridScene(scene)
    step_kw = dict(fixed_dt=cfg.fixed_dt, gravity=cfg.gravity, substeps=cfg.substeps,
                   corotation_iterations=cfg.corotation_iterations, corotation_tolerance=cfg.corotation_tolerance,
                   solve=sp)
    for _ in range(args.warmup):
        gs.step(**step_kw)
    reps = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        reps.append(gs.step(**step_kw))
    wall = time.perf_counter() - t0
    active = scene.active_count()
    iters = sum(r["inner_iterations"] + r["outer_sweeps"] for r in reps)
    solves = sum(r["solves"] for r in reps)
    value = 3.0 * active * iters / wall
    ms_per_step = 1000.0 * wall / args.steps
    print("bench: %d steps of the 64^3 box, %.2f ms per step, %.3e DOF-PCG-iters/s" % (args.steps, ms_per_step, value),
          file=sys.stderr, flush=True)

    # Roofline of the dominant kernel, the FP32 structured bond SpMV, as it
    # runs inside a solve of the next step's assembled system (after the timed
    # steps).
    gs.predict(h, cfg.gravity)
    gs.fit_rotations()
    gs.assemble(h)
    roof = spmv_in_solve(gs, scene, sp)
    roof["traffic"] = measured_traffic(TRAFFIC_FILE)
    roof["cache_resident"] = ("the 64^3 planes (about 60 MB) stay in the 256 MB Infinity Cache between kernels: a "
                              "latency-bound launch of 512 workgroups; roofline_hbm_out_of_cache is the HBM figure")
    roof["isolated_back_to_back_us"] = gs.time_spmv(200)
    # What a plain copy reaches on this device: at 1 GiB (out of every cache)
    # and with the SpMV's own byte count (half read, half written).
    roof["stream_copy"] = {"1GiB": stream_copy(COPY_BYTES), "same_bytes": stream_copy(roof["bytes_per_launch"] // 2)}

    extra = {}
    if not args.no_compare:
        # Plain FP64 GPU solves of the same assembled system: the structured
        # FP64 PCG and the drop-in FP64 solve over the reference row packing
        # (host buffers, PCIe upload included).
        gs.predict(h, cfg.gravity)
        gs.fit_rotations()
        gs.assemble(h)
        t = time.perf_counter()
        rm = gs.solve(tolerance=SOLVE_TOL, inner_tolerance=INNER_TOL, max_inner=8000, max_outer=12)
        tm = time.perf_counter() - t
        t = time.perf_counter()
        r64 = gs.solve(tolerance=SOLVE_TOL, max_inner=20000, mode=grid.F64)
        t64 = time.perf_counter() - t
        # The same system to the reference's default tolerance (beside the
        # CPU solves to 1e-5 in cpu_baseline).
        t = time.perf_counter()
        r5 = gs.solve(tolerance=1e-5, inner_tolerance=INNER_TOL, max_inner=8000, max_outer=12)
        t5 = time.perf_counter() - t
        # The red-black reduced-system preconditioner on the same system.
        rb_sp = grid.solve_params(tolerance=SOLVE_TOL, inner_tolerance=INNER_TOL, max_inner=8000, max_outer=12,
                                  preconditioner=grid.RED_BLACK)
        gs.solve(rb_sp)  # captures its graph
        t = time.perf_counter()
        rrb = gs.solve(rb_sp)
        trb = time.perf_counter() - t
        rb_sp5 = grid.solve_params(tolerance=1e-5, inner_tolerance=INNER_TOL, max_inner=8000, max_outer=12,
                                   preconditioner=grid.RED_BLACK)
        t = time.perf_counter()
        rrb5 = gs.solve(rb_sp5)
        trb5 = time.perf_counter() - t
        # Aggregation multigrid (one V-cycle per iteration) on the same system.
        mg_sp = grid.solve_params(tolerance=SOLVE_TOL, inner_tolerance=INNER_TOL, max_inner=8000, max_outer=12,
                                  preconditioner=grid.MULTIGRID)
        gs.solve(mg_sp)  # builds the hierarchy, captures its graphs
        t = time.perf_counter()
        rmg = gs.solve(mg_sp)
        tmg = time.perf_counter() - t
        mg_sp5 = grid.solve_params(tolerance=1e-5, inner_tolerance=INNER_TOL, max_inner=8000, max_outer=12,
                                   preconditioner=grid.MULTIGRID)
        t = time.perf_counter()
        rmg5 = gs.solve(mg_sp5)
        tmg5 = time.perf_counter() - t
        mixed_rate = 3.0 * active * (rm["inner_iterations"] + rm["outer_sweeps"]) / tm
        f64_rate = 3.0 * active * (r64["inner_iterations"] + r64["outer_sweeps"]) / t64
        rows = gs.export_rows()
        inv = np.linalg.inv(rows["diagonal"].reshape(-1, 3, 3)).reshape(-1, 9)
        # The reference's row arrays as the caller holds them (f32, u32), built
        # before the timed calls so that only the library call is timed.
        row_args = (np.ascontiguousarray(rows["rhs"], np.float32), np.ascontiguousarray(rows["neighbors"], np.uint32),
                    np.ascontiguousarray(rows["diagonal"], np.float32), np.ascontiguousarray(rows["blocks"], np.float32),
                    np.ascontiguousarray(inv, np.float32))
        solver.solve_pcg_f64(*row_args, 20000, 1e-6)  # warm the buffers and graph
        t = time.perf_counter()
        _, it_drop, _ = solver.solve_pcg_f64(*row_args, 20000, 1e-6)
        t_drop = time.perf_counter() - t
        drop_rate = 3.0 * active * it_drop / t_drop
        # The drop-in itself: the reference's f32 block-Jacobi PCG semantics
        # (src/cuda_solver.rs:366-500) at its default tolerance.
        try:
            solver.solve_pcg(*row_args, 20000, 1e-5)
            t = time.perf_counter()
            _, it32, res32 = solver.solve_pcg(*row_args, 20000, 1e-5)
            t32 = time.perf_counter() - t
            drop32 = {"dof_iters_per_s": 3.0 * active * it32 / t32, "ms": 1000 * t32, "iters": it32,
                      "relative_residual": res32, "tolerance": 1e-5,
                      "note": "vx_solve_pcg, the reference's f32 PCG semantics; host buffers, PCIe upload included"}
        except solver.SolverError as e:
            drop32 = {"status": e.status, "error": str(e), "tolerance": 1e-5}
        extra = {
            "mixed_solve": {"dof_iters_per_s": mixed_rate, "ms": 1000 * tm, "inner": rm["inner_iterations"],
                            "sweeps": rm["outer_sweeps"], "rel_residual": rm["relative_residual"]},
            "plain_fp64_structured": {"dof_iters_per_s": f64_rate, "ms": 1000 * t64, "iters": r64["inner_iterations"],
                                      "rel_residual": r64["relative_residual"]},
            "plain_fp64_dropin_rows": {"dof_iters_per_s": drop_rate, "ms": 1000 * t_drop, "iters": it_drop,
                                       "tolerance": 1e-6, "note": "vx_solve_pcg_f64, host buffers, PCIe upload included"},
            "dropin_rows_f32": drop32,
            "mixed_solve_to_1e-5": {"ms": 1000 * t5, "inner": r5["inner_iterations"], "sweeps": r5["outer_sweeps"],
                                    "rel_residual": r5["relative_residual"]},
            "red_black_solve": {"ms": 1000 * trb, "inner": rrb["inner_iterations"], "sweeps": rrb["outer_sweeps"],
                                "rel_residual": rrb["relative_residual"], "tolerance": SOLVE_TOL,
                                "time_vs_block_jacobi": tm / trb},
            "red_black_solve_to_1e-5": {"ms": 1000 * trb5, "inner": rrb5["inner_iterations"],
                                        "sweeps": rrb5["outer_sweeps"], "rel_residual": rrb5["relative_residual"],
                                        "time_vs_block_jacobi": t5 / trb5},
            "multigrid_solve": {"ms": 1000 * tmg, "inner": rmg["inner_iterations"], "sweeps": rmg["outer_sweeps"],
                                "rel_residual": rmg["relative_residual"], "tolerance": SOLVE_TOL,
                                "time_vs_block_jacobi": tm / tmg, "time_vs_red_black": trb / tmg,
                                "note": "solve of an assembled system with the hierarchy already built"},
            "multigrid_solve_to_1e-5": {"ms": 1000 * tmg5, "inner": rmg5["inner_iterations"],
                                        "sweeps": rmg5["outer_sweeps"], "rel_residual": rmg5["relative_residual"],
                                        "time_vs_block_jacobi": t5 / tmg5},
            "mixed_vs_plain_fp64_structured": mixed_rate / f64_rate,
            "mixed_vs_plain_fp64_dropin": mixed_rate / drop_rate,
            "time_to_solution_vs_fp64_structured": t64 / tm,
        }
    gs.close()
    cpu_on = not args.no_cpu_baseline
    if not args.no_compare:
        extra["red_black_step"] = red_black_steps(scene, step_kw, args.steps, args.warmup, ms_per_step)
        extra["multigrid_step"] = red_black_steps(scene, step_kw, args.steps, args.warmup, ms_per_step,
                                                  preconditioner=2)
    if not args.no_batch:
        extra["batched"] = batched(cpu=cpu_on)
        rb = batched(preconditioner=1)
        rb["time_vs_block_jacobi"] = extra["batched"]["wall_s"] / rb["wall_s"]
        extra["batched_red_black"] = rb
        print("bench: batched legs done", file=sys.stderr, flush=True)
    if not args.no_serve:
        extra["served"] = served(args.serve_count, compare=not args.no_compare, cpu=cpu_on)
        # the same requests with every admitted scene in every tick
        fifo = served(args.serve_count, compare=False, cpu=False, policy="fifo")
        fifo["p50_vs_sjf"] = fifo["p50_s"] / extra["served"]["p50_s"]
        fifo["p99_vs_sjf"] = fifo["p99_s"] / extra["served"]["p99_s"]
        extra["served_fifo"] = fifo
        # the same requests with the red-black preconditioner (half the iterations)
        rbs = served(args.serve_count, compare=False, cpu=False, preconditioner=1)
        rbs["p50_vs_block_jacobi"] = extra["served"]["p50_s"] / rbs["p50_s"]
        rbs["p99_vs_block_jacobi"] = extra["served"]["p99_s"] / rbs["p99_s"]
        extra["served_red_black"] = rbs
        print("bench: served legs done", file=sys.stderr, flush=True)
    if not args.no_large_roofline:
        extra["roofline_hbm_out_of_cache"] = large_roofline()
        extra["shard_slabs"] = shard_slabs()
        print("bench: 192x192x96 roofline done", file=sys.stderr, flush=True)
    if not args.no_contact:
        extra["contact_projection"] = contact_bench(cpu=cpu_on)
    if not args.no_cut:
        # the cutting scene (the blade as the rigid tool at 5 m/s, fracture every step), bounded
        extra["cut"] = cut_run(CUT_STEPS, cpu=False, quiet=True)
        print("bench: cut leg done", file=sys.stderr, flush=True)

    cpu = None
    if cpu_on:
        cpu = cpu_baseline(scene)
        extra["error_vs_fp64_oracle"] = error_vs_oracle_first_solve(scene, sp)

    line = {
        "metric": "DOF-PCG-iters/s",
        "value": value,
        "unit": "DOF-PCG-iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 inner / f64 refinement and reductions",
        "data": "synthetic",
        "config": {"workload": "large_64x64x64_mixed: 262144 REVs, %d active rows, 1 ms step, 1 substep, "
                               "<=%d corotation passes, outer rel. residual %g, inner %g" %
                               (active, cfg.corotation_iterations, SOLVE_TOL, INNER_TOL),
                   "active_rows": active, "dof": 3 * active, "dof_total": 3 * scene.n, "solves": solves,
                   "iterations": iters, "parallelism": "single GPU",
                   "launch": "direct" if args.direct_launch else "graphs"},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    line.update(extra)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
