#!/usr/bin/env python3
"""FLIP steps/s of the MSBG step on one MI355X, with the roofline of the
dominant kernel and the FP64 oracle timed beside it on the host cores.

    python bench.py [--steps K] [--warmup W] [--workload pool256]
    python bench.py --traffic-from FETCH.csv WRITE.csv --copy-bytes B

The second form turns two rocprofv3 --pmc runs (FETCH_SIZE, then
WRITE_SIZE, counter-collection CSVs) into HBM bytes per dispatch per kernel,
scaled by what each counter reads for the largest device-to-device copy in
the run (B bytes), and writes profiles/traffic.json for this kernel source.

Default workload: 256^3-effective domain, 3 native levels (cell sizes h, 2h,
4h, h = 1/256, 8^3-cell blocks), level 0 within one block (8 cells) of the
liquid surface, level 1 within four blocks (32 cells); pool y < 0.3 plus a
sphere of radius 0.1 at (0.5, 0.7, 0.5) falling at 2 m/s; 8 particles per
fine cell of the liquid region; gravity -9.81, dt 1/240; every step rebuilds
the topology from the particles, P2G, body force, MGPCG (V(2,2) red-black GS
preconditioner) to relative residual 1e-6, projection, G2P (FLIP 0.95), RK3.
Particles are synthetic (counter-hashed jitter) and resident in HBM before
the timed region.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "FLIP steps/s + smoother GB/s (% HBM peak); adjoint/forward time ratio @1e-8"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload(name):
    from anitowave_adaptive_amd import scenes
    w = dict(scenes.WORKLOADS[name])
    return w


def run_gpu(args, w, pos, vel):
    import anitowave_adaptive_amd as aw
    aw.memory_stats(reset_peak=True)
    ctx = aw.Context(0)
    n = len(pos)
    dpos = aw.DeviceArray.from_numpy(ctx, pos)
    dvel = aw.DeviceArray.from_numpy(ctx, vel)
    dids = aw.DeviceArray.from_numpy(ctx, np.arange(n, dtype=np.int32))
    grid = aw.MSBGManager(ctx, w["dims"], 1.0 / w["n"], num_levels=w["levels"], band=w["band"])
    solver = aw.VCycleSolver(grid)
    step = aw.FlipStep(grid, solver, dt=1.0 / 240.0, gravity=(0.0, -9.81, 0.0), alpha=0.95, tol=args.tol,
                       max_iters=200, rebuild_topology=True)
    iters = []
    for i in range(args.warmup):
        st = step(dpos.ptr, dvel.ptr, dids.ptr, n, allow_not_converged=True)
        log(f"[bench] warmup step {i}: pcg iters {st.solve.iterations} true rel. residual "
            f"{st.solve.rel_residual:.2e} (recurrence {st.solve.recurrence_rel_residual:.2e})")
    ctx.synchronize()
    t0 = time.perf_counter()
    changed = 0
    # the solver's contract: a step whose true residual stays above the
    # tolerance (the FP32 floor) returns AW_ERR_NOT_CONVERGED; count them
    floors, worst_true, worst_rec = 0, 0.0, 0.0
    for i in range(args.steps):
        st = step(dpos.ptr, dvel.ptr, dids.ptr, n, allow_not_converged=True)
        iters.append(st.solve.iterations)
        changed += st.topology_changed
        floors += step.status != aw.AW_OK
        worst_true = max(worst_true, st.solve.rel_residual)
        worst_rec = max(worst_rec, st.solve.recurrence_rel_residual)
    ctx.synchronize()
    t1 = time.perf_counter()
    # per-kernel HIP-event timings from separate steps, so that the events
    # do not slow the timed region down
    prof = {}
    if args.profile:
        ctx.profile(True, runs=True)
        for i in range(args.profile_steps):
            step(dpos.ptr, dvel.ptr, dids.ptr, n, allow_not_converged=True)
        ctx.synchronize()
        prof = ctx.profile_report()
        ctx.profile(False)
    # the level-0 smoother and residual alone, back to back on the last
    # step's hierarchy (one event pair around each run of launches)
    iso = {}
    if args.profile:
        ctx.profile(True, runs=True)
        solver.smooth(0, 20)
        for _ in range(40):
            solver.residual(0)
        ctx.synchronize()
        iso = ctx.profile_report()
        ctx.profile(False)
    lv = list(st.num_blocks_by_level)
    # per multigrid level: blocks, and those the smoother's launches cover
    # (the blocks holding a liquid cell)
    mg_blocks = []
    for k in range(solver.num_levels()):
        L = solver.level_info(k)
        mg_blocks.append([int(L.num_blocks), int(L.smoothed_blocks)])
    # the step's own high-water mark, before the legs below allocate more
    step_peak = aw.memory_stats()[1]
    # nearest-cell resample of the last pressure onto a uniform grid
    resample = None
    if args.resample > 0:
        o = (args.resample,) * 3
        nvox = int(np.prod(o))
        dout = aw.DeviceArray(ctx, 4 * nvox)
        lay = grid.layout()
        grid.resample(lay.d_pressure, o, dout.ptr)
        ctx.synchronize()
        ctx.profile(True, runs=False)
        for _ in range(5):
            grid.resample(lay.d_pressure, o, dout.ptr)
        ctx.synchronize()
        g = ctx.profile_report()["resample"]
        ctx.profile(False)
        dout.free()
        gbs = g["bytes"] / (g["ms"] * 1e-3) / 1e9
        resample = {"out_dims": list(o), "us_per_call": round(1e3 * g["ms"] / g["launches"], 2),
                    "bytes_per_call": round(g["bytes"] / g["launches"]), "gbs": round(gbs, 1),
                    "frac_hbm_peak": round(gbs / HBM_PEAK_GBS, 4)}
    # the particle-radius level rule on the same particles (block-sorted by
    # the steps), radius h / 2 everywhere: the first call changes the
    # topology, the timed ones find it unchanged
    radius = None
    if args.radius_calls > 0:
        drad = aw.DeviceArray.from_numpy(ctx, np.full(n, 0.5 / w["n"], np.float32))
        grid.build_topology_radius(dpos.ptr, drad.ptr, n)
        ctx.synchronize()
        ta = time.perf_counter()
        for _ in range(args.radius_calls):
            grid.build_topology_radius(dpos.ptr, drad.ptr, n)
        tb = time.perf_counter()
        ctx.profile(True, runs=True)
        grid.build_topology_radius(dpos.ptr, drad.ptr, n)
        rp = ctx.profile_report()
        ctx.profile(False)
        radius = {"ms_per_call": round(1e3 * (tb - ta) / args.radius_calls, 3), "particles": n,
                  "radius": "h/2", "levels_ms": round(rp.get("levels", {}).get("ms", 0.0), 4),
                  "bin_ms": round(rp.get("bin", {}).get("ms", 0.0), 4),
                  "blocks_by_level": [int(v) for v in np.bincount(grid.levels(), minlength=w["levels"])]}
    res = dict(seconds=t1 - t0, iters=iters, topology_changes=changed, blocks_by_level=lv, radius_rule=radius,
               resample=resample, steps_above_tol=floors, worst_true_rel_residual=worst_true,
               worst_recurrence_rel_residual=worst_rec, liquid_cells=int(st.liquid_cells),
               total_cells=int(st.total_cells), rel_residual=st.solve.rel_residual, profile=prof, isolated=iso,
               peak_bytes=step_peak, mg_blocks=mg_blocks)
    return res


# profile groups -> the kernel that runs them
KERNEL_OF = {"smooth_l0": "smooth_kernel", "smooth_coarse": "smooth_kernel", "smooth_uniform": "uniform_smooth_kernel",
             "residual_l0": "residual_kernel",
             "residual_coarse": "residual_kernel", "residual_restrict_l0": "residual_kernel",
             "residual_restrict_coarse": "residual_kernel", "g2p_advect": "g2p_block_kernel",
             "p2g_deposit": "p2g_deposit_kernel", "p2g_gather": "p2g_gather_kernel", "to_fine": "to_fine_kernel",
             "apply_l0": "apply_dot_kernel", "sort_particles": "scatter_kernel", "vcycle_tail": "tail_kernel"}


def kernel_source_sha():
    import hashlib
    src = os.path.join(ROOT, "anitowave_adaptive_amd", "csrc", "aw_kernels.hip")
    return hashlib.sha256(open(src, "rb").read()).hexdigest()[:16]


def traffic_for(kernel, workload):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    summary (profiles/traffic.json), only for the workload and the kernel
    source it was measured on (both recorded there); otherwise null."""
    tf = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        t = json.load(open(tf))
    except Exception:
        return None, "no traffic measurement"
    if t.get("kernel_source_sha") != kernel_source_sha():
        return None, "traffic measurement is for another kernel source: not used"
    if t.get("workload") != workload:
        return None, f"traffic measured on workload {t.get('workload')}, not {workload}: not used"
    v = t.get("bytes_per_launch", {}).get(kernel)
    return v, (t.get("measured_by", "profiles/traffic.json") if v is not None else "kernel not measured")


def roofline(prof, workload):
    if not prof:
        return None, {}
    per = {}
    for k, v in prof.items():
        if v["ms"] <= 0:
            continue
        gbs = v["bytes"] / (v["ms"] * 1e-3) / 1e9
        per[k] = dict(launches=v["launches"], ms_total=round(v["ms"], 4),
                      us_per_launch=round(1e3 * v["ms"] / v["launches"], 3),
                      bytes_per_launch=round(v["bytes"] / v["launches"]), gbs=round(gbs, 1),
                      frac=round(gbs / HBM_PEAK_GBS, 4))
    # the dominant kernel: most time over the profiled steps, its groups added
    kern = {}
    for k, v in prof.items():
        if v["ms"] <= 0:
            continue
        a = kern.setdefault(KERNEL_OF.get(k, k), {"ms": 0.0, "bytes": 0.0, "launches": 0, "groups": []})
        a["ms"] += v["ms"]
        a["bytes"] += v["bytes"]
        a["launches"] += v["launches"]
        a["groups"].append(k)
    dom = max(kern, key=lambda k: kern[k]["ms"])
    d = kern[dom]
    gbs = d["bytes"] / (d["ms"] * 1e-3) / 1e9
    traffic, traffic_note = traffic_for(dom, workload)
    rl = {"kernel": dom, "groups": d["groups"], "bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
          "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic, "traffic_source": traffic_note,
          "bytes_per_launch": round(d["bytes"] / d["launches"]), "us_per_launch": round(1e3 * d["ms"] / d["launches"], 3),
          "launches": d["launches"]}
    return rl, per


def run_mixed(args, w, pos, vel, tol=1e-8, steps=None):
    """The same workload with mixed precision (FP32 V-cycle PCG inside FP64
    iterative refinement) to relative residual `tol`.  The first step (from
    the initial particles, untimed) is kept for the error against the FP64
    oracle's step: its FP64 pressure, u_new and the positions by id."""
    import anitowave_adaptive_amd as aw
    aw.memory_stats(reset_peak=True)
    ctx = aw.Context(0)
    n = len(pos)
    dpos = aw.DeviceArray.from_numpy(ctx, pos)
    dvel = aw.DeviceArray.from_numpy(ctx, vel)
    dids = aw.DeviceArray.from_numpy(ctx, np.arange(n, dtype=np.int32))
    grid = aw.MSBGManager(ctx, w["dims"], 1.0 / w["n"], num_levels=w["levels"], band=w["band"])
    solver = aw.VCycleSolver(grid, precision=1)
    step = aw.FlipStep(grid, solver, dt=1.0 / 240.0, gravity=(0.0, -9.81, 0.0), alpha=0.95, tol=tol,
                       max_iters=400, rebuild_topology=True)
    steps = steps or args.mixed_steps
    step(dpos.ptr, dvel.ptr, dids.ptr, n)
    first = None
    if args.cpu:
        P = np.empty((n, 3), np.float32)
        P[dids.download(np.int32, n)] = dpos.download(np.float32, 3 * n).reshape(-1, 3)
        first = {"p": solver.level_get(0, "x64"), "u_new": [grid.face("u_new", a) for a in range(3)], "pos": P,
                 "liquid": grid.cell("liquid")}
    ctx.synchronize()
    t0 = time.perf_counter()
    it, outer, rel = [], [], 0.0
    for _ in range(steps):
        st = step(dpos.ptr, dvel.ptr, dids.ptr, n)
        it.append(st.solve.iterations)
        outer.append(st.solve.outer_iterations)
        rel = max(rel, st.solve.rel_residual)
    ctx.synchronize()
    dt = time.perf_counter() - t0
    return {"steps_per_s": round(steps / dt, 3), "ms_per_step": round(1e3 * dt / steps, 3),
            "rel_tol": tol, "worst_rel_residual": rel, "steps": steps,
            "iterations_mean": round(float(np.mean(it)), 2), "refinements_mean": round(float(np.mean(outer)), 2),
            "peak_hbm_gib": round(aw.memory_stats()[1] / 2**30, 3)}, first


def run_warm(args, w, pos, vel):
    """The headline workload with the pressure solve warm-started from the
    previous step's pressure (remapped onto each new block layout), same
    tolerance.  The reference starts every solve from zero, as the headline does."""
    import anitowave_adaptive_amd as aw
    ctx = aw.Context(0)
    n = len(pos)
    dpos = aw.DeviceArray.from_numpy(ctx, pos)
    dvel = aw.DeviceArray.from_numpy(ctx, vel)
    dids = aw.DeviceArray.from_numpy(ctx, np.arange(n, dtype=np.int32))
    grid = aw.MSBGManager(ctx, w["dims"], 1.0 / w["n"], num_levels=w["levels"], band=w["band"])
    solver = aw.VCycleSolver(grid)
    step = aw.FlipStep(grid, solver, dt=1.0 / 240.0, gravity=(0.0, -9.81, 0.0), alpha=0.95, tol=args.tol,
                       max_iters=200, rebuild_topology=True, warm_start=True)
    for _ in range(args.warmup):
        step(dpos.ptr, dvel.ptr, dids.ptr, n, allow_not_converged=True)
    ctx.synchronize()
    t0 = time.perf_counter()
    it, changed, rel = [], 0, 0.0
    for _ in range(args.warm_steps):
        st = step(dpos.ptr, dvel.ptr, dids.ptr, n, allow_not_converged=True)
        it.append(st.solve.iterations)
        changed += st.topology_changed
        rel = max(rel, st.solve.rel_residual)
    ctx.synchronize()
    dt = time.perf_counter() - t0
    return {"steps_per_s": round(args.warm_steps / dt, 3), "ms_per_step": round(1e3 * dt / args.warm_steps, 3),
            "rel_tol": args.tol, "worst_rel_residual": rel, "steps": args.warm_steps,
            "iterations_mean": round(float(np.mean(it)), 2), "topology_changes": changed}


def run_adjoint(args, name=None, steps=None, scene=None):
    """Differentiable rollout, by default slosh128: 128^3 effective, 2 native
    levels, sloshing pool (y < 0.4 plus a seeded sinusoidal perturbation of
    amplitude 0.05), 16 steps, checkpoint every 4; loss |COM(x_K) - x*|^2,
    gradient with respect to the initial velocities; forward and adjoint
    solves mixed precision to 1e-8.  Ratio = (recompute + backward) / forward."""
    import anitowave_adaptive_amd as aw
    from anitowave_adaptive_amd import scenes
    name = name or args.adjoint_workload
    w = scenes.WORKLOADS[name]
    pos, vel = scene if scene is not None else w["scene"](n=w["n"])
    n = len(pos)
    aw.memory_stats(reset_peak=True)
    ctx = aw.Context(0)
    grid = aw.MSBGManager(ctx, w["dims"], 1.0 / w["n"], num_levels=w["levels"], band=w["band"])
    solver = aw.VCycleSolver(grid, precision=1)
    prm = aw.FlipStep(grid, solver, dt=1.0 / 240.0, gravity=(0.0, -9.81, 0.0), alpha=0.95, tol=1e-8,
                      max_iters=400).params
    steps, every = (steps or args.adjoint_steps or w["steps"]), 4
    ro = aw.Rollout(grid, solver, n, steps, checkpoint_every=every)
    target = np.array([0.6, 0.35, 0.5])
    dpos = aw.DeviceArray.from_numpy(ctx, pos)
    dvel = aw.DeviceArray.from_numpy(ctx, vel)
    dids = aw.DeviceArray(ctx, 4 * n)
    dxb = aw.DeviceArray(ctx, 12 * n)
    dvb = aw.DeviceArray(ctx, 12 * n)
    res = None
    for rep in range(2):  # the first pass warms the allocations up
        dpos.upload(pos)
        dvel.upload(vel)
        dids.upload(np.arange(n, dtype=np.int32))
        st = ro.forward(prm, dpos.ptr, dvel.ptr, dids.ptr)
        fwd = st.forward_ms
        x = dpos.download(np.float32, 3 * n).reshape(-1, 3)
        ids = dids.download(np.int32, n)
        xs = np.empty_like(x)
        xs[ids] = x
        c = xs.astype(np.float64).mean(axis=0) - target
        dxb.upload(np.broadcast_to((2.0 * c / n).astype(np.float32), (n, 3)).copy())
        dvb.upload(np.zeros((n, 3), np.float32))
        st = ro.backward(dxb.ptr, dvb.ptr)
        g = dvb.download(np.float32, 3 * n)
        res = {"workload": f"{name}: {w['n']}^3 effective, {w['levels']} levels, {n} particles, {steps} steps, "
                           f"checkpoint every {every}, mixed precision to 1e-8",
               "forward_ms": round(fwd, 2), "recompute_ms": round(st.recompute_ms, 2),
               "backward_ms": round(st.backward_ms, 2),
               "ratio": round((st.recompute_ms + st.backward_ms) / fwd, 3),
               "forward_iterations": st.forward_iterations, "adjoint_iterations": st.adjoint_iterations,
               "worst_adjoint_rel_residual": st.adjoint_rel_residual,
               "replay_mismatched_steps": st.replay_mismatched_steps,
               "replay_max_state_diff": st.replay_max_state_diff,
               "grad_norm": float(np.linalg.norm(g)), "loss": float(c @ c),
               "peak_hbm_gib": round(aw.memory_stats()[1] / 2**30, 3)}
        log(f"[bench] adjoint pass {rep}: {res}")
    return res


def run_large(args):
    """The 512^3-effective, 4-level double dam break (two 0.3 x 0.6 x 0.3
    columns at opposite corners, 8 particles per fine cell): plain FP32 steps
    with per-kernel timings, mixed precision to 1e-8, and the adjoint over an
    8-step window with checkpoints every 4 steps; peak library HBM per leg."""
    import copy
    w = workload("double512")
    t = time.perf_counter()
    pos, vel = w["scene"](n=w["n"])
    log(f"[bench] large: seeded {len(pos)} particles in {time.perf_counter() - t:.1f}s")
    a = copy.copy(args)
    a.steps, a.warmup, a.profile_steps, a.mixed_steps, a.radius_calls, a.resample = 5, 2, 2, 3, 0, 0
    a.cpu = False
    gpu = run_gpu(a, w, pos, vel)
    rl, per = roofline(gpu["profile"], "double512")
    log(f"[bench] large: {a.steps / gpu['seconds']:.3f} steps/s")
    mixed, _ = run_mixed(a, w, pos, vel)
    log(f"[bench] large mixed precision 1e-8: {mixed}")
    adj = run_adjoint(a, name="double512", steps=8, scene=(pos, vel))
    return {"workload": f"double512: 512^3 effective, 4 levels, {len(pos)} particles",
            "steps_per_s": round(a.steps / gpu["seconds"], 3), "ms_per_step": round(1e3 * gpu["seconds"] / a.steps, 3),
            "pcg_iters_mean": round(float(np.mean(gpu["iters"])), 2), "peak_hbm_gib": round(gpu["peak_bytes"] / 2**30, 3),
            "blocks_by_level": gpu["blocks_by_level"][: w["levels"]], "roofline": rl, "kernels": per,
            "mixed_precision_1e-8": mixed, "adjoint": adj}


def run_scene(args, w, pos, vel):
    """The whole scene from its initial state: every step of the workload
    (100 for the pool, so the sphere's impact on the pool and the splash are
    inside), timed as one run, fresh grid and solver (allocations included)."""
    import anitowave_adaptive_amd as aw
    ctx = aw.Context(0)
    n = len(pos)
    dpos = aw.DeviceArray.from_numpy(ctx, pos)
    dvel = aw.DeviceArray.from_numpy(ctx, vel)
    dids = aw.DeviceArray.from_numpy(ctx, np.arange(n, dtype=np.int32))
    grid = aw.MSBGManager(ctx, w["dims"], 1.0 / w["n"], num_levels=w["levels"], band=w["band"])
    solver = aw.VCycleSolver(grid)
    step = aw.FlipStep(grid, solver, dt=1.0 / 240.0, gravity=(0.0, -9.81, 0.0), alpha=0.95, tol=args.tol,
                       max_iters=200, rebuild_topology=True)
    ctx.synchronize()
    t0 = time.perf_counter()
    it, floors, worst, level0 = [], 0, 0.0, []
    for k in range(args.scene_steps):
        st = step(dpos.ptr, dvel.ptr, dids.ptr, n, allow_not_converged=True)
        it.append(st.solve.iterations)
        floors += step.status != aw.AW_OK
        worst = max(worst, st.solve.rel_residual)
        level0.append(int(st.num_blocks_by_level[0]))
    ctx.synchronize()
    dt = time.perf_counter() - t0
    return {"steps": args.scene_steps, "steps_per_s": round(args.scene_steps / dt, 3),
            "ms_per_step": round(1e3 * dt / args.scene_steps, 3), "pcg_iters_mean": round(float(np.mean(it)), 2),
            "pcg_iters_max": int(max(it)), "steps_above_tol": floors, "worst_true_rel_residual": worst,
            "level0_blocks_min_max": [min(level0), max(level0)]}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def rel_max(a, b, mask=None):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    if mask is not None:
        a, b = a[mask], b[mask]
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


def run_cpu_baseline(args, w, pos, vel, firsts):
    """FP64 C++/OpenMP oracle on the host cores, at relative residual 1e-8
    (the tolerance the GPU's mixed-precision leg is compared at): two steps
    of the same workload from the same particles, all threads; the first
    step's pressure, u_new and positions give the error of the GPU's mixed
    1e-8 step, a 1e-10 step (not timed into the value) that of the 1e-10 leg;
    one step of the 64^3 dam break on one thread and on all, for scaling."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    nt = O.num_threads()
    O.set_threads(nt)
    g = O.OracleGrid(w["dims"], num_levels=w["levels"], h=1.0 / w["n"], band=w["band"])
    g.set_params()
    grav = (0.0, -9.81, 0.0)
    errors = {}

    def errs(tol, r):
        f = firsts.get(tol)
        if f is None:
            return None
        liq = f["liquid"] > 0
        return {"pressure_rel": rel_max(f["p"], g.cell("p"), liq),
                "u_new_rel": max(rel_max(f["u_new"][a], g.face("u_new", a)) for a in range(3)),
                "position_abs": float(np.abs(f["pos"] - r["pos"]).max())}

    t0 = time.perf_counter()
    r1 = g.step(pos, vel, 1.0 / 240.0, grav, 0.95, tol=1e-8, max_iters=400)
    t1 = time.perf_counter()
    errors["1e-08"] = errs(1e-8, r1)
    r2 = g.step64(r1["pos"], r1["vel"], 1.0 / 240.0, grav, 0.95, tol=1e-8, max_iters=400)
    t2 = time.perf_counter()
    dt = (t1 - t0) + (t2 - t1)
    r3 = g.step(pos, vel, 1.0 / 240.0, grav, 0.95, tol=1e-10, max_iters=400)
    t3 = time.perf_counter()
    errors["1e-10"] = errs(1e-10, r3)
    # thread scaling on the 64^3 dam break
    wd = workload("dam64")
    pd, vd = wd["scene"](n=wd["n"])
    gd = O.OracleGrid(wd["dims"], num_levels=wd["levels"], h=1.0 / wd["n"], band=wd["band"])
    gd.set_params()
    gd.step(pd, vd, 1.0 / 240.0, grav, 0.95, tol=1e-8)
    ta = time.perf_counter()
    gd.step(pd, vd, 1.0 / 240.0, grav, 0.95, tol=1e-8)
    tb = time.perf_counter()
    O.set_threads(1)
    gd.step(pd, vd, 1.0 / 240.0, grav, 0.95, tol=1e-8)
    tc = time.perf_counter()
    O.set_threads(nt)
    return {"value": round(2.0 / dt, 5), "unit": "steps/s", "cores": nt, "kind": "port",
            "cpu_model": cpu_model(), "rel_tol": 1e-8,
            "sample": f"2 full steps of the same workload ({len(pos)} particles) through the FP64 C++/OpenMP "
                      f"oracle at relative residual 1e-8 on {nt} threads: {t1 - t0:.2f} s + {t2 - t1:.2f} s, "
                      f"{r1['iters']} + {r2['iters']} PCG iterations",
            "step_1e-10_s": round(t3 - t2, 3),
            "dam64_steps_per_s": {"threads_all": round(1.0 / (tb - ta), 3), "threads_1": round(1.0 / (tc - tb), 3)},
            "errors_vs_gpu_mixed": errors}


def _kernel_key(name):
    """Short kernel name from a rocprofv3 kernel name; the two G2P passes apart."""
    base = name.replace("(anonymous namespace)::", "").split("(")[0]
    tmpl = ""
    if "<" in base:
        base, tmpl = base.split("<", 1)
    base = base.replace("void ", "").split("::")[-1]
    if base == "g2p_advect_kernel":
        return base + (" (slow pass)" if tmpl.startswith("true") else "")
    return base


def traffic_from(fetch_csv, write_csv, copy_bytes):
    """bytes_per_launch per kernel from two --pmc runs, each counter scaled so
    that the largest __amd_rocclr_copyBuffer dispatch reads copy_bytes."""
    import csv
    out, scale = {}, {}
    for path, counter in ((fetch_csv, "FETCH_SIZE"), (write_csv, "WRITE_SIZE")):
        rows = [r for r in csv.DictReader(open(path)) if r["Counter_Name"] == counter]
        copies = [r for r in rows if "copyBuffer" in r["Kernel_Name"]]
        big = max(int(r["Grid_Size"]) for r in copies)
        cal = [float(r["Counter_Value"]) for r in copies if int(r["Grid_Size"]) == big]
        kib = sum(cal) / len(cal)
        scale[counter] = copy_bytes / (kib * 1024.0)
        per = {}
        for r in rows:
            if "copyBuffer" in r["Kernel_Name"]:
                continue
            k = _kernel_key(r["Kernel_Name"])
            a = per.setdefault(k, [0.0, 0])
            a[0] += float(r["Counter_Value"]) * 1024.0 * scale[counter]
            a[1] += 1
        for k, (b, n) in per.items():
            o = out.setdefault(k, {"dispatches": n, "FETCH_SIZE": 0.0, "WRITE_SIZE": 0.0})
            o[counter] = b / n
    for o in out.values():
        o["bytes"] = o["FETCH_SIZE"] + o["WRITE_SIZE"]
    return out, scale


def write_traffic(args):
    out, scale = traffic_from(args.traffic_from[0], args.traffic_from[1], args.copy_bytes)
    doc = {"kernel_source_sha": kernel_source_sha(), "workload": args.workload,
           "measured_by": (f"rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE (profiles/pool256_traffic.md); "
                           f"scaled {scale['FETCH_SIZE']:.3f} / {scale['WRITE_SIZE']:.3f} by a "
                           f"{args.copy_bytes:.0f}-byte device copy"),
           "bytes_per_launch": {k: round(v["bytes"]) for k, v in sorted(out.items())}}
    with open(os.path.join(ROOT, "profiles", "traffic.json"), "w") as f:
        json.dump(doc, f, indent=1)
        f.write("\n")
    print("| Kernel | Dispatches | FETCH (scaled, MB) | WRITE (scaled, MB) | Traffic (MB) |")
    print("|---|---:|---:|---:|---:|")
    for k, v in sorted(out.items(), key=lambda kv: -kv[1]["bytes"] * kv[1]["dispatches"]):
        print(f"| {k} | {v['dispatches']} | {v['FETCH_SIZE'] / 1e6:.1f} | {v['WRITE_SIZE'] / 1e6:.1f} | "
              f"{v['bytes'] / 1e6:.1f} |")


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="pool256")
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--no-profile", dest="profile", action="store_false")
    ap.add_argument("--profile-steps", type=int, default=5, help="extra steps timed per kernel with HIP events")
    ap.add_argument("--no-cpu-baseline", dest="cpu", action="store_false")
    ap.add_argument("--no-adjoint", dest="adjoint", action="store_false")
    ap.add_argument("--mixed-steps", type=int, default=5,
                    help="steps of each mixed-precision leg, 1e-8 and 1e-10 (0: skip)")
    ap.add_argument("--scene-steps", type=int, default=100,
                    help="steps of the whole-scene leg from the initial state (0: skip)")
    ap.add_argument("--warm-steps", type=int, default=20,
                    help="steps of the warm-started leg (pressure from the previous step; 0: skip)")
    ap.add_argument("--radius-calls", type=int, default=5,
                    help="timed calls of the particle-radius level rule on the workload's particles (0: skip)")
    ap.add_argument("--resample", type=int, default=512,
                    help="edge of the uniform grid the last pressure is resampled onto (0: skip)")
    ap.add_argument("--adjoint-workload", default="slosh128")
    ap.add_argument("--adjoint-steps", type=int, default=0, help="rollout length (0: the workload's own)")
    ap.add_argument("--no-large", dest="large", action="store_false",
                    help="skip the 512^3 double-dam leg (plain, mixed 1e-8, 8-step adjoint)")
    ap.add_argument("--traffic-from", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="write profiles/traffic.json from two rocprofv3 --pmc counter-collection CSVs")
    ap.add_argument("--copy-bytes", type=float, default=0.0,
                    help="bytes of the largest device copy in the profiled run (the counters' scale)")
    args = ap.parse_args()
    if args.traffic_from:
        write_traffic(args)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is a design note only",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    w = workload(args.workload)
    t = time.perf_counter()
    pos, vel = w["scene"](n=w["n"])
    log(f"[bench] seeded {len(pos)} particles in {time.perf_counter() - t:.1f}s")
    gpu = run_gpu(args, w, pos, vel)
    log(f"[bench] gpu {args.steps} steps in {gpu['seconds']:.3f}s, pcg iters {gpu['iters']}")
    rl, per = roofline(gpu["profile"], args.workload)
    mixed, mixed10, firsts = None, None, {}
    if args.mixed_steps > 0:
        mixed, firsts[1e-8] = run_mixed(args, w, pos, vel, 1e-8)
        log(f"[bench] mixed precision 1e-8: {mixed}")
        mixed10, firsts[1e-10] = run_mixed(args, w, pos, vel, 1e-10)
        log(f"[bench] mixed precision 1e-10: {mixed10}")
    scene = None
    if args.scene_steps > 0:
        scene = run_scene(args, w, pos, vel)
        log(f"[bench] whole scene: {scene}")
    warm = None
    if args.warm_steps > 0:
        warm = run_warm(args, w, pos, vel)
        log(f"[bench] warm-started pressure: {warm}")
    adj = run_adjoint(args) if args.adjoint else None
    large = run_large(args) if args.large else None
    cpu = None
    if args.cpu:
        log("[bench] cpu baseline: oracle steps at 1e-8 and 1e-10, dam64 thread scaling")
        cpu = run_cpu_baseline(args, w, pos, vel, firsts)
        firsts.clear()
        log(f"[bench] cpu baseline {cpu['value']} steps/s on {cpu['cores']} threads")
    sps = args.steps / gpu["seconds"]
    smoother = per.get("smooth_l0")
    resid = per.get("residual_restrict_l0")
    _, iso = roofline(gpu["isolated"], args.workload)
    out = {
        "metric": METRIC,
        "value": round(sps, 3),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * gpu["seconds"] / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{args.workload}: {w['n']}^3 effective, {w['levels']} levels, 8^3-cell blocks, "
                        f"{len(pos)} particles, FP32 MGPCG V(2,2) stopped at recurrence residual {args.tol:g} "
                        f"(true residual: worst {gpu['worst_true_rel_residual']:.2e}, "
                        f"{gpu['steps_above_tol']}/{args.steps} steps above {args.tol:g}, the FP32 floor), "
                        f"FLIP 0.95, RK3",
            "solve": {"stopping": f"PCG recurrence residual <= {args.tol:g} |b|, then the true residual "
                                  f"recomputed; one restart from it; AW_ERR_NOT_CONVERGED when it stays above",
                      "steps_above_tol": gpu["steps_above_tol"],
                      "worst_true_rel_residual": gpu["worst_true_rel_residual"],
                      "worst_recurrence_rel_residual": gpu["worst_recurrence_rel_residual"]},
            "liquid_cells": gpu["liquid_cells"],
            "blocks_by_level": gpu["blocks_by_level"][: w["levels"]],
            "mg0_cells": gpu["total_cells"],
            "mg_blocks_and_smoothed": gpu["mg_blocks"],
            "pcg_iters_mean": round(float(np.mean(gpu["iters"])), 2),
            "topology_changes": gpu["topology_changes"],
            "peak_hbm_gib": round(gpu["peak_bytes"] / 2**30, 3),
        },
        "roofline": rl,
        "smoother": None if smoother is None else {
            "level0_in_step": {"gbs": smoother["gbs"], "frac_hbm_peak": smoother["frac"],
                               "us_per_launch": smoother["us_per_launch"]},
            "level0_back_to_back": iso.get("smooth_l0"),
            "coarse_levels_in_step": per.get("smooth_coarse"),
            "uniform_levels_in_step": per.get("smooth_uniform")},
        "residual": None if resid is None else {
            "level0_restrict_in_step": {"gbs": resid["gbs"], "frac_hbm_peak": resid["frac"],
                                        "us_per_launch": resid["us_per_launch"]},
            "level0_back_to_back": iso.get("residual_l0")},
        "mixed_precision_1e-8": mixed,
        "mixed_precision_1e-10": mixed10,
        "whole_scene": scene,
        "warm_start": warm,
        "topology_radius_rule": gpu["radius_rule"],
        "resample": gpu["resample"],
        "adjoint_forward_ratio": None if adj is None else adj["ratio"],
        "adjoint": adj,
        "kernels": per,
        "large": large,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
