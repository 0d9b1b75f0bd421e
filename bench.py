"""Throughput of the PBD cloth substep with Morton/LBVH self-collision on one MI355X.

Workload: one 256x256-vertex cloth (65,536 vertices, 130,050 triangles, 260,610
structural+shear distance constraints, 194,565 dihedral bends) dropped from 0.5 m
onto a static sphere of radius 0.3 m, thickness 0.002 m, dt 1/120 s with 2
substeps per frame (substep dt 1/240 s), 15 Jacobi iterations, 30-bit Morton codes
in the centroid box, LBVH self-collision with VF and EE contacts every substep,
600 frames (1,200 substeps).

A step is one substep. The drop runs from its initial state; the first
1,200 - steps substeps are an untimed pre-roll and the last `--steps` substeps,
where the cloth lies draped on the sphere with its densest self-contact, are the
timed window (`value`). `full_run` gives the whole 1,200-substep drop's rate and
`early_window` its first 20 substeps. With --steps 1200 or more the whole run is
timed.

Prints one JSON line. ``roofline`` covers the window's dominant kernel (the CCD
plan pass, the BVH traversal or the fused Jacobi iteration, whichever takes the
most time per substep) with its algorithmic bytes per launch over its average
launch time measured live with HIP events on the body's stream;
``roofline_other`` gives the other two.
``cpu_baseline`` times the FP32 CPU oracle seeded with the GPU's state entering the
window, on a bounded sample of the same substeps on this host's cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "PBD substeps/s at 256^2 self-colliding cloth"


FRAMES_SUBSTEPS = 1200  # 600 frames at 2 substeps per frame
EARLY_WINDOW = 20
KERNEL_SOLVE, KERNEL_TRAVERSE, KERNEL_CCD_PLAN = 0, 2, 3
WARN_CANDIDATE_CAP = 1 << 4


def _timed(body, substeps: int) -> float:
    """Wall seconds of `substeps` substeps (graph replays) with the state in HBM."""
    if substeps <= 0:
        return 0.0
    body.sync()
    t0 = time.perf_counter()
    body.fixed_update(substeps)
    body.sync()
    return time.perf_counter() - t0


def _traffic_profiles(size: int) -> dict:
    """HBM bytes per launch measured by rocprofv3 counters (profiles/*.json), scaled
    by a plain copy's reading; (None, None) where no profile exists."""
    out = {}
    for key, name in (("traverse", "k_traverse_sphere_drop_256_traffic.json"),
                      ("solve", "k_solve_sphere_drop_256_traffic.json"),
                      ("plan", "k_ccd_plan_sphere_drop_256_traffic.json")):
        tf = os.path.join(ROOT, "profiles", name)
        if size == 256 and os.path.exists(tf):
            tj = json.load(open(tf))
            out[key] = (tj.get("bytes_per_launch"),
                        f"rocprofv3 FETCH_SIZE x {tj.get('fetch_scale')} + WRITE_SIZE x "
                        f"{tj.get('write_scale')} (256 MiB copy calibration), {tj.get('state')}")
        else:
            out[key] = (None, None)
    return out


def algorithmic_bytes(n, s, bd, t, iters, contacts, pairs):
    per_iter = 36 * n + 16 * s + 24 * bd + 48 * contacts
    substep = 112 * n + iters * per_iter + 300 * t + 32 * n + 8 * pairs
    return per_iter, substep


def cpu_baseline(scene, budget_s: float, state: dict | None = None):
    """The FP32 oracle (the C++ restatement of the reference path) on this host, on
    the substeps the GPU timed: seeded with the GPU's state entering the window, it
    steps substep after substep on all cores (OpenMP over vertices, constraints,
    triangles, leaves, pairs and candidates; results identical to one thread) until
    the time budget is spent, then on one thread when one substep fits the budget."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # the CPU restatement, used only as the reported baseline

    def fresh():
        o = oracle.OracleSim(scene, 32)
        if state is not None:
            o.set_state(state["x"], state["v"], state["p"], state["tri"])
            o.set_parents(state["parents"])
        return o

    def run(threads, budget, max_steps=64):
        used = oracle.OracleSim.set_threads(threads)
        o = fresh()
        done, elapsed = 0, 0.0
        while elapsed < budget and done < max_steps:
            elapsed += o.time_steps(1)
            done += 1
        return used, done, elapsed

    where = (f"substeps {state['substep'] + 1}.. of the drop, seeded with the GPU's state "
             f"entering the timed window" if state is not None else "first substeps of the drop")
    default = oracle.OracleSim.set_threads(0)
    try:
        cores, done, elapsed = run(default, budget_s)
        single = None
        if elapsed / done * cores <= 2 * budget_s:
            _, done1, elapsed1 = run(1, budget_s / 2)
            single = dict(value=done1 / elapsed1, unit="substeps/s", cores=1,
                          sample=f"{done1} substeps, 1 thread, {elapsed1:.1f} s")
    finally:
        oracle.OracleSim.set_threads(default)
    return dict(value=done / elapsed, unit="substeps/s", cores=cores, kind="port",
                sample=f"{done} substeps of the 256x256 sphere-drop workload ({where}), FP32 "
                       f"oracle (C++ restatement of the reference path), {cores} OpenMP "
                       f"threads, {elapsed:.1f} s",
                single_thread=single)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20,
                    help="substeps timed: the last ones of the 1,200-substep drop (all of "
                         "them when 1,200 or more), after an untimed pre-roll from its start")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--kernel-reps", type=int, default=150)
    ap.add_argument("--autotune", type=int, default=0,
                    help="1: pick the single cloth's launch shapes with cm_autotune after the "
                         "warm-up (tuned on the early, flat state; off by default because "
                         "the late state prefers other shapes); the batched leg always tunes")
    ap.add_argument("--batch-cloths", type=int, default=512,
                    help="cloths of the batched-vs-plain leg (0 skips it)")
    ap.add_argument("--batch-steps", type=int, default=300,
                    help="substeps of the batched-vs-plain leg (300 = the workload's 300 "
                         "frames at one substep per frame)")
    ap.add_argument("--large-size", type=int, default=1024,
                    help="side of the crumpled large cloth (0 skips that leg)")
    ap.add_argument("--large-steps", type=int, default=200)
    ap.add_argument("--adj-size", type=int, default=128,
                    help="side of the fitting cloth of the adjoint leg (0 skips it)")
    ap.add_argument("--adj-substeps", type=int, default=60)
    ap.add_argument("--adj-every", type=int, default=10)
    ap.add_argument("--adj-batch", type=int, default=64,
                    help="64x64 cloths of the batched adjoint leg (0 skips it)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 256 MiB device copy three times (the reference "
                         "kernel for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: the batched and single-cloth paths shard by cloth "
              f"with no collectives, so --gpus {args.gpus} is not measured here.",
              file=sys.stderr)
        sys.exit(2)

    from licenta_amd import configs
    from licenta_amd._abi import lib
    from licenta_amd.body import ClothBody, warning_names
    import ctypes as C

    scene = configs.config_scene(1, args.size)
    body = ClothBody.from_scene(scene)
    x0 = scene["xyz"].astype(np.float32)
    tri0 = body.buffer("tri").copy()
    body.fixed_update(args.warmup)  # untimed: graph capture, caches
    body.sync()
    tuned = None
    if args.autotune:
        tuned = body.autotune(3)  # launch shapes only
    # restart the drop from its initial state; the run is the workload's 600 frames
    body.set_state(x0, np.zeros_like(x0), tri=tri0)
    body.sync()
    run_len = max(FRAMES_SUBSTEPS, args.steps)
    preroll = run_len - args.steps
    early_n = min(EARLY_WINDOW, preroll)
    t_early = _timed(body, early_n)
    t_rest = _timed(body, preroll - early_n)
    counts_pre = body.counts()
    warn_pre = body.warnings()
    # the state entering the window, for the CPU baseline on the same substeps
    window_state = None if args.no_cpu else dict(
        x=body.buffer("x")[:, :3].astype(np.float64), p=body.buffer("p")[:, :3].astype(np.float64),
        v=body.buffer("v")[:, :3].astype(np.float64), tri=body.buffer("tri").copy(),
        parents=body.buffer("parents").copy(), substep=preroll)
    # the timed window: the last `steps` substeps of the run (the draped, contact-heavy
    # state when steps is below the run's length)
    t_win = _timed(body, args.steps)
    ms_per_step = t_win * 1e3 / args.steps
    counts = body.counts()
    warn = body.warnings()
    tree = body.tree_coverage()
    # per-stage split of the window's state (eager launches between HIP events, 10 more
    # substeps)
    stage_names = ("predict", "triangles", "sort", "tree", "pairs", "narrow", "solver", "integrate")
    stage = (C.c_float * 8)()
    nstage = 10
    if lib().cm_stage_times(body.h, nstage, stage):
        raise RuntimeError("cm_stage_times failed")
    stage_ms = {k: round(stage[i] / nstage, 4) for i, k in enumerate(stage_names)}
    counts_after = body.counts()
    iters = scene["params"]["solver_iterations"]
    contacts = counts_after["vf"] + counts_after["ee"]
    # live HIP-event timing of the window's three largest kernels on its final state:
    # the CCD plan pass and the BVH traversal (one launch each per substep) and the
    # Jacobi iteration (15 per substep)
    avg = C.c_float()
    if lib().cm_time_kernel(body.h, KERNEL_CCD_PLAN, args.kernel_reps, C.byref(avg)):
        raise RuntimeError("cm_time_kernel(ccd plan) failed")
    plan_ms = float(avg.value)
    if lib().cm_time_kernel(body.h, KERNEL_TRAVERSE, args.kernel_reps, C.byref(avg)):
        raise RuntimeError("cm_time_kernel(traverse) failed")
    trav_ms = float(avg.value)
    if lib().cm_time_kernel(body.h, KERNEL_SOLVE, args.kernel_reps, C.byref(avg)):
        raise RuntimeError("cm_time_kernel(solve) failed")
    solve_ms = float(avg.value)
    n, s, bd, t = counts_after["n"], counts_after["s"], counts_after["b"], counts_after["t"]
    per_iter, per_substep = algorithmic_bytes(n, s, bd, t, iters, contacts, counts_after["pairs"])
    trav_bytes = 88 * t + 8 * counts_after["pairs"]
    trav_gbs = trav_bytes / (trav_ms * 1e-3) / 1e9
    solve_gbs = per_iter / (solve_ms * 1e-3) / 1e9
    # the CCD plan pass: each candidate's 16-byte id record read and its near flag
    # written (positions and velocities are L2 hits; the near candidates' 24-byte plans,
    # about 7% of the candidates, are left out, so this is a lower bound)
    ncand = counts_after["vf_cand"] + counts_after["ee_cand"]
    plan_bytes = 17 * ncand
    plan_gbs = plan_bytes / (plan_ms * 1e-3) / 1e9
    # the dominant kernel of the window by time per substep
    trav_share, solve_share = trav_ms, solve_ms * iters
    shares = dict(plan=plan_ms, traverse=trav_share, solve=solve_share)
    dominant = max(shares, key=shares.get)
    prof = _traffic_profiles(args.size)
    roof_trav = dict(bound="hbm", achieved=trav_gbs, peak=HBM_PEAK_GBS, unit="GB/s",
                     frac=trav_gbs / HBM_PEAK_GBS, traffic=prof["traverse"][0],
                     traffic_source=prof["traverse"][1],
                     kernel="k_traverse (BVH traversal, one launch per substep)",
                     bytes_per_launch=trav_bytes, avg_launch_ms=trav_ms,
                     ms_per_substep=trav_share)
    roof_solve = dict(bound="hbm", achieved=solve_gbs, peak=HBM_PEAK_GBS, unit="GB/s",
                      frac=solve_gbs / HBM_PEAK_GBS, traffic=prof["solve"][0],
                      traffic_source=prof["solve"][1],
                      kernel="k_solve (fused Jacobi iteration, 15 launches per substep)",
                      bytes_per_launch=per_iter, avg_launch_ms=solve_ms,
                      ms_per_substep=solve_share)
    roof_plan = dict(bound="hbm", achieved=plan_gbs, peak=HBM_PEAK_GBS, unit="GB/s",
                     frac=plan_gbs / HBM_PEAK_GBS, traffic=prof["plan"][0],
                     traffic_source=prof["plan"][1],
                     kernel="k_ccd_plan (CCD plan, secant roots and near test, one launch "
                            "per substep; VALU-bound by the secant searches' divergence)",
                     bytes_per_launch=plan_bytes, avg_launch_ms=plan_ms,
                     ms_per_substep=plan_ms, candidates=ncand)
    roofs = dict(plan=roof_plan, traverse=roof_trav, solve=roof_solve)
    roof = roofs[dominant]
    full_s = t_early + t_rest + t_win
    out = dict(
        metric=METRIC,
        value=1e3 / ms_per_step,
        unit="substeps/s",
        n_gpus=1,
        steps=args.steps,
        warmup=args.warmup,
        ms_per_step=ms_per_step,
        higher_is_better=True,
        scaling="weak",
        vs_baseline=None,
        dtype="f32",
        data="synthetic",
        config=dict(workload=f"sphere_drop_{args.size}x{args.size}", vertices=n, triangles=t,
                    distance_constraints=s, bend_constraints=bd, iterations=iters,
                    substep_dt=1 / 240, morton_bits=30, frames=run_len // 2,
                    timed_window=f"substeps {preroll + 1}-{run_len} of {run_len}",
                    parallelism="single cloth, one GPU"),
        roofline=roof,
        roofline_other=[r for k, r in roofs.items() if k != dominant],
        vertex_updates_per_s=n * 1e3 / ms_per_step,
        substep_roofline_frac=per_substep / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS,
        full_run=dict(substeps=run_len, seconds=full_s, substeps_per_s=run_len / full_s,
                      ms_per_substep=full_s * 1e3 / run_len,
                      note="whole 600-frame drop from the initial state, wall clock with a "
                           "sync between the early window, the rest and the timed window"),
        early_window=dict(substeps=early_n, substeps_per_s=(early_n / t_early) if early_n else None,
                          note="first substeps of the drop (flat cloth above the sphere)"),
        window_start=dict(pairs=counts_pre["pairs"], vf=counts_pre["vf"], ee=counts_pre["ee"],
                          warnings=warn_pre),
        last_substep=dict(pairs=counts["pairs"], vf=counts["vf"], ee=counts["ee"],
                          warnings=warn, warning_names=warning_names(warn), tree=tree),
        autotuned=tuned,
        stage_ms_after_run=stage_ms,
    )
    body.close()
    _progress(f"single-cloth leg: {out['value']:.1f} substeps/s")
    out["batched"] = batched_leg(args.batch_cloths, args.batch_steps) if args.batch_cloths > 0 \
        else None
    _progress("batched leg done")
    out["large"] = large_leg(args.large_size, args.large_steps) if args.large_size > 0 else None
    _progress("large-cloth leg done")
    out["adjoint"] = adjoint_leg(args.adj_size, args.adj_substeps, args.adj_every) \
        if args.adj_size > 0 else None
    _progress("adjoint leg done")
    out["batched_adjoint"] = batched_adjoint_leg(args.adj_batch, 30, 10) \
        if args.adj_batch > 0 else None
    _progress("batched adjoint leg done")
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(configs.config_scene(1, args.size), args.cpu_budget,
                                           window_state)
        out["cpu_baseline"]["cpu_model"] = _cpu_model()
        out["cpu_baseline"]["nproc"] = os.cpu_count()
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


def _progress(msg: str) -> None:
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def batched_leg(ncloth: int, steps: int, warmup: int = 3) -> dict:
    """Cloth-steps/s of the 512-cloth batch: one batched body against one body per
    cloth (each on its own stream, all launched before any is waited for)."""
    from licenta_amd import configs
    from licenta_amd.body import ClothBody

    scenes = configs.batched_scene(2, count=ncloth)
    nv = sum(len(s["xyz"]) for s in scenes)
    nt = sum(len(s["tri"]) for s in scenes)
    batch = ClothBody.batch(scenes)
    batch.fixed_update(warmup)
    batch.sync()
    tuned = batch.autotune(3)  # launch shapes for this batch (results unchanged)
    batch.fixed_update(1)  # captures the substep graph again
    batch.sync()
    t0 = time.perf_counter()
    batch.fixed_update(steps)
    batch.sync()
    tb = time.perf_counter() - t0
    bcounts = batch.counts()
    bwarn = batch.warnings()
    batch.close()
    singles = [ClothBody.from_scene(s) for s in scenes]
    for b in singles:
        b.fixed_update(warmup)
    for b in singles:
        b.sync()
    t0 = time.perf_counter()
    for b in singles:
        b.fixed_update(steps)
    for b in singles:
        b.sync()
    tp = time.perf_counter() - t0
    swarn = 0
    for b in singles:
        swarn |= b.warnings()
        b.close()
    batched = ncloth * steps / tb
    plain = ncloth * steps / tp
    return dict(workload=f"{ncloth} cloths, sides U{{16..128}}, seed 2", cloths=ncloth,
                vertices=nv, triangles=nt, steps=steps, unit="cloth-steps/s",
                batched_cloth_steps_per_s=batched, plain_cloth_steps_per_s=plain,
                batched_over_plain=batched / plain, batched_ms_per_step=tb * 1e3 / steps,
                plain_ms_per_step=tp * 1e3 / steps, autotuned=tuned,
                last_substep=dict(pairs=bcounts["pairs"], vf=bcounts["vf"], ee=bcounts["ee"]),
                warnings=dict(batched=bwarn, plain=swarn,
                              candidate_cap_hit=bool((bwarn | swarn) & WARN_CANDIDATE_CAP),
                              note="CM_WARN_* bits accumulated over the run; with "
                                   "candidate_cap_hit the two sides may skip different "
                                   "narrow-phase work and the ratio is not comparable"))


def _large_run(sc: dict, steps: int, warmup: int) -> dict:
    from licenta_amd.body import ClothBody, warning_names

    t0 = time.perf_counter()
    body = ClothBody.from_scene(sc)
    t_create = time.perf_counter() - t0
    body.fixed_update(warmup)
    body.sync()
    t0 = time.perf_counter()
    body.fixed_update(steps)
    body.sync()
    dt = time.perf_counter() - t0
    c = body.counts()
    w = body.warnings()
    tree = body.tree_coverage()
    body.close()
    return dict(c=c, dt=dt, create=t_create,
                last=dict(pairs=c["pairs"], vf=c["vf"], ee=c["ee"], warnings=w,
                          warning_names=warning_names(w), tree=tree))


def large_leg(size: int, steps: int, warmup: int = 3) -> dict:
    """The crumpled large cloth: 1024x1024 vertices, spacing 1 mm, y displaced by
    8 random sinusoids (seed 4), thickness 0.5 mm, 20 iterations, 1 substep per
    frame at dt 1/60 s, 200 frames, full Morton/sort/LBVH rebuild every substep.
    The reference's split (midpoint on tied 30-bit codes) leaves leaves out of the
    tree on this grid; the same run with the index-augmented Karras split is
    reported beside it (corrected_split)."""
    from licenta_amd import configs

    sc = configs.config_scene(4, size)
    r = _large_run(sc, steps, warmup)
    c, dt = r["c"], r["dt"]
    params = dict(sc["params"], ref_compat=sc["params"]["ref_compat"] & ~configs.RC_MIDPOINT_SPLIT)
    k = _large_run(dict(sc, params=params), steps, warmup)
    return dict(workload=f"crumpled_{size}x{size}", vertices=c["n"], triangles=c["t"],
                distance_constraints=c["s"], bend_constraints=c["b"], iterations=20,
                steps=steps, warmup=warmup, substeps_per_s=steps / dt,
                vertex_updates_per_s=c["n"] * steps / dt, ms_per_step=dt * 1e3 / steps,
                create_s=r["create"], last_substep=r["last"],
                corrected_split=dict(substeps_per_s=steps / k["dt"],
                                     ms_per_step=k["dt"] * 1e3 / steps, last_substep=k["last"],
                                     note="index-augmented Karras split (CM_RC_MIDPOINT_SPLIT "
                                          "cleared): every leaf in the tree"))


def _time_rollout(body, substeps, every, x0, v0, tri0, lx_of, reps=3):
    """Median wall times of: plain forward (cm_step), checkpointed forward, adjoint."""
    plain, fwd, adj = [], [], []
    for rep in range(reps + 1):  # the first round warms up (graph capture, buffers)
        body.set_state(x0, v0, tri=tri0)
        body.sync()
        t0 = time.perf_counter()
        body.fixed_update(substeps)
        body.sync()
        t1 = time.perf_counter()
        body.set_state(x0, v0, tri=tri0)
        body.sync()
        t2 = time.perf_counter()
        body.forward_ckpt(substeps, every)
        body.sync()
        t3 = time.perf_counter()
        lx = lx_of(body.positions())
        t4 = time.perf_counter()
        grads = body.adjoint(lx)
        t5 = time.perf_counter()
        if rep:
            plain.append(t1 - t0)
            fwd.append(t3 - t2)
            adj.append(t5 - t4)
    med = lambda a: float(np.median(a))
    return med(plain), med(fwd), med(adj), grads


def adjoint_leg(size: int, substeps: int, every: int) -> dict:
    """Adjoint/forward time ratio of the fitting workload: a 128x128 pinned cloth,
    60 substeps, 10 iterations, self-collision, from the fit's start (stiffness 0.65
    everywhere, zero velocity), loss = mean squared vertex error against the
    ground-truth run's final positions, checkpoints every 10 substeps."""
    from licenta_amd import configs
    from licenta_amd.body import ClothBody

    sc = configs.config_scene(3, size)
    body = ClothBody.from_scene(sc)  # ground-truth stiffness and velocities
    tri0 = body.buffer("tri").copy()
    body.set_state(sc["xyz"], sc["v0"], tri=tri0)
    body.fixed_update(substeps)
    target = body.positions().astype(np.float64)
    c = body.counts()
    body.set_stiffness(np.full(c["s"], 0.65, np.float32))
    zero = np.zeros_like(sc["xyz"])
    n = len(target)
    plain, fwd, adj, grads = _time_rollout(
        body, substeps, every, sc["xyz"], zero, tri0,
        lambda x: (2.0 * (x - target) / n).astype(np.float32))
    counts = body.counts()
    body.close()
    return dict(workload=f"fit_{size}x{size}: {substeps} substeps, checkpoint every {every}",
                vertices=n, distance_constraints=c["s"], bend_constraints=c["b"],
                forward_ms=plain * 1e3, forward_ckpt_ms=fwd * 1e3, adjoint_ms=adj * 1e3,
                adjoint_over_forward=adj / plain,
                forward_plus_adjoint_over_forward=(fwd + adj) / plain,
                last_substep_contacts=counts["vf"] + counts["ee"],
                grad_norm_k_dist=float(np.linalg.norm(grads[2])))


def batched_adjoint_leg(ncloth: int, substeps: int, every: int) -> dict:
    """Forward + adjoint of 64 cloths of 64x64 in one batched body (per-cloth bend
    stiffness), 30 substeps, checkpoints every 10; loss against the run's own
    final positions shifted by 1 mm."""
    from licenta_amd import configs
    from licenta_amd.body import ClothBody

    scenes = configs.batched_scene(4, count=ncloth)
    body = ClothBody.batch(scenes)
    x0 = np.concatenate([s["xyz"] for s in scenes])
    tri0 = body.buffer("tri").copy()
    zero = np.zeros_like(x0)
    n = len(x0)
    plain, fwd, adj, _ = _time_rollout(
        body, substeps, every, x0, zero, tri0,
        lambda x: (2.0 * (x - (x + 1e-3)) / n).astype(np.float32))
    body.close()
    return dict(workload=f"{ncloth} cloths of 64x64, {substeps} substeps, checkpoint every "
                         f"{every}", vertices=n, forward_ms=plain * 1e3,
                forward_ckpt_ms=fwd * 1e3, adjoint_ms=adj * 1e3,
                adjoint_over_forward=adj / plain,
                forward_plus_adjoint_over_forward=(fwd + adj) / plain)


def calibrate_copy():
    import torch

    n = 64 * 1024 * 1024  # floats: 256 MiB
    a = torch.empty(n, device="cuda", dtype=torch.float32).fill_(1.0)
    c = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        c.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps(dict(calibration="device copy", copies=3, bytes_read_per_copy=4 * n,
                          bytes_written_per_copy=4 * n)))


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


if __name__ == "__main__":
    main()
