"""Throughput of the PBD cloth substep with Morton/LBVH self-collision on one MI355X.

Workload: one 256x256-vertex cloth (65,536 vertices, 130,050 triangles, 260,610
structural+shear distance constraints, 194,565 dihedral bends) dropped from 0.5 m
onto a static sphere of radius 0.3 m, thickness 0.002 m, dt 1/120 s with 2
substeps per frame (substep dt 1/240 s), 15 Jacobi iterations, 30-bit Morton codes
in the centroid box, LBVH self-collision with VF and EE contacts every substep.
A step is one substep; the default run times the full 600 frames (1,200 substeps).

Prints one JSON line. ``roofline`` covers the dominant kernel (the fused Jacobi
iteration, 15 launches per substep) with its algorithmic bytes per launch
(36 N + 16 S + 24 Bd + 48 C) over its average launch time measured with HIP events
on the body's stream; ``cpu_baseline`` times the FP32 CPU oracle on a bounded
sample of the same workload on this host's cores.
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


def algorithmic_bytes(n, s, bd, t, iters, contacts, pairs):
    per_iter = 36 * n + 16 * s + 24 * bd + 48 * contacts
    substep = 112 * n + iters * per_iter + 300 * t + 32 * n + 8 * pairs
    return per_iter, substep


def cpu_baseline(scene, budget_s: float):
    """The FP32 oracle (the C++ restatement of the reference path) on this host:
    all cores (OpenMP over vertices, constraints, triangles, leaves, pairs and
    candidates; results identical to one thread), then one thread, each on the
    first substeps of the same workload within the time budget."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # the CPU restatement, used only as the reported baseline

    def run(threads, budget):
        used = oracle.OracleSim.set_threads(threads)
        o = oracle.OracleSim(scene, 32)
        done, elapsed = 0, 0.0
        while elapsed < budget and done < 64:
            elapsed += o.time_steps(1)
            done += 1
        return used, done, elapsed

    default = oracle.OracleSim.set_threads(0)
    try:
        cores, done, elapsed = run(default, budget_s)
        _, done1, elapsed1 = run(1, budget_s / 2)
    finally:
        oracle.OracleSim.set_threads(default)
    return dict(value=done / elapsed, unit="substeps/s", cores=cores, kind="port",
                sample=f"first {done} substeps of the same 256x256 sphere-drop workload, FP32 "
                       f"oracle (C++ restatement of the reference path), {cores} OpenMP "
                       f"threads, {elapsed:.1f} s",
                single_thread=dict(value=done1 / elapsed1, unit="substeps/s", cores=1,
                                   sample=f"first {done1} substeps, 1 thread, {elapsed1:.1f} s"))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1200)
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
    ap.add_argument("--batch-steps", type=int, default=30)
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
    from licenta_amd.body import ClothBody
    import ctypes as C

    scene = configs.config_scene(1, args.size)
    body = ClothBody.from_scene(scene)
    body.fixed_update(args.warmup)
    body.sync()
    tuned = None
    if args.autotune:
        tuned = body.autotune(3)  # launch shapes only
        body.fixed_update(1)  # captures the tuned substep graph
        body.sync()
    t0 = time.perf_counter()
    body.fixed_update(args.steps)
    body.sync()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    counts = body.counts()
    warn = body.warnings()
    # per-stage split of the last substeps (eager launches between HIP events)
    stage_names = ("predict", "triangles", "sort", "tree", "pairs", "narrow", "solver", "integrate")
    stage = (C.c_float * 8)()
    nstage = 10
    if lib().cm_stage_times(body.h, nstage, stage):
        raise RuntimeError("cm_stage_times failed")
    stage_ms = {k: round(stage[i] / nstage, 4) for i, k in enumerate(stage_names)}
    # dominant kernel, timed live with HIP events on the body's stream
    avg = C.c_float()
    rc = lib().cm_time_kernel(body.h, 0, args.kernel_reps, C.byref(avg))
    if rc:
        raise RuntimeError("cm_time_kernel failed")
    iters = scene["params"]["solver_iterations"]
    contacts = counts["vf"] + counts["ee"]
    per_iter, per_substep = algorithmic_bytes(counts["n"], counts["s"], counts["b"], counts["t"],
                                              iters, contacts, counts["pairs"])
    solve_ms = float(avg.value)
    achieved = per_iter / (solve_ms * 1e-3) / 1e9
    traffic, traffic_note = None, None
    tf = os.path.join(ROOT, "profiles", "k_solve_sphere_drop_256_traffic.json")
    if os.path.exists(tf) and args.size == 256:
        tj = json.load(open(tf))
        traffic = tj.get("bytes_per_launch")
        traffic_note = (f"rocprofv3 FETCH_SIZE x {tj.get('fetch_scale')} + WRITE_SIZE x "
                        f"{tj.get('write_scale')} (256 MiB copy calibration), {tj.get('state')}")
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
        config=dict(workload=f"sphere_drop_{args.size}x{args.size}", vertices=counts["n"],
                    triangles=counts["t"], distance_constraints=counts["s"],
                    bend_constraints=counts["b"], iterations=iters, substep_dt=1 / 240,
                    morton_bits=30, parallelism="single cloth, one GPU"),
        roofline=dict(bound="hbm", achieved=achieved, peak=HBM_PEAK_GBS, unit="GB/s",
                      frac=achieved / HBM_PEAK_GBS, traffic=traffic, traffic_source=traffic_note,
                      kernel="k_solve (fused Jacobi iteration)", bytes_per_launch=per_iter,
                      avg_launch_ms=solve_ms),
        vertex_updates_per_s=counts["n"] * 1e3 / ms_per_step,
        substep_roofline_frac=per_substep / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS,
        last_substep=dict(pairs=counts["pairs"], vf=counts["vf"], ee=counts["ee"],
                          warnings=warn),
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
        out["cpu_baseline"] = cpu_baseline(configs.config_scene(1, args.size), args.cpu_budget)
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
    for b in singles:
        b.close()
    batched = ncloth * steps / tb
    plain = ncloth * steps / tp
    return dict(workload=f"{ncloth} cloths, sides U{{16..128}}, seed 2", cloths=ncloth,
                vertices=nv, triangles=nt, steps=steps, unit="cloth-steps/s",
                batched_cloth_steps_per_s=batched, plain_cloth_steps_per_s=plain,
                batched_over_plain=batched / plain, batched_ms_per_step=tb * 1e3 / steps,
                plain_ms_per_step=tp * 1e3 / steps, autotuned=tuned,
                last_substep=dict(pairs=bcounts["pairs"], vf=bcounts["vf"], ee=bcounts["ee"]))


def large_leg(size: int, steps: int, warmup: int = 3) -> dict:
    """The crumpled large cloth: 1024x1024 vertices, spacing 1 mm, y displaced by
    8 random sinusoids (seed 4), thickness 0.5 mm, 20 iterations, 1 substep per
    frame at dt 1/60 s, 200 frames, full Morton/sort/LBVH rebuild every substep."""
    from licenta_amd import configs
    from licenta_amd.body import ClothBody

    sc = configs.config_scene(4, size)
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
    body.close()
    return dict(workload=f"crumpled_{size}x{size}", vertices=c["n"], triangles=c["t"],
                distance_constraints=c["s"], bend_constraints=c["b"], iterations=20,
                steps=steps, warmup=warmup, substeps_per_s=steps / dt,
                vertex_updates_per_s=c["n"] * steps / dt, ms_per_step=dt * 1e3 / steps,
                create_s=t_create,
                last_substep=dict(pairs=c["pairs"], vf=c["vf"], ee=c["ee"], warnings=w))


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
