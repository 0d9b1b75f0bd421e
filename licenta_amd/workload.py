"""Run the sphere-drop workload for a number of substeps (for profilers).

    python -m licenta_amd.workload --size 256 --substeps 640 --kernel-reps 150
    python -m licenta_amd.workload --substeps 1190 --kernel-reps 0 --save s.npz
    python -m licenta_amd.workload --load s.npz --substeps 10 --time-traverse 50

    python -m licenta_amd.workload --adjoint --size 128 --substeps 60 --every 10
    python -m licenta_amd.workload --batch 512 --substeps 8
    python -m licenta_amd.workload --batch 512 --substeps 290 --kernel-reps 0 --save b.npz

Creates the 256x256 sphere-drop cloth, writes its initial state back with
cm_set_state (which synchronises the body's stream before the substep graph is
captured), runs ``--substeps`` substeps through the captured substep graph, then
launches the Jacobi kernel ``--kernel-reps`` times on the final state with HIP events
around it (as bench.py does for its roofline) and prints one JSON line with the
counts and the average launch time. --save / --load carry a state between runs, so
a profiler can cover a few substeps of a late state without the thousand substeps
that lead there. With --adjoint: the fitting cloth (pinned
corners, stiffness 0.65, zero velocity), then ``--reps`` rounds of a checkpointed
forward of ``--substeps`` substeps and its adjoint (loss against the ground-truth
run's final positions).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json

import numpy as np

from licenta_amd import configs
from licenta_amd._abi import lib
from licenta_amd.body import ClothBody


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--substeps", type=int, default=640)
    ap.add_argument("--kernel-reps", type=int, default=150)
    ap.add_argument("--adjoint", action="store_true")
    ap.add_argument("--every", type=int, default=10)
    ap.add_argument("--reps", type=int, default=2)
    ap.add_argument("--batch", type=int, default=0,
                    help="step the N-cloth batch of uneven cloths instead (seed 2)")
    ap.add_argument("--save", default=None,
                    help="after the substeps, write the state (x, v, p, triangle order, "
                         "parents) to this .npz")
    ap.add_argument("--load", default=None,
                    help="start from a state written by --save instead of the initial one")
    ap.add_argument("--time-traverse", type=int, default=0,
                    help="also launch the BVH traversal this many times on the final state "
                         "(HIP events around them)")
    args = ap.parse_args()
    if args.adjoint:
        adjoint(args)
        return
    if args.batch:
        batch(args)
        return
    scene = configs.config_scene(1, args.size)
    body = ClothBody.from_scene(scene)
    start = 0
    if args.load:
        st = np.load(args.load)
        body.set_state(st["x"], st["v"], p=st["p"], tri=st["tri"])
        start = int(st["substep"])
    else:
        body.set_state(scene["xyz"], np.zeros_like(scene["xyz"]))
    body.fixed_update(args.substeps)
    body.sync()
    if args.save:
        np.savez(args.save, x=body.buffer("x")[:, :3], v=body.buffer("v")[:, :3],
                 p=body.buffer("p")[:, :3], tri=body.buffer("tri"),
                 parents=body.buffer("parents"), substep=start + args.substeps)
    trav = C.c_float()
    if args.time_traverse > 0 and lib().cm_time_kernel(body.h, 2, args.time_traverse,
                                                       C.byref(trav)):
        raise RuntimeError("cm_time_kernel(traverse) failed")
    avg = C.c_float()
    if args.kernel_reps > 0 and lib().cm_time_kernel(body.h, 0, args.kernel_reps, C.byref(avg)):
        raise RuntimeError("cm_time_kernel failed")
    out = dict(workload=f"sphere_drop_{args.size}x{args.size}", start_substep=start,
               substeps=args.substeps, solve_avg_ms=float(avg.value),
               traverse_avg_ms=float(trav.value), counts=body.counts(), warnings=body.warnings())
    body.close()
    print(json.dumps(out))


def batch(args) -> None:
    scenes = configs.batched_scene(2, count=args.batch)
    body = ClothBody.batch(scenes)
    start = 0
    if args.load:
        st = np.load(args.load)
        body.set_state(st["x"], st["v"], p=st["p"], tri=st["tri"])
        start = int(st["substep"])
    else:
        x0 = np.concatenate([s["xyz"] for s in scenes])
        body.set_state(x0, np.zeros_like(x0))
    body.fixed_update(args.substeps)
    body.sync()
    if args.save:
        np.savez(args.save, x=body.buffer("x")[:, :3], v=body.buffer("v")[:, :3],
                 p=body.buffer("p")[:, :3], tri=body.buffer("tri"),
                 substep=start + args.substeps)
    trav = C.c_float()
    if args.time_traverse > 0 and lib().cm_time_kernel(body.h, 2, args.time_traverse,
                                                       C.byref(trav)):
        raise RuntimeError("cm_time_kernel(traverse) failed")
    avg = C.c_float()
    if args.kernel_reps > 0 and lib().cm_time_kernel(body.h, 0, args.kernel_reps, C.byref(avg)):
        raise RuntimeError("cm_time_kernel failed")
    print(json.dumps(dict(workload=f"batch of {args.batch} cloths", start_substep=start,
                          substeps=args.substeps, solve_avg_ms=float(avg.value),
                          traverse_avg_ms=float(trav.value), counts=body.counts(),
                          warnings=body.warnings())))
    body.close()


def adjoint(args) -> None:
    import time

    sc = configs.config_scene(3, args.size)
    body = ClothBody.from_scene(sc)
    tri0 = body.buffer("tri").copy()
    body.set_state(sc["xyz"], sc["v0"], tri=tri0)
    body.fixed_update(args.substeps)
    target = body.positions().astype(np.float64)
    body.set_stiffness(np.full(body.counts()["s"], 0.65, np.float32))
    times = []
    for _ in range(args.reps):
        body.set_state(sc["xyz"], np.zeros_like(sc["xyz"]), tri=tri0)
        t0 = time.perf_counter()
        body.forward_ckpt(args.substeps, args.every)
        x = body.positions()
        t1 = time.perf_counter()
        body.adjoint((2.0 * (x - target) / len(x)).astype(np.float32))
        times.append((t1 - t0, time.perf_counter() - t1))
    body.close()
    print(json.dumps(dict(workload=f"fit_{args.size}x{args.size}", substeps=args.substeps,
                          every=args.every, forward_and_adjoint_s=times)))


if __name__ == "__main__":
    main()
