"""Run the sphere-drop workload for a number of substeps (for profilers).

    python -m licenta_amd.workload --size 256 --substeps 640 --kernel-reps 150

Creates the 256x256 sphere-drop cloth, writes its initial state back with
cm_set_state (which synchronises the body's stream before the substep graph is
captured), runs ``--substeps`` substeps through the captured substep graph, then
launches the Jacobi kernel ``--kernel-reps`` times on the final state with HIP events
around it (as bench.py does for its roofline) and prints one JSON line with the
counts and the average launch time.
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
    args = ap.parse_args()
    scene = configs.config_scene(1, args.size)
    body = ClothBody.from_scene(scene)
    body.set_state(scene["xyz"], np.zeros_like(scene["xyz"]))
    body.fixed_update(args.substeps)
    body.sync()
    avg = C.c_float()
    if args.kernel_reps > 0 and lib().cm_time_kernel(body.h, 0, args.kernel_reps, C.byref(avg)):
        raise RuntimeError("cm_time_kernel failed")
    out = dict(workload=f"sphere_drop_{args.size}x{args.size}", substeps=args.substeps,
               solve_avg_ms=float(avg.value), counts=body.counts(), warnings=body.warnings())
    body.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
