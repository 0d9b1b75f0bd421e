"""Synthetic cloth scenes: grid meshes, constraint sets and parameter blocks.

Two families:

* ``reference_scene`` restates the reference's own scene setup
  (LicentaV2/LicentaV2/Managers/MastersSimulationManager.cpp:25-95 and the grid of
  Rendering/VisualBodyFactory.cpp:343-391: alternating quad diagonal, xy-plane
  placed in [0.25, 0.5]^2 at z = 0.5, two pinned corners, thickness |p0-p1|/3 as
  Physics/CudaDeformableBody.cpp:86), with every reference default switched on
  (ground, world clamp, damping, LRA, 63-bit raw Morton codes). Wind is off: the
  reference draws it from std::rand each frame (Physics/CudaPBD.cu:613-619).
* ``config_scene(k)`` builds the five benchmark workloads (plain 64^2, plain 256^2
  over a sphere, 512 batched cloths, 128^2 differentiable, 1024^2 crumpled /
  64 x 64^2 batched-gradient) from one numpy PCG64 stream per workload.

A scene is a plain dict of numpy arrays plus a ``cm_params``-shaped dict; the
same dict feeds the HIP library and the CPU oracle, so both see identical
FP32 inputs.
"""
from __future__ import annotations

import math

import numpy as np

# flag values mirror include/clothmi.h
F_SELF_COLLISION = 1 << 0
F_GROUND = 1 << 1
F_WORLD_CLAMP = 1 << 2
F_DAMPING = 1 << 3
F_LRA = 1 << 4
F_SPHERE = 1 << 5
F_EE_CONTACTS = 1 << 6
F_FRICTION = 1 << 7
F_BEND_FP64 = 1 << 8
F_VERTEX_NORMALS = 1 << 9
F_REVERT_TOC = 1 << 10
F_DCD = 1 << 11

RC_STALE_LEAF_AABB = 1 << 0
RC_BEND_ABS = 1 << 1
RC_MIDPOINT_SPLIT = 1 << 2
RC_RIGID_FRICTION_TAIL = 1 << 3
RC_ALL = RC_STALE_LEAF_AABB | RC_BEND_ABS | RC_MIDPOINT_SPLIT | RC_RIGID_FRICTION_TAIL


def default_params() -> dict:
    """The reference's ClothParams literal (Managers/MastersSimulationManager.cpp:33-62)
    with its compile-time constants (Core/Utils.hpp:16-19,
    Physics/CudaDeformableUtils.cuh:16-17)."""
    return dict(
        timestep=1.0 / 60.0,
        gravity=(0.0, -0.0981, 0.0),
        k_stretch=0.99,
        k_bend=0.1,
        thickness=0.0,
        solver_iterations=6,
        ground_height=0.0,
        world_min=(0.0, 0.0, 0.0),
        world_max=(1.0, 1.0, 1.0),
        global_vel_damp=0.01,
        k_damp=0.05,
        strain_limit=0.15,
        k_friction=0.1,
        pinned_mass=1000.0,
        sphere_center=(0.0, 0.0, 0.0),
        sphere_radius=0.0,
        morton_bits=63,
        morton_normalize=0,
        pair_cap=64,
        flags=(F_SELF_COLLISION | F_GROUND | F_WORLD_CLAMP | F_DAMPING | F_LRA | F_EE_CONTACTS
               | F_BEND_FP64 | F_VERTEX_NORMALS),
        ref_compat=RC_ALL,
    )


# ---------------------------------------------------------------------------
# meshes
# ---------------------------------------------------------------------------
def grid_triangles(rows: int, cols: int, alternating: bool) -> np.ndarray:
    """Triangle indices of a rows x cols vertex grid, vertex (i, j) = i*cols + j.

    alternating=True restates Rendering/VisualBodyFactory.cpp:363-385 (diagonal flips
    with (i+j)%2); False uses the even-quad split for every quad (same diagonal)."""
    i, j = np.meshgrid(np.arange(rows - 1), np.arange(cols - 1), indexing="ij")
    i = i.ravel()
    j = j.ravel()
    a = i * cols + j          # (i, j)
    b = (i + 1) * cols + j    # (i+1, j)
    c = i * cols + j + 1      # (i, j+1)
    d = (i + 1) * cols + j + 1
    even = np.stack([a, b, c, c, b, d], axis=1)
    if alternating:
        odd = np.stack([a, b, d, d, c, a], axis=1)
        sel = ((i + j) % 2 == 0)[:, None]
        quads = np.where(sel, even, odd)
    else:
        quads = even
    return quads.reshape(-1, 3).astype(np.uint32)


def grid_distance_pairs(rows: int, cols: int) -> np.ndarray:
    """Structural (4-neighbour) and shear (both diagonals) distance constraints.

    Order: for each vertex (i, j) in row-major order, (i,j)-(i+1,j), (i,j)-(i,j+1),
    (i,j)-(i+1,j+1), (i+1,j)-(i,j+1) where they exist."""
    out = []
    idx = np.arange(rows * cols).reshape(rows, cols)
    for i in range(rows):
        for j in range(cols):
            if i + 1 < rows:
                out.append((idx[i, j], idx[i + 1, j]))
            if j + 1 < cols:
                out.append((idx[i, j], idx[i, j + 1]))
            if i + 1 < rows and j + 1 < cols:
                out.append((idx[i, j], idx[i + 1, j + 1]))
                out.append((idx[i + 1, j], idx[i, j + 1]))
    return np.asarray(out, dtype=np.int32)


def grid_distance_pairs_fast(rows: int, cols: int) -> np.ndarray:
    """Vectorised grid_distance_pairs (identical order)."""
    idx = np.arange(rows * cols, dtype=np.int64).reshape(rows, cols)
    blocks = []
    # per vertex up to 4 entries; build a (rows, cols, 4, 2) table with -1 for absent
    tab = -np.ones((rows, cols, 4, 2), dtype=np.int64)
    tab[:-1, :, 0, 0] = idx[:-1, :]
    tab[:-1, :, 0, 1] = idx[1:, :]
    tab[:, :-1, 1, 0] = idx[:, :-1]
    tab[:, :-1, 1, 1] = idx[:, 1:]
    tab[:-1, :-1, 2, 0] = idx[:-1, :-1]
    tab[:-1, :-1, 2, 1] = idx[1:, 1:]
    tab[:-1, :-1, 3, 0] = idx[1:, :-1]
    tab[:-1, :-1, 3, 1] = idx[:-1, 1:]
    flat = tab.reshape(-1, 2)
    flat = flat[flat[:, 0] >= 0]
    blocks.append(flat)
    return np.concatenate(blocks).astype(np.int32)


def _jitter(rng: np.random.Generator, n: int, amp: float = 1e-4) -> np.ndarray:
    return rng.uniform(-amp, amp, size=(n, 3))


# ---------------------------------------------------------------------------
# scenes
# ---------------------------------------------------------------------------
def _scene(params, xyz, tri, inv_mass, pinned=None, dist_pairs=None, k_dist=None,
           k_bend=None, name="", **extra) -> dict:
    sc = dict(
        name=name,
        params=params,
        xyz=np.ascontiguousarray(xyz, dtype=np.float32),
        tri=np.ascontiguousarray(tri, dtype=np.uint32),
        inv_mass=np.ascontiguousarray(inv_mass, dtype=np.float32),
        pinned=None if pinned is None else np.ascontiguousarray(pinned, dtype=np.uint8),
        dist_pairs=None if dist_pairs is None else np.ascontiguousarray(dist_pairs, dtype=np.int32),
        k_dist=None if k_dist is None else np.ascontiguousarray(k_dist, dtype=np.float32),
        k_bend=None if k_bend is None else np.ascontiguousarray(k_bend, dtype=np.float32),
    )
    sc.update(extra)
    return sc


def reference_scene(dim: int = 60) -> dict:
    """The thesis scene (Managers/MastersSimulationManager.cpp:25-95) without wind."""
    i, j = np.meshgrid(np.arange(dim), np.arange(dim), indexing="ij")
    u = (i / dim).astype(np.float32).ravel()
    w = (j / dim).astype(np.float32).ravel()
    # model matrix T(0.5) R(pi, z) S(0.25) baked in (Rendering/VisualBody.cpp:25-58)
    xyz = np.stack([0.5 - 0.25 * u, 0.5 - 0.25 * w, np.full_like(u, 0.5)], axis=1).astype(np.float32)
    tri = grid_triangles(dim, dim, alternating=True)
    pinned = np.zeros(dim * dim, dtype=np.uint8)
    pinned[0] = 1
    pinned[(dim - 1) * dim] = 1
    params = default_params()
    vmass = np.float32(0.525) / np.float32(dim * dim)
    inv_mass = np.where(pinned == 1, 0.0, np.float32(1.0) / vmass).astype(np.float32)
    d = xyz[0] - xyz[1]
    params["thickness"] = float(np.float32(np.sqrt(np.float32(d @ d))) / np.float32(3.0))
    return _scene(params, xyz, tri, inv_mass, pinned, name=f"reference_{dim}")


def _plain_params(dt, iters, thickness, morton_bits=30):
    p = default_params()
    p.update(timestep=dt, gravity=(0.0, -9.81, 0.0), k_stretch=1.0, k_bend=0.1,
             thickness=thickness, solver_iterations=iters, morton_bits=morton_bits,
             morton_normalize=1, flags=F_SELF_COLLISION | F_EE_CONTACTS | F_BEND_FP64)
    return p


def flat_grid(n: int, spacing: float, rng: np.random.Generator, y: float = 0.0,
              centered: bool = False, jitter: float = 1e-4, m: int | None = None) -> np.ndarray:
    m = n if m is None else m
    i, j = np.meshgrid(np.arange(n), np.arange(m), indexing="ij")
    off_i = (n - 1) / 2.0 if centered else 0.0
    off_j = (m - 1) / 2.0 if centered else 0.0
    xyz = np.stack([(i.ravel() - off_i) * spacing, np.full(n * m, y),
                    (j.ravel() - off_j) * spacing], axis=1)
    return (xyz + _jitter(rng, n * m, jitter)).astype(np.float32)


def config_scene(k: int, size: int | None = None) -> dict:
    """Benchmark workloads by number: 0 = plain 64x64 cloth with two pinned corners,
    dt 1/60 s, 10 iterations, 120 frames; 1 = 256x256 cloth dropped onto a sphere of
    radius 0.3 m, 2 substeps of 1/240 s, 15 iterations, 600 frames; 3 = 128x128
    fitting cloth with corners pinned, per-constraint stiffness and random initial
    velocities, 60 substeps; 4 = crumpled 1024x1024 cloth, 20 iterations, 200 frames.
    ``size`` overrides the grid side for quick tests."""
    if k == 0:
        n = size or 64
        rng = np.random.default_rng(0)
        xyz = flat_grid(n, 0.01, rng)
        pinned = np.zeros(n * n, dtype=np.uint8)
        pinned[0] = 1
        pinned[(n - 1) * n] = 1
        inv_mass = np.where(pinned == 1, 0.0, 1000.0).astype(np.float32)
        params = _plain_params(1.0 / 60.0, 10, 0.005)
        return _scene(params, xyz, grid_triangles(n, n, False), inv_mass, pinned,
                      grid_distance_pairs_fast(n, n), name=f"plain_{n}x{n}",
                      frames=120, substeps_per_frame=1)
    if k == 1:
        n = size or 256
        rng = np.random.default_rng(1)
        xyz = flat_grid(n, 0.004, rng, y=0.5, centered=True)
        inv_mass = np.full(n * n, 1000.0, dtype=np.float32)
        params = _plain_params(1.0 / 240.0, 15, 0.002)
        params["flags"] |= F_SPHERE
        params["sphere_center"] = (0.0, 0.0, 0.0)
        params["sphere_radius"] = 0.3
        return _scene(params, xyz, grid_triangles(n, n, False), inv_mass, None,
                      grid_distance_pairs_fast(n, n), name=f"sphere_drop_{n}x{n}",
                      frames=600, substeps_per_frame=2)
    if k == 3:
        n = size or 128
        rng = np.random.default_rng(3)
        xyz = flat_grid(n, 0.008, rng)
        pinned = np.zeros(n * n, dtype=np.uint8)
        for c in (0, n - 1, (n - 1) * n, n * n - 1):
            pinned[c] = 1
        inv_mass = np.where(pinned == 1, 0.0, 1000.0).astype(np.float32)
        pairs = grid_distance_pairs_fast(n, n)
        k_true = rng.uniform(0.3, 1.0, size=len(pairs)).astype(np.float32)
        v0 = rng.normal(0.0, 0.05, size=(n * n, 3)).astype(np.float32)
        params = _plain_params(1.0 / 60.0, 10, 0.008 / 3.0)
        return _scene(params, xyz, grid_triangles(n, n, False), inv_mass, pinned, pairs,
                      k_dist=k_true, name=f"fit_{n}x{n}", v0=v0, substeps=60)
    if k == 4:
        n = size or 1024
        rng = np.random.default_rng(4)
        xyz = flat_grid(n, 0.001, rng)
        amp = rng.uniform(0.0, 0.02, 8)
        lam = rng.uniform(0.05, 0.3, 8)
        theta = rng.uniform(0.0, 2 * math.pi, 8)
        phase = rng.uniform(0.0, 2 * math.pi, 8)
        y = np.zeros(n * n)
        for a, l, th, ph in zip(amp, lam, theta, phase):
            y += a * np.sin(2 * math.pi / l * (xyz[:, 0] * math.cos(th) + xyz[:, 2] * math.sin(th)) + ph)
        xyz[:, 1] += y.astype(np.float32)
        inv_mass = np.full(n * n, 1000.0, dtype=np.float32)
        params = _plain_params(1.0 / 60.0, 20, 0.0005)
        return _scene(params, xyz, grid_triangles(n, n, False), inv_mass, None,
                      grid_distance_pairs_fast(n, n), name=f"crumpled_{n}x{n}", frames=200,
                      substeps_per_frame=1)
    raise ValueError(f"config {k} is a batch; use batched_scene({k})")


def batched_scene(k: int = 2, count: int | None = None) -> list[dict]:
    """Batches by number: 2 = 512 cloths with sides drawn from 16..128 per axis, per-cloth
    stiffness, top edge or two corners pinned; 4 = 64 cloths of 64x64 with per-cloth
    bend stiffness (the batched-gradient workload)."""
    scenes = []
    if k == 2:
        count = count or 512
        rng = np.random.default_rng(2)
        sides = rng.integers(16, 129, size=(count, 2))
        for c in range(count):
            nx, ny = int(sides[c, 0]), int(sides[c, 1])
            kd = float(rng.uniform(0.5, 1.0))
            kb = float(rng.uniform(0.01, 0.2))
            edge = bool(rng.integers(0, 2))
            xyz = flat_grid(nx, 0.01, rng, m=ny)
            pinned = np.zeros(nx * ny, dtype=np.uint8)
            if edge:
                pinned[0:ny] = 1          # top edge (row i = 0)
            else:
                pinned[0] = 1
                pinned[ny - 1] = 1
            inv_mass = np.where(pinned == 1, 0.0, 1000.0).astype(np.float32)
            params = _plain_params(1.0 / 60.0, 10, 0.003)
            params["k_stretch"] = kd
            params["k_bend"] = kb
            # candidate / contact budgets per triangle: over the 300 frames the largest
            # counts of any cloth were 30.9 VF and 74.3 EE candidates and 13.9 EE
            # contacts per triangle (each budget also has a fixed slack, 1,024
            # candidates and 65,536 contacts per cloth)
            params["reserved"] = (40, 96, 16)
            scenes.append(_scene(params, xyz, grid_triangles(nx, ny, False), inv_mass, pinned,
                                 grid_distance_pairs_fast(nx, ny), name=f"batch{c}_{nx}x{ny}"))
        return scenes
    if k == 4:
        count = count or 64
        rng = np.random.default_rng(5)
        for c in range(count):
            n = 64
            kb = float(rng.uniform(0.01, 0.2))
            xyz = flat_grid(n, 0.01, rng)
            pinned = np.zeros(n * n, dtype=np.uint8)
            pinned[0] = 1
            pinned[(n - 1) * n] = 1
            inv_mass = np.where(pinned == 1, 0.0, 1000.0).astype(np.float32)
            params = _plain_params(1.0 / 60.0, 10, 0.003)
            params["k_bend"] = kb
            scenes.append(_scene(params, xyz, grid_triangles(n, n, False), inv_mass, pinned,
                                 grid_distance_pairs_fast(n, n), name=f"gradbatch{c}"))
        return scenes
    raise ValueError(k)
