"""GPU parity of the dead-path entry points against the oracle, bit for bit on the
same FP32 inputs:

* cm_ms_step: a jittered 24 x 24 grid with its spring network (cm_ms_grid_springs),
  two pinned corners and an initial sag, 40 mass-spring steps with the reference's
  5 adjustment passes; positions, previous positions and velocities after every
  step;
* cm_impulses: the contacts of a folded cloth's substep with its positions and
  velocities, with and without fixed vertices, with no contacts, and with a contact
  whose vertex id is out of range.
"""
import ctypes as C

import numpy as np
import pytest

from licenta_amd import configs
from licenta_amd._abi import CONTACT_DTYPE

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def lib():
    from licenta_amd import _abi
    return _abi.lib()


def check(rc):
    from licenta_amd._abi import check as chk
    chk(rc)


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def vp(t):
    return C.c_void_p(t.data_ptr())


def stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def f4(a3):
    out = np.zeros((len(a3), 4), np.float32)
    out[:, :3] = a3
    return out


def grid_springs(rows, cols, xyz, kst, ksh, kb):
    L = lib()
    ns = C.c_int32()
    p = xyz.ctypes.data_as(C.c_void_p)
    check(L.cm_ms_grid_springs(rows, cols, p, kst, ksh, kb, None, None, None, None, None, None,
                               C.byref(ns)))
    S = ns.value
    out = dict(a=np.zeros(S, np.int32), b=np.zeros(S, np.int32), ks=np.zeros(S, np.float32),
               l0=np.zeros(S, np.float32), info=np.zeros(S, np.int32),
               ids=np.zeros(rows * cols + 1, np.int32))
    check(L.cm_ms_grid_springs(rows, cols, p, kst, ksh, kb,
                               *[out[k].ctypes.data_as(C.c_void_p)
                                 for k in ("a", "b", "ks", "l0", "info", "ids")], C.byref(ns)))
    return out


def test_mass_spring_steps_match_oracle(oracle_mod):
    rows = cols = 24
    rng = np.random.default_rng(11)
    i, j = np.meshgrid(np.arange(rows), np.arange(cols), indexing="ij")
    xyz = np.stack([j.ravel() * 0.02, np.zeros(rows * cols), i.ravel() * 0.02], 1)
    xyz = (xyz + rng.uniform(-2e-4, 2e-4, xyz.shape)).astype(np.float32)
    kst, ksh, kb = 80.0, 40.0, 10.0
    S = grid_springs(rows, cols, xyz, kst, ksh, kb)
    n = rows * cols
    fixed = np.zeros(n, np.uint8)
    fixed[0] = fixed[cols - 1] = 1
    mass = np.full(n, 0.5 / n, np.float32)
    pos = xyz.copy()
    pos[:, 1] -= 0.01 * (i.ravel() / rows).astype(np.float32)  # a sag in y
    pos[fixed == 1] = xyz[fixed == 1]
    prev = xyz.copy()
    vel = np.zeros((n, 3), np.float32)
    g = np.array([0.0, -9.81, 0.0], np.float32)
    kd, dt = 0.02, 1.0 / 240.0
    d = {k: dev(f4(a)) for k, a in (("pos", pos), ("prev", prev), ("vel", vel))}
    dm, dfx = dev(mass), dev(fixed)
    ds = {k: dev(v) for k, v in S.items()}
    tmp = torch.empty((n, 4), dtype=torch.float32, device="cuda")
    L = lib()
    print(f"mass-spring {rows}x{cols}, {len(S['a'])} directed springs", flush=True)
    for step in range(40):
        check(L.cm_ms_step(n, vp(d["pos"]), vp(d["prev"]), vp(d["vel"]), vp(dm), vp(dfx),
                           vp(ds["a"]), vp(ds["b"]), vp(ds["ks"]), vp(ds["l0"]), vp(ds["info"]),
                           vp(ds["ids"]), g.ctypes.data_as(C.c_void_p), kd, dt, kb, 5, 0.1,
                           vp(tmp), stream()))
        pos, prev, vel = oracle_mod.ms_step(pos, prev, vel, mass, fixed, S, g, kd, dt, kb, 5, 0.1)
        torch.cuda.synchronize()
        for k, ref in (("pos", pos), ("prev", prev), ("vel", vel)):
            got = d[k].cpu().numpy()[:, :3]
            assert np.array_equal(got, ref), (step, k, np.abs(got - ref).max())
    assert np.isfinite(pos).all() and np.abs(pos - xyz).max() > 1e-3


def _folded_contacts():
    from licenta_amd.body import ClothBody
    n = 24
    sc = configs.config_scene(0, n)
    xyz = sc["xyz"].reshape(n, n, 3).copy()
    h = n // 2
    for i in range(h, n):
        src = 2 * h - 1 - i
        xyz[i, :, 0] = xyz[src, :, 0] + 0.002
        xyz[i, :, 1] = 0.003 + 0.0005 * (i - h)
    sc["xyz"] = xyz.reshape(-1, 3).astype(np.float32)
    sc["pinned"] = None
    sc["inv_mass"] = np.full(n * n, 1000.0, np.float32)
    sc = dict(sc, params=dict(sc["params"], thickness=0.004))
    v = np.zeros((n * n, 3), np.float32)
    v.reshape(n, n, 3)[h:, :, 1] = -0.3
    body = ClothBody.from_scene(sc)
    body.set_state(sc["xyz"], v)
    body.fixed_update(1)
    vf, ee = body.contacts()
    x = body.buffer("x")[:, :3].copy()
    vv = body.buffer("v")[:, :3].copy()
    p = body.buffer("p")[:, :3].copy()
    body.close()
    return x, vv, p, vf, ee


def run_impulses(x, v, p, fixed, vf, ee, stiffness, m, dt, th):
    L = lib()
    n = len(x)
    nc = len(vf) + len(ee)
    nbytes = L.cm_impulses_temp_bytes(n, nc)
    tmp = torch.empty(max(nbytes, 1), dtype=torch.uint8, device="cuda")
    dx, dv, dp = dev(f4(x)), dev(f4(v)), dev(f4(p))
    dfx = dev(np.asarray(fixed, np.uint8))
    dvf = dev(np.frombuffer(vf.tobytes(), np.uint8)) if len(vf) else None
    dee = dev(np.frombuffer(ee.tobytes(), np.uint8)) if len(ee) else None
    check(L.cm_impulses(n, vp(dx), vp(dv), vp(dp), vp(dfx), vp(dvf) if dvf is not None else None,
                        len(vf), vp(dee) if dee is not None else None, len(ee), stiffness, m, dt,
                        th, vp(tmp), nbytes, stream()))
    torch.cuda.synchronize()
    return dv.cpu().numpy()[:, :3], dp.cpu().numpy()[:, :3]


def test_impulses_match_oracle(oracle_mod):
    x, v, p, vf, ee = _folded_contacts()
    n = len(x)
    print(f"impulses: {len(vf)} VF and {len(ee)} EE contacts on {n} vertices", flush=True)
    assert len(vf) > 0 and len(ee) > 0
    none = np.zeros(0, CONTACT_DTYPE)
    fixed_some = np.zeros(n, np.uint8)
    fixed_some[::7] = 1
    bad = ee.copy()
    bad["v"][len(bad) // 2, 1] = n + 5
    cases = [("all", np.zeros(n, np.uint8), vf, ee), ("fixed", fixed_some, vf, ee),
             ("vf only", np.zeros(n, np.uint8), vf, none), ("none", fixed_some, none, none),
             ("invalid id", np.zeros(n, np.uint8), vf, bad)]
    for name, fixed, cvf, cee in cases:
        for stiffness, m in ((1.0, 1e-3), (200.0, 0.05)):
            got_v, got_p = run_impulses(x, v, p, fixed, cvf, cee, stiffness, m, 1.0 / 60.0, 0.004)
            ref_v, ref_p = oracle_mod.impulses(x, v, p, fixed, cvf, cee, stiffness, m, 1.0 / 60.0,
                                               0.004)
            assert np.array_equal(got_v, ref_v), (name, stiffness, np.abs(got_v - ref_v).max())
            assert np.array_equal(got_p, ref_p), (name, stiffness, np.abs(got_p - ref_p).max())
            if name == "all":
                assert not np.array_equal(got_v, v)
