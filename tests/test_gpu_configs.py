"""GPU parity on the named workloads at their own sizes and budgets:

* the 512-cloth batch (sides 16..128 per axis, seed 2) with its reduced candidate
  budgets (16 VF / 24 EE candidates and 4 contacts per triangle): every cloth equal
  bit for bit to its own single body, sampled cloths against the FP32 oracle, and no
  candidate buffer overflow on either side;
* the batched-gradient workload (64 cloths of 64 x 64, per-cloth bend stiffness,
  30 substeps, checkpoints every 10): every cloth's gradients equal its single
  body's, and the first cloth's the FP64 oracle adjoint's within 1e-3;
* the 128 x 128 fitting workload at the fit's start (stiffness 0.65, zero velocity,
  60 substeps, checkpoints every 10, loss against the ground-truth run): the
  gradient against the FP64 oracle adjoint within 1e-3, and against central finite
  differences (h = 1e-3) of the GPU forward on 64 random coordinates;
* the reference's FP32 bend arithmetic (CM_F_BEND_FP64 cleared) against the FP32
  oracle.
"""
import numpy as np
import pytest

from licenta_amd import configs

pytestmark = pytest.mark.gpu

WARN_CANDIDATE_CAP = 1 << 4


def scaled_err(g, ref):
    ref = np.asarray(ref, np.float64)
    return float(np.abs(np.asarray(g, np.float64) - ref).max() / max(np.abs(ref).max(), 1e-30))


@pytest.mark.timeout(600)
def test_batch_512_own_budgets(oracle_mod):
    from licenta_amd.body import ClothBody
    scenes = configs.batched_scene(2)
    assert len(scenes) == 512 and tuple(scenes[0]["params"]["reserved"][:3]) == (16, 24, 4)
    print("[batch512] create", flush=True)
    batch = ClothBody.batch(scenes)
    singles = [ClothBody.from_scene(s) for s in scenes]
    steps = 3
    # the FP32 oracle on the smallest, the largest and two drawn cloths
    sizes = [len(s["xyz"]) for s in scenes]
    sample = sorted({int(np.argmin(sizes)), int(np.argmax(sizes)), 17, 300})
    oracles = {c: oracle_mod.OracleSim(scenes[c], 32) for c in sample}
    for step in range(steps):
        print(f"[batch512] substep {step}", flush=True)
        batch.fixed_update(1)
        for b in singles:
            b.fixed_update(1)
        for o in oracles.values():
            o.step(1)
    x = batch.positions()
    bp = batch.pairs()
    bvf, bee = batch.contacts()
    vo = to = 0
    for c, (s, body) in enumerate(zip(scenes, singles)):
        nv, nt = len(s["xyz"]), len(s["tri"])
        xs = body.positions()
        assert np.array_equal(x[vo:vo + nv], xs), s["name"]
        sel = (bp[:, 0] >= to) & (bp[:, 0] < to + nt)
        assert np.array_equal(bp[sel] - to, body.pairs()), s["name"]
        for bc, sc in zip((bvf, bee), body.contacts()):
            mine = bc[(bc["v"][:, 0] >= vo) & (bc["v"][:, 0] < vo + nv)]
            assert np.array_equal(mine["v"] - vo, sc["v"]), s["name"]
        if c in oracles:
            o = oracles[c]
            x32, _, _ = o.state()
            err = float(np.abs(xs - x32).max() / np.abs(x32).max())
            print(f"  cloth {c} ({nv} vertices): rel err vs FP32 oracle {err:.2e}", flush=True)
            assert err <= 2e-6
            assert np.array_equal(body.pairs(), o.pairs())
            gvf, gee = body.contacts()
            assert np.array_equal(gvf["v"], o.contacts(False)["v"])
            assert np.array_equal(gee["v"], o.contacts(True)["v"])
        vo += nv
        to += nt
    bw = batch.warnings()
    sw = 0
    for b in singles:
        sw |= b.warnings()
    print(f"  pairs {batch.counts()['pairs']}, warnings batch {bw:#x} singles {sw:#x}", flush=True)
    assert batch.counts()["pairs"] > 1_000_000
    assert not bw & WARN_CANDIDATE_CAP and not sw & WARN_CANDIDATE_CAP
    batch.close()
    for b in singles:
        b.close()


@pytest.mark.timeout(600)
def test_batched_adjoint_64_cloths(oracle_mod):
    from licenta_amd.body import ClothBody
    scenes = configs.batched_scene(4)
    assert len(scenes) == 64 and len(scenes[0]["xyz"]) == 64 * 64
    substeps, every = 30, 10
    rng = np.random.default_rng(8)
    v0s = [rng.normal(0.0, 0.05, size=(len(s["xyz"]), 3)).astype(np.float32) for s in scenes]
    targets = [s["xyz"] + rng.normal(0.0, 0.01, size=(len(s["xyz"]), 3)) for s in scenes]
    print("[adj64] single bodies", flush=True)
    singles = []
    for c, (s, v0, tg) in enumerate(zip(scenes, v0s, targets)):
        b = ClothBody.from_scene(s)
        b.set_stiffness(np.full(b.counts()["s"], 0.9, np.float32))
        if c == 0:
            tri0 = b.buffer("tri").copy()
        b.set_state(s["xyz"], v0)
        b.forward_ckpt(substeps, every)
        lx = 2.0 * (b.positions() - tg) / len(tg)
        singles.append((b.counts(), b.adjoint(lx)))
        b.close()
    print("[adj64] batch", flush=True)
    batch = ClothBody.batch(scenes)
    batch.set_stiffness(np.full(batch.counts()["s"], 0.9, np.float32))
    batch.set_state(np.concatenate([s["xyz"] for s in scenes]), np.concatenate(v0s))
    batch.forward_ckpt(substeps, every)
    x = batch.positions()
    n0 = np.cumsum([0] + [len(s["xyz"]) for s in scenes])
    lx = np.concatenate([2.0 * (x[n0[c]:n0[c + 1]] - targets[c]) / len(targets[c])
                         for c in range(len(scenes))])
    gx, gv, gkd, gkb = batch.adjoint(lx)
    s0 = np.cumsum([0] + [c["s"] for c, _ in singles])
    b0 = np.cumsum([0] + [c["b"] for c, _ in singles])
    for c, (_, (sx, sv, skd, skb)) in enumerate(singles):
        for g, ref in ((gx[n0[c]:n0[c + 1]], sx), (gv[n0[c]:n0[c + 1]], sv),
                       (gkd[s0[c]:s0[c + 1]], skd), (gkb[b0[c]:b0[c + 1]], skb)):
            assert scaled_err(g, ref) <= 1e-5, c
    batch.close()
    # the first cloth against the FP64 oracle adjoint
    print("[adj64] FP64 oracle adjoint of cloth 0", flush=True)
    s = scenes[0]
    ns = singles[0][0]["s"]
    kd = np.full(ns, 0.9, np.float32)
    o = oracle_mod.OracleSim(dict(s, k_dist=kd), 64)
    o.set_state(s["xyz"].astype(np.float64), v0s[0].astype(np.float64), None, tri0)
    o.record(True)
    o.step(substeps)
    x64, _, _ = o.state()
    lx64 = 2.0 * (x64 - targets[0]) / len(x64)
    ref = o.adjoint(lx64, k_dist=kd, k_bend=np.full(singles[0][0]["b"],
                                                     s["params"]["k_bend"], np.float32))
    errs = [scaled_err(g, r) for g, r in zip(singles[0][1], ref)]
    print(f"  cloth 0 vs FP64 adjoint: {['%.2e' % e for e in errs]}", flush=True)
    assert max(errs) <= 1e-3


def _fit_problem():
    from licenta_amd.body import ClothBody
    sc = configs.config_scene(3)
    body = ClothBody.from_scene(sc)  # ground-truth stiffness and velocities
    tri0 = body.buffer("tri").copy()
    body.set_state(sc["xyz"], sc["v0"], tri=tri0)
    body.fixed_update(60)
    target = body.positions().astype(np.float64)
    c = body.counts()
    k0 = np.full(c["s"], 0.65, np.float32)
    body.set_stiffness(k0)
    return sc, body, tri0, target, k0, c


@pytest.mark.timeout(900)
def test_fit_128_gradients(oracle_mod):
    sc, body, tri0, target, k0, c = _fit_problem()
    n = c["n"]
    zero = np.zeros_like(sc["xyz"])
    body.set_state(sc["xyz"], zero, tri=tri0)
    body.forward_ckpt(60, 10)
    x = body.positions().astype(np.float64)
    lx = 2.0 * (x - target) / n
    gx, gv, gkd, gkb = body.adjoint(lx)
    # FP64 oracle adjoint from the same start
    print("[fit128] FP64 oracle forward + adjoint", flush=True)
    o = oracle_mod.OracleSim(dict(sc, k_dist=k0), 64)
    o.set_state(sc["xyz"].astype(np.float64), zero.astype(np.float64), None, tri0)
    o.record(True)
    o.step(60)
    x64, _, _ = o.state()
    assert np.abs(x - x64).max() <= 1e-5 * np.abs(x64).max()
    rx, rv, rkd, rkb = o.adjoint(2.0 * (x64 - target) / n, k_dist=k0,
                                 k_bend=np.full(c["b"], sc["params"]["k_bend"], np.float32))
    errs = dict(kd=scaled_err(gkd, rkd), kb=scaled_err(gkb, rkb), x=scaled_err(gx, rx),
                v=scaled_err(gv, rv))
    print(f"  vs FP64 adjoint: { {k: '%.2e' % e for k, e in errs.items()} }", flush=True)
    assert max(errs.values()) <= 1e-3
    # central differences of the GPU forward on 64 random coordinates: 48 distance
    # stiffnesses and 16 initial-velocity components
    rng = np.random.default_rng(9)
    ks = rng.choice(c["s"], 48, replace=False)
    vs = rng.choice(3 * n, 16, replace=False)

    def loss(kd, v0):
        body.set_stiffness(kd)
        body.set_state(sc["xyz"], v0, tri=tri0)
        body.fixed_update(60)
        xx = body.positions().astype(np.float64)
        return float(((xx - target) ** 2).sum(axis=1).mean())

    h = 1e-3
    fd, ad = [], []
    for j in ks:
        hi, lo = k0.copy(), k0.copy()
        hi[j] += h
        lo[j] -= h
        fd.append((loss(hi, zero) - loss(lo, zero)) / (2 * h))
        ad.append(gkd[j])
    for j in vs:
        hi, lo = zero.copy(), zero.copy()
        hi.reshape(-1)[j] += h
        lo.reshape(-1)[j] -= h
        fd.append((loss(k0, hi) - loss(k0, lo)) / (2 * h))
        ad.append(gv.reshape(-1)[j])
    fd, ad = np.array(fd), np.array(ad)
    # per group: the largest difference against the group's largest finite difference
    e_k = float(np.abs(fd[:48] - ad[:48]).max() / np.abs(fd[:48]).max())
    e_v = float(np.abs(fd[48:] - ad[48:]).max() / np.abs(fd[48:]).max())
    print(f"  vs central differences: stiffness {e_k:.2e}, velocity {e_v:.2e}", flush=True)
    assert e_k <= 2e-2 and e_v <= 2e-2
    body.close()


def test_fp32_bend_arithmetic(oracle_mod):
    """The reference's FP32 dihedral bend (CM_F_BEND_FP64 cleared) on a plain cloth
    and a folded one with contacts: integer stages bit-exact, positions within 2e-6
    of the FP32 oracle (acosf and sqrtf of the device and host libraries may differ
    in the last ulp)."""
    from licenta_amd.body import ClothBody
    for name, sc in (("plain", configs.config_scene(0, 24)), ("folded", _folded(16))):
        sc["params"] = dict(sc["params"], flags=sc["params"]["flags"] & ~configs.F_BEND_FP64)
        body = ClothBody.from_scene(sc)
        if name == "folded":
            v = np.zeros((len(sc["xyz"]), 3), np.float32)
            v.reshape(16, 16, 3)[8:, :, 1] = -0.3
            body.set_state(sc["xyz"], v)
        total = 0
        for s in range(4):
            x = body.buffer("x")[:, :3].astype(np.float64)
            p = body.buffer("p")[:, :3].astype(np.float64)
            vv = body.buffer("v")[:, :3].astype(np.float64)
            o = oracle_mod.OracleSim(sc, 32)
            o.set_state(x, vv, p, body.buffer("tri"))
            o.set_parents(body.buffer("parents"))
            body.fixed_update(1)
            o.step(1)
            assert np.array_equal(body.pairs(), o.pairs())
            gvf, gee = body.contacts()
            assert np.array_equal(gvf["v"], o.contacts(False)["v"])
            assert np.array_equal(gee["v"], o.contacts(True)["v"])
            total += len(gvf) + len(gee)
            x32, _, _ = o.state()
            err = float(np.abs(body.positions() - x32).max() / np.abs(x32).max())
            print(f"[fp32 bend {name}] substep {s}: rel err {err:.2e}", flush=True)
            assert err <= 2e-6
        if name == "folded":
            assert total > 0
        body.close()


def _folded(n):
    sc = configs.config_scene(0, n)
    xyz = sc["xyz"].reshape(n, n, 3).copy()
    h = n // 2
    for i in range(h, n):
        xyz[i, :, 0] = xyz[2 * h - 1 - i, :, 0] + 0.002
        xyz[i, :, 1] = 0.003 + 0.0005 * (i - h)
    sc["xyz"] = xyz.reshape(-1, 3).astype(np.float32)
    sc["params"] = dict(sc["params"], thickness=0.004)
    sc["pinned"] = None
    sc["inv_mass"] = np.full(n * n, 1000.0, np.float32)
    return sc
