"""GPU parity on the named workloads at their own sizes and budgets:

* the 512-cloth batch (sides 16..128 per axis, seed 2) with its candidate budgets
  (40 VF / 96 EE candidates and 16 contacts per triangle): every cloth equal
  bit for bit to its own single body, sampled cloths against the FP32 oracle, and no
  candidate buffer overflow on either side;
* the batched-gradient workload (64 cloths of 64 x 64, per-cloth bend stiffness,
  30 substeps, checkpoints every 10): every cloth's gradients equal its single
  body's, and the first cloth's gradients over one checkpoint segment within 1e-4
  of the FP64 oracle adjoint evaluated on the FP32 forward's iterates (the FP64
  oracle's own trajectory, on the FP32 run's contact decisions, is compared and
  printed too);
* the 128 x 128 fitting workload at the fit's start (stiffness 0.65, zero velocity,
  loss against the ground-truth run): over 10 and over 60 substeps the gradient
  within 1e-4 of the FP64 oracle adjoint on the FP32 forward's iterates; over 10
  substeps, directional derivatives against central finite differences of the GPU
  forward;
* the reference's FP32 bend arithmetic (CM_F_BEND_FP64 cleared) against the FP32
  oracle.
"""
import numpy as np
import pytest

from licenta_amd import configs

pytestmark = pytest.mark.gpu

WARN_CANDIDATE_CAP = 1 << 4


def fp64_adjoints(oracle_mod, scene, x0, v0, tri0, substeps, target, k_dist, k_bend):
    """FP64 oracle adjoints of the same rollout, which the GPU's FP32 forward
    reproduces bit for bit in the FP32 oracle:
    * "tape": the FP64 adjoint evaluated on the FP32 oracle's recorded iterates,
      contact sets and activation decisions (the FP32 forward's own points);
    * "own": the FP64 oracle's own forward and adjoint, taking the FP32 run's
      contact activation decisions (cc < 0) in every iteration; its trajectory
      drifts from the FP32 one by the forward's rounding.
    Returns (tape gradients, own gradients, FP64 final positions, decisions
    replayed, substeps whose FP64 contact set differed from the FP32 one); gradients
    are (dL/dx0, dL/dv0, dL/dk_dist, dL/dk_bend) of the mean squared error against
    target at each forward's own final positions."""
    o32 = oracle_mod.OracleSim(scene, 32)
    o64 = oracle_mod.OracleSim(scene, 64)
    for o in (o32, o64):
        o.set_state(np.asarray(x0, np.float64), np.asarray(v0, np.float64), None, tri0)
        o.record(True)
    o32.activation_mode(1)
    replayed = mismatched = 0
    for _ in range(substeps):
        o32.step(1)
        act = o32.activations()
        o64.activation_mode(2)
        o64.set_activations(act)
        o64.step(1)
        mismatched += int(o64.activation_mismatch())
        replayed += len(act)
    x64, _, _ = o64.state()
    own = o64.adjoint(2.0 * (x64 - target) / len(x64), k_dist=k_dist, k_bend=k_bend)
    x32, _, _ = o32.state()
    o64.copy_tape_from(o32)
    tape = o64.adjoint(2.0 * (x32 - target) / len(x32), k_dist=k_dist, k_bend=k_bend)
    return tape, own, x64, replayed, mismatched


def scaled_err(g, ref):
    ref = np.asarray(ref, np.float64)
    return float(np.abs(np.asarray(g, np.float64) - ref).max() / max(np.abs(ref).max(), 1e-30))


@pytest.mark.timeout(600)
def test_batch_512_own_budgets(oracle_mod):
    from licenta_amd.body import ClothBody
    scenes = configs.batched_scene(2)
    assert len(scenes) == 512 and tuple(scenes[0]["params"]["reserved"][:3]) == (40, 96, 16)
    steps = 3
    print("[batch512] batched body", flush=True)
    batch = ClothBody.batch(scenes)
    try:
        batch.fixed_update(steps)
        x = batch.positions()
        bp = batch.pairs()
        bvf, bee = batch.contacts()
        bw = batch.warnings()
        npairs = batch.counts()["pairs"]
    finally:
        batch.close()
    # the FP32 oracle on the smallest, the largest and two drawn cloths
    sizes = [len(s["xyz"]) for s in scenes]
    sample = sorted({int(np.argmin(sizes)), int(np.argmax(sizes)), 17, 300})
    print("[batch512] single bodies, one at a time", flush=True)
    vo = to = 0
    sw = 0
    for c, s in enumerate(scenes):
        nv, nt = len(s["xyz"]), len(s["tri"])
        body = ClothBody.from_scene(s)
        try:
            body.fixed_update(steps)
            xs = body.positions()
            assert np.array_equal(x[vo:vo + nv], xs), s["name"]
            sel = (bp[:, 0] >= to) & (bp[:, 0] < to + nt)
            assert np.array_equal(bp[sel] - to, body.pairs()), s["name"]
            for bc, sc in zip((bvf, bee), body.contacts()):
                mine = bc[(bc["v"][:, 0] >= vo) & (bc["v"][:, 0] < vo + nv)]
                assert np.array_equal(mine["v"] - vo, sc["v"]), s["name"]
            sw |= body.warnings()
            if c in sample:
                o = oracle_mod.OracleSim(s, 32)
                o.step(steps)
                x32, _, _ = o.state()
                err = float(np.abs(xs - x32).max() / np.abs(x32).max())
                print(f"  cloth {c} ({nv} vertices): rel err vs FP32 oracle {err:.2e}", flush=True)
                assert err <= 2e-6
                assert np.array_equal(body.pairs(), o.pairs())
                gvf, gee = body.contacts()
                assert np.array_equal(gvf["v"], o.contacts(False)["v"])
                assert np.array_equal(gee["v"], o.contacts(True)["v"])
        finally:
            body.close()
        vo += nv
        to += nt
    print(f"  pairs {npairs}, warnings batch {bw:#x} singles {sw:#x}", flush=True)
    assert npairs > 1_000_000
    assert not bw & WARN_CANDIDATE_CAP and not sw & WARN_CANDIDATE_CAP


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
        try:
            b.set_stiffness(np.full(b.counts()["s"], 0.9, np.float32))
            if c == 0:
                tri0 = b.buffer("tri").copy()
            b.set_state(s["xyz"], v0)
            b.forward_ckpt(substeps, every)
            lx = 2.0 * (b.positions() - tg) / len(tg)
            singles.append((b.counts(), b.adjoint(lx)))
        finally:
            b.close()
    print("[adj64] batch", flush=True)
    batch = ClothBody.batch(scenes)
    try:
        batch.set_stiffness(np.full(batch.counts()["s"], 0.9, np.float32))
        batch.set_state(np.concatenate([s["xyz"] for s in scenes]), np.concatenate(v0s))
        batch.forward_ckpt(substeps, every)
        x = batch.positions()
        n0 = np.cumsum([0] + [len(s["xyz"]) for s in scenes])
        lx = np.concatenate([2.0 * (x[n0[c]:n0[c + 1]] - targets[c]) / len(targets[c])
                             for c in range(len(scenes))])
        gx, gv, gkd, gkb = batch.adjoint(lx)
    finally:
        batch.close()
    s0 = np.cumsum([0] + [c["s"] for c, _ in singles])
    b0 = np.cumsum([0] + [c["b"] for c, _ in singles])
    for c, (_, (sx, sv, skd, skb)) in enumerate(singles):
        for g, ref in ((gx[n0[c]:n0[c + 1]], sx), (gv[n0[c]:n0[c + 1]], sv),
                       (gkd[s0[c]:s0[c + 1]], skd), (gkb[b0[c]:b0[c + 1]], skb)):
            assert scaled_err(g, ref) <= 1e-5, c
    # the first cloth against the FP64 oracle adjoint over one checkpoint segment
    # (10 substeps), where the FP32 and FP64 forward runs agree to 1e-5
    print("[adj64] cloth 0, 10 substeps, against the FP64 oracle adjoint", flush=True)
    s = scenes[0]
    ns, nb = singles[0][0]["s"], singles[0][0]["b"]
    kd = np.full(ns, 0.9, np.float32)
    b = ClothBody.from_scene(s)
    try:
        b.set_stiffness(kd)
        b.set_state(s["xyz"], v0s[0], tri=tri0)
        b.forward_ckpt(10, every)
        xg = b.positions().astype(np.float64)
        g = b.adjoint(2.0 * (xg - targets[0]) / len(xg))
    finally:
        b.close()
    tape, own, x64, nact, mism = fp64_adjoints(
        oracle_mod, dict(s, k_dist=kd), s["xyz"], v0s[0], tri0, 10, targets[0], kd,
        np.full(nb, s["params"]["k_bend"], np.float32))
    assert mism == 0  # the same contact sets in FP32 and FP64
    assert np.abs(xg - x64).max() <= 1e-5 * np.abs(x64).max()
    e_tape = [scaled_err(a, r) for a, r in zip(g, tape)]
    e_own = [scaled_err(a, r) for a, r in zip(g, own)]
    print(f"  cloth 0 (x, v, k_dist, k_bend; {nact} contact decisions): vs FP64 adjoint on the "
          f"FP32 iterates {['%.2e' % e for e in e_tape]}, vs the FP64 oracle's own "
          f"trajectory {['%.2e' % e for e in e_own]}", flush=True)
    assert max(e_tape) <= 1e-4


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


def _sine_bump(n_vertices):
    """v = (0, sin(pi i / (m - 1)) sin(pi j / (m - 1)), 0) over the m x m grid."""
    m = int(round(np.sqrt(n_vertices)))
    t = np.sin(np.pi * np.arange(m) / (m - 1))
    v = np.zeros((m * m, 3), np.float32)
    v[:, 1] = np.outer(t, t).ravel()
    return v


@pytest.mark.timeout(900)
def test_fit_128_gradients(oracle_mod):
    """The adjoint of the 128 x 128 fit at its start, over one checkpoint segment (10
    substeps, the same contact sets in FP32 and FP64) and over the workload's 60
    substeps (checkpoints every 10): within 1e-4 of the largest component of the FP64
    oracle adjoint evaluated on the FP32 forward's iterates. The FP64 oracle's own
    trajectory is printed beside: it drifts from the FP32 one by 1e-4 within 10
    substeps, because the cloth starts flat in the plane y = 0 and the bend in
    absolute coordinates (the reference's form) then measures the angle between two
    nearly coincident planes through the origin, which rounding moves.

    Over the 10 substeps also against central differences of the GPU forward along
    four directions: all distance stiffnesses together, random non-negative weights
    on them, a uniform initial velocity and a smooth initial velocity field (one
    sine bump over the grid). One coordinate at a time would not do: the loss moves
    by about 1e-11 per unit of a single stiffness, below the FP32 forward's rounding
    of the loss (about 1e-10); a white-noise velocity field would not either, for
    the bend's conditioning above. Within 1e-2 of each directional derivative; the
    white-noise field, and over the 60 substeps (where the perturbed runs also switch
    contacts on and off) every direction, are printed."""
    from licenta_amd.body import ClothBody
    sc, body, tri0, target, k0, c = _fit_problem()
    n = c["n"]
    zero = np.zeros_like(sc["xyz"])
    kb = np.full(c["b"], sc["params"]["k_bend"], np.float32)
    rng = np.random.default_rng(9)
    dirs = [("k_dist, all", np.ones(c["s"], np.float32), None),
            ("k_dist, random weights", rng.uniform(0.0, 1.0, c["s"]).astype(np.float32), None),
            ("v0, uniform +y", None, np.tile(np.float32([0.0, 1.0, 0.0]), (n, 1))),
            ("v0, sine bump", None, _sine_bump(n)),
            ("v0, white noise (printed)", None, rng.normal(0.0, 1.0, (n, 3)).astype(np.float32))]
    msgs, fails = [], []
    try:
        for horizon in (10, 60):
            body.set_stiffness(k0)
            body.set_state(sc["xyz"], zero, tri=tri0)
            body.forward_ckpt(horizon, 10)
            x = body.positions().astype(np.float64)
            gx, gv, gkd, gkb = body.adjoint(2.0 * (x - target) / n)
            print(f"[fit128] {horizon} substeps: FP64 oracle forward + adjoint", flush=True)
            tape, own, x64, _, mism = fp64_adjoints(
                oracle_mod, dict(sc, k_dist=k0), sc["xyz"], zero, tri0, horizon, target, k0, kb)
            drift = float(np.abs(x - x64).max() / np.abs(x64).max())
            names = ("x", "v", "kd", "kb")
            errs = {k: scaled_err(a, r) for k, a, r in zip(names, (gx, gv, gkd, gkb), tape)}
            eown = {k: scaled_err(a, r) for k, a, r in zip(names, (gx, gv, gkd, gkb), own)}
            msgs.append(f"{horizon} substeps: forward drift {drift:.2e}, adjoint vs FP64 on the "
                        f"FP32 iterates { {k: '%.2e' % e for k, e in errs.items()} }, vs the "
                        f"FP64 oracle's own trajectory { {k: '%.2e' % e for k, e in eown.items()} }"
                        f", substeps with another FP64 contact set {mism}")
            print("  " + msgs[-1], flush=True)
            if max(errs.values()) > 1e-4 or (horizon == 10 and mism):
                fails.append(msgs[-1])

            def loss(kd, v0):
                body.set_stiffness(kd)
                body.set_state(sc["xyz"], v0, tri=tri0)
                body.fixed_update(horizon)
                xx = body.positions().astype(np.float64)
                return float(((xx - target) ** 2).sum(axis=1).mean())

            for name, uk, uv in dirs:
                h = 1e-2 if uk is not None else 1e-3
                if uk is not None:
                    f = [loss(np.clip(k0 + sgn * h * uk, 0.0, 1.0), zero) for sgn in (1, -1)]
                    ad = float(np.dot(gkd.astype(np.float64), uk))
                else:
                    f = [loss(k0, zero + sgn * h * uv) for sgn in (1, -1)]
                    ad = float(np.dot(gv.astype(np.float64).ravel(), uv.ravel()))
                fd = (f[0] - f[1]) / (2 * h)
                e = abs(fd - ad) / max(abs(fd), 1e-30)
                msgs.append(f"{horizon} substeps, direction {name}: adjoint {ad:.6e}, central "
                            f"difference {fd:.6e} (h = {h}), relative {e:.2e}")
                print("  " + msgs[-1], flush=True)
                if horizon == 10 and e > 1e-2 and "printed" not in name:
                    fails.append(msgs[-1])
    finally:
        body.close()
    assert not fails, fails


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
