"""GPU adjoint of the differentiable substep (cm_forward_ckpt / cm_adjoint through the
C ABI) against the oracle's FP64 adjoint on the same starting state, stiffness and
loss, on small pinned cloths of the fitting workload (128 x 128 in full; 8 x 8 and
16 x 16 here so the FP64 oracle runs in seconds).

* Gradients w.r.t. every distance and bend stiffness and every initial position and
  velocity component agree with the FP64 oracle adjoint to 1e-3 of the largest
  component: max |g_gpu - g_64| <= 1e-3 max |g_64| per gradient.
* With self-collision the GPU and oracle contact counts are checked to be the same
  in every substep first, so both differentiate the same frozen-set function.
* The result does not depend on the checkpoint interval (bit-exact: every segment
  is recomputed by the same kernels), and the body is left in its forward state.
* The torch.autograd.Function gives the same gradients, and a central finite
  difference (h = 1e-3) of the GPU forward agrees for the largest components.
"""
import numpy as np
import pytest

from licenta_amd import configs

pytestmark = pytest.mark.gpu

TOL = 1e-3


def scaled_err(g, ref):
    ref = np.asarray(ref, np.float64)
    return float(np.abs(np.asarray(g, np.float64) - ref).max() / max(np.abs(ref).max(), 1e-30))


def _problem(size, selfcoll, seed=11):
    sc = configs.config_scene(3, size)
    if not selfcoll:
        sc["params"] = dict(sc["params"], flags=sc["params"]["flags"] & ~configs.F_SELF_COLLISION)
    from licenta_amd.body import ClothBody
    body = ClothBody.from_scene(sc)
    c = body.counts()
    rng = np.random.default_rng(seed)
    k_dist = rng.uniform(0.3, 0.95, size=c["s"]).astype(np.float32)
    k_bend = rng.uniform(0.05, 0.3, size=c["b"]).astype(np.float32)
    v0 = rng.normal(0.0, 0.05, size=(c["n"], 3)).astype(np.float32)
    target = sc["xyz"].astype(np.float64) + rng.normal(0.0, 0.01, size=(c["n"], 3))
    body.set_stiffness(k_dist, k_bend)
    body.set_state(sc["xyz"], v0)
    return sc, body, k_dist, k_bend, v0, target


def _loss_grad(x, target):
    x = np.asarray(x, np.float64)
    return float(((x - target) ** 2).sum(axis=1).mean()), 2.0 * (x - target) / len(x)


def _oracle_grads(oracle_mod, sc, body, k_dist, k_bend, v0, target, substeps):
    s = dict(sc, k_dist=k_dist, k_bend=k_bend)
    o = oracle_mod.OracleSim(s, 64)
    o.set_state(sc["xyz"].astype(np.float64), v0.astype(np.float64), None, body.buffer("tri"))
    o.record(True)
    counts = []
    for _ in range(substeps):
        o.step(1)
        counts.append((o.counts()["vf"], o.counts()["ee"]))
    x, _, _ = o.state()
    _, lx = _loss_grad(x, target)
    gx0, gv0, gkd, gkb = o.adjoint(lx, k_dist=k_dist, k_bend=k_bend)
    return dict(x=gx0, v=gv0, kd=gkd, kb=gkb), counts, x


def _gpu_counts(body, sc, v0, substeps):
    tri0 = body.buffer("tri").copy()
    counts = []
    for _ in range(substeps):
        body.fixed_update(1)
        c = body.counts()
        counts.append((c["vf"], c["ee"]))
    body.set_state(sc["xyz"], v0, tri=tri0)
    return counts


def _gpu_grads(body, target, substeps, every):
    body.forward_ckpt(substeps, every)
    x = body.positions()
    _, lx = _loss_grad(x, target)
    gx0, gv0, gkd, gkb = body.adjoint(lx)
    return dict(x=gx0, v=gv0, kd=gkd, kb=gkb), x


@pytest.mark.parametrize("size,substeps,every", [(8, 4, 3), (16, 6, 4)])
def test_adjoint_matches_fp64_oracle(oracle_mod, size, substeps, every):
    print(f"adjoint smooth size={size} substeps={substeps} every={every}", flush=True)
    sc, body, kd, kb, v0, target = _problem(size, selfcoll=False)
    ref, _, x64 = _oracle_grads(oracle_mod, sc, body, kd, kb, v0, target, substeps)
    g, x = _gpu_grads(body, target, substeps, every)
    assert np.abs(x - x64).max() <= 1e-5 * np.abs(x64).max()
    print({k: f"{scaled_err(g[k], ref[k]):.2e}" for k in ("kd", "kb", "x", "v")}, flush=True)
    for k in ("kd", "kb", "x", "v"):
        assert np.abs(ref[k]).max() > 0
        assert scaled_err(g[k], ref[k]) <= TOL, (k, scaled_err(g[k], ref[k]))
    body.close()


def test_adjoint_with_contacts(oracle_mod):
    print("adjoint contacts size=8 substeps=4", flush=True)
    sc, body, kd, kb, v0, target = _problem(8, selfcoll=True)
    substeps = 4
    gpu_counts = _gpu_counts(body, sc, v0, substeps)
    ref, counts, _ = _oracle_grads(oracle_mod, sc, body, kd, kb, v0, target, substeps)
    assert gpu_counts == counts
    assert sum(a + b for a, b in counts) > 0  # contacts take part
    g, _ = _gpu_grads(body, target, substeps, 2)
    print({k: f"{scaled_err(g[k], ref[k]):.2e}" for k in ("kd", "kb", "x", "v")}, flush=True)
    for k in ("kd", "kb", "x", "v"):
        assert scaled_err(g[k], ref[k]) <= TOL, (k, scaled_err(g[k], ref[k]))
    body.close()


def test_adjoint_plant_fails(oracle_mod):
    """The comparison has teeth: the oracle's gradient with the contact cotangents
    left out (self-collision off in the oracle only) differs from the GPU's."""
    print("adjoint plant size=8", flush=True)
    sc, body, kd, kb, v0, target = _problem(8, selfcoll=True)
    nocoll = dict(sc, params=dict(sc["params"],
                                  flags=sc["params"]["flags"] & ~configs.F_SELF_COLLISION))
    ref, _, _ = _oracle_grads(oracle_mod, nocoll, body, kd, kb, v0, target, 4)
    g, _ = _gpu_grads(body, target, 4, 2)
    assert max(scaled_err(g[k], ref[k]) for k in ("kd", "kb", "x", "v")) > 10 * TOL
    body.close()


def test_checkpoint_interval_and_state():
    print("adjoint checkpoint invariance size=16", flush=True)
    sc, body, kd, kb, v0, target = _problem(16, selfcoll=True)
    tri0 = body.buffer("tri").copy()
    out = []
    for every in (1, 3, 7):
        body.set_state(sc["xyz"], v0, tri=tri0)
        body.forward_ckpt(7, every)
        xf, vf = body.positions(), body.velocities()
        _, lx = _loss_grad(xf, target)
        out.append(body.adjoint(lx, 0.1 * lx))
        assert np.array_equal(body.positions(), xf)  # left in the forward's state
        assert np.array_equal(body.velocities(), vf)
    for o in out[1:]:
        for a, b in zip(out[0], o):
            assert np.array_equal(a, b)
    body.close()


def test_autograd_function_and_fd():
    import torch
    from licenta_amd.diff import Rollout
    print("adjoint autograd + fd size=16", flush=True)
    sc, body, kd, kb, v0, target = _problem(16, selfcoll=False)
    substeps = 5
    roll = Rollout(body, substeps, every=2)
    tk = torch.tensor(kd, dtype=torch.float64, requires_grad=True)
    tb = torch.tensor(kb, dtype=torch.float64, requires_grad=True)
    tx = torch.tensor(sc["xyz"], dtype=torch.float64, requires_grad=True)
    tv = torch.tensor(v0, dtype=torch.float64, requires_grad=True)
    tt = torch.tensor(target)
    x, v = roll(tk, tb, tx, tv)
    loss = ((x - tt) ** 2).sum(dim=1).mean()
    loss.backward()
    body.set_state(sc["xyz"], v0, tri=roll.tri0)
    g, _ = _gpu_grads(body, target, substeps, 2)
    # the same call underneath; dL/dx reaches it through torch's float64 autograd
    np.testing.assert_allclose(tk.grad.numpy(), g["kd"], rtol=1e-5, atol=1e-6 * np.abs(g["kd"]).max())
    np.testing.assert_allclose(tv.grad.numpy(), g["v"], rtol=1e-5, atol=1e-6 * np.abs(g["v"]).max())

    def loss_at(kdv):
        x, _ = roll(torch.tensor(kdv), tb.detach(), tx.detach(), tv.detach())
        return _loss_grad(x.numpy(), target)[0]

    h = 1e-3
    for i in np.argsort(-np.abs(g["kd"]))[:3]:
        lo, hi = kd.copy(), kd.copy()
        lo[i] -= h
        hi[i] += h
        fd = (loss_at(hi) - loss_at(lo)) / (2 * h)
        assert abs(fd - g["kd"][i]) <= 2e-2 * abs(fd), (i, fd, g["kd"][i])
    body.close()


def test_adjoint_rejects():
    from licenta_amd.body import ClothBody
    print("adjoint rejects", flush=True)
    sc = configs.config_scene(3, 8)
    body = ClothBody.from_scene(sc)
    with pytest.raises(RuntimeError, match="error -4"):
        body.adjoint(np.zeros((body.n, 3), np.float32))  # no forward on record
    with pytest.raises(RuntimeError, match="error -1"):
        body.set_stiffness(np.full(body.counts()["s"], 1.5, np.float32))
    body.close()
    # a sphere set after the forward: the record is dropped, the adjoint refuses
    body = ClothBody.from_scene(sc)
    body.forward_ckpt(2, 1)
    body.set_sphere_pos((0.0, -1.0, 0.0), 0.1)
    with pytest.raises(RuntimeError, match="error -4"):
        body.adjoint(np.zeros((body.n, 3), np.float32))
    body.close()
    sc = dict(sc, params=dict(sc["params"], flags=sc["params"]["flags"] | configs.F_GROUND))
    body = ClothBody.from_scene(sc)
    with pytest.raises(RuntimeError, match="error -1"):
        body.forward_ckpt(2, 1)
    body.close()


def test_batched_adjoint_matches_single_bodies():
    """Three 64 x 64 cloths of the batched-gradient workload stepped as one body give
    each cloth the gradients of its own single-cloth body."""
    from licenta_amd.body import ClothBody
    print("adjoint batch of 3 x 64^2", flush=True)
    scenes = configs.batched_scene(4, count=3)
    substeps, every = 6, 4
    rng = np.random.default_rng(3)
    v0s = [rng.normal(0.0, 0.05, size=(len(s["xyz"]), 3)).astype(np.float32) for s in scenes]
    targets = [s["xyz"] + rng.normal(0.0, 0.01, size=(len(s["xyz"]), 3)) for s in scenes]
    singles = []
    for s, v0, tg in zip(scenes, v0s, targets):
        b = ClothBody.from_scene(s)
        b.set_stiffness(np.full(b.counts()["s"], 0.9, np.float32))  # dk/dkp is finite below 1
        b.set_state(s["xyz"], v0)
        b.forward_ckpt(substeps, every)
        lx = 2.0 * (b.positions() - tg) / len(tg)
        singles.append((b.counts(), b.adjoint(lx)))
        b.close()
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
        assert scaled_err(gx[n0[c]:n0[c + 1]], sx) <= 1e-5
        assert scaled_err(gv[n0[c]:n0[c + 1]], sv) <= 1e-5
        assert scaled_err(gkd[s0[c]:s0[c + 1]], skd) <= 1e-5
        assert scaled_err(gkb[b0[c]:b0[c + 1]], skb) <= 1e-5
    batch.close()


def test_contact_tape_growth():
    """With a contact tape too small for a segment (one record per substep at first),
    cm_adjoint grows it and recomputes the segment: the gradients equal those of a
    body whose tape never overflows."""
    from licenta_amd.body import ClothBody
    print("adjoint tape growth size=8", flush=True)
    sc, body, kd, kb, v0, target = _problem(8, selfcoll=True)
    assert sum(a + b for a, b in _gpu_counts(body, sc, v0, 4)) > 0  # contacts occur
    g_ref, _ = _gpu_grads(body, target, 4, 2)
    body.close()
    small = dict(sc, params=dict(sc["params"], reserved=(0, 0, 0, 1, 0, 0, 0)))
    b = ClothBody.from_scene(small)
    b.set_stiffness(kd, kb)
    b.set_state(sc["xyz"], v0)
    g, _ = _gpu_grads(b, target, 4, 2)
    for k in ("kd", "kb", "x", "v"):
        assert np.array_equal(g[k], g_ref[k]), k
    b.close()
