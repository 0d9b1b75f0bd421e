"""GPU parity of the unit-level LBVH entry points (cm_tri_update, cm_lbvh_sort,
cm_lbvh_build, cm_lbvh_refit, cm_lbvh_pairs) against the oracle on identical
FP32 inputs: bit-exact keys, order, tree arrays, node AABBs and pair lists."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def lib():
    from licenta_amd import _abi
    return _abi.lib()


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def vp(t):
    return C.c_void_p(t.data_ptr())


def stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def check(rc):
    from licenta_amd._abi import check as chk
    chk(rc)


def f4(a3):
    out = np.zeros((len(a3), 4), np.float32)
    out[:, :3] = a3
    return out


def run_sort(keys, vals, bits):
    t = len(keys)
    L = lib()
    dk = dev(keys.view(np.int64))
    dv = dev(vals.view(np.int32))
    nbytes = L.cm_lbvh_sort_temp_bytes(t)
    tmp = torch.empty(max(nbytes, 1), dtype=torch.uint8, device="cuda")
    check(L.cm_lbvh_sort(t, bits, vp(dk), vp(dv), vp(tmp), nbytes, stream()))
    torch.cuda.synchronize()
    return dk.cpu().numpy().view(np.uint64), dv.cpu().numpy().view(np.uint32)


@pytest.mark.parametrize("t", [1, 2, 255, 2048, 2049, 70001])
@pytest.mark.parametrize("bits", [30, 63])
def test_sort_stable(oracle_mod, t, bits):
    rng = np.random.default_rng(t + bits)
    keys = rng.integers(0, 1 << bits, size=t, dtype=np.uint64)
    keys[: t // 3] = keys[0]  # many duplicates
    rng.shuffle(keys)
    vals = np.arange(t, dtype=np.uint32)
    gk, gv = run_sort(keys, vals, bits)
    ok, ov = oracle_mod.stable_sort(keys, vals)
    assert np.array_equal(gk, ok)
    assert np.array_equal(gv, ov)


def make_tree_inputs(t, seed, ties):
    rng = np.random.default_rng(seed)
    if ties:
        keys = np.sort(rng.integers(0, 64, size=t).astype(np.uint64) * np.uint64(1 << 20))
    else:
        keys = np.unique(rng.integers(0, 1 << 62, size=2 * t, dtype=np.int64))[:t].astype(np.uint64)
    lo = rng.uniform(0, 1, size=(t, 3)).astype(np.float32)
    hi = lo + rng.uniform(0, 0.08, size=(t, 3)).astype(np.float32)
    return keys, lo, hi


# 300,000 leaves: more than 2,048 deferred nodes, so the refit's middle pass runs
# (with tied codes the midpoint split makes a broken tree, whose shared nodes can
# leave parents to the top pass)
@pytest.mark.parametrize("t,ties,midpoint", [(2, False, True), (777, False, True),
                                             (5000, True, True), (5000, True, False),
                                             (40000, False, True), (300000, False, True),
                                             (300000, True, True)])
def test_build_refit_pairs(oracle_mod, t, ties, midpoint):
    print(f"build/refit/pairs t={t} ties={ties} midpoint={midpoint}", flush=True)
    keys, lo, hi = make_tree_inputs(t, t, ties)
    L = lib()
    s = stream()
    m = 2 * t - 1
    parents0 = np.full(m, 7, np.int32)
    dk = dev(keys.view(np.int64))
    dl = torch.zeros(t, dtype=torch.int32, device="cuda")
    dr = torch.zeros(t, dtype=torch.int32, device="cuda")
    dp = dev(parents0)
    dw = torch.zeros(1, dtype=torch.int32, device="cuda")
    nb = L.cm_lbvh_build_temp_bytes(t)
    tmp = torch.empty(nb, dtype=torch.uint8, device="cuda")
    check(L.cm_lbvh_build(t, vp(dk), int(midpoint), vp(dl), vp(dr), vp(dp), vp(dw), vp(tmp), nb, s))
    amin = np.zeros((m, 3), np.float32)
    amax = np.zeros((m, 3), np.float32)
    amin[t - 1:] = lo
    amax[t - 1:] = hi
    dmin, dmax = dev(f4(amin)), dev(f4(amax))
    drl = torch.zeros(t, dtype=torch.int32, device="cuda")
    drr = torch.zeros(t, dtype=torch.int32, device="cuda")
    check(L.cm_lbvh_refit(t, vp(dl), vp(dr), vp(dp), vp(dmin), vp(dmax), vp(drl), vp(drr), vp(dw),
                          vp(tmp), nb, s))
    cap = 16
    pb = L.cm_lbvh_pairs_temp_bytes(t, cap)
    ptmp = torch.empty(pb, dtype=torch.uint8, device="cuda")
    dpairs = torch.zeros(2 * t * cap, dtype=torch.int32, device="cuda")
    dnp = torch.zeros(1, dtype=torch.int32, device="cuda")
    check(L.cm_lbvh_pairs(t, vp(dl), vp(dr), vp(dmin), vp(dmax), vp(drl), vp(drr), cap, vp(dpairs),
                          vp(dnp), vp(dw), vp(ptmp), pb, s))
    torch.cuda.synchronize()
    ol, orr, op, ow = oracle_mod.build(keys, midpoint, parents0)
    assert np.array_equal(dl.cpu().numpy()[: t - 1], ol)
    assert np.array_equal(dr.cpu().numpy()[: t - 1], orr)
    assert np.array_equal(dp.cpu().numpy(), op)
    omin, omax, orml, ormr = oracle_mod.refit(ol, orr, amin, amax)
    assert np.array_equal(dmin.cpu().numpy()[:, :3], omin)
    assert np.array_equal(dmax.cpu().numpy()[:, :3], omax)
    assert np.array_equal(drl.cpu().numpy()[: t - 1], orml)
    assert np.array_equal(drr.cpu().numpy()[: t - 1], ormr)
    opairs, owarn = oracle_mod.pairs(ol, orr, omin, omax, orml, ormr, cap)
    npairs = int(dnp.cpu().numpy()[0])
    gp = dpairs.cpu().numpy()[: 2 * npairs].reshape(-1, 2)
    assert np.array_equal(gp, opairs)
    gw = int(dw.cpu().numpy()[0]) & 0xFFFF
    assert gw == (owarn | ow)


@pytest.mark.parametrize("bits,normalize", [(63, 0), (30, 1), (30, 0)])
def test_tri_update_keys(oracle_mod, bits, normalize):
    from licenta_amd import configs
    sc = configs.config_scene(0, 32)
    n = len(sc["xyz"])
    rng = np.random.default_rng(bits)
    x = sc["xyz"].astype(np.float32)
    p = x + rng.uniform(-0.002, 0.002, x.shape).astype(np.float32)
    tri = np.zeros((len(sc["tri"]), 4), np.int32)
    tri[:, :3] = sc["tri"]
    t = len(tri)
    th = np.float32(0.003)
    L = lib()
    dmin = torch.zeros((2 * t - 1, 4), dtype=torch.float32, device="cuda")
    dmax = torch.zeros((2 * t - 1, 4), dtype=torch.float32, device="cuda")
    dk = torch.zeros(t, dtype=torch.int64, device="cuda")
    dx, dp, dt = dev(f4(x)), dev(f4(p)), dev(tri)  # keep the inputs alive until the sync
    check(L.cm_tri_update(n, t, vp(dx), vp(dp), vp(dt), float(th), bits, normalize, vp(dmin),
                          vp(dmax), vp(dk), stream()))
    torch.cuda.synchronize()
    # oracle restatement of the swept AABB (float32 arithmetic in numpy, exact ops only)
    P, X = p[tri[:, :3]], x[tri[:, :3]]
    mn = np.minimum(P.min(1) - th, X.min(1) - th)
    mx = np.maximum(P.max(1) + th, X.max(1) + th)
    assert np.array_equal(dmin.cpu().numpy()[t - 1:, :3], mn)
    assert np.array_equal(dmax.cpu().numpy()[t - 1:, :3], mx)
    cen = (mn + (mx - mn) / np.float32(2)).astype(np.float32)
    if bits == 63:
        want = oracle_mod.morton63(cen)
    elif normalize:
        want = oracle_mod.morton30(cen, cen.min(0), cen.max(0))
    else:
        want = oracle_mod.morton30(cen, [0, 0, 0], [1, 1, 1])
    assert np.array_equal(dk.cpu().numpy().view(np.uint64), want)
