"""cm_tri_update_noalloc (no device allocation, static scratch) against
cm_tri_update through the C ABI: bit-identical leaf AABBs and Morton keys on two
calls in a row, the second with a smaller input, for the normalised 30-bit path
that uses the scratch and the 63-bit path that does not; sizes beyond the static
scratch are refused."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _inputs(size, seed):
    from licenta_amd import configs
    sc = configs.config_scene(0, size)
    rng = np.random.default_rng(seed)
    x = np.zeros((len(sc["xyz"]), 4), np.float32)
    x[:, :3] = sc["xyz"]
    p = x.copy()
    p[:, :3] += rng.uniform(-0.002, 0.002, (len(x), 3)).astype(np.float32)
    tri = np.zeros((len(sc["tri"]), 4), np.int32)
    tri[:, :3] = sc["tri"]
    return x, p, tri


def _run(fn, x, p, tri, bits, normalize):
    from licenta_amd._abi import check
    t = len(tri)
    dx, dp, dt = (torch.from_numpy(a).cuda() for a in (x, p, tri))
    dmin = torch.full((2 * t - 1, 4), 7.0, dtype=torch.float32, device="cuda")
    dmax = torch.full((2 * t - 1, 4), 7.0, dtype=torch.float32, device="cuda")
    dk = torch.zeros(t, dtype=torch.int64, device="cuda")
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    vp = lambda a: C.c_void_p(a.data_ptr())
    check(fn(len(x), t, vp(dx), vp(dp), vp(dt), 0.003, bits, normalize, vp(dmin), vp(dmax),
             vp(dk), s))
    torch.cuda.synchronize()
    return dmin.cpu().numpy(), dmax.cpu().numpy(), dk.cpu().numpy()


@pytest.mark.parametrize("bits,normalize", [(30, 1), (63, 0)])
def test_noalloc_matches_tri_update(bits, normalize):
    from licenta_amd._abi import lib
    L = lib()
    for size, seed in ((96, 1), (40, 2)):  # 18,050 then 3,042 triangles
        print(f"[noalloc] bits={bits} size={size}", flush=True)
        x, p, tri = _inputs(size, seed)
        want = _run(L.cm_tri_update, x, p, tri, bits, normalize)
        got = _run(L.cm_tri_update_noalloc, x, p, tri, bits, normalize)
        for a, b in zip(want, got):
            assert np.array_equal(a, b)


def test_noalloc_rejects_beyond_scratch():
    from licenta_amd._abi import lib
    L = lib()
    t = (1 << 18) + 1
    dummy = torch.zeros(16, dtype=torch.float32, device="cuda")
    vp = C.c_void_p(dummy.data_ptr())
    rc = L.cm_tri_update_noalloc(3, t, vp, vp, vp, 0.003, 30, 1, vp, vp, vp,
                                 C.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == -1
    assert b"static scratch" in L.cm_last_error()
