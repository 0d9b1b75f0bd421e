"""GPU parity of whole substeps (cm_create / cm_step through the C ABI) against
the oracle on identical starting states.

Per substep, re-seeded from the GPU state:
* integer stages bit-exact vs the FP32 oracle: Morton keys, sort permutation,
  sorted triangles with feature bits, lefts/rights/parents, rightmost leaves,
  node AABBs, the triangle-pair list and the VF/EE contact lists;
* positions vs the FP32 oracle within 2e-6 relative (only acosf/powf rounding
  can differ), and vs the FP64 oracle within 1e-5 relative (the stated bar);
  relative error = max |x_gpu - x_ref| / max |x_ref|.
"""
import numpy as np
import pytest

from licenta_amd import configs

pytestmark = pytest.mark.gpu

TOL64 = 1e-5
TOL32 = 2e-6


def rel_err(a, b):
    return float(np.abs(np.asarray(a, np.float64) - b).max() / max(np.abs(b).max(), 1e-30))


def seeded_pair(oracle_mod, scene, body):
    """Oracles (FP32, FP64) holding exactly the GPU body's current state."""
    from licenta_amd.body import ClothBody  # noqa: F401
    x = body.buffer("x")[:, :3].astype(np.float64)
    p = body.buffer("p")[:, :3].astype(np.float64)
    v = body.buffer("v")[:, :3].astype(np.float64)
    tri = body.buffer("tri")
    out = []
    for prec in (32, 64):
        o = oracle_mod.OracleSim(scene, prec)
        o.set_state(x, v, p, tri)
        o.set_parents(body.buffer("parents"))
        out.append(o)
    return out


def compare_substep(oracle_mod, scene, body, check_tree=True, tol64=TOL64, dev=False,
                    replay64=False):
    """One substep of the GPU body against the FP32 and FP64 oracles seeded with its
    state. With replay64 a third oracle, in FP64, takes the FP32 oracle's contact
    activation decisions of every iteration (cc < 0, Physics/CudaPBD.cu:1295,1400)
    instead of evaluating them: the FP64 arithmetic on the path the FP32 run took."""
    o32, o64 = seeded_pair(oracle_mod, scene, body)
    if replay64:
        o32.activation_mode(1)
        o64r = seeded_pair(oracle_mod, scene, body)[1]
    body.fixed_update(1)
    o32.step(1)
    o64.step(1)
    if replay64:
        o64r.activation_mode(2)
        o64r.set_activations(o32.activations())
        o64r.step(1)
        assert not o64r.activation_mismatch()
    t = body.t
    res = {}
    if scene["params"]["flags"] & configs.F_SELF_COLLISION:
        assert np.array_equal(body.buffer("keys"), o32.keys())
        assert np.array_equal(body.buffer("perm"), o32.perm())
        assert np.array_equal(body.buffer("tri"), o32.get_i32("tri").reshape(-1, 4))
        if check_tree:
            assert np.array_equal(body.buffer("lefts"), o32.get_i32("lefts"))
            assert np.array_equal(body.buffer("rights"), o32.get_i32("rights"))
            assert np.array_equal(body.buffer("parents"), o32.get_i32("parents"))
            assert np.array_equal(body.buffer("rmll"), o32.get_i32("rmll"))
            assert np.array_equal(body.buffer("rmlr"), o32.get_i32("rmlr"))
            mn, mx = o32.aabbs()
            assert np.array_equal(body.buffer("aabb_min")[:, :3], mn)
            assert np.array_equal(body.buffer("aabb_max")[:, :3], mx)
        gp = body.pairs()
        assert np.array_equal(gp, o32.pairs())
        gvf, gee = body.contacts()
        ovf, oee = o32.contacts(False), o32.contacts(True)
        assert np.array_equal(gvf["v"], ovf["v"]) and np.array_equal(gee["v"], oee["v"])
        assert np.array_equal(gvf["w"], ovf["w"]) and np.array_equal(gee["w"], oee["w"])
        assert np.array_equal(gvf["t"], ovf["t"]) and np.array_equal(gee["t"], oee["t"])
        assert np.array_equal(gvf["n"], ovf["n"]) and np.array_equal(gee["n"], oee["n"])
        res.update(pairs=len(gp), vf=len(gvf), ee=len(gee))
        # the FP64 oracle's discrete stages run on the FP32 shadow: same pairs
        assert np.array_equal(o64.pairs(), gp)
    x = body.positions()
    x32, _, _ = o32.state()
    x64, _, _ = o64.state()
    res["err32"] = rel_err(x, x32)
    res["err64"] = rel_err(x, x64)
    assert res["err32"] <= TOL32, res
    assert res["err64"] <= tol64, res
    if dev:  # per-vertex deviation from FP64, relative to the largest coordinate
        res["dev64"] = np.abs(np.asarray(x, np.float64) - x64).max(axis=1) / np.abs(x64).max()
    if replay64:
        xr, _, _ = o64r.state()
        res["err64_replay"] = rel_err(x, xr)
        # per-vertex relative error: |x_gpu - x_64| / |x_64| per vertex
        d = np.linalg.norm(np.asarray(x, np.float64) - xr, axis=1)
        res["vertex_rel64_replay"] = float((d / np.maximum(np.linalg.norm(xr, axis=1),
                                                            1e-30)).max())
        res["active_decisions"] = int(o32.activations().sum())
    return res


def run_scene(oracle_mod, scene, substeps, check_tree=True):
    from licenta_amd.body import ClothBody
    body = ClothBody.from_scene(scene)
    out = []
    for s in range(substeps):
        print(f"[{scene['name']}] substep {s}", flush=True)
        out.append(compare_substep(oracle_mod, scene, body, check_tree))
    w = body.warnings()
    body.close()
    return out, w


def test_plain_small_cloth(oracle_mod):
    sc = configs.config_scene(0, 24)
    res, w = run_scene(oracle_mod, sc, 4)
    assert all(r["pairs"] > 0 for r in res)


def test_plain_64(oracle_mod):
    sc = configs.config_scene(0, 64)
    res, w = run_scene(oracle_mod, sc, 2)
    print(res, hex(w))


def test_sphere_drop_small(oracle_mod):
    sc = configs.config_scene(1, 40)
    # start just above the sphere so the sphere projection acts in the first substeps
    sc["xyz"][:, 1] = sc["xyz"][:, 1] - np.float32(0.5 - 0.305)
    res, w = run_scene(oracle_mod, sc, 3)


def test_fit_scene_small(oracle_mod):
    """The fitting workload's forward (pinned corners, per-constraint stiffness,
    random initial velocities) at 24 x 24."""
    sc = configs.config_scene(3, 24)
    from licenta_amd.body import ClothBody
    body = ClothBody.from_scene(sc)
    body.set_state(sc["xyz"], sc["v0"])
    for s in range(3):
        print(f"[{sc['name']}] substep {s}", flush=True)
        compare_substep(oracle_mod, sc, body)
    body.close()


def test_crumpled_small(oracle_mod):
    """The crumpled large-cloth workload (sinusoid-displaced, dense self-contact at
    thickness 0.5 mm, 20 iterations) at 40 x 40."""
    sc = configs.config_scene(4, 40)
    res, w = run_scene(oracle_mod, sc, 3)
    assert all(r["pairs"] > 0 for r in res)


def test_reference_scene_defaults(oracle_mod):
    """Ground, world clamp, damping, LRA and 63-bit raw Morton codes on."""
    sc = configs.reference_scene(24)
    res, w = run_scene(oracle_mod, sc, 3)


def test_fixed_leaf_aabb_and_relative_bend(oracle_mod):
    sc = configs.config_scene(0, 20)
    sc["params"] = dict(sc["params"], ref_compat=configs.RC_MIDPOINT_SPLIT)
    res, w = run_scene(oracle_mod, sc, 3)


def test_folded_cloth_contacts(oracle_mod):
    """A cloth folded onto itself so the narrow phase produces VF/EE contacts."""
    n = 24
    sc = configs.config_scene(0, n)
    xyz = sc["xyz"].reshape(n, n, 3).copy()
    # fold the second half of the rows back over the first, 3 mm above it
    h = n // 2
    for i in range(h, n):
        src = 2 * h - 1 - i
        xyz[i, :, 0] = xyz[src, :, 0] + 0.002
        xyz[i, :, 1] = 0.003 + 0.0005 * (i - h)
    sc["xyz"] = xyz.reshape(-1, 3).astype(np.float32)
    sc["params"] = dict(sc["params"], thickness=0.004)
    sc["pinned"] = None
    sc["inv_mass"] = np.full(n * n, 1000.0, np.float32)
    from licenta_amd.body import ClothBody
    body = ClothBody.from_scene(sc)
    v = np.zeros((n * n, 3), np.float32)
    v.reshape(n, n, 3)[h:, :, 1] = -0.3
    body.set_state(sc["xyz"], v)
    total = 0
    for s in range(3):
        r = compare_substep(oracle_mod, sc, body)
        total += r["vf"] + r["ee"]
    body.close()
    assert total > 0


def test_folded_cloth_dcd(oracle_mod):
    """The discrete narrow phase (CM_F_DCD): the folded cloth's candidates tested
    once at the start positions, three substeps against the oracle (contact sets bit
    for bit, positions), with other contacts than the continuous phase finds on the
    same first substep (the flag switches the path: 103 discrete against 133
    continuous contacts on the first substep)."""
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
    from licenta_amd.body import ClothBody
    v = np.zeros((n * n, 3), np.float32)
    v.reshape(n, n, 3)[h:, :, 1] = -0.3
    first = {}
    for name, flags in (("ccd", sc["params"]["flags"]),
                        ("dcd", sc["params"]["flags"] | configs.F_DCD)):
        s2 = dict(sc, params=dict(sc["params"], thickness=0.004, flags=flags))
        body = ClothBody.from_scene(s2)
        body.set_state(s2["xyz"], v)
        for k in range(3):
            r = compare_substep(oracle_mod, s2, body)
            if k == 0:
                first[name] = r["vf"] + r["ee"]
        body.close()
    print(f"  first-substep contacts: continuous {first['ccd']}, discrete {first['dcd']}")
    assert first["dcd"] > 0 and first["ccd"] > 0 and first["dcd"] != first["ccd"]


@pytest.mark.parametrize("tail", [True, False])
def test_friction_folded_on_ground(oracle_mod, tail):
    """Friction (ground step, then the substep's VF / EE contacts) on a 16 x 16 cloth
    folded onto itself on the ground: 256 vertices, so the reference's 255-thread
    ground step misses the last one when that behaviour is on. Friction changes only
    the velocities, which each substep here compares with the oracle's (relative to
    the largest speed) besides everything compare_substep checks."""
    n = 16
    sc = configs.config_scene(0, n)
    xyz = sc["xyz"].reshape(n, n, 3).copy()
    h = n // 2
    for i in range(h, n):
        src = 2 * h - 1 - i
        xyz[i, :, 0] = xyz[src, :, 0] + 0.002
        xyz[i, :, 1] = 0.003 + 0.0005 * (i - h)
    sc["xyz"] = xyz.reshape(-1, 3).astype(np.float32)
    rc = configs.RC_ALL if tail else configs.RC_ALL & ~configs.RC_RIGID_FRICTION_TAIL
    sc["params"] = dict(sc["params"], thickness=0.004, k_friction=0.25, ref_compat=rc,
                        flags=sc["params"]["flags"] | configs.F_GROUND | configs.F_FRICTION)
    sc["pinned"] = None
    sc["inv_mass"] = np.full(n * n, 1000.0, np.float32)
    from licenta_amd.body import ClothBody
    body = ClothBody.from_scene(sc)
    rng = np.random.default_rng(11)
    v = rng.normal(0, 0.05, (n * n, 3)).astype(np.float32)
    v.reshape(n, n, 3)[h:, :, 1] = -0.3
    body.set_state(sc["xyz"], v)
    total = 0
    for s in range(4):
        print(f"[friction tail={tail}] substep {s}", flush=True)
        o32, o64 = seeded_pair(oracle_mod, sc, body)
        body.fixed_update(1)
        o32.step(1)
        o64.step(1)
        gvf, gee = body.contacts()
        ovf, oee = o32.contacts(False), o32.contacts(True)
        assert np.array_equal(gvf["v"], ovf["v"]) and np.array_equal(gee["v"], oee["v"])
        total += len(gvf) + len(gee)
        x, vg = body.positions(), body.velocities()
        x32, _, v32 = o32.state()
        x64, _, v64 = o64.state()
        assert rel_err(x, x32) <= TOL32 and rel_err(x, x64) <= TOL64
        # v = (p - x) / dt turns a position rounding difference into a larger
        # velocity one: 1e-3 of the largest speed
        e32, e64 = rel_err(vg, v32), rel_err(vg, v64)
        print(f"  contacts {len(gvf)} + {len(gee)}, velocity rel err {e32:.2e} / {e64:.2e}",
              flush=True)
        assert e32 <= 1e-3 and e64 <= 1e-3
    body.close()
    assert total > 0


def test_revert_to_toc(oracle_mod):
    """CM_F_REVERT_TOC: two triangles meeting inside a substep at three contact times
    (the oracle's closed-form case), then a folded cloth whose contacts start in
    contact (t = 0, so contact vertices return to their start)."""
    from licenta_amd.body import ClothBody
    prm = configs.default_params()
    prm.update(gravity=(0.0, 0.0, 0.0), k_stretch=0.5, k_bend=0.0, thickness=0.001,
               solver_iterations=2, morton_bits=30, morton_normalize=1,
               flags=configs.F_SELF_COLLISION | configs.F_EE_CONTACTS | configs.F_REVERT_TOC)
    xyz = np.array([[0, 0, 0], [1, 0, 0], [0, 0, 1], [0.2, 0.00101, 0.2], [0.3, 0.001012, 0.2],
                    [0.2, 0.001015, 0.3]], np.float32)
    sc = dict(name="toc_pair", params=prm, xyz=xyz, tri=np.array([[0, 2, 1], [3, 5, 4]], np.uint32),
              inv_mass=np.ones(6, np.float32), pinned=None, dist_pairs=None, k_dist=None,
              k_bend=None)
    body = ClothBody.from_scene(sc)
    v = np.zeros((6, 3), np.float32)
    v[:3, 1], v[3:, 1] = 0.3, -0.3
    body.set_state(xyz, v)
    print("[toc_pair] substep 0", flush=True)
    r = compare_substep(oracle_mod, sc, body)
    assert r["vf"] == 3
    body.close()
    n = 24
    sc = configs.config_scene(0, n)
    xyz = sc["xyz"].reshape(n, n, 3).copy()
    h = n // 2
    for i in range(h, n):
        xyz[i, :, 0] = xyz[2 * h - 1 - i, :, 0] + 0.002
        xyz[i, :, 1] = 0.003 + 0.0005 * (i - h)
    sc["xyz"] = xyz.reshape(-1, 3).astype(np.float32)
    sc["params"] = dict(sc["params"], thickness=0.004,
                        flags=sc["params"]["flags"] | configs.F_REVERT_TOC)
    res, w = run_scene(oracle_mod, sc, 3)
    assert sum(r["vf"] + r["ee"] for r in res) > 0


def test_carried_vertex_normals(oracle_mod):
    """CM_F_VERTEX_NORMALS: the creation normals, then the normals carried through
    four substeps of a wavy cloth (from each substep's predicted positions) against
    the FP32 oracle run
    alongside from the same start (not re-seeded, so the carried normals of both
    build up over the whole run)."""
    from licenta_amd.body import ClothBody
    n = 24
    sc = configs.config_scene(0, n)
    xyz = sc["xyz"].reshape(n, n, 3).copy()
    xyz[:, :, 1] += 0.01 * np.sin(np.linspace(0, 3, n))[:, None]
    sc["xyz"] = xyz.reshape(-1, 3).astype(np.float32)
    sc["params"] = dict(sc["params"], flags=sc["params"]["flags"] | configs.F_VERTEX_NORMALS)
    body = ClothBody.from_scene(sc)
    o = oracle_mod.OracleSim(sc, 32)
    assert float(np.abs(body.normals() - o.normals()).max()) <= 1e-6  # creation normals
    for s in range(4):
        print(f"[normals] substep {s}", flush=True)
        body.fixed_update(1)
        o.step(1)
        x32, _, _ = o.state()
        assert rel_err(body.positions(), x32) <= 1e-5
        err = float(np.abs(body.normals() - o.normals()).max())
        print(f"  normal max abs err {err:.2e}", flush=True)
        assert err <= 1e-5
    body.close()


def test_fresh_normals_and_vertex_format(oracle_mod):
    """Without CM_F_VERTEX_NORMALS, cm_read_normals gives the normals of the current
    positions; cm_read_vertices writes positions and normals into 48-byte
    VertexFormat records and keeps their color and uv."""
    from licenta_amd._abi import VERTEX_DTYPE
    from licenta_amd.body import ClothBody
    sc = configs.config_scene(0, 20)
    sc["params"] = dict(sc["params"], flags=sc["params"]["flags"] & ~configs.F_VERTEX_NORMALS)
    body = ClothBody.from_scene(sc)
    body.fixed_update(5)
    x = body.positions()
    nrm = body.normals()
    o = oracle_mod.OracleSim(sc, 32)
    assert float(np.abs(nrm - o.normals(x)).max()) <= 1e-6
    rng = np.random.default_rng(3)
    verts = np.zeros(body.n, VERTEX_DTYPE)
    verts["color"] = rng.random((body.n, 4), dtype=np.float32)
    verts["uv"] = rng.random((body.n, 2), dtype=np.float32)
    keep = verts.copy()
    body.read_vertices(verts)
    assert np.array_equal(verts["position"], x) and np.array_equal(verts["normal"], nrm)
    assert np.array_equal(verts["color"], keep["color"]) and np.array_equal(verts["uv"], keep["uv"])
    with pytest.raises(ValueError):
        body.read_vertices(np.zeros(body.n - 1, VERTEX_DTYPE))
    body.close()


@pytest.mark.timeout(300)
def test_full_size_sphere_drop(oracle_mod):
    """The benchmark workload at its full size (256 x 256 cloth onto the sphere).
    First two substeps: every stage against the oracles, positions within 2e-6 of
    FP32 and 1e-5 of FP64. Substep 1,201, after the cloth has draped over the sphere
    and self-contact is at its densest (millions of pairs, tens of thousands of
    contacts): every integer stage bit-exact, positions within 2e-6 of the FP32
    oracle, and within 1e-5 relative (max |dx| over max |x|) of the FP64 oracle run
    on the FP32 run's contact activation decisions, with the per-vertex relative
    error |dx| / |x| reported beside it. A contact's inequality (distance
    below the thickness) can switch on in one iteration in FP32 and in another in
    FP64 when the distance sits within rounding of the thickness; the FP64 oracle
    that evaluates its own decisions then differs by about one contact correction
    (bounded here by 1e-4, with every vertex beyond 1e-5 in a contact)."""
    from licenta_amd.body import ClothBody
    sc = configs.config_scene(1)
    body = ClothBody.from_scene(sc)
    for s in range(2):
        print(f"[{sc['name']}] substep {s}", flush=True)
        compare_substep(oracle_mod, sc, body)
    body.fixed_update(1198)
    print(f"[{sc['name']}] substep 1200", flush=True)
    r = compare_substep(oracle_mod, sc, body, tol64=1e-4, dev=True, replay64=True)
    dev = r.pop("dev64")
    print(f"  {r}; vertices beyond 1e-5 of FP64: {int((dev > TOL64).sum())}", flush=True)
    assert r["pairs"] > 5_000_000 and r["vf"] + r["ee"] > 10_000
    # the stated bar on the global figure; per vertex (|dx| / |x| of each vertex, the
    # cloth lying 0.3-0.5 m from the origin) about 1.1e-5 was measured
    assert r["err64_replay"] <= TOL64 and r["vertex_rel64_replay"] <= 2 * TOL64
    gvf, gee = body.contacts()
    touched = np.zeros(body.n, bool)
    for c in (gvf, gee):
        touched[c["v"].ravel()] = True
    assert touched[dev > TOL64].all()
    body.close()


def test_graph_replay_matches_eager_state(oracle_mod):
    """Several substeps in one cm_step call equal the same substeps one by one."""
    from licenta_amd.body import ClothBody
    sc = configs.config_scene(0, 16)
    a = ClothBody.from_scene(sc)
    b = ClothBody.from_scene(sc)
    a.fixed_update(5)
    for _ in range(5):
        b.fixed_update(1)
    assert np.array_equal(a.positions(), b.positions())
    assert np.array_equal(a.buffer("tri"), b.buffer("tri"))


def test_high_valence_fan_slot_parallel_solve(oracle_mod):
    """A fan whose centre holds more constraint slots (3,400) than a Jacobi tile
    (3,072), so the solver runs its slot-parallel form instead of tiles."""
    m = 1700
    ang = np.linspace(0.0, 1.9 * np.pi, m + 1)
    rim = np.stack([0.3 + 0.05 * np.cos(ang), np.full_like(ang, 0.4), 0.3 + 0.05 * np.sin(ang)], 1)
    xyz = np.vstack([[0.3, 0.4, 0.3], rim]).astype(np.float32)
    tri = np.array([[0, i, i + 1] for i in range(1, m + 1)], dtype=np.uint32)
    params = dict(configs.config_scene(0, 8)["params"])
    params["flags"] = configs.F_BEND_FP64
    inv_mass = np.full(len(xyz), 1000.0, np.float32)
    sc = configs._scene(params, xyz, tri, inv_mass, name="fan")
    res, w = run_scene(oracle_mod, sc, 3)


def test_kernel_timing_leaves_the_step_unchanged():
    """cm_time_kernel on the traversal and on the CCD plan pass (CM_KERNEL_TRAVERSE,
    CM_KERNEL_CCD_PLAN) re-runs them on the last substep's tree and candidates: the
    contact lists and pairs stay as they were, and the following substep equals that
    of a twin body that was not timed, bit for bit."""
    import ctypes as C

    from licenta_amd._abi import lib
    from licenta_amd.body import ClothBody
    n = 24
    sc = configs.config_scene(0, n)
    xyz = sc["xyz"].reshape(n, n, 3).copy()
    h = n // 2
    for i in range(h, n):
        xyz[i, :, 0] = xyz[2 * h - 1 - i, :, 0] + 0.002
        xyz[i, :, 1] = 0.003 + 0.0005 * (i - h)
    sc["xyz"] = xyz.reshape(-1, 3).astype(np.float32)
    sc["pinned"] = None
    sc["inv_mass"] = np.full(n * n, 1000.0, np.float32)
    sc = dict(sc, params=dict(sc["params"], thickness=0.004))
    v = np.zeros((n * n, 3), np.float32)
    v.reshape(n, n, 3)[h:, :, 1] = -0.3
    bodies = [ClothBody.from_scene(sc) for _ in range(2)]
    for b in bodies:
        b.set_state(sc["xyz"], v)
        b.fixed_update(1)
    vf0, ee0 = bodies[0].contacts()
    pairs0 = bodies[0].pairs()
    assert len(vf0) + len(ee0) > 0
    avg = C.c_float()
    for which in (2, 3):
        assert lib().cm_time_kernel(bodies[0].h, which, 3, C.byref(avg)) == 0
        assert avg.value > 0
    vf1, ee1 = bodies[0].contacts()
    assert np.array_equal(vf1, vf0) and np.array_equal(ee1, ee0)
    assert np.array_equal(bodies[0].pairs(), pairs0)
    for b in bodies:
        b.fixed_update(1)
    assert np.array_equal(bodies[0].positions(), bodies[1].positions())
    for b in bodies:
        b.close()


def test_sorted_record_csr_matches_atomic_csr(monkeypatch):
    """The per-vertex record CSR built by the radix sort of all records (the default)
    equals the one built by atomic placement and per-vertex sorts (CLOTHMI_SORT_CSR=0):
    twin bodies, one on each way, step to the same state bit for bit. Cases: the
    folded cloth, the crumpled 40 x 40 cloth (vertices with long record lists) and a
    batch of 8 cloths."""
    from licenta_amd.body import ClothBody
    n = 24
    fold = configs.config_scene(0, n)
    xyz = fold["xyz"].reshape(n, n, 3).copy()
    h = n // 2
    for i in range(h, n):
        xyz[i, :, 0] = xyz[2 * h - 1 - i, :, 0] + 0.002
        xyz[i, :, 1] = 0.003 + 0.0005 * (i - h)
    fold["xyz"] = xyz.reshape(-1, 3).astype(np.float32)
    fold["pinned"] = None
    fold["inv_mass"] = np.full(n * n, 1000.0, np.float32)
    fold = dict(fold, params=dict(fold["params"], thickness=0.004))
    v = np.zeros((n * n, 3), np.float32)
    v.reshape(n, n, 3)[h:, :, 1] = -0.3

    def make(kind, threshold):
        monkeypatch.setenv("CLOTHMI_SORT_CSR", threshold)
        if kind == "fold":
            b = ClothBody.from_scene(fold)
            b.set_state(fold["xyz"], v)
        elif kind == "crumpled":
            b = ClothBody.from_scene(configs.config_scene(4, 40))
        else:
            b = ClothBody.batch(configs.batched_scene(2, count=8))
        return b

    for kind in ("fold", "crumpled", "batch"):
        a, b = make(kind, "1"), make(kind, "0")
        contacts = 0
        for s in range(4):
            print(f"[sorted csr] {kind} substep {s}", flush=True)
            a.fixed_update(1)
            b.fixed_update(1)
            vf, ee = a.contacts()
            contacts += len(vf) + len(ee)
            for name in ("x", "p", "v"):
                assert np.array_equal(a.buffer(name), b.buffer(name)), (kind, s, name)
        if kind == "fold":
            assert contacts > 0
        a.close()
        b.close()
