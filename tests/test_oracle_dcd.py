"""Pins the oracle's discrete narrow phase (CM_F_DCD: CudaPrimitiveTests::
DCDTriangleTests, LicentaV2/LicentaV2/Collision/CudaPrimitiveTests.cu:7-94) on two
triangles with closed-form answers:

* a vertex half a thickness above a triangle's interior point with barycentrics
  (1/2, 1/4, 1/4) gives exactly one VF contact, weights (1, -1/2, -1/4, -1/4),
  normal +z, time 0 (TestVF sets it); at one and a half thicknesses none;
* two edges crossing at their midpoints half a thickness apart give an EE contact
  with weights (1/2, 1/2, -1/2, -1/2) and time -1, the candidate's initial value,
  which the reference's TestEE leaves unset on its main branch (that branch
  measures the distance between x1 + a x21 and x3 + a x43, the first edge's
  parameter on both, so the crossing sits at both midpoints);
* the tests run at the substep's start positions x (the reference passes
  m_dPrevPositions, Physics/CudaDeformableBody.cpp:428): a vertex moving away at
  one and a half thicknesses per substep still gives the contact;
* the continuous narrow phase on the same static scenes finds none (the
  coplanarity function is constant and not zero), so the flag switches the path.
"""
import numpy as np

from licenta_amd import configs

TH = 0.01


def _two_triangles(b_xyz, flags):
    a = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]])
    xyz = np.vstack([a, np.asarray(b_xyz, float)]).astype(np.float32)
    tri = np.array([[0, 1, 2], [3, 4, 5]], np.uint32)
    params = configs._plain_params(1.0 / 240.0, 4, TH)
    params.update(gravity=(0.0, 0.0, 0.0), flags=flags, k_bend=0.0)
    return configs._scene(params, xyz, tri, np.full(6, 1000.0, np.float32), name="two_tris")


def _step(oracle_mod, sc, v=None):
    o = oracle_mod.OracleSim(sc, 32)
    x = sc["xyz"].astype(np.float64)
    o.set_state(x, np.zeros_like(x) if v is None else v)
    o.step(1)
    return o.contacts(False), o.contacts(True)


DCD = configs.F_SELF_COLLISION | configs.F_EE_CONTACTS | configs.F_DCD
CCD = configs.F_SELF_COLLISION | configs.F_EE_CONTACTS


def _vf_scene(d, flags):
    # vertex 3 at (1/4, 1/4, d) above triangle A; triangle B tilts away from A
    return _two_triangles([[0.25, 0.25, d], [0.6, 0.25, 0.3], [0.25, 0.6, 0.3]], flags)


def test_dcd_vf_closed_form(oracle_mod):
    vf, ee = _step(oracle_mod, _vf_scene(0.5 * TH, DCD))
    assert len(vf) == 1 and len(ee) == 0
    c = vf[0]
    assert c["v"].tolist() == [3, 0, 1, 2]
    np.testing.assert_allclose(c["w"], [1.0, -0.5, -0.25, -0.25], atol=1e-6)
    np.testing.assert_allclose(c["n"], [0.0, 0.0, 1.0], atol=1e-6)
    assert c["t"] == 0.0
    # beyond the thickness: nothing
    vf, ee = _step(oracle_mod, _vf_scene(1.5 * TH, DCD))
    assert len(vf) == 0 and len(ee) == 0
    # the continuous narrow phase on the static scene: no coplanarity root, no contact
    vf, ee = _step(oracle_mod, _vf_scene(0.5 * TH, CCD))
    assert len(vf) == 0 and len(ee) == 0


def test_dcd_tests_start_positions(oracle_mod):
    """Vertex 3 moves away from A by 1.5 thicknesses over the substep: tested at x
    (the reference's previous positions) it is in contact; at the predicted
    positions it would not be."""
    sc = _vf_scene(0.5 * TH, DCD)
    v = np.zeros((6, 3))
    v[3, 2] = 1.5 * TH / sc["params"]["timestep"]
    vf, _ = _step(oracle_mod, sc, v)
    assert len(vf) == 1 and vf[0]["v"].tolist() == [3, 0, 1, 2]
    # planted mistake: the same static test at the predicted positions finds nothing
    moved = _vf_scene(0.5 * TH + 1.5 * TH, DCD)
    vf_p, _ = _step(oracle_mod, moved)
    assert len(vf_p) == 0


def test_dcd_ee_crossing_keeps_time(oracle_mod):
    # B's edge (3, 4) runs along y at x = 1/2, d above A's edge (0, 1) on the x axis;
    # both cross at their midpoints
    d = 0.5 * TH
    b = [[0.5, -0.35, d], [0.5, 0.35, d], [0.9, -0.35, 0.3]]
    vf, ee = _step(oracle_mod, _two_triangles(b, DCD))
    sets = [frozenset(c["v"].tolist()) for c in ee]
    assert frozenset((0, 1, 3, 4)) in sets
    c = ee[sets.index(frozenset((0, 1, 3, 4)))]
    assert c["v"].tolist() == [0, 1, 3, 4]
    np.testing.assert_allclose(c["w"], [0.5, 0.5, -0.5, -0.5], atol=1e-6)
    assert c["t"] == -1.0
    np.testing.assert_allclose(np.abs(c["n"]), [0.0, 0.0, 1.0], atol=1e-6)
    # B's vertex 4 lies above A's interior as well: a VF contact at time 0
    assert any(
        cv["v"].tolist() == [4, 0, 1, 2] and cv["t"] == 0.0 for cv in vf)
    # farther than the thickness: no EE contact
    far = [[0.5, -0.35, 1.5 * TH], [0.5, 0.35, 1.5 * TH], [0.9, -0.35, 0.3]]
    _, ee = _step(oracle_mod, _two_triangles(far, DCD))
    assert frozenset((0, 1, 3, 4)) not in [frozenset(c["v"].tolist()) for c in ee]
    # the continuous narrow phase on the static scene finds no contact
    vf, ee = _step(oracle_mod, _two_triangles(b, CCD))
    assert len(vf) == 0 and len(ee) == 0
