"""Pins the oracle's vertex normals (Physics/CudaDeformableUtils.cu:190-249: unit
face normals cross(b-a, c-a)/|.| summed per vertex, then normalised; carried from
substep to substep; raw cross products at creation) with closed forms: a flat
grid, a two-triangle tent whose vertex normal is the bisector of the two unit face
normals whatever the areas, the tent's area-weighted creation normal, and the
carried normal after the tent's second triangle is turned over."""
import numpy as np
import pytest

from licenta_amd import configs

R2 = np.sqrt(0.5)


def _tent(flags):
    p = configs.default_params()
    p.update(gravity=(0.0, 0.0, 0.0), k_stretch=0.0, k_bend=0.0, thickness=0.001,
             solver_iterations=1, flags=flags)
    # triangle 0 (big, z = 0 plane, normal +z), triangle 1 (small, x = 0 plane, normal +x)
    xyz = np.array([[0, 0, 0], [10, 0, 0], [0, 10, 0], [0, 0.1, 0], [0, 0, 0.1]], np.float32)
    tri = np.array([[0, 1, 2], [0, 3, 4]], np.uint32)
    return dict(params=p, xyz=xyz, tri=tri, inv_mass=np.ones(5, np.float32), pinned=None,
                dist_pairs=None, k_dist=None, k_bend=None)


@pytest.mark.parametrize("prec", [32, 64])
def test_flat_grid_normals(oracle_mod, prec):
    sc = configs.config_scene(0, 8)
    xyz = sc["xyz"].reshape(-1, 3).copy()
    xyz[:, 1] = 0.25
    o = oracle_mod.OracleSim(sc, prec)
    nrm = o.normals(xyz)
    s = nrm[0, 1]
    assert abs(abs(s) - 1.0) < 1e-7
    np.testing.assert_allclose(nrm, np.tile([0.0, s, 0.0], (64, 1)), atol=1e-7)


@pytest.mark.parametrize("prec", [32, 64])
def test_tent_normals_are_area_independent(oracle_mod, prec):
    sc = _tent(0)
    o = oracle_mod.OracleSim(sc, prec)
    nrm = o.normals(sc["xyz"])
    tol = 1e-6 if prec == 32 else 1e-14
    np.testing.assert_allclose(nrm[0], [R2, 0, R2], atol=tol)
    np.testing.assert_allclose(nrm[1], [0, 0, 1], atol=tol)
    np.testing.assert_allclose(nrm[3], [1, 0, 0], atol=tol)
    # the carried normals are not touched by a query
    assert np.array_equal(o.normals(), np.zeros((5, 3)))


@pytest.mark.parametrize("prec", [32, 64])
def test_carried_normals(oracle_mod, prec):
    """Substep 1 adds the tent's unit face normals to the creation normal. Then vertex 3 moves to (0, 0, 0.1) and vertex 4
    to (0, 0.1, 0), which turns triangle 1's normal to -x; substep 2 adds the new unit
    face normals to the carried unit normal before normalising."""
    sc = _tent(configs.F_VERTEX_NORMALS)
    o = oracle_mod.OracleSim(sc, prec)
    o.step(1)
    tol = 1e-6 if prec == 32 else 1e-14
    n1 = o.normals()
    c = np.array([0.01, 0.0, 100.0])
    want1 = c / np.linalg.norm(c) + np.array([1.0, 0, 0]) + np.array([0, 0, 1.0])
    np.testing.assert_allclose(n1[0], want1 / np.linalg.norm(want1), atol=tol)
    xyz = sc["xyz"].astype(np.float64).copy()
    xyz[3], xyz[4] = [0, 0, 0.1], [0, 0.1, 0]
    o.set_state(xyz, np.zeros((5, 3)))
    o.step(1)
    n2 = o.normals()
    u1 = want1 / np.linalg.norm(want1)
    want0 = u1 + np.array([-1.0, 0.0, 0.0]) + np.array([0.0, 0.0, 1.0])
    np.testing.assert_allclose(n2[0], want0 / np.linalg.norm(want0), atol=tol)
    np.testing.assert_allclose(n2[1], [0, 0, 1], atol=tol)


@pytest.mark.parametrize("prec", [32, 64])
def test_creation_normals_are_area_weighted(oracle_mod, prec):
    """At creation the reference sums the raw cross products (twice the areas): the
    tent's shared vertex leans to the big triangle, (0.01, 0, 100) normalised."""
    o = oracle_mod.OracleSim(_tent(configs.F_VERTEX_NORMALS), prec)
    n0 = o.normals()
    want = np.array([0.01, 0.0, 100.0])
    tol = 1e-6 if prec == 32 else 1e-14
    np.testing.assert_allclose(n0[0], want / np.linalg.norm(want), atol=tol)
    np.testing.assert_allclose(n0[3], [1, 0, 0], atol=tol)
