"""Pins the oracle's CCD narrow phase (Physics/CudaCubicSolver.cu,
Collision/CudaPrimitiveTests.cu:252-425) with closed forms: a vertex approaching
a fixed plane linearly, a vertex above a triangle centroid, crossing edges and
parallel edges (the degenerate-midpoint branch)."""
import numpy as np
import pytest


@pytest.mark.parametrize("h,u", [(0.3, 1.0), (0.05, 0.5), (0.9, 1.0)])
def test_coplanarity_time_linear_approach(oracle_mod, h, u):
    x = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0.2, 0.2, h]], np.float32)
    v = np.array([[0, 0, 0], [0, 0, 0], [0, 0, 0], [0, 0, -u]], np.float32)
    t = oracle_mod.coplanarity_times(x, v)
    assert len(t) == 1
    assert abs(t[0] - h / u) < 1e-5


def test_coplanarity_no_root_when_receding(oracle_mod):
    x = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0.2, 0.2, 0.3]], np.float32)
    v = np.array([[0, 0, 0], [0, 0, 0], [0, 0, 0], [0, 0, 0.5]], np.float32)
    assert len(oracle_mod.coplanarity_times(x, v)) == 0


def test_coplanarity_starts_coplanar(oracle_mod):
    x = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0.2, 0.2, 0.0]], np.float32)
    v = np.array([[0, 0, 0], [0, 0, 0], [0, 0, 0], [0, 0, 0.5]], np.float32)
    t = oracle_mod.coplanarity_times(x, v)
    assert len(t) >= 1 and t[0] == 0.0


def test_vf_centroid(oracle_mod):
    th = 0.01
    tri = [[1, 0, 0], [0, 1, 0], [0, 0, 0]]
    hit, c = oracle_mod.test_vf(np.array([[1 / 3, 1 / 3, 0.005]] + tri, np.float32), th)
    assert hit
    np.testing.assert_allclose(c["w"], [1, -1 / 3, -1 / 3, -1 / 3], atol=1e-6)
    np.testing.assert_allclose(c["n"], [0, 0, 1], atol=1e-6)
    hit, _ = oracle_mod.test_vf(np.array([[1 / 3, 1 / 3, 0.02]] + tri, np.float32), th)
    assert not hit
    # outside the triangle beyond the 1+delta slack
    hit, _ = oracle_mod.test_vf(np.array([[1.0, 1.0, 0.0]] + tri, np.float32), th)
    assert not hit


def test_ee_crossing(oracle_mod):
    x = np.array([[-1, 0, 0], [1, 0, 0], [0, -1, 0.004], [0, 1, 0.004]], np.float32)
    hit, c = oracle_mod.test_ee(x, 0.005)
    assert hit
    np.testing.assert_allclose(c["w"], [0.5, 0.5, -0.5, -0.5], atol=1e-6)
    np.testing.assert_allclose(c["n"], [0, 0, -1], atol=1e-6)
    hit, _ = oracle_mod.test_ee(x, 0.003)
    assert not hit


def test_ee_parallel_degenerate(oracle_mod):
    x = np.array([[0, 0, 0], [1, 0, 0], [0, 0.002, 0], [1, 0.002, 0]], np.float32)
    hit, c = oracle_mod.test_ee(x, 0.005)
    assert hit
    np.testing.assert_allclose(c["w"], [0.5, 0.5, 0.5, 0.5])
    np.testing.assert_allclose(c["n"], [0, -1, 0], atol=1e-6)
