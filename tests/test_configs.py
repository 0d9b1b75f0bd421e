"""Scene builders: grid meshes and constraint sets have the sizes the workloads
name (an n x n grid split along one diagonal has 2 (n-1)^2 triangles and
2 n (n-1) + 2 (n-1)^2 structural+shear constraints: 64^2 -> 7,938 and 16,002;
256^2 -> 130,050 and 260,610)."""
import numpy as np

from licenta_amd import configs


def test_grid_counts():
    for n, t, s in ((64, 7938, 16002), (256, 130050, 260610), (128, 32258, 64770)):
        tri = configs.grid_triangles(n, n, False)
        assert len(tri) == t
        assert len(configs.grid_distance_pairs_fast(n, n)) == s


def test_distance_pairs_fast_matches_loop():
    for r, c in ((5, 7), (16, 16), (2, 9)):
        assert np.array_equal(configs.grid_distance_pairs(r, c), configs.grid_distance_pairs_fast(r, c))


def test_triangles_consistent_orientation():
    for alt in (False, True):
        tri = configs.grid_triangles(6, 6, alt)
        i, j = np.meshgrid(np.arange(6), np.arange(6), indexing="ij")
        xy = np.stack([i.ravel(), j.ravel()], 1).astype(float)
        a, b, c = xy[tri[:, 0]], xy[tri[:, 1]], xy[tri[:, 2]]
        area2 = (b[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) - (b[:, 1] - a[:, 1]) * (c[:, 0] - a[:, 0])
        assert np.allclose(np.abs(area2), 1.0)
        # every quad split into two triangles covering area 1
        assert np.isclose(np.abs(area2).sum() / 2, 25)


def test_reference_scene_shape():
    sc = configs.reference_scene(60)
    assert sc["xyz"].shape == (3600, 3) and len(sc["tri"]) == 2 * 59 * 59
    assert sc["pinned"].sum() == 2 and sc["pinned"][0] == 1 and sc["pinned"][59 * 60] == 1
    assert np.isclose(sc["params"]["thickness"], (0.25 / 60) / 3, rtol=1e-5)
    assert (sc["xyz"][:, 2] == 0.5).all()


def test_config_scenes_deterministic():
    a = configs.config_scene(1, 32)
    b = configs.config_scene(1, 32)
    assert np.array_equal(a["xyz"], b["xyz"])
    assert np.abs(a["xyz"][:, 1] - 0.5).max() <= 1e-4 + 1e-7
    batch = configs.batched_scene(2, count=8)
    sides = [len(s["xyz"]) for s in batch]
    assert len(set(sides)) > 1
