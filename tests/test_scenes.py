"""Scene generator: deterministic, round-trips through .tri, and shaped as
the benchmark configurations say."""
import hashlib

import numpy as np

from radgpu import scenes


def test_c1_scene_is_pinned(goldens, tmp_path):
    path = str(tmp_path / "c1.tri")
    scenes.write_tri(path, scenes.room_scene(1))
    assert hashlib.sha256(open(path, "rb").read()).hexdigest() == goldens["generator"]["c1_seed1.tri"]


def test_tri_round_trip(tmp_path):
    s = scenes.room_scene(7)
    path = str(tmp_path / "r.tri")
    scenes.write_tri(path, s)
    t = scenes.read_tri(path)
    assert t.p.tobytes() == s.p.tobytes()
    assert np.array_equal(t.frgb, s.frgb) and np.array_equal(t.brgb, s.brgb)


def test_room_scene_shape():
    s = scenes.room_scene(1)
    assert 1800 <= s.n <= 2200
    lights = (s.brgb == 255).any(1)
    assert lights.sum() > 0 and (s.brgb[lights] == 255).all()
    rho = s.frgb / 510.0
    assert rho.min() >= 0.2 - 1e-3 and rho.max() <= 0.8 + 1e-3


def test_hall_scene_shape():
    s = scenes.hall_scene(2, n_boxes=400)
    assert 180_000 <= s.n <= 230_000
    lights = (s.brgb == 255).any(1)
    assert lights.sum() == 64
    assert (s.brgb[lights, 0] == 255).all()
    lo, hi = s.p.reshape(-1, 3).min(0), s.p.reshape(-1, 3).max(0)
    assert np.allclose(lo, 0) and np.allclose(hi, [100, 100, 30])


def test_patch_room_exact_counts_and_geometry():
    from radgpu import scenes
    import numpy as np
    room = scenes.room_scene(1)
    room_area = float(np.sum(0.5 * np.linalg.norm(np.cross(room.p[:, 1] - room.p[:, 0], room.p[:, 2] - room.p[:, 0]),
                                                  axis=1)))
    for n in (64, 500, 2000, 7777):
        sc = scenes.patch_room(1, n)
        assert sc.n == n
        area = 0.5 * np.linalg.norm(np.cross(sc.p[:, 1] - sc.p[:, 0], sc.p[:, 2] - sc.p[:, 0]), axis=1)
        assert (area > 0).all()
        # same surfaces as room_scene(1): total area and the reflectance set
        assert abs(area.sum() - room_area) < 1e-3 * room_area
        assert set(map(tuple, sc.frgb.tolist())) == set(map(tuple, room.frgb.tolist()))
        if n >= 500:  # at the coarsest grids no ceiling cell centre lies in the light rectangle
            assert (sc.brgb == 255).all(axis=1).sum() > 0
    a, b = scenes.patch_room(3, 1234), scenes.patch_room(3, 1234)
    assert np.array_equal(a.p, b.p) and np.array_equal(a.frgb, b.frgb)


def test_matrix_workload_helpers():
    from radgpu import scenes
    import numpy as np
    rho = np.full((50, 3), 0.5)
    E, r = scenes.lighting_vectors(rho, count=4, lights=8, seed=7)
    assert E.shape == (50, 12) and r.shape == (50, 12)
    assert ((E > 0).sum(axis=0) == 8).all() and ((E == 0) | ((E >= 0.5) & (E <= 2.0))).all()
    sizes = scenes.batch_sizes(256, 500, 8000, seed=4)
    assert len(sizes) == 256 and min(sizes) >= 500 and max(sizes) <= 8000
    assert sizes == scenes.batch_sizes(256, 500, 8000, seed=4)
    h = scenes.hidden_reflectance(10, seed=1)
    assert h.shape == (10, 3) and ((h >= 0.2) & (h <= 0.8)).all()
    E1, rho1 = scenes.emission_reflectance([[408, 102, 0]], [[255, 0, 255]])
    assert np.allclose(rho1, [[0.8, 0.2, 0.0]]) and np.allclose(E1, [[1.0, 0.0, 1.0]])
