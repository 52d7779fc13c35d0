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
