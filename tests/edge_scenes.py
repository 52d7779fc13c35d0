"""Small edge-case scenes for the parity tests: no triangles, a lone light,
two facing triangles, no light at all, and a room with zero-area
triangles (their normals are 0/0 = NaN, whose comparisons the reference's
tests treat as passing)."""
import numpy as np

from radgpu import scenes


def _scene(p, frgb, brgb):
    return scenes.Scene(p=np.asarray(p, np.float32).reshape(-1, 3, 3),
                        frgb=np.asarray(frgb, np.int32).reshape(-1, 3),
                        brgb=np.asarray(brgb, np.int32).reshape(-1, 3))


def edge_scenes():
    tri = [[0, 0, 0], [1, 0, 0], [0, 1, 0]]
    up = [[0, 0, 1], [0, 1, 1], [1, 0, 1]]   # faces the first one
    room = scenes.room_scene(1)
    deg = np.array([[[5, 4, 3]] * 3, [[2, 2, 1], [2, 2, 1], [3, 3, 3]], [[7, 1, 2], [8, 2, 3], [9, 3, 4]]],
                   np.float32)   # a point, a segment, three collinear points
    with_deg = _scene(np.concatenate([room.p, deg]), np.concatenate([room.frgb, np.full((3, 3), 300)]),
                      np.concatenate([room.brgb, np.zeros((3, 3))]))
    return {
        "empty": _scene(np.zeros((0, 3, 3)), np.zeros((0, 3)), np.zeros((0, 3))),
        "lone_light": _scene([tri], [[200, 200, 200]], [[255, 255, 255]]),
        "facing_pair": _scene([tri, up], [[200, 100, 50], [300, 200, 100]], [[255, 255, 255], [0, 0, 0]]),
        "no_light": _scene([tri, up], [[200, 200, 200]] * 2, [[254, 0, 0], [0, 0, 0]]),
        "room_zero_area": with_deg,
    }
