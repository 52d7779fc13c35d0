"""Seeded synthetic scenes in the reference's ``.tri`` format.

The ``.tri`` format is the one ``ReadTriangle`` parses (rad.cpp:47-111): per
triangle the word ``Triangle``, three ``Frgb`` ints (reflectance, rho =
Frgb * 0.5 / 255, shade.cpp:9 and shade.cpp:246), three ``Brgb`` ints (a
triangle is a light iff one channel equals 255, rad.h:47-49) and three vertices.

Every coordinate is computed with IEEE +, -, *, / in float64 and rounded once
to float32 (no libm transcendentals: a random yaw uses the rational
parametrisation cos = (1-u^2)/(1+u^2), sin = 2u/(1+u^2)), so a seed gives the
same bits on every host.  Grid vertices of one face are computed once and
shared, so coplanar neighbours match exactly, as the neighbour search in
InitRadiosity requires (rad.cpp:303-322).

Scenes
  room_scene(seed)  closed 10 x 8 x 6 room, 6 gridded walls, 2 axis-aligned
                    boxes (sides U[1,3]) with one quad per face (so some
                    triangles have area >= 1 and pre-partition has work),
                    one emissive ceiling quad (Brgb 255 255 255), per-face
                    reflectance U[0.2,0.8]; about 2k triangles.
  hall_scene(seed)  100 x 100 x 30 hall, 400 yawed boxes (sides U[1,8]),
                    8 emissive ceiling quads with emission U[0.5,2] RGB,
                    reflectance U[0.1,0.9]; about 200k triangles.
  patch_room(seed, n)  the room_scene geometry (same walls, boxes,
                    reflectances and light rectangle) triangulated to exactly
                    n patches, for the explicit form-factor matrix.
"""
from __future__ import annotations

import dataclasses
import gzip
import io

import numpy as np

MASK64 = (1 << 64) - 1


class SplitMix64:
    """splitmix64 (Steele, Lea, Flood 2014) -- a tiny, portable PRNG."""

    def __init__(self, seed: int):
        self.state = seed & MASK64

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self, lo: float = 0.0, hi: float = 1.0) -> float:
        return lo + (hi - lo) * ((self.next_u64() >> 11) * (1.0 / (1 << 53)))


@dataclasses.dataclass
class Scene:
    p: np.ndarray       # float32 [N, 3, 3] vertices
    frgb: np.ndarray    # int32 [N, 3]
    brgb: np.ndarray    # int32 [N, 3]

    @property
    def n(self) -> int:
        return int(self.p.shape[0])


class _Builder:
    def __init__(self):
        self.p, self.f, self.b = [], [], []

    def grid(self, origin, du, dv, nu, nv, frgb, brgb_of=None, flip=False):
        """Triangulate the parallelogram origin + s*du + t*dv, s,t in [0,1],
        into nu x nv cells of two right triangles.  The normal
        (p1-p0) x (p2-p0) is du x dv, or its opposite with ``flip``.
        Triangles come cell by cell (a outer, c inner), two per cell."""
        o = np.asarray(origin, np.float64)
        du = np.asarray(du, np.float64)
        dv = np.asarray(dv, np.float64)
        i = np.arange(nu + 1, dtype=np.float64)[:, None, None]
        j = np.arange(nv + 1, dtype=np.float64)[None, :, None]
        verts = (o + du * (i / nu) + dv * (j / nv)).astype(np.float32)
        v00, v10 = verts[:-1, :-1], verts[1:, :-1]
        v01, v11 = verts[:-1, 1:], verts[1:, 1:]
        t1 = np.stack([v00, v10, v11], axis=2)  # (nu, nv, 3, 3)
        t2 = np.stack([v00, v11, v01], axis=2)
        tris = np.stack([t1, t2], axis=2).reshape(-1, 3, 3)
        if flip:
            tris = tris[:, [0, 2, 1]]
        self.p.append(tris)
        self.f.append(np.tile(np.asarray(frgb, np.int32), (len(tris), 1)))
        if brgb_of is None:
            self.b.append(np.zeros((len(tris), 3), np.int32))
        else:
            br = np.array([brgb_of(a, c) for a in range(nu) for c in range(nv)], np.int32).reshape(-1, 3)
            self.b.append(np.repeat(br, 2, axis=0))

    def scene(self) -> Scene:
        return Scene(np.concatenate(self.p).astype(np.float32),
                     np.concatenate(self.f).astype(np.int32).reshape(-1, 3),
                     np.concatenate(self.b).astype(np.int32).reshape(-1, 3))


def _frgb(rng: SplitMix64, lo: float, hi: float):
    # rho = Frgb * 0.5 / 255  =>  Frgb = round(510 * rho)
    return tuple(int(510.0 * rng.uniform(lo, hi) + 0.5) for _ in range(3))


def _cells(length: float, h: float) -> int:
    return max(1, int(length / h + 0.5))


def _room_walls(bld, rng, X, Y, Z, h, rho, light_rect=None, light_brgb=None):
    """Six inward-facing walls of [0,X]x[0,Y]x[0,Z]; returns nothing.
    light_rect: list of (i0, i1, j0, j1) cell ranges on the ceiling grid."""
    nx, ny, nz = _cells(X, h), _cells(Y, h), _cells(Z, h)
    # floor z=0, normal +z: du=x, dv=y
    bld.grid((0, 0, 0), (X, 0, 0), (0, Y, 0), nx, ny, _frgb(rng, *rho))
    # ceiling z=Z, normal -z
    def ceil_b(a, c):
        for k, (i0, i1, j0, j1) in enumerate(light_rect or ()):
            if i0 <= a < i1 and j0 <= c < j1:
                return light_brgb[k]
        return (0, 0, 0)
    bld.grid((0, 0, Z), (X, 0, 0), (0, Y, 0), nx, ny, _frgb(rng, *rho),
             brgb_of=ceil_b, flip=True)
    # x=0 wall normal +x: du=y, dv=z
    bld.grid((0, 0, 0), (0, Y, 0), (0, 0, Z), ny, nz, _frgb(rng, *rho))
    bld.grid((X, 0, 0), (0, Y, 0), (0, 0, Z), ny, nz, _frgb(rng, *rho), flip=True)
    # y=0 wall normal +y: du=z, dv=x
    bld.grid((0, 0, 0), (0, 0, Z), (X, 0, 0), nz, nx, _frgb(rng, *rho))
    bld.grid((0, Y, 0), (0, 0, Z), (X, 0, 0), nz, nx, _frgb(rng, *rho), flip=True)
    return nx, ny


def _box(bld, rng, cx, cy, a, b, c, cs, sn, h, rho):
    """A box standing on z=0, footprint a x b rotated by (cs, sn) about its
    centre, height c; 4 side faces and a top, outward normals, faces gridded
    with cell size about h (h=None: one cell per face)."""
    corners = []
    for dx, dy in ((-a / 2, -b / 2), (a / 2, -b / 2), (a / 2, b / 2), (-a / 2, b / 2)):
        corners.append((cx + cs * dx - sn * dy, cy + sn * dx + cs * dy, 0.0))
    for k in range(4):
        A, B = corners[k], corners[(k + 1) % 4]
        du = (B[0] - A[0], B[1] - A[1], 0.0)
        side = a if k % 2 == 0 else b
        nu = 1 if h is None else _cells(side, h)
        nv = 1 if h is None else _cells(c, h)
        # du x (0,0,c) points outward for counter-clockwise corners
        bld.grid(A, du, (0, 0, c), nu, nv, _frgb(rng, *rho))
    A = (corners[0][0], corners[0][1], c)
    du = (corners[1][0] - corners[0][0], corners[1][1] - corners[0][1], 0.0)
    dv = (corners[3][0] - corners[0][0], corners[3][1] - corners[0][1], 0.0)
    nu = 1 if h is None else _cells(a, h)
    nv = 1 if h is None else _cells(b, h)
    bld.grid(A, du, dv, nu, nv, _frgb(rng, *rho))


def room_scene(seed: int = 1) -> Scene:
    """Closed 10 x 8 x 6 room with two boxes and one ceiling light quad."""
    rng = SplitMix64(seed)
    bld = _Builder()
    h = 0.625
    nx, ny = _cells(10, h), _cells(8, h)
    i0, j0 = nx // 2 - 2, ny // 2 - 1
    _room_walls(bld, rng, 10.0, 8.0, 6.0, h, (0.2, 0.8),
                light_rect=[(i0, i0 + 4, j0, j0 + 3)], light_brgb=[(255, 255, 255)])
    for _ in range(2):
        a, b, c = rng.uniform(1, 3), rng.uniform(1, 3), rng.uniform(1, 3)
        cx, cy = rng.uniform(1 + a / 2, 9 - a / 2), rng.uniform(1 + b / 2, 7 - b / 2)
        _box(bld, rng, cx, cy, a, b, c, 1.0, 0.0, None, (0.2, 0.8))
    return bld.scene()


def v3_room(seed: int = 4, h: float = 1.25) -> Scene:
    """A 10 x 8 x 6 room gridded coarsely (cells of about h, so most
    triangles have every edge >= 1 and area < 1: pre-partition leaves them
    whole, and the Delta-ff splitting of the reference's V3 variant, which
    needs squared edges >= 1, can split them), two boxes gridded alike as
    occluders, and a coloured ceiling light (255, 200, 150) whose vertices,
    not saturated, are receivers too."""
    rng = SplitMix64(seed)
    bld = _Builder()
    nx, ny = _cells(10, h), _cells(8, h)
    i0, j0 = nx // 2 - 1, ny // 2 - 1
    _room_walls(bld, rng, 10.0, 8.0, 6.0, h, (0.2, 0.8),
                light_rect=[(i0, i0 + 2, j0, j0 + 2)], light_brgb=[(255, 200, 150)])
    for _ in range(2):
        a, b, c = rng.uniform(1.5, 3), rng.uniform(1.5, 3), rng.uniform(1.5, 3)
        cx, cy = rng.uniform(1 + a / 2, 9 - a / 2), rng.uniform(1 + b / 2, 7 - b / 2)
        _box(bld, rng, cx, cy, a, b, c, 1.0, 0.0, h, (0.2, 0.8))
    return bld.scene()


def _room_cells(h: float, boxes) -> int:
    """Grid cells of the room walls and box faces at cell size h."""
    nx, ny, nz = _cells(10.0, h), _cells(8.0, h), _cells(6.0, h)
    t = 2 * (nx * ny + ny * nz + nz * nx)
    for a, b, c, _, _ in boxes:
        t += 2 * _cells(a, h) * _cells(c, h) + 2 * _cells(b, h) * _cells(c, h) + _cells(a, h) * _cells(b, h)
    return t


def patch_room(seed: int = 1, n: int = 32768) -> Scene:
    """room_scene(seed) re-triangulated to exactly n patches.

    Walls and box faces are gridded with the largest common cell size whose
    triangle count 2 * cells stays <= n; then the largest triangle (ties: the
    lowest index) is bisected at the midpoint of its longest edge until there
    are n.  Ceiling cells whose centre lies in room_scene's light rectangle
    (x in [3.75, 6.25], y in [40/13, 64/13]) emit Brgb 255 255 255 (at very
    coarse grids, n below a few hundred, no centre may fall inside it)."""
    if n < 64:
        raise ValueError("patch_room needs n >= 64")
    rng = SplitMix64(seed)
    rho = (0.2, 0.8)
    wall_f = [_frgb(rng, *rho) for _ in range(6)]
    boxes, box_f = [], []
    for _ in range(2):
        a, b, c = rng.uniform(1, 3), rng.uniform(1, 3), rng.uniform(1, 3)
        cx, cy = rng.uniform(1 + a / 2, 9 - a / 2), rng.uniform(1 + b / 2, 7 - b / 2)
        boxes.append((a, b, c, cx, cy))
        box_f.append([_frgb(rng, *rho) for _ in range(5)])
    lo, hi = 0.001, 10.0
    for _ in range(200):  # largest h with 2 * cells(h) <= n
        mid = 0.5 * (lo + hi)
        if 2 * _room_cells(mid, boxes) <= n:
            hi = mid
        else:
            lo = mid
    h = hi
    bld = _Builder()

    class _Feed:  # replays the drawn reflectances in room_scene's order
        def __init__(self, seq):
            self.seq = list(seq)

        def uniform(self, lo_, hi_):
            return self.seq.pop(0)

    X, Y, Z = 10.0, 8.0, 6.0
    nx, ny = _cells(X, h), _cells(Y, h)
    lx0, lx1, ly0, ly1 = 3.75, 6.25, 40.0 / 13.0, 64.0 / 13.0

    def ceil_b(a, c):
        x, y = (a + 0.5) * X / nx, (c + 0.5) * Y / ny
        return (255, 255, 255) if lx0 <= x <= lx1 and ly0 <= y <= ly1 else (0, 0, 0)

    # the wall reflectances were drawn already: feed them back as Frgb
    def inv(f):
        return [(v - 0.5 + 0.25) / 510.0 for v in f]  # round(510 * u + 0.5) gives f back

    feed = _Feed([u for f in wall_f for u in inv(f)])
    _room_walls(bld, feed, X, Y, Z, h, (0.0, 1.0))
    # the ceiling (the builder's second face) emits where its cells are lit
    bld.b[1] = np.repeat(np.array([ceil_b(a, c) for a in range(nx) for c in range(ny)], np.int32), 2, axis=0)
    for (a, b, c, cx, cy), fs in zip(boxes, box_f):
        _box(bld, _Feed([u for f in fs for u in inv(f)]), cx, cy, a, b, c, 1.0, 0.0, h, (0.0, 1.0))
    sc = bld.scene()
    assert sc.n <= n and sc.n == 2 * _room_cells(h, boxes), (sc.n, n)
    # bisect the largest triangles up to exactly n
    import heapq
    import math
    p = [tuple(tuple(float(x) for x in v) for v in t) for t in sc.p]
    f, b = [tuple(x) for x in sc.frgb.tolist()], [tuple(x) for x in sc.brgb.tolist()]

    def sub(u, v):
        return (u[0] - v[0], u[1] - v[1], u[2] - v[2])

    def area(t):
        a, c = sub(t[1], t[0]), sub(t[2], t[0])
        x = (a[1] * c[2] - a[2] * c[1], a[2] * c[0] - a[0] * c[2], a[0] * c[1] - a[1] * c[0])
        return 0.5 * math.sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2])

    heap = [(-area(t), k) for k, t in enumerate(p)]
    heapq.heapify(heap)
    while len(p) < n:
        _, k = heapq.heappop(heap)
        t = p[k]
        e = []
        for v in range(3):
            d = sub(t[(v + 1) % 3], t[v])
            e.append(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])
        v = e.index(max(e))
        a, bb, c = t[v], t[(v + 1) % 3], t[(v + 2) % 3]
        m = tuple(float(np.float32((a[q] + bb[q]) * 0.5)) for q in range(3))
        t1, t2 = (a, m, c), (m, bb, c)
        p[k] = t1
        p.append(t2)
        f.append(f[k])
        b.append(b[k])
        heapq.heappush(heap, (-area(t1), k))
        heapq.heappush(heap, (-area(t2), len(p) - 1))
    return Scene(np.asarray(p, np.float64).astype(np.float32), np.asarray(f, np.int32), np.asarray(b, np.int32))


def emission_reflectance(frgb, brgb, light: float = 1.0):
    """Right-hand side and reflectance of the matrix system for a .tri scene:
    rho = Frgb * 0.5 / 255 (shade.cpp:246); a triangle emits iff a Brgb
    channel is 255 (rad.h:47-49), with E = light * Brgb / 255.  (n, 3) each."""
    frgb = np.asarray(frgb, np.float64)
    brgb = np.asarray(brgb, np.float64)
    rho = frgb * 0.5 / 255.0
    emits = (brgb == 255).any(axis=1)
    return np.where(emits[:, None], light * brgb / 255.0, 0.0), rho


def lighting_vectors(rho, count: int = 64, lights: int = 8, seed: int = 7, lo: float = 0.5, hi: float = 2.0):
    """count emission vectors x RGB = 3 * count right-hand sides: vector k
    lights `lights` patches drawn without replacement (seeded), each with
    emission U[lo, hi] per channel.  Returns (E, rho) of shape
    (n, 3 * count); column 3k + c is channel c of vector k and carries
    rho[:, c]."""
    rho = np.asarray(rho, np.float64)
    n = rho.shape[0]
    rng = SplitMix64(seed)
    E = np.zeros((n, 3 * count))
    for k in range(count):
        chosen = []
        while len(chosen) < lights:
            j = int(rng.next_u64() % n)
            if j not in chosen:
                chosen.append(j)
        for j in chosen:
            for c in range(3):
                E[j, 3 * k + c] = rng.uniform(lo, hi)
    return E, np.tile(rho, (1, count))


def hidden_reflectance(n: int, seed: int, lo: float = 0.2, hi: float = 0.8):
    """A seeded per-patch RGB reflectance U[lo, hi], shape (n, 3)."""
    rng = SplitMix64(seed)
    return np.array([[rng.uniform(lo, hi) for _ in range(3)] for _ in range(n)])


def batch_sizes(count: int = 256, lo: int = 500, hi: int = 8000, seed: int = 4):
    """count scene sizes drawn log-uniform in [lo, hi] (seeded)."""
    rng = SplitMix64(seed)
    import math
    return [int(round(math.exp(rng.uniform(math.log(lo), math.log(hi))))) for _ in range(count)]


def hall_scene(seed: int = 2, n_boxes: int = 400, box_h: float = 0.65,
               wall_h: float = 2.0) -> Scene:
    """100 x 100 x 30 hall with ``n_boxes`` yawed boxes and 8 light quads."""
    rng = SplitMix64(seed)
    bld = _Builder()
    X, Y, Z = 100.0, 100.0, 30.0
    nx, ny = _cells(X, wall_h), _cells(Y, wall_h)
    rects, brgbs = [], []
    for k in range(8):
        # 2 x 4 layout of 2x2-cell quads, away from the walls
        gi, gj = k % 4, k // 4
        i0 = int((gi + 0.5) * nx / 4) - 1
        j0 = int((gj + 0.5) * ny / 2) - 1
        rects.append((i0, i0 + 2, j0, j0 + 2))
        e = [rng.uniform(0.5, 2.0) for _ in range(3)]
        # a light needs one channel equal to 255 (rad.h:47-49): R = 255 and
        # G, B keep the emission ratios; the absolute magnitude is lost
        brgbs.append((255, int(255.0 * e[1] / e[0] + 0.5), int(255.0 * e[2] / e[0] + 0.5)))
    _room_walls(bld, rng, X, Y, Z, wall_h, (0.1, 0.9), light_rect=rects, light_brgb=brgbs)
    for _ in range(n_boxes):
        a, b, c = rng.uniform(1, 8), rng.uniform(1, 8), rng.uniform(1, 8)
        u = rng.uniform(-1.0, 1.0)
        cs, sn = (1.0 - u * u) / (1.0 + u * u), 2.0 * u / (1.0 + u * u)
        cx, cy = rng.uniform(8.0, X - 8.0), rng.uniform(8.0, Y - 8.0)
        _box(bld, rng, cx, cy, a, b, c, cs, sn, box_h, (0.1, 0.9))
    return bld.scene()


def format_tri(scene: Scene) -> str:
    """Text in the .tri layout; '%.9g' round-trips every float32 exactly
    through fscanf("%f") (rad.cpp:66-73)."""
    out = io.StringIO()
    p = scene.p.astype(np.float64)
    for k in range(scene.n):
        f, b = scene.frgb[k], scene.brgb[k]
        out.write("Triangle\n%d %d %d %d %d %d\n" % (f[0], f[1], f[2], b[0], b[1], b[2]))
        for v in range(3):
            out.write("%.9g %.9g %.9g\n" % (p[k, v, 0], p[k, v, 1], p[k, v, 2]))
    return out.getvalue()


def write_tri(path: str, scene: Scene) -> None:
    data = format_tri(scene)
    if path.endswith(".gz"):
        with gzip.open(path, "wt") as fh:
            fh.write(data)
    else:
        with open(path, "w") as fh:
            fh.write(data)


def read_tri(path: str) -> Scene:
    """Parse a .tri file the way ReadTriangle does: a keyword, 6 ints, 9 floats
    per triangle, whitespace-separated (rad.cpp:59-75)."""
    opener = gzip.open if path.endswith(".gz") else open
    with opener(path, "rt") as fh:
        toks = fh.read().split()
    n = len(toks) // 16
    p = np.empty((n, 3, 3), np.float32)
    f = np.empty((n, 3), np.int32)
    b = np.empty((n, 3), np.int32)
    for k in range(n):
        t = toks[16 * k: 16 * k + 16]
        f[k] = [int(x) for x in t[1:4]]
        b[k] = [int(x) for x in t[4:7]]
        p[k] = np.asarray([float(x) for x in t[7:16]], np.float64).astype(np.float32).reshape(3, 3)
    return Scene(p, f, b)
