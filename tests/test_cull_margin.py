"""The ray-tracing kernel's early rejection by cull radius is exact: on rays
aimed at points just inside and just outside the edges and corners of the
benchmark hall's triangles (the cases where rounding decides), from near and
far, head-on and grazing, no ray/triangle pair that TriHitted accepts
(raycast.cpp:283-313, float arithmetic with the reference compiler's fused
multiply-adds) is rejected by the margin test of test_entry in
hw-radiosity_amd/csrc/device.hip, for every reciprocal estimate of t within
its 2^-21 error bound.  The float operations are emulated with numpy (a fused
multiply-add as one rounding of the exact product plus the addend)."""
import numpy as np

from radgpu import scenes

F32 = np.float32


def fma(a, b, c):
    return (a.astype(np.float64) * b.astype(np.float64) + c.astype(np.float64)).astype(F32)


def dot3(x, y):
    """fma(x2, y2, fma(x0, y0, x1 * y1)): the compiler's contraction of a dot product."""
    return fma(x[..., 2], y[..., 2], fma(x[..., 0], y[..., 0], (x[..., 1] * y[..., 1]).astype(F32)))


def smallest_float_at_least(x):
    f = np.float32(x)
    while float(f) < x:
        f = np.nextafter(f, F32(np.inf))
    return f


def test_cull_never_rejects_an_accepted_hit():
    import radgpu
    s = scenes.hall_scene(2, n_boxes=60)
    with radgpu.Engine("-quiet", "1", "-triangle", "1000000", "-octree", "host") as e:
        e.set_scene(s)                      # ReadTriangle's records: every one consistent with its vertices
        rec, _ = e.records()
    rng = np.random.default_rng(7)
    tri = rng.choice(len(rec), 4000, replace=False)
    P = rec["p"][tri].astype(F32)           # (T, 3, 3)
    c, n, se = rec["c"][tri], rec["n"][tri], rec["se"][tri]
    r64 = np.sqrt(((P.astype(np.float64) - c[:, None, :].astype(np.float64)) ** 2).sum(-1)).max(-1)
    r = np.nextafter(r64.astype(F32), F32(np.inf))
    CL = np.nextafter(F32(np.abs(rec["c"].astype(np.float64)).sum(-1).max()), F32(np.inf))
    # targets: barycentric points at and around the edges and corners
    eps = np.array([0.0, 1e-7, 1e-6, 1e-5, 1e-4, 1e-3], np.float64)
    T = len(tri)
    rows = []
    for e0 in eps:
        for sgn in (-1.0, 1.0):
            for kind in range(3):
                b = np.zeros((T, 3))
                if kind == 0:       # near edge 0-1
                    w = rng.random(T)
                    b[:, 0], b[:, 1], b[:, 2] = w, 1 - w, sgn * e0
                elif kind == 1:     # near vertex 2
                    b[:, 2] = 1 - sgn * e0 * 2
                    b[:, 0] = b[:, 1] = sgn * e0
                else:               # near edge 2-0, random side
                    w = rng.random(T)
                    b[:, 2], b[:, 0], b[:, 1] = w, 1 - w, sgn * e0
                rows.append(b)
    B = np.concatenate(rows)                           # (K*T, 3)
    idx = np.tile(np.arange(T), len(rows))
    Pd = P[idx].astype(np.float64)
    target = (B[:, :, None] * Pd).sum(1)
    m = len(idx)
    # origins: near and far, head-on and grazing (close to the triangle's plane)
    dist = 10.0 ** rng.uniform(-1, 2, m)
    dirs = rng.normal(size=(m, 3))
    graze = rng.random(m) < 0.3
    nn = n[idx].astype(np.float64)
    dirs[graze] -= (dirs[graze] * nn[graze]).sum(1)[:, None] * nn[graze] * (1 - 1e-3 * rng.random(graze.sum()))[:, None]
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    p = (target + dirs * dist[:, None]).astype(F32)
    scale = rng.uniform(0.5, 2.0, m)                   # the target at t = 1 / scale
    v = ((target - p.astype(np.float64)) * scale[:, None]).astype(F32)
    ci, ni, sei, ri = c[idx], n[idx], se[idx], r[idx]
    # TriHitted
    a1 = dot3(v, ni)
    a2 = dot3((ci - p).astype(F32), ni)
    with np.errstate(all="ignore"):
        tt = (a2 / a1).astype(F32)
        h = fma(tt[:, None], v, p)
        side = np.stack([dot3(h, sei[:, k, :3]) + sei[:, k, 3] for k in range(3)], 1)
        thr = smallest_float_at_least(1e-4)
        accepted = (tt >= thr) & (side >= 0).all(1)
        ck = ((np.abs(v).sum(1)) * F32(2.0 ** -11)).astype(F32)
        cb = ((np.abs(p).sum(1) + CL) * F32(2.0 ** -12)).astype(F32)
        culled_any = np.zeros(m, bool)
        for rel in (-2.0 ** -21, 0.0, 2.0 ** -21):
            ta = (tt.astype(np.float64) * (1 + rel)).astype(F32)
            d = (fma(ta[:, None], v, p) - ci).astype(F32)
            d2 = fma(d[:, 2], d[:, 2], fma(d[:, 1], d[:, 1], (d[:, 0] * d[:, 0]).astype(F32)))
            lim = fma(ri, np.full(m, 1 + 2.0 ** -11, F32), fma(np.abs(ta), ck, cb))
            culled_any |= d2 > (lim * lim).astype(F32)
    assert accepted.sum() > m // 4                       # the inside targets are hit
    bad = np.nonzero(accepted & culled_any)[0]
    assert len(bad) == 0, "accepted hits the cull would reject: %s" % bad[:5]
    # and it does reject: most targets placed outside by 1e-3 of the triangle
    outside = ~accepted & np.isfinite(tt) & (tt >= thr)
    assert culled_any[outside].mean() > 0.0
