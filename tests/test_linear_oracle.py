"""Pins the FP64 matrix oracle (oracle/linear.py) and the centroid
form-factor rows of the C oracle, on the CPU.

Every check here is independent of the formula it checks: closed forms,
brute-force re-solves and brute-force finite differences, reciprocity."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "hw-radiosity_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import linear as lo  # noqa: E402
import oracle_lib as ol  # noqa: E402


def _random_system(n, seed, cols=3, scale=0.9):
    rng = np.random.default_rng(seed)
    F = rng.random((n, n)) * (rng.random((n, n)) < 0.7)
    np.fill_diagonal(F, 0.0)
    F *= scale / F.sum(axis=1, keepdims=True)
    rho = rng.uniform(0.2, 0.8, (n, cols))
    E = np.where(rng.random((n, cols)) < 0.1, rng.uniform(0.5, 2.0, (n, cols)), 0.0)
    E[0] = 1.0
    return F, rho, E


def test_dense_solve_uniform_rank_one_closed_form():
    # F_ij = s / n for all i, j; rho constant:  B = E + rho s m 1 with
    # m = mean(B) = mean(E) / (1 - rho s)
    n, s, r = 50, 0.9, 0.6
    F = np.full((n, n), s / n)
    rng = np.random.default_rng(3)
    E = rng.random((n, 2))
    rho = np.full((n, 2), r)
    m = E.mean(axis=0) / (1 - r * s)
    np.testing.assert_allclose(lo.dense_solve(F, rho, E), E + r * s * m[None, :], rtol=1e-13)


def test_dense_solve_two_patches():
    # B1 = E1 + r1 f B2, B2 = E2 + r2 g B1
    f, g, r1, r2, e1, e2 = 0.3, 0.7, 0.5, 0.8, 1.0, 0.25
    F = np.array([[0.0, f], [g, 0.0]])
    det = 1 - r1 * f * r2 * g
    B = lo.dense_solve(F, np.array([[r1], [r2]]), np.array([[e1], [e2]]))
    np.testing.assert_allclose(B[:, 0], [(e1 + r1 * f * e2) / det, (e2 + r2 * g * e1) / det], rtol=1e-14)


def test_adjoint_matches_brute_force_differences():
    n = 7
    F, rho, E = _random_system(n, 11, cols=2)
    Bt = lo.dense_solve(F, np.full_like(rho, 0.5), E)
    B = lo.dense_solve(F, rho, E)
    dE, drho = lo.adjoint(F, rho, B, 2 * (B - Bt))
    h = 1e-6
    for i in range(n):
        for c in range(2):
            for arr, grad in ((rho, drho), (E, dE)):
                p, m = arr.copy(), arr.copy()
                p[i, c] += h
                m[i, c] -= h
                args_p = (F, p, E) if arr is rho else (F, rho, p)
                args_m = (F, m, E) if arr is rho else (F, rho, m)
                fd = (lo.loss(lo.dense_solve(*args_p), Bt) - lo.loss(lo.dense_solve(*args_m), Bt)) / (2 * h)
                assert abs(fd - grad[i, c]) <= 1e-7 * max(1.0, abs(fd)), (i, c, fd, grad[i, c])


def test_sherman_morrison_fd_matches_re_solves():
    n = 40
    F, rho, E = _random_system(n, 5)
    Bt = lo.dense_solve(F, rho * 0.9, E)
    idx = [(3, 0), (17, 2), (39, 1), (0, 0)]
    for wrt in ("rho", "E"):
        got = lo.fd_grad(F, rho, E, Bt, idx, h=1e-5, wrt=wrt)
        for k, (i, c) in enumerate(idx):
            arr = rho if wrt == "rho" else E
            vals = []
            for d in (1e-5, -1e-5):
                a = arr.copy()
                a[i, c] += d
                B = lo.dense_solve(F, a, E) if wrt == "rho" else lo.dense_solve(F, rho, a)
                vals.append(lo.loss(B, Bt))
            want = (vals[0] - vals[1]) / 2e-5
            assert abs(got[k] - want) <= 1e-6 * max(1.0, abs(want)), (wrt, i, c, got[k], want)


def test_shooting_first_shot_and_limit():
    n = 30
    F, rho, E = _random_system(n, 9)
    area = np.random.default_rng(1).uniform(0.5, 2.0, n)
    # one shot from the single largest-power source: B_i = E_i + rho_i F_ij E_j
    B1, dB1, src = lo.shoot32(F, rho, E, area, 1)
    j = int(np.argmax((E.sum(axis=1)) * area))
    assert src == [j]
    want = E + rho * F[:, j:j + 1] * E[j][None, :]
    np.testing.assert_allclose(B1, want, rtol=1e-6)
    assert np.all(dB1[j] == (rho[j] * F[j, j] * E[j]).astype(np.float32))
    # FP64 schedule converges to the dense solve
    B64, _ = lo.shoot64(F, rho, E, area, 4000)
    np.testing.assert_allclose(B64, lo.dense_solve(F, rho, E), rtol=1e-9, atol=1e-12)
    # the float32 restatement follows it to float32 accuracy
    B32, _, _ = lo.shoot32(F, rho, E, area, 600)
    B64b, _ = lo.shoot64(F, rho, E, area, 600)
    assert np.linalg.norm(B32 - B64b) <= 1e-5 * np.linalg.norm(B64b)


def test_shooting_ties_go_to_the_lowest_index():
    F = np.zeros((4, 4))
    E = np.array([[0.0], [1.0], [1.0], [0.5]])
    _, _, src = lo.shoot32(F, np.full((4, 1), 0.5), E, np.ones(4), 3)
    assert src == [1, 2, 3]


def test_emission_reflectance():
    E, rho = lo.emission_reflectance([[510, 255, 0]], [[255, 128, 0]], light=2.0)
    np.testing.assert_allclose(rho, [[1.0, 0.5, 0.0]])
    np.testing.assert_allclose(E, [[2.0, 2.0 * 128 / 255, 0.0]])
    E, _ = lo.emission_reflectance([[1, 1, 1]], [[254, 254, 254]])
    assert not E.any()


@pytest.fixture(scope="module")
def c1_small():
    from radgpu import scenes
    sc = scenes.patch_room(1, 600)
    o = ol.Oracle("-triangle", "600", "-quiet", "1", "-jobs", "8")
    o.load(sc)
    o.build_tree()
    o.init_radiosity()
    o.prepartition()
    return o, sc


def _all_pairs(n):
    r = np.arange(n)
    return np.stack(np.meshgrid(r, r, indexing="ij"), -1).reshape(-1, 2)


def test_segment_visibility_matches_brute_force(c1_small):
    """The grid walk over the complete leaf lists finds exactly the occluders
    that testing every original finds (same float test), and that test agrees
    with a float64 segment/triangle test up to pairs whose decision lies
    within 1e-5 of a triangle edge or of a segment end."""
    o, _ = c1_small
    n = o.trinum
    pairs = _all_pairs(n)
    seg = o.ff_centroid(pairs, ol.VIS_SEGMENT).reshape(n, n)
    brute = o.ff_centroid(pairs, ol.VIS_BRUTE).reshape(n, n)
    none = o.ff_centroid(pairs, ol.VIS_NONE).reshape(n, n)
    assert np.array_equal(seg, brute)
    assert np.all((seg == none) | (seg == 0))
    assert 0.3 < (seg > 0).mean() < 0.9
    rec, _ = o.dump()
    c = rec["c"].astype(np.float64)
    nn = rec["n"].astype(np.float64)
    se = rec["se"].astype(np.float64)
    differ, border = 0, 0
    for i in range(n):
        js = np.nonzero(none[i] > 0)[0]
        if len(js) == 0:
            continue
        v = c[js] - c[i]
        a1 = v @ nn.T
        a2 = np.einsum("kd,kd->k", c - c[i], nn)[None, :]
        with np.errstate(divide="ignore", invalid="ignore"):
            t = a2 / a1
            q = c[i][None, None, :] + t[..., None] * v[:, None, :]
            side = np.einsum("jkd,ksd->jks", q, se[:, :, :3]) + se[None, :, :, 3]
        ok_t = (t > 1e-4) & (t < 1 - 1e-4)
        hit = ok_t & (side >= 0).all(axis=-1)
        hit[:, i] = False
        hit[np.arange(len(js)), js] = False
        occ = hit.any(axis=1)
        near = (np.abs(side) < 1e-5).any(axis=-1) | (np.abs(t - 1e-4) < 1e-5) | (np.abs(t - (1 - 1e-4)) < 1e-5)
        amb = (near & (t > 0) & (t < 1)).any(axis=1)
        d = occ != (seg[i, js] == 0)
        differ += int((d & ~amb).sum())
        border += int((d & amb).sum())
    assert differ == 0
    assert border <= 1e-4 * n * n


def test_centroid_rows_reciprocity_and_clamp(c1_small):
    o, sc = c1_small
    n = o.trinum
    assert n == 600
    rows = np.arange(n)
    F = o.ff_rows(rows)
    raw = o.ff_centroid(_all_pairs(n)).reshape(n, n)
    assert (np.diag(raw) == 0).all() and (raw >= 0).all()
    sums = raw.astype(np.float64).sum(axis=1)
    scaled = sums > 1.0
    # rows that sum to <= 1 are the raw form factors, bit for bit
    assert np.array_equal(F[~scaled], raw[~scaled])
    # scaled rows sum to 1 within float rounding
    assert np.all(np.abs(F[scaled].astype(np.float64).sum(axis=1) - 1.0) < 1e-5)
    # reciprocity of the point-to-patch estimate: A_i F_ij ~ A_j F_ji for
    # well-separated pairs (the estimate takes the source's area and the
    # receiver's point, so it is symmetric up to its discretisation)
    rec, _ = o.dump()
    area = rec["area"].astype(np.float64)
    c = rec["c"].astype(np.float64)
    d = np.linalg.norm(c[:, None, :] - c[None, :, :], axis=-1)
    far = (d > 4 * np.sqrt(area.max())) & (raw > 0) & (raw.T > 0)
    lhs = area[:, None] * raw
    assert far.sum() > 1000
    rel = np.abs(lhs - lhs.T)[far] / lhs[far]
    assert np.median(rel) < 0.05
    # segment visibility is symmetric: wherever both directions have a
    # non-zero unoccluded value, both are visible or both are blocked
    none = o.ff_centroid(_all_pairs(n), ol.VIS_NONE).reshape(n, n)
    both = (none > 0) & (none.T > 0)
    assert both.sum() > 1000
    assert np.array_equal((raw > 0) & both, (raw.T > 0) & both)
    # the reference's RayHitted from a centroid is a different (and, on
    # this octree, mostly blind) visibility: it reaches far fewer patches
    ref = o.ff_centroid(_all_pairs(n), ol.VIS_REFERENCE).reshape(n, n)
    assert np.all((ref == raw) | (ref == 0) | (raw == 0))
