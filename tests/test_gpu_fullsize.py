"""The configurations at their stated sizes, through the C ABI.

- hall-200k (the benchmark's shooting workload): the whole 64-shot solve at
  -converge 1200 bit-identical to the oracle, and the first 100 shots of the
  interreflection solve at -converge 175 (64 emitter shots, then 36 shots
  from receivers) bit-identical too.
- 192 right-hand sides on the 32768-patch matrix, FP16 and BF16: the device's
  FP64 residual meets the tolerance and an FP64 residual recomputed on the
  host over sampled rows agrees.
- 256 rooms of 500-8000 patches in one graph launch: every scene meets the
  tolerance, the graph result equals the scene-by-scene launches bit for bit,
  and a host FP64 residual over eight scenes spanning the size range agrees.
- The adjoint at 16384 patches: the transposed system's FP64 residual and
  dL/drho = lambda o (F B), recomputed on the host over sampled rows.
"""
import os
import sys

import numpy as np
import pytest

import oracle_lib as ol
import radgpu
from radgpu import scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import linear as lo  # noqa: E402

pytestmark = pytest.mark.gpu
THREADS = str(min(16, len(os.sched_getaffinity(0))))


def _describe(rec, ref):
    out = []
    for f in ref.dtype.names:
        if rec[f].tobytes() != ref[f].tobytes():
            bad = np.nonzero((rec[f] != ref[f]).reshape(len(ref), -1).any(1))[0]
            out.append("%s: %d slots (first %s)" % (f, len(bad), bad[:4]))
    return "; ".join(out)


@pytest.fixture(scope="module")
def hall():
    return scenes.hall_scene(2)


def _oracle(scene, opts):
    o = ol.Oracle(*(opts + ["-jobs", THREADS]))
    o.load(scene)
    o.build_tree()
    o.init_radiosity()
    o.prepartition()
    return o


def test_hall_full_solve_matches_oracle(hall):
    """Every shot of the benchmark's direct-light solve, every record."""
    opts = ["-converge", "1200", "-triangle", "1000000", "-quiet", "1"]
    print("hall-200k full solve, %d triangles" % hall.n, flush=True)
    with radgpu.Engine(*opts) as eng:
        eng.set_scene(hall).prepare_host().upload()
        shots, conv = eng.shoot()
        rec, _ = eng.records()
        cnt = eng.counters()
    o = _oracle(hall, opts)
    oshots, oconv = o.iterate(1 << 30)
    orec, _ = o.dump()
    assert conv and oconv and shots == oshots == 64
    assert cnt["ref_rays"] == int(o.counters()[ol.CNT_RAYS])
    assert rec.tobytes() == orec.tobytes(), _describe(rec, orec)


def test_hall_interreflection_first_100_shots_match_oracle(hall):
    """At -converge 175 receivers become sources after the emitters: the
    first 100 shots (36 of them from lit receivers) bit-identical."""
    opts = ["-converge", "175", "-triangle", "1000000", "-quiet", "1"]
    print("hall-200k interreflection, first 100 shots", flush=True)
    with radgpu.Engine(*opts) as eng:
        eng.set_scene(hall).prepare_host().upload()
        shots, conv = eng.shoot(100)
        rec, _ = eng.records()
    o = _oracle(hall, opts)
    oshots, _ = o.iterate(100)
    orec, _ = o.dump()
    assert shots == oshots == 100 and not conv
    lights = rec["Brgb"].max(axis=1) == 255
    assert lights.any() and (rec["deltaB"][lights].sum(axis=1) < 175).all()   # the emitters have shot
    assert rec.tobytes() == orec.tobytes(), _describe(rec, orec)


# ---- explicit matrix at full size -------------------------------------------

def _sampled_residual(F, rho, E, B, rows):
    """||E + rho o (F B) - B|| / ||E|| over the sampled rows, in FP64 on the host."""
    r = E[rows] + rho[rows] * (F[rows].astype(np.float64) @ B) - B[rows]
    return float(np.linalg.norm(r) / np.linalg.norm(E[rows]))


@pytest.fixture(scope="module")
def room32k():
    n = 32768
    print("patch room %d" % n, flush=True)
    sc = scenes.patch_room(1, n)
    eng = radgpu.Engine("-triangle", str(n), "-quiet", "1")
    eng.set_scene(sc).prepare_host().upload()
    lin = radgpu.LinearSystem(n)
    lin.from_scene(eng)
    F = lin.matrix()
    E, rho = lo.emission_reflectance(sc.frgb, sc.brgb)
    yield lin, F, E, rho
    lin.close()
    eng.close()


@pytest.mark.parametrize("prec", ["f16", "bf16"])
def test_gather_192_rhs_full_size(room32k, prec):
    """64 lighting vectors x RGB on the 32768-patch matrix in one GEMM."""
    lin, F, E, rho = room32k
    E192, rho192 = scenes.lighting_vectors(rho, count=64, lights=8, seed=7)
    B, info = lin.gather(rho192, E192, prec=prec, tol=1e-6, max_outer=40, sweeps=8)
    print(prec, info, flush=True)
    assert B.shape == (32768, 192)
    assert info["rel_res"] <= 1e-6
    rows = np.sort(np.random.default_rng(17).choice(len(F), 1024, replace=False))
    assert _sampled_residual(F, rho192, E192, B, rows) <= 2e-6
    # each column's own residual is small too (no column left behind)
    r = E192[rows] + rho192[rows] * (F[rows].astype(np.float64) @ B) - B[rows]
    col = np.linalg.norm(r, axis=0) / np.maximum(np.linalg.norm(E192[rows], axis=0), 1e-300)
    lit = np.linalg.norm(E192[rows], axis=0) > 0
    assert col[lit].max() <= 1e-5


def test_adjoint_full_size():
    """(I - diag(rho) F)^T lambda = dL/dB at 16384 patches, and
    dL/drho = lambda o (F B), recomputed on the host over sampled rows."""
    n = 16384
    print("patch room %d adjoint" % n, flush=True)
    sc = scenes.patch_room(5, n)
    with radgpu.Engine("-triangle", str(n), "-quiet", "1") as eng:
        eng.set_scene(sc).prepare_host().upload()
        with radgpu.LinearSystem(n) as lin:
            lin.from_scene(eng)
            F = lin.matrix()
            E, rho = lo.emission_reflectance(sc.frgb, sc.brgb)
            rho_t = scenes.hidden_reflectance(n, seed=11)
            Bt, _ = lin.gather(rho_t, E, prec="f16", tol=1e-10, max_outer=60, sweeps=8)
            B, _ = lin.gather(rho, E, prec="f16", tol=1e-10, max_outer=60, sweeps=8)
            g = 2.0 * (B - Bt)
            dE, drho, info = lin.adjoint(rho, B, g, tol=1e-8, max_outer=40, sweeps=8)
    print(info, flush=True)
    assert info["rel_res"] <= 1e-8
    cols = np.sort(np.random.default_rng(23).choice(n, 1024, replace=False))
    # row j of the transposed system: lambda_j - sum_i F_ij rho_i lambda_i = g_j
    FT = F[:, cols].astype(np.float64).T
    rT = g[cols] + FT @ (rho * dE) - dE[cols]
    assert np.linalg.norm(rT) <= 2e-8 * np.linalg.norm(g[cols])
    want = dE[cols] * (F[cols].astype(np.float64) @ B)
    assert np.linalg.norm(drho[cols] - want) <= 1e-12 * np.linalg.norm(want) + 1e-300


def test_batch_256_rooms_full_size():
    """Config 4b as stated: 256 rooms, sizes log-uniform in [500, 8000], all
    in one graph launch."""
    sizes = scenes.batch_sizes(256, 500, 8000, seed=4)
    order = np.argsort(sizes)
    sample = sorted(set(int(order[k]) for k in np.linspace(0, len(sizes) - 1, 8).round().astype(int)))
    print("batch of 256 rooms, %d patches; host check on scenes %s" % (sum(sizes), sample), flush=True)
    b = radgpu.SceneBatch(sizes)
    E_all, rho_all, Fs = [], [], {}
    for s, n in enumerate(sizes):
        sc = scenes.patch_room(1000 + s, n)
        with radgpu.Engine("-triangle", str(n), "-quiet", "1") as e:
            e.set_scene(sc).prepare_host().upload()
            b.from_scene(s, e)
            if s in sample:
                with radgpu.LinearSystem(n) as lin:
                    lin.from_scene(e)
                    Fs[s] = lin.matrix()
        E, rho = lo.emission_reflectance(sc.frgb, sc.brgb)
        E_all.append(E)
        rho_all.append(rho)
    E, rho = np.vstack(E_all), np.vstack(rho_all)
    off = b.offsets
    try:
        for prec in ("f16", "bf16"):
            Bg, relg, info = b.solve(rho, E, prec=prec, graph=True, tol=1e-6, outer=12, sweeps=6)
            print(prec, "graph", info, float(relg.max()), flush=True)
            assert info["above_tol"] == 0 and np.all(relg <= 1e-6)
            Bs, rels, _ = b.solve(rho, E, prec=prec, graph=False, tol=1e-6, outer=12, sweeps=6)
            assert np.array_equal(Bg, Bs) and np.array_equal(relg, rels)
            for s in sample:
                sl = slice(off[s], off[s + 1])
                r = lo.residual(Fs[s], rho[sl], E[sl], Bg[sl])
                host = float(np.linalg.norm(r) / np.linalg.norm(E[sl]))
                assert host <= 1.5e-6 and abs(host - relg[s]) <= 1e-7 + 1e-3 * relg[s], (prec, s, host, relg[s])
    finally:
        b.close()
