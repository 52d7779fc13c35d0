"""GPU parity of the explicit-matrix solvers (rad_lin_* through the C ABI).

The matrix built on the device equals the oracle's rows bit for bit; the
MFMA products agree with an FP64 product of the same rounded operands to
FP32-accumulation accuracy (which pins the MFMA lane layout); the
refined gather meets its FP64 residual and agrees with the FP64 dense-LU
oracle; plain shooting equals the oracle's float32 restatement bit for bit;
the adjoint agrees with the oracle's FP64 adjoint and with FP64 central
finite differences; full-size systems are checked through their FP64
residual, recomputed independently on the host."""
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


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _random_system(n, seed, cols, scale=0.9):
    """A float32 F with zero diagonal and rows summing to `scale`, rho in
    [0.2, 0.8], sparse emission with row 0 lit in every column."""
    rng = np.random.default_rng(seed)
    F = (rng.random((n, n)) * (rng.random((n, n)) < 0.7)).astype(np.float32)
    np.fill_diagonal(F, 0.0)
    F *= np.float32(scale) / F.sum(axis=1, keepdims=True).astype(np.float32)
    rho = rng.uniform(0.2, 0.8, (n, cols))
    E = np.where(rng.random((n, cols)) < 0.1, rng.uniform(0.5, 2.0, (n, cols)), 0.0)
    E[0] = 1.0
    return F, rho, E


def scene_system(n, seed=1, vis="segment"):
    """(engine, LinearSystem, scene, E, rho, area) for patch_room(seed, n)."""
    print("scene patch_room(%d, %d) vis=%s" % (seed, n, vis), flush=True)
    sc = scenes.patch_room(seed, n)
    eng = radgpu.Engine("-triangle", str(n), "-quiet", "1")
    eng.set_scene(sc).prepare_host().upload()
    lin = radgpu.LinearSystem(n)
    lin.from_scene(eng, vis)
    rec, _ = eng.records()
    assert len(rec) == n
    E, rho = lo.emission_reflectance(sc.frgb, sc.brgb)
    return eng, lin, sc, E, rho, rec["area"].astype(np.float64)


def oracle_for(sc, n):
    o = ol.Oracle("-triangle", str(n), "-quiet", "1", "-jobs", "16")
    o.load(sc)
    o.build_tree()
    o.init_radiosity()
    o.prepartition()
    return o


@pytest.fixture(scope="module")
def sys2k():
    eng, lin, sc, E, rho, area = scene_system(2000)
    F = lin.matrix()
    yield eng, lin, sc, E, rho, area, F
    lin.close()
    eng.close()


# ---- the matrix ------------------------------------------------------------

def test_matrix_matches_oracle_bit_for_bit(sys2k):
    eng, lin, sc, E, rho, area, F = sys2k
    o = oracle_for(sc, 2000)
    ref = o.ff_rows(np.arange(2000), ol.VIS_SEGMENT)
    bad = np.nonzero((F != ref).any(axis=1))[0]
    assert len(bad) == 0, "rows differ: %s" % bad[:8]
    assert 0.3 < (F > 0).mean() < 0.9
    assert np.all(F.astype(np.float64).sum(axis=1) <= 1.0 + 1e-6)


def test_matrix_reference_visibility_matches_oracle():
    n = 1200
    eng, lin, sc, E, rho, area = scene_system(n, vis="reference")
    F = lin.matrix()
    ref = oracle_for(sc, n).ff_rows(np.arange(n), ol.VIS_REFERENCE)
    assert np.array_equal(F, ref)


def test_segment_visibility_refuses_split_scene(inputs):
    with radgpu.Engine("-triangle", "20000", "-quiet", "1") as eng:
        eng.load_tri(inputs["c1_seed1.tri"]).prepare_host().upload()
        rec, _ = eng.records()
        with radgpu.LinearSystem(len(rec)) as lin:
            with pytest.raises(radgpu.RadError, match="split"):
                lin.from_scene(eng)
            lin.from_scene(eng, "reference")  # the reference walk has no such need


# ---- MFMA products -----------------------------------------------------------

def _f16_scale(F):
    m = float(np.abs(F).max())
    return 2.0 ** np.floor(np.log2(16384.0 / m))


def _to_bf16(x):
    """Round float32 to bfloat16 (nearest even), returned as float64."""
    b = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000
    return b.astype(np.uint32).view(np.float32).astype(np.float64)


@pytest.mark.parametrize("nrhs", [3, 17, 40, 100, 192])
def test_mfma_product_matches_fp64_of_rounded_operands(nrhs):
    n = 700
    print("multiply n=%d nrhs=%d" % (n, nrhs), flush=True)
    rng = np.random.default_rng(nrhs)
    F = (rng.random((n, n)) * (rng.random((n, n)) < 0.7)).astype(np.float32)
    F *= np.float32(0.9) / F.sum(axis=1, keepdims=True).astype(np.float32)
    X = rng.uniform(-1, 1, (n, nrhs))
    with radgpu.LinearSystem(n) as lin:
        lin.set_matrix(F)
        Y = lin.multiply(X, "f32")
        assert rel(Y, F.astype(np.float64) @ X) < 1e-13
        bound = np.abs(F.astype(np.float64)) @ np.abs(X)
        s = _f16_scale(F)
        F16 = (F.astype(np.float64) * s).astype(np.float16).astype(np.float64)
        X16 = X.astype(np.float32).astype(np.float16).astype(np.float64)
        for prec, Fr, Xr, sc in (("f16", F16, X16, s), ("bf16", _to_bf16(F), _to_bf16(X.astype(np.float32)), 1.0)):
            want = (Fr @ Xr) / sc
            # split-K 8 at n = 700 (ld 768): 3 k-blocks per chunk, an odd
            # count for the wide kernels' two fragment sets
            for rt, sk in ((1, 1), (2, 1), (2, 2), (4, 1), (4, 2), (1, 8), (2, 8), (4, 8)):
                try:
                    lin.set_tile(nrhs, prec, rt, sk)
                except radgpu.RadError:
                    continue
                got = lin.multiply(X, prec)
                err = np.abs(got - want) / (bound + 1e-30)
                assert err.max() < 2e-5, (prec, rt, sk, err.max())
        # the low-precision product is close to the FP32 one
        assert rel(lin.multiply(X, "f16"), Y) < 2e-3
        assert rel(lin.multiply(X, "bf16"), Y) < 1e-2


# ---- gather -----------------------------------------------------------------

@pytest.mark.parametrize("prec", ["f16", "bf16"])
def test_gather_matches_dense_lu(sys2k, prec):
    eng, lin, sc, E, rho, area, F = sys2k
    ref = lo.dense_solve(F, rho, E)
    B, info = lin.gather(rho, E, prec=prec, tol=1e-12, max_outer=60, sweeps=6, Bref=ref, target=1e-3)
    print(prec, info, flush=True)
    assert info["rel_res"] <= 1e-12
    assert rel(B, ref) < 1e-10
    assert info["err"] < 1e-10 and 0 < info["ms_target"] <= info["ms"]
    # residual recomputed in FP64 on the host
    assert np.linalg.norm(lo.residual(F, rho, E, B)) <= 1e-11 * np.linalg.norm(E)
    # the energy sums
    assert abs(info["power_E"] - float(np.sum(area[:, None] * E))) <= 1e-9 * info["power_E"]
    assert abs(info["power_B"] - float(np.sum(area[:, None] * B))) <= 1e-9 * info["power_B"]


def test_gather_192_rhs_matches_dense_lu(sys2k):
    eng, lin, sc, E, rho, area, F = sys2k
    E192, rho192 = scenes.lighting_vectors(rho, count=64, lights=8, seed=7)
    ref = lo.dense_solve(F, rho192, E192)
    B, info = lin.gather(rho192, E192, prec="f16", tol=1e-10, max_outer=60, sweeps=6)
    print(info, flush=True)
    assert rel(B, ref) < 1e-8
    # one column at a time gives the same solution
    B1, _ = lin.gather(rho192[:, 5:6], E192[:, 5:6], prec="f16", tol=1e-10, max_outer=60, sweeps=6)
    assert rel(B1[:, 0], ref[:, 5]) < 1e-8


@pytest.mark.parametrize("nrhs", [1, 3, 4, 5, 17])
def test_gather_column_counts_on_ragged_size(nrhs):
    """n = 700 (padded to 768, not a multiple of 1024) with 1-17 columns:
    the FP64 residual takes the vector-unit kernel in 256-k chunks for at most
    4 columns and the FP64 MFMA kernel above; both reach the dense solve."""
    n = 700
    F, rho, E = _random_system(n, 100 + nrhs, nrhs)
    print("gather n=%d nrhs=%d" % (n, nrhs), flush=True)
    ref = lo.dense_solve(F.astype(np.float64), rho, E)
    with radgpu.LinearSystem(n) as lin:
        lin.set_matrix(F)
        for prec in ("f16", "bf16"):
            B, info = lin.gather(rho, E, prec=prec, tol=1e-12, max_outer=60, sweeps=6)
            assert info["rel_res"] <= 1e-12, (prec, info)
            assert rel(B, ref) < 1e-10, prec
            assert np.linalg.norm(lo.residual(F.astype(np.float64), rho, E, B)) <= 1e-11 * np.linalg.norm(E)


def test_gather_rejects_bad_arguments(sys2k):
    eng, lin, sc, E, rho, area, F = sys2k
    with pytest.raises(radgpu.RadError):
        lin.gather(rho, E, prec="f32")
    with pytest.raises(radgpu.RadError):
        lin.gather(rho, E, sweeps=0)


# ---- plain shooting -------------------------------------------------------------

def test_shoot_is_bit_exact_with_float32_restatement(sys2k):
    eng, lin, sc, E, rho, area, F = sys2k
    B, info = lin.shoot(rho, E, max_shots=300)
    want, _, src = lo.shoot32(F, rho, E, area, 300)
    assert info["outer"] == len(src) == 300
    assert np.array_equal(B.astype(np.float32), want)


def test_shoot_reaches_target_against_dense_lu(sys2k):
    eng, lin, sc, E, rho, area, F = sys2k
    ref = lo.dense_solve(F, rho, E)
    B, info = lin.shoot(rho, E, Bref=ref, target=1e-3)
    print(info, flush=True)
    assert info["err"] <= 1e-3 and info["ms_target"] > 0
    assert rel(B, ref) <= 1e-3


# ---- adjoint ----------------------------------------------------------------

def test_adjoint_matches_fp64_adjoint_and_finite_differences(sys2k):
    eng, lin, sc, E, rho, area, F = sys2k
    rho_t = scenes.hidden_reflectance(len(rho), seed=11)
    Bt = lo.dense_solve(F, rho_t, E)
    B = lo.dense_solve(F, rho, E)
    g = 2.0 * (B - Bt)
    dE, drho, info = lin.adjoint(rho, B, g, tol=1e-12, max_outer=60, sweeps=6)
    print(info, flush=True)
    wE, wrho = lo.adjoint(F, rho, B, g)
    assert rel(dE, wE) < 1e-10
    assert rel(drho, wrho) < 1e-10
    rng = np.random.default_rng(5)
    idx = [(int(i), int(c)) for i, c in zip(rng.choice(len(rho), 64, replace=False), rng.integers(0, 3, 64))]
    fd = lo.fd_grad(F, rho, E, Bt, idx, h=1e-6, wrt="rho")
    got = np.array([drho[i, c] for i, c in idx])
    scale = np.abs(wrho).max()
    assert np.max(np.abs(got - fd)) <= 1e-6 * scale
    fdE = lo.fd_grad(F, rho, E, Bt, idx, h=1e-6, wrt="E")
    gotE = np.array([dE[i, c] for i, c in idx])
    assert np.max(np.abs(gotE - fdE)) <= 1e-6 * np.abs(wE).max()


def test_adjoint_through_torch_autograd(sys2k):
    import torch
    from autograd_fn import RadiositySolve
    eng, lin, sc, E, rho, area, F = sys2k
    rho_t = torch.tensor(rho, dtype=torch.float64, requires_grad=True)
    E_t = torch.tensor(E, dtype=torch.float64, requires_grad=True)
    Bt = torch.tensor(lo.dense_solve(F, scenes.hidden_reflectance(len(rho), seed=3), E))
    B = RadiositySolve.apply(lin, rho_t, E_t)
    loss = ((B - Bt) ** 2).sum()
    loss.backward()
    wE, wrho = lo.adjoint(F, rho, lo.dense_solve(F, rho, E), 2 * (B.detach().numpy() - Bt.numpy()))
    assert rel(rho_t.grad.numpy(), wrho) < 1e-9
    assert rel(E_t.grad.numpy(), wE) < 1e-9


# ---- full sizes -------------------------------------------------------------

def _host_residual(F, rho, E, B, chunk=2048):
    """||E + rho o (F B) - B|| / ||E|| in float64 on the host, by row chunks
    (independent of the device)."""
    ss = 0.0
    for r0 in range(0, F.shape[0], chunk):
        FB = F[r0:r0 + chunk].astype(np.float64) @ B
        r = E[r0:r0 + chunk] + rho[r0:r0 + chunk] * FB - B[r0:r0 + chunk]
        ss += float(np.sum(r * r))
    return np.sqrt(ss) / np.linalg.norm(E)


@pytest.mark.parametrize("n", [32768, 16384])
def test_full_size_gather_residual_and_sampled_rows(n):
    eng, lin, sc, E, rho, area = scene_system(n)
    F = lin.matrix()
    rows = np.random.default_rng(n).choice(n, 16, replace=False)
    ref = oracle_for(sc, n).ff_rows(rows, ol.VIS_SEGMENT)
    assert np.array_equal(F[rows], ref)
    nz = (F > 0).mean()
    print("n=%d nonzero %.3f" % (n, nz), flush=True)
    for prec in ("f16", "bf16"):
        B, info = lin.gather(rho, E, prec=prec, tol=1e-6, max_outer=40, sweeps=8)
        print(prec, info, flush=True)
        assert info["rel_res"] <= 1e-6
        assert _host_residual(F, rho, E, B) <= 1.5e-6
        assert abs(info["power_B"] - float(np.sum(area[:, None] * B))) <= 1e-9 * info["power_B"]
    lin.close()
    eng.close()


# ---- many scenes in one launch ----------------------------------------------

def _batch(sizes, seed0=1000):
    print("batch of %d scenes" % len(sizes), flush=True)
    b = radgpu.SceneBatch(sizes)
    E_all, rho_all, Fs = [], [], []
    for s, n in enumerate(sizes):
        sc = scenes.patch_room(seed0 + s, n)
        with radgpu.Engine("-triangle", str(n), "-quiet", "1") as e:
            e.set_scene(sc).prepare_host().upload()
            b.from_scene(s, e)
            with radgpu.LinearSystem(n) as lin:
                lin.from_scene(e)
                Fs.append(lin.matrix())
        E, rho = lo.emission_reflectance(sc.frgb, sc.brgb)
        E_all.append(E)
        rho_all.append(rho)
    return b, Fs, np.vstack(E_all), np.vstack(rho_all)


def test_batch_graph_equals_sequential_and_dense_lu():
    sizes = [700, 64, 1500, 300, 2100, 999]
    b, Fs, E, rho = _batch(sizes)
    for prec in ("f16", "bf16"):
        Bg, relg, info = b.solve(rho, E, prec=prec, graph=True, tol=1e-11, outer=12, sweeps=6)
        print(prec, "graph", info, relg, flush=True)
        Bs, rels, info_s = b.solve(rho, E, prec=prec, graph=False, tol=1e-11, outer=12, sweeps=6)
        print(prec, "sequential", info_s, flush=True)
        assert np.array_equal(Bg, Bs) and np.array_equal(relg, rels)
        assert info["above_tol"] == 0 and np.all(relg <= 1e-11)
        # replaying the cached graph gives the same answer
        Bg2, _, info2 = b.solve(rho, E, prec=prec, graph=True, tol=1e-11, outer=12, sweeps=6)
        assert info2["build_ms"] == 0 and np.array_equal(Bg, Bg2)
        off = b.offsets
        for s, F in enumerate(Fs):
            ref = lo.dense_solve(F, rho[off[s]:off[s + 1]], E[off[s]:off[s + 1]])
            assert rel(Bg[off[s]:off[s + 1]], ref) < 1e-9, (prec, s)
    b.close()


@pytest.mark.parametrize("tile", [(1, 1024, 512), (2, 4096, 64), (4, 8192, 2048)])
def test_batch_tile_schedules_agree(tile):
    """Every sweep-product schedule (rows per wave, k per tile, persistent
    grid) solves to the dense LU answer, graph and scene-by-scene alike;
    k chunks of 1024 split the 2100-patch scene three ways, 8192 not at all."""
    sizes = [700, 64, 2100, 999]
    b, Fs, E, rho = _batch(sizes, seed0=5000)
    b.set_tile(*tile)
    Bg, relg, info = b.solve(rho, E, prec="f16", graph=True, tol=1e-11, outer=12, sweeps=6)
    Bs, rels, _ = b.solve(rho, E, prec="f16", graph=False, tol=1e-11, outer=12, sweeps=6)
    print(tile, info, flush=True)
    assert np.array_equal(Bg, Bs) and info["above_tol"] == 0
    off = b.offsets
    for s, F in enumerate(Fs):
        ref = lo.dense_solve(F, rho[off[s]:off[s + 1]], E[off[s]:off[s + 1]])
        assert rel(Bg[off[s]:off[s + 1]], ref) < 1e-9, (tile, s)
    b.close()


def test_batch_autotune_and_bad_tile():
    b, Fs, E, rho = _batch([300, 1200], seed0=6000)
    for prec in ("f16", "bf16"):
        rt, kc, grid = b.autotune(prec)
        assert rt in (1, 2, 4) and kc in (1024, 2048, 4096, 8192) and 512 <= grid <= 4096
        B, r, info = b.solve(rho, E, prec=prec, graph=True, tol=1e-10, outer=12, sweeps=6)
        assert info["above_tol"] == 0
    for bad in ((3, 1024, 512), (1, 512, 512), (1, 1024, 0)):
        with pytest.raises(radgpu.RadError):
            b.set_tile(*bad)
    b.close()


def test_batch_scenes_stop_at_tolerance():
    sizes = [400, 900]
    b, Fs, E, rho = _batch(sizes, seed0=3000)
    B, r, info = b.solve(rho, E, prec="f16", graph=True, tol=1e-4, outer=10, sweeps=4)
    assert np.all(r <= 1e-4) and info["above_tol"] == 0
    B2, r2, info2 = b.solve(rho, E, prec="f16", graph=True, tol=1e-14, outer=2, sweeps=2)
    assert info2["above_tol"] == 2 and np.all(r2 > 1e-14)
    b.close()


def test_benchmark_dense_reference_matches_cpu_lu(sys2k, tmp_path):
    """radgpu.dense_ref (torch LU on the GPU, the benchmark's measuring
    stick at 32768 patches) agrees with the oracle's CPU LU."""
    import subprocess
    eng, lin, sc, E, rho, area, F = sys2k
    np.save(tmp_path / "F.npy", F)
    np.save(tmp_path / "rho.npy", rho)
    np.save(tmp_path / "E.npy", E)
    env = dict(os.environ, PYTHONPATH=os.path.join(ROOT, "hw-radiosity_amd"))
    r = subprocess.run([sys.executable, "-m", "radgpu.dense_ref", str(tmp_path)], env=env, capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    assert rel(np.load(tmp_path / "B.npy"), lo.dense_solve(F, rho, E)) < 1e-12


def test_batch_graph_recaptured_after_matrix_change():
    """A cached graph bakes in the FP16 scale of the matrices it was captured
    for; replacing the matrices by a quarter of themselves (the power-of-two
    FP16 scale moves by 2^2) must give the dense
    LU answer of the new system, not Jacobi sweeps with a stale scale."""
    sizes = [300, 500]
    rng = np.random.default_rng(41)
    Fs = []
    for n in sizes:
        F, _, _ = _random_system(n, int(rng.integers(1 << 30)), 3, scale=0.9)
        Fs.append(F)
    rho = rng.uniform(0.2, 0.8, (sum(sizes), 3))
    E = np.where(rng.random((sum(sizes), 3)) < 0.2, 1.0, 0.0)
    E[0] = E[sizes[0]] = 1.0
    b = radgpu.SceneBatch(sizes)
    off = b.offsets
    try:
        for s, F in enumerate(Fs):
            b.set_matrix(s, F)
        B1, r1, i1 = b.solve(rho, E, prec="f16", graph=True, tol=1e-11, outer=12, sweeps=6)
        assert i1["above_tol"] == 0
        F2 = [Fs[0] * np.float32(0.25), Fs[1] * np.float32(0.25)]
        for s, F in enumerate(F2):
            b.set_matrix(s, F)
        B2, r2, i2 = b.solve(rho, E, prec="f16", graph=True, tol=1e-11, outer=12, sweeps=6)
        print(i1, i2, flush=True)
        assert i2["build_ms"] > 0           # captured again
        assert i2["above_tol"] == 0 and np.all(r2 <= 1e-11)
        for s, F in enumerate(F2):
            ref = lo.dense_solve(F.astype(np.float64), rho[off[s]:off[s + 1]], E[off[s]:off[s + 1]])
            assert rel(B2[off[s]:off[s + 1]], ref) < 1e-9, s
    finally:
        b.close()
