"""Profiling driver: one hall-200k solve at -converge 175 (with --direct at
-converge 1200; or, with --linear, the explicit
matrix workload: the 32768-patch room's form factors, FP16 gathers with 3
and 192 right-hand sides, 2000 plain shots; with --gemm, the same matrix and
five FP16 sweeps with 192 right-hand sides at the tile RT = 2 (the LDS DMA
kernel), split-K 2;
with --batch, 256 rooms of 500-8000 patches, the autotuned FP16 batch solve
twice in one graph launch; with --adjoint, the 16384-patch room's FP16
adjoint solve twice), then a plain 256 MiB device copy that calibrates the
FETCH_SIZE / WRITE_SIZE counters in the same run.

    rocprofv3 --pmc FETCH_SIZE -d <dir> -o p -- python3 hw-radiosity_amd/radgpu/prof_driver.py [--linear]
"""
import ctypes
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import radgpu  # noqa: E402
from radgpu import scenes  # noqa: E402

COPY_BYTES = 256 << 20


def linear():
    n = 32768
    sc = scenes.patch_room(1, n)
    e = radgpu.Engine("-triangle", str(n), "-quiet", "1")
    e.set_scene(sc).prepare_host().upload()
    lin = radgpu.LinearSystem(n)
    print("form factors %.1f ms" % lin.from_scene(e), flush=True)
    E, rho = scenes.emission_reflectance(sc.frgb, sc.brgb)
    for prec in ("f16", "bf16"):
        B, info = lin.gather(rho, E, prec=prec, tol=1e-6, max_outer=40, sweeps=8)
        print(prec, info, flush=True)
    E192, rho192 = scenes.lighting_vectors(rho, count=64, lights=8, seed=7)
    B, info = lin.gather(rho192, E192, prec="f16", tol=1e-6, max_outer=40, sweeps=8)
    print("192", info, flush=True)
    B, info = lin.shoot(rho, E, max_shots=2000)
    print("shoot", info, flush=True)
    lin.close()
    e.close()


def gemm():
    n = 32768
    sc = scenes.patch_room(1, n)
    e = radgpu.Engine("-triangle", str(n), "-quiet", "1")
    e.set_scene(sc).prepare_host().upload()
    lin = radgpu.LinearSystem(n)
    print("form factors %.1f ms" % lin.from_scene(e), flush=True)
    lin.set_tile(192, "f16", 4, 2)
    print("gemm 192 f16 %.3f ms" % lin.gemm_ms(192, "f16", reps=5), flush=True)
    lin.close()
    e.close()


def batch():
    import numpy as np
    sizes = scenes.batch_sizes(256, 500, 8000, seed=4)
    b = radgpu.SceneBatch(sizes)
    E_all, rho_all = [], []
    for s, n in enumerate(sizes):
        sc = scenes.patch_room(1000 + s, n)
        with radgpu.Engine("-triangle", str(n), "-quiet", "1") as e:
            e.set_scene(sc).prepare_host().upload()
            b.from_scene(s, e)
        E, rho = scenes.emission_reflectance(sc.frgb, sc.brgb)
        E_all.append(E)
        rho_all.append(rho)
    E, rho = np.vstack(E_all), np.vstack(rho_all)
    print("tile", b.autotune("f16"), flush=True)
    for _ in range(2):
        B, rel, info = b.solve(rho, E, prec="f16", graph=True, tol=1e-6, outer=12, sweeps=6)
        print("batch", info, float(rel.max()), flush=True)
    b.close()


def adjoint():
    import numpy as np
    n = 16384
    sc = scenes.patch_room(1, n)
    e = radgpu.Engine("-triangle", str(n), "-quiet", "1")
    e.set_scene(sc).prepare_host().upload()
    lin = radgpu.LinearSystem(n)
    print("form factors %.1f ms" % lin.from_scene(e), flush=True)
    E, rho = scenes.emission_reflectance(sc.frgb, sc.brgb)
    B, info = lin.gather(rho, E, prec="f16", tol=1e-12, max_outer=40, sweeps=8)
    g = np.random.default_rng(3).standard_normal(B.shape)
    for _ in range(2):
        dE, drho, info = lin.adjoint(rho, B, g, tol=1e-8, max_outer=40, sweeps=8)
        print("adjoint", info, flush=True)
    lin.close()
    e.close()


def main():
    if "--linear" in sys.argv:
        linear()
    elif "--gemm" in sys.argv:
        gemm()
    elif "--batch" in sys.argv:
        batch()
    elif "--adjoint" in sys.argv:
        adjoint()
    else:
        # the benchmark's hall-200k solve: -converge 175 (emitters, then the
        # lit receivers), or with --direct -converge 1200 (emitters only)
        s = scenes.hall_scene(2)
        conv = "1200" if "--direct" in sys.argv else "175"
        # every kernel launched from the host (-graph 0): counter collection
        # then sees each dispatch of the same kernels the graphs replay
        e = radgpu.Engine("-converge", conv, "-triangle", "1000000", "-quiet", "1", "-graph", "0")
        e.set_scene(s).prepare_host().upload()
        n, conv = e.shoot()
        print("shots", n, conv, e.counters(), flush=True)
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(COPY_BYTES)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(COPY_BYTES)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(COPY_BYTES)) == 0
    assert hip.hipMemcpy(b, a, ctypes.c_size_t(COPY_BYTES), 3) == 0  # device to device
    assert hip.hipDeviceSynchronize() == 0
    print("copy of %d bytes done" % COPY_BYTES, flush=True)


if __name__ == "__main__":
    main()
