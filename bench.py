#!/usr/bin/env python3
"""Benchmark: plain progressive shooting at scale on one MI355X.

Workload ("hall-200k"): the seeded 100 x 100 x 30 hall with 400 yawed boxes
and 8 emissive quads, about 206k triangles, run with -converge 1200
-delta_ff 5e-6 -adapt_area 5 -triangle 1000000.  A step is one full solve:
the prepared scene is uploaded to HBM (outside the timed region), then the
device-resident shooting loop runs to convergence (timed).

value = rays traced per second over the timed steps.  The JSON line also
carries the reference-equivalent ray rate (rays the reference would cast for
the same solve, including AdaptCalFF's discarded full-source ray), the
time-to-converge, the roofline of the fused shoot kernel and a CPU baseline:
the oracle (a faithful CPU restatement of the reference, OpenMP over
receivers) timed on this host's cores over a bounded sample of the same
solve.
"""
import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "hw-radiosity_amd"))

METRIC = "rays/s + time-to-converge at 1e-3 rel-L2 vs FP64; mixed/batched vs plain, 1 GPU"
OPTS = ["-converge", "1200", "-delta_ff", "5e-6", "-adapt_area", "5", "-triangle", "1000000",
        "-write_cycle", "1000000"]
HBM_PEAK_GBS = 8000.0
MFMA_PEAK_TFLOPS = 2500.0  # dense FP16/BF16
SWEEP_CHOICES = (4, 6, 8, 12, 16)
# measured HBM bytes per k_trace dispatch on this workload (rocprofv3 --pmc
# FETCH_SIZE + WRITE_SIZE, copy-calibrated); see profiles/
PMC_SUMMARY = os.path.join(ROOT, "profiles", "k_trace_hall200k_pmc.json")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes(cnt):
    """Bytes the ray-tracing kernel must read per the path's definition: per
    ray 40 (origin, source patch) + 32 per leaf (box, list offset) + 8 per
    descent level + 76 per triangle test (n, c, 3 side planes, list entry)."""
    return 40 * cnt["rays"] + 32 * cnt["cells"] + 8 * cnt["levels"] + 76 * cnt["tests"]


def reference_drop_in(cores):
    """The shipped reference binary (oracle/_ref/rad, built from the
    reference's sources by oracle/ref.mk) against the engine's drop-in `rad`
    on the reference's own CPU-runnable case (the seeded C1 room): the
    "Radiosity Method took" span of each (InitRadiosity .. final .fin) and
    whether the .fin files are byte-identical."""
    import hashlib
    import re
    import subprocess
    from radgpu import scenes
    ref_bin = os.path.join(ROOT, "oracle", "_ref", "rad")
    gpu_bin = os.path.join(ROOT, "hw-radiosity_amd", "_bin", "rad")
    if not os.path.exists(ref_bin):
        return {"skipped": "oracle/_ref/rad not built"}
    out = {}
    with tempfile.TemporaryDirectory() as tmp:
        tri = os.path.join(tmp, "c1.tri")
        scenes.write_tri(tri, scenes.room_scene(1))
        for name, opts in (("c1 (-converge 1200 -delta_ff 5e-6 -adapt_area 5 -triangle 20000)",
                            ["-converge", "1200", "-delta_ff", "5e-6", "-adapt_area", "5", "-triangle", "20000"]),
                           ("c1 at -converge 100", ["-converge", "100", "-triangle", "20000"])):
            row = {}
            for who, exe in (("reference", ref_bin), ("engine", gpu_bin)):
                prefix = os.path.join(tmp, who)
                t0 = time.perf_counter()
                try:
                    r = subprocess.run([exe] + opts + ["-jobs", str(cores), tri, "-o", prefix], capture_output=True,
                                       text=True, timeout=300)
                except (OSError, subprocess.TimeoutExpired) as e:
                    row[who] = {"error": str(e)}
                    continue
                wall = time.perf_counter() - t0
                m = re.search(r"Radiosity Method took \S*?([0-9.]+)", re.sub(r"\x1b\[[0-9;]*m", "", r.stderr))
                it = re.search(r"#Iterations (\d+)", r.stderr)
                fin = prefix + ".fin"
                row[who] = {"rc": r.returncode, "wall_s": wall, "method_s": float(m.group(1)) if m else None,
                            "iterations": int(it.group(1)) if it else None,
                            "fin_md5": hashlib.md5(open(fin, "rb").read()).hexdigest() if os.path.exists(fin) else None}
            if "method_s" in row.get("reference", {}) and "method_s" in row.get("engine", {}):
                a, b = row["reference"]["method_s"], row["engine"]["method_s"]
                row["speedup_method_span"] = a / b if a and b else None
                row["fin_identical"] = row["reference"]["fin_md5"] == row["engine"]["fin_md5"]
            out[name] = row
            log("drop-in %s: %s" % (name, row))
    return out


def cpu_baseline(scene, seconds, cores):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol
    o = ol.Oracle(*(OPTS + ["-quiet", "1", "-jobs", str(cores)]))
    o.load(scene)
    o.build_tree()
    # InitRadiosity's O(N^2) neighbour pass as the reference runs it: a
    # sample of rows, extrapolated to all of them
    rows = 400
    t0 = time.perf_counter()
    o.neighbor_rows(rows)
    init_est = (time.perf_counter() - t0) * o.trinum / rows
    o.init_radiosity()
    o.prepartition()
    t0 = time.perf_counter()
    shots = 0
    while time.perf_counter() - t0 < seconds:
        d, conv = o.iterate(1)
        shots += d
        if conv:
            break
    dt = time.perf_counter() - t0
    c = o.counters()
    return {"rays": int(c[ol.CNT_RAYS]), "seconds": dt, "shots": shots, "init_est_s": init_est, "rows": rows}


def dense_reference(F, rho, E, tmp):
    """FP64 dense LU solve on the GPU in a child process (radgpu.dense_ref)."""
    import subprocess
    np.save(os.path.join(tmp, "F.npy"), F)
    np.save(os.path.join(tmp, "rho.npy"), rho)
    np.save(os.path.join(tmp, "E.npy"), E)
    env = dict(os.environ, PYTHONPATH=os.path.join(ROOT, "hw-radiosity_amd"))
    r = subprocess.run([sys.executable, "-m", "radgpu.dense_ref", tmp], env=env, capture_output=True, text=True,
                       timeout=600)
    log(r.stdout.strip() or r.stderr[-2000:])
    if r.returncode != 0:
        raise RuntimeError("dense reference failed: %s" % r.stderr[-2000:])
    return np.load(os.path.join(tmp, "B.npy"))


# (rows per wave / 16, k per tile, persistent grid) of the scene-by-scene solve
SEQ_TILES = ((4, 1024, 1024), (4, 2048, 1024), (2, 1024, 1024))


def batch_leg(count=256):
    import radgpu
    from radgpu import scenes
    sizes = scenes.batch_sizes(count, 500, 8000, seed=4)
    t0 = time.perf_counter()
    b = radgpu.SceneBatch(sizes)
    E_all, rho_all, ff_ms = [], [], 0.0
    for s, n in enumerate(sizes):
        sc = scenes.patch_room(1000 + s, n)
        with radgpu.Engine("-triangle", str(n), "-quiet", "1") as e:
            e.set_scene(sc).prepare_host().upload()
            ff_ms += b.from_scene(s, e)
        E, rho = scenes.emission_reflectance(sc.frgb, sc.brgb)
        E_all.append(E)
        rho_all.append(rho)
    E, rho = np.vstack(E_all), np.vstack(rho_all)
    log("batch: %d scenes, %d patches, built in %.1f s (form factors %.0f ms on the device)" % (
        count, sum(sizes), time.perf_counter() - t0, ff_ms))
    res = {"scenes": count, "patches": int(sum(sizes)), "sizes": "log-uniform [500, 8000], seed 4; rooms seeds 1000+i",
           "matrix_entries": int(sum(n * n for n in sizes)), "form_factor_ms": ff_ms}
    # up to 12 refinements of 6 sweeps, each scene stopping at 1e-6
    kw = dict(tol=1e-6, outer=12, sweeps=6)
    for prec in ("f16", "bf16"):
        t0 = time.perf_counter()
        tile = b.autotune(prec)
        tune_s = time.perf_counter() - t0
        b.solve(rho, E, prec=prec, graph=True, **kw)  # capture + warm
        runs = [b.solve(rho, E, prec=prec, graph=True, **kw) for _ in range(3)]
        Bg, relg, ig = min(runs, key=lambda r: r[2]["ms"])
        Bs, rels, is_ = b.solve(rho, E, prec=prec, graph=False, **kw)
        identical = bool(np.array_equal(Bg, Bs))
        # the unbatched comparison: scene after scene, under the schedule
        # that suits one scene at a time (split-K keeps a lone scene's
        # launch wide), the best of these
        seq = {}
        for t in SEQ_TILES:
            b.set_tile(*t)
            seq[t] = b.solve(rho, E, prec=prec, graph=False, **kw)[2]
        ts = min(seq, key=lambda t: seq[t]["ms"])
        res[prec] = {"graph_ms": ig["ms"], "graph_nodes": ig["launches"], "tile": list(tile),
                     "autotune_s": tune_s, "sequential_ms": seq[ts]["ms"], "sequential_tile": list(ts),
                     "sequential_ms_by_tile": {"%d/%d/%d" % t: seq[t]["ms"] for t in seq},
                     "sequential_launches": seq[ts]["launches"],
                     "speedup_graph_vs_sequential": seq[ts]["ms"] / ig["ms"],
                     "identical_to_sequential_same_tile": identical, "max_rel_residual": float(relg.max()),
                     "scenes_above_tol": ig["above_tol"]}
        log("batch %s: %s" % (prec, res[prec]))
    b.close()
    return res


def linear_leg(n=32768, n_adj=16384):
    """Explicit-matrix solvers on the patch room: the mixed-precision gather
    (FP16 / BF16 MFMA, FP64 refinement) against plain FP32 progressive
    shooting, time to 1e-3 relative L2 against an FP64 dense solve; 192
    right-hand sides in one GEMM; the adjoint."""
    import radgpu
    from radgpu import scenes
    out = {}
    t0 = time.perf_counter()
    sc = scenes.patch_room(1, n)
    eng = radgpu.Engine("-triangle", str(n), "-quiet", "1")
    eng.set_scene(sc).prepare_host().upload()
    lin = radgpu.LinearSystem(n)
    ff_ms = lin.from_scene(eng)
    F = lin.matrix()
    E, rho = scenes.emission_reflectance(sc.frgb, sc.brgb)
    area = eng.records()[0]["area"].astype(np.float64)
    log("linear: n=%d, F built in %.1f ms (%.3f nonzero), setup %.1f s" % (n, ff_ms, (F > 0).mean(),
                                                                          time.perf_counter() - t0))
    E192, rho192 = scenes.lighting_vectors(rho, count=64, lights=8, seed=7)
    with tempfile.TemporaryDirectory() as tmp:
        ref = dense_reference(F, np.hstack([rho, rho192]), np.hstack([E, E192]), tmp)
    Bref, Bref192 = ref[:, :3], ref[:, 3:]
    out["scene"] = {"workload": "patch room (room_scene seed 1 geometry) at %d patches" % n,
                    "nonzero_fraction": float((F > 0).mean()), "form_factor_ms": ff_ms,
                    "form_factor_rays_per_s": float(eng.counters()["rays"]) / (ff_ms * 1e-3),
                    "reference": "FP64 dense LU (torch.linalg.solve on the GPU, separate process)"}
    res = {}
    for prec in ("f16", "bf16"):
        lin.autotune(3, prec)
        lin.gather(rho, E, prec=prec, tol=1e-6, max_outer=40, sweeps=8)  # warm
        tried = {}
        for sw in SWEEP_CHOICES:  # sweeps per refinement: the solver's one free parameter
            _, info = lin.gather(rho, E, prec=prec, tol=1e-6, max_outer=40, sweeps=sw, Bref=Bref, target=1e-3)
            tried[sw] = info
        best = min(tried, key=lambda k: tried[k]["ms_target"] if tried[k]["ms_target"] > 0 else 1e30)
        info = dict(tried[best], sweeps_per_refinement=best, tile=lin.autotune(3, prec),
                    ms_target_by_sweeps={k: v["ms_target"] for k, v in tried.items()},
                    ms_to_1e6_by_sweeps={k: v["ms"] for k, v in tried.items()})
        res["gather_" + prec] = info
        log("gather %s: %s" % (prec, info))
    B, info = lin.shoot(rho, E, Bref=Bref, target=1e-3)
    res["shoot_f32"] = info
    log("shoot: %s" % info)
    res["speedup_gather_f16_vs_shoot_at_1e-3"] = info["ms_target"] / res["gather_f16"]["ms_target"]
    res["speedup_gather_bf16_vs_shoot_at_1e-3"] = info["ms_target"] / res["gather_bf16"]["ms_target"]
    out["precision"] = res
    # roofline of the low-precision product, per sweep: algorithmic bytes
    # 2 N^2 (F in FP16/BF16) + 8 N R (X in, Y out) and flops 2 N^2 R
    roof = {}
    for nrhs in (3, 192):
        for prec in ("f16", "bf16"):
            tile = lin.autotune(nrhs, prec)
            ms = lin.gemm_ms(nrhs, prec, reps=30)
            byts = 2.0 * n * n + 8.0 * n * nrhs
            flops = 2.0 * n * n * nrhs
            roof["%s_%d" % (prec, nrhs)] = {"ms": ms, "achieved_GBps": byts / ms / 1e6,
                                           "hbm_frac": byts / ms / 1e6 / HBM_PEAK_GBS,
                                           "achieved_TFLOPs": flops / ms / 1e9,
                                           "mfma_frac": flops / ms / 1e9 / MFMA_PEAK_TFLOPS,
                                           "tile_rows_per_wave_x16_splitk": tile}
    out["gemm_roofline"] = roof
    # batching: 64 lighting vectors x RGB in one GEMM vs one vector at a time,
    # each with its own best sweeps per refinement (with 192 columns the FP64
    # residual costs as much as ~16 sweeps)
    def best_of(rho_, E_, Bref_, choices):
        tried = {}
        for sw in choices:
            _, inf = lin.gather(rho_, E_, prec="f16", tol=1e-6, max_outer=40, sweeps=sw, Bref=Bref_, target=1e-3)
            tried[sw] = inf
        sw = min(tried, key=lambda k: tried[k]["ms_target"] if tried[k]["ms_target"] > 0 else 1e30)
        return sw, tried
    lin.gather(rho192, E192, prec="f16", tol=1e-6, max_outer=40, sweeps=8)  # warm (autotune 192)
    sw, tried = best_of(rho192, E192, Bref192, SWEEP_CHOICES)
    info = tried[sw]
    sw1, tried1 = best_of(rho192[:, :3], E192[:, :3], Bref192[:, :3], SWEEP_CHOICES)
    one = tried1[sw1]
    out["batching_192_rhs"] = {"gather_f16": info, "one_vector_ms_target": one["ms_target"],
                               "sweeps_per_refinement": sw, "one_vector_sweeps_per_refinement": sw1,
                               "ms_target_by_sweeps": {k: v["ms_target"] for k, v in tried.items()},
                               "tile": lin.autotune(192, "f16"),
                               "speedup_vs_64_single_vector_solves": 64 * one["ms_target"] / info["ms_target"]}
    log("batched 192: %s" % info)
    lin.close()
    eng.close()
    # many scenes: 256 rooms of log-uniform size in one graph launch vs the
    # same kernels launched scene after scene
    out["batched_scenes"] = batch_leg()
    # adjoint at n_adj
    sc = scenes.patch_room(5, n_adj)
    eng = radgpu.Engine("-triangle", str(n_adj), "-quiet", "1")
    eng.set_scene(sc).prepare_host().upload()
    lin = radgpu.LinearSystem(n_adj)
    lin.from_scene(eng)
    E, rho = scenes.emission_reflectance(sc.frgb, sc.brgb)
    rho_t = scenes.hidden_reflectance(n_adj, seed=11)
    Bt, _ = lin.gather(rho_t, E, prec="f16", tol=1e-10, max_outer=60, sweeps=8)
    B, _ = lin.gather(rho, E, prec="f16", tol=1e-10, max_outer=60, sweeps=8)
    by_sweeps = {}
    for sw in SWEEP_CHOICES:  # FP32 sweeps per FP64 refinement, as for the gather
        dE, drho, inf = lin.adjoint(rho, B, 2 * (B - Bt), tol=1e-8, max_outer=40, sweeps=sw)
        by_sweeps[sw] = inf
    best = min(by_sweeps, key=lambda k: by_sweeps[k]["ms"])
    info = by_sweeps[best]
    out["adjoint"] = {"n": n_adj, "loss": float(np.sum((B - Bt) ** 2)), "info": info, "sweeps_per_refinement": best,
                      "ms_by_sweeps": {k: v["ms"] for k, v in by_sweeps.items()}}
    log("adjoint: %s" % info)
    lin.close()
    eng.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--boxes", type=int, default=400)
    ap.add_argument("--no-linear", action="store_true", help="skip the explicit-matrix leg")
    ap.add_argument("--no-e2e", action="store_true",
                    help="skip the end-to-end run (so a kernel trace holds the timed steps' launches only)")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: the shooting loop runs on a single MI355X "
              "(multi-GPU sharding is a design note, not part of this build)", file=sys.stderr)
        sys.exit(2)

    import radgpu
    from radgpu import scenes

    t0 = time.perf_counter()
    scene = scenes.hall_scene(2, n_boxes=args.boxes)
    log("scene: %d triangles (%.1f s)" % (scene.n, time.perf_counter() - t0))

    # HIP runtime start-up (a once-per-process cost) happens before any timing
    import ctypes
    ctypes.CDLL("libamdhip64.so").hipFree(ctypes.c_void_p(0))
    # end-to-end once: InitRadiosity through the final .fin, as the reference times it
    e2e_s, e2e_t = None, {"setup_ms": None, "upload_ms": None}
    if not args.no_e2e:
        with tempfile.TemporaryDirectory() as tmp:
            # the other once-per-process device costs (loading the device code,
            # the first allocations, a cold GPU's first launches) are start-up
            # too: one run of the small room through the same path first
            with radgpu.Engine(*(OPTS + ["-quiet", "1"])) as warm:
                warm.set_scene(scenes.room_scene(1))
                warm.run(os.path.join(tmp, "warm"))
            with radgpu.Engine(*(OPTS + ["-quiet", "1"])) as e2e:
                e2e.set_scene(scene)
                t1 = time.perf_counter()
                iters = e2e.run(os.path.join(tmp, "hall"))
                e2e_s = time.perf_counter() - t1
                e2e_t = e2e.timings()
        log("end-to-end solve (init .. .fin): %.3f s, %d iterations" % (e2e_s, iters))

    eng = radgpu.Engine(*(OPTS + ["-quiet", "1", "-kernel_timing", "1"]))
    eng.set_scene(scene).prepare_host()
    rec, _ = eng.records()
    slots = len(rec)
    for w in range(args.warmup):
        eng.upload()
        eng.shoot()
    total_s, rays, ref_rays, shots_all, kms, kn = 0.0, 0, 0, 0, 0.0, 0
    cnt_sum = None
    for k in range(args.steps):
        eng.upload()
        ts = time.perf_counter()
        shots, conv = eng.shoot()
        dt = time.perf_counter() - ts
        assert conv, "did not converge"
        c = eng.counters()
        t = eng.timings()
        total_s += dt
        shots_all += shots
        kms += t["kernel_ms"]
        kn += t["kernels_timed"]
        cnt_sum = c if cnt_sum is None else {x: cnt_sum[x] + c[x] for x in c}
        log("step %d: %d shots, %.3f s, %d rays" % (k, shots, dt, c["rays"]))
    rays, ref_rays = cnt_sum["rays"], cnt_sum["ref_rays"]
    kernel_avg_ms = kms / max(kn, 1)
    bytes_per_launch = algorithmic_bytes(cnt_sum) / max(kn, 1)
    achieved = bytes_per_launch / (kernel_avg_ms * 1e-3) / 1e9 if kernel_avg_ms > 0 else 0.0
    value = rays / total_s
    traffic = None
    if os.path.exists(PMC_SUMMARY):
        with open(PMC_SUMMARY) as fh:
            traffic = json.load(fh).get("traffic_bytes_per_dispatch")
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "rays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * total_s / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded hall scene, see hw-radiosity_amd/radgpu/scenes.py)",
        "config": {"workload": "hall-200k: plain progressive shooting to convergence (config 2)",
                   "triangles": scene.n, "slots_after_prepartition": slots, "shots_per_solve": shots_all // args.steps,
                   "options": " ".join(OPTS), "parallelism": "single GPU"},
        "ref_equivalent_rays_per_s": ref_rays / total_s,
        "time_to_converge_s": {"device_shooting": total_s / args.steps, "end_to_end_init_to_fin": e2e_s,
                               "host_setup_ms": e2e_t["setup_ms"], "upload_ms": e2e_t["upload_ms"]},
        "roofline": {"kernel": "k_trace (octree-DDA visibility of the iteration's rays)",
                     "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "avg_kernel_ms": kernel_avg_ms, "launches": kn,
                     "algorithmic_bytes_per_launch": bytes_per_launch,
                     "note": "algorithmic bytes count every octree and triangle byte a ray reads; most "
                             "are L2/MALL hits (traffic = measured HBM bytes per launch), so frac above "
                             "1 means the kernel is latency/divergence-bound, not HBM-bound; rocprofv3 "
                             "SQ counters of the same workload put its VALU lane use at 0.44 "
                             "(profiles/k_trace_hall200k_sq.md)"},
        "counters": cnt_sum,
    }
    eng.close()
    if not args.no_linear:
        out["linear"] = linear_leg()
    if not args.no_cpu:
        cores = min(16, os.cpu_count() or 1)
        cb = cpu_baseline(scene, args.cpu_seconds, cores)
        cpu_rate = cb["rays"] / cb["seconds"]
        per_shot = cb["seconds"] / max(cb["shots"], 1)
        out["cpu_baseline"] = {"value": cpu_rate, "unit": "rays/s", "cores": cores, "kind": "port",
                               "sample": "oracle (refsem, -O2 OpenMP) on the same hall scene: first %d shots of the "
                                         "solve, %.1f s; faithful to the reference, so it also casts AdaptCalFF's "
                                         "discarded full-source rays" % (cb["shots"], cb["seconds"]),
                               "time_to_converge_est_s": per_shot * shots_all / args.steps}
        out["speedup_time_to_converge_device"] = out["cpu_baseline"]["time_to_converge_est_s"] / (total_s / args.steps)
        out["cpu_baseline"]["init_neighbour_pass_est_s"] = cb["init_est_s"]
        out["cpu_baseline"]["init_sample"] = ("the reference's exact O(N^2) neighbour pass (rad.cpp:295-334, OpenMP "
                                              "over j) on the first %d rows, extrapolated to all %d" %
                                              (cb["rows"], scene.n))
        est = cb["init_est_s"] + out["cpu_baseline"]["time_to_converge_est_s"]
        out["cpu_baseline"]["end_to_end_est_s"] = est
        out["speedup_end_to_end_vs_cpu_est"] = est / e2e_s if e2e_s else None
        out["reference_drop_in_c1"] = reference_drop_in(cores)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
