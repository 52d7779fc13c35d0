#!/usr/bin/env python3
"""Benchmark: plain progressive shooting at scale on one MI355X.

Workload ("hall-200k"): the seeded 100 x 100 x 30 hall with 400 yawed boxes
and 8 emissive quads, about 206k triangles.  The headline solve runs with
-converge 175 -delta_ff 5e-6 -adapt_area 5 -triangle 1000000: the 64 emitter
shots, then the lit receivers shooting on (interreflection), over a thousand
shots in all.  A step is one full solve: the prepared scene is uploaded to
HBM (outside the timed region), then the device-resident shooting loop runs
to convergence (timed).

value = rays traced per second over the timed steps.  The JSON line also
carries the reference-equivalent ray rate (rays the reference would cast for
the same solve, including AdaptCalFF's discarded full-source ray), the
time-to-converge, the roofline of the ray-tracing kernel and a CPU baseline:
the oracle (a faithful CPU restatement of the reference, OpenMP over
receivers) timed on this host's cores over a bounded sample of the same
solve, with its records compared to the engine's; the reference binary built
from its sources timed beside the oracle and the engine's drop-in `rad` on
the reference's own models; and the direct-light solve (-converge 1200, the
emitters only) run to the end on both sides and compared record for record.
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
COMMON = ["-delta_ff", "5e-6", "-adapt_area", "5", "-triangle", "1000000", "-write_cycle", "1000000"]
HEADLINE_CONVERGE = "175"    # receivers shoot on after the emitters: interreflection
DIRECT_CONVERGE = "1200"     # the emitters only (64 shots on this scene)
OPTS = ["-converge", HEADLINE_CONVERGE] + COMMON
DIRECT_OPTS = ["-converge", DIRECT_CONVERGE] + COMMON
HBM_PEAK_GBS = 8000.0
MFMA_PEAK_TFLOPS = 2500.0  # dense FP16/BF16
# VALU issue peak: 256 CUs x 4 SIMDs x 16 lanes per clock at 2.4 GHz (a
# wave64 VALU instruction occupies its SIMD for 4 clocks), in G lane-ops/s
VALU_PEAK_GLANES = 256 * 4 * 16 * 2.4
SWEEP_CHOICES = (4, 6, 8, 12, 16)
# rocprofv3 summaries of k_trace on the headline workload: SQ counters per
# dispatch (VALU thread-instructions) and HBM bytes per dispatch (FETCH_SIZE
# + WRITE_SIZE, copy-calibrated)
VALU_SUMMARY = os.path.join(ROOT, "profiles", "k_trace_hall200k_c175_sq.json")
PMC_SUMMARY = os.path.join(ROOT, "profiles", "k_trace_hall200k_c175_pmc.json")
MODELS = os.path.join(ROOT, "tests", "golden", "models")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_threads():
    """Host threads for the CPU legs: the CPUs this process may run on,
    capped by OMP_NUM_THREADS where the machine sets it (a GPU box sets its
    share of the host that way)."""
    aff = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS", "")
    return (min(aff, int(omp)) if omp.isdigit() and int(omp) > 0 else aff), aff


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for ln in fh:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def algorithmic_bytes(cnt):
    """Bytes the ray-tracing kernel must read per the path's definition: per
    ray 40 (origin, source patch) + 32 per leaf (box, list offset) + 8 per
    descent level + 76 per triangle test (n, c, 3 side planes, list entry)."""
    return 40 * cnt["rays"] + 32 * cnt["cells"] + 8 * cnt["levels"] + 76 * cnt["tests"]


def _load_json(path):
    if os.path.exists(path):
        with open(path) as fh:
            return json.load(fh)
    return None


def _run_timed(exe, args, timeout=600, cwd=None):
    """One `rad`-style run: its "Radiosity Method took" span, iterations and
    the .fin md5."""
    import hashlib
    import re
    import subprocess
    t0 = time.perf_counter()
    try:
        r = subprocess.run([exe] + args, capture_output=True, text=True, timeout=timeout, cwd=cwd)
    except (OSError, subprocess.TimeoutExpired) as e:
        return {"error": str(e)}
    wall = time.perf_counter() - t0
    err = re.sub(r"\x1b\[[0-9;]*m", "", r.stderr)
    m = re.search(r"Radiosity Method took \S*?([0-9.]+)", err)
    it = re.search(r"#Iterations (\d+)", err) or re.search(r"iterations (\d+)", err)
    prefix = args[args.index("-o") + 1]
    fin = prefix + ".fin"
    return {"rc": r.returncode, "wall_s": wall, "method_s": float(m.group(1)) if m else None,
            "iterations": int(it.group(1)) if it else None,
            "fin_md5": hashlib.md5(open(fin, "rb").read()).hexdigest() if os.path.exists(fin) else None}


def reference_beside(threads):
    """The reference binary built from its sources (oracle/_ref/rad), the
    oracle's CLI (oracle/_bin/refsem, with the reference's exact O(N^2)
    neighbour pass) and the engine's drop-in `rad` on the reference's own
    models at -triangle 40000 (its published workloads, doc/README.md:404-408)
    and on the seeded C1 room; and the reference's V3 build (oracle/_ref/rad_v3,
    Delta-ff splitting live, at -jobs 1) beside `rad -variant v3` on the
    coarse room: the "Radiosity Method took" span of each (InitRadiosity ..
    final .fin), iterations, and whether the .fin files are byte-identical.
    Gives the oracle/reference speed ratio that anchors the CPU baseline, and
    the drop-in's speed-up over the reference."""
    import gzip
    import shutil
    from radgpu import scenes
    ref_bin = os.path.join(ROOT, "oracle", "_ref", "rad")
    sem_bin = os.path.join(ROOT, "oracle", "_bin", "refsem")
    gpu_bin = os.path.join(ROOT, "hw-radiosity_amd", "_bin", "rad")
    if not os.path.exists(ref_bin):
        return {"skipped": "oracle/_ref/rad not built"}
    out = {}
    with tempfile.TemporaryDirectory() as tmp:
        cases = []
        for m in ("room", "hall", "church"):
            tri = os.path.join(tmp, m + ".tri")
            with gzip.open(os.path.join(MODELS, m + ".tri.gz"), "rb") as fi, open(tri, "wb") as fo:
                shutil.copyfileobj(fi, fo)
            cases.append(("%s.tri -triangle 40000" % m, tri, ["-triangle", "40000"]))
        c1 = os.path.join(tmp, "c1.tri")
        scenes.write_tri(c1, scenes.room_scene(1))
        cases.append(("c1 (-converge 1200 -delta_ff 5e-6 -adapt_area 5 -triangle 20000)", c1,
                      ["-converge", "1200", "-delta_ff", "5e-6", "-adapt_area", "5", "-triangle", "20000"]))
        cases.append(("c1 at -converge 100", c1, ["-converge", "100", "-triangle", "20000"]))
        # the V3 build (Delta-ff splitting live) at -jobs 1, its deterministic
        # order, against -variant v3, on the coarse room where receivers split
        v3room = os.path.join(tmp, "v3room.tri")
        scenes.write_tri(v3room, scenes.v3_room())
        v3_bin = os.path.join(ROOT, "oracle", "_ref", "rad_v3")
        if os.path.exists(v3_bin):
            cases.append(("v3 v3room -adapt_area 0.5 -triangle 50000", v3room,
                          ["-adapt_area", "0.5", "-triangle", "50000"]))
            cases.append(("v3 v3room -adapt_area 0.5 -delta_ff 1e-4 -sample_area 0.1 -converge 300 -triangle 50000",
                          v3room, ["-adapt_area", "0.5", "-delta_ff", "1e-4", "-sample_area", "0.1",
                                   "-converge", "300", "-triangle", "50000"]))
        for name, tri, opts in cases:
            row = {}
            v3 = name.startswith("v3 ")
            runs = ((("reference", v3_bin, []), ("engine", gpu_bin, ["-variant", "v3"])) if v3 else
                    (("reference", ref_bin, []), ("oracle", sem_bin, ["-neighbors", "exact"]), ("engine", gpu_bin, [])))
            for who, exe, extra in runs:
                if who == "oracle" and name.startswith("c1"):
                    continue
                args = opts + extra + ["-write_cycle", "1000000", "-jobs", "1" if v3 else str(threads),
                                       tri, "-o", os.path.join(tmp, who)]
                if who != "engine":
                    row[who] = _run_timed(exe, args)
                    continue
                # the engine's span is tens of ms, where a single fresh
                # process can be off by a few ms: the median of three runs
                tries = [_run_timed(exe, args) for _ in range(3)]
                good = sorted((t for t in tries if t.get("method_s")), key=lambda t: t["method_s"])
                row[who] = dict(good[len(good) // 2]) if good else tries[0]
                row[who]["method_s_runs"] = [t.get("method_s") for t in tries]
                row[who]["fin_md5_all_equal"] = len({t.get("fin_md5") for t in tries}) == 1
            ref, eng = row["reference"], row["engine"]
            if ref.get("method_s") and eng.get("method_s"):
                row["speedup_method_span"] = ref["method_s"] / eng["method_s"]
                row["fin_identical"] = ref["fin_md5"] == eng["fin_md5"]
            if "oracle" in row and ref.get("method_s") and row["oracle"].get("method_s"):
                row["oracle_over_reference_time"] = row["oracle"]["method_s"] / ref["method_s"]
                row["oracle_fin_identical"] = row["oracle"]["fin_md5"] == ref["fin_md5"]
            out[name] = row
            log("beside the reference, %s: %s" % (name, row))
    return out


def _oracle_for(scene, opts, threads):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol
    o = ol.Oracle(*(opts + ["-quiet", "1", "-jobs", str(threads)]))
    o.load(scene)
    o.build_tree()
    return o, ol


def cpu_baseline(scene, seconds, threads, engine_ref_rays, engine_shots):
    """The oracle on the headline solve: InitRadiosity's O(N^2) neighbour
    pass sampled on 400 rows, then as many shots as fit in `seconds`; its
    RayHitted calls per second, the time-to-converge that rate gives for the
    whole solve's reference-equivalent rays, and its records after those
    shots (for the comparison with the engine's)."""
    o, ol = _oracle_for(scene, OPTS, threads)
    rows = 400
    t0 = time.perf_counter()
    o.neighbor_rows(rows)
    init_est = (time.perf_counter() - t0) * o.trinum / rows
    o.init_radiosity()
    o.prepartition()
    t0 = time.perf_counter()
    shots, conv = 0, False
    step = 4
    while time.perf_counter() - t0 < seconds:
        d, conv = o.iterate(step)
        shots += d
        if conv:
            break
    dt = time.perf_counter() - t0
    c = o.counters()
    rays = int(c[ol.CNT_RAYS])
    rec, _ = o.dump()
    return {"rays": rays, "seconds": dt, "shots": shots, "init_est_s": init_est, "rows": rows,
            "converged": bool(conv), "records": rec,
            "time_to_converge_est_s": dt * engine_ref_rays / max(rays, 1) if shots < engine_shots else dt}


def direct_light_leg(scene, threads, steps):
    """The direct-light solve (-converge 1200: the 64 emitter shots) on the
    device and, to the end, on the oracle: both times measured, records
    compared."""
    import radgpu
    eng = radgpu.Engine(*(DIRECT_OPTS + ["-quiet", "1"]))
    eng.set_scene(scene).prepare_host()
    eng.upload()
    eng.shoot()
    tot, shots = 0.0, 0
    for _ in range(steps):
        eng.upload()
        t0 = time.perf_counter()
        shots, conv = eng.shoot()
        tot += time.perf_counter() - t0
        assert conv
    cnt = eng.counters()
    rec, _ = eng.records()
    eng.close()
    dev_s = tot / steps
    out = {"options": " ".join(DIRECT_OPTS), "shots": shots, "device_time_to_converge_s": dev_s,
           "rays_per_s": cnt["rays"] / dev_s}
    o, ol = _oracle_for(scene, DIRECT_OPTS, threads)
    o.init_radiosity()
    o.prepartition()
    t0 = time.perf_counter()
    oshots, oconv = o.iterate(1 << 30)
    odt = time.perf_counter() - t0
    orec, _ = o.dump()
    out.update({"oracle_time_to_converge_s": odt, "oracle_shots": oshots, "oracle_threads": threads,
                "speedup_device_vs_oracle": odt / dev_s,
                "output_identical": bool(oconv and oshots == shots and rec.tobytes() == orec.tobytes())})
    log("direct light: %s" % out)
    return out


def c1_iteration_latency():
    """Per-iteration device time of the small room (C1, 2k triangles) at
    -converge 100 (1000 shots), where launch latency, not work, bounds the
    loop."""
    import radgpu
    from radgpu import scenes
    out = {}
    for conv in ("100", "1200"):
        with radgpu.Engine("-converge", conv, "-triangle", "20000", "-quiet", "1") as e:
            e.set_scene(scenes.room_scene(1)).prepare_host()
            best = None
            for _ in range(3):
                e.upload()
                t0 = time.perf_counter()
                n, conv_ = e.shoot()
                wall = time.perf_counter() - t0
                t = e.timings()
                row = {"shots": n, "wall_ms": 1e3 * wall, "device_ms": t["shoot_ms"],
                       "device_us_per_iteration": 1e3 * t["shoot_ms"] / max(t["iterations_launched"], 1)}
                if best is None or row["wall_ms"] < best["wall_ms"]:
                    best = row
        out["converge_" + conv] = best
    log("c1 iteration latency: %s" % out)
    return out


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


def k_trace_roofline(kernel_ms, launches, bytes_per_launch):
    """The ray-tracing kernel's roofline.  Its working set (octree, lists,
    triangle records: ~14 MB) is served by L2 and the MALL, so the bound it
    can reach is VALU issue: lane-operations per launch (SQ_THREAD_CYCLES_VALU
    per dispatch from the committed rocprofv3 summary of this workload) over
    the launch's duration (HIP events around every launch, an untimed solve),
    against 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz.  The HBM view is kept
    beside it: the algorithmic bytes (which exceed HBM peak because they are
    cache hits) and the measured HBM traffic."""
    valu = _load_json(VALU_SUMMARY) or {}
    pmc = _load_json(PMC_SUMMARY) or {}
    traffic = pmc.get("traffic_bytes_per_dispatch")
    lanes = valu.get("thread_valu_per_dispatch")
    sec = kernel_ms * 1e-3
    achieved = lanes / sec / 1e9 if (lanes and sec > 0) else None
    alg_gbs = bytes_per_launch / sec / 1e9 if sec > 0 else None
    return {"kernel": "k_trace (octree-DDA visibility of one source's rays)",
            "bound": "valu", "achieved": achieved, "peak": VALU_PEAK_GLANES, "unit": "G lane-ops/s",
            "frac": achieved / VALU_PEAK_GLANES if achieved else None,
            "traffic": traffic, "avg_kernel_ms": kernel_ms, "launches": launches,
            "valu_issue_frac": valu.get("valu_issue_frac"), "valu_lane_use": valu.get("valu_lane_use"),
            "hbm": {"algorithmic_bytes_per_launch": bytes_per_launch, "algorithmic_GBps": alg_gbs,
                    "algorithmic_frac_of_peak": alg_gbs / HBM_PEAK_GBS if alg_gbs else None,
                    "measured_traffic_GBps": traffic / sec / 1e9 if (traffic and sec > 0) else None,
                    "measured_traffic_frac_of_peak": traffic / sec / 1e9 / HBM_PEAK_GBS if (traffic and sec > 0) else None,
                    "peak": HBM_PEAK_GBS, "fetch_scale": (pmc.get("FETCH_SIZE") or {}).get("scale"),
                    "write_scale": (pmc.get("WRITE_SIZE") or {}).get("scale")},
            "profiles": [os.path.relpath(VALU_SUMMARY, ROOT), os.path.relpath(PMC_SUMMARY, ROOT)],
            "note": "frac = VALU lane-ops per launch / duration / VALU issue peak; the algorithmic bytes "
                    "(40 + 32 C + 8 D + 76 T per ray) are cache hits, so their rate is not an HBM roofline"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU legs")
    ap.add_argument("--boxes", type=int, default=400)
    ap.add_argument("--no-linear", action="store_true", help="skip the explicit-matrix leg")
    ap.add_argument("--no-e2e", action="store_true",
                    help="skip every run but the timed steps (so a kernel trace holds their launches only)")
    ap.add_argument("--no-graph", action="store_true",
                    help="the timed steps launch every kernel from the host (-graph 0, the engine's default) "
                         "even where --graph is given")
    ap.add_argument("--graph", action="store_true",
                    help="the timed steps replay captured graphs of shooting iterations (-graph 1)")
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
    e2e_s, e2e_t, iters = None, {"setup_ms": None, "upload_ms": None}, None
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

    # the timed steps: no per-kernel events, no host wait inside a batch
    eng = radgpu.Engine(*(OPTS + ["-quiet", "1"] + (["-graph", "1"] if args.graph and not args.no_graph else [])))
    eng.set_scene(scene).prepare_host()
    rec, _ = eng.records()
    slots = len(rec)
    for w in range(args.warmup):
        eng.upload()
        eng.shoot()
    total_s, shots_all = 0.0, 0
    cnt_sum = None
    for k in range(args.steps):
        eng.upload()
        ts = time.perf_counter()
        shots, conv = eng.shoot()
        dt = time.perf_counter() - ts
        assert conv, "did not converge"
        c = eng.counters()
        total_s += dt
        shots_all += shots
        cnt_sum = c if cnt_sum is None else {x: cnt_sum[x] + c[x] for x in c}
        log("step %d: %d shots, %.3f s, %d rays" % (k, shots, dt, c["rays"]))
    final_records, _ = eng.records()
    eng.close()
    rays, ref_rays = cnt_sum["rays"], cnt_sum["ref_rays"]
    value = rays / total_s
    solve_ref_rays = ref_rays // args.steps
    solve_shots = shots_all // args.steps

    # the ray-tracing kernel's own duration: HIP events around every launch
    # on the engine's stream, in one more (untimed) solve of the same workload
    roofline = None
    if not args.no_e2e:
        with radgpu.Engine(*(OPTS + ["-quiet", "1", "-kernel_timing", "1"])) as ek:
            ek.set_scene(scene).prepare_host().upload()
            ek.shoot()
            t = ek.timings()
            ck = ek.counters()
        kn = max(t["kernels_timed"], 1)
        roofline = k_trace_roofline(t["kernel_ms"] / kn, t["kernels_timed"], algorithmic_bytes(ck) / kn)
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
        "config": {"workload": "hall-200k: plain progressive shooting to convergence with interreflection "
                               "(config 2, -converge %s)" % HEADLINE_CONVERGE,
                   "triangles": scene.n, "slots_after_prepartition": slots, "shots_per_solve": solve_shots,
                   "options": " ".join(OPTS), "parallelism": "single GPU"},
        "ref_equivalent_rays_per_s": ref_rays / total_s,
        "time_to_converge_s": {"device_shooting": total_s / args.steps, "end_to_end_init_to_fin": e2e_s,
                               "host_setup_ms": e2e_t["setup_ms"], "upload_ms": e2e_t["upload_ms"]},
        "roofline": roofline,
        "counters": cnt_sum,
    }
    if not args.no_linear:
        out["linear"] = linear_leg()
    if not args.no_cpu and not args.no_e2e:
        threads, aff = host_threads()
        cb = cpu_baseline(scene, args.cpu_seconds, threads, solve_ref_rays, solve_shots)
        cpu_rate = cb["rays"] / cb["seconds"]
        # the engine after the same number of shots, for the record comparison
        with radgpu.Engine(*(OPTS + ["-quiet", "1"])) as ep:
            ep.set_scene(scene).prepare_host().upload()
            done, _ = ep.shoot(cb["shots"])
            erec, _ = ep.records()
        same = done == cb["shots"] and erec.tobytes() == cb["records"].tobytes()
        if cb["converged"]:
            same = same and erec.tobytes() == final_records.tobytes()
        out["cpu_baseline"] = {"value": cpu_rate, "unit": "rays/s", "cores": threads, "kind": "port",
                               "cpu_model": cpu_model(), "host_cpus_available": aff,
                               "sample": "oracle (refsem, -O2 OpenMP, %d threads) on the same hall solve: its first "
                                         "%d shots (%d from receivers), %.1f s; faithful to the reference, so it also "
                                         "casts AdaptCalFF's discarded full-source rays (value counts them)" % (
                                             threads, cb["shots"], max(cb["shots"] - 64, 0), cb["seconds"]),
                               "output_identical_after_sample_shots": bool(same),
                               "time_to_converge_est_s": cb["time_to_converge_est_s"],
                               "time_to_converge_est_method": "sample seconds x (the solve's reference-equivalent "
                                                              "rays / the rays of the sample)"}
        out["speedup_time_to_converge_device"] = cb["time_to_converge_est_s"] / (total_s / args.steps)
        out["cpu_baseline"]["init_neighbour_pass_est_s"] = cb["init_est_s"]
        out["cpu_baseline"]["init_sample"] = ("the reference's exact O(N^2) neighbour pass (rad.cpp:295-334, OpenMP "
                                              "over j) on the first %d rows, extrapolated to all %d" %
                                              (cb["rows"], scene.n))
        est = cb["init_est_s"] + cb["time_to_converge_est_s"]
        out["cpu_baseline"]["end_to_end_est_s"] = est
        out["speedup_end_to_end_vs_cpu_est"] = est / e2e_s if e2e_s else None
        out["direct_light"] = direct_light_leg(scene, threads, max(1, min(args.steps, 5)))
        beside = reference_beside(threads)
        out["reference_beside"] = beside
        ratios = [r["oracle_over_reference_time"] for r in beside.values()
                  if isinstance(r, dict) and "oracle_over_reference_time" in r]
        out["cpu_baseline"]["oracle_over_reference_time"] = ratios
        out["c1_iteration_latency"] = c1_iteration_latency()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
