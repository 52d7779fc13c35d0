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

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "hw-radiosity_amd"))

METRIC = "rays/s + time-to-converge at 1e-3 rel-L2 vs FP64; mixed/batched vs plain, 1 GPU"
OPTS = ["-converge", "1200", "-delta_ff", "5e-6", "-adapt_area", "5", "-triangle", "1000000",
        "-write_cycle", "1000000"]
HBM_PEAK_GBS = 8000.0
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


def cpu_baseline(scene, seconds, cores):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol
    o = ol.Oracle(*(OPTS + ["-quiet", "1", "-jobs", str(cores)]))
    o.load(scene)
    o.build_tree()
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
    return {"rays": int(c[ol.CNT_RAYS]), "seconds": dt, "shots": shots}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--boxes", type=int, default=400)
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

    # end-to-end once: InitRadiosity through the final .fin, as the reference times it
    with tempfile.TemporaryDirectory() as tmp:
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
                             "1 means the kernel is latency/divergence-bound, not HBM-bound"},
        "counters": cnt_sum,
    }
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
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
