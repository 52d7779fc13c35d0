"""GPU parity: the HIP path through the C ABI against the reference's goldens
and the oracle.  Integer/index results (hits, iteration counts, topology) are
bit-exact; floating-point state (accB, deltaB, deltaaccB) is also held to
bit-exactness because the engine reproduces the reference's operation order
and contraction sites -- relative L2 is reported as a diagnostic only."""
import hashlib
import os

import numpy as np
import pytest

import oracle_lib as ol
import radgpu
from radgpu import scenes

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def sha(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def case_keys():
    import json
    with open(os.path.join(GOLDEN, "goldens.json")) as fh:
        return sorted(json.load(fh)["cases"])


def opts_of(e, *extra):
    return e["opts"] + ["-fma", "1" if e["variant"] == "fma" else "0", "-quiet", "1"] + list(extra)


def rel_l2(a, b):
    a, b = np.asarray(a, np.float64).ravel(), np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def describe_diff(rec, ref):
    out = []
    for f in ref.dtype.names:
        if rec[f].tobytes() != ref[f].tobytes():
            bad = np.nonzero((rec[f] != ref[f]).reshape(len(ref), -1).any(1))[0]
            out.append("%s: %d slots (first %s)" % (f, len(bad), bad[:4]))
    return "; ".join(out) + " | accB rel-L2 %.3g" % rel_l2(rec["accB"], ref["accB"])


@pytest.mark.parametrize("key", [k for k in case_keys()
                                 if os.path.exists(os.path.join(GOLDEN, "vectors_%s.npz" % k.replace("/", "_")))])
def test_cast_and_ff_match_reference(key, goldens, inputs):
    print("case", key, flush=True)
    e = goldens["cases"][key]
    v = np.load(os.path.join(GOLDEN, "vectors_%s.npz" % key.replace("/", "_")))
    rays = v["rays"].view(radgpu.RAY_DTYPE)
    with radgpu.Engine(*opts_of(e)) as eng:
        eng.load_tri(inputs[e["input"]]).prepare_host().upload()
        hit, cells, tests = eng.cast(rays)
        ff = eng.ff_pairs(v["pairs"])
    assert np.array_equal(hit, v["hits"]), "hits differ on %d rays" % int((hit != v["hits"]).sum())
    o = ol.Oracle(*opts_of(e))
    o.load_file(inputs[e["input"]])
    o.build_tree()
    o.init_radiosity()
    o.prepartition()
    ohit, ocells, otests = o.raycast(rays.view(ol.RAY_DTYPE))
    assert np.array_equal(cells, ocells) and np.array_equal(tests, otests)
    assert ff.tobytes() == o.ff_pairs(v["pairs"]).tobytes()
    # Shade's increments from the reference: lo = 255*k*dB/3, d = lo*ff
    k = np.float32(0.5 / 255.0)
    lo = (np.float32(255.0) * k * np.array([1000, 2000, 3000], np.float32)) / np.float32(3)
    d = (lo[None, None, :] * ff[:, :, None]).astype(np.float32).reshape(-1, 9)
    assert d.tobytes() == v["shade"].astype(np.float32).tobytes()


@pytest.mark.parametrize("key", case_keys())
def test_full_solve_matches_reference(key, goldens, inputs, tmp_path):
    print("case", key, flush=True)
    e = goldens["cases"][key]
    with radgpu.Engine(*opts_of(e, "-write_cycle", "1000000")) as eng:
        eng.load_tri(inputs[e["input"]])
        it = eng.run(str(tmp_path / "gpu"))
        rec, g = eng.records()
    ref = np.load(os.path.join(GOLDEN, "final_%s.npz" % key.replace("/", "_")))
    assert it == e["iterations"]
    assert len(rec) == e["final"]["trinum"]
    assert np.array_equal(rec["parent"], ref["parent"])
    assert rec["accB"].tobytes() == ref["accB"].tobytes(), "accB rel-L2 %.3g" % rel_l2(rec["accB"], ref["accB"])
    assert rec["deltaB"].tobytes() == ref["deltaB"].tobytes()
    assert sha(rec, g) == e["final"]["sha256"]
    assert hashlib.md5(open(tmp_path / "gpu.fin", "rb").read()).hexdigest() == e["fin_md5"]


def test_cli_drop_in(goldens, inputs, tmp_path):
    e = goldens["cases"]["c1/fma"]
    r = radgpu.run_cli(e["opts"] + ["-write_cycle", "10", inputs["c1_seed1.tri"], "-o", str(tmp_path / "c1")],
                       timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "Complete in #Iterations %d" % e["iterations"] in r.stderr
    assert hashlib.md5(open(tmp_path / "c1.fin", "rb").read()).hexdigest() == e["fin_md5"]
    # snapshots every 10 iterations, as -write_cycle asks
    assert sorted(p.name for p in tmp_path.iterdir() if p.name.startswith("c1.") and p.name != "c1.fin") == \
        ["c1.10", "c1.20"]
    # the viewer's model file from the engine's .fin, as the reference converter writes it
    import subprocess
    v = goldens["viewer_json"]["c1"]
    exe = os.path.join(os.path.dirname(radgpu.LIB_PATH), "tri2json")
    r = subprocess.run([exe, "-i", str(tmp_path / "c1.fin"), "-o", str(tmp_path / "c1.json")], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0 and r.stderr == v["stderr"]
    assert hashlib.md5(open(tmp_path / "c1.json", "rb").read()).hexdigest() == v["json_md5"]


@pytest.mark.parametrize("case", ["c1", "room_t5000", "box_t2000"])
def test_cli_old_format(case, goldens, inputs, tmp_path):
    """-old_format 1 writes the OLD_FORMAT build's per-vertex .fin."""
    print("case", case, flush=True)
    e = goldens["old_format"][case]
    r = radgpu.run_cli(e["opts"] + ["-old_format", "1", "-write_cycle", "1000000", inputs[e["input"]],
                                    "-o", str(tmp_path / "o")], timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    assert hashlib.md5(open(tmp_path / "o.fin", "rb").read()).hexdigest() == e["fin_md5"]


@pytest.mark.parametrize("select", ["high", "low"])
def test_hall_subset_matches_oracle(select):
    """The hall generator (yawed boxes, pre-partition at scale) with 40 boxes:
    the full solve is bit-identical to the oracle, with either tie-break."""
    s = scenes.hall_scene(2, n_boxes=40)
    opts = ["-converge", "1200", "-triangle", "1000000", "-select", select, "-quiet", "1"]
    print("hall40 select", select, s.n, flush=True)
    with radgpu.Engine(*opts) as eng:
        eng.set_scene(s).prepare_host().upload()
        shots, conv = eng.shoot()
        rec, _ = eng.records()
    o = ol.Oracle(*(opts + ["-jobs", "16"]))
    o.load(s)
    o.build_tree()
    o.init_radiosity()
    o.prepartition()
    oshots, oconv = o.iterate(1 << 30)
    orec, _ = o.dump()
    assert conv and oconv and shots == oshots
    assert rec.tobytes() == orec.tobytes(), describe_diff(rec, orec)


def test_hall_full_first_shots_match_oracle():
    """The 200k-triangle hall: the first 3 shots bit-identical to the oracle,
    and random rays through the full octree identical: 200000 rays take one
    lane each, the first 100000 of them cast alone two lanes each and the
    first 50000 four (each lane tests every second or fourth list entry, and
    the lanes combine), with the same hits, cells and tests."""
    s = scenes.hall_scene(2)
    opts = ["-converge", "1200", "-triangle", "1000000", "-quiet", "1"]
    print("hall full", s.n, flush=True)
    with radgpu.Engine(*opts) as eng:
        eng.set_scene(s).prepare_host().upload()
        rec0, g = eng.records()
        rng = np.random.default_rng(5)
        live = np.nonzero(rec0["parent"] >= 0)[0]
        n = 200000
        src, dst, v = rng.choice(live, n), rng.choice(live, n), rng.integers(0, 3, n)
        rays = np.zeros(n, radgpu.RAY_DTYPE)
        rays["p"] = rec0["p"][dst, v]
        rays["v"] = rec0["c"][src] - rec0["p"][dst, v]
        rays["ban"] = rec0["parent"][dst]
        rays["target"] = rec0["parent"][src]
        hit, cells, tests = eng.cast(rays)
        shared = [eng.cast(rays[:m]) for m in (100000, 50000)]
        shots, _ = eng.shoot(3)
        rec, _ = eng.records()
    o = ol.Oracle(*(opts + ["-jobs", "16"]))
    o.load(s)
    o.build_tree()
    o.init_radiosity()
    o.prepartition()
    orec0, _ = o.dump()
    assert rec0.tobytes() == orec0.tobytes()
    ohit, ocells, otests = o.raycast(rays.view(ol.RAY_DTYPE))
    assert np.array_equal(hit, ohit) and np.array_equal(cells, ocells) and np.array_equal(tests, otests)
    for h2, c2, t2 in shared:
        m = len(h2)
        assert np.array_equal(h2, hit[:m]) and np.array_equal(c2, cells[:m]) and np.array_equal(t2, tests[:m])
    o.iterate(3)
    orec, _ = o.dump()
    assert shots == 3
    assert rec.tobytes() == orec.tobytes(), describe_diff(rec, orec)


def test_repeat_runs_are_bitwise_equal():
    s = scenes.room_scene(3)
    outs = []
    for _ in range(2):
        with radgpu.Engine("-triangle", "20000", "-converge", "300", "-quiet", "1") as eng:
            eng.set_scene(s).prepare_host().upload()
            eng.shoot()
            outs.append(eng.records()[0].tobytes())
    assert outs[0] == outs[1]


def test_shoot_in_pieces_equals_one_call():
    s = scenes.room_scene(4)
    opts = ("-triangle", "20000", "-converge", "400", "-quiet", "1")
    with radgpu.Engine(*opts) as a, radgpu.Engine(*(opts + ("-batch", "3"))) as b:
        a.set_scene(s).prepare_host().upload()
        b.set_scene(s).prepare_host().upload()
        na, _ = a.shoot()
        nb = 0
        while True:
            k, conv = b.shoot(7)
            nb += k
            if conv:
                break
        assert na == nb
        assert a.records()[0].tobytes() == b.records()[0].tobytes()


def test_device_float_ops_are_ieee():
    """Correctly rounded division and square root and IEEE NaN handling in
    the walk: rays built to give a1 = 0 (NaN and inf hit parameters)."""
    s = scenes.room_scene(1)
    with radgpu.Engine("-quiet", "1", "-triangle", "20000") as eng:
        eng.set_scene(s).prepare_host().upload()
        rec, g = eng.records()
        rng = np.random.default_rng(9)
        n = 50000
        rays = np.zeros(n, radgpu.RAY_DTYPE)
        rays["p"] = (g[:3] + (g[3:] - g[:3]) * rng.random((n, 3))).astype(np.float32)
        d = rng.normal(size=(n, 3)).astype(np.float32)
        d[: n // 4, 2] = 0.0   # parallel to floor/ceiling planes
        d[n // 4: n // 2, :2] = 0.0
        rays["v"] = d
        rays["ban"] = -1
        hit, cells, tests = eng.cast(rays)
    o = ol.Oracle("-quiet", "1", "-triangle", "20000")
    o.load(s)
    o.build_tree()
    o.init_radiosity()
    o.prepartition()
    ohit, ocells, otests = o.raycast(rays.view(ol.RAY_DTYPE))
    assert np.array_equal(hit, ohit) and np.array_equal(cells, ocells) and np.array_equal(tests, otests)


@pytest.mark.parametrize("name", ["empty", "lone_light", "facing_pair", "no_light", "room_zero_area"])
def test_edge_scenes_match_oracle(name):
    """Full solves of the edge-case scenes (no triangles, a lone light with no
    receiver, two facing triangles, no light, zero-area triangles whose NaN
    normals make the reference's tests pass): shot counts and every record
    bit-identical to the oracle."""
    from edge_scenes import edge_scenes
    s = edge_scenes()[name]
    opts = ["-converge", "1200", "-triangle", "20000", "-quiet", "1"]
    print("edge scene", name, s.n, flush=True)
    with radgpu.Engine(*opts) as eng:
        eng.set_scene(s).prepare_host().upload()
        shots, conv = eng.shoot()
        rec, _ = eng.records()
    o = ol.Oracle(*opts)
    o.load(s)
    o.build_tree()
    o.init_radiosity()
    o.prepartition()
    oshots, oconv = o.iterate(1 << 30)
    orec, _ = o.dump()
    assert conv and oconv and shots == oshots, (shots, oshots)
    assert rec.tobytes() == orec.tobytes(), describe_diff(rec, orec)


def test_cli_zip_and_interactive_match_reference_binary(goldens, inputs, tmp_path):
    """-zip and -interactive against the reference binary built from its
    sources (oracle/_ref/rad), each run in its own directory holding a
    WebGL_demo/run.sh that records its argument: same files (gzip'd .fin
    with the same content), same script calls, same status lines."""
    import gzip
    import subprocess
    ref = os.path.join(os.path.dirname(GOLDEN), "..", "oracle", "_ref", "rad")
    if not os.path.exists(ref):
        pytest.skip("reference binary not built here")
    opts = ["-converge", "1200", "-triangle", "20000", "-zip", "-interactive", "-write_cycle", "10", "-jobs", "2"]
    runs = {}
    for name in ("ref", "eng"):
        d = tmp_path / name
        (d / "WebGL_demo").mkdir(parents=True)
        script = d / "WebGL_demo" / "run.sh"
        script.write_text("#!/bin/bash\necho \"$1\" >> calls.txt\n")
        script.chmod(0o755)
        args = opts + [inputs["c1_seed1.tri"], "-o", "out"]
        if name == "ref":
            r = subprocess.run([ref] + args, cwd=d, capture_output=True, text=True, timeout=300)
        else:
            r = radgpu.run_cli(args, cwd=str(d), timeout=300)
        assert r.returncode == 0, (name, r.stderr[-2000:])
        files = sorted(p.name for p in d.iterdir() if p.name.startswith("out"))
        content = {f: gzip.open(d / f).read() for f in files}
        calls = (d / "WebGL_demo" / "calls.txt").read_text().split()
        status = [ln for ln in (r.stdout + r.stderr).splitlines()
                  if ln.startswith(("WebGL:", "Compress output file")) or ln in ("OK", "Failed")]
        runs[name] = (files, content, calls, status)
    assert runs["eng"][0] == runs["ref"][0] and runs["eng"][0] == ["out.10.gz", "out.20.gz", "out.fin.gz"]
    assert runs["eng"][1] == runs["ref"][1]
    assert runs["eng"][2] == runs["ref"][2] == ["../out.10", "../out.20", "../out.fin"]
    assert sorted(runs["eng"][3]) == sorted(runs["ref"][3])


@pytest.mark.gpu
@pytest.mark.parametrize("scene", ["room", "hall_subset"])
def test_reference_ray_count_matches_oracle(scene):
    """The engine traces only the rays whose result is used; its ref_rays
    counter adds the full-source rays AdaptCalFF casts and discards
    (shade.cpp:442-450), so it equals the oracle's RayHitted count over the
    same shots."""
    if scene == "room":
        s, opts, shots = scenes.room_scene(1), ["-converge", "1200", "-quiet", "1"], 1 << 30
    else:
        s, opts, shots = scenes.hall_scene(2, n_boxes=40), ["-converge", "1200", "-triangle", "1000000", "-quiet", "1"], 12
    print("ref rays", scene, s.n, flush=True)
    with radgpu.Engine(*opts) as eng:
        eng.set_scene(s).prepare_host().upload()
        n, _ = eng.shoot(shots)
        c = eng.counters()
    o = ol.Oracle(*(opts + ["-jobs", "16"]))
    o.load(s)
    o.build_tree()
    o.init_radiosity()
    o.prepartition()
    on, _ = o.iterate(shots)
    oc = o.counters()
    assert n == on and c["ref_rays"] == int(oc[0]) and 0 < c["rays"] <= c["ref_rays"]


def test_parallel_src_zero_shoots_nothing(inputs, tmp_path):
    """-parallel_src 0: listCanditate takes min(ceil(|C|/5), 0) = 0 sources
    (rad.cpp:400), so the first iteration finds none and the run ends at
    iteration 1 -- as the reference binary built from its sources and the
    oracle do."""
    import re
    import subprocess
    opts = ["-converge", "1200", "-triangle", "20000", "-parallel_src", "0", "-write_cycle", "1000000"]
    r = radgpu.run_cli(opts + [inputs["c1_seed1.tri"], "-o", str(tmp_path / "eng")], timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "Complete in #Iterations 1\n" in re.sub(r"\x1b\[[0-9;]*m", "", r.stderr)
    fin = open(tmp_path / "eng.fin", "rb").read()
    o = ol.Oracle(*(opts + ["-quiet", "1"]))
    o.load_file(inputs["c1_seed1.tri"])
    o.build_tree()
    o.init_radiosity()
    o.prepartition()
    shots, conv = o.iterate(1 << 30)
    assert shots == 0 and conv
    o.write_fin(str(tmp_path / "o.fin"))
    assert fin == open(tmp_path / "o.fin", "rb").read()
    ref = os.path.join(os.path.dirname(GOLDEN), "..", "oracle", "_ref", "rad")
    if os.path.exists(ref):
        rr = subprocess.run([ref] + opts + ["-jobs", "2", inputs["c1_seed1.tri"], "-o", str(tmp_path / "ref")],
                            capture_output=True, text=True, timeout=300)
        assert rr.returncode == 0
        assert "Complete in #Iterations 1\n" in re.sub(r"\x1b\[[0-9;]*m", "", rr.stderr)
        assert fin == open(tmp_path / "ref.fin", "rb").read()


@pytest.mark.parametrize("scene", ["c1", "hall_subset"])
def test_graph_replays_equal_eager_launches(scene):
    """With -graph 1 the shooting loop replays captured graphs of 8
    iterations between host checks; by default (-graph 0), and with
    -kernel_timing 1, it launches every kernel eagerly instead.  Solves in
    pieces (graph replays, eager tails and host synchronisations interleaved)
    give the same shots and records as the eager solves."""
    if scene == "c1":
        s, opts = scenes.room_scene(1), ["-converge", "100", "-triangle", "20000", "-quiet", "1"]
    else:
        s, opts = scenes.hall_scene(2, n_boxes=40), ["-converge", "175", "-triangle", "1000000", "-quiet", "1"]
    print("graph vs eager", scene, flush=True)
    with radgpu.Engine(*(opts + ["-kernel_timing", "1"])) as a, radgpu.Engine(*(opts + ["-graph", "1"])) as b:
        a.set_scene(s).prepare_host().upload()
        b.set_scene(s).prepare_host().upload()
        na, ca = a.shoot()
        nb = 0
        for chunk in (5, 8, 17, 1 << 30):
            k, cb = b.shoot(chunk)
            nb += k
            if cb:
                break
        assert ca and cb and na == nb and na > 24
        assert a.records()[0].tobytes() == b.records()[0].tobytes()
        assert a.counters() == b.counters()
        rec_b, cnt_b = b.records()[0].tobytes(), b.counters()
    with radgpu.Engine(*(opts + ["-graph", "0"])) as c:
        c.set_scene(s).prepare_host().upload()
        nc, cc = c.shoot()
        assert cc and nc == na
        assert c.records()[0].tobytes() == rec_b
        assert c.counters() == cnt_b


@pytest.mark.parametrize("key", case_keys())
def test_device_neighbour_search_matches_reference(key, goldens, inputs):
    """InitRadiosity's neighbour search on the device (the default): the
    records after pre-partition (which reads and rewrites the neighbour
    links) hash to the reference's, as the host search's do."""
    e = goldens["cases"][key]
    with radgpu.Engine(*opts_of(e)) as eng:
        eng.load_tri(inputs[e["input"]]).prepare_host()
        rec, g = eng.records()
    assert len(rec) == e["prepart"]["trinum"]
    assert sha(rec, g) == e["prepart"]["sha256"]


def test_device_neighbour_search_on_the_hall():
    """The 200k-triangle hall: device and host neighbour searches give the
    same records after InitRadiosity and pre-partition, equal to the
    oracle's; and with the device search the set-up takes less host time."""
    s = scenes.hall_scene(2)
    opts = ["-converge", "1200", "-triangle", "1000000", "-quiet", "1"]
    print("hall neighbours", s.n, flush=True)
    with radgpu.Engine(*opts) as dev, radgpu.Engine(*(opts + ["-neighbors", "host"])) as host:
        rd, gd = dev.set_scene(s).prepare_host().records()
        rh, gh = host.set_scene(s).prepare_host().records()
        td, th = dev.timings()["setup_ms"], host.timings()["setup_ms"]
    print("setup ms: device search %.1f, host search %.1f" % (td, th), flush=True)
    assert (rd["neighbor"] >= 0).mean() > 0.9
    assert rd.tobytes() == rh.tobytes() and np.array_equal(gd, gh)
    o = ol.Oracle(*(opts + ["-jobs", "16"]))
    o.load(s)
    o.init_radiosity()
    o.prepartition()
    orec, _ = o.dump()
    assert rd.tobytes() == orec.tobytes()


def test_device_octree_matches_oracle(inputs):
    """BuildTree on the device (the default): the reference's TreeNodeStore
    and TriListStore (node boxes and child links in depth-first order, every
    leaf's list) byte-identical to the oracle's and to the host builder's, on
    the reference's models and the 200k-triangle hall; the hall's build
    times are printed."""
    import time
    cases = [(m, None) for m in ("box.tri", "room.tri", "blocks.tri", "c1_seed1.tri")] + [("hall-200k", scenes.hall_scene(2))]
    for m, sc in cases:
        print("octree", m, flush=True)
        opts = ["-quiet", "1", "-triangle", "1000000"]
        o = ol.Oracle(*(opts + ["-jobs", "16"]))
        ms = {}
        trees = {}
        for where in ("device", "host"):
            with radgpu.Engine(*(opts + ["-octree", where])) as eng:
                t0 = time.perf_counter()
                if sc is None:
                    eng.load_tri(inputs[m])
                else:
                    eng.set_scene(sc)
                ms[where] = 1e3 * (time.perf_counter() - t0)
                trees[where] = (eng.tree_info(), eng.tree_arrays())
        if sc is None:
            o.load_file(inputs[m])
        else:
            o.load(sc)
        o.build_tree()
        onodes, olists = o.tree_arrays()
        info, (nodes, lists) = trees["device"]
        print("%s: %d nodes, %d list entries; set_scene %.1f ms (device octree) %.1f ms (host)" % (
            m, info[0], info[1], ms["device"], ms["host"]), flush=True)
        assert info == trees["host"][0] and info[:3] == o.tree_info()[:3]
        assert nodes.tobytes() == onodes.tobytes(), m
        assert np.array_equal(lists, olists), m
        assert nodes.tobytes() == trees["host"][1][0].tobytes()
