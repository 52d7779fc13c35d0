"""The engine's host stages (libradgpu.so, no GPU needed) against the
reference goldens and the oracle: ReadTriangle-derived records, octree facts,
InitRadiosity's neighbour links and pre-partition's slot contents must be
byte-identical; the .fin writer's text must match the oracle's."""
import hashlib
import math
import os
import struct

import numpy as np
import pytest

import oracle_lib as ol
import radgpu

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


def engine_for(entry, *extra):
    return radgpu.Engine(*(entry["opts"] + ["-fma", "1" if entry["variant"] == "fma" else "0", "-quiet", "1",
                           "-neighbors", "host", "-octree", "host"]
                           + list(extra)))


@pytest.mark.parametrize("key", case_keys())
def test_host_stages_match_reference(key, goldens, inputs):
    e = goldens["cases"][key]
    with engine_for(e) as eng:
        eng.load_tri(inputs[e["input"]])
        nn, nl, _, _ = eng.tree_info()
        assert (nn, nl) == (e["tree_nodes"], e["tree_lists"])
        assert sha(*eng.records()) == e["read"]["sha256"]
        eng.prepare_host()
        rec, g = eng.records()
        assert len(rec) == e["prepart"]["trinum"]
        assert sha(rec, g) == e["prepart"]["sha256"]


def test_set_triangles_equals_file(inputs):
    from radgpu import scenes
    s = scenes.read_tri(inputs["c1_seed1.tri"])
    with radgpu.Engine("-quiet", "1", "-octree", "host") as a, radgpu.Engine("-quiet", "1", "-octree", "host") as b:
        a.load_tri(inputs["c1_seed1.tri"])
        b.set_scene(s)
        assert a.records()[0].tobytes() == b.records()[0].tobytes()


def test_tree_matches_oracle(inputs):
    for m in ("room.tri", "blocks.tri", "c1_seed1.tri"):
        o = ol.Oracle("-quiet", "1")
        o.load_file(inputs[m])
        o.build_tree()
        with radgpu.Engine("-quiet", "1", "-octree", "host") as eng:
            eng.load_tri(inputs[m])
            a, b = eng.tree_info(), o.tree_info()
            assert a[:3] == b[:3] and struct.pack("f", a[3]) == struct.pack("f", b[3]), m


def test_fin_text_matches_oracle(inputs, tmp_path):
    """After pre-partition (lights carry accB), both writers print the same bytes."""
    for m, opts in (("room.tri", ["-triangle", "4000"]), ("c1_seed1.tri", ["-triangle", "20000"])):
        o = ol.Oracle(*(opts + ["-quiet", "1"]))
        o.load_file(inputs[m])
        o.build_tree()
        o.init_radiosity()
        o.prepartition()
        o.write_fin(str(tmp_path / "o.fin"))
        with radgpu.Engine(*(opts + ["-quiet", "1", "-neighbors", "host", "-octree", "host"])) as eng:
            eng.load_tri(inputs[m]).prepare_host()
            eng.write_fin(str(tmp_path / "e.fin"))
        assert open(tmp_path / "o.fin", "rb").read() == open(tmp_path / "e.fin", "rb").read(), m


def _ref_format(x):
    """fp::tostring<float> (float2string.h:103-271) restated from its definition:
    the exact value of x truncated to 8 fraction digits, trailing zeros removed."""
    bits = struct.unpack("<I", struct.pack("<f", x))[0]
    e = (bits >> 23) & 0xFF
    if e == 0xFF or (e == 0 and bits & 0x7FFFFF):
        return None
    sign = "-" if bits >> 31 else ""
    if e == 0:
        return sign + "0.0"
    ex = e - 150
    if ex >= 41:
        return None
    from fractions import Fraction
    v = Fraction((bits & 0x7FFFFF) | 0x800000) * (Fraction(2) ** ex)
    ip = int(v)
    if ex >= 0 or ex < -60:
        frac = ""
    else:
        digits = int((v - ip) * 10 ** 8)
        frac = ("%08d" % digits).rstrip("0")
    return "%s%d.%s" % (sign, ip, frac or "0")


def test_number_format_known_answers(tmp_path):
    """The engine's formatter, through a written .fin, against the definition
    and hand-checked values."""
    from radgpu import scenes
    assert _ref_format(1.0) == "1.0"
    assert _ref_format(-0.0) == "-0.0"
    assert _ref_format(0.1) == "0.1"            # 0.100000001490116...
    assert _ref_format(1.0 / 3.0) == "0.33333334"  # 0.3333333432674408
    assert _ref_format(2.0 ** -30) == "0.0"      # below 1e-8 after truncation
    assert _ref_format(123456.789) == "123456.78906250"[:-1].rstrip("0")
    rng = np.random.default_rng(3)
    vals = np.concatenate([rng.normal(size=300) * 10.0 ** rng.integers(-9, 12, 300),
                           [0.0, -0.0, 1e-45, 3e38, 1e-7, 5e-9, 16777217.0, -2.5]]).astype(np.float32)
    n = len(vals) // 9 * 9
    p = vals[:n].reshape(-1, 3, 3)
    s = scenes.Scene(p, np.full((len(p), 3), 100, np.int32), np.zeros((len(p), 3), np.int32))
    path = str(tmp_path / "v.tri")
    scenes.write_tri(path, s)
    with radgpu.Engine("-quiet", "1", "-octree", "host") as eng:
        eng.load_tri(path)
        eng.write_fin(str(tmp_path / "v.fin"))
    lines = open(tmp_path / "v.fin").read().split("\n")
    got = []
    for k in range(len(p)):
        for v in range(3):
            got.extend(lines[5 * k + 2 + v].split()[:3])
    last = ""
    for x, txt in zip(p.reshape(-1), got):
        want = _ref_format(float(x))
        if want is None:  # the reference reprints the previous text
            want = last
        assert txt == want, (float(x), txt, want)
        last = want


def test_hall_host_stages_match_oracle():
    """The 200k-triangle hall (the benchmark scene): records after
    InitRadiosity and pre-partition, neighbour links included, identical to
    the oracle's (its neighbour search is an independent hashed pass)."""
    from radgpu import scenes
    s = scenes.hall_scene(2)
    opts = ["-converge", "1200", "-triangle", "1000000", "-quiet", "1"]
    with radgpu.Engine(*(opts + ["-neighbors", "host", "-octree", "host"])) as eng:
        rec, g = eng.set_scene(s).prepare_host().records()
    o = ol.Oracle(*(opts + ["-jobs", "8"]))
    o.load(s)
    o.init_radiosity()
    o.prepartition()
    orec, og = o.dump()
    assert len(rec) == len(orec) == 253728
    assert (rec["neighbor"] >= 0).mean() > 0.9
    assert rec.tobytes() == orec.tobytes()
    assert np.array_equal(g, og)


def test_hall_tree_matches_oracle():
    """The benchmark scene's octree: node count, list entries, grid size and
    MinGridLen equal to the oracle's (the ray parity tests check the lists)."""
    from radgpu import scenes
    s = scenes.hall_scene(2)
    o = ol.Oracle("-quiet", "1", "-triangle", "1000000", "-jobs", "8")
    o.load(s)
    o.build_tree()
    with radgpu.Engine("-quiet", "1", "-triangle", "1000000", "-octree", "host") as eng:
        eng.set_scene(s)
        a, b = eng.tree_info(), o.tree_info()
    assert a[:3] == b[:3] and struct.pack("f", a[3]) == struct.pack("f", b[3])


def test_fin_unconvertible_values_across_chunks(tmp_path):
    """Values the reference's formatter does not convert (a degenerate
    triangle's NaN normal) print the previous value's text again; the engine
    formats records in parallel chunks of 4096, so degenerate records at and
    around the chunk starts must carry the text over exactly as the oracle's
    sequential writer does."""
    from radgpu import scenes
    rng = np.random.default_rng(12)
    n = 9000
    base = rng.uniform(0, 50, (n, 1, 3)).astype(np.float32)
    p = (base + rng.uniform(0, 1, (n, 3, 3))).astype(np.float32)
    for i in (0, 4095, 4096, 4097, 8191, 8192, n - 1):
        p[i, 1] = p[i, 0]
        p[i, 2] = p[i, 0]          # zero area: normal 0/0 = NaN
    s = scenes.Scene(p=p, frgb=rng.integers(0, 255, (n, 3)).astype(np.int32),
                     brgb=np.zeros((n, 3), np.int32))
    opts = ["-triangle", "20000", "-quiet", "1"]
    o = ol.Oracle(*opts)
    o.load(s)
    o.write_fin(str(tmp_path / "o.fin"))
    with radgpu.Engine(*(opts + ["-octree", "host"])) as eng:
        eng.set_scene(s)
        eng.write_fin(str(tmp_path / "e.fin"))
    a, b = open(tmp_path / "o.fin", "rb").read(), open(tmp_path / "e.fin", "rb").read()
    assert b"nan" not in a.lower() and a.count(b"Triangle") == n
    recs = a.split(b"Triangle \n")[1:]
    prev_last = recs[4095].split()[-4]           # record 4095's p[2][2] (before its normal)
    assert recs[4096].split(b"\n")[1].split()[3:6] == [prev_last] * 3  # 4096's normal repeats it
    assert a == b


def test_fin_wholly_unconvertible_record_at_chunk_end(tmp_path):
    """A record none of whose 12 values converts (coordinates of 2^64 or
    more and the NaN normal of a zero-area triangle) at the end of a 4096-record
    chunk: the next chunk starts from the text of the latest earlier record
    that does convert, as the sequential writer carries it."""
    from radgpu import scenes
    rng = np.random.default_rng(13)
    n = 8400
    base = rng.uniform(0, 50, (n, 1, 3)).astype(np.float32)
    p = (base + rng.uniform(0, 1, (n, 3, 3))).astype(np.float32)
    huge = np.array([1e20, 2e20, 3e20], np.float32)
    for i in (4094, 4095, 8190, 8191):
        p[i, :] = huge                 # every coordinate unconvertible, normal 0/0
    for i in (4093, 4096, 8192):
        p[i, 1] = p[i, 0]
        p[i, 2] = p[i, 0]              # convertible coordinates, NaN normal
    s = scenes.Scene(p=p, frgb=rng.integers(0, 255, (n, 3)).astype(np.int32),
                     brgb=np.zeros((n, 3), np.int32))
    opts = ["-triangle", "20000", "-quiet", "1"]
    o = ol.Oracle(*opts)
    o.load(s)
    o.write_fin(str(tmp_path / "o.fin"))
    with radgpu.Engine(*(opts + ["-octree", "host"])) as eng:
        eng.set_scene(s)
        eng.write_fin(str(tmp_path / "e.fin"))
    a, b = open(tmp_path / "o.fin", "rb").read(), open(tmp_path / "e.fin", "rb").read()
    recs = a.split(b"Triangle \n")[1:]
    assert len(recs) == n
    carried = recs[4093].split()[-4]   # 4093's last coordinate, before its repeated normal
    assert recs[4095].split(b"\n")[1].split() == [carried] * 6   # 4094/4095 print it throughout
    assert recs[4096].split(b"\n")[1].split()[3:6] == [carried] * 3  # so does 4096's normal
    assert a == b


@pytest.mark.parametrize("name", ["empty", "lone_light", "facing_pair", "no_light", "room_zero_area"])
def test_edge_scene_host_stages_match_oracle(name):
    """Host stages of the edge-case scenes (no triangles, a lone light, no
    light, zero-area triangles with NaN normals) equal the oracle's."""
    from edge_scenes import edge_scenes
    s = edge_scenes()[name]
    opts = ["-converge", "1200", "-triangle", "20000", "-quiet", "1"]
    with radgpu.Engine(*(opts + ["-neighbors", "host", "-octree", "host"])) as eng:
        a = eng.set_scene(s).tree_info()
        rec, g = eng.prepare_host().records()
    o = ol.Oracle(*opts)
    o.load(s)
    o.build_tree()
    b = o.tree_info()
    o.init_radiosity()
    o.prepartition()
    orec, og = o.dump()
    assert a[:3] == b[:3] and struct.pack("f", a[3]) == struct.pack("f", b[3])
    assert len(rec) == len(orec) and rec.tobytes() == orec.tobytes()
    assert g.tobytes() == og.tobytes()


def _ref_converter():
    exe = os.path.join(os.path.dirname(GOLDEN), "..", "oracle", "_ref", "tri2json")
    return exe if os.path.exists(exe) else None


@pytest.mark.parametrize("case", ["c1", "room_t5000", "empty"])
def test_viewer_json_matches_reference_converter(case, goldens, inputs, tmp_path):
    """tri2json's JSON from the oracle's .fin (itself byte-identical to the
    reference's, checked by md5) equals the reference converter's output
    (golden md5 from the reference's own .fin)."""
    g = goldens["viewer_json"][case]
    fin = str(tmp_path / "x.fin")
    if g["input"] is None:
        open(fin, "wb").close()
    else:
        o = ol.Oracle(*(g["opts"] + ["-quiet", "1", "-jobs", "8"]))
        o.load_file(inputs[g["input"]])
        o.build_tree()
        o.init_radiosity()
        o.prepartition()
        o.iterate(1 << 30)
        o.write_fin(fin)
    assert hashlib.md5(open(fin, "rb").read()).hexdigest() == g["fin_md5"]
    with radgpu.Engine("-quiet", "1", "-octree", "host") as eng:
        n = eng.fin_to_json(fin, str(tmp_path / "e.json"))
    js = open(tmp_path / "e.json", "rb").read()
    assert n == int(g["stderr"].split()[1])
    assert len(js) == g["json_bytes"] and hashlib.md5(js).hexdigest() == g["json_md5"]
    ref = _ref_converter()
    if ref:  # the reference converter itself, where it is built
        import subprocess
        subprocess.run([ref, "-i", fin, "-o", str(tmp_path / "r.json")], check=True, capture_output=True, timeout=120)
        assert open(tmp_path / "r.json", "rb").read() == js


def test_tri2json_cli_matches_reference_messages(goldens, inputs, tmp_path):
    """The command-line converter: same output file and stderr lines as the
    reference tool, usage errors with status 1, unreadable records rejected."""
    import subprocess
    exe = os.path.join(os.path.dirname(radgpu.LIB_PATH), "tri2json")
    fin = str(tmp_path / "x.fin")
    o = ol.Oracle("-quiet", "1", "-triangle", "20000")
    o.load_file(inputs["c1_seed1.tri"])
    o.init_radiosity()
    o.prepartition()
    o.write_fin(fin)
    r = subprocess.run([exe, "-i", fin, "-o", str(tmp_path / "a.json"), "--format", "COLOR"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0 and r.stderr.endswith("Model-Converter Success\n")
    with radgpu.Engine("-quiet", "1", "-octree", "host") as eng:
        eng.fin_to_json(fin, str(tmp_path / "b.json"))
    assert open(tmp_path / "a.json", "rb").read() == open(tmp_path / "b.json", "rb").read()
    assert subprocess.run([exe, "--help"], capture_output=True, timeout=60).returncode == 0
    assert subprocess.run([exe, "-i", fin], capture_output=True, timeout=60).returncode == 1
    assert subprocess.run([exe, "-i", fin, "-o", str(tmp_path / "c.json"), "--format", "SIMPLE"],
                          capture_output=True, timeout=60).returncode == 1
    bad = tmp_path / "bad.fin"
    bad.write_bytes(b"Triangle \n1 2 3 0 0 0\n0 0 0 0 0 1\n1 0 0 0 0 1\n")  # cut short
    assert subprocess.run([exe, "-i", str(bad), "-o", str(tmp_path / "d.json")], capture_output=True,
                          timeout=60).returncode == 1
    with radgpu.Engine("-quiet", "1", "-octree", "host") as eng:
        with pytest.raises(radgpu.RadError):
            eng.fin_to_json(str(bad), str(tmp_path / "e.json"))


def test_tree_arrays_match_oracle(inputs):
    """The host octree as the reference's TreeNodeStore / TriListStore
    (depth-first node order, every leaf list and its -1) equals the
    oracle's arrays byte for byte."""
    from radgpu import scenes
    cases = [(m, None) for m in ("room.tri", "blocks.tri", "c1_seed1.tri")] + [("hall", scenes.hall_scene(2, n_boxes=80))]
    for m, sc in cases:
        o = ol.Oracle("-quiet", "1", "-triangle", "1000000")
        with radgpu.Engine("-quiet", "1", "-triangle", "1000000", "-octree", "host") as eng:
            if sc is None:
                o.load_file(inputs[m])
                eng.load_tri(inputs[m])
            else:
                o.load(sc)
                eng.set_scene(sc)
            o.build_tree()
            nodes, lists = eng.tree_arrays()
        onodes, olists = o.tree_arrays()
        assert nodes.tobytes() == onodes.tobytes(), m
        assert np.array_equal(lists, olists), m


def test_prepartition_at_scale_matches_oracle():
    """Pre-partition keeps its queue as a binary heap (a std::set of (area,
    slot) in the reference and the oracle); on the hall scaled 3x, where it
    grows 206k originals to 444k slots through stale entries and neighbour
    splits, every record equals the oracle's."""
    from radgpu import scenes
    s = scenes.hall_scene(2)
    s3 = scenes.Scene(p=(s.p * 3).astype(np.float32), frgb=s.frgb, brgb=s.brgb)
    opts = ["-triangle", "4000000", "-max_tri", "4000000", "-quiet", "1"]
    with radgpu.Engine(*(opts + ["-neighbors", "host", "-octree", "host"])) as eng:
        rec, g = eng.set_scene(s3).prepare_host().records()
    o = ol.Oracle(*(opts + ["-jobs", "8"]))
    o.load(s3)
    o.init_radiosity()
    o.prepartition()
    orec, og = o.dump()
    assert len(rec) == len(orec) == 443890
    assert rec.tobytes() == orec.tobytes()


@pytest.mark.parametrize("octree", ["host", "device"])
def test_nonfinite_coordinates_refused(tmp_path, octree):
    """A .tri whose vertices are not finite has no octree grid (GridNum is
    (G1[0]-G0[0])/MinGrid, raycast.cpp): rad_load_tri refuses it with
    RAD_EINVAL before any octree is built, on either builder."""
    p = tmp_path / "nonfinite.tri"
    p.write_text("Triangle\n1 2 3 4 5 6\nnan 0 0\n1 inf 0\n0 1 1e39\n"
                 "Triangle\n1 1 1 0 0 0\n0 0 0\n1 0 0\n0 1 0\n")
    with radgpu.Engine("-quiet", "1", "-octree", octree) as eng:
        with pytest.raises(radgpu.RadError) as ei:
            eng.load_tri(str(p))
        assert ei.value.status == radgpu.RAD_EINVAL and "not finite" in str(ei.value)
