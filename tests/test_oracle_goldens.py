"""The oracle (oracle/refsem) against the reference's own outputs.

Every golden was produced by the unmodified reference (tests/golden/
make_goldens.py): sha256 of the reference's whole TriStore after each stage,
the .fin md5, the iteration count, RayHitted results for 6000 rays and
Shade's per-vertex increments for ~400 slot pairs per case."""
import hashlib
import os

import numpy as np
import pytest

import oracle_lib as ol

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


def oracle_for(entry, *extra):
    return ol.Oracle(*(entry["opts"] + ["-fma", "1" if entry["variant"] == "fma" else "0", "-quiet", "1",
                                        "-jobs", "4"] + list(extra)))


@pytest.mark.parametrize("key", case_keys())
def test_stages_match_reference(key, goldens, inputs):
    e = goldens["cases"][key]
    o = oracle_for(e)
    o.load_file(inputs[e["input"]])
    o.build_tree()
    nn, nl, _, _ = o.tree_info()
    assert (nn, nl) == (e["tree_nodes"], e["tree_lists"])
    rec, g = o.dump()
    assert sha(rec, g) == e["read"]["sha256"]
    o.init_radiosity()
    assert sha(*o.dump()) == e["init"]["sha256"]
    o.prepartition()
    rec, g = o.dump()
    assert len(rec) == e["prepart"]["trinum"]
    assert sha(rec, g) == e["prepart"]["sha256"]


@pytest.mark.parametrize("key", case_keys())
def test_full_run_matches_reference(key, goldens, inputs, tmp_path):
    e = goldens["cases"][key]
    o = oracle_for(e, "-write_cycle", "1000000")
    o.load_file(inputs[e["input"]])
    o.build_tree()
    it = o.run(str(tmp_path / "out"))
    assert it == e["iterations"]
    assert sha(*o.dump()) == e["final"]["sha256"]
    fin = open(tmp_path / "out.fin", "rb").read()
    assert hashlib.md5(fin).hexdigest() == e["fin_md5"]


@pytest.mark.parametrize("case", ["c1", "room_t5000", "box_t2000"])
def test_old_format_matches_reference(case, goldens, inputs, tmp_path):
    """The OLD_FORMAT build's per-vertex .fin (rad.cpp:150-159), from the
    reference compiled with its own switch (oracle/ref.mk rad_oldfmt)."""
    e = goldens["old_format"][case]
    o = ol.Oracle(*(e["opts"] + ["-old_format", "1", "-quiet", "1", "-write_cycle", "1000000"]))
    o.load_file(inputs[e["input"]])
    o.build_tree()
    o.run(str(tmp_path / "out"))
    fin = open(tmp_path / "out.fin", "rb").read()
    assert len(fin) == e["fin_bytes"]
    assert hashlib.md5(fin).hexdigest() == e["fin_md5"]


def test_neighbor_search_hash_equals_exact(inputs):
    """The hashed neighbour search reproduces the O(N^2) scan in loop order."""
    for m in ("box.tri", "room.tri", "blocks.tri", "c1_seed1.tri"):
        dumps = []
        for mode in ("exact", "hash"):
            o = ol.Oracle("-neighbors", mode, "-quiet", "1", "-triangle", "3000")
            o.load_file(inputs[m])
            o.build_tree()
            o.init_radiosity()
            dumps.append(o.dump()[0]["neighbor"].copy())
        assert np.array_equal(dumps[0], dumps[1]), m


@pytest.mark.parametrize("key", [k for k in case_keys()
                                 if os.path.exists(os.path.join(GOLDEN, "vectors_%s.npz" % k.replace("/", "_")))])
def test_rays_and_shade_match_reference(key, goldens, inputs):
    e = goldens["cases"][key]
    v = np.load(os.path.join(GOLDEN, "vectors_%s.npz" % key.replace("/", "_")))
    o = oracle_for(e)
    o.load_file(inputs[e["input"]])
    o.build_tree()
    o.init_radiosity()
    o.prepartition()
    rays = v["rays"].view(ol.RAY_DTYPE)
    hit, _, _ = o.raycast(rays)
    assert np.array_equal(hit, v["hits"])
    shade = o.shade_pairs(v["pairs"])
    assert shade.tobytes() == v["shade"].astype(np.float32).tobytes()


def test_fma_sites_matter(inputs):
    """Guards the contraction map: box differs between REF-FMA and REF-NOFMA
    (62 vs 86 iterations in the reference), so a wrong site shows up."""
    its = []
    for f in ("1", "0"):
        o = ol.Oracle("-triangle", "2000", "-fma", f, "-quiet", "1")
        o.load_file(inputs["box.tri"])
        o.build_tree()
        o.init_radiosity()
        o.prepartition()
        n = 0
        while True:
            d, c = o.iterate(1000)
            n += d
            if c:
                break
        its.append(n + 1)
    assert its == [62, 86]


def v3_keys():
    import json
    with open(os.path.join(GOLDEN, "goldens.json")) as fh:
        return sorted(json.load(fh)["v3"])


@pytest.mark.parametrize("key", v3_keys())
def test_v3_matches_reference(key, goldens, inputs, tmp_path):
    """The reference's V3 build (Delta-ff splitting live: config switches
    without DONT_PARTITION, oracle/ref.mk refdump_v3) at -jobs 1: iterations,
    the whole final TriStore (split topology, parents, neighbours, accB) and
    the .fin bytes.  On box and room pre-partition fills the pool, so V3
    equals V0 there."""
    e = goldens["v3"][key]
    o = ol.Oracle(*(e["opts"] + ["-variant", "v3", "-quiet", "1"]))
    o.load_file(inputs[e["input"]])
    o.build_tree()
    it = o.run(str(tmp_path / "out"))
    assert it == e["iterations"]
    rec, g = o.dump()
    assert len(rec) == e["final"]["trinum"]
    assert sha(rec, g) == e["final"]["sha256"]
    fin = open(tmp_path / "out.fin", "rb").read()
    assert hashlib.md5(fin).hexdigest() == e["fin_md5"]


def test_v3_input_is_the_generated_scene(goldens, inputs):
    import hashlib as h
    assert h.sha256(open(inputs["v3room.tri"], "rb").read()).hexdigest() == goldens["generator"]["v3room.tri"]


def test_glibc_rand_restatement():
    """The oracle's rand() restatement (TYPE_3, default seed) against the C
    library's own rand() in a fresh process, and the first published outputs."""
    import subprocess
    import sys
    code = "import ctypes; l = ctypes.CDLL('libc.so.6'); print(' '.join(str(l.rand()) for _ in range(2000)))"
    libc = [int(x) for x in subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                                           check=True).stdout.split()]
    assert libc[:3] == [1804289383, 846930886, 1681692777]
    assert ol.rand_outputs(2000).tolist() == libc
