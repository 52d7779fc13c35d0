"""The engine's host stages under AddressSanitizer and UndefinedBehavior-
Sanitizer: hw-radiosity_amd/sanitize.mk builds them with g++ (no device code)
into _bin/host_check, which reads a .tri as rad_load_tri does, runs the host
octree, InitRadiosity's neighbour search, pre-partition and the .fin and JSON
writers, or parses a given .fin.  Inputs: the reference's models and the
generated test scenes, and malformed files (truncated records, text where
numbers belong, coordinates that are not finite, a .fin cut mid-record).
Every run must end normally with no sanitizer report; the malformed inputs
must be refused with a message, not a crash."""
import gzip
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "hw-radiosity_amd")
BIN = os.path.join(PKG, "_bin", "host_check")
MODELS = os.path.join(ROOT, "tests", "golden", "models")


@pytest.fixture(scope="module")
def host_check():
    if shutil.which("g++") is None:
        pytest.skip("g++ not installed")
    r = subprocess.run(["make", "-C", PKG, "-f", "sanitize.mk"], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return BIN


def run(binary, *args):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:exitcode=99",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1:exitcode=98")
    r = subprocess.run([binary, *args], capture_output=True, text=True, timeout=300, env=env)
    text = r.stdout + r.stderr
    assert r.returncode == 0, "exit %d:\n%s" % (r.returncode, text[-4000:])
    assert "runtime error" not in text and "AddressSanitizer" not in text and "LeakSanitizer" not in text, text[-4000:]
    return r.stdout


@pytest.mark.parametrize("model,triangle", [("box.tri", 2000), ("room.tri", 5000), ("blocks.tri", 3000),
                                            ("hall.tri", 20000), ("church.tri", 20000)])
def test_models_clean(host_check, model, triangle, tmp_path):
    src = str(tmp_path / model)
    with gzip.open(os.path.join(MODELS, model + ".gz"), "rb") as fi, open(src, "wb") as fo:
        shutil.copyfileobj(fi, fo)
    out = run(host_check, "-triangle", str(triangle), src, str(tmp_path / "out"))
    assert out.startswith("triangles ") and "json rc 0" in out, out
    n = int(out.split()[1])
    assert n > 0
    # the .fin written is read back by the -fin path too
    out = run(host_check, "-fin", str(tmp_path / "out.fin"), str(tmp_path / "back.json"))
    assert out.startswith("fin_to_json rc 0"), out


def test_generated_scenes_clean(host_check, tmp_path):
    from radgpu import scenes
    for name, tris in (("c1", scenes.room_scene(1)), ("v3room", scenes.v3_room()),
                       ("hall", scenes.hall_scene(1, n_boxes=8))):
        p = str(tmp_path / (name + ".tri"))
        scenes.write_tri(p, tris)
        out = run(host_check, "-triangle", "20000", p, str(tmp_path / ("out_" + name)))
        assert "json rc 0" in out, out


MALFORMED_TRI = {
    "trunc.tri": ("Triangle\n1 2 3 4 5 6\n0 0 0\n1 0 0\n", "malformed"),
    "text.tri": ("Triangle\nxx yy\n", None),
    "nonfinite.tri": ("Triangle\n1 2 3 4 5 6\nnan 0 0\n1 inf 0\n0 1 1e39\n", "not finite"),
    "empty.tri": ("", None),
    "cut.tri": ("Triangle\n" + "1 1 1 0 0 0\n0 0 0\n1 0 0\n0 1 0\n" * 20 + "1 1 1 0 0 0\n0 0", None),
}


@pytest.mark.parametrize("name", sorted(MALFORMED_TRI))
def test_malformed_tri_refused_or_clean(host_check, name, tmp_path):
    body, expect = MALFORMED_TRI[name]
    p = tmp_path / name
    p.write_text(body)
    out = run(host_check, "-triangle", "2000", str(p), str(tmp_path / "out"))
    if expect:
        assert expect in out, out


@pytest.mark.parametrize("body", ["Triangle \n1 2 3 0 0 0\n0 0 0 0 0 1\nxx\n",
                                  "Triangle \n1 2 3 0 0 0\n0 0 0 0 0 1\n1 0 0 0 0",
                                  "", "Triangle \n"])
def test_malformed_fin_refused_or_clean(host_check, body, tmp_path):
    p = tmp_path / "bad.fin"
    p.write_text(body)
    out = run(host_check, "-fin", str(p), str(tmp_path / "bad.json"))
    assert out.startswith("fin_to_json rc "), out
