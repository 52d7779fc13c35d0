import gzip
import json
import os
import shutil
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "hw-radiosity_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs on the GPU box)")


@pytest.fixture(scope="session")
def goldens():
    with open(os.path.join(GOLDEN, "goldens.json")) as fh:
        return json.load(fh)


@pytest.fixture(scope="session")
def inputs(tmp_path_factory):
    """Input files of the golden cases, materialised in a temp dir."""
    from radgpu import scenes
    d = tmp_path_factory.mktemp("inputs")
    out = {}
    for m in ("box.tri", "room.tri", "blocks.tri"):
        dst = str(d / m)
        with gzip.open(os.path.join(GOLDEN, "models", m + ".gz"), "rb") as fi, open(dst, "wb") as fo:
            shutil.copyfileobj(fi, fo)
        out[m] = dst
    c1 = str(d / "c1_seed1.tri")
    scenes.write_tri(c1, scenes.room_scene(1))
    out["c1_seed1.tri"] = c1
    v3 = str(d / "v3room.tri")
    scenes.write_tri(v3, scenes.v3_room())
    out["v3room.tri"] = v3
    return out
