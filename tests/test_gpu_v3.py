"""The Delta-ff adaptive splitting of the reference's V3 build on the device
(-variant v3) against the reference itself (goldens from oracle/_ref/refdump_v3
at -jobs 1): iterations, the whole final TriStore -- split topology, parents,
neighbours, accB bits -- and the .fin bytes.  On box and room pre-partition
fills the pool, so no split can happen and V3 equals V0; the coarse room
splits hundreds of receivers, some after rand() samples."""
import hashlib
import json
import os

import numpy as np
import pytest

import radgpu

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def sha(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def v3_keys():
    with open(os.path.join(GOLDEN, "goldens.json")) as fh:
        return sorted(json.load(fh)["v3"])


@pytest.mark.parametrize("key", v3_keys())
def test_v3_matches_reference(key, goldens, inputs, tmp_path):
    e = goldens["v3"][key]
    print("v3 case", key, flush=True)
    with radgpu.Engine(*(e["opts"] + ["-variant", "v3", "-quiet", "1"])) as eng:
        eng.load_tri(inputs[e["input"]])
        it = eng.run(str(tmp_path / "gpu"))
        rec, g = eng.records()
    ref = np.load(os.path.join(GOLDEN, "final_%s.npz" % key))
    assert it == e["iterations"]
    assert len(rec) == e["final"]["trinum"]
    assert np.array_equal(rec["parent"], ref["parent"])
    assert rec["accB"].tobytes() == ref["accB"].tobytes()
    assert sha(rec, g) == e["final"]["sha256"]
    assert hashlib.md5(open(tmp_path / "gpu.fin", "rb").read()).hexdigest() == e["fin_md5"]


def test_v3_refuses_several_sources(inputs):
    with radgpu.Engine("-variant", "v3", "-parallel_src", "2", "-quiet", "1") as eng:
        eng.load_tri(inputs["v3room.tri"]).prepare_host()
        with pytest.raises(radgpu.RadError):
            eng.upload()
