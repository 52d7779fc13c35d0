"""The C ABI library loads without a GPU and exports every declared symbol."""
import os
import re
import subprocess

import radgpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "radiosity_gpu.h")).read()
    return sorted(set(re.findall(r"\b(rad_[a-z_0-9]+)\s*\(", text)))


def test_exports_every_declared_symbol():
    names = declared()
    assert set(names) == set(radgpu.EXPORTS)
    lib = radgpu.library()
    for n in names:
        assert hasattr(lib, n), n
    out = subprocess.run(["nm", "-D", "--defined-only", radgpu.LIB_PATH], capture_output=True, text=True).stdout
    for n in names:
        assert re.search(r"\bT %s$" % n, out, re.M), n


def test_errors_are_statuses_not_exits(tmp_path):
    import pytest
    with pytest.raises(radgpu.RadError):
        radgpu.Engine("-write_cycle")          # missing value
    with pytest.raises(radgpu.RadError):
        radgpu.Engine("-nonsense", "1")
    with radgpu.Engine("-quiet", "1") as e:
        with pytest.raises(radgpu.RadError) as ei:
            e.load_tri(str(tmp_path / "missing.tri"))
        assert ei.value.status == radgpu.RAD_EIO
        with pytest.raises(radgpu.RadError) as ei:
            e.prepare_host()
        assert ei.value.status == radgpu.RAD_ESTATE


def test_cli_usage_and_open_error(tmp_path):
    r = radgpu.run_cli([])
    assert r.returncode == 1 and "Usage" in r.stderr
    r = radgpu.run_cli([str(tmp_path / "none.tri"), "-o", str(tmp_path / "x")])
    assert r.returncode == 1 and "File open error" in r.stderr
