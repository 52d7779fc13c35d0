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
    with pytest.raises(radgpu.RadError) as ei:
        radgpu.Engine("-parallel_src", "-1")   # undefined behaviour in listCanditate (rad.cpp:400)
    assert ei.value.status == radgpu.RAD_EINVAL
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


def test_run_without_a_device_is_a_device_error(tmp_path):
    """The product path has no CPU fallback: with no HIP device the octree
    build, and then the run, stop with RAD_EDEVICE (the device start-up
    thread does not start) and write no .fin."""
    import pytest
    from radgpu import scenes
    probe = subprocess.run(["bash", "-c", "ls /dev/kfd"], capture_output=True)
    if probe.returncode == 0:
        pytest.skip("a GPU device node is present")
    with pytest.raises(radgpu.RadError) as ei:   # the octree is built on the device
        radgpu.Engine("-quiet", "1").set_scene(scenes.room_scene(1))
    assert ei.value.status == radgpu.RAD_EDEVICE and "GPU only" in str(ei.value)
    with radgpu.Engine("-quiet", "1", "-octree", "host") as e:
        e.set_scene(scenes.room_scene(1))
        with pytest.raises(radgpu.RadError) as ei:   # so is the neighbour search, and the rest
            e.run(str(tmp_path / "out"))
        assert ei.value.status == radgpu.RAD_EDEVICE and "GPU only" in str(ei.value)
    assert not os.path.exists(str(tmp_path / "out.fin"))


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus above 1 exits with a message before
    touching a device."""
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr and r.stdout == ""
