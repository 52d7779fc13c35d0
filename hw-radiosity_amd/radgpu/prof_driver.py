"""Profiling driver: one hall-200k solve, then a plain 256 MiB device copy
that calibrates the FETCH_SIZE / WRITE_SIZE counters in the same run.

    rocprofv3 --pmc FETCH_SIZE -d <dir> -o p -- python3 hw-radiosity_amd/radgpu/prof_driver.py
"""
import ctypes
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import radgpu  # noqa: E402
from radgpu import scenes  # noqa: E402

COPY_BYTES = 256 << 20


def main():
    s = scenes.hall_scene(2)
    e = radgpu.Engine("-converge", "1200", "-triangle", "1000000", "-quiet", "1")
    e.set_scene(s).prepare_host().upload()
    n, conv = e.shoot()
    print("shots", n, conv, e.counters(), flush=True)
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(COPY_BYTES)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(COPY_BYTES)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(COPY_BYTES)) == 0
    assert hip.hipMemcpy(b, a, ctypes.c_size_t(COPY_BYTES), 3) == 0  # device to device
    assert hip.hipDeviceSynchronize() == 0
    print("copy of %d bytes done" % COPY_BYTES, flush=True)


if __name__ == "__main__":
    main()
