"""Python binding of the MI355X radiosity engine (libradgpu.so, C ABI in
include/radiosity_gpu.h).

Mirrors the reference's path (morris821028/hw-radiosity): load a ``.tri``
scene (InitModelFromFile, rad.cpp:244-252), prepare (InitRadiosity +
PrePartitionTriangles), shoot on the device (parallelRadiosity loop,
rad.cpp:406-451), write ``.fin`` (PrintOut, rad.cpp:117-242).  The compute
runs in the HIP library only; if the library is missing this raises.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.environ.get("RADGPU_LIB") or os.path.join(PKG_DIR, "_bin", "libradgpu.so")
RAD_BIN = os.path.join(PKG_DIR, "_bin", "rad")

RAD_OK, RAD_EINVAL, RAD_EIO, RAD_ESTATE, RAD_EDEVICE = range(5)
(CNT_RAYS, CNT_REF_RAYS, CNT_CELLS, CNT_TESTS, CNT_LEVELS, CNT_PRECISION,
 CNT_STALLED, CNT_RECEIVERS) = range(8)
COUNTER_NAMES = ("rays", "ref_rays", "cells", "tests", "levels", "precision", "stalled", "receivers")

# The reference Triangle record (rad.h:25-39), 236 bytes.
RECORD_DTYPE = np.dtype([
    ("p", "<f4", (3, 3)), ("Frgb", "<i4", 3), ("Brgb", "<i4", 3),
    ("accB", "<f4", (3, 3)), ("deltaaccB", "<f4", (3, 3)), ("deltaB", "<f4", 3),
    ("se", "<f4", (3, 4)), ("c", "<f4", 3), ("n", "<f4", 3), ("area", "<f4"),
    ("parent", "<i4"), ("neighbor", "<i4", 3)])
assert RECORD_DTYPE.itemsize == 236
# RayHitted's arguments (raycast.cpp:335).
RAY_DTYPE = np.dtype([("p", "<f4", 3), ("v", "<f4", 3), ("ban", "<i4"), ("target", "<i4")])

EXPORTS = ("rad_create", "rad_destroy", "rad_last_error", "rad_load_tri", "rad_set_triangles",
           "rad_prepare_host", "rad_upload", "rad_shoot", "rad_run", "rad_write_fin", "rad_fin_to_json", "rad_records",
           "rad_tree_info", "rad_tree_arrays", "rad_cast", "rad_cast_static", "rad_ff_pairs", "rad_counters", "rad_timings",
           "rad_lin_create", "rad_lin_destroy", "rad_lin_last_error", "rad_lin_from_scene",
           "rad_lin_set_matrix", "rad_lin_get_matrix", "rad_lin_set_areas", "rad_lin_autotune", "rad_lin_set_tile",
           "rad_lin_gemm_ms", "rad_lin_multiply", "rad_lin_gather", "rad_lin_shoot", "rad_lin_adjoint",
           "rad_batch_create", "rad_batch_destroy", "rad_batch_last_error", "rad_batch_from_scene",
           "rad_batch_set_matrix", "rad_batch_autotune", "rad_batch_set_tile", "rad_batch_solve")

PREC_F32, PREC_F16, PREC_BF16 = 0, 1, 2
CAST_STATIC_MAX = 1 << 18
# TreeNode (rad.h:41-45)
NODE_DTYPE = np.dtype([("g0", "<f4", 3), ("g1", "<f4", 3), ("sub", "<i4", 8), ("list", "<i4")])
VIS_SEGMENT, VIS_REFERENCE = 0, 1
PRECISIONS = {"f32": PREC_F32, "f16": PREC_F16, "bf16": PREC_BF16}
INFO_NAMES = ("outer", "sweeps", "rel_res", "ms", "ms_target", "err", "power_E", "power_B")


class RadError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("status %d: %s" % (status, msg))
        self.status = status


_lib = None


def library() -> ctypes.CDLL:
    """Load libradgpu.so (build it with ``make -C hw-radiosity_amd``)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libradgpu.so is not built: run `make -C hw-radiosity_amd` (%s)" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        vp, ip, cp, dp = ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_double
        sig = {
            "rad_create": [ctypes.POINTER(vp), ip, ctypes.POINTER(cp)],
            "rad_destroy": [vp],
            "rad_load_tri": [vp, cp],
            "rad_set_triangles": [vp, ip, vp, vp, vp],
            "rad_prepare_host": [vp],
            "rad_upload": [vp],
            "rad_shoot": [vp, ip, ctypes.POINTER(ip), ctypes.POINTER(ip)],
            "rad_run": [vp, cp, ctypes.POINTER(ip)],
            "rad_write_fin": [vp, cp],
            "rad_fin_to_json": [vp, cp, cp, ctypes.POINTER(ip)],
            "rad_records": [vp, ctypes.POINTER(ip), vp, vp],
            "rad_tree_info": [vp, vp, ctypes.POINTER(ctypes.c_float)],
            "rad_tree_arrays": [vp, vp, vp],
            "rad_cast": [vp, ip, vp, vp, vp, vp],
            "rad_cast_static": [vp, ip, vp, vp, vp, vp],
            "rad_ff_pairs": [vp, ip, vp, vp],
            "rad_counters": [vp, vp],
            "rad_timings": [vp, vp],
            "rad_lin_create": [ctypes.POINTER(vp), ip],
            "rad_lin_destroy": [vp],
            "rad_lin_from_scene": [vp, vp, ip, vp],
            "rad_lin_set_matrix": [vp, vp],
            "rad_lin_get_matrix": [vp, vp],
            "rad_lin_set_areas": [vp, vp],
            "rad_lin_autotune": [vp, ip, ip, ctypes.POINTER(ip), ctypes.POINTER(ip)],
            "rad_lin_gemm_ms": [vp, ip, ip, ip, vp],
            "rad_lin_set_tile": [vp, ip, ip, ip, ip],
            "rad_lin_multiply": [vp, ip, ip, vp, vp],
            "rad_lin_gather": [vp, ip, ip, vp, vp, dp, ip, ip, vp, dp, vp, vp],
            "rad_lin_shoot": [vp, ip, vp, vp, vp, dp, ip, vp, vp],
            "rad_lin_adjoint": [vp, ip, vp, vp, vp, dp, ip, ip, vp, vp, vp],
            "rad_batch_create": [ctypes.POINTER(vp), ip, vp],
            "rad_batch_destroy": [vp],
            "rad_batch_from_scene": [vp, ip, vp, ip, vp],
            "rad_batch_set_matrix": [vp, ip, vp],
            "rad_batch_autotune": [vp, ip, vp, vp, vp],
            "rad_batch_set_tile": [vp, ip, ip, ip],
            "rad_batch_solve": [vp, ip, ip, vp, vp, dp, ip, ip, vp, vp, vp],
        }
        for name, args in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = None if name in ("rad_destroy", "rad_lin_destroy", "rad_batch_destroy") else ip
        for name in ("rad_last_error", "rad_lin_last_error", "rad_batch_last_error"):
            getattr(L, name).argtypes = [vp]
            getattr(L, name).restype = cp
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class Engine:
    """One engine context.  ``opts`` are rad options, e.g.
    ``Engine("-triangle", "20000", "-converge", "1200")``."""

    def __init__(self, *opts: str):
        L = library()
        argv = (ctypes.c_char_p * max(1, len(opts)))(*[o.encode() for o in opts])
        h = ctypes.c_void_p()
        st = L.rad_create(ctypes.byref(h), len(opts), argv)
        if st != RAD_OK:
            raise RadError(st, "bad options %r" % (opts,))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            library().rad_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _ok(self, st):
        if st != RAD_OK:
            raise RadError(st, library().rad_last_error(self.h).decode())

    # ---- stages -------------------------------------------------------
    def load_tri(self, path: str):
        self._ok(library().rad_load_tri(self.h, path.encode()))
        return self

    def set_scene(self, scene):
        p = np.ascontiguousarray(scene.p, np.float32)
        f = np.ascontiguousarray(scene.frgb, np.int32)
        b = np.ascontiguousarray(scene.brgb, np.int32)
        self._ok(library().rad_set_triangles(self.h, int(p.shape[0]), _p(p), _p(f), _p(b)))
        return self

    def prepare_host(self):
        self._ok(library().rad_prepare_host(self.h))
        return self

    def upload(self):
        self._ok(library().rad_upload(self.h))
        return self

    def shoot(self, max_shots: int = 1 << 30):
        shots, conv = ctypes.c_int(0), ctypes.c_int(0)
        self._ok(library().rad_shoot(self.h, int(max_shots), ctypes.byref(shots), ctypes.byref(conv)))
        return shots.value, bool(conv.value)

    def run(self, prefix: str) -> int:
        it = ctypes.c_int(0)
        self._ok(library().rad_run(self.h, prefix.encode(), ctypes.byref(it)))
        return it.value

    def write_fin(self, path: str):
        self._ok(library().rad_write_fin(self.h, path.encode()))

    def fin_to_json(self, fin_path: str, json_path: str) -> int:
        """The viewer's model file from a .fin (the reference's tri2json);
        returns the number of triangles converted."""
        n = ctypes.c_int(0)
        self._ok(library().rad_fin_to_json(self.h, fin_path.encode(), json_path.encode(), ctypes.byref(n)))
        return n.value

    # ---- state / checks ------------------------------------------------
    def records(self):
        n = ctypes.c_int(0)
        self._ok(library().rad_records(self.h, ctypes.byref(n), None, None))
        rec = np.zeros(n.value, RECORD_DTYPE)
        g = np.zeros(6, np.float32)
        self._ok(library().rad_records(self.h, ctypes.byref(n), _p(rec), _p(g)))
        return rec, g

    def tree_info(self):
        out = np.zeros(3, np.int32)
        mgl = ctypes.c_float(0)
        self._ok(library().rad_tree_info(self.h, _p(out), ctypes.byref(mgl)))
        return int(out[0]), int(out[1]), int(out[2]), mgl.value

    def tree_arrays(self):
        """(nodes, lists): the octree as the reference's TreeNodeStore
        (NODE_DTYPE records, depth-first preorder) and TriListStore."""
        nn, nl, _, _ = self.tree_info()
        nodes = np.zeros(nn, NODE_DTYPE)
        lists = np.zeros(nl, np.int32)
        self._ok(library().rad_tree_arrays(self.h, _p(nodes), _p(lists)))
        return nodes, lists

    def cast(self, rays):
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        n = len(rays)
        hit, cells, tests = (np.zeros(n, np.int32) for _ in range(3))
        self._ok(library().rad_cast(self.h, n, _p(rays), _p(hit), _p(cells), _p(tests)))
        return hit, cells, tests

    def cast_static(self, rays):
        """cast() through rad_cast_static (no device allocation; at most
        CAST_STATIC_MAX rays)."""
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        n = len(rays)
        hit, cells, tests = (np.zeros(n, np.int32) for _ in range(3))
        self._ok(library().rad_cast_static(self.h, n, _p(rays), _p(hit), _p(cells), _p(tests)))
        return hit, cells, tests

    def ff_pairs(self, pairs):
        pairs = np.ascontiguousarray(pairs, np.int32)
        out = np.zeros((len(pairs), 3), np.float32)
        self._ok(library().rad_ff_pairs(self.h, len(pairs), _p(pairs), _p(out)))
        return out

    def counters(self) -> dict:
        out = np.zeros(8, np.int64)
        self._ok(library().rad_counters(self.h, _p(out)))
        return dict(zip(COUNTER_NAMES, (int(x) for x in out)))

    def timings(self) -> dict:
        out = np.zeros(7, np.float64)
        self._ok(library().rad_timings(self.h, _p(out)))
        return {"setup_ms": out[0], "upload_ms": out[1], "shoot_ms": out[2], "iterations_launched": int(out[3]),
                "kernel_ms": out[4], "kernels_timed": int(out[5]), "cast_ms": out[6]}


def _prec(p):
    return PRECISIONS[p] if isinstance(p, str) else int(p)


def _rhs(a, n):
    """n x nrhs float64, C-contiguous (a 1-D array is one column)."""
    a = np.asarray(a, np.float64)
    if a.ndim == 1:
        a = a[:, None]
    if a.shape[0] != n:
        raise ValueError("expected %d rows, got %s" % (n, a.shape))
    return np.ascontiguousarray(a)


class LinearSystem:
    """The radiosity system B = E + diag(rho) F B on an explicit form-factor
    matrix held on the device (rad_lin_* in include/radiosity_gpu.h).

    Right-hand sides are (n, nrhs) float64 arrays; rho has the same shape
    (per patch and column, e.g. one column per RGB channel)."""

    def __init__(self, n: int):
        L = library()
        h = ctypes.c_void_p()
        st = L.rad_lin_create(ctypes.byref(h), int(n))
        if st != RAD_OK:
            raise RadError(st, "rad_lin_create(%d) failed" % n)
        self.h, self.n = h, int(n)

    def close(self):
        if getattr(self, "h", None):
            library().rad_lin_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _ok(self, st):
        if st != RAD_OK:
            raise RadError(st, library().rad_lin_last_error(self.h).decode())

    def from_scene(self, engine: "Engine", visibility: str = "segment") -> float:
        """F and areas from an uploaded scene; returns the kernel's ms.
        visibility: "segment" (clear segment between centroids) or
        "reference" (RayHitted from the centroid)."""
        ms = ctypes.c_double(0)
        vis = {"segment": VIS_SEGMENT, "reference": VIS_REFERENCE}[visibility]
        self._ok(library().rad_lin_from_scene(self.h, engine.h, vis, ctypes.byref(ms)))
        return ms.value

    def set_matrix(self, F):
        F = np.ascontiguousarray(F, np.float32)
        assert F.shape == (self.n, self.n)
        self._ok(library().rad_lin_set_matrix(self.h, _p(F)))
        return self

    def matrix(self):
        F = np.zeros((self.n, self.n), np.float32)
        self._ok(library().rad_lin_get_matrix(self.h, _p(F)))
        return F

    def set_areas(self, area):
        a = np.ascontiguousarray(area, np.float64)
        assert a.shape == (self.n,)
        self._ok(library().rad_lin_set_areas(self.h, _p(a)))
        return self

    def autotune(self, nrhs: int, prec="f16"):
        rt, sk = ctypes.c_int(0), ctypes.c_int(0)
        self._ok(library().rad_lin_autotune(self.h, int(nrhs), _prec(prec), ctypes.byref(rt), ctypes.byref(sk)))
        return rt.value, sk.value

    def set_tile(self, nrhs: int, prec, rt: int, splitk: int):
        self._ok(library().rad_lin_set_tile(self.h, int(nrhs), _prec(prec), int(rt), int(splitk)))
        return self

    def gemm_ms(self, nrhs: int, prec="f16", reps: int = 20) -> float:
        ms = np.zeros(1, np.float64)
        self._ok(library().rad_lin_gemm_ms(self.h, int(nrhs), _prec(prec), int(reps), _p(ms)))
        return float(ms[0])

    def multiply(self, X, prec="f32"):
        X = _rhs(X, self.n)
        Y = np.zeros_like(X)
        self._ok(library().rad_lin_multiply(self.h, _prec(prec), X.shape[1], _p(X), _p(Y)))
        return Y

    def gather(self, rho, E, prec="f16", tol=1e-6, max_outer=50, sweeps=8, Bref=None, target=1e-3):
        E = _rhs(E, self.n)
        rho = _rhs(rho, self.n)
        assert rho.shape == E.shape
        ref = None if Bref is None else _rhs(Bref, self.n)
        B = np.zeros_like(E)
        info = np.zeros(8, np.float64)
        self._ok(library().rad_lin_gather(self.h, _prec(prec), E.shape[1], _p(rho), _p(E), float(tol),
                                          int(max_outer), int(sweeps), _p(ref), float(target), _p(B), _p(info)))
        return B, dict(zip(INFO_NAMES, info.tolist()))

    def shoot(self, rho, E, Bref=None, target=1e-3, max_shots=1 << 30):
        E = _rhs(E, self.n)
        rho = _rhs(rho, self.n)
        assert rho.shape == E.shape
        ref = None if Bref is None else _rhs(Bref, self.n)
        B = np.zeros_like(E)
        info = np.zeros(8, np.float64)
        self._ok(library().rad_lin_shoot(self.h, E.shape[1], _p(rho), _p(E), _p(ref), float(target),
                                         int(max_shots), _p(B), _p(info)))
        return B, dict(zip(INFO_NAMES, info.tolist()))

    def adjoint(self, rho, B, dLdB, tol=1e-10, max_outer=50, sweeps=8):
        """Returns (dL/dE, dL/drho, info)."""
        B = _rhs(B, self.n)
        rho = _rhs(rho, self.n)
        g = _rhs(dLdB, self.n)
        assert rho.shape == B.shape == g.shape
        dE, drho = np.zeros_like(B), np.zeros_like(B)
        info = np.zeros(8, np.float64)
        self._ok(library().rad_lin_adjoint(self.h, B.shape[1], _p(rho), _p(B), _p(g), float(tol), int(max_outer),
                                           int(sweeps), _p(dE), _p(drho), _p(info)))
        return dE, drho, dict(zip(INFO_NAMES, info.tolist()))


class SceneBatch:
    """count independent RGB systems of sizes n[s], solved together
    (rad_batch_* in include/radiosity_gpu.h)."""

    def __init__(self, sizes):
        L = library()
        self.sizes = [int(x) for x in sizes]
        arr = np.ascontiguousarray(self.sizes, np.int32)
        h = ctypes.c_void_p()
        st = L.rad_batch_create(ctypes.byref(h), len(arr), _p(arr))
        if st != RAD_OK:
            raise RadError(st, "rad_batch_create failed")
        self.h = h
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)]).astype(np.int64)

    def close(self):
        if getattr(self, "h", None):
            library().rad_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _ok(self, st):
        if st != RAD_OK:
            raise RadError(st, library().rad_batch_last_error(self.h).decode())

    def from_scene(self, s: int, engine: "Engine", visibility: str = "segment") -> float:
        ms = ctypes.c_double(0)
        vis = {"segment": VIS_SEGMENT, "reference": VIS_REFERENCE}[visibility]
        self._ok(library().rad_batch_from_scene(self.h, int(s), engine.h, vis, ctypes.byref(ms)))
        return ms.value

    def set_matrix(self, s: int, F):
        F = np.ascontiguousarray(F, np.float32)
        assert F.shape == (self.sizes[s], self.sizes[s])
        self._ok(library().rad_batch_set_matrix(self.h, int(s), _p(F)))

    def autotune(self, prec="f16"):
        """Times every sweep-product schedule; returns (rt, kc, grid)."""
        v = [ctypes.c_int(0) for _ in range(3)]
        self._ok(library().rad_batch_autotune(self.h, _prec(prec), *[ctypes.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def set_tile(self, rt, kc, grid):
        self._ok(library().rad_batch_set_tile(self.h, int(rt), int(kc), int(grid)))

    def solve(self, rho, E, prec="f16", graph=True, tol=1e-6, outer=4, sweeps=8):
        """rho, E: (sum n, 3).  Returns (B, rel_res per scene, info)."""
        rho = np.ascontiguousarray(rho, np.float64)
        E = np.ascontiguousarray(E, np.float64)
        assert rho.shape == E.shape == (int(self.offsets[-1]), 3)
        B = np.zeros_like(E)
        rel = np.zeros(len(self.sizes), np.float64)
        info = np.zeros(4, np.float64)
        self._ok(library().rad_batch_solve(self.h, _prec(prec), 1 if graph else 0, _p(rho), _p(E), float(tol),
                                           int(outer), int(sweeps), _p(B), _p(rel), _p(info)))
        return B, rel, {"ms": info[0], "build_ms": info[1], "launches": int(info[2]), "above_tol": int(info[3])}


def run_cli(args, cwd=None, timeout=None):
    """Run the drop-in ``rad`` binary; returns the CompletedProcess."""
    if not os.path.exists(RAD_BIN):
        raise ImportError("rad is not built: run `make -C hw-radiosity_amd`")
    return subprocess.run([RAD_BIN] + list(args), cwd=cwd, capture_output=True, text=True, timeout=timeout)
