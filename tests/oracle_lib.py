"""ctypes binding of the oracle (oracle/_bin/librefsem.so).

Test infrastructure: the oracle is the checker, never the thing measured or
shipped.  Record layout = the reference Triangle (rad.h:25-39)."""
from __future__ import annotations

import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "_bin", "librefsem.so")

REC_DTYPE = np.dtype([
    ("p", "<f4", (3, 3)), ("Frgb", "<i4", 3), ("Brgb", "<i4", 3),
    ("accB", "<f4", (3, 3)), ("deltaaccB", "<f4", (3, 3)), ("deltaB", "<f4", 3),
    ("se", "<f4", (3, 4)), ("c", "<f4", 3), ("n", "<f4", 3), ("area", "<f4"),
    ("parent", "<i4"), ("neighbor", "<i4", 3)])
assert REC_DTYPE.itemsize == 236
RAY_DTYPE = np.dtype([("p", "<f4", 3), ("v", "<f4", 3), ("ban", "<i4"), ("target", "<i4")])
NODE_DTYPE = np.dtype([("g0", "<f4", 3), ("g1", "<f4", 3), ("sub", "<i4", 8), ("list", "<i4")])

VIS_SEGMENT, VIS_REFERENCE, VIS_NONE, VIS_BRUTE = 0, 1, 2, 3
CNT_RAYS, CNT_CELLS, CNT_TESTS, CNT_LEVELS, CNT_PRECISION, CNT_SHADES, CNT_ITERS = range(7)

_lib = None


def lib():
    global _lib
    if _lib is None:
        L = ctypes.CDLL(LIB_PATH)
        vp, ip, cp = ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p
        L.rs_new.restype = vp
        L.rs_new.argtypes = [ip, ctypes.POINTER(cp)]
        L.rs_free.argtypes = [vp]
        L.rs_error.restype = cp
        L.rs_error.argtypes = [vp]
        for name, args in [("rs_load_file", [vp, cp]), ("rs_load_arrays", [vp, ip, vp, vp, vp]),
                           ("rs_build_tree", [vp]), ("rs_init_radiosity", [vp]), ("rs_prepartition", [vp]),
                           ("rs_iterate", [vp, ip, ctypes.POINTER(ip)]), ("rs_run", [vp, cp]),
                           ("rs_write_fin", [vp, cp]), ("rs_trinum", [vp]), ("rs_dump", [vp, vp, vp]),
                           ("rs_tree_info", [vp, vp, ctypes.POINTER(ctypes.c_float)]),
                           ("rs_descend", [vp, ip, ip, ip]), ("rs_tree_arrays", [vp, vp, vp]),
                           ("rs_raycast", [vp, ip, vp, vp, vp, vp]), ("rs_shade_pairs", [vp, ip, vp, vp]),
                           ("rs_ff_pairs", [vp, ip, vp, vp]), ("rs_ff_centroid", [vp, ip, ip, vp, vp]),
                           ("rs_ff_rows", [vp, ip, ip, vp, vp])]:
            f = getattr(L, name)
            f.restype = ip
            f.argtypes = args
        L.rs_neighbor_rows.argtypes = [vp, ip]
        L.rs_neighbor_rows.restype = ctypes.c_longlong
        L.rs_counters.argtypes = [vp, vp]
        L.rs_reset_counters.argtypes = [vp]
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class Oracle:
    """One oracle context.  ``opts``: the rad options plus oracle switches."""

    def __init__(self, *opts: str):
        L = lib()
        argv = (ctypes.c_char_p * max(1, len(opts)))(*[o.encode() for o in opts])
        self.h = L.rs_new(len(opts), argv)
        if not self.h:
            raise ValueError("bad oracle options %r" % (opts,))

    def __del__(self):
        if getattr(self, "h", None):
            lib().rs_free(self.h)
            self.h = None

    def _check(self, rc):
        if rc < 0:
            raise RuntimeError(lib().rs_error(self.h).decode())
        return rc

    def load_file(self, path):
        return self._check(lib().rs_load_file(self.h, path.encode()))

    def load(self, scene):
        p = np.ascontiguousarray(scene.p, np.float32)
        f = np.ascontiguousarray(scene.frgb, np.int32)
        b = np.ascontiguousarray(scene.brgb, np.int32)
        return self._check(lib().rs_load_arrays(self.h, scene.n, _ptr(p), _ptr(f), _ptr(b)))

    def build_tree(self):
        return lib().rs_build_tree(self.h)

    def init_radiosity(self):
        return lib().rs_init_radiosity(self.h)

    def prepartition(self):
        return lib().rs_prepartition(self.h)

    def iterate(self, max_shots):
        conv = ctypes.c_int(0)
        done = lib().rs_iterate(self.h, max_shots, ctypes.byref(conv))
        return done, bool(conv.value)

    def run(self, prefix):
        return self._check(lib().rs_run(self.h, prefix.encode()))

    def write_fin(self, path):
        return self._check(lib().rs_write_fin(self.h, path.encode()))

    @property
    def trinum(self):
        return lib().rs_trinum(self.h)

    def dump(self):
        n = self.trinum
        rec = np.zeros(n, REC_DTYPE)
        g = np.zeros(6, np.float32)
        lib().rs_dump(self.h, _ptr(rec), _ptr(g))
        return rec, g

    def tree_info(self):
        out = np.zeros(3, np.int32)
        mgl = ctypes.c_float(0)
        lib().rs_tree_info(self.h, _ptr(out), ctypes.byref(mgl))
        return int(out[0]), int(out[1]), int(out[2]), mgl.value

    def tree_arrays(self):
        nn, nl, _, _ = self.tree_info()
        nodes = np.zeros(nn, NODE_DTYPE)
        lists = np.zeros(nl, np.int32)
        lib().rs_tree_arrays(self.h, _ptr(nodes), _ptr(lists))
        return nodes, lists

    def descend(self, gx, gy, gz):
        return lib().rs_descend(self.h, gx, gy, gz)

    def raycast(self, rays):
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        n = len(rays)
        hit, cells, tests = (np.zeros(n, np.int32) for _ in range(3))
        lib().rs_raycast(self.h, n, _ptr(rays), _ptr(hit), _ptr(cells), _ptr(tests))
        return hit, cells, tests

    def shade_pairs(self, pairs):
        pairs = np.ascontiguousarray(pairs, np.int32)
        out = np.zeros((len(pairs), 9), np.float32)
        lib().rs_shade_pairs(self.h, len(pairs), _ptr(pairs), _ptr(out))
        return out

    def ff_pairs(self, pairs):
        pairs = np.ascontiguousarray(pairs, np.int32)
        out = np.zeros((len(pairs), 3), np.float32)
        lib().rs_ff_pairs(self.h, len(pairs), _ptr(pairs), _ptr(out))
        return out

    def ff_centroid(self, pairs, vis=VIS_SEGMENT):
        """Unscaled gathering form factors F[i][j] for (i, j) pairs."""
        pairs = np.ascontiguousarray(pairs, np.int32)
        out = np.zeros(len(pairs), np.float32)
        lib().rs_ff_centroid(self.h, int(vis), len(pairs), _ptr(pairs), _ptr(out))
        return out

    def ff_rows(self, rows, vis=VIS_SEGMENT):
        """Rows of the explicit matrix (scaled to sum <= 1), shape (len(rows), trinum)."""
        rows = np.ascontiguousarray(rows, np.int32)
        out = np.zeros((len(rows), self.trinum), np.float32)
        lib().rs_ff_rows(self.h, int(vis), len(rows), _ptr(rows), _ptr(out))
        return out

    def neighbor_rows(self, rows):
        return lib().rs_neighbor_rows(self.h, int(rows))

    def counters(self):
        out = np.zeros(8, np.int64)
        lib().rs_counters(self.h, _ptr(out))
        return out


def rand_outputs(n):
    """glibc rand()'s first n outputs as the oracle restates them."""
    out = np.zeros(n, np.int32)
    lib().rs_rand(n, _ptr(out))
    return out


def read_dump(path):
    """Parse a state dump written by oracle/refdump (the reference harness)."""
    raw = np.fromfile(path, np.uint8)
    hdr = raw[:8].view("<i4")
    assert hdr[0] == 0x504D4452
    n = int(hdr[1])
    g = raw[8:32].view("<f4").copy()
    assert int(raw[32:36].view("<i4")[0]) == 236
    rec = raw[36:36 + 236 * n].view(REC_DTYPE).copy()
    return rec, g
