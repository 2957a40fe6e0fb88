"""ORACLE - test infrastructure only.

ctypes binding of ``oracle/liboracle.so``, the CPU restatement of the
reference pipeline (oracle.cpp / oracle_funcspace.h).  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg import this
module, as the checker; the product package never does.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

ORACLE_LIB = Path(__file__).resolve().parent / "liboracle.so"

_BUF = {  # prg_buffer ids -> (dtype, width)
    0: (np.float32, 3), 1: (np.float32, 3), 2: (np.int32, 1), 3: (np.uint32, 1), 4: (np.int32, 1),
    5: (np.int32, 1), 6: (np.int32, 1), 7: (np.int32, 8), 8: (np.int32, 27), 9: (np.int32, 1),
    10: (np.int32, 1), 11: (np.float32, 3), 12: (np.float32, 1), 13: (np.float32, 1), 14: (np.float32, 1),
    15: (np.float32, 8),
}

_lib: C.CDLL | None = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not ORACLE_LIB.exists():
            raise ImportError(f"{ORACLE_LIB} missing: run `make -C oracle`")
        L = C.CDLL(str(ORACLE_LIB))
        P = C.c_void_p
        sig = {
            "orc_create": (P, [C.c_int32]),
            "orc_destroy": (None, [P]),
            "orc_set_points": (C.c_int32, [P, P, P, C.c_int64]),
            "orc_get_transform": (None, [P, P, P]),
            "orc_build_octree": (C.c_int32, [P]),
            "orc_level_info": (None, [P, P, P]),
            "orc_compute_divergence": (C.c_int32, [P]),
            "orc_solve": (C.c_int32, [P, C.c_int32, C.c_float, C.c_int32, C.c_int32, C.c_int32, P, P]),
            "orc_solve_fp64": (C.c_int32, [P, C.c_double, C.c_int32, P]),
            "orc_isovalue": (C.c_float, [P]),
            "orc_set_isovalue": (None, [P, C.c_float]),
            "orc_extract_leaf_mesh": (C.c_int32, [P, P, P]),
            "orc_copy_mesh": (C.c_int32, [P, P, P]),
            "orc_extract_subdivision": (C.c_int32, [P, P, P]),
            "orc_extract_subdivision_from": (C.c_int32, [P, C.c_int32, P, P]),
            "orc_subdivision_info": (None, [P, P, P]),
            "orc_subdivision_roots": (C.c_int32, [P, P]),
            "orc_buffer_size": (C.c_int64, [P, C.c_int32]),
            "orc_read_buffer": (C.c_int32, [P, C.c_int32, P, C.c_int64]),
            "orc_write_buffer": (C.c_int32, [P, C.c_int32, P, C.c_int64]),
            "orc_fs_entry": (C.c_double, [C.c_int32, C.c_int32, C.c_int32, C.c_int64, C.c_int32, C.c_int64]),
            "orc_fs_base_function": (None, [C.c_int32, C.c_int64, P]),
            "orc_fs_base_maxdepth": (None, [C.c_int32, P]),
            "orc_fs_value": (C.c_float, [C.c_int32, C.c_int64, C.c_float]),
            "orc_fs_stencil": (None, [C.c_int32, C.c_int32, P]),
            "orc_num_threads": (C.c_int32, []),
            "orc_set_num_threads": (None, [C.c_int32]),
        }
        for k, (r, a) in sig.items():
            f = getattr(L, k)
            f.restype, f.argtypes = r, a
        _lib = L
    return _lib


def _p(a: np.ndarray) -> C.c_void_p:
    return C.c_void_p(a.ctypes.data)


class Oracle:
    """CPU restatement of the pipeline; mirrors poissonrecon_gpu_amd.Reconstructor."""

    def __init__(self, max_depth: int):
        self.max_depth = max_depth
        self._h = lib().orc_create(max_depth)
        if not self._h:
            raise ValueError("max_depth out of range")

    def close(self):
        if getattr(self, "_h", None):
            lib().orc_destroy(self._h)
            self._h = None

    __del__ = close

    def set_points(self, xyz, nxyz):
        xyz = np.ascontiguousarray(xyz, np.float32)
        nxyz = np.ascontiguousarray(nxyz, np.float32)
        if lib().orc_set_points(self._h, _p(xyz), _p(nxyz), len(xyz)) != 0:
            raise ValueError("bad point count")

    def transform(self):
        c = np.zeros(3, np.float32)
        s = np.zeros(1, np.float32)
        lib().orc_get_transform(self._h, _p(c), _p(s))
        return c, float(s[0])

    def build_octree(self):
        assert lib().orc_build_octree(self._h) == 0

    def level_info(self):
        c = np.zeros(self.max_depth + 1, np.int64)
        b = np.zeros(self.max_depth + 1, np.int64)
        lib().orc_level_info(self._h, _p(c), _p(b))
        return c, b

    def compute_divergence(self):
        assert lib().orc_compute_divergence(self._h) == 0

    def solve(self, mode=0, tol=1e-6, max_iters=2000, refresh=20, fixed_iters=0):
        it = np.zeros(self.max_depth + 1, np.int32)
        rr = np.zeros(self.max_depth + 1, np.float64)
        assert lib().orc_solve(self._h, mode, tol, max_iters, refresh, fixed_iters, _p(it), _p(rr)) == 0
        return it, rr

    def solve_fp64(self, tol=1e-12, max_iters=5000):
        c, _ = self.level_info()
        x = np.zeros(int(c.sum()), np.float64)
        assert lib().orc_solve_fp64(self._h, tol, max_iters, _p(x)) == 0
        return x

    def isovalue(self) -> float:
        return float(lib().orc_isovalue(self._h))

    def set_isovalue(self, iso: float):
        lib().orc_set_isovalue(self._h, iso)

    def extract_leaf_mesh(self):
        nv = np.zeros(1, np.int64)
        nt = np.zeros(1, np.int64)
        assert lib().orc_extract_leaf_mesh(self._h, _p(nv), _p(nt)) == 0
        v = np.zeros((int(nv[0]), 3), np.float32)
        t = np.zeros((int(nt[0]), 3), np.int32)
        lib().orc_copy_mesh(self._h, _p(v), _p(t))
        return v, t

    def extract_subdivision(self, min_root_depth=0):
        """Appends the coarse-leaf subdivision batches (those of the groups whose
        roots lie at min_root_depth or deeper); returns the whole mesh."""
        nv = np.zeros(1, np.int64)
        nt = np.zeros(1, np.int64)
        assert lib().orc_extract_subdivision_from(self._h, int(min_root_depth), _p(nv), _p(nt)) == 0
        v = np.zeros((int(nv[0]), 3), np.float32)
        t = np.zeros((int(nt[0]), 3), np.int32)
        lib().orc_copy_mesh(self._h, _p(v), _p(t))
        return v, t

    def subdivision_info(self):
        """(number of subdivided coarse leaves, number of virtual leaves)."""
        a = np.zeros(1, np.int64)
        b = np.zeros(1, np.int64)
        lib().orc_subdivision_info(self._h, _p(a), _p(b))
        return int(a[0]), int(b[0])

    def subdivision_roots(self):
        """Subdivision roots per depth 0..D (after extract_leaf_mesh; nothing is meshed)."""
        a = np.zeros(self.max_depth + 1, np.int64)
        assert lib().orc_subdivision_roots(self._h, _p(a)) == 0
        return a

    def read_buffer(self, which: int) -> np.ndarray:
        n = lib().orc_buffer_size(self._h, which)
        dt, w = _BUF[which]
        a = np.empty(n // np.dtype(dt).itemsize, dt)
        assert lib().orc_read_buffer(self._h, which, _p(a), n) == 0
        return a.reshape(-1, w) if w > 1 else a

    def write_buffer(self, which: int, a: np.ndarray):
        a = np.ascontiguousarray(a, _BUF[which][0])
        assert lib().orc_write_buffer(self._h, which, _p(a), a.nbytes) == 0


def fs_entry(max_depth, which, da, oa, db, ob) -> float:
    return lib().orc_fs_entry(max_depth, which, da, oa, db, ob)


def fs_stencil(max_depth, depth) -> np.ndarray:
    out = np.zeros(27, np.float32)
    lib().orc_fs_stencil(max_depth, depth, _p(out))
    return out


def fs_base_function(max_depth, idx) -> np.ndarray:
    out = np.zeros(20, np.float32)
    lib().orc_fs_base_function(max_depth, idx, _p(out))
    return out.reshape(4, 5)


def fs_base_maxdepth(max_depth) -> np.ndarray:
    out = np.zeros(16, np.float32)
    lib().orc_fs_base_maxdepth(max_depth, _p(out))
    return out.reshape(4, 4)


def fs_value(max_depth, idx, x) -> float:
    return float(lib().orc_fs_value(max_depth, idx, x))


def num_threads() -> int:
    return int(lib().orc_num_threads())


def set_num_threads(n: int) -> None:
    lib().orc_set_num_threads(n)
