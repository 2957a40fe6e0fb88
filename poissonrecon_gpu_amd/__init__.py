"""MI355X-native Poisson surface reconstruction.

Python mirror of the C ABI in ``include/prg.h`` (ctypes over the in-tree
``libprg.so``).  The stage methods keep the reference's stage boundaries
(main.cu of DavidXu-JJ/PoissonRecon_GPU): octree build, divergence,
per-depth Laplacian solve, isovalue, marching cubes.  There is no CPU
fallback: every method runs the HIP library, and importing this package
raises if the library is missing.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from pathlib import Path

import numpy as np

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / "libprg.so"

# prg_status
OK, E_ARG, E_NOMEM, E_HIP, E_IO, E_NOT_CONVERGED, E_STATE, E_NODEVICE = range(8)
STATUS_NAMES = {0: "PRG_OK", 1: "PRG_E_ARG", 2: "PRG_E_NOMEM", 3: "PRG_E_HIP", 4: "PRG_E_IO",
                5: "PRG_E_NOT_CONVERGED", 6: "PRG_E_STATE", 7: "PRG_E_NODEVICE"}

# prg_buffer
BUF_POINTS, BUF_NORMALS, BUF_POINT_LEAF, BUF_KEY, BUF_PIDX, BUF_PNUM, BUF_PARENT, BUF_CHILDREN, \
    BUF_NEIGHS, BUF_DIDX, BUF_DNUM, BUF_VECTOR_FIELD, BUF_DIVERGENCE, BUF_SOLUTION, BUF_POINT_VALUES, \
    BUF_LEAF_CORNER_VALUES = range(16)
BUFFER_DTYPES = {
    BUF_POINTS: (np.float32, 3), BUF_NORMALS: (np.float32, 3), BUF_POINT_LEAF: (np.int32, 1),
    BUF_KEY: (np.uint32, 1), BUF_PIDX: (np.int32, 1), BUF_PNUM: (np.int32, 1), BUF_PARENT: (np.int32, 1),
    BUF_CHILDREN: (np.int32, 8), BUF_NEIGHS: (np.int32, 27), BUF_DIDX: (np.int32, 1), BUF_DNUM: (np.int32, 1),
    BUF_VECTOR_FIELD: (np.float32, 3), BUF_DIVERGENCE: (np.float32, 1), BUF_SOLUTION: (np.float32, 1),
    BUF_POINT_VALUES: (np.float32, 1), BUF_LEAF_CORNER_VALUES: (np.float32, 8),
}

STAGES = ["normalize", "octree", "tables", "vector_field", "divergence", "solve", "isovalue",
          "mc_topology", "mc_eval", "subdivide", "mesh_copy"]

WARM_NONE, WARM_LEARNED = 0, 2


class PrgConfig(C.Structure):
    _fields_ = [("max_depth", C.c_int32), ("device", C.c_int32), ("cg_tol", C.c_float),
                ("cg_max_iters", C.c_int32), ("cg_refresh", C.c_int32), ("cg_fixed_iters", C.c_int32),
                ("warm_start", C.c_int32), ("timing", C.c_int32)]


class PrgSolveStats(C.Structure):
    _fields_ = [("depth", C.c_int32), ("iters", C.c_int32), ("nodes", C.c_int64),
                ("rel_residual", C.c_double), ("true_rel_residual", C.c_double), ("ms", C.c_double)]


class PrgMesh(C.Structure):
    _fields_ = [("n_vertices", C.c_int64), ("n_triangles", C.c_int64),
                ("vertices", C.POINTER(C.c_float)), ("triangles", C.POINTER(C.c_int32)),
                ("center", C.c_float * 3), ("scale", C.c_float),
                ("n_leaf_vertices", C.c_int64), ("n_leaf_triangles", C.c_int64)]


class PrgBatchItem(C.Structure):
    _fields_ = [("xyz", C.c_void_p), ("nxyz", C.c_void_p), ("n", C.c_int64), ("max_depth", C.c_int32),
                ("status", C.c_int32), ("mesh", PrgMesh), ("ms", C.c_double),
                ("max_true_rel_residual", C.c_double), ("total_iters", C.c_int32)]


class PrgTimings(C.Structure):
    _fields_ = [("ms", C.c_double * 11)]


class PrgLevelDesc(C.Structure):
    _fields_ = [("depth", C.c_int32), ("count", C.c_int64), ("base", C.c_int64),
                ("key", C.c_void_p), ("pidx", C.c_void_p), ("pnum", C.c_void_p), ("parent", C.c_void_p),
                ("child_base", C.c_void_p), ("didx", C.c_void_p), ("dnum", C.c_void_p),
                ("neighs", C.c_void_p), ("neigh_stride", C.c_int64)]


# Every function of include/prg.h and include/prg_synth.h: name -> (restype, argtypes)
_P = C.c_void_p
_F32P = C.POINTER(C.c_float)
SIGNATURES = {
    "prg_config_default": (C.c_int, [C.POINTER(PrgConfig)]),
    "prg_create": (C.c_int, [C.POINTER(_P), C.POINTER(PrgConfig)]),
    "prg_set_max_depth": (C.c_int, [_P, C.c_int32]),
    "prg_destroy": (None, [_P]),
    "prg_last_error": (C.c_char_p, [_P]),
    "prg_set_points": (C.c_int, [_P, _P, _P, C.c_int64, C.c_int32]),
    "prg_build_octree": (C.c_int, [_P]),
    "prg_level_info": (C.c_int, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "prg_level_view": (C.c_int, [_P, C.c_int32, C.POINTER(PrgLevelDesc)]),
    "prg_compute_divergence": (C.c_int, [_P]),
    "prg_solve": (C.c_int, [_P, C.POINTER(PrgSolveStats)]),
    "prg_isovalue": (C.c_int, [_P, _F32P]),
    "prg_set_isovalue": (C.c_int, [_P, C.c_float]),
    "prg_extract_mesh": (C.c_int, [_P, C.c_int32, C.POINTER(PrgMesh)]),
    "prg_mesh_release": (None, [C.POINTER(PrgMesh)]),
    "prg_weld_mesh": (C.c_int, [_P, C.POINTER(PrgMesh), C.POINTER(C.c_int64)]),
    "prg_decimate_mesh": (C.c_int, [C.POINTER(PrgMesh), C.c_int32, C.c_float]),
    "prg_weld_mesh_noalloc": (C.c_int, [_P, C.POINTER(PrgMesh), C.POINTER(C.c_int64)]),
    "prg_extract_mesh_into": (C.c_int, [_P, C.c_int32, _P, C.c_int64, _P, C.c_int64, C.POINTER(PrgMesh)]),
    "prg_subdivision_info": (C.c_int, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "prg_subdivision_bricks": (C.c_int, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "prg_reconstruct": (C.c_int, [_P, C.c_int32, C.POINTER(PrgMesh), C.POINTER(PrgSolveStats)]),
    "prg_stage_times": (C.c_int, [_P, C.POINTER(PrgTimings)]),
    "prg_buffer_size": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_int64)]),
    "prg_read_buffer": (C.c_int, [_P, C.c_int32, _P, C.c_int64]),
    "prg_write_buffer": (C.c_int, [_P, C.c_int32, _P, C.c_int64]),
    "prg_kernel_stats": (C.c_int, [_P, C.c_char_p, C.POINTER(C.c_int64), C.POINTER(C.c_double),
                                   C.POINTER(C.c_double)]),
    "prg_kernel_stats_reset": (C.c_int, [_P]),
    "prg_get_transform": (C.c_int, [_P, _F32P, _F32P]),
    "prg_read_points": (C.c_int, [C.c_char_p, C.POINTER(_F32P), C.POINTER(_F32P), C.POINTER(C.c_int64)]),
    "prg_points_release": (None, [_F32P, _F32P]),
    "prg_write_mesh_ply": (C.c_int, [C.c_char_p, C.POINTER(PrgMesh)]),
    "prg_reconstruct_batch": (C.c_int, [C.POINTER(PrgBatchItem), C.c_int64, C.POINTER(PrgConfig), C.c_int32,
                                         C.c_int32]),
    "prg_unet_create": (C.c_int, [C.POINTER(_P), C.c_int32, C.c_int32, C.c_uint64]),
    "prg_unet_destroy": (None, [_P]),
    "prg_unet_last_error": (C.c_char_p, [_P]),
    "prg_unet_num_params": (C.c_int64, [_P]),
    "prg_unet_grid": (C.c_int32, [_P]),
    "prg_unet_get_params": (C.c_int, [_P, _P, C.c_int64]),
    "prg_unet_set_params": (C.c_int, [_P, _P, C.c_int64]),
    "prg_unet_predict": (C.c_int, [_P, _P, _P, C.c_float, _P]),
    "prg_unet_gradients": (C.c_int, [_P, _P, _P, _P, C.c_float, _P, C.POINTER(C.c_double)]),
    "prg_unet_train_step": (C.c_int, [_P, _P, _P, _P, _P, C.c_int32, C.c_float, C.POINTER(C.c_double)]),
    "prg_dense_grid": (C.c_int, [_P, C.c_int32, C.c_int32, _P, C.POINTER(C.c_float)]),
    "prg_set_warm_start_net": (C.c_int, [_P, _P]),
    "prg_warm_start_info": (C.c_int, [_P, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "prg_synth_sphere": (C.c_int, [C.c_int64, C.c_uint64, _P, _P]),
    "prg_synth_torus": (C.c_int, [C.c_int64, C.c_uint64, _P, _P]),
    "prg_synth_superquadrics": (C.c_int, [C.c_int64, C.c_int32, C.c_float, C.c_uint64, _P, _P]),
}


def load_library(path: Path | str = LIB_PATH) -> C.CDLL:
    path = Path(path)
    if not path.exists():
        raise ImportError(f"{path} is missing: build it with `python -m poissonrecon_gpu_amd.build`")
    lib = C.CDLL(str(path))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_lib: C.CDLL | None = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        _lib = load_library()
    return _lib


class PrgError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


def _ptr(a: np.ndarray) -> C.c_void_p:
    return C.c_void_p(a.ctypes.data)


# ---- synthetic workloads ------------------------------------------------------------
def synth_sphere(n: int, seed: int = 1) -> tuple[np.ndarray, np.ndarray]:
    p = np.empty((n, 3), np.float32)
    q = np.empty((n, 3), np.float32)
    _check(lib().prg_synth_sphere(n, seed, _ptr(p), _ptr(q)), "prg_synth_sphere")
    return p, q


def synth_torus(n: int, seed: int = 2) -> tuple[np.ndarray, np.ndarray]:
    p = np.empty((n, 3), np.float32)
    q = np.empty((n, 3), np.float32)
    _check(lib().prg_synth_torus(n, seed, _ptr(p), _ptr(q)), "prg_synth_torus")
    return p, q


def synth_superquadrics(n: int, shapes: int = 3, sigma: float = 0.001, seed: int = 3) -> tuple[np.ndarray, np.ndarray]:
    p = np.empty((n, 3), np.float32)
    q = np.empty((n, 3), np.float32)
    _check(lib().prg_synth_superquadrics(n, shapes, sigma, seed, _ptr(p), _ptr(q)), "prg_synth_superquadrics")
    return p, q


def _check(st: int, what: str, ctx=None) -> int:
    if st != OK:
        msg = lib().prg_last_error(ctx).decode() if ctx else what
        raise PrgError(st, msg)
    return st


@dataclass
class Mesh:
    vertices: np.ndarray   # (V, 3) normalised unit-cube coordinates
    triangles: np.ndarray  # (T, 3) int32
    center: np.ndarray
    scale: float
    n_leaf_vertices: int
    n_leaf_triangles: int

    def world_vertices(self) -> np.ndarray:
        """v * scale + center in float32, as PlyWriteTriangles (plyfile.cu:2800-2806)."""
        return (self.vertices * np.float32(self.scale)) + self.center.astype(np.float32)


class Reconstructor:
    """One reconstruction context (prg_ctx) on one HIP device."""

    def __init__(self, max_depth: int = 8, device: int = 0, cg_tol: float = 1e-6, cg_max_iters: int = 2000,
                 cg_refresh: int = 20, cg_fixed_iters: int = 0, warm_start: int = WARM_NONE, timing: bool = False):
        L = lib()
        cfg = PrgConfig()
        L.prg_config_default(C.byref(cfg))
        cfg.max_depth, cfg.device, cfg.cg_tol = max_depth, device, cg_tol
        cfg.cg_max_iters, cfg.cg_refresh, cfg.cg_fixed_iters = cg_max_iters, cg_refresh, cg_fixed_iters
        cfg.warm_start, cfg.timing = warm_start, int(timing)
        h = C.c_void_p()
        st = L.prg_create(C.byref(h), C.byref(cfg))
        if st != OK:
            raise PrgError(st, "prg_create")
        self._h = h
        self.max_depth = max_depth

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().prg_destroy(self._h)
            self._h = None

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _call(self, name: str, *args) -> int:
        st = getattr(lib(), name)(self._h, *args)
        if st not in (OK, E_NOT_CONVERGED):
            raise PrgError(st, lib().prg_last_error(self._h).decode())
        return st

    def set_max_depth(self, max_depth: int) -> None:
        """prg_set_max_depth: a new depth for this context, keeping its buffers."""
        self._call("prg_set_max_depth", int(max_depth))
        self.max_depth = int(max_depth)

    # ---- stages ----
    def set_points(self, xyz: np.ndarray, nxyz: np.ndarray) -> None:
        xyz = np.ascontiguousarray(xyz, np.float32)
        nxyz = np.ascontiguousarray(nxyz, np.float32)
        self._call("prg_set_points", _ptr(xyz), _ptr(nxyz), len(xyz), 0)

    def build_octree(self) -> None:
        self._call("prg_build_octree")

    def compute_divergence(self) -> None:
        self._call("prg_compute_divergence")

    def solve(self) -> list[dict]:
        stats = (PrgSolveStats * (self.max_depth + 1))()
        st = self._call("prg_solve", stats)
        out = [dict(depth=s.depth, iters=s.iters, nodes=s.nodes, rel_residual=s.rel_residual,
                    true_rel_residual=s.true_rel_residual, ms=s.ms) for s in stats]
        if st == E_NOT_CONVERGED:
            for o in out:
                o["not_converged"] = True
        return out

    def isovalue(self) -> float:
        v = C.c_float()
        self._call("prg_isovalue", C.byref(v))
        return float(v.value)

    def set_isovalue(self, iso: float) -> None:
        self._call("prg_set_isovalue", C.c_float(iso))

    def extract_mesh(self, subdivide: bool = True) -> Mesh:
        m = PrgMesh()
        self._call("prg_extract_mesh", int(subdivide), C.byref(m))
        try:
            v = np.ctypeslib.as_array(m.vertices, shape=(max(m.n_vertices, 1) * 3,))[: 3 * m.n_vertices]
            t = np.ctypeslib.as_array(m.triangles, shape=(max(m.n_triangles, 1) * 3,))[: 3 * m.n_triangles]
            return Mesh(v.reshape(-1, 3).copy(), t.reshape(-1, 3).copy(), np.array(m.center[:], np.float32),
                        float(m.scale), m.n_leaf_vertices, m.n_leaf_triangles)
        finally:
            lib().prg_mesh_release(C.byref(m))

    def weld_mesh(self, mesh: Mesh, noalloc: bool = False) -> tuple[Mesh, int]:
        """prg_weld_mesh (or, with noalloc, prg_weld_mesh_noalloc) on a copy of
        `mesh`: bit-identical vertices merged into their first occurrence;
        returns (welded mesh, vertices removed)."""
        v = np.ascontiguousarray(mesh.vertices, np.float32).copy()
        t = np.ascontiguousarray(mesh.triangles, np.int32).copy()
        m = PrgMesh()
        m.n_vertices, m.n_triangles = len(v), len(t)
        m.vertices = v.ctypes.data_as(C.POINTER(C.c_float))
        m.triangles = t.ctypes.data_as(C.POINTER(C.c_int32))
        m.n_leaf_vertices, m.n_leaf_triangles = int(mesh.n_leaf_vertices), int(mesh.n_leaf_triangles)
        merged = C.c_int64()
        self._call("prg_weld_mesh_noalloc" if noalloc else "prg_weld_mesh", C.byref(m), C.byref(merged))
        return (Mesh(v[: m.n_vertices].copy(), t, mesh.center, mesh.scale, m.n_leaf_vertices, m.n_leaf_triangles),
                merged.value)

    def subdivision_info(self) -> tuple[int, int]:
        """(coarse leaves refined, virtual leaves) of the last subdivided mesh."""
        a, b = C.c_int64(), C.c_int64()
        self._call("prg_subdivision_info", C.byref(a), C.byref(b))
        return a.value, b.value

    def subdivision_bricks(self) -> tuple[int, int, int]:
        """(bricks, left by the interval certificate, evaluated) of the last subdivided mesh."""
        a, b, e = C.c_int64(), C.c_int64(), C.c_int64()
        self._call("prg_subdivision_bricks", C.byref(a), C.byref(b), C.byref(e))
        return a.value, b.value, e.value

    def dense_grid(self, depth: int, which: int) -> tuple[np.ndarray, float]:
        """Dense (G, G, G) grid of one depth (0 divergence, 1 solution, 2 node mask), and
        the divergence's root-mean-square over the depth's nodes."""
        G = 1 << depth
        out = np.zeros((G, G, G), np.float32)
        rms = C.c_float()
        self._call("prg_dense_grid", depth, which, _ptr(out), C.byref(rms))
        return out, rms.value

    def set_warm_start_net(self, net: "UNet | None") -> None:
        self._call("prg_set_warm_start_net", net._h if net is not None else None)

    def warm_start_info(self) -> tuple[int, bool]:
        """(depth the network warm-started or -1, whether its start was rejected) of the last solve."""
        d, r = C.c_int32(), C.c_int32()
        self._call("prg_warm_start_info", C.byref(d), C.byref(r))
        return d.value, bool(r.value)

    # ---- introspection ----
    def level_info(self) -> tuple[np.ndarray, np.ndarray]:
        counts = (C.c_int64 * (self.max_depth + 1))()
        bases = (C.c_int64 * (self.max_depth + 1))()
        self._call("prg_level_info", counts, bases)
        return np.array(counts[:], np.int64), np.array(bases[:], np.int64)

    def read_buffer(self, which: int) -> np.ndarray:
        nbytes = C.c_int64()
        self._call("prg_buffer_size", which, C.byref(nbytes))
        dt, width = BUFFER_DTYPES[which]
        a = np.empty(nbytes.value // np.dtype(dt).itemsize, dt)
        self._call("prg_read_buffer", which, _ptr(a), nbytes.value)
        return a.reshape(-1, width) if width > 1 else a

    def write_buffer(self, which: int, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, BUFFER_DTYPES[which][0])
        self._call("prg_write_buffer", which, _ptr(a), a.nbytes)

    def stage_times(self) -> dict[str, float]:
        t = PrgTimings()
        self._call("prg_stage_times", C.byref(t))
        return dict(zip(STAGES, t.ms[:]))

    def kernel_stats(self, name: str) -> tuple[int, float, float]:
        n, ms, b = C.c_int64(), C.c_double(), C.c_double()
        self._call("prg_kernel_stats", name.encode(), C.byref(n), C.byref(ms), C.byref(b))
        return n.value, ms.value, b.value

    def kernel_stats_reset(self) -> None:
        self._call("prg_kernel_stats_reset")

    def transform(self) -> tuple[np.ndarray, float]:
        c = (C.c_float * 3)()
        s = C.c_float()
        self._call("prg_get_transform", c, C.byref(s))
        return np.array(c[:], np.float32), float(s.value)


def read_points(path: str) -> tuple[np.ndarray, np.ndarray]:
    x, n = _F32P(), _F32P()
    cnt = C.c_int64()
    _check(lib().prg_read_points(path.encode(), C.byref(x), C.byref(n), C.byref(cnt)), f"prg_read_points({path})")
    try:
        xyz = np.ctypeslib.as_array(x, shape=(cnt.value * 3,)).reshape(-1, 3).copy()
        nxyz = np.ctypeslib.as_array(n, shape=(cnt.value * 3,)).reshape(-1, 3).copy()
    finally:
        lib().prg_points_release(x, n)
    return xyz, nxyz


def decimate_mesh(mesh: Mesh, mode: int, edge_ratio: float) -> Mesh:
    """prg_decimate_mesh on a copy of `mesh`: EdgeCollapse (mode 0) or
    TriangleCollapse (mode 1) with the given edge ratio."""
    v = np.ascontiguousarray(mesh.vertices, np.float32).copy()
    t = np.ascontiguousarray(mesh.triangles, np.int32).copy()
    m = PrgMesh()
    m.n_vertices, m.n_triangles = len(v), len(t)
    m.vertices = v.ctypes.data_as(_F32P)
    m.triangles = t.ctypes.data_as(C.POINTER(C.c_int32))
    _check(lib().prg_decimate_mesh(C.byref(m), int(mode), C.c_float(edge_ratio)), "prg_decimate_mesh")
    return Mesh(v[: m.n_vertices].copy(), t[: m.n_triangles].copy(), mesh.center, mesh.scale, 0, 0)


def write_mesh_ply(path: str, mesh: Mesh) -> None:
    m = PrgMesh()
    v = np.ascontiguousarray(mesh.vertices, np.float32)
    t = np.ascontiguousarray(mesh.triangles, np.int32)
    m.n_vertices, m.n_triangles = len(v), len(t)
    m.vertices = v.ctypes.data_as(_F32P)
    m.triangles = t.ctypes.data_as(C.POINTER(C.c_int32))
    m.center[:] = list(mesh.center)
    m.scale = mesh.scale
    _check(lib().prg_write_mesh_ply(path.encode(), C.byref(m)), f"prg_write_mesh_ply({path})")


class UNet:
    """The learned warm start's network (include/prg_unet.h), trained and run on the GPU."""

    def __init__(self, grid: int = 64, device: int = 0, seed: int = 1):
        self._h = C.c_void_p()
        st = lib().prg_unet_create(C.byref(self._h), device, grid, seed)
        if st != OK:
            raise PrgError(st, "prg_unet_create")
        self.grid = grid

    def close(self) -> None:
        if self._h:
            lib().prg_unet_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _call(self, name: str, *args) -> None:
        st = getattr(lib(), name)(self._h, *args)
        if st != OK:
            raise PrgError(st, f"{name}: {lib().prg_unet_last_error(self._h).decode()}")

    @property
    def num_params(self) -> int:
        return int(lib().prg_unet_num_params(self._h))

    def get_params(self) -> np.ndarray:
        out = np.zeros(self.num_params, np.float32)
        self._call("prg_unet_get_params", _ptr(out), out.size)
        return out

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, np.float32)
        self._call("prg_unet_set_params", _ptr(p), p.size)

    def predict(self, div_grid: np.ndarray, mask_grid: np.ndarray, scale: float) -> np.ndarray:
        b = np.ascontiguousarray(div_grid, np.float32)
        m = np.ascontiguousarray(mask_grid, np.float32)
        out = np.zeros_like(b)
        self._call("prg_unet_predict", _ptr(b), _ptr(m), C.c_float(scale), _ptr(out))
        return out

    def gradients(self, div_grid, mask_grid, target_grid, scale: float) -> tuple[float, np.ndarray]:
        b = np.ascontiguousarray(div_grid, np.float32)
        m = np.ascontiguousarray(mask_grid, np.float32)
        t = np.ascontiguousarray(target_grid, np.float32)
        g = np.zeros(self.num_params, np.float32)
        loss = C.c_double()
        self._call("prg_unet_gradients", _ptr(b), _ptr(m), _ptr(t), C.c_float(scale), _ptr(g), C.byref(loss))
        return loss.value, g

    def train_step(self, div_grids, mask_grids, target_grids, scales, lr: float) -> float:
        b = np.ascontiguousarray(div_grids, np.float32)
        m = np.ascontiguousarray(mask_grids, np.float32)
        t = np.ascontiguousarray(target_grids, np.float32)
        s = np.ascontiguousarray(scales, np.float32)
        loss = C.c_double()
        self._call("prg_unet_train_step", _ptr(b), _ptr(m), _ptr(t), _ptr(s), len(s), C.c_float(lr), C.byref(loss))
        return loss.value


def reconstruct_batch(clouds, depths, concurrency: int = 4, subdivide: bool = True, device: int = 0,
                      cg_tol: float = 1e-6, return_meshes: bool = False):
    """Reconstructs independent clouds [(xyz, nxyz), ...] in one pass
    (prg_reconstruct_batch).  Returns per-cloud dicts: status, mesh sizes, ms,
    the largest per-depth true relative residual and the CG iterations."""
    n = len(clouds)
    items = (PrgBatchItem * n)()
    keep = []
    for i, ((p, q), d) in enumerate(zip(clouds, depths)):
        p = np.ascontiguousarray(p, np.float32)
        q = np.ascontiguousarray(q, np.float32)
        keep.append((p, q))
        items[i].xyz = p.ctypes.data
        items[i].nxyz = q.ctypes.data
        items[i].n = len(p)
        items[i].max_depth = int(d)
    cfg = PrgConfig()
    lib().prg_config_default(C.byref(cfg))
    cfg.device, cfg.cg_tol = device, cg_tol
    st = lib().prg_reconstruct_batch(items, n, C.byref(cfg), concurrency, int(subdivide))
    out = []
    for it in items:
        r = dict(status=int(it.status), vertices=int(it.mesh.n_vertices), triangles=int(it.mesh.n_triangles),
                 ms=float(it.ms), max_true_rel_residual=float(it.max_true_rel_residual),
                 total_iters=int(it.total_iters))
        if return_meshes and it.mesh.vertices:
            nv, nt = it.mesh.n_vertices, it.mesh.n_triangles
            r["mesh_vertices"] = np.ctypeslib.as_array(it.mesh.vertices, shape=(max(nv, 1) * 3,))[: 3 * nv].reshape(-1, 3).copy()
            r["mesh_triangles"] = np.ctypeslib.as_array(it.mesh.triangles, shape=(max(nt, 1) * 3,))[: 3 * nt].reshape(-1, 3).copy()
        out.append(r)
        lib().prg_mesh_release(C.byref(it.mesh))
    return st, out
