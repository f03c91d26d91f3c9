"""CPU oracle for the cpvs build and lookup path -- TEST INFRASTRUCTURE ONLY.

ctypes bindings to ``oracle/libcpvs_oracle.so`` (a C restatement of TobiasRp/cpvs
``MinMaxHierarchy``, ``CompressedShadow::create``/``traverse`` and ``shader/traverse.cs``).
Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py``
import this module, and only to check or time against; the product never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libcpvs_oracle.so")

SHADOW, VISIBLE, PARTIAL = 0, 1, 2


def build() -> None:
    """Compile the oracle (gcc) in place."""
    subprocess.run(["make", "-B", "-C", _HERE], check=True, stdout=subprocess.DEVNULL)


def _load():
    if not os.path.exists(_LIB_PATH):
        build()
    lib = C.CDLL(_LIB_PATH)
    f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
    u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
    u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
    u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
    i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
    sig = {
        "or_num_levels": (C.c_int, [C.c_int]),
        "or_pyramid_floats": (C.c_size_t, [C.c_int]),
        "or_pyramid_level_offset": (C.c_size_t, [C.c_int, C.c_int]),
        "or_pyramid_build": (None, [f32p, C.c_int, f32p]),
        "or_create_childmask": (C.c_uint32, [f32p, f32p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint32]),
        "or_create_leaf_node": (C.c_uint32, [f32p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint32, u64p, C.POINTER(C.c_int)]),
        "or_child_coordinates": (C.c_int, [C.c_uint32, C.c_int, C.c_int, C.c_int, i32p]),
        "or_merge_level": (C.c_uint32, [u32p, C.c_size_t, C.c_uint32, u32p, u32p, C.c_int]),
        "or_create": (C.c_int, [f32p, f32p, C.c_int, C.c_uint32, C.c_uint32, C.c_int,
                                C.POINTER(C.POINTER(C.c_uint32)), C.POINTER(C.c_size_t), C.POINTER(C.c_int)]),
        "or_free": (None, [C.c_void_p]),
        "or_last_node_counts": (None, [u64p, C.c_int]),
        "or_total_visibility": (C.c_int, [u32p]),
        "or_traverse_cpu": (C.c_int, [u32p, C.c_int, f32p, C.c_int]),
        "or_set_leafmasks": (None, [C.c_int]),
        "or_traverse_path_cs": (C.c_int, [u32p, u32p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
        "or_path_from_ndc": (None, [f32p, C.c_int, C.c_int, i32p]),
        "or_evaluate": (None, [u32p, u32p, C.c_int, C.c_int, f32p, C.c_uint32, C.c_size_t, C.c_void_p, u8p, C.c_int]),
        "or_brute_visible": (C.c_int, [f32p, C.c_int, C.c_uint32, C.c_int, C.c_int, C.c_int]),
        "or_gen_depth_tile": (None, [C.c_int, C.c_uint64, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, f32p]),
        "or_gen_queries_uniform": (None, [C.c_uint64, C.c_size_t, C.c_uint32, f32p]),
        "or_gen_queries_gbuffer": (None, [f32p, C.c_int, C.c_uint64, C.c_int, C.c_int, f32p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def num_levels(N: int) -> int:
    return lib().or_num_levels(N)


class Pyramid:
    """MinMaxHierarchy (src/MinMaxHierarchy.{h,cpp}) restated on the CPU."""

    def __init__(self, depth: np.ndarray):
        depth = _f32(depth)
        assert depth.ndim == 2 and depth.shape[0] == depth.shape[1]
        self.N = depth.shape[0]
        self.depth = depth
        self.P = np.zeros(max(1, lib().or_pyramid_floats(self.N)), np.float32)
        lib().or_pyramid_build(depth, self.N, self.P)

    def getNumLevels(self) -> int:
        return lib().or_num_levels(self.N)

    def level(self, L: int) -> np.ndarray:
        """Level L >= 1 as an (w, w, 2) array of (min, max)."""
        w = self.N >> L
        off = lib().or_pyramid_level_offset(self.N, L)
        return self.P[off: off + w * w * 2].reshape(w, w, 2)

    def getMin(self, L, x, y):
        return float(self.depth[y, x]) if L == 0 else float(self.level(L)[y, x, 0])

    def getMax(self, L, x, y):
        return float(self.depth[y, x]) if L == 0 else float(self.level(L)[y, x, 1])


def create_childmask(pyr: Pyramid, level: int, off, zTileNum: int = 1) -> int:
    return lib().or_create_childmask(pyr.depth, pyr.P, pyr.N, level, off[0], off[1], off[2], zTileNum)


def create_leaf_node(depth: np.ndarray, off, zTileNum: int = 1):
    depth = _f32(depth)
    masks = np.zeros(8, np.uint64)
    k = C.c_int(0)
    m = lib().or_create_leaf_node(depth, depth.shape[0], off[0], off[1], off[2], zTileNum, masks, C.byref(k))
    return m, [int(v) for v in masks[: k.value]]


def child_coordinates(mask: int, parent):
    out = np.zeros(24, np.int32)
    k = lib().or_child_coordinates(mask, parent[0], parent[1], parent[2], out)
    return [tuple(int(v) for v in out[3 * i: 3 * i + 3]) for i in range(k)]


def merge_level(words: np.ndarray, node_size: int, mode: int = 0):
    words = np.ascontiguousarray(words, np.uint32)
    n = len(words) // node_size
    out = np.zeros(max(1, len(words)), np.uint32)
    mapping = np.zeros(max(1, n), np.uint32)
    u = lib().or_merge_level(words, n, node_size, out, mapping, mode)
    return out[: u * node_size], {i * node_size: int(mapping[i]) for i in range(n)}


class leafmasks_off:
    """The reference built without LEAFMASKS (CompressedShadow.cpp:17, traverse.cs:8-9): inside the
    block create() and evaluate() build and walk DAGs that go down to level 0."""

    def __enter__(self):
        lib().or_set_leafmasks(0)
        return self

    def __exit__(self, *exc):
        lib().or_set_leafmasks(1)
        return False


def create(pyr: Pyramid, zTileIndex: int = 0, zTileNum: int = 1, merge_mode: int = 1, leafmasks: bool = True):
    """CompressedShadow::create -> (words uint32 array, numLevels); leafmasks=False: the reference
    compiled without LEAFMASKS."""
    ptr = C.POINTER(C.c_uint32)()
    ln = C.c_size_t(0)
    nl = C.c_int(0)
    lib().or_set_leafmasks(int(leafmasks))
    try:
        rc = lib().or_create(pyr.depth, pyr.P, pyr.N, zTileIndex, zTileNum, merge_mode,
                             C.byref(ptr), C.byref(ln), C.byref(nl))
    finally:
        lib().or_set_leafmasks(1)
    if rc != 0:
        raise ValueError(f"or_create failed rc={rc}")
    words = np.ctypeslib.as_array(ptr, shape=(ln.value,)).copy()
    lib().or_free(ptr)
    return words, nl.value


def last_node_counts(n: int) -> np.ndarray:
    out = np.zeros(n, np.uint64)
    lib().or_last_node_counts(out, n)
    return out


def total_visibility(words) -> int:
    return lib().or_total_visibility(np.ascontiguousarray(words, np.uint32))


def traverse_cpu(words, num_levels: int, ndc, tryLeafmasks: bool = True) -> int:
    return lib().or_traverse_cpu(np.ascontiguousarray(words, np.uint32), num_levels, _f32(ndc), int(tryLeafmasks))


def path_from_ndc(ndc, dag_levels: int, grid_levels: int = 0):
    out = np.zeros(3, np.int32)
    lib().or_path_from_ndc(_f32(ndc), dag_levels, grid_levels, out)
    return tuple(int(v) for v in out)


def traverse_path_cs(dag, grid, dag_levels, grid_levels, path) -> int:
    return lib().or_traverse_path_cs(np.ascontiguousarray(dag, np.uint32), np.ascontiguousarray(grid, np.uint32),
                                     dag_levels, grid_levels, int(path[0]), int(path[1]), int(path[2]))


def evaluate(dag, grid, dag_levels, grid_levels, positions: np.ndarray, matrix=None, nthreads: int = 1,
             leafmasks: bool = True):
    """traverse.cs main() over an (count, 3|4) float32 array -> uint8 visibilities (leafmasks=False:
    the shader compiled without LEAFMASKS)."""
    pos = _f32(positions)
    stride = pos.shape[1]
    vis = np.zeros(pos.shape[0], np.uint8)
    m = None if matrix is None else _f32(matrix).reshape(16)
    mp = None if m is None else m.ctypes.data_as(C.c_void_p)
    lib().or_set_leafmasks(int(leafmasks))
    try:
        lib().or_evaluate(np.ascontiguousarray(dag, np.uint32), np.ascontiguousarray(grid, np.uint32),
                          dag_levels, grid_levels, pos, stride, pos.shape[0], mp, vis, nthreads)
    finally:
        lib().or_set_leafmasks(1)
    return vis


def top_level_grid(dags) -> np.ndarray:
    """createTopLevelGrid with the 0xFFFFFFFF / 0xFFFFFFFE sentinels traverse.cs tests."""
    grid = np.zeros(len(dags), np.uint32)
    off = 0
    for i, d in enumerate(dags):
        v = total_visibility(d)
        grid[i] = 0xFFFFFFFF if v == SHADOW else 0xFFFFFFFE if v == VISIBLE else off
        off += len(d)
    return grid


def brute_visible(depth: np.ndarray, zTileNum: int, px, py, pz):
    """Vectorised leaf-resolution midpoint test float(pz)+0.5 <= D[py,px]*float(N*T)."""
    depth = _f32(depth)
    N = depth.shape[0]
    h = np.float32(N * zTileNum)
    pz = np.asarray(pz, np.int64)
    mid = (pz.astype(np.float32) + (pz + 1).astype(np.float32)) * np.float32(0.5)
    return (mid <= depth[np.asarray(py), np.asarray(px)] * h).astype(np.uint8)


def gen_depth(N: int, seed: int, nboxes: int, noise: bool, x0: int = 0, y0: int = 0, tileN: int | None = None):
    tileN = N if tileN is None else tileN
    out = np.zeros((tileN, tileN), np.float32)
    lib().or_gen_depth_tile(N, seed, nboxes, int(noise), x0, y0, tileN, out)
    return out


def gen_queries_uniform(seed: int, count: int, stride: int = 4) -> np.ndarray:
    out = np.zeros((count, stride), np.float32)
    lib().or_gen_queries_uniform(seed, count, stride, out)
    return out


def gen_queries_gbuffer(depth: np.ndarray, seed: int, W: int = 3840, H: int = 2160) -> np.ndarray:
    depth = _f32(depth)
    out = np.zeros((2 * W * H, 4), np.float32)
    lib().or_gen_queries_gbuffer(depth, depth.shape[0], seed, W, H, out)
    return out
