"""cpvs_amd -- precomputed voxelized shadows (TobiasRp/cpvs build + lookup path) on MI355X.

Python mirror of the reference's C++ interface, over the C ABI of ``libcpvs_amd.so``
(include/cpvs.h).  Names, argument meaning and ownership follow the reference:

* ``MinMaxHierarchy(image)``                      -- src/MinMaxHierarchy.h:23-72
* ``CompressedShadow.create(minMax, zTileIndex, zTileNum)`` and ``traverse``, ``getDAG``,
  ``getNumLevels``, ``getTotalVisibility``         -- src/CompressedShadow.h:20-125
* ``CompressedShadowContainer(length | shadow)`` with ``set``/``get``/``copyToGPU``/
  ``moveToGPU``/``freeOnCPU``/``setFilterSize``/``evaluate`` -- src/CompressedShadowContainer.h:18-92

Every computation runs as HIP kernels on the GPU; DAGs stay in device memory.
"""
from __future__ import annotations

import ctypes as C
import json
import os
import weakref
from typing import Optional, Sequence

import numpy as np

from . import _lib
from ._lib import CpvsError, call

__all__ = ["Context", "DeviceArray", "MinMaxHierarchy", "CompressedShadow", "CompressedShadowContainer",
           "DagSet", "Server", "CpvsError", "default_context", "build_batch", "tune_batch", "dag_check", "gen_depth",
           "gen_queries_uniform", "gen_queries_uniform_range", "gen_queries_gbuffer"]


class Context:
    """One HIP device, its stream and build scratch (cpvs_ctx)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        call("cpvs_ctx_create", device, C.byref(h))
        self.handle = h
        self.device = device
        self._objects = weakref.WeakSet()  # objects made on this context that are still open

    def synchronize(self):
        call("cpvs_ctx_synchronize", self.handle)

    @property
    def stream(self):
        s = C.c_void_p()
        call("cpvs_ctx_stream", self.handle, C.byref(s))
        return s

    def profile(self, mode: int):
        """1: time every kernel with HIP events, 0: stop, -1: clear totals."""
        call("cpvs_ctx_profile", self.handle, mode)

    def profile_read(self) -> dict:
        """{kernel name: (total ms, launches)} accumulated while profiling was on."""
        buf = C.create_string_buffer(1 << 16)
        call("cpvs_ctx_profile_read", self.handle, buf, len(buf))
        out = {}
        for line in buf.value.decode().splitlines():
            name, ms, cnt = line.split("\t")
            out[name] = (float(ms), int(cnt))
        return out

    def device_name(self) -> str:
        buf = C.create_string_buffer(256)
        call("cpvs_ctx_device_name", self.handle, buf, len(buf))
        return buf.value.decode()

    def batch_lanes(self) -> int:
        v = C.c_uint32()
        call("cpvs_ctx_batch_lanes", self.handle, C.byref(v))
        return int(v.value)

    def set_batch_lanes(self, lanes: int):
        call("cpvs_ctx_set_batch_lanes", self.handle, lanes)

    def set_merge(self, mode: str):
        """'table' (default) or 'sort': how builds merge common subtrees (same words)."""
        call("cpvs_ctx_set_merge", self.handle, {"table": _lib.CPVS_MERGE_TABLE, "sort": _lib.CPVS_MERGE_SORT}[mode])

    def set_hash_bits(self, leaf_bits: int, inner_bits: int = 0):
        """Test hook: cut the table merge's key hashes to their low bits (0 = whole), at the leaf
        level and above it, so that distinct keys collide; the key compares must catch each
        collision (the build then sorts: same words)."""
        call("cpvs_ctx_set_hash_bits", self.handle, leaf_bits, inner_bits)

    def set_leafmasks(self, on: bool):
        """The reference's LEAFMASKS compile switch (src/CompressedShadow.cpp:17, shader/traverse.cs:8-9)
        as a mode of this context, on by default.  Off, builds of maps of N >= 16 go down to level 0
        and pyramids made afterwards keep levels 1 and 2; DAGs keep the mode they were built in."""
        call("cpvs_ctx_set_leafmasks", self.handle, int(bool(on)))

    def leafmasks(self) -> bool:
        v = C.c_int()
        call("cpvs_ctx_leafmasks", self.handle, C.byref(v))
        return bool(v.value)

    def release_scratch(self) -> int:
        """Frees the build scratch, the batch lanes' arenas and the cache of freed buffers; returns
        the bytes freed.  Call it between builds."""
        n = C.c_uint64()
        call("cpvs_ctx_release_scratch", self.handle, C.byref(n))
        return int(n.value)

    def event(self, stream=None):
        e = C.c_void_p()
        call("cpvs_event_record", self.handle, stream, C.byref(e))
        return e

    @staticmethod
    def elapsed_ms(start, stop) -> float:
        ms = C.c_float()
        call("cpvs_event_elapsed_ms", start, stop, C.byref(ms))
        return float(ms.value)

    @staticmethod
    def destroy_event(e):
        call("cpvs_event_destroy", e)

    def close(self):
        """Destroys the context after closing every object still open on it (DAG sets and servers
        first, then containers, DAGs, pyramids and device arrays): none of them may outlive it."""
        if self.handle:
            live = list(self._objects)
            for kind in (DagSet, Server, CompressedShadowContainer, CompressedShadow, MinMaxHierarchy, DeviceArray):
                for o in live:
                    if type(o) is kind:
                        o.close()
            _lib.load().cpvs_ctx_destroy(self.handle)
            self.handle = None


def _track(obj):
    obj.ctx._objects.add(obj)
    return obj


_default: Optional[Context] = None


def default_context() -> Context:
    global _default
    if _default is None:
        _default = Context(0)
    return _default


class DeviceArray:
    """A block of device memory with a numpy dtype/shape (allocated through the C ABI)."""

    def __init__(self, shape, dtype, ctx: Optional[Context] = None):
        self.ctx = ctx or default_context()
        self.shape = tuple(shape) if isinstance(shape, (tuple, list)) else (int(shape),)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        call("cpvs_device_alloc", self.ctx.handle, self.nbytes, C.byref(p))
        self.ptr = p
        _track(self)

    @classmethod
    def from_numpy(cls, a: np.ndarray, ctx: Optional[Context] = None) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype, ctx)
        if a.nbytes:
            call("cpvs_memcpy_h2d", d.ctx.handle, d.ptr, a.ctypes.data_as(C.c_void_p), a.nbytes)
        return d

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            call("cpvs_memcpy_d2h", self.ctx.handle, out.ctypes.data_as(C.c_void_p), self.ptr, self.nbytes)
        return out

    def free(self):
        if self.ptr:
            call("cpvs_device_free", self.ctx.handle, self.ptr)
            self.ptr = None

    close = free

    def __del__(self):
        try:
            if self.ptr and self.ctx.handle:
                _lib.load().cpvs_device_free(self.ctx.handle, self.ptr)
        except Exception:
            pass


class MinMaxHierarchy:
    """MinMaxHierarchy(const ImageF&) (src/MinMaxHierarchy.h:23, .cpp:9-97) on the GPU."""

    def __init__(self, image, all_levels: bool = False, borrow: bool = False, ctx: Optional[Context] = None,
                 stream=None):
        self.ctx = ctx or default_context()
        flags = _lib.CPVS_PYR_ALL_LEVELS if all_levels else 0
        h = C.c_void_p()
        if isinstance(image, DeviceArray):
            N = image.shape[0]
            assert image.dtype == np.float32 and image.shape == (N, N)
            if borrow:
                flags |= _lib.CPVS_PYR_BORROW_DEPTH
            self._keep = image if borrow else None
            call("cpvs_pyramid_build", self.ctx.handle, image.ptr, 1, N, flags, stream, C.byref(h))
        else:
            img = np.ascontiguousarray(image, np.float32)
            if img.ndim != 2 or img.shape[0] != img.shape[1]:
                raise ValueError("MinMaxHierarchy needs a square 2-D image")
            N = img.shape[0]
            call("cpvs_pyramid_build", self.ctx.handle, img.ctypes.data_as(C.c_void_p), 0, N, flags, stream,
                 C.byref(h))
        self.handle = h
        self.N = N
        _track(self)

    def getNumLevels(self) -> int:
        v = C.c_int()
        call("cpvs_pyramid_num_levels", self.handle, C.byref(v))
        return v.value

    def _get(self, level, x, y):
        mn, mx = C.c_float(), C.c_float()
        call("cpvs_pyramid_get", self.handle, level, x, y, C.byref(mn), C.byref(mx))
        return mn.value, mx.value

    def getMin(self, level: int, x: int, y: int) -> float:
        return self._get(level, x, y)[0]

    def getMax(self, level: int, x: int, y: int) -> float:
        return self._get(level, x, y)[1]

    def getLevel(self, level: int) -> np.ndarray:
        """Level 0 as (N, N); level L >= 1 as (w, w, 2) of (min, max)."""
        w = self.N >> level
        out = np.empty((w, w) if level == 0 else (w, w, 2), np.float32)
        call("cpvs_pyramid_copy_level", self.handle, level, out.ctypes.data_as(C.c_void_p), out.size)
        return out

    def estimate(self, zTileIndex: int = 0, zTileNum: int = 1):
        """(nodes per level before merging, scratch bytes, max DAG bytes) of a build."""
        n = self.getNumLevels()
        nodes = np.zeros(max(n, 1), np.uint64)
        sb, db = C.c_uint64(), C.c_uint64()
        call("cpvs_estimate_build", self.ctx.handle, self.handle, zTileIndex, zTileNum,
             nodes.ctypes.data_as(C.c_void_p), C.byref(sb), C.byref(db))
        return nodes, sb.value, db.value

    def close(self):
        if getattr(self, "handle", None):
            _lib.load().cpvs_pyramid_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class CompressedShadow:
    """CompressedShadow (src/CompressedShadow.h:20-125); the DAG words live on the GPU."""

    SHADOW, VISIBLE, PARTIAL = 0, 1, 2

    def __init__(self, handle, ctx: Context):
        self.handle = handle
        self.ctx = ctx
        _track(self)

    @staticmethod
    def create(minMax: MinMaxHierarchy, zTileIndex: int = 0, zTileNum: int = 1, stream=None) -> "CompressedShadow":
        h = C.c_void_p()
        call("cpvs_build", minMax.ctx.handle, minMax.handle, zTileIndex, zTileNum, stream, C.byref(h))
        return CompressedShadow(h, minMax.ctx)

    def _info(self):
        nl, nw, tv = C.c_uint32(), C.c_uint64(), C.c_int()
        call("cpvs_dag_info", self._h(), C.byref(nl), C.byref(nw), C.byref(tv))
        return nl.value, nw.value, tv.value

    def _h(self):
        if not self.handle:
            raise CpvsError("CompressedShadow", _lib.CPVS_ESTATE, "moved into a container")
        return self.handle

    def getNumLevels(self) -> int:
        return self._info()[0]

    def getTotalVisibility(self) -> int:
        return self._info()[2]

    def num_words(self) -> int:
        return self._info()[1]

    def getDAG(self) -> np.ndarray:
        n = self.num_words()
        out = np.empty(max(n, 1), np.uint32)
        call("cpvs_dag_copy_words", self._h(), out.ctypes.data_as(C.c_void_p), out.size)
        return out[:n]

    def level_stats(self):
        nodes = np.zeros(24, np.uint64)
        uniq = np.zeros(24, np.uint64)
        call("cpvs_dag_level_stats", self._h(), nodes.ctypes.data_as(C.c_void_p), uniq.ctypes.data_as(C.c_void_p), 24)
        n = self.getNumLevels()
        return nodes[:n], uniq[:n]

    def merge_info(self):
        """('table' | 'sort', 32-bit sort-key collisions resolved by scanning)."""
        mode, coll = C.c_int(), C.c_uint64()
        call("cpvs_dag_merge_info", self._h(), C.byref(mode), C.byref(coll))
        return ("table" if mode.value == _lib.CPVS_MERGE_TABLE else "sort"), coll.value

    def has_leafmasks(self) -> bool:
        """True when level 2 holds leafmask nodes (useLeafmasks, src/CompressedShadow.cpp:20-27, in the
        mode the DAG was built in)."""
        v = C.c_int()
        call("cpvs_dag_leafmasks", self._h(), C.byref(v))
        return bool(v.value)

    def traverse_many(self, ndc: np.ndarray, tryLeafmasks: bool = True) -> np.ndarray:
        ndc = np.ascontiguousarray(np.asarray(ndc, np.float32).reshape(-1, 3))
        out = np.zeros(ndc.shape[0], np.int32)
        call("cpvs_dag_traverse", self._h(), ndc.ctypes.data_as(C.c_void_p), ndc.shape[0], int(tryLeafmasks),
             out.ctypes.data_as(C.c_void_p))
        return out

    def traverse(self, position, tryLeafmasks: bool = True) -> int:
        """CompressedShadow::traverse(vec3 ndc, bool tryLeafmasks) (src/CompressedShadow.cpp:404-463)."""
        return int(self.traverse_many(np.asarray(position, np.float32).reshape(1, 3), tryLeafmasks)[0])

    def setTopLevels(self, levels: int):
        """The DAG's own lookup top table (cpvs_dag_set_top_levels): -1 automatic, 0 none, 1..9."""
        call("cpvs_dag_set_top_levels", self._h(), levels)

    def top_levels(self) -> int:
        k = C.c_int()
        call("cpvs_dag_top_levels", self._h(), C.byref(k))
        return k.value

    def evaluate(self, ndc: DeviceArray, visibilities: Optional[DeviceArray] = None, count: Optional[int] = None,
                 offset: int = 0, stride: Optional[int] = None, stream=None) -> DeviceArray:
        """Device lookup of NDC points as a one-DAG container with an identity light matrix
        (src/CompressedShadowContainer.h:27-32 + shader/traverse.cs); 0/1 per point."""
        st = ndc.shape[1] if stride is None else stride
        n = (ndc.shape[0] - offset) if count is None else count
        if visibilities is None:
            visibilities = DeviceArray((n,), np.uint8, self.ctx)
        base = (ndc.ptr.value or 0) + 4 * offset * (ndc.shape[1] if len(ndc.shape) > 1 else 1)
        call("cpvs_dag_evaluate", self._h(), C.c_void_p(base), st, n, visibilities.ptr, stream)
        return visibilities

    def serialize(self) -> bytes:
        """The DAG in the library's byte format (64-byte header + words)."""
        n = C.c_uint64()
        call("cpvs_dag_serialize", self._h(), None, 0, C.byref(n))
        buf = C.create_string_buffer(n.value)
        call("cpvs_dag_serialize", self._h(), buf, n.value, C.byref(n))
        return buf.raw[:n.value]

    @staticmethod
    def deserialize(data: bytes, ctx: Optional[Context] = None) -> "CompressedShadow":
        ctx = ctx or default_context()
        h = C.c_void_p()
        call("cpvs_dag_deserialize", ctx.handle, data, len(data), C.byref(h))
        return CompressedShadow(h, ctx)

    def save(self, path: str):
        call("cpvs_dag_save", self._h(), os.fsencode(path))

    @staticmethod
    def load(path: str, ctx: Optional[Context] = None) -> "CompressedShadow":
        ctx = ctx or default_context()
        h = C.c_void_p()
        call("cpvs_dag_load", ctx.handle, os.fsencode(path), C.byref(h))
        return CompressedShadow(h, ctx)

    def _release(self):
        h, self.handle = self.handle, None
        return h

    def close(self):
        if getattr(self, "handle", None):
            _lib.load().cpvs_dag_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class CompressedShadowContainer:
    """CompressedShadowContainer (src/CompressedShadowContainer.h:18-92)."""

    def __init__(self, length_or_shadow, ctx: Optional[Context] = None):
        if isinstance(length_or_shadow, CompressedShadow):
            self.ctx = length_or_shadow.ctx
            self._create(1)
            self.set(length_or_shadow, 0, 0, 0)
        else:
            self.ctx = ctx or default_context()
            self._create(int(length_or_shadow))

    def _create(self, length):
        h = C.c_void_p()
        call("cpvs_container_create", self.ctx.handle, length, C.byref(h))
        self.handle = h
        self.length = length
        _track(self)

    def set(self, shadow: CompressedShadow, x: int, y: int, z: int):
        call("cpvs_container_set", self.handle, shadow._h(), x, y, z)
        shadow._release()  # ownership moves into the container, as with unique_ptr

    def get(self, x: int, y: int, z: int) -> Optional[int]:
        """The slot's DAG handle (an int), or None after freeOnCPU."""
        h = C.c_void_p()
        call("cpvs_container_get", self.handle, x, y, z, C.byref(h))
        return h.value

    def setFilterSize(self, size: int):
        call("cpvs_container_set_filter_size", self.handle, size)

    def setTopLevels(self, levels: int):
        """Levels of the lookup top table built by copyToGPU: -1 automatic, 0 none, 1..9."""
        call("cpvs_container_set_top_levels", self.handle, levels)

    def top_levels(self) -> int:
        """Levels of the top table the last copyToGPU built (0: none)."""
        k = C.c_int()
        call("cpvs_container_top_levels", self.handle, C.byref(k))
        return k.value

    def tuneTopLevels(self, positions: DeviceArray, candidates=(0, 3, 4, 5, 6, 7), lightViewProj=None,
                      visibilities: Optional[DeviceArray] = None, reps: int = 3):
        """Times evaluate() on these positions with each top-table setting (cpvs_container_tune_top_levels),
        keeps the fastest.  Returns (best setting, {setting: ms})."""
        count, stride = positions.shape
        if visibilities is None:
            visibilities = DeviceArray((count,), np.uint8, self.ctx)
        m = None
        if lightViewProj is not None:
            m = np.ascontiguousarray(np.asarray(lightViewProj, np.float32).T.reshape(16))
        cand = (C.c_int * len(candidates))(*candidates)
        ms = np.zeros(len(candidates), np.float64)
        best = C.c_int()
        call("cpvs_container_tune_top_levels", self.handle, positions.ptr, stride, count,
             None if m is None else m.ctypes.data_as(C.c_void_p), visibilities.ptr, cand, len(candidates), reps,
             C.byref(best), ms.ctypes.data_as(C.c_void_p))
        return int(candidates[best.value]), {int(k): float(t) for k, t in zip(candidates, ms)}

    def copyToGPU(self):
        call("cpvs_container_copy_to_gpu", self.handle)

    def freeOnCPU(self):
        call("cpvs_container_free_on_cpu", self.handle)

    def moveToGPU(self):
        call("cpvs_container_move_to_gpu", self.handle)

    def info(self):
        dl, gl, nw = C.c_int(), C.c_int(), C.c_uint64()
        call("cpvs_container_info", self.handle, C.byref(dl), C.byref(gl), C.byref(nw))
        return dl.value, gl.value, nw.value

    def words_and_grid(self):
        _, _, nw = self.info()
        words = np.empty(max(nw, 1), np.uint32)
        grid = np.empty(self.length ** 3, np.uint32)
        call("cpvs_container_copy_words", self.handle, words.ctypes.data_as(C.c_void_p), words.size,
             grid.ctypes.data_as(C.c_void_p), grid.size)
        return words[:nw], grid

    def evaluate(self, positions: DeviceArray, lightViewProj=None, visibilities: Optional[DeviceArray] = None,
                 stream=None) -> DeviceArray:
        """evaluate(positionsWS, lightViewProj, visibilities) (src/CompressedShadowContainer.cpp:93-124).

        positions: DeviceArray (count, 3|4) float32; lightViewProj: 4x4 (numpy, applied as
        lightViewProj @ (x, y, z, 1)) or None for identity; returns a (count,) uint8 DeviceArray.
        """
        count, stride = positions.shape
        if visibilities is None:
            visibilities = DeviceArray((count,), np.uint8, self.ctx)
        m = None
        if lightViewProj is not None:
            m = np.ascontiguousarray(np.asarray(lightViewProj, np.float32).T.reshape(16))  # column-major
        call("cpvs_container_evaluate", self.handle, positions.ptr, stride, count,
             None if m is None else m.ctypes.data_as(C.c_void_p), visibilities.ptr, stream)
        return visibilities

    def evaluate_pcf(self, positions: DeviceArray, lightViewProj=None, visibilities: Optional[DeviceArray] = None,
                     stream=None) -> DeviceArray:
        """Percentage-closer filtered visibility (float32 in [0, 1]) with the filter size of
        setFilterSize: the fraction of lit leaf cells over the f x f neighbourhood of each path."""
        n = positions.shape[0]
        if visibilities is None:
            visibilities = DeviceArray((n,), np.float32, self.ctx)
        m = None
        if lightViewProj is not None:
            m = np.ascontiguousarray(np.asarray(lightViewProj, np.float32).T.reshape(16))
        call("cpvs_container_evaluate_pcf", self.handle, positions.ptr, positions.shape[1], n,
             None if m is None else m.ctypes.data_as(C.c_void_p), visibilities.ptr, stream)
        return visibilities

    def evaluate_image(self, positions: DeviceArray, lightViewProj=None, visibilities: Optional[DeviceArray] = None,
                       stream=None) -> DeviceArray:
        """Same over an (H, W, 4) rgba32f position image; returns an (H, W) uint8 DeviceArray."""
        H, W, _ = positions.shape
        if visibilities is None:
            visibilities = DeviceArray((H, W), np.uint8, self.ctx)
        m = None
        if lightViewProj is not None:
            m = np.ascontiguousarray(np.asarray(lightViewProj, np.float32).T.reshape(16))
        call("cpvs_container_evaluate_image", self.handle, positions.ptr, W, H,
             None if m is None else m.ctypes.data_as(C.c_void_p), visibilities.ptr, stream)
        return visibilities

    def close(self):
        if getattr(self, "handle", None):
            _lib.load().cpvs_container_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DagSet:
    """Many DAGs queried in one launch (cpvs_dagset): query i uses DAG ids[i], or i mod len(set)."""

    def __init__(self, shadows: Sequence[CompressedShadow], ctx: Optional[Context] = None):
        self.ctx = ctx or (shadows[0].ctx if len(shadows) else default_context())
        self.handle = None
        self.shadows = list(shadows)  # kept alive: the set borrows their words
        arr = (C.c_void_p * len(self.shadows))(*[sh._h() for sh in self.shadows])
        h = C.c_void_p()
        call("cpvs_dagset_create", self.ctx.handle, arr, len(self.shadows), C.byref(h))
        self.handle = h
        _track(self)

    def __len__(self):
        return len(self.shadows)

    def evaluate(self, ndc: DeviceArray, ids: Optional[DeviceArray] = None, visibilities: Optional[DeviceArray] = None,
                 stream=None) -> DeviceArray:
        n = ndc.shape[0]
        if visibilities is None:
            visibilities = DeviceArray((n,), np.uint8, self.ctx)
        call("cpvs_dagset_evaluate", self.handle, ndc.ptr, ndc.shape[1], n, ids.ptr if ids is not None else None,
             visibilities.ptr, stream)
        return visibilities

    def close(self):
        if getattr(self, "handle", None):
            _lib.load().cpvs_dagset_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Serving loop (cpvs_server): builds on concurrent stream lanes, lookups batched
    continuously, DAGs resident under an HBM budget.  serial=True: one request at a time in
    arrival order (the plain path).  Requests borrow their device buffers until waited for."""

    STAT_NAMES = ("builds", "lookups", "batches", "queries", "evicted", "resident", "resident_bytes",
                  "scratch_bytes", "reserved_bytes", "budget", "trims", "cached_bytes")

    def __init__(self, lanes: int = 4, budget_bytes: int = 200 << 30, serial: bool = False,
                 ctx: Optional[Context] = None):
        self.ctx = ctx or default_context()
        self.handle = None
        h = C.c_void_p()
        call("cpvs_server_create", self.ctx.handle, lanes, budget_bytes, int(serial), C.byref(h))
        self.handle = h
        _track(self)

    def build(self, depth: DeviceArray, zTileIndex: int = 0, zTileNum: int = 1) -> int:
        t = C.c_uint64()
        call("cpvs_server_submit_build", self.handle, depth.ptr, depth.shape[0], zTileIndex, zTileNum, C.byref(t))
        return t.value

    def lookup(self, dag: int, ndc: DeviceArray, vis: DeviceArray, count: Optional[int] = None, offset: int = 0,
               vis_offset: int = 0) -> int:
        """Queries ndc[offset:offset+count] (stride = ndc's row length) against the resident DAG
        of build ticket `dag`; results to vis[vis_offset:...]."""
        stride = ndc.shape[1]
        n = (ndc.shape[0] - offset) if count is None else count
        t = C.c_uint64()
        call("cpvs_server_submit_lookup", self.handle, dag, C.c_void_p(ndc.ptr.value + 4 * stride * offset), stride, n,
             C.c_void_p(vis.ptr.value + vis_offset), C.byref(t))
        return t.value

    def wait(self, ticket: int) -> float:
        """Latency (ms, submission to completion) of a finished request; raises on failure."""
        ms = C.c_double()
        call("cpvs_server_wait", self.handle, ticket, C.byref(ms))
        return ms.value

    def poll(self, ticket: int) -> bool:
        d = C.c_int()
        call("cpvs_server_poll", self.handle, ticket, C.byref(d))
        return bool(d.value)

    def dag_words(self, dag: int) -> np.ndarray:
        h = C.c_void_p()
        call("cpvs_server_dag", self.handle, dag, C.byref(h))
        nw = C.c_uint64()
        call("cpvs_dag_info", h, None, C.byref(nw), None)
        out = np.empty(max(nw.value, 1), np.uint32)
        call("cpvs_dag_copy_words", h, out.ctypes.data_as(C.c_void_p), out.size)
        return out[:nw.value]

    def release(self, dag: int):
        call("cpvs_server_release", self.handle, dag)

    def resident(self):
        n = C.c_uint32()
        call("cpvs_server_resident", self.handle, None, 0, C.byref(n))
        ids = np.zeros(max(n.value, 1), np.uint64)
        call("cpvs_server_resident", self.handle, ids.ctypes.data_as(C.c_void_p), ids.size, C.byref(n))
        return sorted(int(v) for v in ids[:min(n.value, ids.size)])

    def stats(self) -> dict:
        v = np.zeros(len(self.STAT_NAMES), np.uint64)
        call("cpvs_server_stats", self.handle, v.ctypes.data_as(C.c_void_p), len(v))
        return {k: int(x) for k, x in zip(self.STAT_NAMES, v)}

    def close(self):
        if getattr(self, "handle", None):
            _lib.load().cpvs_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def dag_check(data: bytes):
    """Checks a serialized DAG without a GPU; returns (numLevels, words, nodes per level)."""
    nl, nw = C.c_uint32(), C.c_uint64()
    nodes = np.zeros(24, np.uint64)
    call("cpvs_dag_check", data, len(data), C.byref(nl), C.byref(nw), nodes.ctypes.data_as(C.c_void_p), 24)
    return nl.value, nw.value, nodes[:nl.value]


def _batch_items(depths: Sequence[DeviceArray], tiles=None):
    items = (_lib.BatchItem * len(depths))()
    for i, d in enumerate(depths):
        t, T = (0, 1) if tiles is None else tiles[i]
        items[i] = _lib.BatchItem(d.ptr.value, d.shape[0], t, T, 0)
    return items


def batch_config(lanes: int = 4, forest_mtexels: int = 256, leaf_cols: int = 8, leaf_waves: int = 1,
                 graph: bool = False) -> "_lib.BatchConfig":
    """A cpvs_batch_config (forest_mtexels = 0: one map per launch chain on `lanes` streams)."""
    return _lib.BatchConfig(lanes, forest_mtexels, leaf_cols, leaf_waves, int(graph))


def build_batch(depths: Sequence[DeviceArray], tiles=None, lanes: int = 0, ctx: Optional[Context] = None,
                config=None):
    """MinMaxHierarchy + CompressedShadow::create for every depth map at once (cpvs_build_batch).

    depths: square float32 DeviceArrays; tiles: optional (zTileIndex, zTileNum) per map; lanes: 0 =
    the context's tuned count; config: a batch_config() (the context's when None).  Returns
    (shadows, done_ms) with done_ms[i] the completion time of map i in ms since the call began."""
    ctx = ctx or default_context()
    n = len(depths)
    items = _batch_items(depths, tiles)
    out = (C.c_void_p * max(n, 1))()
    done = np.zeros(max(n, 1), np.float64)
    if config is not None:
        if lanes:
            config = _lib.BatchConfig(lanes, config.forest_mtexels, config.leaf_cols, config.leaf_waves, config.graph)
        call("cpvs_build_batch_config", ctx.handle, items, n, C.byref(config), out, done.ctypes.data_as(C.c_void_p))
    else:
        call("cpvs_build_batch", ctx.handle, items, n, lanes, out, done.ctypes.data_as(C.c_void_p))
    return [CompressedShadow(C.c_void_p(out[i]), ctx) for i in range(n)], done[:n]


def get_batch_config(ctx: Optional[Context] = None) -> dict:
    ctx = ctx or default_context()
    c = _lib.BatchConfig()
    call("cpvs_ctx_batch_config", ctx.handle, C.byref(c))
    return c.as_dict()


def set_batch_config(config, ctx: Optional[Context] = None):
    ctx = ctx or default_context()
    call("cpvs_ctx_set_batch_config", ctx.handle, C.byref(config))


def tune_batch_config(depths: Sequence[DeviceArray], candidates, tiles=None, reps: int = 1,
                      ctx: Optional[Context] = None):
    """The autotuner (cpvs_tune_batch_config): times every candidate batch_config() on this batch,
    makes the fastest the context's configuration.  Returns (best index, [ms per candidate])."""
    ctx = ctx or default_context()
    items = _batch_items(depths, tiles)
    arr = (_lib.BatchConfig * len(candidates))(*candidates)
    ms = np.zeros(len(candidates), np.float64)
    best = C.c_uint32()
    call("cpvs_tune_batch_config", ctx.handle, items, len(depths), arr, len(candidates), reps, C.byref(best),
         ms.ctypes.data_as(C.c_void_p))
    return int(best.value), [float(m) for m in ms]


def _tune_key(ctx: Context, depths) -> str:
    import collections
    hist = collections.Counter(int(d.shape[0]) for d in depths)
    return ctx.device_name() + "|" + ",".join(f"{k}x{hist[k]}" for k in sorted(hist))


def tune_batch(depths: Sequence[DeviceArray], tiles=None, candidates=(1, 2, 3, 4, 6, 8), reps: int = 1,
               cache: Optional[str] = None, ctx: Optional[Context] = None):
    """Pick the batch lane count for this mix of map sizes (cpvs_tune_batch) and make it the
    context's default.  With cache (a JSON file path), a result stored for the same device and
    size mix is reused without timing.  Returns (best, {lanes: ms} or None when cached)."""
    ctx = ctx or default_context()
    key = _tune_key(ctx, depths)
    table = {}
    if cache and os.path.exists(cache):
        try:
            with open(cache) as f:
                table = json.load(f)
        except (OSError, ValueError):
            table = {}
        if key in table:
            best = int(table[key]["lanes"])
            call("cpvs_ctx_set_batch_lanes", ctx.handle, best)
            return best, None
    items = _batch_items(depths, tiles)
    cand = (C.c_uint32 * len(candidates))(*candidates)
    ms = np.zeros(len(candidates), np.float64)
    best = C.c_uint32()
    call("cpvs_tune_batch", ctx.handle, items, len(depths), cand, len(candidates), reps, C.byref(best),
         ms.ctypes.data_as(C.c_void_p))
    timings = {int(c): float(m) for c, m in zip(candidates, ms)}
    if cache:
        table[key] = {"lanes": int(best.value), "ms": timings}
        tmp = cache + ".tmp"
        with open(tmp, "w") as f:
            json.dump(table, f, indent=1, sort_keys=True)
        os.replace(tmp, cache)
    return int(best.value), timings


def gen_depth(N: int, seed: int, nboxes: int, noise: bool, x0: int = 0, y0: int = 0, tileN: Optional[int] = None,
              ctx: Optional[Context] = None) -> DeviceArray:
    """Synthetic depth tile [x0, x0+tileN)^2 of an N x N map, generated on the device."""
    ctx = ctx or default_context()
    tileN = N if tileN is None else tileN
    out = DeviceArray((tileN, tileN), np.float32, ctx)
    call("cpvs_gen_depth", ctx.handle, out.ptr, N, seed, nboxes, int(noise), x0, y0, tileN, None)
    return out


def gen_queries_uniform(seed: int, count: int, stride: int = 4, ctx: Optional[Context] = None) -> DeviceArray:
    ctx = ctx or default_context()
    out = DeviceArray((count, stride), np.float32, ctx)
    call("cpvs_gen_queries_uniform", ctx.handle, out.ptr, seed, count, stride, None)
    ctx.synchronize()
    return out


def gen_queries_uniform_range(seed: int, first: int, step: int, count: int, stride: int = 3,
                              ctx: Optional[Context] = None) -> DeviceArray:
    """Points first, first+step, ... of the gen_queries_uniform stream."""
    ctx = ctx or default_context()
    out = DeviceArray((count, stride), np.float32, ctx)
    call("cpvs_gen_queries_uniform_range", ctx.handle, out.ptr, seed, first, step, count, stride, None)
    ctx.synchronize()
    return out


def gen_queries_gbuffer(depth: DeviceArray, seed: int, W: int = 3840, H: int = 2160,
                        ctx: Optional[Context] = None) -> DeviceArray:
    ctx = ctx or default_context()
    out = DeviceArray((2 * W * H, 4), np.float32, ctx)
    call("cpvs_gen_queries_gbuffer", ctx.handle, out.ptr, depth.ptr, depth.shape[0], seed, W, H, None)
    ctx.synchronize()
    return out
