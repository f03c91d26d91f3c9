"""ctypes loader for the in-tree HIP library ``cpvs_amd/libcpvs_amd.so`` (C ABI: include/cpvs.h).

Loading fails loudly when the library is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcpvs_amd.so")

CPVS_OK, CPVS_EINVAL, CPVS_ENOMEM, CPVS_EHIP, CPVS_EINTERNAL, CPVS_ESTATE, CPVS_EBUDGET = range(7)
STATUS_NAMES = {0: "CPVS_OK", 1: "CPVS_EINVAL", 2: "CPVS_ENOMEM", 3: "CPVS_EHIP", 4: "CPVS_EINTERNAL",
                5: "CPVS_ESTATE", 6: "CPVS_EBUDGET"}
CPVS_MERGE_TABLE, CPVS_MERGE_SORT = 0, 1
CPVS_PYR_ALL_LEVELS = 1
CPVS_PYR_BORROW_DEPTH = 2


class CpvsError(RuntimeError):
    def __init__(self, fn: str, code: int, msg: str):
        super().__init__(f"{fn} -> {STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code


vp = C.c_void_p
u32, u64, i32 = C.c_uint32, C.c_uint64, C.c_int
PP = C.POINTER(C.c_void_p)

# name -> argtypes (all return int status except cpvs_last_error)
_SIGS = {
    "cpvs_ctx_create": [i32, PP],
    "cpvs_ctx_destroy": [vp],
    "cpvs_ctx_synchronize": [vp],
    "cpvs_ctx_stream": [vp, PP],
    "cpvs_ctx_profile": [vp, i32],
    "cpvs_ctx_profile_read": [vp, C.c_char_p, u64],
    "cpvs_ctx_set_merge": [vp, i32],
    "cpvs_ctx_set_hash_bits": [vp, u32, u32],
    "cpvs_dag_merge_info": [vp, C.POINTER(i32), C.POINTER(u64)],
    "cpvs_ctx_set_leafmasks": [vp, i32],
    "cpvs_ctx_release_scratch": [vp, C.POINTER(u64)],
    "cpvs_ctx_leafmasks": [vp, C.POINTER(i32)],
    "cpvs_dag_leafmasks": [vp, C.POINTER(i32)],
    "cpvs_pyramid_build": [vp, vp, i32, u32, u32, vp, PP],
    "cpvs_pyramid_num_levels": [vp, C.POINTER(i32)],
    "cpvs_pyramid_get": [vp, u32, u32, u32, C.POINTER(C.c_float), C.POINTER(C.c_float)],
    "cpvs_pyramid_copy_level": [vp, u32, vp, u64],
    "cpvs_pyramid_destroy": [vp],
    "cpvs_estimate_build": [vp, vp, u32, u32, vp, C.POINTER(u64), C.POINTER(u64)],
    "cpvs_build": [vp, vp, u32, u32, vp, PP],
    "cpvs_build_noalloc": [vp, vp, u32, u32, vp, vp, u64, C.POINTER(u64), C.POINTER(u32)],
    "cpvs_build_noalloc_scratch_bytes": [],
    "cpvs_dag_info": [vp, C.POINTER(u32), C.POINTER(u64), C.POINTER(i32)],
    "cpvs_dag_level_stats": [vp, vp, vp, i32],
    "cpvs_dag_copy_words": [vp, vp, u64],
    "cpvs_dag_device_words": [vp, PP],
    "cpvs_dag_traverse": [vp, vp, u64, i32, vp],
    "cpvs_dag_destroy": [vp],
    "cpvs_container_create": [vp, u32, PP],
    "cpvs_container_set": [vp, vp, u32, u32, u32],
    "cpvs_container_get": [vp, u32, u32, u32, PP],
    "cpvs_container_set_filter_size": [vp, u32],
    "cpvs_container_set_top_levels": [vp, C.c_int],
    "cpvs_dag_set_top_levels": [vp, C.c_int],
    "cpvs_dag_top_levels": [vp, C.POINTER(C.c_int)],
    "cpvs_container_top_levels": [vp, C.POINTER(C.c_int)],
    "cpvs_container_tune_top_levels": [vp, vp, u32, u64, vp, vp, vp, u32, u32, C.POINTER(C.c_int), vp],
    "cpvs_container_copy_to_gpu": [vp],
    "cpvs_container_free_on_cpu": [vp],
    "cpvs_container_move_to_gpu": [vp],
    "cpvs_container_info": [vp, C.POINTER(i32), C.POINTER(i32), C.POINTER(u64)],
    "cpvs_container_copy_words": [vp, vp, u64, vp, u64],
    "cpvs_container_evaluate": [vp, vp, u32, u64, vp, vp, vp],
    "cpvs_container_evaluate_image": [vp, vp, u32, u32, vp, vp, vp],
    "cpvs_container_destroy": [vp],
    "cpvs_device_alloc": [vp, u64, PP],
    "cpvs_device_free": [vp, vp],
    "cpvs_memcpy_h2d": [vp, vp, vp, u64],
    "cpvs_memcpy_d2h": [vp, vp, vp, u64],
    "cpvs_memcpy_d2d": [vp, vp, vp, u64, vp],
    "cpvs_device_copy": [vp, vp, vp, u64, vp],
    "cpvs_event_record": [vp, vp, PP],
    "cpvs_event_elapsed_ms": [vp, vp, C.POINTER(C.c_float)],
    "cpvs_event_destroy": [vp],
    "cpvs_gen_depth": [vp, vp, u32, u64, u32, i32, u32, u32, u32, vp],
    "cpvs_gen_queries_uniform": [vp, vp, u64, u64, u32, vp],
    "cpvs_gen_queries_gbuffer": [vp, vp, vp, u32, u64, u32, u32, vp],
    "cpvs_gen_queries_uniform_range": [vp, vp, u64, u64, u64, u64, u32, vp],
    "cpvs_build_batch": [vp, vp, u32, u32, vp, vp],
    "cpvs_build_batch_config": [vp, vp, u32, vp, vp, vp],
    "cpvs_ctx_set_batch_config": [vp, vp],
    "cpvs_ctx_batch_config": [vp, vp],
    "cpvs_tune_batch_config": [vp, vp, u32, vp, u32, u32, C.POINTER(u32), vp],
    "cpvs_ctx_set_batch_lanes": [vp, u32],
    "cpvs_ctx_batch_lanes": [vp, C.POINTER(u32)],
    "cpvs_ctx_device_name": [vp, C.c_char_p, u64],
    "cpvs_tune_batch": [vp, vp, u32, vp, u32, u32, C.POINTER(u32), vp],
    "cpvs_dagset_create": [vp, vp, u32, PP],
    "cpvs_dagset_evaluate": [vp, vp, u32, u64, vp, vp, vp],
    "cpvs_dagset_destroy": [vp],
    "cpvs_dag_evaluate": [vp, vp, u32, u64, vp, vp],
    "cpvs_dag_serialize": [vp, vp, u64, C.POINTER(u64)],
    "cpvs_dag_check": [vp, u64, C.POINTER(u32), C.POINTER(u64), vp, i32],
    "cpvs_dag_deserialize": [vp, vp, u64, PP],
    "cpvs_dag_save": [vp, C.c_char_p],
    "cpvs_dag_load": [vp, C.c_char_p, PP],
    "cpvs_container_evaluate_pcf": [vp, vp, u32, u64, vp, vp, vp],
    "cpvs_server_create": [vp, u32, u64, i32, PP],
    "cpvs_server_submit_build": [vp, vp, u32, u32, u32, C.POINTER(u64)],
    "cpvs_server_submit_lookup": [vp, u64, vp, u32, u64, vp, C.POINTER(u64)],
    "cpvs_server_wait": [vp, u64, C.POINTER(C.c_double)],
    "cpvs_server_poll": [vp, u64, C.POINTER(i32)],
    "cpvs_server_dag": [vp, u64, PP],
    "cpvs_server_release": [vp, u64],
    "cpvs_server_resident": [vp, vp, u32, C.POINTER(u32)],
    "cpvs_server_stats": [vp, vp, u32],
    "cpvs_server_destroy": [vp],
}


class BatchItem(C.Structure):
    """cpvs_batch_item: one map of a batch build (device depth, N, zTileIndex, zTileNum)."""
    _fields_ = [("depth", C.c_void_p), ("N", C.c_uint32), ("zTileIndex", C.c_uint32), ("zTileNum", C.c_uint32),
                ("reserved", C.c_uint32)]

class BatchConfig(C.Structure):
    """cpvs_batch_config: how a batch runs (lanes, forest size in Mtexels (0 = one map per launch
    chain), leaf kernel columns per wave and waves per workgroup, hipGraph capture)."""
    _fields_ = [("lanes", C.c_uint32), ("forest_mtexels", C.c_uint32), ("leaf_cols", C.c_uint32),
                ("leaf_waves", C.c_uint32), ("graph", C.c_uint32), ("reserved", C.c_uint32 * 3)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k in ("lanes", "forest_mtexels", "leaf_cols", "leaf_waves", "graph")}


_lib = None


def load():
    """Load the HIP library (raises OSError with a clear message when it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OSError(f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
                      f"(make -C cpvs_amd/csrc); cpvs_amd has no CPU fallback")
    lib = C.CDLL(LIB_PATH)
    lib.cpvs_last_error.restype = C.c_char_p
    lib.cpvs_last_error.argtypes = []
    for name, args in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = C.c_int
        fn.argtypes = args
    lib.cpvs_build_noalloc_scratch_bytes.restype = C.c_size_t
    _lib = lib
    return lib


def exported_names():
    return ["cpvs_last_error"] + list(_SIGS)


def call(name: str, *args) -> None:
    """Call an entry point and raise CpvsError on a non-zero status."""
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != CPVS_OK:
        raise CpvsError(name, rc, lib.cpvs_last_error().decode(errors="replace"))
