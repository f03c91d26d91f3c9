"""Maps with scattered NaN and infinite depths, whose level-by-level node counts exceed the octree's
nodes (cells partial under a parent cell that is not; see tests/nanmaps.py): every build path --
plain with the table and the sort merge, forests, without leafmasks, the server and the
caller-buffer build -- gives the oracle's words.  Needs a GPU."""
import numpy as np
import pytest

import cpvs_amd as cp
import oracle as O
from tests.nanmaps import excess, sparse_nan_map

pytestmark = pytest.mark.gpu
SEED0 = 0x43505653


def _say(*a):
    print("[case]", *a, flush=True)


SPECS = [(64, 1, 0.02, 0.01), (256, 1, 0.002, 0.001), (512, 2, 0.002, 0.001), (1024, 4, 0.001, 0.0005)]


@pytest.fixture(scope="module")
def nan_maps():
    out = []
    for k, (N, T, fn, fi) in enumerate(SPECS):
        D = sparse_nan_map(N, SEED0 + 37 * k + N, fn, fi)
        pyr = O.Pyramid(D)
        exp = {lm: [O.create(pyr, t, T, leafmasks=lm)[0] for t in range(T)] for lm in (True, False)}
        ex = {lm: sum(excess(D, t, T, lm) for t in range(T)) for lm in (True, False)}
        out.append((N, T, D, exp, ex))
    # the maps reach the exact-count fallback in both modes
    assert sum(ex[True] for *_, ex in out) > 0 and sum(ex[False] for *_, ex in out) > 0
    return out


def test_plain_builds_match_oracle(ctx, nan_maps):
    for N, T, D, exp, ex in nan_maps:
        mm = cp.MinMaxHierarchy(D)
        for t in range(T):
            _say("nan map plain", N, t, T, "excess", ex[True])
            cs = cp.CompressedShadow.create(mm, t, T)
            assert np.array_equal(cs.getDAG(), exp[True][t]), (N, t, T)


def test_sort_merge_matches_oracle(nan_maps):
    c2 = cp.Context(0)
    try:
        c2.set_merge("sort")
        for N, T, D, exp, _ in nan_maps:
            mm = cp.MinMaxHierarchy(D, ctx=c2)
            for t in range(T):
                _say("nan map sort", N, t, T)
                assert np.array_equal(cp.CompressedShadow.create(mm, t, T).getDAG(), exp[True][t]), (N, t, T)
    finally:
        c2.close()


def test_forests_match_oracle(ctx, nan_maps):
    depths, tiles, want = [], [], []
    for N, T, D, exp, _ in nan_maps:
        d = cp.DeviceArray.from_numpy(D)
        for t in range(T):
            depths.append(d)
            tiles.append((t, T))
            want.append(exp[True][t])
    for mt in (1, 4096):
        _say("nan map forest", mt)
        shadows, _ = cp.build_batch(depths, tiles, config=cp.batch_config(lanes=2, forest_mtexels=mt))
        for k, sh in enumerate(shadows):
            assert np.array_equal(sh.getDAG(), want[k]), (mt, k)


def test_without_leafmasks_matches_oracle(nan_maps):
    c2 = cp.Context(0)
    try:
        c2.set_leafmasks(False)
        for N, T, D, exp, ex in nan_maps:
            mm = cp.MinMaxHierarchy(D, ctx=c2)
            for t in range(T):
                _say("nan map no leafmasks", N, t, T, "excess", ex[False])
                assert np.array_equal(cp.CompressedShadow.create(mm, t, T).getDAG(), exp[False][t]), (N, t, T)
    finally:
        c2.close()


def test_server_matches_oracle(ctx, nan_maps):
    srv = cp.Server(lanes=2)
    try:
        ds = [cp.DeviceArray.from_numpy(D) for _, _, D, _, _ in nan_maps]
        tickets = [(k, t, srv.build(ds[k], t, T)) for k, (N, T, _, _, _) in enumerate(nan_maps) for t in range(T)]
        for k, t, tk in tickets:
            srv.wait(tk)
            _say("nan map server", nan_maps[k][0], t)
            assert np.array_equal(srv.dag_words(tk), nan_maps[k][3][True][t]), (k, t)
    finally:
        srv.close()


def test_caller_buffer_build_matches_oracle(ctx, nan_maps):
    """cpvs_build_noalloc (fixed scratch, the caller's output buffer) on the same maps."""
    import ctypes as C
    from cpvs_amd import _lib
    for N, T, D, exp, _ in nan_maps[:3]:
        mm = cp.MinMaxHierarchy(D)
        for t in range(T):
            _say("nan map noalloc", N, t, T)
            cap = len(exp[True][t]) + 64
            out = cp.DeviceArray((cap,), np.uint32)
            nw, root = C.c_uint64(), C.c_uint32()
            _lib.call("cpvs_build_noalloc", ctx.handle, mm.handle, t, T, None, out.ptr, cap, C.byref(nw), C.byref(root))
            assert np.array_equal(out.numpy()[:nw.value], exp[True][t]), (N, t, T)
            out.free()
        mm.close()
