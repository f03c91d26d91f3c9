"""cpvs_build_noalloc (the build that allocates no device memory) against cpvs_build, through the
C ABI: identical words on two calls in a row, the second with a smaller map (the fixed scratch
keeps the first call's contents), plus the capacity errors.  Needs a GPU."""
import ctypes as C

import numpy as np
import pytest

import cpvs_amd as cp
from cpvs_amd import _lib
import oracle as O

pytestmark = pytest.mark.gpu
SEED0 = 0x43505653


def _say(*a):
    print("[case]", *a, flush=True)


def _noalloc(mm, t, T, cap):
    out = cp.DeviceArray((max(cap, 1),), np.uint32)
    nw, root = C.c_uint64(), C.c_uint32()
    _lib.call("cpvs_build_noalloc", mm.ctx.handle, mm.handle, t, T, None, out.ptr, cap, C.byref(nw), C.byref(root))
    words = out.numpy()[:nw.value]
    out.free()
    return words, root.value


def test_noalloc_matches_build_on_two_calls():
    cases = [(1024, 32, 0, 1), (256, 8, 1, 2)]  # the second call's map is smaller
    for N, boxes, t, T in cases:
        _say("noalloc", N, t, T)
        D = O.gen_depth(N, SEED0 + N + 5, boxes, True)
        mm = cp.MinMaxHierarchy(D)
        ref = cp.CompressedShadow.create(mm, t, T).getDAG()
        words, root = _noalloc(mm, t, T, len(ref) + 64)
        assert np.array_equal(words, ref), (N, len(words), len(ref))
        assert root == int(ref[0])
        mm.close()
    assert _lib.load().cpvs_build_noalloc_scratch_bytes() == 64 << 20


def test_noalloc_root_only_and_limits():
    _say("noalloc limits")
    mm = cp.MinMaxHierarchy(np.full((64, 64), 1.0, np.float32))  # fully lit: the DAG is one word
    words, root = _noalloc(mm, 0, 1, 1)
    assert words.tolist() == [0x5555] and root == 0x5555
    mm.close()
    D = O.gen_depth(512, SEED0 + 9, 16, True)
    mm = cp.MinMaxHierarchy(D)
    n = len(cp.CompressedShadow.create(mm).getDAG())
    with pytest.raises(cp.CpvsError) as e:
        _noalloc(mm, 0, 1, n - 1)  # one word short of the output
    assert e.value.code == _lib.CPVS_EINVAL
    words, _ = _noalloc(mm, 0, 1, n)  # exactly enough
    assert len(words) == n
    mm.close()
    # White noise at 2048^2: ~4 M distinct leaves need far more than the fixed scratch.
    big = cp.MinMaxHierarchy(np.random.default_rng(3).random((2048, 2048), dtype=np.float32))
    with pytest.raises(cp.CpvsError) as e:
        _noalloc(big, 0, 1, 1 << 26)
    assert e.value.code == _lib.CPVS_EINVAL and "scratch" in str(e.value)
    big.close()
