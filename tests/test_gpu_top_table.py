"""The container's lookup top table (per grid cell, the walk's state after its first K levels):
lookups, PCF and image lookups are the same with every K as with none and as the oracle's walk
(shader/traverse.cs:75-133), with leafmasks and without, over containers whose grid holds
sentinel cells, and with positions outside the unit cube.  Needs a GPU."""
import numpy as np
import pytest

import cpvs_amd as cp
import oracle as O

pytestmark = pytest.mark.gpu
SEED0 = 0x43505653
MAX_K = 9
AUTO_BYTES = 4 << 20


def _say(*a):
    print("[case]", *a, flush=True)


def _expected_k(req, dl, cells, leaf):
    lead = dl - 1 - (3 if leaf else 1)
    if req < 0:
        k = min(lead, MAX_K)
        while k > 0 and cells * 8 ** k * 8 > AUTO_BYTES:
            k -= 1
        if k < 3:
            k = 0
    else:
        k = min(req, lead)
    return max(k, 0)


def _queries(seed, n, finite=False):
    """Positions in [-1.1, 1.1)^3 (w = 1), 5% of the coordinates replaced by special values (only
    finite ones with finite=True: the DAG-set lookups take NDC as given, while the oracle's identity
    product turns an infinity into NaN)."""
    rng = np.random.default_rng(seed)
    q = (rng.random((n, 4), dtype=np.float32) * 2.2 - 1.1).astype(np.float32)
    special = np.array([-1.0, 1.0, 0.0] if finite else [np.nan, np.inf, -np.inf, -1.0, 1.0, 0.0], np.float32)
    pick = rng.random(q.shape) < 0.05
    q[pick] = special[rng.integers(0, special.size, q.shape)][pick]
    q[:, 3] = 1
    return q


def _sweep(c, q, want, dl, cells, leaf, ctx=None, pcf=False, image=False):
    """Every top-table setting gives `want` (and one PCF / image answer across settings)."""
    dq = cp.DeviceArray.from_numpy(q, ctx=ctx)
    pcf0 = None
    for req in [0, 1, 2, 3, 4, 5, 6, 7, 8, 9, -1]:
        if cells * 8 ** min(req, MAX_K) * 8 > (1 << 30):
            continue
        c.setTopLevels(req)
        c.copyToGPU()
        k = c.top_levels()
        _say("top table", req, "->", k, "dl", dl, "cells", cells, "leaf", leaf)
        assert k == _expected_k(req, dl, cells, leaf), (req, k)
        assert np.array_equal(c.evaluate(dq).numpy(), want), req
        if pcf:
            got = c.evaluate_pcf(dq).numpy()
            if pcf0 is None:
                pcf0 = got
            assert np.array_equal(got, pcf0), req
        if image:
            W = 97
            H = q.shape[0] // W
            got = c.evaluate_image(cp.DeviceArray.from_numpy(q[:W * H].reshape(H, W, 4), ctx=ctx)).numpy()
            assert np.array_equal(got.reshape(-1), want[:W * H]), req
    dq.free()


@pytest.mark.parametrize("N", [16, 64, 256, 1024])
def test_single_dag_every_k(ctx, N):
    D = O.gen_depth(N, SEED0 + 700 + N, max(1, N // 32), True)
    cs = cp.CompressedShadow.create(cp.MinMaxHierarchy(D))
    w, n = cs.getDAG(), cs.getNumLevels()
    c = cp.CompressedShadowContainer(cs)
    q = _queries(N, 1 << 16)
    want = O.evaluate(w, O.top_level_grid([w]), n, 0, q, nthreads=8)
    c.setFilterSize(3)
    _sweep(c, q, want, n, 1, N >= 16, pcf=True, image=True)
    c.close()


@pytest.mark.parametrize("N", [64, 512])
def test_single_dag_without_leafmasks(N):
    nctx = cp.Context(0)
    try:
        nctx.set_leafmasks(False)
        D = O.gen_depth(N, SEED0 + 710 + N, max(1, N // 32), True)
        cs = cp.CompressedShadow.create(cp.MinMaxHierarchy(D, ctx=nctx))
        w, n = cs.getDAG(), cs.getNumLevels()
        c = cp.CompressedShadowContainer(cs)
        q = _queries(N + 1, 1 << 16)
        want = O.evaluate(w, O.top_level_grid([w]), n, 0, q, nthreads=8, leafmasks=False)
        c.setFilterSize(2)
        _sweep(c, q, want, n, 1, False, ctx=nctx, pcf=True, image=True)
        c.close()
    finally:
        nctx.close()


def test_tiled_container_with_sentinel_cells(ctx):
    """2x2 xy tiles x 2 z-tiles: two tiles constant (all-lit, all-shadowed cells in the grid),
    two generator tiles; uniform queries across every cell."""
    L, tile = 2, 128
    c = cp.CompressedShadowContainer(L)
    dags = [None] * L ** 3
    n = None
    for ty in range(L):
        for tx in range(L):
            k = ty * L + tx
            if k == 0:
                D = np.full((tile, tile), 1.0, np.float32)  # nothing in front: lit
            elif k == 3:
                D = np.full((tile, tile), 0.25, np.float32)  # lit above, shadowed below
            else:
                D = O.gen_depth(tile, SEED0 + 720 + k, 6, True)
            mm = cp.MinMaxHierarchy(D)
            for tz in range(L):
                cs = cp.CompressedShadow.create(mm, tz, L)
                w, n = cs.getDAG(), cs.getNumLevels()
                dags[(tz * L + ty) * L + tx] = w
                c.set(cs, tx, ty, tz)
    grid = O.top_level_grid(dags)
    assert (grid >= 0xFFFFFFFE).any()  # the container has sentinel cells
    words = np.concatenate(dags)
    q = _queries(77, 1 << 17)
    want = O.evaluate(words, grid, n, 1, q, nthreads=8)
    c.setFilterSize(3)
    _sweep(c, q, want, n, L ** 3, True, pcf=True)
    c.close()


def test_top_levels_arguments(ctx):
    D = O.gen_depth(64, SEED0 + 730, 2, True)
    c = cp.CompressedShadowContainer(cp.CompressedShadow.create(cp.MinMaxHierarchy(D)))
    for bad in (-2, MAX_K + 1, 100):
        with pytest.raises(cp.CpvsError):
            c.setTopLevels(bad)
    assert c.top_levels() == 0  # nothing built before copyToGPU
    c.close()
    big = cp.CompressedShadowContainer(8)  # 512 cells x 8^9 entries: beyond the 1 GiB limit
    with pytest.raises(cp.CpvsError):
        big.setTopLevels(MAX_K)
    big.setTopLevels(5)
    big.close()


def test_tuner_keeps_the_fastest_and_results(ctx):
    D = O.gen_depth(1024, SEED0 + 740, 32, True)
    cs = cp.CompressedShadow.create(cp.MinMaxHierarchy(D))
    w, n = cs.getDAG(), cs.getNumLevels()
    c = cp.CompressedShadowContainer(cs)
    q = _queries(5, 1 << 18)
    dq = cp.DeviceArray.from_numpy(q)
    best, ms = c.tuneTopLevels(dq, candidates=(0, 3, 6, -1), reps=2)
    _say("tuner", best, ms)
    assert best in (0, 3, 6, -1) and all(t > 0 for t in ms.values())
    assert ms[best] == min(ms.values())
    assert c.top_levels() == _expected_k(best, n, 1, True)
    assert np.array_equal(c.evaluate(dq).numpy(), O.evaluate(w, O.top_level_grid([w]), n, 0, q, nthreads=8))
    with pytest.raises(cp.CpvsError):
        c.tuneTopLevels(dq, candidates=(0, MAX_K + 1))  # beyond the maximum: rejected before any timing
    assert c.top_levels() == _expected_k(best, n, 1, True)
    c.freeOnCPU()
    with pytest.raises(cp.CpvsError):
        c.tuneTopLevels(dq)  # nothing left to rebuild the table from
    assert np.array_equal(c.evaluate(dq).numpy(), O.evaluate(w, O.top_level_grid([w]), n, 0, q, nthreads=8))
    c.close()


def test_dag_tables_in_dag_lookups_and_sets(ctx):
    """A DAG's own table (cpvs_dag_set_top_levels): one-DAG lookups and DAG sets, round-robin and
    by id, for every setting; a set made before a table is replaced keeps its answers."""
    specs = [(32, 1), (64, 2), (256, 6), (1024, 24)]
    shadows, words, levels = [], [], []
    for k, (N, boxes) in enumerate(specs):
        D = O.gen_depth(N, SEED0 + 750 + k, boxes, True)
        cs = cp.CompressedShadow.create(cp.MinMaxHierarchy(D))
        shadows.append(cs)
        words.append(cs.getDAG())
        levels.append(cs.getNumLevels())
    flat = cp.CompressedShadow.create(cp.MinMaxHierarchy(np.full((64, 64), 1.0, np.float32)))  # all lit
    shadows.append(flat)
    words.append(flat.getDAG())
    levels.append(flat.getNumLevels())
    m = len(shadows)
    q = _queries(91, 1 << 16, finite=True)[:, :3].copy()
    dq = cp.DeviceArray.from_numpy(q)
    want = [O.evaluate(w, O.top_level_grid([w]), n, 0, q, nthreads=8) for w, n in zip(words, levels)]
    ids = np.random.default_rng(4).integers(0, m, len(q)).astype(np.uint32)
    dids = cp.DeviceArray.from_numpy(ids)
    want_rr = np.array([want[i % m][i] for i in range(len(q))], np.uint8)
    want_id = np.array([want[ids[i]][i] for i in range(len(q))], np.uint8)
    before = None
    for req in [0, 2, 3, 5, 7, -1]:
        for cs, n in zip(shadows, levels):
            cs.setTopLevels(req)
            k = 0 if cs is flat else _expected_k(req, n, 1, True)
            _say("dag top table", req, "->", cs.top_levels(), "levels", n)
            assert cs.top_levels() == k, (req, n)
        for cs, w in zip(shadows, want):
            assert np.array_equal(cs.evaluate(dq).numpy(), w), req
        s = cp.DagSet(shadows)
        assert np.array_equal(s.evaluate(dq).numpy(), want_rr), req
        assert np.array_equal(s.evaluate(dq, ids=dids).numpy(), want_id), req
        if before is None:
            before = s  # made with no tables; every later setting replaces them
        else:
            s.close()
        assert np.array_equal(before.evaluate(dq).numpy(), want_rr), req
    before.close()
    for cs in shadows:
        cs.close()


def test_dag_table_without_leafmasks():
    nctx = cp.Context(0)
    try:
        nctx.set_leafmasks(False)
        D = O.gen_depth(512, SEED0 + 760, 16, True)
        cs = cp.CompressedShadow.create(cp.MinMaxHierarchy(D, ctx=nctx))
        w, n = cs.getDAG(), cs.getNumLevels()
        q = _queries(92, 1 << 16, finite=True)[:, :3].copy()
        want = O.evaluate(w, O.top_level_grid([w]), n, 0, q, nthreads=8, leafmasks=False)
        dq = cp.DeviceArray.from_numpy(q, ctx=nctx)
        for req in [0, 4, 7, -1]:
            cs.setTopLevels(req)
            assert cs.top_levels() == _expected_k(req, n, 1, False)
            assert np.array_equal(cs.evaluate(dq).numpy(), want), req
        cs.close()
    finally:
        nctx.close()
