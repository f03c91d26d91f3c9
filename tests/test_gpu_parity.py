"""Parity of the HIP path (through the C ABI) with the CPU oracle: bit-exact pyramids, DAG
words, visibilities and generator output.  Needs a GPU."""
import ctypes
import hashlib
import json
import os

import numpy as np
import pytest

import cpvs_amd as cp
import oracle as O
from tests.dagcheck import check_dag

pytestmark = pytest.mark.gpu
SEED0 = 0x43505653
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _say(*a):
    print("[case]", *a, flush=True)


def _gpu_dag(D, t=0, T=1):
    mm = cp.MinMaxHierarchy(D)
    cs = cp.CompressedShadow.create(mm, t, T)
    return cs, cs.getDAG(), cs.getNumLevels()


def _oracle_dag(D, t=0, T=1):
    return O.create(O.Pyramid(D), t, T, merge_mode=1)


# --- pyramid -------------------------------------------------------------------------------

@pytest.mark.parametrize("N", [8, 16, 32, 64, 128, 256, 1024, 2048])
def test_pyramid_bit_exact(ctx, N):
    _say("pyramid", N)
    rng = np.random.default_rng(N)
    D = rng.random((N, N), dtype=np.float32)
    D[rng.random((N, N)) < 0.1] = 0.25
    mm = cp.MinMaxHierarchy(D, all_levels=True)
    ref = O.Pyramid(D)
    assert mm.getNumLevels() == ref.getNumLevels()
    assert np.array_equal(mm.getLevel(0), D)
    for L in range(1, mm.getNumLevels()):
        assert np.array_equal(mm.getLevel(L).view(np.uint32), ref.level(L).view(np.uint32)), L


def test_pyramid_reference_known_answers(ctx, ref_fixtures):  # MinMaxTest.cpp:41-90
    for e in ref_fixtures["minmax"]:
        mm = cp.MinMaxHierarchy(ref_fixtures["arrays"][e[0]], all_levels=True)
        if e[1] == "numLevels":
            assert mm.getNumLevels() == e[2]
            continue
        _, L, x, y, which, exp, tol = e
        v = mm.getMin(L, x, y) if which == "min" else mm.getMax(L, x, y)
        assert (v == np.float32(exp)) if tol == 0 else abs(v - exp) < tol, e


def test_borrowed_pyramid_waits_for_its_levels(ctx):
    # A pyramid over a borrowed device depth map is built on the context's stream without a host
    # wait.  Level reads and builds on another stream must still see the finished levels.
    _say("borrowed pyramid, other stream")
    N = 1024
    D = O.gen_depth(N, SEED0 + 7, 32, True)
    ref = O.Pyramid(D)
    wr, nr = O.create(ref, 0, 1, merge_mode=1)
    d = cp.DeviceArray.from_numpy(D)
    mm = cp.MinMaxHierarchy(d, borrow=True, all_levels=True)
    for L in range(1, mm.getNumLevels()):
        assert np.array_equal(mm.getLevel(L).view(np.uint32), ref.level(L).view(np.uint32)), L
    mm.close()
    hip = ctypes.CDLL("libamdhip64.so")
    st = ctypes.c_void_p()
    assert hip.hipStreamCreate(ctypes.byref(st)) == 0
    try:
        for _ in range(3):
            mm = cp.MinMaxHierarchy(d, borrow=True)
            cs = cp.CompressedShadow.create(mm, stream=st)
            assert cs.getNumLevels() == nr and np.array_equal(cs.getDAG(), wr)
            cs.close()
            mm.close()
    finally:
        hip.hipStreamDestroy(st)


def test_pyramid_levels_1_2_need_flag(ctx):
    mm = cp.MinMaxHierarchy(np.full((64, 64), 0.5, np.float32))
    with pytest.raises(cp.CpvsError):
        mm.getLevel(1)
    assert mm.getLevel(3).shape == (8, 8, 2)


# --- node counts (admission estimate) --------------------------------------------------------

@pytest.mark.parametrize("N,T", [(64, 1), (256, 2), (1024, 1), (1024, 4)])
def test_exact_node_counts(ctx, N, T):
    _say("counts", N, T)
    D = O.gen_depth(N, SEED0 + N, max(1, 32 * N * N // (1 << 20)), True)
    mm = cp.MinMaxHierarchy(D)
    ref = O.Pyramid(D)
    n = ref.getNumLevels()
    for t in range(T):
        nodes, scratch, maxdag = mm.estimate(t, T)
        w, _ = O.create(ref, t, T)
        exp = O.last_node_counts(n)
        exp[n - 2] = 1
        assert np.array_equal(nodes[:n - 1], exp[:n - 1]), (t, nodes, exp)
        assert maxdag >= 4 * len(w) and scratch > 0


# --- DAG words --------------------------------------------------------------------------------

def test_reference_images_bit_exact(ctx, ref_fixtures):
    for name in ("depths8x8", "depths16x16", "depths32x32", "minmax8x8"):
        _say("ref image", name)
        D = ref_fixtures["arrays"][name]
        _, w, n = _gpu_dag(D)
        wr, nr = _oracle_dag(D)
        assert n == nr and np.array_equal(w, wr), name


def test_reference_traverse_known_answers(ctx, ref_fixtures):  # CompressedShadowTest.cpp:50-151
    shadows = {}
    for name, ndc, try_leaf, exp in ref_fixtures["traverse"]:
        if name not in shadows:
            shadows[name] = _gpu_dag(ref_fixtures["arrays"][name])[0]
        assert shadows[name].traverse(ndc, try_leaf) == exp, (name, ndc)
    for sw in ref_fixtures["sweeps"]:
        cs = shadows[sw["image"]]
        pts = []
        for y in range(*sw["y"]):
            for x in range(*sw["x"]):
                v = np.array([x / 32.0, y / 32.0, sw["z"]], np.float32)
                pts.append(v * np.float32(2.0) + np.float32(-1.0))
        got = cs.traverse_many(np.array(pts, np.float32), True)
        assert np.all(got == sw["expected"])


def test_traverse_mode_must_match_dag(ctx, ref_fixtures):
    cs8 = _gpu_dag(ref_fixtures["arrays"]["depths8x8"])[0]
    cs16 = _gpu_dag(ref_fixtures["arrays"]["depths16x16"])[0]
    with pytest.raises(cp.CpvsError):
        cs8.traverse([0, 0, 0], True)
    with pytest.raises(cp.CpvsError):
        cs16.traverse([0, 0, 0], False)


@pytest.mark.parametrize("N", [16, 32, 64, 128, 256, 512, 1024])
@pytest.mark.parametrize("T", [1, 2, 4, 8])
def test_generator_maps_bit_exact(ctx, N, T):
    D = O.gen_depth(N, SEED0 + 7 * N + T, max(1, 32 * N * N // (1 << 20)), N >= 128)
    mm = cp.MinMaxHierarchy(D)
    pyr = O.Pyramid(D)
    for t in range(T):
        _say("generator map", N, T, t)
        cs = cp.CompressedShadow.create(mm, t, T)
        w = cs.getDAG()
        wr, nr = O.create(pyr, t, T, merge_mode=1)
        assert cs.getNumLevels() == nr
        assert np.array_equal(w, wr), (N, T, t, len(w), len(wr))
        assert cs.getTotalVisibility() == O.total_visibility(wr)


@pytest.fixture(scope="module")
def sort_ctx():
    c = cp.Context(0)
    c.set_merge("sort")
    yield c
    c.close()


@pytest.mark.parametrize("N,T", [(16, 1), (64, 2), (256, 1), (1024, 1), (1024, 4)])
def test_sort_merge_path_bit_exact(sort_ctx, N, T):
    """The radix-sort merge gives the same words as the hash-table merge and the oracle."""
    D = O.gen_depth(N, SEED0 + 3 * N + T, max(1, 32 * N * N // (1 << 20)), N >= 128)
    pyr = O.Pyramid(D)
    mm_sort = cp.MinMaxHierarchy(D, ctx=sort_ctx)
    mm_table = cp.MinMaxHierarchy(D)
    for t in range(T):
        _say("sort merge", N, T, t)
        a = cp.CompressedShadow.create(mm_sort, t, T)
        b = cp.CompressedShadow.create(mm_table, t, T)
        assert a.merge_info()[0] == "sort" and b.merge_info()[0] == "table"
        wr, _ = O.create(pyr, t, T, merge_mode=1)
        assert np.array_equal(a.getDAG(), wr)
        assert np.array_equal(b.getDAG(), wr)


@pytest.mark.parametrize("N", [16, 64, 256])
def test_random_maps_bit_exact(ctx, N):
    _say("random map", N)
    rng = np.random.default_rng(100 + N)
    for D in (rng.random((N, N), dtype=np.float32),
              np.round(rng.random((N, N), dtype=np.float32) * 8) / 8,  # many exact ties
              (rng.random((N, N), dtype=np.float32) * 0.01 + 0.5).astype(np.float32)):
        _, w, n = _gpu_dag(D)
        wr, nr = _oracle_dag(D)
        assert np.array_equal(w, wr)
        check_dag(w, n)


@pytest.mark.parametrize("value", [0.0, 1.0, 0.5, 0.625, 0.515625])
def test_uniform_and_early_ending_maps(ctx, value):
    # 0.625 * 8 is an integer at level 3 while 0.625 * 4 is not at level 4: the octree has
    # level-3 nodes but no leaves (undefined in the reference, a shorter DAG here).
    _say("uniform", value)
    D = np.full((64, 64), value, np.float32)
    cs, w, n = _gpu_dag(D)
    wr, nr = _oracle_dag(D)
    assert np.array_equal(w, wr)
    if value == 1.0:
        assert cs.getTotalVisibility() == cp.CompressedShadow.VISIBLE and len(w) == 1
    if value == 0.0:
        assert cs.getTotalVisibility() == cp.CompressedShadow.SHADOW and len(w) == 1


def _special_map(rng, N):
    D = rng.random((N, N), dtype=np.float32)
    for frac, val in ((0.05, np.nan), (0.03, np.inf), (0.03, -np.inf), (0.05, -0.5), (0.05, 3.0),
                      (0.02, 1e30), (0.02, -1e30), (0.03, 1.0), (0.03, 0.0)):
        D[rng.random((N, N)) < frac] = val
    return D


@pytest.mark.parametrize("N,T", [(64, 1), (64, 2), (256, 1), (256, 3)])
def test_special_depth_values_bit_exact(ctx, N, T):
    # NaN, infinities, negative, above-one and huge depths.  The pyramid's min/max keep a NaN only
    # in the first operand, a NaN bound fails every classification compare (the cell is partial),
    # and a NaN texel is never lit.
    _say("special values", N, T)
    rng = np.random.default_rng(7 + N + T)
    nan_first = rng.random((N, N), dtype=np.float32)
    nan_first[::2, ::2] = np.nan  # NaN where it reaches every pyramid level
    maps = [_special_map(rng, N), nan_first, np.full((N, N), np.nan, np.float32),
            np.full((N, N), np.inf, np.float32), np.full((N, N), -1e30, np.float32)]
    for mi, D in enumerate(maps):
        pyr = O.Pyramid(D)
        mm = cp.MinMaxHierarchy(D, all_levels=True)
        for L in range(1, mm.getNumLevels()):
            assert np.array_equal(mm.getLevel(L).view(np.uint32), pyr.level(L).view(np.uint32)), (mi, L)
        for t in range(T):
            cs = cp.CompressedShadow.create(mm, t, T)
            wr, nr = O.create(pyr, t, T, merge_mode=1)
            assert cs.getNumLevels() == nr
            assert np.array_equal(cs.getDAG(), wr), (mi, t)
            check_dag(wr, nr)


def test_tall_wall_columns_bit_exact(ctx):
    # Columns whose 8x8 texel block holds both 0 and 1 are partial over the whole z range: 128
    # leaf nodes per column at N = 1024, more than one 64-node chunk of the leaf kernel.
    _say("tall walls")
    N = 1024
    rng = np.random.default_rng(11)
    D = rng.random((N, N), dtype=np.float32)
    walls = rng.random((N // 8, N // 8)) < 0.2
    by, bx = np.nonzero(walls)
    D[8 * by, 8 * bx] = 0.0
    D[8 * by + 7, 8 * bx + 7] = 1.0
    pyr = O.Pyramid(D)
    for T in (1, 2):
        for t in range(T):
            _, w, n = _gpu_dag(D, t, T)
            wr, nr = O.create(pyr, t, T, merge_mode=1)
            assert np.array_equal(w, wr), (T, t)


def test_lookups_with_special_coordinates(ctx):
    # traverse.cs getPathFromNDC: NaN and negative positions go to cell 0, positions at or past
    # the far edge (infinities included) to the last cell.
    _say("special lookups")
    N = 256
    D = O.gen_depth(N, SEED0 + 2, 8, True)
    cs = cp.CompressedShadow.create(cp.MinMaxHierarchy(D))
    n, words = cs.getNumLevels(), cs.getDAG()
    c = cp.CompressedShadowContainer(cs)
    c.moveToGPU()
    rng = np.random.default_rng(3)
    q = (rng.random((65536, 4), dtype=np.float32) * 2.4 - 1.2).astype(np.float32)
    special = np.array([np.nan, np.inf, -np.inf, -1.0, 1.0, 1e30, -1e30, 0.0], np.float32)
    pick = rng.random(q.shape) < 0.2
    q[pick] = special[rng.integers(0, special.size, q.shape)][pick]
    vis = c.evaluate(cp.DeviceArray.from_numpy(q)).numpy()
    exp = O.evaluate(words, O.top_level_grid([words]), n, 0, q, nthreads=8)
    assert np.array_equal(vis, exp)


def test_hash_collision_path_exercised_by_many_duplicates(ctx):
    # A periodic map yields huge groups of identical subtrees at every level.
    N = 512
    y, x = np.mgrid[0:N, 0:N]
    D = (0.3 + 0.4 * (((x // 8 + y // 8) % 2) == 0) + 0.001 * (x % 8)).astype(np.float32)
    _say("periodic")
    cs, w, n = _gpu_dag(D)
    wr, _ = _oracle_dag(D)
    assert np.array_equal(w, wr)
    nodes, uniq = cs.level_stats()
    assert nodes[2] > 100 * uniq[2]


def test_leaf_table_overflow_builds_again(ctx):
    # White noise: almost every leaf node is distinct, more keys than the leaf level's merge table
    # (sized for a quarter of its nodes from 2^20 leaf nodes on) has slots, so inserts give up and
    # the build runs again with a table for every node.
    N = 1024
    D = np.random.default_rng(7).random((N, N), dtype=np.float32)
    _say("noise", N)
    cs, w, n = _gpu_dag(D)
    wr, _ = _oracle_dag(D)
    assert np.array_equal(w, wr)
    nodes, uniq = cs.level_stats()
    assert nodes[2] >= 1 << 20
    slots = 1024
    while slots < nodes[2] // 4 + nodes[2] // 8 + 1:
        slots <<= 1
    assert uniq[2] > slots


def test_invalid_arguments_are_rejected(ctx):
    with pytest.raises(cp.CpvsError):
        cp.MinMaxHierarchy(np.zeros((48, 48), np.float32))  # not a power of two
    mm = cp.MinMaxHierarchy(np.full((32, 32), 0.5, np.float32))
    with pytest.raises(cp.CpvsError):
        cp.CompressedShadow.create(mm, 2, 2)  # zTileIndex >= zTileNum
    with pytest.raises(cp.CpvsError):
        cp.CompressedShadow.create(mm, 0, 0)
    small = cp.MinMaxHierarchy(np.full((4, 4), 0.5, np.float32))  # a pyramid, but no octree (N < 8)
    assert small.getNumLevels() == 3
    with pytest.raises(cp.CpvsError):
        cp.CompressedShadow.create(small)


# --- lookups ------------------------------------------------------------------------------------

def test_evaluate_uniform_queries_c0(ctx):
    _say("evaluate C0")
    N = 1024
    D = O.gen_depth(N, SEED0, 32, False)
    mm = cp.MinMaxHierarchy(D)
    cs = cp.CompressedShadow.create(mm)
    n = cs.getNumLevels()
    words = cs.getDAG()
    c = cp.CompressedShadowContainer(cs)
    c.moveToGPU()
    q = O.gen_queries_uniform(SEED0, 1 << 20, 4)
    dq = cp.DeviceArray.from_numpy(q)
    vis = c.evaluate(dq).numpy()
    grid = O.top_level_grid([words])
    exp = O.evaluate(words, grid, n, 0, q, nthreads=8)
    assert np.array_equal(vis, exp)
    # brute force on every query
    paths = np.array([O.path_from_ndc(v[:3], n, 0) for v in q[:20000]])
    bf = O.brute_visible(D, 1, paths[:, 0], paths[:, 1], paths[:, 2])
    assert np.array_equal(vis[:20000], bf)
    # float3 layout gives the same answers
    dq3 = cp.DeviceArray.from_numpy(np.ascontiguousarray(q[:, :3]))
    assert np.array_equal(c.evaluate(dq3).numpy(), vis)


def test_evaluate_with_light_matrix(ctx):
    _say("evaluate matrix")
    N = 256
    D = O.gen_depth(N, SEED0 + 1, 8, True)
    mm = cp.MinMaxHierarchy(D)
    cs = cp.CompressedShadow.create(mm)
    n = cs.getNumLevels()
    words = cs.getDAG()
    c = cp.CompressedShadowContainer(cs)
    c.copyToGPU()
    rng = np.random.default_rng(1)
    M = np.array([[0.5, 0.0, 0.0, -0.25], [0.0, 0.25, 0.0, 0.1], [0.0, 0.0, 0.125, 0.3], [0.0, 0.0, 0.0, 1.0]],
                 np.float32)
    pos = (rng.random((100000, 4), dtype=np.float32) * 4 - 2).astype(np.float32)
    pos[:, 3] = 1
    vis = c.evaluate(cp.DeviceArray.from_numpy(pos), M).numpy()
    exp = O.evaluate(words, O.top_level_grid([words]), n, 0, pos, matrix=M.T.copy(), nthreads=8)
    assert np.array_equal(vis, exp)


def test_tiled_container_bit_exact(ctx):
    """L=2: 2x2 xy tiles x 2 z-tiles of a 512^2 map (CompressedShadowContainer.cpp:52-91)."""
    N, L = 256, 2
    full = O.gen_depth(N * L, SEED0 + 5, 64, True)
    c = cp.CompressedShadowContainer(L)
    oracle_dags = [None] * (L ** 3)
    n = None
    for z in range(L):
        for y in range(L):
            for x in range(L):
                _say("tile", x, y, z)
                tile = np.ascontiguousarray(full[y * N:(y + 1) * N, x * N:(x + 1) * N])
                mm = cp.MinMaxHierarchy(tile)
                cs = cp.CompressedShadow.create(mm, z, L)
                w, n = O.create(O.Pyramid(tile), z, L)
                assert np.array_equal(cs.getDAG(), w)
                oracle_dags[(z * L + y) * L + x] = w
                c.set(cs, x, y, z)
    c.moveToGPU()
    dl, gl, nw = c.info()
    assert (dl, gl) == (n, 1)
    words, grid = c.words_and_grid()
    assert np.array_equal(words, np.concatenate(oracle_dags))
    assert np.array_equal(grid, O.top_level_grid(oracle_dags))
    q = O.gen_queries_uniform(9, 1 << 18, 4)
    vis = c.evaluate(cp.DeviceArray.from_numpy(q)).numpy()
    assert np.array_equal(vis, O.evaluate(words, grid, n, 1, q, nthreads=8))
    paths = np.array([O.path_from_ndc(v[:3], n, 1) for v in q[:20000]])
    assert np.array_equal(vis[:20000], O.brute_visible(full, 1, paths[:, 0], paths[:, 1], paths[:, 2]))


def test_evaluate_image_matches_flat(ctx):
    N = 512
    D = O.gen_depth(N, SEED0 + 2, 16, True)
    cs = cp.CompressedShadow.create(cp.MinMaxHierarchy(D))
    words, n = cs.getDAG(), cs.getNumLevels()
    c = cp.CompressedShadowContainer(cs)
    c.moveToGPU()
    W, H = 333, 77  # ragged: neither a multiple of the 32-pixel blocks nor of the 8 rows
    img = O.gen_queries_uniform(3, W * H, 4).reshape(H, W, 4)
    d_img = cp.DeviceArray.from_numpy(img)
    a = c.evaluate_image(d_img).numpy().reshape(-1)
    flat = cp.DeviceArray.from_numpy(img.reshape(-1, 4))
    b = c.evaluate(flat).numpy()
    assert np.array_equal(a, b)
    assert np.array_equal(b, O.evaluate(words, O.top_level_grid([words]), n, 0, img.reshape(-1, 4), nthreads=8))


# --- generator ------------------------------------------------------------------------------------

def test_device_generator_matches_fixtures(ctx):
    fx = json.load(open(os.path.join(GOLDEN, "generator.json")))
    for e in fx["depth"]:
        d = cp.gen_depth(e["N"], e["seed"], e["boxes"], e["noise"], e["x0"], e["y0"], e["tile"]).numpy()
        assert hashlib.sha256(d.tobytes()).hexdigest() == e["sha256"], e
    for e in fx["uniform"]:
        q = cp.gen_queries_uniform(e["seed"], e["count"], e["stride"]).numpy()
        assert hashlib.sha256(q.tobytes()).hexdigest() == e["sha256"], e


def test_device_gbuffer_queries_match_oracle(ctx):
    N = 1024
    D = O.gen_depth(N, SEED0, 32, False)
    dD = cp.DeviceArray.from_numpy(D)
    q = cp.gen_queries_gbuffer(dD, 11, 320, 180).numpy()
    assert np.array_equal(q, O.gen_queries_gbuffer(D, 11, 320, 180))


# --- full size (8192^2, the headline workload) ------------------------------------------------

def test_full_size_c1_properties(ctx):
    """8192^2 + 512 boxes + noise: words identical to the oracle, layout invariants, and every one
    of 1M G-buffer queries equal to the brute-force depth compare."""
    N = 8192
    _say("C1 generate")
    d = cp.gen_depth(N, SEED0 + 1, 512, True)
    mm = cp.MinMaxHierarchy(d, borrow=True)
    _say("C1 build")
    cs = cp.CompressedShadow.create(mm)
    n = cs.getNumLevels()
    w = cs.getDAG()
    assert cs.merge_info()[0] == "table"  # no 64-bit key collision forced the sort path
    _say("C1 check layout", len(w))
    check_dag(w, n)
    D = d.numpy()
    _say("C1 oracle")
    wr, _ = O.create(O.Pyramid(D), 0, 1, merge_mode=1)
    assert np.array_equal(w, wr)
    c = cp.CompressedShadowContainer(cs)
    c.moveToGPU()
    q = cp.gen_queries_gbuffer(d, 17, 1024, 512)
    vis = c.evaluate(q).numpy()
    qh = q.numpy()
    R = np.float32((1 << (n - 1)) - 1)
    p = (((qh[:, :3] + np.float32(1)) * np.float32(0.5)) * R).astype(np.int64)
    bf = O.brute_visible(D, 1, p[:, 0], p[:, 1], p[:, 2])
    assert np.array_equal(vis, bf)
    assert 0.2 < vis.mean() < 0.8


# --- tiled 65536^2 (8x8 xy tiles of 8192^2 x 8 z-tiles = 512 DAGs) ----------------------------

def test_tiled_65536_container_properties(ctx):
    """The 65536^2 configuration (DeferredRenderer.cpp:150-187): per-tile device generation, 512
    z-tile DAGs in an 8^3 grid, and lookups equal to the brute-force depth compare on 2 tiles."""
    full, tile, L = 65536, 8192, 8
    seed, boxes = SEED0 + 2, 16384
    c = cp.CompressedShadowContainer(L)
    kept = {}
    for ty in range(L):
        for tx in range(L):
            d = cp.gen_depth(full, seed, boxes, True, tx * tile, ty * tile, tile)
            mm = cp.MinMaxHierarchy(d, borrow=True)
            for tz in range(L):
                c.set(cp.CompressedShadow.create(mm, tz, L), tx, ty, tz)
            if (tx, ty) in ((0, 0), (5, 3)):
                kept[(tx, ty)] = d.numpy()
            mm.close()
            d.free()
        _say("tiled row", ty)
    c.moveToGPU()
    dl, gl, nw = c.info()
    assert (dl, gl) == (14, 3)
    words, grid = c.words_and_grid()
    assert (grid == 0xFFFFFFFF).any()  # fully shadowed z-tiles use the sentinel the lookup tests
    rng = np.random.default_rng(5)
    R = np.float32(full - 1)
    for (tx, ty), D in kept.items():
        n = 200000
        px = rng.integers(tx * tile, (tx + 1) * tile, n)
        py = rng.integers(ty * tile, (ty + 1) * tile, n)
        # z near the surface so that both outcomes occur
        z = np.clip(D[py - ty * tile, px - tx * tile] + rng.uniform(-0.002, 0.002, n).astype(np.float32), 0, 1)
        q = np.stack([(px + 0.5) / full, (py + 0.5) / full, z, np.ones(n)], 1).astype(np.float32)
        q[:, :3] = q[:, :3] * np.float32(2) - np.float32(1)
        vis = c.evaluate(cp.DeviceArray.from_numpy(q)).numpy()
        p = (((q[:, :3] + np.float32(1)) * np.float32(0.5)) * R).astype(np.int64)
        inside = (p[:, 0] // tile == tx) & (p[:, 1] // tile == ty)  # quantisation can step off the tile edge
        assert inside.mean() > 0.99
        p, v = p[inside], vis[inside]
        # z resolution of a tile DAG is tile * L = full; brute force at leaf resolution
        mid = (p[:, 2].astype(np.float32) + (p[:, 2] + 1).astype(np.float32)) * np.float32(0.5)
        bf = (mid <= D[p[:, 1] - ty * tile, p[:, 0] - tx * tile] * np.float32(full)).astype(np.uint8)
        assert np.array_equal(v, bf), (tx, ty)
        assert 0.05 < v.mean() < 0.95


def test_context_close_closes_its_objects():
    """Closing a context closes the pyramids, DAGs, containers, DAG sets, servers and device arrays
    still open on it, so none of them later reaches into the destroyed context."""
    c2 = cp.Context(0)
    D = O.gen_depth(64, SEED0 + 8, 4, True)
    d = cp.DeviceArray.from_numpy(D, ctx=c2)
    mm = cp.MinMaxHierarchy(d, ctx=c2)
    a = cp.CompressedShadow.create(mm)
    b = cp.CompressedShadow.create(mm)
    s = cp.DagSet([a, b])
    cont = cp.CompressedShadowContainer(cp.CompressedShadow.create(mm))
    srv = cp.Server(lanes=1, ctx=c2)
    c2.close()
    assert not any(getattr(o, "handle", None) for o in (mm, a, b, s, cont, srv)) and not d.ptr
    for o in (mm, a, b, s, cont, srv, d):
        o.close()  # a second close is a no-op
    del mm, a, b, s, cont, srv, d
    import gc
    gc.collect()
