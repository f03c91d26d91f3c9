"""Batched builds (cpvs_build_batch), the lane-count tuner and DAG sets (many DAGs in one lookup
launch), checked against plain single builds and the CPU oracle.  Needs a GPU."""
import numpy as np
import pytest

import cpvs_amd as cp
import oracle as O

pytestmark = pytest.mark.gpu
SEED0 = 0x43505653


def _say(*a):
    print("[case]", *a, flush=True)


def _maps():
    """(name, depth, zTileIndex, zTileNum): generator maps of several sizes and z-tiles, a map
    that ends early (uniform 0.625), all-shadow and all-lit maps."""
    out = []
    for k, (N, boxes, noise, t, T) in enumerate([(8, 1, False, 0, 1), (16, 2, True, 0, 1), (64, 4, True, 1, 2),
                                                 (256, 8, True, 0, 1), (1024, 32, False, 0, 1), (512, 16, True, 3, 4),
                                                 (128, 4, False, 0, 1), (2048, 64, True, 0, 1), (32, 2, True, 0, 2)]):
        out.append((f"gen{N}_t{t}of{T}", O.gen_depth(N, SEED0 + 1000 * (k + 1), boxes, noise), t, T))
    out.append(("uniform0.625", np.full((64, 64), 0.625, np.float32), 0, 1))
    out.append(("shadow", np.zeros((64, 64), np.float32), 0, 1))
    out.append(("lit", np.ones((64, 64), np.float32), 0, 1))
    return out


def _plain(D, t, T):
    cs = cp.CompressedShadow.create(cp.MinMaxHierarchy(D), t, T)
    return cs.getNumLevels(), cs.getDAG(), cs.level_stats()


@pytest.fixture(scope="module")
def mix(ctx):
    maps = _maps()
    depths = [cp.DeviceArray.from_numpy(D) for _, D, _, _ in maps]
    tiles = [(t, T) for _, _, t, T in maps]
    plain = [_plain(D, t, T) for _, D, t, T in maps]
    return maps, depths, tiles, plain


@pytest.mark.parametrize("lanes", [1, 3, 4, 8])
def test_batch_matches_plain_builds(ctx, mix, lanes):
    maps, depths, tiles, plain = mix
    _say("batch lanes", lanes, "maps", len(maps))
    shadows, done = cp.build_batch(depths, tiles, lanes=lanes)
    assert len(shadows) == len(maps)
    for (name, D, t, T), sh, (n, words, stats) in zip(maps, shadows, plain):
        assert sh.getNumLevels() == n, name
        assert np.array_equal(sh.getDAG(), words), name
        got = sh.level_stats()
        assert np.array_equal(got[0], stats[0]) and np.array_equal(got[1], stats[1]), name
    assert np.all(np.isfinite(done)) and np.all(done > 0.0)


def test_batch_matches_oracle_words(ctx, mix):
    maps, depths, tiles, _ = mix
    _say("batch vs oracle")
    shadows, _ = cp.build_batch(depths, tiles, lanes=4)
    for (name, D, t, T), sh in zip(maps, shadows):
        if D.shape[0] > 512:
            continue  # the oracle is slow there; plain-vs-batch equality covers it
        words, n = O.create(O.Pyramid(D), t, T, merge_mode=1)
        assert sh.getNumLevels() == n, name
        assert np.array_equal(sh.getDAG(), words), name


def test_batch_with_sort_merge(mix):
    maps, depths, tiles, plain = mix
    _say("batch sort merge")
    c2 = cp.Context(0)
    try:
        c2.set_merge("sort")
        shadows, _ = cp.build_batch(depths, tiles, lanes=3, ctx=c2)
        try:
            for (name, *_), sh, (n, words, _) in zip(maps, shadows, plain):
                assert np.array_equal(sh.getDAG(), words), name
                assert sh.merge_info()[0] == "sort"
        finally:
            for sh in shadows:  # the DAGs refer to c2: destroy them first
                sh.close()
    finally:
        c2.close()


def test_batch_z_tiles_share_pyramids(ctx):
    """Every z-tile of three maps in one batch, listed in shuffled order (the tiled configuration's
    shape): a lane builds one pyramid per map and all of its z-tiles from it.  Words equal plain
    builds for 1 and 3 lanes."""
    _say("batch z-tiles of shared maps")
    specs = [(512, 8, 8), (256, 4, 4), (1024, 16, 2)]  # N, boxes, zTileNum
    Ds = [O.gen_depth(N, SEED0 + 77 * k, b, True) for k, (N, b, _) in enumerate(specs)]
    depths = [cp.DeviceArray.from_numpy(D) for D in Ds]
    items = [(k, t) for k, (_, _, T) in enumerate(specs) for t in range(T)]
    np.random.default_rng(11).shuffle(items)
    plain = {(k, t): _plain(Ds[k], t, specs[k][2])[1] for k, t in items}
    for lanes in (1, 3):
        shadows, _ = cp.build_batch([depths[k] for k, _ in items], [(t, specs[k][2]) for k, t in items], lanes=lanes)
        for (k, t), sh in zip(items, shadows):
            assert np.array_equal(sh.getDAG(), plain[(k, t)]), (k, t, lanes)


def test_batch_many_small_maps(ctx):
    _say("batch 96 small maps")
    rng = np.random.default_rng(7)
    Ds = [O.gen_depth(int(N), SEED0 + 31 * i, 2, True) for i, N in enumerate(rng.choice([8, 16, 32, 64], 96))]
    depths = [cp.DeviceArray.from_numpy(D) for D in Ds]
    shadows, done = cp.build_batch(depths, lanes=8)
    for D, sh in zip(Ds, shadows):
        n, words, _ = _plain(D, 0, 1)
        assert np.array_equal(sh.getDAG(), words)
    assert len(done) == 96


def test_batch_rejects_bad_items(ctx):
    _say("batch errors")
    good = cp.DeviceArray.from_numpy(np.full((16, 16), 0.5, np.float32))
    bad = cp.DeviceArray.from_numpy(np.full((12, 12), 0.5, np.float32))
    with pytest.raises(cp.CpvsError) as e:
        cp.build_batch([good, bad])
    assert e.value.code == 1 and "item 1" in str(e.value)
    with pytest.raises(cp.CpvsError):
        cp.build_batch([good], tiles=[(2, 2)])
    with pytest.raises(cp.CpvsError):
        cp.build_batch([good], lanes=17)
    # empty batch is fine
    sh, done = cp.build_batch([])
    assert sh == [] and len(done) == 0


def test_tune_batch_and_cache(ctx, mix, tmp_path):
    _, depths, tiles, plain = mix
    _say("tune batch")
    cache = str(tmp_path / "tune.json")
    try:
        best, ms = cp.tune_batch(depths, tiles, candidates=(1, 2, 4), cache=cache)
        assert best in (1, 2, 4) and set(ms) == {1, 2, 4} and all(v > 0 for v in ms.values())
        assert ctx.batch_lanes() == best
        best2, ms2 = cp.tune_batch(depths, tiles, candidates=(1, 2, 4), cache=cache)
        assert best2 == best and ms2 is None
        shadows, _ = cp.build_batch(depths, tiles)  # lanes = the tuned count
        for sh, (n, words, _) in zip(shadows, plain):
            assert np.array_equal(sh.getDAG(), words)
    finally:
        ctx.set_batch_lanes(0)


# --- DAG sets -------------------------------------------------------------------------------

def _paths(q, R):
    """traverse.cs quantisation with float32 rounding at every step, clamped to [0, R-1]."""
    v = (q[:, :3].astype(np.float32) + np.float32(1.0)) * np.float32(0.5)
    v = v * np.float32(R - 1)
    p = np.where(v > 0, v, 0).astype(np.float64)
    p = np.minimum(np.floor(p), R - 1)
    return p.astype(np.int64)


@pytest.fixture(scope="module")
def dset(ctx):
    specs = [(32, 2, True), (64, 4, False), (256, 8, True), (1024, 32, True)]
    Ds = [O.gen_depth(N, SEED0 + 77 * (i + 1), b, nz) for i, (N, b, nz) in enumerate(specs)]
    Ds += [np.zeros((64, 64), np.float32), np.ones((64, 64), np.float32)]
    shadows = [cp.CompressedShadow.create(cp.MinMaxHierarchy(D)) for D in Ds]
    return Ds, shadows


def test_paths_helper_matches_oracle():
    q = O.gen_queries_uniform(SEED0 + 5, 2000, 3)
    q[:7] = [[-1, -1, -1], [1, 1, 1], [-2, 3, 0.5], [0.999999, -0.999999, 0], [0, 0, 0], [-0.0, 0.5, -0.5],
             [1.5, -1.5, 0.25]]
    for n in (6, 7, 11):
        p = _paths(q, 1 << (n - 1))
        exp = np.array([O.path_from_ndc(v, n, 0) for v in q])
        assert np.array_equal(p, exp)


def test_dagset_matches_oracle_and_brute_force(ctx, dset):
    Ds, shadows = dset
    _say("dagset", len(shadows))
    k = len(shadows)
    nq = 300_000
    q = O.gen_queries_uniform(SEED0 + 9, nq, 3)
    rng = np.random.default_rng(3)
    ids = rng.integers(0, k, nq, dtype=np.uint32)
    s = cp.DagSet(shadows)
    dq = cp.DeviceArray.from_numpy(q)
    vis = s.evaluate(dq, cp.DeviceArray.from_numpy(ids)).numpy()
    rr = s.evaluate(dq).numpy()
    for j, (D, sh) in enumerate(zip(Ds, shadows)):
        words, n = sh.getDAG(), sh.getNumLevels()
        grid = O.top_level_grid([words])
        sel = ids == j
        exp = O.evaluate(words, grid, n, 0, q[sel], nthreads=8)
        assert np.array_equal(vis[sel], exp), j
        p = _paths(q[sel], D.shape[0])
        assert np.array_equal(vis[sel], O.brute_visible(D, 1, p[:, 0], p[:, 1], p[:, 2])), j
        selr = (np.arange(nq) % k) == j
        expr = O.evaluate(words, grid, n, 0, q[selr], nthreads=8)
        assert np.array_equal(rr[selr], expr), j
        # one DAG alone
        one = sh.evaluate(dq).numpy()
        assert np.array_equal(one, O.evaluate(words, grid, n, 0, q, nthreads=8)), j
    assert set(np.unique(vis)) <= {0, 1}
    # ids outside the set are flagged
    bad = ids.copy()
    bad[::1000] = k + 5
    vb = s.evaluate(dq, cp.DeviceArray.from_numpy(bad)).numpy()
    assert np.all(vb[::1000] == 255)
    assert np.array_equal(np.delete(vb, np.arange(0, nq, 1000)), np.delete(vis, np.arange(0, nq, 1000)))
    s.close()


def test_dagset_stride4_and_equals_container(ctx, dset):
    Ds, shadows = dset
    _say("dagset stride 4")
    q4 = O.gen_queries_uniform(SEED0 + 11, 100_000, 4)
    s = cp.DagSet(shadows[:4])
    v4 = s.evaluate(cp.DeviceArray.from_numpy(q4)).numpy()
    v3 = s.evaluate(cp.DeviceArray.from_numpy(np.ascontiguousarray(q4[:, :3]))).numpy()
    assert np.array_equal(v3, v4)
    # DAG 2 through a one-DAG container with an identity matrix (a separate build: the
    # container takes ownership)
    D = Ds[2]
    c = cp.CompressedShadowContainer(cp.CompressedShadow.create(cp.MinMaxHierarchy(D)))
    c.moveToGPU()
    sel = np.arange(2, 100_000, 4)
    vc = c.evaluate(cp.DeviceArray.from_numpy(np.ascontiguousarray(q4[sel]))).numpy()
    assert np.array_equal(vc, v4[sel])
    s.close()


def test_dagset_rejects_bad_arguments(ctx, dset):
    _, shadows = dset
    small = cp.CompressedShadow.create(cp.MinMaxHierarchy(np.full((8, 8), 0.5, np.float32)))
    with pytest.raises(cp.CpvsError):
        cp.DagSet([shadows[0], small])  # numLevels 4: no leafmasks
    with pytest.raises(cp.CpvsError):
        cp.DagSet([])


def test_uniform_range_generator(ctx):
    _say("uniform range")
    got = cp.gen_queries_uniform_range(SEED0 + 3, 5, 7, 1000, 3).numpy()
    full = O.gen_queries_uniform(SEED0 + 3, 5 + 7 * 1000, 3)
    assert np.array_equal(got, full[5::7][:1000])
    got4 = cp.gen_queries_uniform_range(SEED0 + 3, 0, 1, 500, 4).numpy()
    assert np.array_equal(got4, O.gen_queries_uniform(SEED0 + 3, 500, 4))


def test_dagset_round_robin_many_dags(ctx):
    """Round-robin lookups over 37 DAGs (three tiles of 16, the last partial; groups on several XCDs)
    and a ragged query count, with 3- and 4-float points: each DAG's queries as the oracle's walk."""
    k = 37
    Ds = [O.gen_depth(16 << (j % 4), SEED0 + 900 + j, 1 + j % 5, j % 2 == 0) for j in range(k)]
    shadows = [cp.CompressedShadow.create(cp.MinMaxHierarchy(D)) for D in Ds]
    s = cp.DagSet(shadows)
    nq = k * 1500 + 17
    for stride in (3, 4):
        _say("dagset round robin", k, nq, stride)
        q = O.gen_queries_uniform(SEED0 + 910 + stride, nq, stride)
        rr = s.evaluate(cp.DeviceArray.from_numpy(q)).numpy()
        for j, sh in enumerate(shadows):
            words, n = sh.getDAG(), sh.getNumLevels()
            sel = (np.arange(nq) % k) == j
            assert np.array_equal(rr[sel], O.evaluate(words, O.top_level_grid([words]), n, 0, q[sel], nthreads=8)), j
    s.close()
    for sh in shadows:
        sh.close()
