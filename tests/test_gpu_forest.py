"""Forest batches: many maps (and the z-tiles of one map) built through one launch chain
(cpvs_build_batch_config with forest_mtexels > 0), checked word for word against plain builds
and the CPU oracle, in every leaf-kernel shape, with and without hipGraph capture, with forests
small enough to split a batch into many, and through the configuration autotuner.  Needs a GPU."""
import numpy as np
import pytest

import cpvs_amd as cp
import oracle as O

pytestmark = pytest.mark.gpu
SEED0 = 0x43505653


def _say(*a):
    print("[case]", *a, flush=True)


def _maps():
    """(name, depth, zTileIndex, zTileNum): generator maps of many sizes and z-tiles, a map with no
    leafmask levels (8^2), a map that ends early (uniform 0.625), all-shadow, all-lit, maps of
    special values and tall walls."""
    out = []
    specs = [(8, 1, False, 0, 1), (16, 2, True, 0, 1), (64, 4, True, 1, 2), (256, 8, True, 0, 1), (1024, 32, False, 0, 1),
             (512, 16, True, 3, 4), (128, 4, False, 0, 1), (2048, 64, True, 0, 1), (32, 2, True, 0, 2), (16, 1, True, 0, 8),
             (4096, 128, True, 5, 8)]
    for k, (N, boxes, noise, t, T) in enumerate(specs):
        out.append((f"gen{N}_t{t}of{T}", O.gen_depth(N, SEED0 + 1000 * (k + 1), boxes, noise), t, T))
    out.append(("uniform0.625", np.full((64, 64), 0.625, np.float32), 0, 1))
    out.append(("shadow", np.zeros((64, 64), np.float32), 0, 1))
    out.append(("lit", np.ones((64, 64), np.float32), 0, 1))
    rng = np.random.default_rng(5)
    sp = rng.random((128, 128), dtype=np.float32)
    for frac, val in ((0.05, np.nan), (0.03, np.inf), (0.03, -np.inf), (0.05, -0.5), (0.05, 3.0)):
        sp[rng.random((128, 128)) < frac] = val
    out.append(("special", sp, 1, 2))
    walls = rng.random((256, 256), dtype=np.float32)
    by, bx = np.nonzero(rng.random((32, 32)) < 0.2)
    walls[8 * by, 8 * bx] = 0.0
    walls[8 * by + 7, 8 * bx + 7] = 1.0
    out.append(("walls", walls, 0, 1))
    return out


@pytest.fixture(scope="module")
def mix(ctx):
    maps = _maps()
    depths = [cp.DeviceArray.from_numpy(D) for _, D, _, _ in maps]
    tiles = [(t, T) for _, _, t, T in maps]
    plain = []
    for name, D, t, T in maps:
        cs = cp.CompressedShadow.create(cp.MinMaxHierarchy(D), t, T)
        plain.append((cs.getNumLevels(), cs.getDAG(), cs.level_stats()))
    return maps, depths, tiles, plain


def _check(maps, shadows, plain, what):
    for (name, D, t, T), sh, (n, words, stats) in zip(maps, shadows, plain):
        assert sh.getNumLevels() == n, (what, name)
        assert np.array_equal(sh.getDAG(), words), (what, name)
        got = sh.level_stats()
        assert np.array_equal(got[0], stats[0]) and np.array_equal(got[1], stats[1]), (what, name)


CONFIGS = [
    dict(lanes=1, forest_mtexels=256),
    dict(lanes=3, forest_mtexels=256),
    dict(lanes=2, forest_mtexels=1),          # many forests of about one map each
    dict(lanes=1, forest_mtexels=4096),       # the whole batch as one forest
    dict(lanes=1, forest_mtexels=256, leaf_cols=4),
    dict(lanes=1, forest_mtexels=256, leaf_waves=2),
    dict(lanes=1, forest_mtexels=256, leaf_cols=4, leaf_waves=2),
    dict(lanes=2, forest_mtexels=256, graph=True),
    dict(lanes=2, forest_mtexels=0),          # one map per launch chain
]


@pytest.mark.parametrize("cfg", CONFIGS, ids=lambda c: "-".join(f"{k}{int(v)}" for k, v in c.items()))
def test_forest_matches_plain_builds(ctx, mix, cfg):
    maps, depths, tiles, plain = mix
    _say("forest config", cfg)
    shadows, done = cp.build_batch(depths, tiles, config=cp.batch_config(**cfg))
    _check(maps, shadows, plain, cfg)
    assert np.all(np.isfinite(done)) and np.all(done > 0.0)
    for sh in shadows:
        sh.close()


def test_forest_matches_oracle(ctx, mix):
    maps, depths, tiles, _ = mix
    _say("forest vs oracle")
    shadows, _ = cp.build_batch(depths, tiles, config=cp.batch_config(lanes=1, forest_mtexels=4096))
    for (name, D, t, T), sh in zip(maps, shadows):
        if D.shape[0] > 1024:
            continue  # covered by the plain-build comparison
        words, n = O.create(O.Pyramid(D), t, T, merge_mode=1)
        assert sh.getNumLevels() == n, name
        assert np.array_equal(sh.getDAG(), words), name


def test_forest_with_sort_merge(mix):
    maps, depths, tiles, plain = mix
    _say("forest sort merge")
    c2 = cp.Context(0)
    try:
        c2.set_merge("sort")
        dd = [cp.DeviceArray.from_numpy(D, ctx=c2) for _, D, _, _ in maps]
        shadows, _ = cp.build_batch(dd, tiles, ctx=c2, config=cp.batch_config(lanes=2, forest_mtexels=64))
        try:
            _check(maps, shadows, plain, "sort")
            assert all(sh.merge_info()[0] == "sort" for sh in shadows if sh.getNumLevels() >= 5)
        finally:
            for sh in shadows:
                sh.close()
            for d in dd:
                d.free()
    finally:
        c2.close()


def test_forest_of_z_tiles_shares_pyramids(ctx):
    """Every z-tile of three maps, shuffled (the tiled configuration's shape): one forest per
    map's z-tiles, or all of them in one."""
    _say("forest z-tiles")
    specs = [(1024, 32, 8), (512, 8, 4), (2048, 64, 2)]  # N, boxes, zTileNum
    Ds = [O.gen_depth(N, SEED0 + 91 * k, b, True) for k, (N, b, _) in enumerate(specs)]
    depths = [cp.DeviceArray.from_numpy(D) for D in Ds]
    items = [(k, t) for k, (_, _, T) in enumerate(specs) for t in range(T)]
    np.random.default_rng(13).shuffle(items)
    pyrs = [O.Pyramid(D) for D in Ds]
    exp = {}
    for k, t in items:
        if specs[k][0] <= 1024:
            exp[(k, t)] = O.create(pyrs[k], t, specs[k][2], merge_mode=1)[0]
        else:
            exp[(k, t)] = cp.CompressedShadow.create(cp.MinMaxHierarchy(Ds[k]), t, specs[k][2]).getDAG()
    for mt in (1, 4096):
        shadows, _ = cp.build_batch([depths[k] for k, _ in items], [(t, specs[k][2]) for k, t in items],
                                    config=cp.batch_config(lanes=2, forest_mtexels=mt))
        for (k, t), sh in zip(items, shadows):
            assert np.array_equal(sh.getDAG(), exp[(k, t)]), (k, t, mt)


def test_forest_tile_of_the_65536_map(ctx):
    """The eight z-tiles of one 8192^2 tile of the 65536^2 configuration as one forest, against
    eight plain builds."""
    full, tile, L = 65536, 8192, 8
    _say("forest 65536^2 tile")
    d = cp.gen_depth(full, SEED0 + 2, 16384, True, 3 * tile, 5 * tile, tile)
    mm = cp.MinMaxHierarchy(d, borrow=True)
    plain = [cp.CompressedShadow.create(mm, t, L).getDAG() for t in range(L)]
    mm.close()
    shadows, _ = cp.build_batch([d] * L, [(t, L) for t in range(L)], config=cp.batch_config(lanes=1, forest_mtexels=256))
    for t, sh in enumerate(shadows):
        assert np.array_equal(sh.getDAG(), plain[t]), t
    assert sum(len(w) for w in plain) > 100000
    d.free()


def test_forest_lookups_through_dagset(ctx, mix):
    """DAGs from a forest answer lookups exactly as the oracle's traverse.cs transcription."""
    maps, depths, tiles, _ = mix
    _say("forest lookups")
    shadows, _ = cp.build_batch(depths, tiles, config=cp.batch_config(lanes=1, forest_mtexels=4096))
    q = O.gen_queries_uniform(77, 1 << 16, 4)
    dq = cp.DeviceArray.from_numpy(q)
    for (name, D, t, T), sh in zip(maps, shadows):
        if sh.getNumLevels() < 5 or T != 1:
            continue
        w, n = sh.getDAG(), sh.getNumLevels()
        c = cp.CompressedShadowContainer(sh)
        c.copyToGPU()
        vis = c.evaluate(dq).numpy()
        exp = O.evaluate(w, O.top_level_grid([w]), n, 0, q, nthreads=8)
        assert np.array_equal(vis, exp), name


def test_tune_batch_config_picks_a_candidate(ctx, mix):
    maps, depths, tiles, plain = mix
    _say("tune batch config")
    cands = [cp.batch_config(lanes=2, forest_mtexels=0), cp.batch_config(lanes=2, forest_mtexels=256),
             cp.batch_config(lanes=1, forest_mtexels=4096, leaf_cols=4)]
    before = cp.get_batch_config()
    try:
        best, ms = cp.tune_batch_config(depths, cands, tiles)
        assert 0 <= best < len(cands) and len(ms) == 3 and all(m > 0 for m in ms)
        assert cp.get_batch_config() == cands[best].as_dict()
        shadows, _ = cp.build_batch(depths, tiles)  # with the tuned configuration
        _check(maps, shadows, plain, "tuned")
    finally:
        cp.set_batch_config(cp.batch_config(**{k: v for k, v in before.items()}))


def test_batch_config_errors(ctx, mix):
    _, depths, tiles, _ = mix
    with pytest.raises(cp.CpvsError):
        cp.build_batch(depths[:2], tiles[:2], config=cp.batch_config(leaf_cols=16))
    with pytest.raises(cp.CpvsError):
        cp.build_batch(depths[:2], tiles[:2], config=cp.batch_config(lanes=17))


def test_release_scratch_between_batches(mix):
    """cpvs_ctx_release_scratch frees the lanes' grown arenas and the buffer cache; the next batch
    grows them again and builds the same words."""
    maps, _, tiles, plain = mix
    c2 = cp.Context(0)
    try:
        dd = [cp.DeviceArray.from_numpy(D, ctx=c2) for _, D, _, _ in maps]
        cfg = cp.batch_config(lanes=2, forest_mtexels=256)
        for rep in range(2):
            _say("release scratch, batch", rep)
            shadows, _ = cp.build_batch(dd, tiles, ctx=c2, config=cfg)
            _check(maps, shadows, plain, ("release", rep))
            for sh in shadows:
                sh.close()
            assert c2.release_scratch() > 0
        assert c2.release_scratch() == 0  # nothing left to free
    finally:
        c2.close()
