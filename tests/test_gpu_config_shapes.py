"""DAG words and lookups at the benchmark configurations' own shapes, checked against the CPU
oracle (and the brute-force depth compare where the oracle is not run): zTileNum = 8 (the tiled
65536^2 map's z-tiles), a tile of that map itself, 4096^2 and 16384^2 maps (the batched and
serving mixes' sizes), batches and a server that include 4096^2 and 8192^2 maps, and every one of
the 16.6 M G-buffer queries of the 8192^2 headline workload.  Needs a GPU."""
import numpy as np
import pytest

import cpvs_amd as cp
import oracle as O
from tests.dagcheck import check_dag

pytestmark = pytest.mark.gpu
SEED0 = 0x43505653


def _say(*a):
    print("[case]", *a, flush=True)


def _device_map(N, seed, boxes, noise=True, full=None, x0=0, y0=0):
    """A generator map made on the device, and its host copy for the oracle."""
    d = cp.gen_depth(full or N, seed, boxes, noise, x0, y0, N)
    return d, d.numpy()


def _brute(D, T, n, q):
    """Brute-force visibility (midpoint depth compare at leaf resolution) of NDC queries q against
    an N x N map split into T z-tiles with one grid cell per tile: the value every lookup must
    return (traverse.cs quantisation, src/CompressedShadowUtil.h:51-57)."""
    R = np.float32((1 << (n - 1)) - 1)
    p = (((q[:, :3] + np.float32(1)) * np.float32(0.5)) * R).astype(np.int64)
    return O.brute_visible(D, T, p[:, 0], p[:, 1], p[:, 2])


def test_every_z_tile_of_eight_bit_exact(ctx):
    """zTileNum = 8 (the tiled configuration's T), all eight z-tiles of a 1024^2 map."""
    N, T = 1024, 8
    d, D = _device_map(N, SEED0 + 808, 32)
    mm = cp.MinMaxHierarchy(d, borrow=True)
    pyr = O.Pyramid(D)
    for t in range(T):
        _say("z-tile", t, "of", T, "N", N)
        cs = cp.CompressedShadow.create(mm, t, T)
        wr, nr = O.create(pyr, t, T, merge_mode=1)
        assert cs.getNumLevels() == nr
        assert np.array_equal(cs.getDAG(), wr), t


def test_tile_of_the_65536_map_bit_exact(ctx):
    """Tile (0, 0) of the 65536^2 configuration (8192^2, generated with the full map's 16,384
    boxes), z-tiles of eight: the plane's slab (4), a slab of box tops (1) and the empty top slab
    (7)."""
    full, tile, T = 65536, 8192, 8
    d, D = _device_map(tile, SEED0 + 2, 16384, True, full, 0, 0)
    mm = cp.MinMaxHierarchy(d, borrow=True)
    pyr = O.Pyramid(D)
    sizes = {}
    for t in (4, 1, 7):
        _say("65536^2 tile (0,0) z-tile", t)
        cs = cp.CompressedShadow.create(mm, t, T)
        w = cs.getDAG()
        wr, nr = O.create(pyr, t, T, merge_mode=1)
        assert cs.getNumLevels() == nr == 14
        assert np.array_equal(w, wr), t
        sizes[t] = len(w)
    assert sizes[4] > 100000 and sizes[1] > 1000 and sizes[7] == 1, sizes


def test_4096_map_bit_exact(ctx):
    N = 4096
    _say("4096^2")
    d, D = _device_map(N, SEED0 + 4096, 512)
    cs = cp.CompressedShadow.create(cp.MinMaxHierarchy(d, borrow=True))
    w = cs.getDAG()
    wr, nr = O.create(O.Pyramid(D), 0, 1, merge_mode=1)
    assert cs.getNumLevels() == nr == 13
    assert np.array_equal(w, wr)
    check_dag(w, nr)


def test_16384_map_bit_exact(ctx):
    """The largest map of the serving mix: words identical to the oracle, layout invariants, and
    1 M lookups against the brute force."""
    N = 16384
    _say("16384^2 generate")
    d, D = _device_map(N, SEED0 + 16384, 32 * 256)
    _say("16384^2 build")
    cs = cp.CompressedShadow.create(cp.MinMaxHierarchy(d, borrow=True))
    w = cs.getDAG()
    n = cs.getNumLevels()
    assert n == 15 and cs.merge_info()[0] == "table"
    _say("16384^2 oracle")
    wr, nr = O.create(O.Pyramid(D), 0, 1, merge_mode=1)
    assert nr == n and np.array_equal(w, wr), (len(w), len(wr))
    check_dag(w, n)
    c = cp.CompressedShadowContainer(cs)
    c.moveToGPU()
    q = cp.gen_queries_gbuffer(d, 23, 1024, 512)
    vis = c.evaluate(q).numpy()
    assert np.array_equal(vis, _brute(D, 1, n, q.numpy()))
    c.close()
    d.free()


def test_batch_with_large_maps_matches_oracle(ctx):
    """A batch of the batched configuration's sizes, its largest (8192^2) and 4096^2 maps
    included, every DAG compared with the oracle (not with plain builds)."""
    specs = [(8192, 512), (256, 2), (4096, 512), (1024, 32), (512, 8), (2048, 128), (256, 2)]
    maps = [_device_map(N, SEED0 + 3 + 1000 * (i + 1), b) for i, (N, b) in enumerate(specs)]
    _say("batch of", [N for N, _ in specs])
    shadows, done = cp.build_batch([d for d, _ in maps], lanes=4)
    assert len(done) == len(specs)
    for (N, _), (d, D), sh in zip(specs, maps, shadows):
        _say("batch oracle", N)
        wr, nr = O.create(O.Pyramid(D), 0, 1, merge_mode=1)
        assert sh.getNumLevels() == nr
        assert np.array_equal(sh.getDAG(), wr), N
    for sh in shadows:
        sh.close()
    for d, _ in maps:
        d.free()


def test_server_with_large_maps_matches_oracle(ctx):
    """The concurrent server building a 4096^2 and an 8192^2 map while lookups on a resident DAG
    run: words equal the oracle's, and each lookup request equals the brute force."""
    specs = [(8192, 512), (4096, 512), (512, 8)]
    maps = [_device_map(N, SEED0 + 4 + 1000 * (i + 1), b) for i, (N, b) in enumerate(specs)]
    q = cp.gen_queries_uniform(SEED0 + 41, 1 << 20, 3)
    qh = q.numpy()
    srv = cp.Server(lanes=2)
    try:
        first = srv.build(maps[2][0])
        srv.wait(first)
        vis = cp.DeviceArray((8 << 20,), np.uint8)
        _say("server builds", [N for N, _ in specs[:2]], "with lookups in flight")
        tickets = [srv.build(maps[0][0]), srv.build(maps[1][0])]
        looks = [srv.lookup(first, q, vis, count=1 << 18, offset=(k << 18) % (1 << 20), vis_offset=k << 18)
                 for k in range(8)]
        for t in tickets + looks:
            srv.wait(t)
        n512 = 10
        v = vis.numpy()
        for k in range(8):
            off = (k << 18) % (1 << 20)
            exp = _brute(maps[2][1], 1, n512, qh[off:off + (1 << 18)])
            assert np.array_equal(v[k << 18:(k + 1) << 18], exp), k
        for (N, _), (d, D), t in zip(specs[:2], maps[:2], tickets):
            _say("server oracle", N)
            wr, nr = O.create(O.Pyramid(D), 0, 1, merge_mode=1)
            assert np.array_equal(srv.dag_words(t), wr), N
    finally:
        srv.close()
    for d, _ in maps:
        d.free()


def test_headline_gbuffer_lookups_all_queries(ctx):
    """All 16,588,800 G-buffer queries of the 8192^2 headline workload (3840 x 2160 x 2) against
    the brute-force depth compare."""
    N = 8192
    _say("C1 full G-buffer")
    d, D = _device_map(N, SEED0 + 1, 512)
    cs = cp.CompressedShadow.create(cp.MinMaxHierarchy(d, borrow=True))
    n = cs.getNumLevels()
    c = cp.CompressedShadowContainer(cs)
    c.moveToGPU()
    q = cp.gen_queries_gbuffer(d, SEED0 + 17, 3840, 2160)
    assert q.shape[0] == 16588800
    vis = c.evaluate(q).numpy()
    qh = q.numpy()
    for lo in range(0, qh.shape[0], 1 << 22):  # in slices: the brute force holds int64 paths
        hi = min(lo + (1 << 22), qh.shape[0])
        assert np.array_equal(vis[lo:hi], _brute(D, 1, n, qh[lo:hi])), lo
    assert 0.1 < vis.mean() < 0.95  # both outcomes occur (about 80% lit)
    c.close()
    q.free()
    d.free()
