"""The leafmask-off mode (the reference built without #define LEAFMASKS, src/CompressedShadow.cpp:17,
shader/traverse.cs:8-9): DAG words against the oracle in the same mode, lookups, traversals and
PCF against the oracle's walks and the brute-force depth compare, containers and DAG sets of one
mode, the byte format's flag, batches and the server.  Needs a GPU."""
import numpy as np
import pytest

import cpvs_amd as cp
import oracle as O
from tests.dagcheck import check_dag
from tests.test_format import write

pytestmark = pytest.mark.gpu
SEED0 = 0x43505653


def _say(*a):
    print("[case]", *a, flush=True)


@pytest.fixture(scope="module")
def nctx():
    c = cp.Context(0)
    c.set_leafmasks(False)
    assert not c.leafmasks()
    yield c
    c.close()


def _maps():
    out = []
    for k, (N, boxes, noise) in enumerate([(8, 1, False), (16, 1, True), (32, 2, True), (64, 4, True), (128, 4, False),
                                            (256, 8, True), (512, 16, True), (1024, 32, True)]):
        out.append((f"gen{N}", O.gen_depth(N, SEED0 + 500 + k, boxes, noise)))
    rng = np.random.default_rng(3)
    out.append(("noise64", rng.random((64, 64), dtype=np.float32)))
    sp = rng.random((128, 128), dtype=np.float32)
    for frac, val in ((0.05, np.nan), (0.03, np.inf), (0.03, -np.inf), (0.05, -0.5), (0.05, 3.0)):
        sp[rng.random((128, 128)) < frac] = val
    out.append(("special128", sp))
    out.append(("uniform0.625", np.full((64, 64), 0.625, np.float32)))
    return out


@pytest.mark.parametrize("T", [1, 2, 4])
def test_words_match_oracle_without_leafmasks(nctx, T):
    for name, D in _maps():
        mm = cp.MinMaxHierarchy(D, ctx=nctx)
        pyr = O.Pyramid(D)
        for t in range(T):
            _say("leafmasks off", name, t, T)
            cs = cp.CompressedShadow.create(mm, t, T)
            w, n = cs.getDAG(), cs.getNumLevels()
            wr, nr = O.create(pyr, t, T, leafmasks=False)
            assert n == nr and np.array_equal(w, wr), (name, t, T)
            assert not cs.has_leafmasks()
            if t == 0:
                check_dag(w, n, leafmasks=False)


def test_reference_images_without_leafmasks(nctx, ref_fixtures):
    for name in ("depths8x8", "depths16x16", "depths32x32"):
        D = ref_fixtures["arrays"][name]
        cs = cp.CompressedShadow.create(cp.MinMaxHierarchy(D, ctx=nctx))
        assert np.array_equal(cs.getDAG(), O.create(O.Pyramid(D), leafmasks=False)[0]), name


def test_sort_merge_without_leafmasks():
    c2 = cp.Context(0)
    try:
        c2.set_leafmasks(False)
        c2.set_merge("sort")
        D = O.gen_depth(512, SEED0 + 77, 16, True)
        cs = cp.CompressedShadow.create(cp.MinMaxHierarchy(D, ctx=c2))
        assert cs.merge_info()[0] == "sort"
        assert np.array_equal(cs.getDAG(), O.create(O.Pyramid(D), leafmasks=False)[0])
    finally:
        c2.close()


def test_traverse_without_leafmasks(nctx):
    """CompressedShadow::traverse(ndc, false) (src/CompressedShadow.cpp:404-463) on DAGs built without
    leafmasks; asking for leafmasks is an error."""
    D = O.gen_depth(256, SEED0 + 31, 8, True)
    cs = cp.CompressedShadow.create(cp.MinMaxHierarchy(D, ctx=nctx))
    w, n = cs.getDAG(), cs.getNumLevels()
    pts = O.gen_queries_uniform(SEED0 + 32, 20000, 3)
    got = cs.traverse_many(pts, False)
    exp = np.array([O.traverse_cpu(w, n, q, False) for q in pts])
    assert np.array_equal(got, exp)
    res = (1 << (n - 1)) - 1
    p = (((pts + np.float32(1)) * np.float32(0.5)) * np.float32(res)).astype(np.int64)
    assert np.array_equal(got, O.brute_visible(D, 1, p[:, 0], p[:, 1], p[:, 2]))
    with pytest.raises(cp.CpvsError):
        cs.traverse_many(pts[:4], True)


def test_container_lookups_without_leafmasks(nctx):
    """traverse.cs with MIN_LEVEL 0 through a container, against the oracle's walk in the same mode
    and the brute force."""
    N = 1024
    D = O.gen_depth(N, SEED0 + 41, 32, True)
    cs = cp.CompressedShadow.create(cp.MinMaxHierarchy(D, ctx=nctx))
    w, n = cs.getDAG(), cs.getNumLevels()
    c = cp.CompressedShadowContainer(cs)
    c.moveToGPU()
    q = O.gen_queries_uniform(SEED0 + 42, 1 << 18, 4)
    vis = c.evaluate(cp.DeviceArray.from_numpy(q, ctx=nctx)).numpy()
    assert np.array_equal(vis, O.evaluate(w, O.top_level_grid([w]), n, 0, q, nthreads=8, leafmasks=False))
    p = np.array([O.path_from_ndc(x[:3], n, 0) for x in q[:20000]])
    assert np.array_equal(vis[:20000], O.brute_visible(D, 1, p[:, 0], p[:, 1], p[:, 2]))
    # the image form over a G-buffer of the same map (ragged: 333 x 77 pixels)
    g = cp.gen_queries_gbuffer(cp.DeviceArray.from_numpy(D, ctx=nctx), SEED0 + 43, 333, 77, ctx=nctx).numpy()
    g = np.ascontiguousarray(g[:77 * 333])  # the generator gives two points per pixel: the first image
    img = c.evaluate_image(cp.DeviceArray.from_numpy(g.reshape(77, 333, 4), ctx=nctx)).numpy().reshape(-1)
    flat = c.evaluate(cp.DeviceArray.from_numpy(g, ctx=nctx)).numpy()
    assert np.array_equal(img, flat)
    assert np.array_equal(flat, O.evaluate(w, O.top_level_grid([w]), n, 0, g, nthreads=8, leafmasks=False))
    c.close()


def test_tiled_container_and_pcf_without_leafmasks(nctx):
    """A 2x2x2 tiled container: every lookup against the brute force, PCF taps across tiles."""
    L, tile = 2, 128
    full = L * tile
    D = O.gen_depth(full, SEED0 + 51, 8, True)
    c = cp.CompressedShadowContainer(L, ctx=nctx)
    for ty in range(L):
        for tx in range(L):
            mm = cp.MinMaxHierarchy(np.ascontiguousarray(D[ty * tile:(ty + 1) * tile, tx * tile:(tx + 1) * tile]), ctx=nctx)
            for tz in range(L):
                c.set(cp.CompressedShadow.create(mm, tz, L), tx, ty, tz)
    c.setFilterSize(3)
    c.moveToGPU()
    rng = np.random.default_rng(9)
    px = rng.integers(0, full, 40000)
    py = rng.integers(0, full, 40000)
    z = D[py, px] + rng.uniform(-3, 3, len(px)).astype(np.float32) / full
    q = np.stack([(px + 0.5) / full * 2 - 1, (py + 0.5) / full * 2 - 1, np.clip(z, 0, 1) * 2 - 1], 1).astype(np.float32)
    dq = cp.DeviceArray.from_numpy(q, ctx=nctx)
    n = c.info()[0]
    pp = np.array([O.path_from_ndc(x, n, 1) for x in q])
    vis = c.evaluate(dq).numpy()
    assert np.array_equal(vis, O.brute_visible(D, 1, pp[:, 0], pp[:, 1], pp[:, 2]))
    got = c.evaluate_pcf(dq).numpy()
    lit = np.zeros(len(q), np.int64)
    for a in range(3):
        for b in range(3):
            x = np.clip(pp[:, 0] - 1 + b, 0, full - 1)
            y = np.clip(pp[:, 1] - 1 + a, 0, full - 1)
            lit += O.brute_visible(D, 1, x, y, pp[:, 2])
    assert np.array_equal(got, lit.astype(np.float32) / np.float32(9))
    c.close()


def test_one_mode_per_container_and_set(nctx):
    D = O.gen_depth(128, SEED0 + 61, 4, True)
    off = cp.CompressedShadow.create(cp.MinMaxHierarchy(D, ctx=nctx))
    nctx.set_leafmasks(True)
    try:
        on = cp.CompressedShadow.create(cp.MinMaxHierarchy(D, ctx=nctx))
    finally:
        nctx.set_leafmasks(False)
    assert on.has_leafmasks() and not off.has_leafmasks()
    assert not np.array_equal(on.getDAG(), off.getDAG())
    with pytest.raises(cp.CpvsError):
        cp.DagSet([off, on])
    c = cp.CompressedShadowContainer(2, ctx=nctx)
    nctx.set_leafmasks(True)
    try:
        c.set(cp.CompressedShadow.create(cp.MinMaxHierarchy(D, ctx=nctx)), 0, 0, 0)
    finally:
        nctx.set_leafmasks(False)
    for k in range(1, 8):
        c.set(cp.CompressedShadow.create(cp.MinMaxHierarchy(D, ctx=nctx)), k & 1, (k >> 1) & 1, k >> 2)
    with pytest.raises(cp.CpvsError):
        c.copyToGPU()
    c.close()
    # each DAG's own lookup follows its own mode: both give the brute-force answer
    q = O.gen_queries_uniform(SEED0 + 62, 50000, 3)
    dq = cp.DeviceArray.from_numpy(q, ctx=nctx)
    n = off.getNumLevels()
    p = np.array([O.path_from_ndc(x, n, 0) for x in q])
    exp = O.brute_visible(D, 1, p[:, 0], p[:, 1], p[:, 2])
    assert np.array_equal(off.evaluate(dq).numpy(), exp)
    assert np.array_equal(on.evaluate(dq).numpy(), exp)
    s = cp.DagSet([off, cp.CompressedShadow.create(cp.MinMaxHierarchy(O.gen_depth(64, 5, 2, True), ctx=nctx))])
    v = s.evaluate(dq).numpy()
    assert np.array_equal(v[0::2], exp[0::2])
    s.close()


def test_format_round_trip_without_leafmasks(nctx, tmp_path):
    D = O.gen_depth(256, SEED0 + 71, 8, True)
    cs = cp.CompressedShadow.create(cp.MinMaxHierarchy(D, ctx=nctx), 1, 2)
    w, n = cs.getDAG(), cs.getNumLevels()
    data = cs.serialize()
    assert data == write(w, n, 1, 2, flags=1)
    back = cp.CompressedShadow.deserialize(data, ctx=nctx)
    assert not back.has_leafmasks() and np.array_equal(back.getDAG(), w)
    pts = O.gen_queries_uniform(SEED0 + 72, 5000, 3)
    assert np.array_equal(back.traverse_many(pts, False), cs.traverse_many(pts, False))
    path = str(tmp_path / "nl.cpvs")
    cs.save(path)
    # a context in the default mode loads it in the DAG's own mode
    loaded = cp.CompressedShadow.load(path)
    assert not loaded.has_leafmasks() and np.array_equal(loaded.getDAG(), w)


def test_batch_and_server_without_leafmasks(nctx):
    """Batches (a forest configuration is run map by map) and the server build in the context's
    mode: words equal plain builds, server lookups equal the brute force."""
    specs = [(512, 16), (64, 2), (1024, 32), (16, 1), (256, 8)]
    Ds = [O.gen_depth(N, SEED0 + 80 + k, b, True) for k, (N, b) in enumerate(specs)]
    depths = [cp.DeviceArray.from_numpy(D, ctx=nctx) for D in Ds]
    plain = [cp.CompressedShadow.create(cp.MinMaxHierarchy(D, ctx=nctx)).getDAG() for D in Ds]
    for cfg in (cp.batch_config(lanes=2, forest_mtexels=256), cp.batch_config(lanes=3, forest_mtexels=0)):
        _say("batch leafmasks off", cfg.as_dict())
        shadows, _ = cp.build_batch(depths, ctx=nctx, config=cfg)
        for k, sh in enumerate(shadows):
            assert not sh.has_leafmasks()
            assert np.array_equal(sh.getDAG(), plain[k]), k
    srv = cp.Server(lanes=2, ctx=nctx)
    try:
        q = cp.gen_queries_uniform(SEED0 + 88, 1 << 16, 3, ctx=nctx)
        qh = q.numpy()
        tickets = [srv.build(depths[k]) for k in (0, 2, 4)]
        vis = cp.DeviceArray((3 << 16,), np.uint8, nctx)
        for t in tickets:
            srv.wait(t)
        looks = [srv.lookup(t, q, vis, vis_offset=i << 16) for i, t in enumerate(tickets)]
        for t in looks:
            srv.wait(t)
        v = vis.numpy()
        for i, k in enumerate((0, 2, 4)):
            assert np.array_equal(srv.dag_words(tickets[i]), plain[k]), k
            n = int(np.log2(specs[k][0])) + 1
            R = np.float32((1 << (n - 1)) - 1)
            p = (((qh + np.float32(1)) * np.float32(0.5)) * R).astype(np.int64)
            assert np.array_equal(v[i << 16:(i + 1) << 16], O.brute_visible(Ds[k], 1, p[:, 0], p[:, 1], p[:, 2])), k
    finally:
        srv.close()


def test_large_map_without_leafmasks(nctx):
    """A 4096^2 map down to level 0: words equal the oracle's, layout invariants hold, and 1 M
    G-buffer lookups equal the brute force."""
    N = 4096
    _say("4096^2 leafmasks off")
    d = cp.gen_depth(N, SEED0 + 4097, 512, True, ctx=nctx)
    D = d.numpy()
    cs = cp.CompressedShadow.create(cp.MinMaxHierarchy(d, borrow=True, ctx=nctx))
    w, n = cs.getDAG(), cs.getNumLevels()
    wr, nr = O.create(O.Pyramid(D), leafmasks=False)
    assert n == nr and np.array_equal(w, wr)
    check_dag(w, n, leafmasks=False)
    c = cp.CompressedShadowContainer(cs)
    c.moveToGPU()
    q = cp.gen_queries_gbuffer(d, SEED0 + 4098, 1024, 1024, ctx=nctx)
    vis = c.evaluate(q).numpy()
    qh = q.numpy()
    R = np.float32((1 << (n - 1)) - 1)
    p = (((qh[:, :3] + np.float32(1)) * np.float32(0.5)) * R).astype(np.int64)
    assert np.array_equal(vis, O.brute_visible(D, 1, p[:, 0], p[:, 1], p[:, 2]))
    c.close()
    d.free()
