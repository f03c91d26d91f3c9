"""The CPU oracle pinned against the reference's own tests, closed forms and brute force.

CPU only.  Reference test citations: test/MinMaxTest.cpp, test/CompressedShadowUtilTest.cpp,
test/CompressedShadowTest.cpp (values recorded in tests/golden/reference_tests.json).
"""
import numpy as np
import pytest

import oracle as O
from oracle import structural
from tests.dagcheck import DagError, check_dag

SEED0 = 0x43505653


def _img(fx, name):
    return fx["arrays"][name]


# --- reference known answers -------------------------------------------------------------

def test_minmax_known_answers(ref_fixtures):  # MinMaxTest.cpp:41-90
    for e in ref_fixtures["minmax"]:
        p = O.Pyramid(_img(ref_fixtures, e[0]))
        if e[1] == "numLevels":
            assert p.getNumLevels() == e[2]
            continue
        _, L, x, y, which, exp, tol = e
        v = p.getMin(L, x, y) if which == "min" else p.getMax(L, x, y)
        if tol == 0:
            assert v == np.float32(exp), e
        else:
            assert abs(v - exp) < tol, e


def test_childmask_known_answer(ref_fixtures):  # CompressedShadowUtilTest.cpp:14-19
    for name, level, off, exp in ref_fixtures["childmask"]:
        assert O.create_childmask(O.Pyramid(_img(ref_fixtures, name)), level, off) == exp


def test_child_coordinates_known_answers(ref_fixtures):  # CompressedShadowUtilTest.cpp:21-65
    for mask, exp in ref_fixtures["num_children"]:
        assert len(O.child_coordinates(mask, (0, 0, 0))) == exp
    for mask, parent, exp in ref_fixtures["child_coordinates"]:
        assert O.child_coordinates(mask, parent)[0] == tuple(exp)


@pytest.mark.parametrize("mode", [0, 1])
def test_merge_level_known_answer(ref_fixtures, mode):  # CompressedShadowUtilTest.cpp:83-97
    m = ref_fixtures["merge_level"]
    out, mapping = O.merge_level(np.array(m["words"], np.uint32), 9, mode)
    for old, new in m["mapping"]:
        assert mapping[old] == new
    assert out[9] == m["res9"]
    assert mapping[18] == 0  # the third node merges into the first


@pytest.mark.parametrize("mode", [0, 1])
def test_is_equal_subtree_semantics(ref_fixtures, mode):  # CompressedShadowUtilTest.cpp:67-81
    for a, b, equal in ref_fixtures["is_equal_subtree"]:
        out, mapping = O.merge_level(np.array(a + b, np.uint32), 9, mode)
        assert (mapping[9] == 0) == equal


@pytest.mark.parametrize("mode", [0, 1])
def test_traverse_known_answers(ref_fixtures, mode):  # CompressedShadowTest.cpp:50-136
    for name, ndc, try_leaf, exp in ref_fixtures["traverse"]:
        w, n = O.create(O.Pyramid(_img(ref_fixtures, name)), merge_mode=mode)
        assert O.traverse_cpu(w, n, ndc, try_leaf) == exp, (name, ndc)


def test_traverse_sweeps(ref_fixtures):  # CompressedShadowTest.cpp:138-151
    for sw in ref_fixtures["sweeps"]:
        w, n = O.create(O.Pyramid(_img(ref_fixtures, sw["image"])))
        for y in range(*sw["y"]):
            for x in range(*sw["x"]):
                v = np.array([x / 32.0, y / 32.0, sw["z"]], np.float32)
                ndc = v * np.float32(2.0) + np.float32(-1.0)
                assert O.traverse_cpu(w, n, ndc, True) == sw["expected"], (x, y)


# --- closed forms ------------------------------------------------------------------------

@pytest.mark.parametrize("N", [8, 16, 64, 256, 512])
def test_pyramid_closed_form(N):
    rng = np.random.default_rng(N)
    D = rng.random((N, N), dtype=np.float32)
    D[rng.random((N, N)) < 0.05] = 0.5  # ties
    p = O.Pyramid(D)
    assert p.getNumLevels() == N.bit_length()
    for L in range(1, p.getNumLevels()):
        w = N >> L
        blk = D.reshape(w, N // w, w, N // w)
        assert np.array_equal(p.level(L)[..., 0], blk.min(axis=(1, 3)))
        assert np.array_equal(p.level(L)[..., 1], blk.max(axis=(1, 3)))


def test_leafmask_is_midpoint_test():
    rng = np.random.default_rng(3)
    D = rng.random((32, 32), dtype=np.float32)
    for (x, y, z) in [(0, 0, 0), (2, 4, 6), (6, 6, 2), (4, 0, 4)]:
        mask, payload = O.create_leaf_node(D, (x, y, z), 1)
        k = 0
        for s in range(8):
            zz = 4 * z + s
            blk = D[4 * y:4 * y + 8, 4 * x:4 * x + 8]
            bits = ((np.float32(zz) + np.float32(zz + 1)) * np.float32(0.5) <= blk * np.float32(32)).reshape(-1)
            m = sum(int(b) << i for i, b in enumerate(bits))
            code = (mask >> (2 * s)) & 3
            if m == (1 << 64) - 1:
                assert code == 1
            elif m == 0:
                assert code == 0
            else:
                assert code == 2 and payload[k] == m
                k += 1
        assert k == len(payload)


# --- independent layout and merge checks -------------------------------------------------

def _maps():
    out = []
    for N, nb, noise in [(16, 1, False), (32, 2, True), (64, 4, True), (128, 6, False), (128, 8, True)]:
        out.append((N, O.gen_depth(N, SEED0 + N, nb, noise)))
    rng = np.random.default_rng(7)
    for N in (16, 32, 64):
        out.append((N, rng.random((N, N), dtype=np.float32)))
    return out


@pytest.mark.parametrize("T", [1, 2, 4, 8])
def test_structural_builder_agrees(T):
    for N, D in _maps():
        p = O.Pyramid(D)
        for t in range(T):
            w, n = O.create(p, t, T, merge_mode=0)
            ws, ns = structural.build(D, t, T)
            assert n == ns
            assert np.array_equal(w, ws), (N, T, t)


@pytest.mark.parametrize("T", [1, 2, 4, 8])
def test_faithful_and_hashed_merge_agree(T):
    for N in (64, 256, 1024):
        D = O.gen_depth(N, SEED0 + 11 * N, max(1, 32 * N * N // (1024 * 1024)), N >= 256)
        p = O.Pyramid(D)
        for t in range(T):
            w0, _ = O.create(p, t, T, merge_mode=0)
            w1, _ = O.create(p, t, T, merge_mode=1)
            assert np.array_equal(w0, w1), (N, T, t)


def test_reference_images_layout(ref_fixtures):
    for name in ("depths8x8", "depths16x16", "depths32x32", "minmax8x8"):
        D = _img(ref_fixtures, name)
        w, n = O.create(O.Pyramid(D))
        ws, _ = structural.build(D)
        assert np.array_equal(w, ws), name
        check_dag(w, n)


def test_layout_invariants_on_generator_maps():
    for N in (64, 256, 1024):
        D = O.gen_depth(N, SEED0, max(1, 32 * N * N // (1024 * 1024)), True)
        w, n = O.create(O.Pyramid(D))
        counts = check_dag(w, n)
        assert len(counts) == n - 3  # root .. leaf level all present on a sloped map


# --- brute force: visibility of every voxel ----------------------------------------------

def _all_paths(N, zlo, zhi):
    z, y, x = np.meshgrid(np.arange(zlo, zhi), np.arange(N), np.arange(N), indexing="ij")
    return x.ravel(), y.ravel(), z.ravel()


@pytest.mark.parametrize("N,T", [(16, 1), (32, 1), (32, 2), (64, 1), (64, 4)])
def test_every_voxel_matches_brute_force(N, T):
    D = O.gen_depth(N, SEED0 + 5 * N + T, max(1, N // 16), N >= 32)
    if N <= 16:
        D = np.minimum(D, np.float32(0.9))  # keep the plane inside [0,1)
    p = O.Pyramid(D)
    for t in range(T):
        w, n = O.create(p, t, T)
        grid = O.top_level_grid([w])
        x, y, z = _all_paths(N, t * N, (t + 1) * N)
        got = np.array([O.traverse_path_cs(w, grid, n, 0, (xi, yi, zi)) for xi, yi, zi in zip(x, y, z - t * N)],
                       np.uint8)
        exp = O.brute_visible(D, T, x, y, z)
        assert np.array_equal(got, exp), (N, T, t)


def test_tiled_container_matches_brute_force():
    """2x2 xy tiles x 2 z-tiles behind a top-level grid (CompressedShadowContainer.cpp:52-91)."""
    N, L = 16, 2
    full = O.gen_depth(N * L, SEED0 + 99, 3, True)
    dags = [None] * (L ** 3)
    n = None
    for z in range(L):
        for y in range(L):
            for x in range(L):
                tile = np.ascontiguousarray(full[y * N:(y + 1) * N, x * N:(x + 1) * N])
                w, n = O.create(O.Pyramid(tile), z, L)
                dags[(z * L + y) * L + x] = w
    grid = O.top_level_grid(dags)
    words = np.concatenate(dags)
    R = N * L
    x, y, zz = _all_paths(R, 0, R)
    got = np.array([O.traverse_path_cs(words, grid, n, 1, p) for p in zip(x, y, zz)], np.uint8)
    exp = O.brute_visible(full, 1, x, y, zz)  # z resolution N*L = the full map's side
    assert np.array_equal(got, exp)
    # and through NDC quantisation (traverse.cs:43-48)
    pts = O.gen_queries_uniform(5, 2000, 4)
    vis = O.evaluate(words, grid, n, 1, pts)
    px = [O.path_from_ndc(q[:3], n, 1) for q in pts]
    exp2 = np.array([O.brute_visible(full, 1, [a], [b], [c])[0] for a, b, c in px], np.uint8)
    assert np.array_equal(vis, exp2)


# --- without leafmasks (the reference built without #define LEAFMASKS) ------------------------

@pytest.mark.parametrize("T", [1, 2])
def test_leafmasks_off_structural_builder_agrees(T):
    """CompressedShadow.cpp:17-41 with LEAFMASKS undefined: 1+k-word nodes down to level 0."""
    for N, D in _maps():
        if N > 64:
            continue
        p = O.Pyramid(D)
        for t in range(T):
            w, n = O.create(p, t, T, merge_mode=0, leafmasks=False)
            ws, ns = structural.build(D, t, T, leafmasks=False)
            assert n == ns
            assert np.array_equal(w, ws), (N, T, t)
            assert np.array_equal(O.create(p, t, T, merge_mode=1, leafmasks=False)[0], w), (N, T, t)
            counts = check_dag(w, n, leafmasks=False)
            if N >= 16 and len(w) > 1:
                assert not np.array_equal(w, O.create(p, t, T)[0])  # the leafmask DAG differs


@pytest.mark.parametrize("N,T", [(16, 1), (32, 1), (32, 2)])
def test_leafmasks_off_every_voxel_matches_brute_force(N, T):
    """traverse.cs with MIN_LEVEL 0 (traverse.cs:34-38) and CompressedShadow::traverse with
    tryLeafmasks false (CompressedShadow.cpp:404-463), every voxel against the depth compare."""
    D = O.gen_depth(N, SEED0 + 7 * N + T, max(1, N // 16), True)
    p = O.Pyramid(D)
    for t in range(T):
        w, n = O.create(p, t, T, leafmasks=False)
        grid = O.top_level_grid([w])
        x, y, z = _all_paths(N, t * N, (t + 1) * N)
        exp = O.brute_visible(D, T, x, y, z)
        with O.leafmasks_off():
            got = np.array([O.traverse_path_cs(w, grid, n, 0, (xi, yi, zi)) for xi, yi, zi in zip(x, y, z - t * N)],
                           np.uint8)
        assert np.array_equal(got, exp), (N, T, t)
        # the same walk with the leafmask lookup reads payload as pointers: never equal here
        if T == 1:
            pts = O.gen_queries_uniform(11, 4000, 3)
            vis = O.evaluate(w, grid, n, 0, pts, leafmasks=False)
            px = np.array([O.path_from_ndc(q, n, 0) for q in pts])
            assert np.array_equal(vis, O.brute_visible(D, 1, px[:, 0], px[:, 1], px[:, 2]))
            res = (1 << (n - 1)) - 1
            pc = (((pts + np.float32(1)) * np.float32(0.5)) * np.float32(res)).astype(np.int64)
            cpu = np.array([O.traverse_cpu(w, n, q, False) for q in pts])
            assert np.array_equal(cpu, O.brute_visible(D, 1, pc[:, 0], pc[:, 1], pc[:, 2]))


def test_leafmasks_off_planted_mistake_is_caught():
    """A level-0 child flipped from lit to shadow: the brute-force comparison sees it."""
    N = 32
    D = O.gen_depth(N, SEED0 + 21, 2, True)
    w, n = O.create(O.Pyramid(D), leafmasks=False)
    counts = check_dag(w, n, leafmasks=False)
    last = len(w) - 1
    while w[last] & 0x5555 == 0:  # a level-0 node (masks only: no pointers there) with a lit child
        last -= 1
    bad = w.copy()
    bit = int(w[last] & 0x5555).bit_length() - 1
    bad[last] ^= np.uint32(1 << bit)
    grid = O.top_level_grid([w])
    x, y, z = _all_paths(N, 0, N)
    exp = O.brute_visible(D, 1, x, y, z)
    with O.leafmasks_off():
        got = np.array([O.traverse_path_cs(bad, grid, n, 0, q) for q in zip(x, y, z)], np.uint8)
    assert not np.array_equal(got, exp)
    assert counts[-1] > 0


# --- NaN depths: cells partial under a parent that is not -------------------------------------

@pytest.mark.parametrize("leafmasks", [True, False])
def test_nan_maps_structural_builder_agrees(leafmasks):
    """Maps with scattered NaN and infinite texels, against the structural restatement with the
    reference's std::min/std::max pair order; the maps make partial cells that no partial parent
    reaches (which the top-down build never visits)."""
    from tests.nanmaps import excess, sparse_nan_map
    seen = 0
    for N in (16, 32, 64):
        for k in range(3):
            D = sparse_nan_map(N, 900 + 10 * N + k, 0.02, 0.01)
            p = O.Pyramid(D)
            for T in (1, 2):
                for t in range(T):
                    w, n = O.create(p, t, T, leafmasks=leafmasks)
                    ws, _ = structural.build(D, t, T, leafmasks=leafmasks)
                    assert np.array_equal(w, ws), (N, k, t, T)
                    check_dag(w, n, leafmasks=leafmasks)
                    seen += excess(D, t, T, leafmasks)
    assert seen > 0


# --- planted mistakes: each check above must catch a plausible bug ------------------------

def test_planted_mistakes_are_caught():
    D = O.gen_depth(64, SEED0 + 3, 4, True)
    w, n = O.create(O.Pyramid(D))
    check_dag(w, n)
    # (1) two root-children swapped in layout (first-occurrence order broken)
    root_k = int(bin(w[0] & 0xAAAA).count("1"))
    assert root_k >= 2
    bad = w.copy()
    bad[1], bad[2] = w[2], w[1]
    with pytest.raises(DagError):
        check_dag(bad, n)
    # (2) a leafmask bit flipped: the brute-force comparison sees it
    bad = w.copy()
    bad[-1] ^= 1
    grid = O.top_level_grid([w])
    x, y, z = _all_paths(64, 0, 64)
    exp = O.brute_visible(D, 1, x, y, z)
    got = np.array([O.traverse_path_cs(bad, grid, n, 0, p) for p in zip(x, y, z)], np.uint8)
    assert not np.array_equal(got, exp)
    # (3) child index transposed (z|y<<1|x<<2 instead of x|y<<1|z<<2): traversal disagrees
    got_t = np.array([O.traverse_path_cs(w, grid, n, 0, (c, b, a)) for a, b, c in zip(x, y, z)], np.uint8)
    assert not np.array_equal(got_t, exp)
    # (4) last-occurrence merging instead of first: the structural builder disagrees
    words = np.array([1, 0, 0, 0, 0, 0, 0, 0, 0, 2, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0], np.uint32)
    out, mapping = O.merge_level(words, 9, 0)
    assert mapping[18] == 0 and out[0] == 1  # first occurrence keeps slot 0


# --- generator ------------------------------------------------------------------------------

def test_generator_fixture_checksums():
    import hashlib
    import json
    import os
    path = os.path.join(os.path.dirname(__file__), "golden", "generator.json")
    fx = json.load(open(path))
    for e in fx["depth"]:
        D = O.gen_depth(e["N"], e["seed"], e["boxes"], e["noise"], e["x0"], e["y0"], e["tile"])
        assert hashlib.sha256(D.tobytes()).hexdigest() == e["sha256"], e
    for e in fx["uniform"]:
        q = O.gen_queries_uniform(e["seed"], e["count"], e["stride"])
        assert hashlib.sha256(q.tobytes()).hexdigest() == e["sha256"], e
