"""DAG byte format round trips through the device, and PCF lookups against the brute-force depth
compare over the filter taps.  Needs a GPU."""
import numpy as np
import pytest

import cpvs_amd as cp
import oracle as O
from tests.test_format import write

pytestmark = pytest.mark.gpu
SEED0 = 0x43505653


def _say(*a):
    print("[case]", *a, flush=True)


@pytest.mark.parametrize("N,t,T", [(8, 0, 1), (64, 0, 1), (256, 1, 2), (1024, 0, 1)])
def test_serialize_round_trip(ctx, N, t, T, tmp_path):
    _say("format round trip", N, t, T)
    D = O.gen_depth(N, SEED0 + 21 * N, max(1, 32 * N * N >> 20), N >= 64)
    cs = cp.CompressedShadow.create(cp.MinMaxHierarchy(D), t, T)
    words, n = cs.getDAG(), cs.getNumLevels()
    data = cs.serialize()
    assert data == write(words, n, t, T)  # the documented layout, byte for byte
    back = cp.CompressedShadow.deserialize(data)
    assert back.getNumLevels() == n and np.array_equal(back.getDAG(), words)
    assert back.getTotalVisibility() == cs.getTotalVisibility()
    path = str(tmp_path / "dag.cpvs")
    cs.save(path)
    loaded = cp.CompressedShadow.load(path)
    assert np.array_equal(loaded.getDAG(), words)
    if n >= 5:
        q = cp.DeviceArray.from_numpy(O.gen_queries_uniform(SEED0 + 5, 50_000, 3))
        assert np.array_equal(loaded.evaluate(q).numpy(), cs.evaluate(q).numpy())
    pts = O.gen_queries_uniform(SEED0 + 6, 2000, 3)
    assert np.array_equal(loaded.traverse_many(pts, n >= 5), cs.traverse_many(pts, n >= 5))


def test_damaged_buffers_never_reach_the_device(ctx):
    D = O.gen_depth(64, SEED0 + 1, 4, True)
    data = bytearray(cp.CompressedShadow.create(cp.MinMaxHierarchy(D)).serialize())
    data[64 + 8] ^= 0x40  # a word changes: the checksum no longer matches
    with pytest.raises(cp.CpvsError) as e:
        cp.CompressedShadow.deserialize(bytes(data))
    assert "checksum" in str(e.value)
    with pytest.raises(cp.CpvsError):
        cp.CompressedShadow.load("/nonexistent/dag.cpvs")


def _paths(q, R):
    v = (q[:, :3].astype(np.float32) + np.float32(1.0)) * np.float32(0.5)
    v = v * np.float32(R - 1)
    return np.minimum(np.floor(np.where(v > 0, v, 0)), R - 1).astype(np.int64)


def _pcf_brute(D, p, f, R):
    """Fraction of lit taps: float32(lit) / float32(f*f) over (px+dx, py+dy), dx, dy in
    [-(f-1)//2, f//2], clamped to [0, R-1]; lit by the midpoint test on the depth map."""
    lo = -((f - 1) // 2)
    lit = np.zeros(len(p), np.int64)
    for a in range(f):
        for b in range(f):
            x = np.clip(p[:, 0] + lo + b, 0, R - 1)
            y = np.clip(p[:, 1] + lo + a, 0, R - 1)
            lit += O.brute_visible(D, 1, x, y, p[:, 2])
    return lit.astype(np.float32) / np.float32(f * f)


@pytest.mark.parametrize("f", [1, 2, 3, 5, 8])
def test_pcf_matches_brute_force(ctx, f):
    _say("pcf", f)
    N = 512
    D = O.gen_depth(N, SEED0 + 99, 8, True)
    c = cp.CompressedShadowContainer(cp.CompressedShadow.create(cp.MinMaxHierarchy(D)))
    c.setFilterSize(f)
    c.moveToGPU()
    # points near the surface, where the taps disagree
    rng = np.random.default_rng(f)
    px = rng.integers(0, N, 60_000)
    py = rng.integers(0, N, 60_000)
    z = D[py, px] + rng.uniform(-4, 4, 60_000).astype(np.float32) / N
    q = np.stack([(px + 0.5) / N * 2 - 1, (py + 0.5) / N * 2 - 1, np.clip(z, 0, 1) * 2 - 1], 1).astype(np.float32)
    dq = cp.DeviceArray.from_numpy(q)
    got = c.evaluate_pcf(dq).numpy()
    p = _paths(q, N)
    exp = _pcf_brute(D, p, f, N)
    assert np.array_equal(got, exp)
    assert 0.01 < np.mean((got > 0) & (got < 1)) or f == 1  # filtering actually mixes taps
    if f == 1:
        assert np.array_equal(got, c.evaluate(dq).numpy().astype(np.float32))


def test_pcf_across_tiles(ctx):
    """Taps cross tile boundaries of a 2x2x2 tiled container (the walk takes any path)."""
    _say("pcf tiled")
    L, tile = 2, 256
    full = L * tile
    D = O.gen_depth(full, SEED0 + 123, 16, True)
    c = cp.CompressedShadowContainer(L)
    for ty in range(L):
        for tx in range(L):
            mm = cp.MinMaxHierarchy(np.ascontiguousarray(D[ty * tile:(ty + 1) * tile, tx * tile:(tx + 1) * tile]))
            for tz in range(L):
                c.set(cp.CompressedShadow.create(mm, tz, L), tx, ty, tz)
    c.setFilterSize(4)
    c.moveToGPU()
    rng = np.random.default_rng(5)
    px = np.concatenate([rng.integers(tile - 3, tile + 3, 5000), rng.integers(0, full, 20_000)])
    py = np.concatenate([rng.integers(0, full, 5000), rng.integers(0, full, 20_000)])
    z = D[py, px] + rng.uniform(-3, 3, len(px)).astype(np.float32) / full
    q = np.stack([(px + 0.5) / full * 2 - 1, (py + 0.5) / full * 2 - 1, np.clip(z, 0, 1) * 2 - 1], 1).astype(np.float32)
    got = c.evaluate_pcf(cp.DeviceArray.from_numpy(q)).numpy()
    p = _paths(q, full)
    assert np.array_equal(got, _pcf_brute(D, p, 4, full))


def test_pcf_rejects_bad_filter_size(ctx):
    D = O.gen_depth(64, SEED0 + 4, 2, True)
    c = cp.CompressedShadowContainer(cp.CompressedShadow.create(cp.MinMaxHierarchy(D)))
    c.moveToGPU()
    q = cp.DeviceArray.from_numpy(O.gen_queries_uniform(SEED0, 100, 3))
    for f in (0, 65):
        c.setFilterSize(f)
        with pytest.raises(cp.CpvsError):
            c.evaluate_pcf(q)
