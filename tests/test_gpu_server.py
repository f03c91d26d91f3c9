"""The serving loop (cpvs_server): results of builds and lookups are the same served alone, in the
mix of a concurrent server, and through the plain entry points; admission keeps the HBM
budget by evicting least recently used DAGs; continuous batching merges lookups.  Needs a GPU."""
import numpy as np
import pytest

import cpvs_amd as cp
import oracle as O

pytestmark = pytest.mark.gpu
SEED0 = 0x43505653
NQ = 120_000
CHUNKS = [1, 255, 256, 1000, 30_000, 88_488]  # one lookup request per chunk (sums to NQ)


def _say(*a):
    print("[case]", *a, flush=True)


@pytest.fixture(scope="module")
def work(ctx):
    specs = [(64, 0, 1), (128, 0, 1), (256, 1, 2), (256, 0, 1), (512, 0, 1), (1024, 0, 1), (2048, 0, 1), (32, 0, 1)]
    maps = [(O.gen_depth(N, SEED0 + 500 + i, max(1, 32 * N * N >> 20), True), t, T) for i, (N, t, T) in enumerate(specs)]
    depths = [cp.DeviceArray.from_numpy(D) for D, _, _ in maps]
    qh = O.gen_queries_uniform(SEED0 + 41, NQ, 3)
    q = cp.DeviceArray.from_numpy(qh)
    plain = []
    for (D, t, T) in maps:
        cs = cp.CompressedShadow.create(cp.MinMaxHierarchy(D), t, T)
        plain.append((cs.getDAG(), cs.evaluate(q).numpy()))
        cs.close()
    return maps, depths, qh, q, plain


def _lookups(srv, dag, q, vis):
    tickets, off = [], 0
    for c in CHUNKS:
        tickets.append(srv.lookup(dag, q, vis, count=c, offset=off, vis_offset=off))
        off += c
    assert off == NQ
    return tickets


def _serve(srv, depths, maps, q, interleave):
    """Builds every map and looks up all queries on each DAG.  interleave: submit all builds
    first and each DAG's lookups as soon as its build is done, while the other builds run."""
    vis = [cp.DeviceArray((NQ,), np.uint8) for _ in depths]
    lat = []
    if interleave:
        bt = [srv.build(d, t, T) for d, (_, t, T) in zip(depths, maps)]
        lk = []
        for i, b in enumerate(bt):
            lat.append(srv.wait(b))
            lk += _lookups(srv, b, q, vis[i])
        lat += [srv.wait(t) for t in lk]
    else:
        bt = []
        for i, (d, (_, t, T)) in enumerate(zip(depths, maps)):
            b = srv.build(d, t, T)
            lat.append(srv.wait(b))
            bt.append(b)
            lat += [srv.wait(x) for x in _lookups(srv, b, q, vis[i])]
    words = [srv.dag_words(b) for b in bt]
    return words, [v.numpy() for v in vis], lat


@pytest.mark.parametrize("lanes", [2, 4])
def test_results_alone_and_in_the_mix(ctx, work, lanes):
    maps, depths, qh, q, plain = work
    _say("server alone vs mix, lanes", lanes)
    alone = cp.Server(serial=True)
    try:
        wa, va, la = _serve(alone, depths, maps, q, interleave=False)
    finally:
        alone.close()
    mixed = cp.Server(lanes=lanes)
    try:
        wm, vm, lm = _serve(mixed, depths, maps, q, interleave=True)
        st = mixed.stats()
    finally:
        mixed.close()
    for i, ((w0, v0), w1, v1, w2, v2) in enumerate(zip(plain, wa, va, wm, vm)):
        assert np.array_equal(w1, w0) and np.array_equal(w2, w0), i
        assert np.array_equal(v1, v0) and np.array_equal(v2, v0), i
    assert st["builds"] == len(maps) and st["lookups"] == len(maps) * len(CHUNKS)
    assert all(x > 0 for x in la + lm)
    # one map against the oracle: lookups equal the brute-force depth compare
    D, t, T = maps[4]
    R = D.shape[0]
    v = (qh[:, :3] + np.float32(1)) * np.float32(0.5) * np.float32(R - 1)
    p = np.minimum(np.floor(np.where(v > 0, v, 0)), R - 1).astype(np.int64)
    assert np.array_equal(vm[4], O.brute_visible(D, 1, p[:, 0], p[:, 1], p[:, 2]))


def test_lookups_are_batched(ctx, work):
    maps, depths, qh, q, plain = work
    _say("server batching")
    srv = cp.Server(lanes=2)
    try:
        b = srv.build(depths[3])
        srv.wait(b)
        vis = cp.DeviceArray((NQ,), np.uint8)
        n = 400
        per = NQ // n
        ts = [srv.lookup(b, q, vis, count=per, offset=i * per, vis_offset=i * per) for i in range(n)]
        for t in ts:
            srv.wait(t)
        st = srv.stats()
        assert st["lookups"] == n and st["batches"] < n, st
        assert np.array_equal(vis.numpy()[:n * per], plain[3][1][:n * per])
    finally:
        srv.close()


def test_budget_evicts_least_recently_used(ctx, work):
    maps, depths, qh, q, plain = work
    _say("server eviction")
    mm = cp.MinMaxHierarchy(maps[4][0])  # 512^2
    _, _, dag_max = mm.estimate()
    mm.close()
    probe = cp.Server(serial=True)
    try:
        probe.wait(probe.build(depths[4]))
        ps = probe.stats()
    finally:
        probe.close()
    # lane scratch, one build's reservation (upper bound of its DAG) and 2.5 resident DAGs
    budget = ps["scratch_bytes"] + dag_max + int(2.5 * ps["resident_bytes"])
    srv = cp.Server(serial=True, budget_bytes=budget)
    try:
        tickets = []
        for k in range(8):
            tickets.append(srv.build(depths[4]))
            srv.wait(tickets[-1])
        st = srv.stats()
        assert st["evicted"] > 0, st
        assert st["resident_bytes"] + st["scratch_bytes"] <= budget
        res = srv.resident()
        assert res and res[-1] == tickets[-1]  # the newest stays
        evicted = [t for t in tickets if t not in res]
        assert evicted and evicted[0] == tickets[0]  # the oldest went first
        with pytest.raises(cp.CpvsError) as e:
            srv.lookup(evicted[0], q, cp.DeviceArray((NQ,), np.uint8))
        assert e.value.code == 5  # CPVS_ESTATE
        vis = cp.DeviceArray((NQ,), np.uint8)
        srv.wait(srv.lookup(res[-1], q, vis))
        assert np.array_equal(vis.numpy(), plain[4][1])
        srv.release(res[-1])
        assert res[-1] not in srv.resident()
    finally:
        srv.close()


def test_build_that_cannot_fit_fails(ctx, work):
    maps, depths, qh, q, plain = work
    _say("server budget error")
    srv = cp.Server(lanes=1, budget_bytes=1 << 20)
    try:
        t = srv.build(depths[6])  # 2048^2
        with pytest.raises(cp.CpvsError) as e:
            srv.wait(t)
        assert e.value.code == 6  # CPVS_EBUDGET
        assert srv.resident() == []
    finally:
        srv.close()


def test_budget_capped_at_free_memory(ctx):
    """A budget beyond the device's memory is lowered to what is free at creation, less 1 GiB."""
    _say("server budget cap")
    srv = cp.Server(lanes=1, budget_bytes=1 << 50)
    try:
        b = srv.stats()["budget"]
        assert 0 < b < (1 << 40)  # no device holds 1 TiB
    finally:
        srv.close()
    small = cp.Server(lanes=1, budget_bytes=1 << 30)
    try:
        assert small.stats()["budget"] == 1 << 30
    finally:
        small.close()


def test_bad_requests_are_rejected(ctx, work):
    maps, depths, qh, q, plain = work
    srv = cp.Server(lanes=1)
    try:
        with pytest.raises(cp.CpvsError):
            srv.lookup(12345, q, cp.DeviceArray((NQ,), np.uint8))
        bad = cp.DeviceArray.from_numpy(np.zeros((8, 8), np.float32))
        with pytest.raises(cp.CpvsError):
            srv.build(bad)  # N = 8: no leafmask levels
        with pytest.raises(cp.CpvsError):
            srv.build(depths[0], 2, 2)
        with pytest.raises(cp.CpvsError):
            srv.wait(999999)
    finally:
        srv.close()
