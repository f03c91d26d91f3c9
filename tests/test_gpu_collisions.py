"""Collision detection of the table merge.  With the key hashes cut to a few bits (the context's
hash-bits test hook), distinct subtrees share hashes and table slots.  The full-key compares of
each node with its slot's representative (k_resolve) must catch every such collision and send the
build down the sort path, whose words equal the oracle's.  Each case weakens one kind of level only, so the retry can only come from that
level's compares.  Needs a GPU."""
import numpy as np
import pytest

import cpvs_amd as cp
import oracle as O

pytestmark = pytest.mark.gpu
SEED0 = 0x43505653


def _say(*a):
    print("[case]", *a, flush=True)


@pytest.fixture(scope="module")
def maps():
    """A generator map with noise (many distinct leaves: collisions across windows and within them)
    and a periodic map (few distinct keys, huge groups of equal ones)."""
    N = 1024
    gen = O.gen_depth(N, SEED0 + 77, 32, True)
    y, x = np.mgrid[0:N, 0:N]
    periodic = (0.3 + 0.4 * (((x // 8 + y // 8) % 2) == 0) + 0.001 * (x % 8) + 0.0003 * (y % 16)).astype(np.float32)
    out = []
    for name, D, t, T in (("gen1024", gen, 0, 1), ("gen1024_t3of8", gen, 3, 8), ("periodic", periodic, 0, 1)):
        words, n = O.create(O.Pyramid(D), t, T, merge_mode=1)
        out.append((name, D, t, T, words, n))
    return out


# (leaf bits, inner bits): 2 bits put distinct leaves of one insert window in one hash group (the
# workgroup's LDS merge); 16 bits mostly make distinct keys of different windows meet in the
# global table.
CASES = [(2, 0), (16, 0), (0, 2), (0, 16)]


@pytest.mark.parametrize("bits", CASES, ids=lambda b: f"leaf{b[0]}-inner{b[1]}")
def test_plain_builds_detect_collisions(ctx, maps, bits):
    c = cp.Context(0)
    try:
        for name, D, t, T, words, n in maps:
            _say("collisions", bits, name)
            cs = cp.CompressedShadow.create(cp.MinMaxHierarchy(D, ctx=c), t, T)
            assert cs.merge_info()[0] == "table", name  # whole hashes: no collision
            cs.close()
            c.set_hash_bits(*bits)
            try:
                cs = cp.CompressedShadow.create(cp.MinMaxHierarchy(D, ctx=c), t, T)
            finally:
                c.set_hash_bits(0, 0)
            assert cs.getNumLevels() == n, name
            assert np.array_equal(cs.getDAG(), words), name
            if name == "gen1024":  # distinct keys certainly share cut hashes on its table levels
                assert cs.merge_info()[0] == "sort", name  # the compares caught them
            cs.close()
    finally:
        c.close()


@pytest.mark.parametrize("bits", [(2, 0), (16, 0), (0, 2)], ids=lambda b: f"leaf{b[0]}-inner{b[1]}")
def test_forest_detects_collisions(maps, bits):
    """The same maps as one forest (one launch chain): a collision anywhere sends the forest to
    the sort path, every DAG's words equal to the oracle's."""
    c = cp.Context(0)
    try:
        c.set_hash_bits(*bits)
        dd = [cp.DeviceArray.from_numpy(D, ctx=c) for _, D, *_ in maps]
        _say("forest collisions", bits)
        shadows, _ = cp.build_batch(dd, [(t, T) for _, _, t, T, _, _ in maps], ctx=c,
                                    config=cp.batch_config(lanes=1, forest_mtexels=4096))
        try:
            for (name, D, t, T, words, n), sh in zip(maps, shadows):
                assert sh.getNumLevels() == n, name
                assert np.array_equal(sh.getDAG(), words), name
            assert shadows[0].merge_info()[0] == "sort"
        finally:
            for sh in shadows:
                sh.close()
            for d in dd:
                d.free()
    finally:
        c.close()


def test_hash_bits_arguments(ctx):
    c = cp.Context(0)
    try:
        with pytest.raises(cp.CpvsError):
            c.set_hash_bits(32, 0)
        with pytest.raises(cp.CpvsError):
            c.set_hash_bits(0, 40)
        c.set_hash_bits(0, 0)
    finally:
        c.close()
