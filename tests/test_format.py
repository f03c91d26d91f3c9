"""The DAG byte format, checked on the CPU through cpvs_dag_check: buffers written here from the
documented layout (64-byte header, checksum, word array) around oracle-built DAGs are accepted
with the right level counts, and each kind of damage is rejected.  No GPU needed."""
import struct

import numpy as np
import pytest

import cpvs_amd as cp
import oracle as O
from tests.dagcheck import check_dag

SEED0 = 0x43505653
M64 = (1 << 64) - 1


def checksum(words):
    """h = 0x243F6A8885A308D3; per little-endian pair of words v: h = (h ^ v) * K; h ^= h >> 29;
    finally h ^= W (the last word alone when W is odd)."""
    w = [int(x) for x in np.asarray(words, np.uint32)]
    h = 0x243F6A8885A308D3
    for i in range(0, len(w) - 1, 2):
        v = w[i] | (w[i + 1] << 32)
        h = ((h ^ v) * 0x9E3779B97F4A7C15) & M64
        h ^= h >> 29
    if len(w) % 2:
        h = ((h ^ w[-1]) * 0x9E3779B97F4A7C15) & M64
        h ^= h >> 29
    return h ^ len(w)


def write(words, num_levels, t=0, T=1, *, magic=b"CPVSDAG\0", version=1, header=64, N=None, root=None, ck=None,
          flags=0):
    words = np.asarray(words, np.uint32)
    N = (1 << (num_levels - 1)) if N is None else N
    root = int(words[0]) if root is None else root
    ck = checksum(words) if ck is None else ck
    h = struct.pack("<8sIIIIIIIIQQQ", magic, version, header, num_levels, N, t, T, root, flags, len(words), ck, 0)
    assert len(h) == 64
    return h + words.astype("<u4").tobytes()


def _maps():
    out = []
    for N, t, T in [(16, 0, 1), (32, 1, 2), (64, 0, 1), (128, 2, 4)]:
        D = O.gen_depth(N, SEED0 + 7 * N + t, max(1, N // 16), True)
        out.append((f"gen{N}_t{t}of{T}", D, t, T))
    out.append(("uniform0.625", np.full((64, 64), 0.625, np.float32), 0, 1))
    out.append(("nonleaf8", O.gen_depth(8, SEED0 + 3, 1, False), 0, 1))
    return out


@pytest.mark.parametrize("name,D,t,T", _maps(), ids=[m[0] for m in _maps()])
def test_oracle_dags_are_accepted(name, D, t, T):
    words, n = O.create(O.Pyramid(D), t, T, merge_mode=1)
    nl, nw, nodes = cp.dag_check(write(words, n, t, T))
    assert (nl, nw) == (n, len(words))
    expect = check_dag(words, n)  # per-level node counts, root first
    got = [int(nodes[L]) for L in range(n - 2, -1, -1) if nodes[L]]
    assert got == [int(c) for c in expect if c]


@pytest.fixture(scope="module")
def dag64():
    D = O.gen_depth(64, SEED0 + 11, 4, True)
    words, n = O.create(O.Pyramid(D), 0, 1, merge_mode=1)
    return np.asarray(words, np.uint32), n


def _rejects(buf, what):
    with pytest.raises(cp.CpvsError) as e:
        cp.dag_check(buf)
    assert e.value.code == 1 and what in str(e.value), str(e.value)


def test_header_damage_is_rejected(dag64):
    w, n = dag64
    _rejects(write(w, n, magic=b"CPVSDAX\0"), "bad magic")
    _rejects(write(w, n, version=2), "version 2")
    _rejects(write(w, n, header=60), "header of 60")
    _rejects(write(w, n, N=32), "does not match numLevels")
    _rejects(write(w, n, t=1, T=1), "z tile")
    _rejects(write(w, n)[:-4], "bytes for")
    _rejects(write(w, n) + b"\0\0\0\0", "bytes for")
    _rejects(write(w, n, root=int(w[0]) ^ 4), "root mask")
    bad = w.copy()
    bad[len(bad) // 2] ^= 1
    _rejects(write(bad, n, ck=checksum(w)), "checksum")


def test_structure_damage_is_rejected(dag64):
    w, n = dag64
    # a child pointer of the root moved into the middle of a node
    bad = w.copy()
    bad[1] += 1
    _rejects(write(bad, n), "not referenced")
    # a pointer backwards (to the root itself)
    bad = w.copy()
    bad[1] = 0
    _rejects(write(bad, n), "backwards")
    # trailing words after the last level
    _rejects(write(np.concatenate([w, [0, 0]]).astype(np.uint32), n), "after the last level")
    # a child state 3 (both bits) in the root mask
    bad = w.copy()
    bad[0] |= 3
    _rejects(write(bad, n), "invalid child state")
    # mask bits above 16
    bad = w.copy()
    bad[0] |= 1 << 20
    _rejects(write(bad, n), "bits above 16")
    # the words end inside the leaf level
    _rejects(write(w[:-1], n), "past the end")


@pytest.mark.parametrize("N,t,T", [(16, 0, 1), (64, 1, 2), (128, 0, 1)])
def test_dags_without_leafmasks_carry_their_flag(N, t, T):
    """Flag bit 0 marks a DAG built without leafmasks (nodes down to level 0): accepted with it,
    rejected without it (level 2 then reads as leafmask nodes), and the leafmask DAG of the same
    map the other way round."""
    D = O.gen_depth(N, SEED0 + 5 * N + t, max(1, N // 16), True)
    p = O.Pyramid(D)
    w0, n = O.create(p, t, T, leafmasks=False)
    nl, nw, nodes = cp.dag_check(write(w0, n, t, T, flags=1))
    assert (nl, nw) == (n, len(w0))
    expect = check_dag(w0, n, leafmasks=False)
    assert [int(nodes[L]) for L in range(n - 2, -1, -1) if nodes[L]] == [int(c) for c in expect if c]
    if T == 1:
        assert nodes[0] > 0 and nodes[1] > 0  # the octree reaches level 0
    if nodes[1] > 0:  # read as leafmask nodes, level 2's pointers become payload
        with pytest.raises(cp.CpvsError):
            cp.dag_check(write(w0, n, t, T))
    w1, _ = O.create(p, t, T)
    cp.dag_check(write(w1, n, t, T))
    # A leafmask DAG whose leaf nodes hold partial slices fails when read as plain nodes (its
    # leafmask words read as pointers); one without partial slices is the same either way.
    noise = np.random.default_rng(N + t).random((N, N), dtype=np.float32)
    w2, _ = O.create(O.Pyramid(noise), t, T)
    cp.dag_check(write(w2, n, t, T))
    with pytest.raises(cp.CpvsError):
        cp.dag_check(write(w2, n, t, T, flags=1))


def test_unknown_flags_are_rejected(dag64):
    w, n = dag64
    _rejects(write(w, n, flags=2), "unknown flags")
    _rejects(write(w, n, flags=0x80000000), "unknown flags")
