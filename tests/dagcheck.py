"""Structural checks of a final DAG word array (vectorised; usable at full size).

The reference's compressed layout (src/CompressedShadow.cpp:326-392) is fixed by four facts
that can be checked on the words alone:
  * levels are written root first, each level's nodes contiguous: node = [mask, k pointers]
    (or [mask, lo, hi, ...] at the leafmask level), k = partial children of mask;
  * every pointer of level L lands on a node start of level L-1 and every node of level L-1 is
    pointed to;
  * scanning level L's pointers in order, level L-1's nodes appear for the first time in
    layout order (first-occurrence order of mergeLevel, src/CompressedShadowUtil.h:164-178);
  * no two nodes of a level are equal (all common subtrees merged).
"""
from __future__ import annotations

import numpy as np

_POP16 = np.array([bin(i).count("1") for i in range(1 << 16)], np.uint32)


def popcount16(a):
    return _POP16[np.asarray(a, np.uint32) & 0xFFFF]


class DagError(AssertionError):
    pass


def check_dag(words: np.ndarray, num_levels: int, leafmasks: bool = True):
    """Raises DagError on a violation; returns per-level node counts (root first).  leafmasks=False:
    a DAG built without leafmasks (inner nodes down to level 0)."""
    w = np.asarray(words, np.uint32)
    n = num_levels
    leaf = leafmasks and (n - 3) >= 2
    bottom = 2 if leaf else 0
    R = n - 2
    if w.size == 0:
        raise DagError("empty DAG")
    starts = np.array([0], np.int64)
    counts = []
    L = R
    while True:
        if starts.size == 0:
            break
        if np.any(starts >= w.size):
            raise DagError(f"level {L}: node start beyond the array")
        masks = w[starts]
        if np.any(masks >> 16):
            raise DagError(f"level {L}: mask wider than 16 bits")
        if np.any(masks & (masks >> 1) & 0x5555):
            raise DagError(f"level {L}: a child has both the visible and partial bit")
        k = popcount16(masks & 0xAAAA).astype(np.int64)
        is_leaf_level = leaf and L == 2
        size = 1 + (2 * k if is_leaf_level else k)
        ends = starts + size
        if np.any(starts[1:] != ends[:-1]):
            raise DagError(f"level {L}: nodes are not contiguous")
        level_end = int(ends[-1])
        if level_end > w.size:
            raise DagError(f"level {L}: runs past the end")
        counts.append(int(starts.size))
        # Uniqueness of nodes inside the level.
        width = int(size.max())
        rows = np.full((starts.size, width), 0xFFFFFFFF, np.uint32)
        for j in range(width):
            sel = size > j
            rows[sel, j] = w[starts[sel] + j]
        if np.unique(rows, axis=0).shape[0] != starts.size:
            raise DagError(f"level {L}: two equal nodes (unmerged subtree)")
        if is_leaf_level:
            if L != bottom:
                raise DagError("leaf level misplaced")
            for j in range(int(k.max()) if k.size else 0):
                sel = k > j
                lo = w[starts[sel] + 1 + 2 * j].astype(np.uint64)
                hi = w[starts[sel] + 2 + 2 * j].astype(np.uint64)
                m = lo | (hi << np.uint64(32))
                if np.any(m == 0) or np.any(m == np.uint64(0xFFFFFFFFFFFFFFFF)):
                    raise DagError("leaf level: a partial slice stores an all-0 or all-1 leafmask")
            if level_end != w.size:
                raise DagError(f"{w.size - level_end} words after the leaf level")
            break
        if k.sum() == 0:
            if level_end != w.size:
                raise DagError(f"{w.size - level_end} trailing words after level {L}")
            break
        if L == bottom:
            raise DagError("partial children below the lowest level")
        # Pointers in (parent, slot) order.
        ptr_pos = np.repeat(starts + 1, k) + (np.arange(k.sum()) - np.repeat(np.cumsum(k) - k, k))
        ptrs = w[ptr_pos].astype(np.int64)
        uniq_sorted, first_idx = np.unique(ptrs, return_index=True)
        order = ptrs[np.sort(first_idx)]
        if not np.array_equal(order, uniq_sorted):
            raise DagError(f"level {L - 1}: nodes not in first-occurrence order")
        if uniq_sorted[0] != level_end:
            raise DagError(f"level {L - 1}: does not start right after level {L}")
        starts = uniq_sorted
        L -= 1
    return counts
