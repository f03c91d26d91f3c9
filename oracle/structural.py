"""Independent structural restatement of the final DAG layout -- TEST INFRASTRUCTURE ONLY.

The C oracle follows the reference's own procedure (uncompressed 9/17-word SVO, then
mergeLevel, updateParentPointers, removeUnusedNodes, compress; src/CompressedShadow.cpp:87-392).
This module derives the same words from a different definition, for small maps only:

* two nodes merge iff their subtrees are structurally equal (equal nested tuples), which is
  what word equality after bottom-up pointer rewriting amounts to
  (src/CompressedShadowUtil.h:137-182, src/CompressedShadow.cpp:215-259);
* a level lists its distinct subtrees in order of first occurrence in the breadth-first
  order of the unmerged octree (src/CompressedShadow.cpp:137-158, src/CompressedShadowUtil.h:164-178);
* words are written level by level from the root, node = [mask, absolute child offsets]
  or, at the leafmask level, [mask, lo, hi, ...] (src/CompressedShadow.cpp:326-392).

The pyramid comes from numpy (exact min/max), not from the C oracle: each level pairs the four
texels of a block as pred(pred(a, b), pred(c, d)) with std::min / std::max, a and b the upper row
(find/setRow, src/MinMaxHierarchy.cpp:29-49), so a NaN survives only from the first operand.
"""
from __future__ import annotations

import numpy as np


def _smin(a, b):
    return np.where(b < a, b, a)  # std::min(a, b)


def _smax(a, b):
    return np.where(a < b, b, a)  # std::max(a, b)


def _pyramid(D: np.ndarray):
    N = D.shape[0]
    n = N.bit_length()  # numLevels = log2(N) + 1
    mins, maxs = [D], [D]
    for L in range(1, n):
        m, M = mins[-1], maxs[-1]
        mins.append(_smin(_smin(m[0::2, 0::2], m[0::2, 1::2]), _smin(m[1::2, 0::2], m[1::2, 1::2])))
        maxs.append(_smax(_smax(M[0::2, 0::2], M[0::2, 1::2]), _smax(M[1::2, 0::2], M[1::2, 1::2])))
    return n, mins, maxs


def build(D: np.ndarray, zTileIndex: int = 0, zTileNum: int = 1, leafmasks: bool = True):
    """leafmasks=False: the reference compiled without LEAFMASKS (src/CompressedShadow.cpp:17)."""
    D = np.ascontiguousarray(D, np.float32)
    N = D.shape[0]
    n, mins, maxs = _pyramid(D)
    leaf_mode = leafmasks and n - 3 >= 2
    bottom = 2 if leaf_mode else 0
    f32 = np.float32

    def classify(L, x, y, z):
        h = f32((N >> L) * zTileNum)
        mask, kids = 0, []
        for c in range(8):
            cx, cy, cz = x + (c & 1), y + ((c >> 1) & 1), z + ((c >> 2) & 1)
            mn, mx = mins[L][cy, cx], maxs[L][cy, cx]
            if L > 0:
                if f32(cz + 1) <= mn * h:
                    bits = 1
                elif f32(cz) >= mx * h:
                    bits = 0
                else:
                    bits = 2
            else:
                bits = 1 if (f32(cz) + f32(cz + 1)) * f32(0.5) <= mn * h else 0
            mask |= bits << (2 * c)
            if bits == 2:
                kids.append((2 * cx, 2 * cy, 2 * cz))
        return mask, kids

    def leaf(x, y, z):
        h = f32(N * zTileNum)
        bx, by, bz = 4 * x, 4 * y, 4 * z
        blockd = D[by:by + 8, bx:bx + 8] * h
        mask, payload = 0, []
        for k in range(8):
            zz = bz + k
            mid = (f32(zz) + f32(zz + 1)) * f32(0.5)
            bits = (mid <= blockd).reshape(-1)
            m = int(sum(int(b) << i for i, b in enumerate(bits)))
            if m == (1 << 64) - 1:
                mask |= 1 << (2 * k)
            elif m != 0:
                mask |= 2 << (2 * k)
                payload.append(m)
        return mask, tuple(payload)

    # Breadth-first unmerged octree, each node carrying its structural key.
    root_level = n - 2
    levels = {root_level: [(0, 0, 2 * zTileIndex)]}
    info = {}
    L = root_level
    while L >= bottom and levels.get(L):
        nxt = []
        for (x, y, z) in levels[L]:
            if leaf_mode and L == 2:
                info[(L, x, y, z)] = leaf(x, y, z)
            else:
                m, kids = classify(L, x, y, z)
                info[(L, x, y, z)] = (m, kids)
                nxt.extend(kids)
        if not (leaf_mode and L == 2) and L - 1 >= bottom:
            levels[L - 1] = nxt
        L -= 1

    key_cache = {}

    def key(L, c):
        kk = (L,) + c
        if kk in key_cache:
            return key_cache[kk]
        if leaf_mode and L == 2:
            k = ("leaf",) + info[kk]
        else:
            m, kids = info[kk]
            k = (m, tuple(key(L - 1, ch) for ch in kids))
        key_cache[kk] = k
        return k

    # Distinct subtrees per level in first-occurrence order.
    order = []
    for L in range(root_level, bottom - 1, -1):
        nodes = levels.get(L, [])
        if not nodes:
            break
        seen, uniq = {}, []
        for c in nodes:
            k = key(L, c)
            if k not in seen:
                seen[k] = len(uniq)
                uniq.append(c)
        order.append((L, uniq, seen))

    # Sizes and offsets, then words (keyed by level too: a node without partial children has the
    # same key as an equal-mask node of another level).
    offsets = {}
    pos = 0
    for (L, uniq, seen) in order:
        for c in uniq:
            offsets[(L, key(L, c))] = pos
            if leaf_mode and L == 2:
                pos += 1 + 2 * len(info[(L,) + c][1])
            else:
                pos += 1 + len(info[(L,) + c][1])
    words = np.zeros(pos, np.uint32)
    for (L, uniq, seen) in order:
        for c in uniq:
            o = offsets[(L, key(L, c))]
            if leaf_mode and L == 2:
                m, payload = info[(L,) + c]
                words[o] = m
                for j, p in enumerate(payload):
                    words[o + 1 + 2 * j] = p & 0xFFFFFFFF
                    words[o + 2 + 2 * j] = p >> 32
            else:
                m, kids = info[(L,) + c]
                words[o] = m
                for j, ch in enumerate(kids):
                    words[o + 1 + j] = offsets[(L - 1, key(L - 1, ch))]
    return words, n
