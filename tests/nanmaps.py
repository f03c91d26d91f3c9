"""Depth maps with scattered NaN (and infinite) texels, and a count of the octree cells they make
partial that no partial parent reaches.

The reference's pyramid pairs texels with std::min / std::max (src/MinMaxHierarchy.cpp:29-49), so
a NaN survives into a level only from a block's first texel; such a texel's cell is partial at
every z (a NaN compare fails both tests of visible(), src/CompressedShadowUtil.h:35-43) while its
parent's cell may be lit or shadowed there.  The top-down build never visits those cells; counts
taken level by level, without the ancestors, include them.  excess() measures that difference.
"""
from __future__ import annotations

import numpy as np

import oracle as O


def sparse_nan_map(N: int, seed: int, nan_frac: float = 0.002, inf_frac: float = 0.001) -> np.ndarray:
    """A smooth generator map with a few NaN, +inf and -inf texels."""
    D = O.gen_depth(N, seed, max(1, 32 * N * N >> 20), True)
    rng = np.random.default_rng(seed)
    D[rng.random(D.shape) < nan_frac] = np.nan
    D[rng.random(D.shape) < inf_frac] = np.inf
    D[rng.random(D.shape) < inf_frac] = -np.inf
    return D


def _partial(level: np.ndarray, h: np.float32, zlo: int, zhi: int) -> np.ndarray:
    """(w, w, zhi-zlo) booleans: cell z of each texel partial (visible() fails both tests)."""
    z = np.arange(zlo, zhi, dtype=np.float32)
    mn = level[..., 0][..., None] * h
    mx = level[..., 1][..., None] * h
    return ~((z + np.float32(1)) <= mn) & ~(z >= mx)


def excess(D: np.ndarray, t: int = 0, T: int = 1, leafmasks: bool = True) -> int:
    """Partial cells (levels 3.. or 1.. up to the root's children) with a non-partial ancestor."""
    N = D.shape[0]
    n = N.bit_length()
    R = n - 2
    Lmin = 3 if (leafmasks and n >= 5) else 1
    pyr = O.Pyramid(D)
    reach = None
    total = 0
    for L in range(R, Lmin - 1, -1):
        zlo, zhi = t * (N >> L), (t + 1) * (N >> L)
        part = _partial(pyr.level(L), np.float32((N >> L) * T), zlo, zhi)
        if reach is None:
            ok = part
        else:  # the parent cell (x/2, y/2, z/2) reachable and partial
            par = np.repeat(np.repeat(np.repeat(reach, 2, 0), 2, 1), 2, 2)
            ok = part & par
        total += int(part.sum() - ok.sum())
        reach = ok
    return total
