/*
 * cpvs_oracle.c -- CPU restatement of the TobiasRp/cpvs precomputed-voxelized-shadow
 * build (MinMaxHierarchy + CompressedShadow::create) and lookup (CompressedShadow::traverse,
 * shader/traverse.cs).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product path and the CPU baseline of
 * bench.py.  The product never links or calls it.  See cpvs_oracle.h for what pins each
 * function.  Built with -ffp-contract=off so float expressions round exactly as written.
 */
#include "cpvs_oracle.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>

typedef uint32_t u32;
typedef uint64_t u64;

#define NODE_SIZE 9u  /* CompressedShadowUtil.h:9  childmask + 8 pointers */
#define LEAF_SIZE 17u /* CompressedShadowUtil.h:10 childmask + 8 64-bit leafmasks */

/* ------------------------------------------------------------------------------------ */
/* Min/max pyramid -- MinMaxHierarchy.{h,cpp}                                             */
/* ------------------------------------------------------------------------------------ */

static int ilog2i(int v) {
    int k = 0;
    while ((1 << (k + 1)) <= v) ++k;
    return k;
}

/* MinMaxHierarchy.cpp:17 numLevels = ceil(log2(size)) levels above the root; for the
 * power-of-two sizes the reference accepts (cpvs.h:60-62 log2 via log()*log2(e)) this is
 * exactly k for N = 2^k, k = 3..20.  getNumLevels() = that + 1 (MinMaxHierarchy.h:60-62). */
int or_num_levels(int N) { return ilog2i(N) + 1; }

size_t or_pyramid_level_offset(int N, int L) {
    size_t off = 0;
    for (int l = 1; l < L; ++l) {
        size_t w = (size_t)(N >> l);
        off += w * w * 2;
    }
    return off;
}

size_t or_pyramid_floats(int N) { return or_pyramid_level_offset(N, or_num_levels(N)); }

/* std::min<float> / std::max<float> as used through find() (MinMaxHierarchy.cpp:29-33):
 * min(a,b) = (b < a) ? b : a, max(a,b) = (a < b) ? b : a. */
static inline float smin(float a, float b) { return (b < a) ? b : a; }
static inline float smax(float a, float b) { return (a < b) ? b : a; }

typedef struct {
    const float *in;
    float *out;
    int inW;
    int inC; /* channels of the input: 1 for the root, 2 otherwise */
    int row0, row1;
} pyr_job;

/* constructRange / setRow (MinMaxHierarchy.cpp:35-49): for every 2x2 block, out(min) is
 * min(min(a,b),min(c,d)) over the input min channel, out(max) likewise over the max channel. */
static void *pyr_range(void *arg) {
    const pyr_job *j = (const pyr_job *)arg;
    const int W = j->inW, C = j->inC, oW = W / 2;
    const int cmin = 0, cmax = (C == 1) ? 0 : 1; /* constructLevelFromRoot reads channel 0 twice */
    for (int y = j->row0; y < j->row1; y += 2) {
        for (int x = 0; x < W; x += 2) {
            const float *r0 = j->in + ((size_t)y * W + x) * C;
            const float *r1 = j->in + ((size_t)(y + 1) * W + x) * C;
            float mn = smin(smin(r0[cmin], r0[C + cmin]), smin(r1[cmin], r1[C + cmin]));
            float mx = smax(smax(r0[cmax], r0[C + cmax]), smax(r1[cmax], r1[C + cmax]));
            float *o = j->out + ((size_t)(y / 2) * oW + x / 2) * 2;
            o[0] = mn;
            o[1] = mx;
        }
    }
    return NULL;
}

/* constructLevel / constructLevelFromRoot (MinMaxHierarchy.cpp:51-97): rows are split over
 * 4 threads when the input width is >= PARALLEL_THRESHOLD (256, :7). */
static void pyr_level(const float *in, int inW, int inC, float *out) {
    if (inW >= 256) {
        pthread_t th[3];
        pyr_job jobs[4];
        int per = inW / 4;
        for (int t = 0; t < 4; ++t) {
            jobs[t].in = in;
            jobs[t].out = out;
            jobs[t].inW = inW;
            jobs[t].inC = inC;
            jobs[t].row0 = t * per;
            jobs[t].row1 = (t == 3) ? inW : (t + 1) * per;
        }
        for (int t = 0; t < 3; ++t) pthread_create(&th[t], NULL, pyr_range, &jobs[t]);
        pyr_range(&jobs[3]);
        for (int t = 0; t < 3; ++t) pthread_join(th[t], NULL);
    } else {
        pyr_job j = {in, out, inW, inC, 0, inW};
        pyr_range(&j);
    }
}

void or_pyramid_build(const float *D, int N, float *P) {
    const int n = or_num_levels(N);
    pyr_level(D, N, 1, P + or_pyramid_level_offset(N, 1));
    for (int L = 2; L < n; ++L)
        pyr_level(P + or_pyramid_level_offset(N, L - 1), N >> (L - 1), 2,
                  P + or_pyramid_level_offset(N, L));
}

float or_pyramid_min(const float *D, const float *P, int N, int L, int x, int y) {
    if (L == 0) return D[(size_t)y * N + x];
    const int w = N >> L;
    return P[or_pyramid_level_offset(N, L) + ((size_t)y * w + x) * 2];
}

float or_pyramid_max(const float *D, const float *P, int N, int L, int x, int y) {
    if (L == 0) return D[(size_t)y * N + x];
    const int w = N >> L;
    return P[or_pyramid_level_offset(N, L) + ((size_t)y * w + x) * 2 + 1];
}

/* ------------------------------------------------------------------------------------ */
/* Child masks -- CompressedShadowUtil.{h,cpp}                                            */
/* ------------------------------------------------------------------------------------ */

/* getLevelHeight (CompressedShadowUtil.cpp:16-18): width of the pyramid level times
 * depthOffset, where depthOffset = zTileNum (set by create(), CompressedShadow.cpp:53). */
static inline u32 level_height(int N, int L, u32 T) { return (u32)(N >> L) * T; }

/* visible() (CompressedShadowUtil.h:35-43): VISIBLE=1, SHADOW=0, PARTIAL=2. */
static inline u32 vis_range(float minZ, float maxZ, float minD, float maxD) {
    if (maxZ <= minD) return 1;
    if (minZ >= maxD) return 0;
    return 2;
}

/* absoluteVisible() (CompressedShadowUtil.h:51-57): midpoint compare, never partial. */
static inline u32 vis_abs(float minZ, float maxZ, float d) {
    const float mid = (minZ + maxZ) * 0.5f;
    return (mid <= d) ? 1u : 0u;
}

u32 or_create_childmask(const float *D, const float *P, int N, int level, int ox, int oy,
                        int oz, u32 T) {
    const u32 h = level_height(N, level, T);
    const float hf = (float)h;
    u32 mask = 0;
    for (u32 z = 0; z < 2; ++z)
        for (u32 y = 0; y < 2; ++y)
            for (u32 x = 0; x < 2; ++x) {
                const u32 cz = z + (u32)oz, cy = y + (u32)oy, cx = x + (u32)ox;
                const float mn = or_pyramid_min(D, P, N, level, (int)cx, (int)cy);
                const float mx = or_pyramid_max(D, P, N, level, (int)cx, (int)cy);
                u32 bits;
                if (level > 0)
                    bits = vis_range((float)cz, (float)(cz + 1), mn * hf, mx * hf);
                else
                    bits = vis_abs((float)cz, (float)(cz + 1), mn * hf);
                mask |= bits << ((x | (y << 1) | (z << 2)) * 2);
            }
    return mask;
}

/* createLeafmask (CompressedShadowUtil.cpp:59-78): bit x+8y of an 8x8 texel block at
 * depth slice z, using the level-0 midpoint test with height N*T. */
static u64 leafmask(const float *D, int N, int bx, int by, int bz, u32 T) {
    const float hf = (float)level_height(N, 0, T);
    u64 m = 0;
    for (int y = 0; y < 8; ++y)
        for (int x = 0; x < 8; ++x) {
            const float d = D[(size_t)(by + y) * N + (bx + x)];
            const u64 bit = vis_abs((float)bz, (float)(bz + 1), d * hf);
            m |= bit << (x + 8 * y);
        }
    return m;
}

/* createChildmask1x1x8 (CompressedShadowUtil.cpp:80-99). */
u32 or_create_leaf_node(const float *D, int N, int ox, int oy, int oz, u32 T, u64 *masks,
                        int *count) {
    const int bx = ox * 4, by = oy * 4, bz = oz * 4;
    u32 mask = 0;
    int k = 0;
    for (u32 z = 0; z < 8; ++z) {
        const u64 m = leafmask(D, N, bx, by, bz + (int)z, T);
        if (m == ~(u64)0)
            mask |= 1u << (z * 2);
        else if (m != 0) {
            mask |= 2u << (z * 2);
            masks[k++] = m;
        }
    }
    *count = k;
    return mask;
}

/* getChildCoordinates (CompressedShadowUtil.cpp:101-117). */
int or_child_coordinates(u32 childmask, int px, int py, int pz, int *out) {
    int k = 0;
    for (u32 i = 0; i < 8; ++i) {
        if (childmask & (1u << (i * 2 + 1))) {
            out[3 * k + 0] = (px + (int)(i & 1)) * 2;
            out[3 * k + 1] = (py + (int)((i >> 1) & 1)) * 2;
            out[3 * k + 2] = (pz + (int)((i >> 2) & 1)) * 2;
            ++k;
        }
    }
    return k;
}

static inline u32 num_children(u32 m) { return (u32)__builtin_popcount(m & 0xAAAAu); }

/* ------------------------------------------------------------------------------------ */
/* mergeLevel -- CompressedShadowUtil.h:137-182                                           */
/* ------------------------------------------------------------------------------------ */

static u64 hash_words(const u32 *w, u32 n) {
    u64 h = 0x243F6A8885A308D3ull;
    for (u32 i = 0; i < n; ++i) {
        h ^= w[i];
        h *= 0x100000001B3ull;
        h ^= h >> 29;
    }
    h ^= h >> 33;
    h *= 0xFF51AFD7ED558CCDull;
    h ^= h >> 33;
    return h;
}

u32 or_merge_level(const u32 *lw, size_t num_nodes, u32 ns, u32 *out, u32 *mapping, int mode) {
    u32 uniq = 0;
    if (mode == 0) {
        /* The reference's linear search over the uniques found so far (:164-178). */
        for (size_t i = 0; i < num_nodes; ++i) {
            const u32 *node = lw + i * ns;
            u32 u;
            for (u = 0; u < uniq; ++u)
                if (memcmp(node, out + (size_t)u * ns, ns * sizeof(u32)) == 0) break;
            mapping[i] = u * ns;
            if (u == uniq) {
                memcpy(out + (size_t)uniq * ns, node, ns * sizeof(u32));
                ++uniq;
            }
        }
        return uniq;
    }
    /* Same first-occurrence semantics through an open-addressing table of unique ids. */
    size_t cap = 16;
    while (cap < 2 * num_nodes + 1) cap <<= 1;
    u32 *table = (u32 *)calloc(cap, sizeof(u32)); /* stores unique id + 1 */
    if (!table) return 0xFFFFFFFFu;
    for (size_t i = 0; i < num_nodes; ++i) {
        const u32 *node = lw + i * ns;
        size_t slot = (size_t)hash_words(node, ns) & (cap - 1);
        for (;;) {
            const u32 e = table[slot];
            if (e == 0) {
                table[slot] = uniq + 1;
                memcpy(out + (size_t)uniq * ns, node, ns * sizeof(u32));
                mapping[i] = uniq * ns;
                ++uniq;
                break;
            }
            if (memcmp(node, out + (size_t)(e - 1) * ns, ns * sizeof(u32)) == 0) {
                mapping[i] = (e - 1) * ns;
                break;
            }
            slot = (slot + 1) & (cap - 1);
        }
    }
    free(table);
    return uniq;
}

/* ------------------------------------------------------------------------------------ */
/* CompressedShadow::create -- CompressedShadow.cpp:17-392                                 */
/* ------------------------------------------------------------------------------------ */

typedef struct {
    u32 *w;
    size_t n, cap;
} vu32;

static int vu_reserve(vu32 *v, size_t n) {
    if (n <= v->cap) return 0;
    size_t c = v->cap ? v->cap : 64;
    while (c < n) c *= 2;
    u32 *p = (u32 *)realloc(v->w, c * sizeof(u32));
    if (!p) return -1;
    v->w = p;
    v->cap = c;
    return 0;
}

static int vu_resize0(vu32 *v, size_t n) {
    if (vu_reserve(v, n)) return -1;
    if (n > v->n) memset(v->w + v->n, 0, (n - v->n) * sizeof(u32));
    v->n = n;
    return 0;
}

static int vu_push(vu32 *v, u32 x) {
    if (vu_reserve(v, v->n + 1)) return -1;
    v->w[v->n++] = x;
    return 0;
}

/* useLeafmasks / getMinLevel / getNodeSize (CompressedShadow.cpp:20-41).  The reference compiles
 * leafmasks in or out (#define LEAFMASKS, CompressedShadow.cpp:17, traverse.cs:8-9); g_leafmasks
 * is that switch here (on by default, or_set_leafmasks), for create() and the traverse.cs
 * transcription alike. */
static int g_leafmasks = 1;
void or_set_leafmasks(int on) { g_leafmasks = on != 0; }
static int use_leafmasks(int n) { return g_leafmasks && (n - 3) >= 2; }
static int min_level(int n) { return use_leafmasks(n) ? 2 : 0; }
static u32 node_size(int n, int level) { return (use_leafmasks(n) && level == 2) ? LEAF_SIZE : NODE_SIZE; }

#define MAXLV 32
static __thread u64 g_last_counts[MAXLV];

void or_last_node_counts(u64 *out, int n) {
    for (int i = 0; i < n && i < MAXLV; ++i) out[i] = g_last_counts[i];
}

typedef struct {
    const float *D, *P;
    int N, n;
    u32 T;
    vu32 dag;
    size_t lvl_start[MAXLV]; /* word offset of each level in the uncompressed SVO */
    size_t lvl_count[MAXLV]; /* nodes per level before merging (0 = level absent) */
    size_t lvl_uniq[MAXLV];  /* nodes per level after merging */
} svo_t;

/* constructSvo (CompressedShadow.cpp:87-169) + setChildrenOffsets (:78-85) +
 * constructLastLevels (:171-190).  Uncompressed layout: every inner node has 9 words
 * (mask + 8 absolute pointers, partial children first, unused 0), leaf nodes 17 words.
 * Level presence is tracked by node counts instead of levelOffsets differences, which is
 * the same whenever every level down to the leaves is non-empty (always for a sloped map)
 * and gives a defined result when the octree empties early. */
static int construct_svo(svo_t *s, int rootZ) {
    const int n = s->n;
    const u32 rootmask = or_create_childmask(s->D, s->P, s->N, n - 2, 0, 0, rootZ, s->T);
    u32 k = num_children(rootmask);
    int *coords = (int *)malloc(sizeof(int) * 3 * 8);
    if (!coords) return -1;
    size_t ncoords = (size_t)or_child_coordinates(rootmask, 0, 0, rootZ, coords);

    const u32 csz0 = node_size(n, n - 3);
    if (vu_resize0(&s->dag, NODE_SIZE + (size_t)k * csz0)) return -1;
    s->dag.w[0] = rootmask;
    for (u32 i = 0; i < k; ++i) s->dag.w[1 + i] = NODE_SIZE + i * csz0;
    s->lvl_start[n - 2] = 0;
    s->lvl_count[n - 2] = 1;

    size_t levelOffset = NODE_SIZE;
    size_t numLevelNodes = k;
    int level = n - 3;
    const int lastLevel = use_leafmasks(n) ? 3 : 0;
    u32 *masks = NULL;

    while (level >= lastLevel && numLevelNodes > 0) {
        s->lvl_start[level] = levelOffset;
        s->lvl_count[level] = numLevelNodes;
        const u32 childSize = node_size(n, level - 1);
        const size_t nextLevelOffset = levelOffset + numLevelNodes * NODE_SIZE;

        masks = (u32 *)realloc(masks, numLevelNodes * sizeof(u32));
        if (!masks) return -1;
        size_t newChildren = 0;
        for (size_t i = 0; i < numLevelNodes; ++i) {
            const u32 m = or_create_childmask(s->D, s->P, s->N, level, coords[3 * i],
                                              coords[3 * i + 1], coords[3 * i + 2], s->T);
            s->dag.w[levelOffset + i * NODE_SIZE] = m;
            masks[i] = m;
            newChildren += num_children(m);
        }
        if (level != 0 && vu_resize0(&s->dag, s->dag.n + newChildren * childSize)) return -1;

        int *ncoordsArr = (int *)malloc(sizeof(int) * 3 * (newChildren ? newChildren : 1));
        if (!ncoordsArr) return -1;
        size_t progress = 0, nc = 0;
        for (size_t i = 0; i < numLevelNodes; ++i) {
            const u32 m = masks[i];
            const u32 kc = num_children(m);
            if (kc == 0) continue;
            const size_t nodeOffset = levelOffset + i * NODE_SIZE;
            const size_t childOffset = nextLevelOffset + progress;
            for (u32 c = 0; c < kc; ++c) s->dag.w[nodeOffset + 1 + c] = (u32)(childOffset + c * childSize);
            nc += (size_t)or_child_coordinates(m, coords[3 * i], coords[3 * i + 1], coords[3 * i + 2],
                                               ncoordsArr + 3 * nc);
            progress += (size_t)kc * childSize;
        }
        levelOffset += numLevelNodes * NODE_SIZE;
        numLevelNodes = newChildren;
        free(coords);
        coords = ncoordsArr;
        ncoords = nc;
        --level;
    }
    (void)ncoords;

    if (use_leafmasks(n) && level == 2) {
        s->lvl_start[2] = levelOffset;
        s->lvl_count[2] = numLevelNodes;
        for (size_t i = 0; i < numLevelNodes; ++i) {
            const size_t off = levelOffset + i * LEAF_SIZE;
            u64 lm[8];
            int kc = 0;
            s->dag.w[off] = or_create_leaf_node(s->D, s->N, coords[3 * i], coords[3 * i + 1],
                                                coords[3 * i + 2], s->T, lm, &kc);
            for (int c = 0; c < kc; ++c) {
                s->dag.w[off + 1 + 2 * c] = (u32)lm[c];
                s->dag.w[off + 2 + 2 * c] = (u32)(lm[c] >> 32);
            }
        }
    }
    free(masks);
    free(coords);
    return 0;
}

/* mergeCommonSubtrees (CompressedShadow.cpp:215-241) + updateParentPointers (:243-259). */
static int merge_common_subtrees(svo_t *s, int mode) {
    const int n = s->n;
    for (int level = min_level(n); level < n - 2; ++level) {
        const size_t cnt = s->lvl_count[level];
        if (cnt == 0) continue; /* levelExists (:203-213) */
        const u32 ns = node_size(n, level);
        const size_t off = s->lvl_start[level];
        u32 *tmp = (u32 *)calloc(cnt * ns, sizeof(u32));
        u32 *mapping = (u32 *)malloc(cnt * sizeof(u32));
        if (!tmp || !mapping) return -1;
        const u32 uniq = or_merge_level(s->dag.w + off, cnt, ns, tmp, mapping, mode);
        if (uniq == 0xFFFFFFFFu) return -1;
        memcpy(s->dag.w + off, tmp, cnt * ns * sizeof(u32));
        s->lvl_uniq[level] = uniq;

        /* updateParentPointers: parents at level+1, pointers rebased to merged positions. */
        const int pl = level + 1;
        const size_t poff = s->lvl_start[pl];
        const size_t pcnt = s->lvl_count[pl];
        for (size_t p = 0; p < pcnt; ++p) {
            u32 *node = s->dag.w + poff + p * NODE_SIZE;
            for (u32 c = 1; c < NODE_SIZE; ++c) {
                const u32 old = node[c];
                if (old != 0) node[c] = (u32)(off + mapping[(old - off) / ns]);
            }
        }
        free(tmp);
        free(mapping);
    }
    s->lvl_uniq[n - 2] = 1;
    return 0;
}

/* removeUnusedNodes (CompressedShadow.cpp:271-304) + insertWithCorrectOffset (:262-269). */
static int remove_unused_nodes(svo_t *s) {
    const int n = s->n;
    vu32 nd = {0, 0, 0};
    if (vu_reserve(&nd, s->dag.n)) return -1;
    for (u32 i = 0; i < NODE_SIZE; ++i) vu_push(&nd, s->dag.w[i]);
    size_t levelOffset = NODE_SIZE;
    u32 correction = 0;
    const int start = min_level(n);
    for (int level = n - 3; level >= start; --level) {
        const u32 ns = node_size(n, level);
        const size_t levelSize = s->lvl_count[level] * ns;
        const size_t merged = s->lvl_uniq[level] * ns;
        correction += (u32)(levelSize - merged);
        if (level == start || correction == 0) {
            for (size_t i = levelOffset; i < levelOffset + merged; ++i) vu_push(&nd, s->dag.w[i]);
        } else {
            for (size_t i = levelOffset; i < levelOffset + merged; ++i) {
                u32 v = s->dag.w[i];
                if ((i % NODE_SIZE) != 0 && v != 0) v -= correction;
                vu_push(&nd, v);
            }
        }
        levelOffset += levelSize;
    }
    free(s->dag.w);
    s->dag = nd;
    return 0;
}

/* compress (CompressedShadow.cpp:326-392) + copyNodeInNewDag (:314-324): BFS relayout
 * keeping 1+k words per node (1+2k at the leafmask level), pointers remapped old->new. */
static int compress_dag(svo_t *s) {
    const int n = s->n;
    vu32 nd = {0, 0, 0};
    int level = n - 2;
    size_t numLevelNodes = 1, oldLevel = 0, oldLast = 0, newOff = 0, newLast = 0;
    const int minL = min_level(n);
    u32 *o2n = NULL;      /* old node index in this level -> new offset */
    u32 *prev_o2n = NULL; /* same for the previous (parent) level */
    size_t prevLevelStart = 0;
    (void)prevLevelStart;

    while (level >= minL && numLevelNodes > 0) {
        const size_t newLevelStart = newOff;
        const u32 ns = node_size(n, level);
        o2n = (u32 *)malloc(numLevelNodes * sizeof(u32));
        if (!o2n) return -1;
        for (size_t i = 0; i < numLevelNodes; ++i) {
            const size_t nodeOffset = oldLevel + i * ns;
            const u32 m = s->dag.w[nodeOffset];
            const u32 kc = num_children(m);
            o2n[i] = (u32)newOff;
            const u32 elems = (use_leafmasks(n) && level == minL) ? 1 + 2 * kc : 1 + kc;
            for (u32 e = 0; e < elems; ++e)
                if (vu_push(&nd, s->dag.w[nodeOffset + e])) return -1;
            newOff += elems;
        }
        /* Distinct child pointers of this level = node count of the next level. */
        size_t distinct = 0;
        if (level > minL) {
            const size_t childStart = oldLevel + numLevelNodes * NODE_SIZE;
            const u32 cns = node_size(n, level - 1);
            size_t maxIdx = 0;
            for (size_t i = 0; i < numLevelNodes; ++i) {
                const size_t nodeOffset = oldLevel + i * ns;
                const u32 kc = num_children(s->dag.w[nodeOffset]);
                for (u32 c = 0; c < kc; ++c) {
                    const size_t ci = (s->dag.w[nodeOffset + 1 + c] - childStart) / cns;
                    if (ci + 1 > maxIdx) maxIdx = ci + 1;
                }
            }
            unsigned char *seen = (unsigned char *)calloc(maxIdx ? maxIdx : 1, 1);
            if (!seen) return -1;
            for (size_t i = 0; i < numLevelNodes; ++i) {
                const size_t nodeOffset = oldLevel + i * ns;
                const u32 kc = num_children(s->dag.w[nodeOffset]);
                for (u32 c = 0; c < kc; ++c) {
                    const size_t ci = (s->dag.w[nodeOffset + 1 + c] - childStart) / cns;
                    if (!seen[ci]) {
                        seen[ci] = 1;
                        ++distinct;
                    }
                }
            }
            free(seen);
        }
        /* Parent pointers of the previous level now point at this level's new offsets. */
        if (oldLast != oldLevel) {
            size_t cur = newLast;
            for (size_t pos = oldLast; pos < oldLevel; pos += NODE_SIZE) {
                const u32 m = s->dag.w[pos];
                const u32 kc = num_children(m);
                for (u32 c = 0; c < kc; ++c) {
                    const u32 old = s->dag.w[pos + 1 + c];
                    if (old != 0) nd.w[cur + 1 + c] = o2n[(old - oldLevel) / ns];
                }
                cur += kc + 1;
            }
        }
        free(prev_o2n);
        prev_o2n = o2n;
        o2n = NULL;
        oldLast = oldLevel;
        oldLevel += numLevelNodes * NODE_SIZE;
        newLast = newLevelStart;
        numLevelNodes = distinct;
        --level;
    }
    free(prev_o2n);
    free(s->dag.w);
    s->dag = nd;
    return 0;
}

int or_create(const float *D, const float *P, int N, u32 zTileIndex, u32 zTileNum, int merge_mode,
              u32 **out_words, size_t *out_len, int *num_levels) {
    if (N < 8 || (N & (N - 1)) || zTileNum == 0 || zTileIndex >= zTileNum) return 1;
    svo_t s;
    memset(&s, 0, sizeof(s));
    s.D = D;
    s.P = P;
    s.N = N;
    s.n = or_num_levels(N);
    s.T = zTileNum;
    /* create(): root offset (0,0,2*zTileIndex) (CompressedShadow.cpp:54). */
    if (construct_svo(&s, (int)(2 * zTileIndex))) return 2;
    for (int i = 0; i < MAXLV; ++i) g_last_counts[i] = (i < s.n - 1) ? s.lvl_count[i] : 0;
    if (merge_common_subtrees(&s, merge_mode)) return 2;
    if (remove_unused_nodes(&s)) return 2;
    if (compress_dag(&s)) return 2;
    *out_words = s.dag.w;
    *out_len = s.dag.n;
    *num_levels = s.n;
    return 0;
}

void or_free(void *p) { free(p); }

/* getTotalVisibility (CompressedShadow.cpp:66-72). */
int or_total_visibility(const u32 *w) {
    if (w[0] == 0x5555u) return 1;
    if (w[0] == 0u) return 0;
    return 2;
}

/* ------------------------------------------------------------------------------------ */
/* Lookups                                                                              */
/* ------------------------------------------------------------------------------------ */

/* getChildOffset (CompressedShadow.cpp:394-402): partial children before child c. */
static inline u32 child_offset(u32 mask, u32 c) {
    const u32 bits = c * 2;
    return (u32)__builtin_popcount(mask & (bits ? (0xAAAAu >> (16 - bits)) : 0u));
}

/* CompressedShadow::traverse (CompressedShadow.cpp:404-463) with getPathFromNDC
 * (CompressedShadowUtil.h:62-75): path = int(((ndc + 1) * 0.5f) * float(R-1)). */
int or_traverse_cpu(const u32 *w, int n, const float ndc[3], int tryLeaf) {
    const int res = (1 << (n - 1)) - 1;
    int p[3];
    for (int i = 0; i < 3; ++i) {
        float v = ndc[i] + 1.0f;
        v = v * 0.5f;
        p[i] = (int)(v * (float)res);
    }
    size_t off = 0;
    int level = n - 2;
    const int minL = tryLeaf ? 3 : 0;
    while (level >= minL) {
        const int b = 1 << level;
        const u32 c = ((p[0] & b) ? 1u : 0u) + ((p[1] & b) ? 2u : 0u) + ((p[2] & b) ? 4u : 0u);
        const u32 m = w[off];
        if (m & (1u << (2 * c))) return 1;
        if (!(m & (2u << (2 * c)))) return 0;
        off = w[off + 1 + child_offset(m, c)];
        --level;
    }
    if (tryLeaf) {
        const u32 c = (u32)(p[2] & 7);
        const u32 m = w[off];
        if (m & (1u << (2 * c))) return 1;
        if (!(m & (2u << (2 * c)))) return 0;
        const size_t idx = off + child_offset(m, c) * 2 + 1;
        const int bit = (p[0] & 7) + 8 * (p[1] & 7);
        const u32 v = (bit < 32) ? (w[idx] & (1u << bit)) : (w[idx + 1] & (1u << (bit - 32)));
        return v ? 1 : 0;
    }
    return 2;
}

/* traverse.cs:41-48 getPathFromNDC, clamped to the grid's extent. */
void or_path_from_ndc(const float ndc[3], int dl, int gl, int *path) {
    const u32 R = 1u << (dl + gl - 1);
    const float mx = (float)(R - 1);
    for (int i = 0; i < 3; ++i) {
        float v = ndc[i] + 1.0f;
        v = v * 0.5f;
        v = v * mx;
        int p;
        if (!(v > 0.0f)) p = 0; /* negative and NaN -> 0 */
        else if (v >= (float)R) p = (int)(R - 1);
        else p = (int)v;
        if (p > (int)(R - 1)) p = (int)(R - 1);
        path[i] = p;
    }
}

/* traverse.cs:75-133 traverse(path) with the top-level grid and per-DAG base offsets. */
int or_traverse_path_cs(const u32 *dag, const u32 *grid, int dl, int gl, int px, int py, int pz) {
    const u32 G = 1u << gl;
    const u32 gx = (u32)px >> (dl - 1), gy = (u32)py >> (dl - 1), gz = (u32)pz >> (dl - 1);
    const u32 base = grid[gz * G * G + gy * G + gx];
    if (base == 0xFFFFFFFFu) return 0;
    if (base == 0xFFFFFFFEu) return 1;
    size_t off = base;
    /* Without leafmasks (MIN_LEVEL 0, traverse.cs:34-38) the walk goes down to level 0, whose
     * masks hold no partial child (absoluteVisible, CompressedShadowUtil.cpp:34-40); a walk that
     * passed it would return visible (traverse.cs:130-132). */
    const int last = use_leafmasks(dl) ? 3 : 0;
    for (int level = dl - 2; level >= last; --level) {
        const u32 b = 1u << level;
        const u32 ci = (((u32)px & b) ? 2u : 0u) + (((u32)py & b) ? 4u : 0u) + (((u32)pz & b) ? 8u : 0u);
        const u32 m = dag[off];
        const u32 v = 3u & (m >> ci);
        if (v == 0) return 0;
        if (v == 1) return 1;
        if (level == 0) return 1; /* the shader's fall-through: no partial voxel exists */
        const u32 co = (u32)__builtin_popcount(m & (ci ? (0xAAAAu >> (16 - ci)) : 0u));
        off = (size_t)base + dag[off + 1 + co];
    }
    if (!use_leafmasks(dl)) return 1;
    {
        const u32 ci = ((u32)pz & 7u) * 2u;
        const u32 m = dag[off];
        const u32 v = 3u & (m >> ci);
        if (v == 0) return 0;
        if (v == 1) return 1;
        const u32 co = (u32)__builtin_popcount(m & (ci ? (0xAAAAu >> (16 - ci)) : 0u));
        const size_t idx = off + co * 2 + 1;
        const u32 bit = ((u32)px & 7u) + 8u * ((u32)py & 7u);
        const u32 w = (bit < 32) ? (dag[idx] & (1u << bit)) : (dag[idx + 1] & (1u << (bit - 32)));
        return w ? 1 : 0;
    }
}

typedef struct {
    const u32 *dag, *grid;
    int dl, gl;
    const float *pos;
    u32 stride;
    size_t i0, i1;
    const float *m;
    uint8_t *vis;
} eval_job;

static const float k_identity[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};

/* traverse.cs:135-149 main(): p = lightViewProj * (pos, 1); p /= p.w; path; traverse. */
static void *eval_range(void *arg) {
    const eval_job *j = (const eval_job *)arg;
    const float *m = j->m ? j->m : k_identity;
    for (size_t i = j->i0; i < j->i1; ++i) {
        const float *q = j->pos + i * j->stride;
        float r[4];
        for (int c = 0; c < 4; ++c) {
            float acc = m[0 * 4 + c] * q[0];
            acc = acc + m[1 * 4 + c] * q[1];
            acc = acc + m[2 * 4 + c] * q[2];
            acc = acc + m[3 * 4 + c] * 1.0f;
            r[c] = acc;
        }
        const float ndc[3] = {r[0] / r[3], r[1] / r[3], r[2] / r[3]};
        int p[3];
        or_path_from_ndc(ndc, j->dl, j->gl, p);
        j->vis[i] = (uint8_t)or_traverse_path_cs(j->dag, j->grid, j->dl, j->gl, p[0], p[1], p[2]);
    }
    return NULL;
}

void or_evaluate(const u32 *dag, const u32 *grid, int dl, int gl, const float *pos, u32 stride,
                 size_t count, const float *m, uint8_t *vis, int nthreads) {
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256];
    eval_job jobs[256];
    const size_t per = (count + (size_t)nthreads - 1) / (size_t)nthreads;
    for (int t = 0; t < nthreads; ++t) {
        jobs[t].dag = dag;
        jobs[t].grid = grid;
        jobs[t].dl = dl;
        jobs[t].gl = gl;
        jobs[t].pos = pos;
        jobs[t].stride = stride;
        jobs[t].i0 = (size_t)t * per < count ? (size_t)t * per : count;
        jobs[t].i1 = (size_t)(t + 1) * per < count ? (size_t)(t + 1) * per : count;
        jobs[t].m = m;
        jobs[t].vis = vis;
    }
    for (int t = 1; t < nthreads; ++t) pthread_create(&th[t], NULL, eval_range, &jobs[t]);
    eval_range(&jobs[0]);
    for (int t = 1; t < nthreads; ++t) pthread_join(th[t], NULL);
}

/* createTopLevelGrid (CompressedShadowContainer.cpp:71-91), sentinels as traverse.cs:85-88. */
void or_top_level_grid(const u32 *const *dags, const size_t *lens, size_t count, u32 *grid) {
    u32 off = 0;
    for (size_t i = 0; i < count; ++i) {
        const int v = or_total_visibility(dags[i]);
        grid[i] = (v == 0) ? 0xFFFFFFFFu : (v == 1) ? 0xFFFFFFFEu : off;
        off += (u32)lens[i];
    }
}

int or_brute_visible(const float *D, int N, u32 T, int px, int py, int pz) {
    const float hf = (float)((u32)N * T);
    const float mid = ((float)pz + (float)(pz + 1)) * 0.5f;
    return mid <= D[(size_t)py * N + px] * hf;
}

/* ------------------------------------------------------------------------------------ */
/* Synthetic inputs                                                                      */
/* ------------------------------------------------------------------------------------ */

/* splitmix64 (Steele, Lea & Flood 2014; constants as published by Vigna). */
static inline u64 sm64_next(u64 *s) {
    u64 z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static inline u64 sm64_at(u64 seed, u64 k) {
    u64 s = seed + k * 0x9E3779B97F4A7C15ull;
    return sm64_next(&s);
}
static inline float u01(u64 r) { return (float)(r >> 40) * (1.0f / 16777216.0f); }

typedef struct {
    int s, x0, y0;
    float dz;
} box_t;

static const int k_oct = 4;

/* Depth tile [x0,x0+tileN) x [y0,y0+tileN) of a fullN x fullN map: tilted plane
 * 0.6 + 0.1*u, optional 4-octave value noise (amplitude 0.05), clamp to [0,1], then
 * D = min(D, dz) under each box.  RNG draw order: all boxes, then the noise lattices. */
void or_gen_depth_tile(int fullN, u64 seed, int nboxes, int noise, int x0, int y0, int tileN,
                       float *out) {
    u64 st = seed;
    box_t *boxes = (box_t *)malloc(sizeof(box_t) * (nboxes > 0 ? nboxes : 1));
    for (int b = 0; b < nboxes; ++b) {
        int s = 16 + (int)(sm64_next(&st) % 113u);
        if (s > fullN) s = fullN;
        boxes[b].s = s;
        boxes[b].x0 = (int)(sm64_next(&st) % (u64)(fullN - s + 1));
        boxes[b].y0 = (int)(sm64_next(&st) % (u64)(fullN - s + 1));
        boxes[b].dz = 0.1f + 0.4f * u01(sm64_next(&st));
    }
    float *lat[4] = {0, 0, 0, 0};
    if (noise) {
        for (int o = 0; o < k_oct; ++o) {
            const int cells = 8 << o;
            const int L = cells + 1;
            lat[o] = (float *)malloc(sizeof(float) * (size_t)L * L);
            for (int i = 0; i < L * L; ++i) lat[o][i] = u01(sm64_next(&st));
        }
    }
    const float fN = (float)fullN;
    for (int ty = 0; ty < tileN; ++ty) {
        const int gy = y0 + ty;
        for (int tx = 0; tx < tileN; ++tx) {
            const int gx = x0 + tx;
            float d = 0.6f + 0.1f * (((float)gx + 0.5f) / fN);
            if (noise) {
                float acc = 0.0f, w = 1.0f;
                for (int o = 0; o < k_oct; ++o) {
                    const int cells = 8 << o;
                    const float sc = (float)cells / fN;
                    const float fx = ((float)gx + 0.5f) * sc;
                    const float fy = ((float)gy + 0.5f) * sc;
                    int ix = (int)fx, iy = (int)fy;
                    if (ix > cells - 1) ix = cells - 1;
                    if (iy > cells - 1) iy = cells - 1;
                    const float rx = fx - (float)ix, ry = fy - (float)iy;
                    const float sx = rx * rx * (3.0f - 2.0f * rx);
                    const float sy = ry * ry * (3.0f - 2.0f * ry);
                    const int L = cells + 1;
                    const float a = lat[o][iy * L + ix], b = lat[o][iy * L + ix + 1];
                    const float c = lat[o][(iy + 1) * L + ix], e = lat[o][(iy + 1) * L + ix + 1];
                    const float top = a + (b - a) * sx;
                    const float bot = c + (e - c) * sx;
                    const float v = top + (bot - top) * sy;
                    acc = acc + w * v;
                    w = w * 0.5f;
                }
                d = d + 0.05f * (acc / 1.875f - 0.5f);
            }
            if (d < 0.0f) d = 0.0f;
            if (d > 1.0f) d = 1.0f;
            out[(size_t)ty * tileN + tx] = d;
        }
    }
    for (int b = 0; b < nboxes; ++b) {
        const int bx0 = boxes[b].x0 > x0 ? boxes[b].x0 : x0;
        const int by0 = boxes[b].y0 > y0 ? boxes[b].y0 : y0;
        const int bx1 = (boxes[b].x0 + boxes[b].s) < (x0 + tileN) ? (boxes[b].x0 + boxes[b].s) : (x0 + tileN);
        const int by1 = (boxes[b].y0 + boxes[b].s) < (y0 + tileN) ? (boxes[b].y0 + boxes[b].s) : (y0 + tileN);
        for (int gy = by0; gy < by1; ++gy)
            for (int gx = bx0; gx < bx1; ++gx) {
                float *p = out + (size_t)(gy - y0) * tileN + (gx - x0);
                if (boxes[b].dz < *p) *p = boxes[b].dz;
            }
    }
    for (int o = 0; o < k_oct; ++o) free(lat[o]);
    free(boxes);
}

/* Uniform queries: p = (u01, u01, u01) from a counter-based stream, stored as NDC 2p-1;
 * with stride 4 the fourth component is 1. */
void or_gen_queries_uniform(u64 seed, size_t count, u32 stride, float *out) {
    for (size_t i = 0; i < count; ++i) {
        for (u32 c = 0; c < 3; ++c) {
            const float p = u01(sm64_at(seed, 3 * (u64)i + c));
            out[i * stride + c] = 2.0f * p - 1.0f;
        }
        if (stride == 4) out[i * stride + 3] = 1.0f;
    }
}

/* G-buffer queries: pixel (i,j) of W x H at (u,v) = ((i+.5)/W, (j+.5)/H) on the depth
 * texel under it; query 2p is on the surface, query 2p+1 jittered by U(-2,2) voxels in z.
 * Stored as float4 NDC (x, y, z, 1). */
void or_gen_queries_gbuffer(const float *D, int N, u64 seed, int W, int H, float *out) {
    const float fN = (float)N;
    for (int j = 0; j < H; ++j)
        for (int i = 0; i < W; ++i) {
            const size_t pix = (size_t)j * W + i;
            const float u = ((float)i + 0.5f) / (float)W;
            const float v = ((float)j + 0.5f) / (float)H;
            int tx = (int)(u * fN), ty = (int)(v * fN);
            if (tx > N - 1) tx = N - 1;
            if (ty > N - 1) ty = N - 1;
            const float z0 = D[(size_t)ty * N + tx];
            float z1 = z0 + (2.0f * u01(sm64_at(seed, pix)) - 1.0f) * (2.0f / fN);
            if (z1 < 0.0f) z1 = 0.0f;
            if (z1 > 1.0f) z1 = 1.0f;
            float *q0 = out + (2 * pix) * 4, *q1 = out + (2 * pix + 1) * 4;
            q0[0] = 2.0f * u - 1.0f;
            q0[1] = 2.0f * v - 1.0f;
            q0[2] = 2.0f * z0 - 1.0f;
            q0[3] = 1.0f;
            q1[0] = q0[0];
            q1[1] = q0[1];
            q1[2] = 2.0f * z1 - 1.0f;
            q1[3] = 1.0f;
        }
}
