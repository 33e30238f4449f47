/*
 * prt_oracle.c -- CPU restatement (C99, FP64 where the reference is FP64)
 * of the ParlllelRtree hot path.  TEST INFRASTRUCTURE ONLY: the checker for
 * the HIP library, never linked into it.
 *
 * Reference files (BU = BotomUp-rtree.cu, TD = TopDown-rtree.cu,
 * OV = vector_map_overlay_gpc_kernel.cu) cannot be compiled here (no nvcc,
 * a compile error at BU:391, absent gpc.h/common.h/XMalloc), so this file
 * restates their host semantics.  How each part is pinned (tests/ hold the
 * checks):
 *   orc_interleave_ref4 / orc_morton_ref4d : closed-form bit table KATs.
 *   orc_morton_center30                    : per-bit textbook Morton loop.
 *   orc_stable_order_*                     : numpy argsort(kind="stable").
 *   orc_bu_build                           : hand-derived tiny trees + tree
 *                                            invariants (tight unions, level
 *                                            chain, root last).
 *   orc_td_build                           : hand-derived tiny layout + numpy
 *                                            lexsort restatement + invariants.
 *   orc_join_*                             : brute force (tree-independent
 *                                            hit set).
 *   orc_search_stats_*                     : leaf hits = brute-force hit
 *                                            count; hand-derived counters of a
 *                                            tiny tree; closed forms for
 *                                            single-node trees.
 *   orc_sibling_overlap_*                  : closed forms (identical rects:
 *                                            sum of C(entries, 2); a grid
 *                                            whose nodes are disjoint: 0).
 *   orc_clip_area                          : closed forms (rect/rect, P/P,
 *                                            disjoint, nested), Sutherland-
 *                                            Hodgman on convex inputs, an
 *                                            independent Green's-theorem area
 *                                            on random stars, symmetry and
 *                                            power-of-two translation.
 *   orc_area_grad                          : central finite differences of
 *                                            orc_clip_area and the closed form
 *                                            dA/dx_i = (y_{i+1}-y_{i-1})/2 for
 *                                            vertices deep inside the other
 *                                            polygon.
 * Build with -ffp-contract=off: the HIP kernels evaluate the same FP64
 * expressions in the same order, and contraction would change roundings.
 */
#include "prt_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ====================================================================== */
/* Morton keys                                                            */
/* ====================================================================== */

/* Build-defined CENTER30 quantisation (the configs key rect centres, 15 bits
 * per axis).  c = 0.5f*(lo+hi) in fp32, q = floor(c*2^15) clamped to
 * [0, 32767]; NaN maps to 0.  Replaces the integer scale rule of
 * BU:704-721 + BU:80-83 for fp32 inputs. */
uint32_t orc_quant15(float lo, float hi)
{
    volatile float sum = lo + hi;      /* keep fp32 rounding of the sum */
    float c = 0.5f * sum;
    float s = c * 32768.0f;
    if (!(s >= 0.0f)) return 0u;
    if (s >= 32767.0f) return 32767u;
    return (uint32_t)s;                /* truncation == floor for s >= 0 */
}

/* Spread the low 15 bits of v onto the even bit positions (bit k -> 2k). */
uint32_t orc_spread15(uint32_t v)
{
    uint32_t out = 0;
    for (int k = 0; k < 15; ++k) out |= ((v >> k) & 1u) << (2 * k);
    return out;
}

/* CENTER30 key: x on even bits, y on odd bits. */
void orc_morton_center30(const float *r4, int64_t n, uint32_t *keys)
{
    for (int64_t i = 0; i < n; ++i) {
        uint32_t qx = orc_quant15(r4[4 * i + 0], r4[4 * i + 2]);
        uint32_t qy = orc_quant15(r4[4 * i + 1], r4[4 * i + 3]);
        keys[i] = orc_spread15(qx) | (orc_spread15(qy) << 1);
    }
}

/* Reference 4-D interleave, BU:98-123: bit k (k < 16) of a1 lands on key
 * bit 8*(k/2) + 6 + (k&1), a2 on +4, a3 on +2, a4 on +0; bits >= 16 of each
 * input are dropped by the masks.  Written as an explicit per-bit map. */
uint64_t orc_interleave_ref4(uint64_t a1, uint64_t a2, uint64_t a3, uint64_t a4)
{
    uint64_t in[4] = {a1, a2, a3, a4};
    uint64_t out = 0;
    for (int w = 0; w < 4; ++w) {
        int base = 6 - 2 * w;
        for (int k = 0; k < 16; ++k) {
            uint64_t bit = (in[w] >> k) & 1u;
            out |= bit << (8 * (k / 2) + base + (k & 1));
        }
    }
    return out;
}

/* Reference key on int64 rects, BU:73-94: divide each coordinate by the
 * scale (C truncating division), subtract min (0 after BU:720-721), then
 * interleave (x1, y1, x2, y2). */
void orc_morton_ref4d(const int64_t *r4, int64_t n, int64_t scale, uint64_t *keys)
{
    for (int64_t i = 0; i < n; ++i) {
        uint64_t a[4];
        for (int j = 0; j < 4; ++j) a[j] = (uint64_t)(r4[4 * i + j] / scale);
        keys[i] = orc_interleave_ref4(a[0], a[1], a[2], a[3]);
    }
}

/* Scale rule BU:711-717: the smallest power of ten 10^k with range/10^k
 * <= 1e4 under repeated integer division by ten. */
int64_t orc_ref_scale(int64_t range)
{
    int64_t s = 1;
    while (range > 10000) {
        s *= 10;
        range /= 10;
    }
    return s;
}

/* ====================================================================== */
/* Stable ordering (the sort both builds rely on)                          */
/* ====================================================================== */

/* perm = positions ordered by (key, position): a stable ascending sort, the
 * order the reference's LSD radix sort produces (BU:125-266, TD:162-261) and
 * glibc's merge-sort qsort produces for the 0/1 comparator (BU:638-640,
 * TD:22-24). */
void orc_stable_order_u64(const uint64_t *keys, int64_t n, uint32_t *perm)
{
    if (n <= 0) return;
    uint32_t *tmp = (uint32_t *)malloc((size_t)n * sizeof(uint32_t));
    for (int64_t i = 0; i < n; ++i) perm[i] = (uint32_t)i;
    uint32_t *src = perm, *dst = tmp;
    for (int64_t w = 1; w < n; w *= 2) {
        for (int64_t lo = 0; lo < n; lo += 2 * w) {
            int64_t mid = lo + w < n ? lo + w : n;
            int64_t hi = lo + 2 * w < n ? lo + 2 * w : n;
            int64_t i = lo, j = mid, o = lo;
            while (i < mid && j < hi) {
                if (keys[src[j]] < keys[src[i]]) dst[o++] = src[j++];
                else dst[o++] = src[i++];
            }
            while (i < mid) dst[o++] = src[i++];
            while (j < hi) dst[o++] = src[j++];
        }
        uint32_t *t = src; src = dst; dst = t;
    }
    if (src != perm) memcpy(perm, src, (size_t)n * sizeof(uint32_t));
    free(tmp);
}

void orc_stable_order_u32(const uint32_t *keys, int64_t n, uint32_t *perm)
{
    if (n <= 0) return;
    uint64_t *k = (uint64_t *)malloc((size_t)n * sizeof(uint64_t));
    for (int64_t i = 0; i < n; ++i) k[i] = keys[i];
    orc_stable_order_u64(k, n, perm);
    free(k);
}

/* ====================================================================== */
/* Bottom-up build                                                         */
/* ====================================================================== */

static void set_empty(float *m)
{
    m[0] = INFINITY; m[1] = INFINITY; m[2] = -INFINITY; m[3] = -INFINITY;
}

/* min / max in the total order with -0 below +0 (the GPU's fminf / fmaxf),
 * so that a union does not depend on the order of its boxes even when
 * they hold zeros of both signs */
static float min0(float a, float b)
{
    if (a == b) return signbit(a) ? a : b;
    return b < a ? b : a;
}
static float max0(float a, float b)
{
    if (a == b) return signbit(a) ? b : a;
    return b > a ? b : a;
}

static void merge_into(float *dst, const float *src)
{
    dst[0] = min0(dst[0], src[0]);
    dst[1] = min0(dst[1], src[1]);
    dst[2] = max0(dst[2], src[2]);
    dst[3] = max0(dst[3], src[3]);
}

/* Level chain of BU:788-830: ceil(n/f) leaves, then ceil(count/f) per level
 * until one node remains.  Node ids run level by level, leaves first, root
 * last.  Returns the number of levels. */
int orc_bu_levels(int64_t n, int f, int64_t *level_start, int64_t *level_count)
{
    if (n <= 0 || f < 2) return 0;
    int L = 0;
    int64_t cnt = (n + f - 1) / f, start = 0;
    for (;;) {
        level_start[L] = start;
        level_count[L] = cnt;
        start += cnt;
        ++L;
        if (cnt <= 1) break;
        cnt = (cnt + f - 1) / f;
    }
    return L;
}

/* Host build of BU:776-833 with the GPU twin's start/end fields
 * (BU:373-426): stable sort by key, leaves = consecutive chunks of f sorted
 * rects (entry ref = original id, BU:794-801), each upper entry = union of
 * every entry of its child (BU:816-826, ref = child id).  Unused entry slots
 * hold the empty box (+inf,+inf,-inf,-inf) and ref -1.  Returns #nodes. */
/* Packing of BU:788-830 for a given leaf order perm (sorted position ->
 * rect id). */
static int bu_pack(const float *r4, int64_t n, int f, const uint32_t *perm,
                   int32_t *node_count, int32_t *node_start, int32_t *node_end,
                   float *entry_mbr, int32_t *entry_ref)
{
    int64_t ls[64], lc[64];
    int L = orc_bu_levels(n, f, ls, lc);
    if (L == 0) return -1;
    for (int64_t i = 0; i < lc[0]; ++i) {
        int64_t b = i * f, e = b + f - 1 < n - 1 ? b + f - 1 : n - 1;
        node_count[i] = (int32_t)(e - b + 1);
        node_start[i] = (int32_t)b;
        node_end[i] = (int32_t)e;
        for (int j = 0; j < f; ++j) {
            float *m = entry_mbr + 4 * (i * f + j);
            if (b + j <= e) {
                uint32_t id = perm[b + j];
                memcpy(m, r4 + 4 * (int64_t)id, 4 * sizeof(float));
                entry_ref[i * f + j] = (int32_t)id;
            } else {
                set_empty(m);
                entry_ref[i * f + j] = -1;
            }
        }
    }
    for (int l = 1; l < L; ++l) {
        for (int64_t i = 0; i < lc[l]; ++i) {
            int64_t node = ls[l] + i;
            int64_t b = ls[l - 1] + i * f;
            int64_t last = ls[l - 1] + lc[l - 1] - 1;
            int64_t e = b + f - 1 < last ? b + f - 1 : last;
            node_count[node] = (int32_t)(e - b + 1);
            node_start[node] = node_start[b];
            node_end[node] = node_end[e];
            for (int j = 0; j < f; ++j) {
                float *m = entry_mbr + 4 * (node * f + j);
                set_empty(m);
                if (b + j <= e) {
                    int64_t c = b + j;
                    for (int k = 0; k < node_count[c]; ++k)
                        merge_into(m, entry_mbr + 4 * (c * f + k));
                    entry_ref[node * f + j] = (int32_t)c;
                } else {
                    entry_ref[node * f + j] = -1;
                }
            }
        }
    }
    return (int)(ls[L - 1] + lc[L - 1]);
}

int orc_bu_build(const float *r4, int64_t n, int f, const uint32_t *keys,
                 uint32_t *perm, int32_t *node_count, int32_t *node_start,
                 int32_t *node_end, float *entry_mbr, int32_t *entry_ref)
{
    if (n <= 0) return -1;
    orc_stable_order_u32(keys, n, perm);
    return bu_pack(r4, n, f, perm, node_count, node_start, node_end, entry_mbr, entry_ref);
}

/* The reference's own key order (BU:841-849 sorting the REF4D64 keys of
 * CalcuateZOrderValue, BU:73-94): stable by 64-bit key, then packed. */
int orc_bu_build_keys64(const float *r4, int64_t n, int f, const uint64_t *keys,
                        uint32_t *perm, int32_t *node_count, int32_t *node_start,
                        int32_t *node_end, float *entry_mbr, int32_t *entry_ref)
{
    if (n <= 0) return -1;
    orc_stable_order_u64(keys, n, perm);
    return bu_pack(r4, n, f, perm, node_count, node_start, node_end, entry_mbr, entry_ref);
}

/* ====================================================================== */
/* Top-down build                                                          */
/* ====================================================================== */

/* Slot offsets of TD:504-524: repeat num = ceil(num/f)*f; push; num /= f
 * while num > 1, then push 0, reverse and prefix-sum.  level_offset gets
 * L+1 entries; returns L.  (n == 1 is undefined in the reference; here it
 * takes the n <= f layout: one root block.) */
int orc_td_levels(int64_t n, int f, int64_t *level_offset)
{
    if (n <= 0 || f < 2) return 0;
    int64_t pushed[64];
    int L = 0;
    int64_t num = n;
    do {
        num = ((num + f - 1) / f) * f;
        pushed[L++] = num;
        num /= f;
    } while (num > 1);
    level_offset[0] = 0;
    for (int p = 0; p < L; ++p) level_offset[p + 1] = level_offset[p] + pushed[L - 1 - p];
    return L;
}

/* Order-preserving fp32 -> u32 map; -0 is folded onto +0 so that equal
 * values tie (the reference compares integers, where -0 does not exist). */
static uint32_t ord32(float v)
{
    uint32_t u;
    memcpy(&u, &v, 4);
    if (u == 0x80000000u) u = 0;
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

/* Host build of TD:572-627: for level p, key = (pos / a) << 32 | x1 (p even)
 * or y1 (p odd) -- the exact host form TD:577-586, with ord32 for fp32 --
 * stable sort, a /= f, then entry i of level p covers sorted [i*a,
 * min(n, i*a+a)) with child = offset[p+1] + i*f (0 if empty, TD:591-613).
 * Last level: leaf slot -pos-1, MBR of the rect at pos (TD:616-624); slots
 * past n are 0.  Empty slots carry the empty box.  perm[pos] = original id
 * of the rect at final sorted position pos.  Returns #slots. */
int orc_td_build(const float *r4, int64_t n, int f, int32_t *slot_child,
                 float *slot_mbr, uint32_t *perm)
{
    int64_t off[65];
    int L = orc_td_levels(n, f, off);
    if (L == 0) return -1;
    int64_t a = 1;
    for (int p = 0; p < L; ++p) a *= f;
    uint32_t *cur = perm;
    uint32_t *nxt = (uint32_t *)malloc((size_t)n * sizeof(uint32_t));
    uint32_t *ord = (uint32_t *)malloc((size_t)n * sizeof(uint32_t));
    uint64_t *key = (uint64_t *)malloc((size_t)n * sizeof(uint64_t));
    for (int64_t i = 0; i < n; ++i) cur[i] = (uint32_t)i;
    for (int p = 0; p < L; ++p) {
        int comp = (p & 1) ? 1 : 0;   /* x1 on even levels, y1 on odd */
        for (int64_t i = 0; i < n; ++i)
            key[i] = ((uint64_t)(i / a) << 32) | ord32(r4[4 * (int64_t)cur[i] + comp]);
        orc_stable_order_u64(key, n, ord);
        for (int64_t i = 0; i < n; ++i) nxt[i] = cur[ord[i]];
        memcpy(cur, nxt, (size_t)n * sizeof(uint32_t));
        a /= f;
        for (int64_t i = 0; i < off[p + 1] - off[p]; ++i) {
            int64_t s = off[p] + i;
            int64_t b = i * a, e = b + a < n ? b + a : n;
            float *m = slot_mbr + 4 * s;
            set_empty(m);
            if (b < n) {
                slot_child[s] = (int32_t)(off[p + 1] + i * f);
                for (int64_t k = b; k < e; ++k) merge_into(m, r4 + 4 * (int64_t)cur[k]);
            } else {
                slot_child[s] = 0;
            }
        }
    }
    for (int64_t i = 0; i < off[L] - off[L - 1]; ++i) {
        int64_t s = off[L - 1] + i;
        float *m = slot_mbr + 4 * s;
        if (i < n) {
            slot_child[s] = (int32_t)(-i - 1);
            memcpy(m, r4 + 4 * (int64_t)cur[i], 4 * sizeof(float));
        } else {
            slot_child[s] = 0;
            set_empty(m);
        }
    }
    free(nxt);
    free(ord);
    free(key);
    return (int)off[L];
}

/* ====================================================================== */
/* Spatial join                                                            */
/* ====================================================================== */

/* Closed-interval overlap.  For the reference's integer rects the test
 * (w_a + w_b) + 1e-8 < max(x2) - min(x1) => disjoint (BU:496-506,
 * TD:26-39) is exactly a.x1 <= b.x2 && b.x1 <= a.x2 per axis: touching
 * counts.  On fp32 the closed test is exact. */
int orc_overlap(const float *a, const float *b)
{
    return a[0] <= b[2] && b[0] <= a[2] && a[1] <= b[3] && b[1] <= a[3];
}

static int cmp_u32(const void *x, const void *y)
{
    uint32_t a = *(const uint32_t *)x, b = *(const uint32_t *)y;
    return (a > b) - (a < b);
}

typedef struct {
    uint32_t *buf;
    int64_t len, cap;
} hitvec;

static void hv_push(hitvec *h, uint32_t r)
{
    if (h->len == h->cap) {
        h->cap = h->cap ? 2 * h->cap : 256;
        h->buf = (uint32_t *)realloc(h->buf, (size_t)h->cap * sizeof(uint32_t));
    }
    h->buf[h->len++] = r;
}

/* Recursive search of BU:455-483: test every entry of an internal node and
 * descend into overlapping children; at a leaf every overlapping entry is a
 * hit.  `tested` counts entry tests (the reference's `num`). */
static void bu_dfs(int64_t node, int f, int64_t num_leaves, const int32_t *cnt,
                   const float *mbr, const int32_t *ref, const float *q,
                   hitvec *h, int64_t *tested)
{
    for (int i = 0; i < cnt[node]; ++i) {
        ++*tested;
        const float *m = mbr + 4 * (node * f + i);
        if (!orc_overlap(q, m)) continue;
        if (node < num_leaves) hv_push(h, (uint32_t)ref[node * f + i]);
        else bu_dfs(ref[node * f + i], f, num_leaves, cnt, mbr, ref, q, h, tested);
    }
}

static int64_t emit_sorted(hitvec *h, uint32_t q, uint32_t *pairs, int64_t total, int64_t cap)
{
    qsort(h->buf, (size_t)h->len, sizeof(uint32_t), cmp_u32);
    for (int64_t k = 0; k < h->len; ++k, ++total) {
        if (total < cap) {
            pairs[2 * total] = q;
            pairs[2 * total + 1] = h->buf[k];
        }
    }
    h->len = 0;
    return total;
}

/* Query every window against the BU tree from the root (the last node,
 * BU:867-869); pairs (q, r) sorted by (q, r).  Returns the pair count;
 * writes at most cap pairs. */
int64_t orc_join_bu(int64_t num_nodes, int f, int64_t num_leaves,
                    const int32_t *node_count, const float *entry_mbr,
                    const int32_t *entry_ref, const float *q4, int64_t nq,
                    uint32_t *pairs, int64_t cap, int64_t *tested)
{
    hitvec h = {0, 0, 0};
    int64_t total = 0, t = 0;
    for (int64_t q = 0; q < nq; ++q) {
        bu_dfs(num_nodes - 1, f, num_leaves, node_count, entry_mbr, entry_ref,
               q4 + 4 * q, &h, &t);
        total = emit_sorted(&h, (uint32_t)q, pairs, total, cap);
    }
    free(h.buf);
    if (tested) *tested = t;
    return total;
}

/* Recursive search of TD:56-76 over the implicit slot layout: a block of f
 * slots; child > 0 is the next block, child < 0 a leaf at sorted position
 * -child-1, 0 an empty slot. */
static void td_dfs(int64_t index, int f, const int32_t *child, const float *mbr,
                   const uint32_t *perm, const float *q, hitvec *h, int64_t *tested)
{
    for (int i = 0; i < f; ++i) {
        int64_t s = index + i;
        ++*tested;
        if (child[s] == 0) continue;
        if (!orc_overlap(mbr + 4 * s, q)) continue;
        if (child[s] > 0) td_dfs(child[s], f, child, mbr, perm, q, h, tested);
        else hv_push(h, perm[-(int64_t)child[s] - 1]);
    }
}

int64_t orc_join_td(int f, const int32_t *slot_child, const float *slot_mbr,
                    const uint32_t *perm, const float *q4, int64_t nq,
                    uint32_t *pairs, int64_t cap, int64_t *tested)
{
    hitvec h = {0, 0, 0};
    int64_t total = 0, t = 0;
    for (int64_t q = 0; q < nq; ++q) {
        td_dfs(0, f, slot_child, slot_mbr, perm, q4 + 4 * q, &h, &t);
        total = emit_sorted(&h, (uint32_t)q, pairs, total, cap);
    }
    free(h.buf);
    if (tested) *tested = t;
    return total;
}

/* The hit set is tree independent: {(q, r) : overlap(q, r)}. */
int64_t orc_join_brute(const float *r4, int64_t n, const float *q4, int64_t nq,
                       uint32_t *pairs, int64_t cap)
{
    int64_t total = 0;
    for (int64_t q = 0; q < nq; ++q)
        for (int64_t r = 0; r < n; ++r)
            if (orc_overlap(q4 + 4 * q, r4 + 4 * r)) {
                if (total < cap) {
                    pairs[2 * total] = (uint32_t)q;
                    pairs[2 * total + 1] = (uint32_t)r;
                }
                ++total;
            }
    return total;
}

/* ====================================================================== */
/* Search counters and tree quality (the reference's printed statistics)  */
/* ====================================================================== */

/* RtreeSearchHelp, BU:455-483: every entry of a visited node counts in
 * `num`; an overlapping entry adds one to the node's running value (which
 * starts at -1) and, at an internal node, the child's result; a leaf hit
 * also counts in `num1`; the node returns max(0, value). */
static int64_t bu_stats_rec(int64_t node, int f, int64_t num_leaves, const int32_t *cnt,
                            const float *mbr, const int32_t *ref, const float *q,
                            int64_t *num, int64_t *num1)
{
    int64_t ans = -1;
    for (int i = 0; i < cnt[node]; ++i) {
        ++*num;
        if (!orc_overlap(q, mbr + 4 * (node * f + i))) continue;
        ++ans;
        if (node < num_leaves) ++*num1;
        else ans += bu_stats_rec(ref[node * f + i], f, num_leaves, cnt, mbr, ref, q, num, num1);
    }
    return ans > 0 ? ans : 0;
}

/* Per query q: num[q], count[q] (the search's return value) and num1[q] of
 * RtreeSearch (BU:439-442) started at the root (the last node, BU:1008). */
void orc_search_stats_bu(int64_t num_nodes, int f, int64_t num_leaves,
                         const int32_t *node_count, const float *entry_mbr,
                         const int32_t *entry_ref, const float *q4, int64_t nq,
                         int64_t *num, int64_t *count, int64_t *num1)
{
    for (int64_t q = 0; q < nq; ++q) {
        num[q] = 0;
        num1[q] = 0;
        count[q] = bu_stats_rec(num_nodes - 1, f, num_leaves, node_count, entry_mbr, entry_ref,
                                q4 + 4 * q, &num[q], &num1[q]);
    }
}

/* RtreeSearch, TD:56-76: all f slots of a visited block count in `num`
 * (empty ones too); a non-empty overlapping slot adds one and the child
 * block's result, or counts a leaf hit. */
static int64_t td_stats_rec(int64_t index, int f, const int32_t *child, const float *mbr,
                            const float *q, int64_t *num, int64_t *num1)
{
    int64_t ans = -1;
    for (int i = 0; i < f; ++i) {
        const int64_t s = index + i;
        ++*num;
        if (child[s] == 0) continue;
        if (!orc_overlap(mbr + 4 * s, q)) continue;
        ++ans;
        if (child[s] > 0) ans += td_stats_rec(child[s], f, child, mbr, q, num, num1);
        else ++*num1;
    }
    return ans > 0 ? ans : 0;
}

/* Per query, from the root block at slot 0 (TD:772). */
void orc_search_stats_td(int f, const int32_t *slot_child, const float *slot_mbr,
                         const float *q4, int64_t nq, int64_t *num, int64_t *count,
                         int64_t *num1)
{
    for (int64_t q = 0; q < nq; ++q) {
        num[q] = 0;
        num1[q] = 0;
        count[q] = td_stats_rec(0, f, slot_child, slot_mbr, q4 + 4 * q, &num[q], &num1[q]);
    }
}

/* RtreeSearch1, BU:574-590: pairs of overlapping entries inside each node,
 * summed over the nodes reachable from `node`. */
static int64_t bu_sib_rec(int64_t node, int f, int64_t num_leaves, const int32_t *cnt,
                          const float *mbr, const int32_t *ref)
{
    int64_t ans = 0;
    for (int i = 1; i < cnt[node]; ++i)
        for (int j = i - 1; j >= 0; --j)
            if (orc_overlap(mbr + 4 * (node * f + j), mbr + 4 * (node * f + i))) ++ans;
    if (node >= num_leaves)
        for (int i = 0; i < cnt[node]; ++i)
            ans += bu_sib_rec(ref[node * f + i], f, num_leaves, cnt, mbr, ref);
    return ans;
}

int64_t orc_sibling_overlap_bu(int64_t num_nodes, int f, int64_t num_leaves,
                               const int32_t *node_count, const float *entry_mbr,
                               const int32_t *entry_ref)
{
    return bu_sib_rec(num_nodes - 1, f, num_leaves, node_count, entry_mbr, entry_ref);
}

/* RtreeSearch1, TD:78-93: within a block, pairs among the leading non-empty
 * slots (the loop stops at the first empty slot), then the child blocks. */
static int64_t td_sib_rec(int64_t index, int f, const int32_t *child, const float *mbr)
{
    int64_t ans = 0;
    for (int i = 1; i < f && child[index + i] != 0; ++i)
        for (int j = i - 1; j >= 0; --j)
            if (orc_overlap(mbr + 4 * (index + i), mbr + 4 * (index + j))) ++ans;
    for (int i = 0; i < f && child[index + i] != 0; ++i)
        if (child[index + i] > 0) ans += td_sib_rec(child[index + i], f, child, mbr);
    return ans;
}

int64_t orc_sibling_overlap_td(int f, const int32_t *slot_child, const float *slot_mbr)
{
    return td_sib_rec(0, f, slot_child, slot_mbr);
}

/* ====================================================================== */
/* Polygon intersection area: GPC-style scanbeam sweep                     */
/* ====================================================================== */
/*
 * The reference intersects each candidate pair with GPC (gpc_polygon_clip,
 * OV:115-750, GPC_INT) and keeps the result contours.  The area of those
 * contours is the area swept by the scanbeam loop, which is what this
 * function accumulates:
 *   - minimax pre-test (OV:1656-1709): disjoint closed bounding boxes give
 *     an empty result;
 *   - scanbeam table: the sorted distinct vertex y's (OV:1019-1085,
 *     add_to_sbtree/build_sbt), restricted to the common y range;
 *   - edges bottom->top with dx = (top.x-bot.x)/(top.y-bot.y)
 *     (OV:1241-1242, 1287-1288); horizontal edges span no beam;
 *   - edge x at a beam boundary: top.x when y == top.y, else
 *     bot.x + dx*(y - bot.y) (OV:454-457);
 *   - crossings inside a beam (OV:1395-1445): den = (xt_L-xb_L)-(xt_R-xb_R),
 *     r = (xb_R-xb_L)/den, iy = r*dy + yb (OV:1431-1433, 480); the beam is
 *     cut at every crossing height so that each sub-beam is crossing free;
 *   - even-odd parity per operand (GPC's parity[], OV:140, 313-330): a gap
 *     between two neighbouring edges belongs to the intersection when both
 *     the subject and the clip parity are odd.
 * Each such gap is a trapezoid; its doubled area
 *   ((xa_R - xa_L) + (xc_R - xc_L)) * (yc - ya)
 * is added to the beam's sum (starting from +0), in order of sub-beam, then
 * left to right.  The beam sums s[0..nb) are then added by a fixed pairwise
 * tree: for w = 1, 2, 4, ...: s[i] += s[i + w] for i = 0, 2w, 4w, ... with
 * i + w < nb; area = 0.5 * s[0].  Edges in a sub-beam are ordered by
 * (xa+xc, xa, edge id), a strict total order, so nothing depends on the
 * input order of the active table, and beams may be processed in any order
 * or in parallel.  Edge id = vertex index for P, nP + index for Q.
 */

typedef struct {
    double bx, by, tx, ty, dx;
    int id, poly;
} cedge;

static inline double edge_x(const cedge *e, double y)
{
    if (y == e->ty) return e->tx;
    return e->bx + e->dx * (y - e->by);
}

static int cmp_double(const void *x, const void *y)
{
    double a = *(const double *)x, b = *(const double *)y;
    return (a > b) - (a < b);
}

static int cmp_edge_by(const void *x, const void *y)
{
    const cedge *a = (const cedge *)x, *b = (const cedge *)y;
    if (a->by != b->by) return a->by < b->by ? -1 : 1;
    return a->id - b->id;
}

static void bbox(const double *P, int n, double *bb)
{
    bb[0] = bb[2] = P[0];
    bb[1] = bb[3] = P[1];
    for (int i = 1; i < n; ++i) {
        double x = P[2 * i], y = P[2 * i + 1];
        if (x < bb[0]) bb[0] = x;
        if (x > bb[2]) bb[2] = x;
        if (y < bb[1]) bb[1] = y;
        if (y > bb[3]) bb[3] = y;
    }
}

/* before(i, j): strict order (xa+xc, xa, id) used inside a sub-beam */
static inline int sb_before(const double *km, const double *xa, const cedge *const *act,
                            int i, int j)
{
    if (km[i] != km[j]) return km[i] < km[j];
    if (xa[i] != xa[j]) return xa[i] < xa[j];
    return act[i]->id < act[j]->id;
}

typedef struct {
    cedge *edges;
    double *ys;
    const cedge **act;
    double *xb, *xt, *xa, *xc, *km, *cy, *bsum;
    int *ord;
    int cap_e, cap_y;
} clip_ws;

static void ws_reserve(clip_ws *w, int ne, int ny)
{
    if (ne > w->cap_e) {
        w->cap_e = ne;
        w->edges = (cedge *)realloc(w->edges, (size_t)ne * sizeof(cedge));
        w->act = (const cedge **)realloc(w->act, (size_t)ne * sizeof(cedge *));
        w->xb = (double *)realloc(w->xb, (size_t)ne * sizeof(double));
        w->xt = (double *)realloc(w->xt, (size_t)ne * sizeof(double));
        w->xa = (double *)realloc(w->xa, (size_t)ne * sizeof(double));
        w->xc = (double *)realloc(w->xc, (size_t)ne * sizeof(double));
        w->km = (double *)realloc(w->km, (size_t)ne * sizeof(double));
        w->ord = (int *)realloc(w->ord, (size_t)ne * sizeof(int));
        /* crossings per beam are bounded by m*(m-1)/2 <= ne*ne/2 */
        w->cy = (double *)realloc(w->cy, ((size_t)ne * ne / 2 + 2) * sizeof(double));
    }
    if (ny > w->cap_y) {
        w->cap_y = ny;
        w->ys = (double *)realloc(w->ys, (size_t)ny * sizeof(double));
        w->bsum = (double *)realloc(w->bsum, (size_t)ny * sizeof(double));
    }
}

static void ws_free(clip_ws *w)
{
    free(w->edges); free(w->act); free(w->xb); free(w->xt); free(w->xa);
    free(w->xc); free(w->km); free(w->ord); free(w->cy); free(w->ys); free(w->bsum);
}

/* GPC's operations (gpc_op order DIFF, INT, XOR, UNION; OV:264-306, 492-525)
 * as the set a gap belongs to, from the subject and clip parities. */
static int op_in(int op, int pp, int pq)
{
    switch (op) {
    case 0: return pp && !pq;   /* DIFF */
    case 2: return pp != pq;    /* XOR */
    case 3: return pp || pq;    /* UNION */
    default: return pp && pq;   /* INT */
    }
}

/* The y range the sweep covers: the common range for INT (empty when the
 * boxes are disjoint, GPC's minimax pre-test OV:1656-1709), the subject's
 * range for DIFF, the union of the ranges for XOR and UNION.  A polygon of
 * fewer than 3 vertices is empty.  Returns 0 when the result is empty. */
static int op_range(int op, int eP, int eQ, const double *bp, const double *bq,
                    double *ylo, double *yhi)
{
    if (op == 1) {
        if (eP || eQ) return 0;
        if (bp[2] < bq[0] || bq[2] < bp[0] || bp[3] < bq[1] || bq[3] < bp[1]) return 0;
        *ylo = bp[1] > bq[1] ? bp[1] : bq[1];
        *yhi = bp[3] < bq[3] ? bp[3] : bq[3];
    } else if (op == 0 || eQ) {
        if (eP) return 0;
        *ylo = bp[1];
        *yhi = bp[3];
    } else if (eP) {
        *ylo = bq[1];
        *yhi = bq[3];
    } else {
        *ylo = bp[1] < bq[1] ? bp[1] : bq[1];
        *yhi = bp[3] > bq[3] ? bp[3] : bq[3];
    }
    return *ylo < *yhi;
}

/* One crossing-free sub-beam [ya, yc] with edge x's xa[], xc[] (indexed like
 * act[]).  Returns the doubled-area contribution, summed left to right. */
static double sub_beam(clip_ws *w, int m, double ya, double yc, double acc, int op,
                       uint64_t *fl)
{
    double h = yc - ya;
    for (int i = 0; i < m; ++i) w->km[i] = w->xa[i] + w->xc[i];
    for (int i = 0; i < m; ++i) w->ord[i] = i;
    for (int i = 1; i < m; ++i) {           /* insertion sort, m is small */
        int v = w->ord[i], j = i - 1;
        while (j >= 0 && sb_before(w->km, w->xa, w->act, v, w->ord[j])) {
            w->ord[j + 1] = w->ord[j];
            --j;
        }
        w->ord[j + 1] = v;
    }
    int pp = 0, pq = 0;
    uint64_t ops = 1 + (uint64_t)m;
    for (int t = 0; t + 1 < m; ++t) {
        int l = w->ord[t], r = w->ord[t + 1];
        if (w->act[l]->poly == 0) pp ^= 1; else pq ^= 1;
        if (op_in(op, pp, pq)) {
            acc += ((w->xa[r] - w->xa[l]) + (w->xc[r] - w->xc[l])) * h;
            ops += 5;
        }
    }
    if (fl) *fl += ops;
    return acc;
}

static double clip_area_ws(clip_ws *w, const double *P, int nP, const double *Q, int nQ,
                           int op, uint64_t *fl)
{
    const int eP = nP < 3, eQ = nQ < 3;
    double bp[4] = {0, 0, 0, 0}, bq[4] = {0, 0, 0, 0};
    if (!eP) bbox(P, nP, bp);
    if (!eQ) bbox(Q, nQ, bq);
    double ylo, yhi;
    if (!op_range(op, eP, eQ, bp, bq, &ylo, &yhi)) return 0.0;
    nP = eP ? 0 : nP;   /* an empty polygon contributes no edge and no y */
    nQ = eQ ? 0 : nQ;
    ws_reserve(w, nP + nQ, nP + nQ);
    uint64_t ops = 0;

    /* edge table */
    int ne = 0;
    for (int poly = 0; poly < 2; ++poly) {
        const double *V = poly ? Q : P;
        int n = poly ? nQ : nP;
        for (int i = 0; i < n; ++i) {
            int j = i + 1 == n ? 0 : i + 1;
            double x0 = V[2 * i], y0 = V[2 * i + 1], x1 = V[2 * j], y1 = V[2 * j + 1];
            if (y0 == y1) continue;
            cedge e;
            if (y0 < y1) { e.bx = x0; e.by = y0; e.tx = x1; e.ty = y1; }
            else { e.bx = x1; e.by = y1; e.tx = x0; e.ty = y0; }
            if (e.ty <= ylo || e.by >= yhi) continue;
            e.dx = (e.tx - e.bx) / (e.ty - e.by);
            ops += 3;
            e.id = poly ? nP + i : i;
            e.poly = poly;
            w->edges[ne++] = e;
        }
    }
    qsort(w->edges, (size_t)ne, sizeof(cedge), cmp_edge_by);

    /* scanbeam table: distinct vertex y's within [ylo, yhi] */
    int ny = 0;
    for (int poly = 0; poly < 2; ++poly) {
        const double *V = poly ? Q : P;
        int n = poly ? nQ : nP;
        for (int i = 0; i < n; ++i) {
            double y = V[2 * i + 1];
            if (y >= ylo && y <= yhi) w->ys[ny++] = y;
        }
    }
    qsort(w->ys, (size_t)ny, sizeof(double), cmp_double);
    int u = 0;
    for (int i = 0; i < ny; ++i)
        if (u == 0 || w->ys[i] != w->ys[u - 1]) w->ys[u++] = w->ys[i];
    ny = u;

    if (ny < 2) return 0.0;
    double *bsum = w->bsum;   /* beam sums; the scanbeam table is still needed */
    int m = 0, next = 0;
    for (int k = 0; k + 1 < ny; ++k) {
        double yb = w->ys[k], yt = w->ys[k + 1];
        bsum[k] = 0.0;
        while (next < ne && w->edges[next].by <= yb) w->act[m++] = &w->edges[next++];
        int keep = 0;
        for (int i = 0; i < m; ++i)
            if (w->act[i]->ty > yb) w->act[keep++] = w->act[i];
        m = keep;
        if (m < 2) continue;
        for (int i = 0; i < m; ++i) {
            w->xb[i] = edge_x(w->act[i], yb);
            w->xt[i] = edge_x(w->act[i], yt);
        }
        ops += 6 * (uint64_t)m;
        /* crossing heights inside (yb, yt) */
        int nc = 0;
        double dy = yt - yb;
        for (int i = 0; i < m; ++i) {
            for (int j = i + 1; j < m; ++j) {
                int L, R;
                if (w->xb[i] < w->xb[j] && w->xt[i] > w->xt[j]) { L = i; R = j; }
                else if (w->xb[j] < w->xb[i] && w->xt[j] > w->xt[i]) { L = j; R = i; }
                else continue;
                double den = (w->xt[L] - w->xb[L]) - (w->xt[R] - w->xb[R]);
                ops += 3;
                if (!(den > 0.0)) continue;
                double r = (w->xb[R] - w->xb[L]) / den;
                double iy = r * dy + yb;
                ops += 4;
                if (iy > yb && iy < yt) w->cy[nc++] = iy;
            }
        }
        if (nc == 0) {
            for (int i = 0; i < m; ++i) { w->xa[i] = w->xb[i]; w->xc[i] = w->xt[i]; }
            bsum[k] = sub_beam(w, m, yb, yt, 0.0, op, &ops);
            continue;
        }
        qsort(w->cy, (size_t)nc, sizeof(double), cmp_double);
        u = 0;
        for (int i = 0; i < nc; ++i)
            if (u == 0 || w->cy[i] != w->cy[u - 1]) w->cy[u++] = w->cy[i];
        nc = u;
        double ya = yb, bs = 0.0;
        for (int i = 0; i < m; ++i) w->xa[i] = w->xb[i];
        for (int c = 0; c <= nc; ++c) {
            double yc = c < nc ? w->cy[c] : yt;
            if (c < nc) {
                for (int i = 0; i < m; ++i) w->xc[i] = edge_x(w->act[i], yc);
                ops += 3 * (uint64_t)m;
            } else {
                for (int i = 0; i < m; ++i) w->xc[i] = w->xt[i];
            }
            bs = sub_beam(w, m, ya, yc, bs, op, &ops);
            for (int i = 0; i < m; ++i) w->xa[i] = w->xc[i];
            ya = yc;
        }
        bsum[k] = bs;
    }
    /* fixed pairwise tree over the beam sums */
    const int nb = ny - 1;
    for (int wdt = 1; wdt < nb; wdt *= 2)
        for (int i = 0; i + wdt < nb; i += 2 * wdt) bsum[i] += bsum[i + wdt];
    if (fl) *fl += ops + (uint64_t)nb;
    return 0.5 * bsum[0];
}

double orc_clip_area(const double *P, int nP, const double *Q, int nQ, uint64_t *flops)
{
    return orc_clip_area_op(P, nP, Q, nQ, 1, flops);
}

/* Area of P op Q for op = 0 DIFF, 1 INT, 2 XOR, 3 UNION: the same sweep with
 * the op's y range and gap predicate. */
double orc_clip_area_op(const double *P, int nP, const double *Q, int nQ, int op,
                        uint64_t *flops)
{
    clip_ws w;
    memset(&w, 0, sizeof w);
    double a = clip_area_ws(&w, P, nP, Q, nQ, op, flops);
    ws_free(&w);
    return a;
}

/* ---------------------------------------------------------------------- */
/* Result contours of P op Q (GPC's output polygon, OV:527-750)           */
/* ---------------------------------------------------------------------- */
/*
 * The reference keeps GPC's result contours (vertex lists with hole flags,
 * OV:720-728).  This restatement builds them from the same sweep as the
 * area, with its own assembly rule (GPC's vertex-class tables are not
 * followed; the vertex order is pinned by orientation, closed forms and the
 * areas, not against GPC):
 *   - inside intervals of a sub-beam: maximal runs of neighbouring gaps that
 *     are inside (op_in) and not of zero width (a gap whose two edges have
 *     equal x at both ends is skipped and does not end a run); the interval
 *     keeps its left and right edge and their x at both ends;
 *   - at every sub-beam boundary y the intervals below (x at their top) and
 *     above (x at their bottom) give 2 points each; the points of the two
 *     lists are merged by x (each list keeps its order; on equal x a right
 *     end goes before a left end, then the one below first) and paired
 *     consecutively: every point toggles one of the two inside states, so
 *     consecutive points bound exactly the horizontal parts of the boundary;
 *   - a pair is one of: the left (right) chain continuing upward, a local
 *     minimum (left-right above: new contour), a hole bottom (right-left
 *     above: new contour), a local maximum (left-right below) or a hole top
 *     (right-left below); the last two close a contour when both chain ends
 *     belong to it (outer at a maximum, hole at a hole top), else join the
 *     two contours;
 *   - a contour is a path head..tail: down its left chains, along its
 *     bottoms, up its right chains, so outer contours run counter-clockwise
 *     and holes clockwise (inside on the left);
 *   - vertices: a chain continuing on the same edge adds no vertex; otherwise
 *     the pair adds its two points in path order, or one when the two x's
 *     differ by no more than the sum of their uncertainties: 0 for an edge
 *     end (y is the edge's bottom or top), else
 *     2^-44 * (|bx| + |tx| + |dx| * (|y| + |by|)), which covers the rounding
 *     of x at a crossing height (a crossing is evaluated on each of its two
 *     edges, at a height that is itself rounded).  The one vertex is the
 *     point of smaller uncertainty; on a tie the point below for a chain,
 *     the left point otherwise.
 * Contours come out in closing order.
 */
typedef struct {
    int eL, eR;         /* edge ids */
    double xL, xR;      /* x of the edges at the scanline */
    double uL, uR;      /* their uncertainties */
    int lp, rp;         /* contour of the left / right chain */
} civ;

static double cb_unc(const cedge *e, double y)
{
    if (y == e->ty || y == e->by) return 0.0;
    return 0x1p-44 * (fabs(e->bx) + fabs(e->tx) + fabs(e->dx) * (fabs(y) + fabs(e->by)));
}

typedef struct {
    /* vertex arena and open contours */
    double *vx, *vy;
    int *vnext;
    int64_t nv, cap_v;
    int *phead, *ptail, *pcnt;
    int npoly, cap_p;
    /* interval lists */
    civ *B, *A;
    int nb, na, cap_i;
    /* output */
    double *xy;
    int64_t cap_xy, nout_v;
    int32_t *cnv, *chole;
    int64_t cap_c, nout_c;
} cbuild;

static int cb_vertex(cbuild *b, double x, double y)
{
    if (b->nv == b->cap_v) {
        b->cap_v = b->cap_v ? 2 * b->cap_v : 64;
        b->vx = (double *)realloc(b->vx, (size_t)b->cap_v * sizeof(double));
        b->vy = (double *)realloc(b->vy, (size_t)b->cap_v * sizeof(double));
        b->vnext = (int *)realloc(b->vnext, (size_t)b->cap_v * sizeof(int));
    }
    b->vx[b->nv] = x;
    b->vy[b->nv] = y;
    b->vnext[b->nv] = -1;
    return (int)b->nv++;
}

static int cb_poly(cbuild *b, int head, int tail, int cnt)
{
    if (b->npoly == b->cap_p) {
        b->cap_p = b->cap_p ? 2 * b->cap_p : 16;
        b->phead = (int *)realloc(b->phead, (size_t)b->cap_p * sizeof(int));
        b->ptail = (int *)realloc(b->ptail, (size_t)b->cap_p * sizeof(int));
        b->pcnt = (int *)realloc(b->pcnt, (size_t)b->cap_p * sizeof(int));
    }
    b->phead[b->npoly] = head;
    b->ptail[b->npoly] = tail;
    b->pcnt[b->npoly] = cnt;
    return b->npoly++;
}

static void cb_prepend(cbuild *b, int p, double x, double y)
{
    int v = cb_vertex(b, x, y);
    b->vnext[v] = b->phead[p];
    b->phead[p] = v;
    b->pcnt[p] += 1;
}

static void cb_append(cbuild *b, int p, double x, double y)
{
    int v = cb_vertex(b, x, y);
    b->vnext[b->ptail[p]] = v;
    b->ptail[p] = v;
    b->pcnt[p] += 1;
}

/* path of p, then path of q; every chain end of q now belongs to p */
static void cb_join(cbuild *b, int p, int q)
{
    b->vnext[b->ptail[p]] = b->phead[q];
    b->ptail[p] = b->ptail[q];
    b->pcnt[p] += b->pcnt[q];
    for (int k = 0; k < b->nb; ++k) {
        if (b->B[k].lp == q) b->B[k].lp = p;
        if (b->B[k].rp == q) b->B[k].rp = p;
    }
    for (int k = 0; k < b->na; ++k) {
        if (b->A[k].lp == q) b->A[k].lp = p;
        if (b->A[k].rp == q) b->A[k].rp = p;
    }
}

static void cb_close(cbuild *b, int p, int hole)
{
    if (b->nout_c < b->cap_c) {
        b->cnv[b->nout_c] = b->pcnt[p];
        b->chole[b->nout_c] = hole;
    }
    int v = b->phead[p];
    for (int k = 0; k < b->pcnt[p]; ++k) {
        if (b->nout_v < b->cap_xy) {
            b->xy[2 * b->nout_v] = b->vx[v];
            b->xy[2 * b->nout_v + 1] = b->vy[v];
        }
        b->nout_v++;
        v = b->vnext[v];
    }
    b->nout_c++;
}

typedef struct {
    int below, right, iv;
    double x, unc;
} cpt;

static inline int cb_near(cpt u, cpt v)
{
    return fabs(u.x - v.x) <= u.unc + v.unc;
}

static cpt cb_point(const cbuild *b, int below, int k)
{
    cpt p;
    const civ *I = below ? &b->B[k >> 1] : &b->A[k >> 1];
    p.below = below;
    p.right = k & 1;
    p.iv = k >> 1;
    p.x = p.right ? I->xR : I->xL;
    p.unc = p.right ? I->uR : I->uL;
    return p;
}

static void cb_pair(cbuild *b, double y, cpt u, cpt v)
{
    if (u.below != v.below) {                 /* a chain continues upward */
        cpt lo = u.below ? u : v, hi = u.below ? v : u;
        civ *I = &b->B[lo.iv], *J = &b->A[hi.iv];
        int same = lo.right ? (I->eR == J->eR) : (I->eL == J->eL);
        int p = lo.right ? I->rp : I->lp;
        if (!same) {
            const int two = !cb_near(lo, hi);
            const double x1 = (two || !(hi.unc < lo.unc)) ? lo.x : hi.x;
            if (lo.right) {
                cb_append(b, p, x1, y);
                if (two) cb_append(b, p, hi.x, y);
            } else {
                cb_prepend(b, p, x1, y);
                if (two) cb_prepend(b, p, hi.x, y);
            }
        }
        if (lo.right) J->rp = p; else J->lp = p;
        return;
    }
    const int two = !cb_near(u, v);
    const double x1 = v.unc < u.unc ? v.x : u.x;   /* the one vertex */
    if (!u.below) {                           /* a contour starts */
        int p;
        if (!u.right) {                       /* local minimum: u -> v */
            int h = cb_vertex(b, two ? u.x : x1, y), t = h;
            if (two) { t = cb_vertex(b, v.x, y); b->vnext[h] = t; }
            p = cb_poly(b, h, t, 1 + two);
            b->A[u.iv].lp = p;
            b->A[v.iv].rp = p;
        } else {                              /* hole bottom: v -> u */
            int h = cb_vertex(b, two ? v.x : x1, y), t = h;
            if (two) { t = cb_vertex(b, u.x, y); b->vnext[h] = t; }
            p = cb_poly(b, h, t, 1 + two);
            b->A[v.iv].lp = p;
            b->A[u.iv].rp = p;
        }
        return;
    }
    if (!u.right) {                           /* local maximum: tail -> v -> u -> head */
        int p = b->B[u.iv].lp, q = b->B[v.iv].rp;
        if (two) cb_append(b, q, v.x, y);
        cb_append(b, q, two ? u.x : x1, y);
        if (p == q) cb_close(b, q, 0);
        else cb_join(b, q, p);
    } else {                                  /* hole top: tail -> u -> v -> head */
        int p = b->B[u.iv].rp, q = b->B[v.iv].lp;
        cb_append(b, p, two ? u.x : x1, y);
        if (two) cb_append(b, p, v.x, y);
        if (p == q) cb_close(b, p, 1);
        else cb_join(b, p, q);
    }
}

/* intervals below (B, x at y) -> intervals above (A, x at y) */
static void cb_scanline(cbuild *b, double y)
{
    int i = 0, j = 0, have = 0;
    const int nbp = 2 * b->nb, nap = 2 * b->na;
    cpt pend = {0, 0, 0, 0.0, 0.0};
    while (i < nbp || j < nap) {
        cpt p;
        if (j >= nap) p = cb_point(b, 1, i++);
        else if (i >= nbp) p = cb_point(b, 0, j++);
        else {
            cpt pb = cb_point(b, 1, i), pa = cb_point(b, 0, j);
            int take_b;
            if (pb.x != pa.x) take_b = pb.x < pa.x;
            else take_b = pb.right || !pa.right;
            if (take_b) { p = pb; ++i; } else { p = pa; ++j; }
        }
        if (!have) { pend = p; have = 1; }
        else { cb_pair(b, y, pend, p); have = 0; }
    }
}

static void cb_reserve_iv(cbuild *b, int n)
{
    if (n > b->cap_i) {
        b->cap_i = n;
        b->B = (civ *)realloc(b->B, (size_t)n * sizeof(civ));
        b->A = (civ *)realloc(b->A, (size_t)n * sizeof(civ));
    }
}

static void cb_add_interval(cbuild *b, const clip_ws *w, int s, int e, double ya)
{
    civ *J = &b->A[b->na++];
    J->eL = w->act[s]->id;
    J->eR = w->act[e]->id;
    J->xL = w->xa[s];
    J->xR = w->xa[e];
    J->uL = cb_unc(w->act[s], ya);
    J->uR = cb_unc(w->act[e], ya);
    J->lp = J->rp = -1;
}

/* A <- inside intervals of the sub-beam (ord by sb_before), x at the bottom */
static void cb_intervals(cbuild *b, clip_ws *w, int m, int op, double ya)
{
    for (int i = 0; i < m; ++i) w->km[i] = w->xa[i] + w->xc[i];
    for (int i = 0; i < m; ++i) w->ord[i] = i;
    for (int i = 1; i < m; ++i) {
        int v = w->ord[i], j = i - 1;
        while (j >= 0 && sb_before(w->km, w->xa, w->act, v, w->ord[j])) {
            w->ord[j + 1] = w->ord[j];
            --j;
        }
        w->ord[j + 1] = v;
    }
    int pp = 0, pq = 0, open = 0, start = 0, last = 0;
    b->na = 0;
    for (int t = 0; t + 1 < m; ++t) {
        int l = w->ord[t], r = w->ord[t + 1];
        if (w->act[l]->poly == 0) pp ^= 1; else pq ^= 1;
        if (w->xa[l] == w->xa[r] && w->xc[l] == w->xc[r]) continue;
        int in = op_in(op, pp, pq);
        if (in) {
            if (!open) { open = 1; start = l; }
            last = r;
        } else if (open) {
            cb_add_interval(b, w, start, last, ya);
            open = 0;
        }
    }
    if (open) cb_add_interval(b, w, start, last, ya);
}

/* A (bottom x's) becomes B with the x's at the sub-beam top yc */
static void cb_advance(cbuild *b, const clip_ws *w, int m, double yc)
{
    civ *t = b->B;
    b->B = b->A;
    b->A = t;
    b->nb = b->na;
    b->na = 0;
    for (int k = 0; k < b->nb; ++k) {
        civ *I = &b->B[k];
        for (int i = 0; i < m; ++i) {
            if (w->act[i]->id == I->eL) { I->xL = w->xc[i]; I->uL = cb_unc(w->act[i], yc); }
            if (w->act[i]->id == I->eR) { I->xR = w->xc[i]; I->uR = cb_unc(w->act[i], yc); }
        }
    }
}

static void clip_contours_ws(clip_ws *w, cbuild *b, const double *P, int nP, const double *Q,
                             int nQ, int op)
{
    const int eP = nP < 3, eQ = nQ < 3;
    double bp[4] = {0, 0, 0, 0}, bq[4] = {0, 0, 0, 0};
    if (!eP) bbox(P, nP, bp);
    if (!eQ) bbox(Q, nQ, bq);
    double ylo, yhi;
    if (!op_range(op, eP, eQ, bp, bq, &ylo, &yhi)) return;
    nP = eP ? 0 : nP;
    nQ = eQ ? 0 : nQ;
    ws_reserve(w, nP + nQ, nP + nQ);
    cb_reserve_iv(b, (nP + nQ) / 2 + 1);

    int ne = 0;
    for (int poly = 0; poly < 2; ++poly) {
        const double *V = poly ? Q : P;
        int n = poly ? nQ : nP;
        for (int i = 0; i < n; ++i) {
            int j = i + 1 == n ? 0 : i + 1;
            double x0 = V[2 * i], y0 = V[2 * i + 1], x1 = V[2 * j], y1 = V[2 * j + 1];
            if (y0 == y1) continue;
            cedge e;
            if (y0 < y1) { e.bx = x0; e.by = y0; e.tx = x1; e.ty = y1; }
            else { e.bx = x1; e.by = y1; e.tx = x0; e.ty = y0; }
            if (e.ty <= ylo || e.by >= yhi) continue;
            e.dx = (e.tx - e.bx) / (e.ty - e.by);
            e.id = poly ? nP + i : i;
            e.poly = poly;
            w->edges[ne++] = e;
        }
    }
    qsort(w->edges, (size_t)ne, sizeof(cedge), cmp_edge_by);
    int ny = 0;
    for (int poly = 0; poly < 2; ++poly) {
        const double *V = poly ? Q : P;
        int n = poly ? nQ : nP;
        for (int i = 0; i < n; ++i) {
            double y = V[2 * i + 1];
            if (y >= ylo && y <= yhi) w->ys[ny++] = y;
        }
    }
    qsort(w->ys, (size_t)ny, sizeof(double), cmp_double);
    int u = 0;
    for (int i = 0; i < ny; ++i)
        if (u == 0 || w->ys[i] != w->ys[u - 1]) w->ys[u++] = w->ys[i];
    ny = u;
    if (ny < 2) return;

    b->nb = b->na = 0;
    int m = 0, next = 0;
    for (int k = 0; k + 1 < ny; ++k) {
        double yb = w->ys[k], yt = w->ys[k + 1];
        while (next < ne && w->edges[next].by <= yb) w->act[m++] = &w->edges[next++];
        int keep = 0;
        for (int i = 0; i < m; ++i)
            if (w->act[i]->ty > yb) w->act[keep++] = w->act[i];
        m = keep;
        if (m < 2) {                          /* nothing inside this beam */
            b->na = 0;
            cb_scanline(b, yb);
            b->nb = 0;
            continue;
        }
        for (int i = 0; i < m; ++i) {
            w->xb[i] = edge_x(w->act[i], yb);
            w->xt[i] = edge_x(w->act[i], yt);
        }
        int nc = 0;
        double dy = yt - yb;
        for (int i = 0; i < m; ++i) {
            for (int j = i + 1; j < m; ++j) {
                int L, R;
                if (w->xb[i] < w->xb[j] && w->xt[i] > w->xt[j]) { L = i; R = j; }
                else if (w->xb[j] < w->xb[i] && w->xt[j] > w->xt[i]) { L = j; R = i; }
                else continue;
                double den = (w->xt[L] - w->xb[L]) - (w->xt[R] - w->xb[R]);
                if (!(den > 0.0)) continue;
                double r = (w->xb[R] - w->xb[L]) / den;
                double iy = r * dy + yb;
                if (iy > yb && iy < yt) w->cy[nc++] = iy;
            }
        }
        qsort(w->cy, (size_t)nc, sizeof(double), cmp_double);
        u = 0;
        for (int i = 0; i < nc; ++i)
            if (u == 0 || w->cy[i] != w->cy[u - 1]) w->cy[u++] = w->cy[i];
        nc = u;
        double ya = yb;
        for (int i = 0; i < m; ++i) w->xa[i] = w->xb[i];
        for (int c = 0; c <= nc; ++c) {
            double yc = c < nc ? w->cy[c] : yt;
            if (c < nc) {
                for (int i = 0; i < m; ++i) w->xc[i] = edge_x(w->act[i], yc);
            } else {
                for (int i = 0; i < m; ++i) w->xc[i] = w->xt[i];
            }
            cb_intervals(b, w, m, op, ya);
            cb_scanline(b, ya);
            cb_advance(b, w, m, yc);
            for (int i = 0; i < m; ++i) w->xa[i] = w->xc[i];
            ya = yc;
        }
    }
    b->na = 0;
    cb_scanline(b, w->ys[ny - 1]);
    b->nb = 0;
}

/* Contours of P op Q: up to cap_v vertices into xy (interleaved), up to
 * cap_c contours into nv[] (vertex counts) and hole[]; *nvert / *ncont get
 * the totals (call with caps 0 to size the buffers). */
void orc_clip_contours(const double *P, int nP, const double *Q, int nQ, int op, double *xy,
                       int64_t cap_v, int32_t *nv, int32_t *hole, int64_t cap_c,
                       int64_t *nvert, int64_t *ncont)
{
    clip_ws w;
    memset(&w, 0, sizeof w);
    cbuild b;
    memset(&b, 0, sizeof b);
    b.xy = xy;
    b.cap_xy = xy ? cap_v : 0;
    b.cnv = nv;
    b.chole = hole;
    b.cap_c = (nv && hole) ? cap_c : 0;
    clip_contours_ws(&w, &b, P, nP, Q, nQ, op);
    *nvert = b.nout_v;
    *ncont = b.nout_c;
    ws_free(&w);
    free(b.vx); free(b.vy); free(b.vnext);
    free(b.phead); free(b.ptail); free(b.pcnt);
    free(b.B); free(b.A);
}

/* ---------------------------------------------------------------------- */
/* Independent checks of the area (different algorithms)                  */
/* ---------------------------------------------------------------------- */

/* Signed shoelace area, vertices taken relative to the first vertex. */
double orc_poly_area(const double *P, int n)
{
    if (n < 3) return 0.0;
    double ox = P[0], oy = P[1], s = 0.0;
    for (int i = 0; i < n; ++i) {
        int j = i + 1 == n ? 0 : i + 1;
        double xi = P[2 * i] - ox, yi = P[2 * i + 1] - oy;
        double xj = P[2 * j] - ox, yj = P[2 * j + 1] - oy;
        s += xi * yj - xj * yi;
    }
    return 0.5 * s;
}

/* Even-odd point in polygon (crossing number). */
static int point_in(const double *B, int nB, double px, double py, double ox, double oy)
{
    int in = 0;
    for (int i = 0; i < nB; ++i) {
        int j = i + 1 == nB ? 0 : i + 1;
        double xi = B[2 * i] - ox, yi = B[2 * i + 1] - oy;
        double xj = B[2 * j] - ox, yj = B[2 * j + 1] - oy;
        if ((yi > py) != (yj > py)) {
            double xc = (xj - xi) * (py - yi) / (yj - yi) + xi;
            if (px < xc) in ^= 1;
        }
    }
    return in;
}

/* Parameters t in (0,1) where edge a->b of A crosses the boundary of B,
 * sorted and bracketed by 0 and 1.  Coordinates relative to (ox, oy). */
static int edge_cuts(double ax, double ay, double bx, double by, const double *B, int nB,
                     double ox, double oy, double *ts)
{
    int nt = 0;
    ts[nt++] = 0.0;
    double rx = bx - ax, ry = by - ay;
    for (int i = 0; i < nB; ++i) {
        int j = i + 1 == nB ? 0 : i + 1;
        double cx = B[2 * i] - ox, cy = B[2 * i + 1] - oy;
        double sx = B[2 * j] - ox - cx, sy = B[2 * j + 1] - oy - cy;
        double den = rx * sy - ry * sx;
        if (den == 0.0) continue;
        double qx = cx - ax, qy = cy - ay;
        double t = (qx * sy - qy * sx) / den;
        double u = (qx * ry - qy * rx) / den;
        /* generous on u: a crossing at a vertex of B may be found twice (a
         * zero-length part, harmless) but is never lost to rounding */
        if (t > 0.0 && t < 1.0 && u >= -1e-9 && u <= 1.0 + 1e-9) ts[nt++] = t;
    }
    ts[nt++] = 1.0;
    qsort(ts + 1, (size_t)(nt - 2), sizeof(double), cmp_double);
    return nt;
}

/* Green's theorem: area(A cap B) = sum over the parts of dA inside B and the
 * parts of dB inside A of  integral x dy  (each polygon taken
 * counter-clockwise).  Inside-ness of each part is decided by a point test
 * at its midpoint. */
static double green_part(const double *A, int nA, const double *B, int nB,
                         double ox, double oy, double *ts)
{
    double s = 0.0;
    for (int i = 0; i < nA; ++i) {
        int j = i + 1 == nA ? 0 : i + 1;
        double ax = A[2 * i] - ox, ay = A[2 * i + 1] - oy;
        double bx = A[2 * j] - ox, by = A[2 * j + 1] - oy;
        int nt = edge_cuts(ax, ay, bx, by, B, nB, ox, oy, ts);
        for (int k = 0; k + 1 < nt; ++k) {
            double t0 = ts[k], t1 = ts[k + 1];
            if (!(t1 > t0)) continue;
            double tm = 0.5 * (t0 + t1);
            if (!point_in(B, nB, ax + tm * (bx - ax), ay + tm * (by - ay), ox, oy)) continue;
            double x0 = ax + t0 * (bx - ax), y0 = ay + t0 * (by - ay);
            double x1 = ax + t1 * (bx - ax), y1 = ay + t1 * (by - ay);
            s += 0.5 * (x0 + x1) * (y1 - y0);
        }
    }
    return s;
}

double orc_green_area(const double *P, int nP, const double *Q, int nQ)
{
    if (nP < 3 || nQ < 3) return 0.0;
    double bp[4];
    bbox(P, nP, bp);
    double ox = 0.5 * (bp[0] + bp[2]), oy = 0.5 * (bp[1] + bp[3]);
    double *ts = (double *)malloc((size_t)(nP + nQ + 2) * sizeof(double));
    double sp = orc_poly_area(P, nP) < 0 ? -1.0 : 1.0;
    double sq = orc_poly_area(Q, nQ) < 0 ? -1.0 : 1.0;
    double a = sp * green_part(P, nP, Q, nQ, ox, oy, ts) + sq * green_part(Q, nQ, P, nP, ox, oy, ts);
    free(ts);
    return a;
}

/* Sutherland-Hodgman: clip P by each edge half-plane of the convex polygon
 * Q (orientation detected), shoelace area of the result. */
double orc_sh_convex_area(const double *P, int nP, const double *Q, int nQ)
{
    if (nP < 3 || nQ < 3) return 0.0;
    int cap = 2 * (nP + nQ) + 8;
    double *a = (double *)malloc((size_t)cap * 2 * sizeof(double));
    double *b = (double *)malloc((size_t)cap * 2 * sizeof(double));
    double ox = Q[0], oy = Q[1];
    int na = nP;
    for (int i = 0; i < nP; ++i) { a[2 * i] = P[2 * i] - ox; a[2 * i + 1] = P[2 * i + 1] - oy; }
    double orient = orc_poly_area(Q, nQ) < 0 ? -1.0 : 1.0;
    for (int e = 0; e < nQ && na > 0; ++e) {
        int f = e + 1 == nQ ? 0 : e + 1;
        double cx = Q[2 * e] - ox, cy = Q[2 * e + 1] - oy;
        double dx = Q[2 * f] - ox - cx, dy = Q[2 * f + 1] - oy - cy;
        int nb = 0;
        for (int i = 0; i < na; ++i) {
            int j = i + 1 == na ? 0 : i + 1;
            double px = a[2 * i], py = a[2 * i + 1], qx = a[2 * j], qy = a[2 * j + 1];
            double sp = orient * (dx * (py - cy) - dy * (px - cx));
            double sq = orient * (dx * (qy - cy) - dy * (qx - cx));
            if (sp >= 0) { b[2 * nb] = px; b[2 * nb + 1] = py; ++nb; }
            if ((sp >= 0) != (sq >= 0)) {
                double t = sp / (sp - sq);
                b[2 * nb] = px + t * (qx - px);
                b[2 * nb + 1] = py + t * (qy - py);
                ++nb;
            }
        }
        double *t = a; a = b; b = t;
        na = nb;
    }
    double area = na >= 3 ? fabs(orc_poly_area(a, na)) : 0.0;
    free(a);
    free(b);
    return area;
}

/* ---------------------------------------------------------------------- */
/* Gradient of the intersection area with the clipping topology fixed     */
/* ---------------------------------------------------------------------- */
/*
 * dArea = integral over (dA inside B) of V.n ds, plus the same for B.  For
 * edge i of A (v_i -> v_{i+1}, direction d), the parts inside B form a set
 * I of t in [0,1]; with the outward normal N = s*(d.y, -d.x) (s = +1 for a
 * counter-clockwise polygon, -1 otherwise)
 *   dArea/dv_i     += N * integral_I (1 - t) dt
 *   dArea/dv_{i+1} += N * integral_I t dt
 * The reference has no gradient; this is the build's definition.
 */
static void grad_part(const double *A, int nA, const double *B, int nB, double *gA,
                      double ox, double oy, double *ts)
{
    double s = orc_poly_area(A, nA) < 0 ? -1.0 : 1.0;
    for (int i = 0; i < 2 * nA; ++i) gA[i] = 0.0;
    for (int i = 0; i < nA; ++i) {
        int j = i + 1 == nA ? 0 : i + 1;
        double ax = A[2 * i] - ox, ay = A[2 * i + 1] - oy;
        double bx = A[2 * j] - ox, by = A[2 * j + 1] - oy;
        int nt = edge_cuts(ax, ay, bx, by, B, nB, ox, oy, ts);
        double j0 = 0.0, j1 = 0.0;
        for (int k = 0; k + 1 < nt; ++k) {
            double t0 = ts[k], t1 = ts[k + 1];
            if (!(t1 > t0)) continue;
            double tm = 0.5 * (t0 + t1);
            if (!point_in(B, nB, ax + tm * (bx - ax), ay + tm * (by - ay), ox, oy)) continue;
            double sq = 0.5 * (t1 * t1 - t0 * t0);
            j0 += (t1 - t0) - sq;
            j1 += sq;
        }
        double nx = s * (by - ay), ny = -s * (bx - ax);
        gA[2 * i] += j0 * nx;
        gA[2 * i + 1] += j0 * ny;
        gA[2 * j] += j1 * nx;
        gA[2 * j + 1] += j1 * ny;
    }
}

void orc_area_grad(const double *P, int nP, const double *Q, int nQ, double *gP, double *gQ)
{
    if (nP < 3 || nQ < 3) {
        for (int i = 0; i < 2 * nP; ++i) gP[i] = 0.0;
        for (int i = 0; i < 2 * nQ; ++i) gQ[i] = 0.0;
        return;
    }
    double bp[4];
    bbox(P, nP, bp);
    double ox = 0.5 * (bp[0] + bp[2]), oy = 0.5 * (bp[1] + bp[3]);
    double *ts = (double *)malloc((size_t)(nP + nQ + 2) * sizeof(double));
    grad_part(P, nP, Q, nQ, gP, ox, oy, ts);
    grad_part(Q, nQ, P, nP, gQ, ox, oy, ts);
    free(ts);
}

/* ---------------------------------------------------------------------- */
/* Batch driver (the Execute loop of OV:81-113 on the CPU)                  */
/* ---------------------------------------------------------------------- */

typedef struct {
    const uint32_t *pairs;
    const double *xyA, *xyB;
    const int64_t *offA, *offB;
    double *area;
    int64_t lo, hi;
    int op;
    uint64_t flops;
} ov_job;

static void *ov_run(void *arg)
{
    ov_job *j = (ov_job *)arg;
    clip_ws w;
    memset(&w, 0, sizeof w);
    uint64_t fl = 0;
    for (int64_t t = j->lo; t < j->hi; ++t) {
        uint32_t a = j->pairs[2 * t], b = j->pairs[2 * t + 1];
        const double *P = j->xyA + 2 * j->offA[a];
        const double *Q = j->xyB + 2 * j->offB[b];
        int nP = (int)(j->offA[a + 1] - j->offA[a]);
        int nQ = (int)(j->offB[b + 1] - j->offB[b]);
        j->area[t] = clip_area_ws(&w, P, nP, Q, nQ, j->op, &fl);
    }
    j->flops = fl;
    ws_free(&w);
    return 0;
}

/* pairs[2t] indexes layer A (base), pairs[2t+1] layer B (overlay). */
void orc_overlay(const uint32_t *pairs, int64_t np, const double *xyA, const int64_t *offA,
                 const double *xyB, const int64_t *offB, double *area, uint64_t *flops,
                 int nthreads)
{
    orc_overlay_op(pairs, np, xyA, offA, xyB, offB, 1, area, flops, nthreads);
}

void orc_overlay_op(const uint32_t *pairs, int64_t np, const double *xyA, const int64_t *offA,
                    const double *xyB, const int64_t *offB, int op, double *area,
                    uint64_t *flops, int nthreads)
{
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    ov_job jobs[256];
    pthread_t th[256];
    for (int k = 0; k < nthreads; ++k) {
        jobs[k].pairs = pairs; jobs[k].xyA = xyA; jobs[k].xyB = xyB;
        jobs[k].offA = offA; jobs[k].offB = offB; jobs[k].area = area;
        jobs[k].lo = np * k / nthreads; jobs[k].hi = np * (k + 1) / nthreads;
        jobs[k].flops = 0;
        jobs[k].op = op;
    }
    if (nthreads == 1) ov_run(&jobs[0]);
    else {
        for (int k = 0; k < nthreads; ++k) pthread_create(&th[k], 0, ov_run, &jobs[k]);
        for (int k = 0; k < nthreads; ++k) pthread_join(th[k], 0);
    }
    uint64_t fl = 0;
    for (int k = 0; k < nthreads; ++k) fl += jobs[k].flops;
    if (flops) *flops = fl;
}
