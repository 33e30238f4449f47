/*
 * prt_oracle.h -- CPU restatement of the ParlllelRtree hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * library in parlllelrtree_amd/ is compared against.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * The product never links, imports or calls it.
 *
 * Every function cites the reference file:line it restates.  Abbreviations:
 *   BU = BotomUp-rtree.cu, TD = TopDown-rtree.cu,
 *   OV = vector_map_overlay_gpc_kernel.cu.
 *
 * Layouts
 *   rects  : float r4[4*n] = {xmin, ymin, xmax, ymax} per rect (the reference
 *            keeps rect[0..3] = x1,y1,x2,y2 with x1<=x2, BU:692-695).
 *   polys  : double xy[2*V] interleaved vertices, int64 offs[n+1] (CSR,
 *            OV:800-862 builds the same prefix layout).
 *   pairs  : uint32 pairs[2*np] = {q, r} per pair.
 *
 * Parity status per function (how each one is pinned) is listed in the
 * header of prt_oracle.c.
 */
#ifndef PRT_ORACLE_H
#define PRT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- Morton keys ---------------- */
uint32_t orc_quant15(float lo, float hi);
uint32_t orc_spread15(uint32_t v);
void orc_morton_center30(const float *r4, int64_t n, uint32_t *keys);
uint64_t orc_interleave_ref4(uint64_t a1, uint64_t a2, uint64_t a3, uint64_t a4);
void orc_morton_ref4d(const int64_t *r4, int64_t n, int64_t scale, uint64_t *keys);
int64_t orc_ref_scale(int64_t range);

/* ---------------- stable ordering ---------------- */
void orc_stable_order_u64(const uint64_t *keys, int64_t n, uint32_t *perm);
void orc_stable_order_u32(const uint32_t *keys, int64_t n, uint32_t *perm);

/* ---------------- bottom-up build ---------------- */
int orc_bu_levels(int64_t n, int f, int64_t *level_start, int64_t *level_count);
int orc_bu_build(const float *r4, int64_t n, int f, const uint32_t *keys,
                 uint32_t *perm, int32_t *node_count, int32_t *node_start,
                 int32_t *node_end, float *entry_mbr, int32_t *entry_ref);

int orc_bu_build_keys64(const float *r4, int64_t n, int f, const uint64_t *keys,
                        uint32_t *perm, int32_t *node_count, int32_t *node_start,
                        int32_t *node_end, float *entry_mbr, int32_t *entry_ref);

/* ---------------- top-down build ---------------- */
int orc_td_levels(int64_t n, int f, int64_t *level_offset);
int orc_td_build(const float *r4, int64_t n, int f, int32_t *slot_child,
                 float *slot_mbr, uint32_t *perm);

/* ---------------- spatial join ---------------- */
int orc_overlap(const float *a, const float *b);
int64_t orc_join_bu(int64_t num_nodes, int f, int64_t num_leaves,
                    const int32_t *node_count, const float *entry_mbr,
                    const int32_t *entry_ref, const float *q4, int64_t nq,
                    uint32_t *pairs, int64_t cap, int64_t *tested);
int64_t orc_join_td(int f, const int32_t *slot_child, const float *slot_mbr,
                    const uint32_t *perm, const float *q4, int64_t nq,
                    uint32_t *pairs, int64_t cap, int64_t *tested);
int64_t orc_join_brute(const float *r4, int64_t n, const float *q4, int64_t nq,
                       uint32_t *pairs, int64_t cap);

/* ---------------- search counters, tree quality ---------------- */
void orc_search_stats_bu(int64_t num_nodes, int f, int64_t num_leaves,
                         const int32_t *node_count, const float *entry_mbr,
                         const int32_t *entry_ref, const float *q4, int64_t nq,
                         int64_t *num, int64_t *count, int64_t *num1);
void orc_search_stats_td(int f, const int32_t *slot_child, const float *slot_mbr,
                         const float *q4, int64_t nq, int64_t *num, int64_t *count,
                         int64_t *num1);
int64_t orc_sibling_overlap_bu(int64_t num_nodes, int f, int64_t num_leaves,
                               const int32_t *node_count, const float *entry_mbr,
                               const int32_t *entry_ref);
int64_t orc_sibling_overlap_td(int f, const int32_t *slot_child, const float *slot_mbr);

/* ---------------- polygon intersection area ---------------- */
double orc_clip_area(const double *P, int nP, const double *Q, int nQ,
                     uint64_t *flops);
double orc_clip_area_op(const double *P, int nP, const double *Q, int nQ, int op,
                        uint64_t *flops);
void orc_clip_contours(const double *P, int nP, const double *Q, int nQ, int op, double *xy,
                       int64_t cap_v, int32_t *nv, int32_t *hole, int64_t cap_c,
                       int64_t *nvert, int64_t *ncont);
double orc_poly_area(const double *P, int n);
double orc_green_area(const double *P, int nP, const double *Q, int nQ);
double orc_sh_convex_area(const double *P, int nP, const double *Q, int nQ);
void orc_area_grad(const double *P, int nP, const double *Q, int nQ,
                   double *gP, double *gQ);
void orc_overlay(const uint32_t *pairs, int64_t np, const double *xyA,
                 const int64_t *offA, const double *xyB, const int64_t *offB,
                 double *area, uint64_t *flops, int nthreads);
void orc_overlay_op(const uint32_t *pairs, int64_t np, const double *xyA,
                    const int64_t *offA, const double *xyB, const int64_t *offB, int op,
                    double *area, uint64_t *flops, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
