/*
 * prt.h -- C-ABI of the MI355X (gfx950) R-tree build / spatial join /
 * polygon overlay library (libprt.so).
 *
 * Every compute entry point takes DEVICE pointers (suffix _h marks host
 * pointers), a caller-owned workspace and a hipStream_t passed as void*.
 * Calls are stream-ordered and never synchronise the device, except the
 * reference-shaped shim prt_run(), which synchronises like the reference.
 * No entry point allocates device memory except prt_run() and prt_read_data();
 * prt_pipeline_capture allocates the host-side graph objects it returns;
 * the batched overlay creates, once per device, two helper streams (shared
 * by every call on that device: capture one batched overlay at a time per
 * device).
 * Return value: 0 (PRT_OK) or a negative PRT_E_* code; prt_last_error()
 * returns a message for the last failure on the calling thread.
 *
 * Reference interfaces replaced (BU = BotomUp-rtree.cu, TD =
 * TopDown-rtree.cu, OV = vector_map_overlay_gpc_kernel.cu):
 *   prt_morton_center30 / prt_morton_ref4d  <- CalcuateZOrderValue BU:73-94,
 *                                              Interleave BU:98-123
 *   prt_bu_build                            <- LocalSort/scan1/scan2/Reorder
 *                                              BU:125-266 + Construction_init/
 *                                              Construction1 BU:373-426, driven
 *                                              by main BU:836-869
 *   prt_td_build                            <- changeX/changeY TD:465-481, the
 *                                              radix kernels TD:162-261, fun/
 *                                              fun1/fun2 TD:309-353, driven by
 *                                              main TD:629-679
 *   prt_join_bu / prt_join_td               <- init/query1 BU:592-630,
 *                                              TD:355-377, 450-463 (with the
 *                                              hit semantics of RtreeSearch
 *                                              BU:439-483, TD:56-76)
 *   prt_search_stats_* / prt_sibling_overlap_*
 *                                           <- RtreeSearch / RtreeSearch1
 *                                              BU:439-483, 574-590, TD:56-93
 *   prt_overlay_plain                       <- Execute OV:81-113 +
 *                                              gpc_polygon_clip OV:115-750
 *   prt_overlay_total                       <- count1 of printfOutput OV:45-58
 *   prt_overlay_contours_count / _write     <- gpc_polygon_clip's result
 *                                              polygon OV:527-750, copied
 *                                              out OV:720-728
 *   prt_pipeline_*                          <- the stage sequence of BU:836-987
 *                                              / TD:629-758 feeding run()
 *                                              OV:61-79 (no single reference
 *                                              function: its stages are
 *                                              separate programs)
 */
#ifndef PRT_H
#define PRT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PRT_OK 0
#define PRT_E_INVAL (-1)
#define PRT_E_NOMEM (-2)
#define PRT_E_HIP (-3)
#define PRT_E_OVERFLOW (-4)
#define PRT_E_UNSUPPORTED (-5)
#define PRT_E_DEVICE (-6)     /* a kernel raised a PRT_STATUS_* flag (prt_device_status) */

/* Device status flags (prt_device_status).  Kernels raise them when a
 * result they wrote cannot be trusted; the entry points that launched them
 * have already returned, so the flags are read by prt_device_status. */
#define PRT_STATUS_LOOKBACK 1 /* a radix sort's look-back gave up waiting for an
                                 earlier tile: the sorted order is invalid   */
#define PRT_STATUS_INEXACT 2  /* prt_bu_build_ref4d rounded int64 coordinates
                                 that fp32 cannot hold outward: joins over
                                 that tree return a superset of the exact
                                 int64 Overlap hit set (BU:496-506)          */

/* Polygon set operations, in the order of GPC's gpc_op (OV:264-306,
 * 492-525): the area of A - B, A & B, A ^ B, A | B. */
#define PRT_OP_DIFF 0
#define PRT_OP_INT 1
#define PRT_OP_XOR 2
#define PRT_OP_UNION 3

#define PRT_MAX_LEVELS 40
#define PRT_MAX_FANOUT 64
#define PRT_MAX_POLY_VERTS 4096
#define PRT_MAX_LAYERS 4      /* layers of one prt_td_build_layers call */

/* Bottom-up tree, SoA.  Node ids run level by level, leaves first, root
 * last (BU:788-830).  Entry slot j of node i lives at i*fanout + j. */
typedef struct prt_bu_tree {
    int32_t fanout;
    int32_t num_levels;
    int64_t num_rects;
    int64_t num_nodes;
    int64_t level_start[PRT_MAX_LEVELS];
    int64_t level_count[PRT_MAX_LEVELS];
    /* device arrays, caller allocated */
    uint32_t *keys;       /* [num_rects] Morton keys in sorted order          */
    uint32_t *perm;       /* [num_rects] leaf order: original rect ids        */
    int32_t *node_count;  /* [num_nodes] entries used                        */
    int32_t *node_start;  /* [num_nodes] first sorted position covered       */
    int32_t *node_end;    /* [num_nodes] last sorted position covered        */
    float *entry_mbr;     /* [num_nodes*fanout*4] xmin,ymin,xmax,ymax;
                             unused slots (+inf,+inf,-inf,-inf)              */
    int32_t *entry_ref;   /* [num_nodes*fanout] leaf: rect id; else child
                             node id; unused: -1                             */
} prt_bu_tree;

/* Top-down tree: implicit slot layout of TD:504-524. */
typedef struct prt_td_tree {
    int32_t fanout;
    int32_t num_levels;
    int64_t num_rects;
    int64_t num_slots;
    int64_t level_offset[PRT_MAX_LEVELS + 1];
    int32_t *slot_child;  /* [num_slots] >0 child block, <0 leaf -pos-1, 0 empty */
    float *slot_mbr;      /* [num_slots*4]; empty slots (+inf,+inf,-inf,-inf)   */
    uint32_t *perm;       /* [num_rects] original id at final sorted position   */
} prt_td_tree;

/* Autotuned launch shape of the batched overlay, per vertex bucket. */
#define PRT_TUNE_BUCKETS 6
typedef struct prt_tune {
    int32_t bucket_max_verts[PRT_TUNE_BUCKETS]; /* upper bound of n1+n2 per bucket */
    int32_t lanes_per_pair[PRT_TUNE_BUCKETS];   /* 8, 16, 32, 64 (a slice of one
                                                   wave), or 128 / 256 (a block of 2
                                                   or 4 whole waves per pair)       */
    int32_t waves_per_block[PRT_TUNE_BUCKETS];  /* 1..8; 1 when lanes_per_pair > 64 */
    int32_t list_factor[PRT_TUNE_BUCKETS];      /* beam-list capacity per team, in
                                                   units of the bucket's vertex
                                                   bound: 1..8                      */
} prt_tune;

const char *prt_version(void);
const char *prt_last_error(void);

/* Waits for `stream`, then reads this device's status flags (PRT_STATUS_*
 * bits, OR-ed by every kernel that raised one since the last clear) into
 * *flags_h (optional) and clears them if `clear`.  Returns PRT_OK when no
 * flag is raised, PRT_E_DEVICE (message in prt_last_error) otherwise.  No
 * reference counterpart: the reference checks nothing on the device
 * (checkErrors after run(), OV:78, sees only launch errors). */
int prt_device_status(int32_t *flags_h, int clear, void *stream);

/* Test hook of the radix sort's look-back: polls before it gives up
 * (spin_limit < 0: the default, about a second), and one tile (mute_tile >=
 * 0) that publishes no look-back word in any pass, so that the tiles after
 * it give up and raise PRT_STATUS_LOOKBACK while every kernel still ends.
 * (-1, -1) restores the defaults.  No reference counterpart. */
int prt_debug_lookback(int32_t spin_limit, int64_t mute_tile);

/* Test hook of the contour count pass: one_lane = 1 sends every pair to the
 * one-lane-per-pair kernel instead of the team kernels; tune (null: the
 * defaults) sets the team shapes per vertex-count bucket.  No reference
 * counterpart. */
int prt_debug_contours(int32_t one_lane, const prt_tune *tune);

/* ---- optional per-kernel timing (HIP events around every launch) ----
 * No reference counterpart (the reference times whole phases with
 * cudaEvent, BU:836-874).  prt_profile_read aggregates per kernel name:
 * names = nul-separated, counts[k] launches, total_ms[k] summed duration;
 * returns the number of names and clears the record. */
void prt_profile_enable(int on);
int prt_profile_read(char *names, int names_bytes, int64_t *counts, double *total_ms,
                     int max_kernels);

/* ---- keys ---------------------------------------------------------- */
int prt_morton_center30(const float *rects, int64_t n, uint32_t *keys, void *stream);
int prt_morton_ref4d(const int64_t *rects, int64_t n, int64_t scale, uint64_t *keys,
                     void *stream);

/* ---- stable radix sort of (key, index) ------------------------------ */
size_t prt_sort_workspace_size(int64_t n);
int prt_sort_pairs_u32(const uint32_t *keys_in, int64_t n, int key_bits,
                       uint32_t *keys_out, uint32_t *idx_out, void *ws, size_t ws_bytes,
                       void *stream);

/* ---- bottom-up build ------------------------------------------------ */
int prt_bu_plan(int64_t n, int fanout, prt_bu_tree *tree);
size_t prt_bu_workspace_size(int64_t n, int fanout);
int prt_bu_build(const float *rects, int64_t n, int fanout, prt_bu_tree *tree,
                 void *ws, size_t ws_bytes, void *stream);

/* Bottom-up build in the reference's own key order (mode REF4D64 of
 * CalcuateZOrderValue/Interleave BU:73-123, sorted BU:841-849): rects_i64 =
 * int64 (x1, y1, x2, y2) per rect (device), keys = Interleave of rect/scale,
 * leaf order stable by (64-bit key, id).  Node boxes are the rects in fp32,
 * rounded outward (exact for |v| <= 2^24); a coordinate fp32 cannot hold
 * raises PRT_STATUS_INEXACT (prt_device_status).  tree->keys gets the high
 * word of each sorted key; keys_sorted (optional, device, n) the full keys. */
size_t prt_bu_ref4d_workspace_size(int64_t n, int fanout);
int prt_bu_build_ref4d(const int64_t *rects_i64, int64_t n, int fanout, int64_t scale,
                       prt_bu_tree *tree, uint64_t *keys_sorted, void *ws, size_t ws_bytes,
                       void *stream);

/* ---- top-down build ------------------------------------------------- */
int prt_td_plan(int64_t n, int fanout, prt_td_tree *tree);
size_t prt_td_workspace_size(int64_t n, int fanout);
int prt_td_build(const float *rects, int64_t n, int fanout, prt_td_tree *tree,
                 void *ws, size_t ws_bytes, void *stream);
/* Top-down trees of 1..PRT_MAX_LAYERS layers of n rects each (rects[l]: the
 * device rects of layer l; trees[l]: its tree, output arrays set as for
 * prt_td_build), built by one shared launch sequence: every sort, subtree
 * and top launch covers all layers, so one launch holds num_layers times the
 * work.  trees[l] is exactly what prt_td_build gives for rects[l].  No
 * counterpart in the reference, which builds one layer per call
 * (TopDown-rtree.cu:465-481 per layer); an overlay builds two. */
size_t prt_td_layers_workspace_size(int num_layers, int64_t n, int fanout);
int prt_td_build_layers(int num_layers, const float *const *rects, int64_t n, int fanout,
                        prt_td_tree *trees, void *ws, size_t ws_bytes, void *stream);

/* ---- join: pairs (q, r) sorted by (q, r); *count_d gets the total,
 *      only the first cap pairs are written --------------------------- */
size_t prt_join_workspace_size(int64_t nq);
int prt_join_bu(const prt_bu_tree *tree, const float *queries, int64_t nq,
                uint32_t *pairs, int64_t cap, int64_t *count_d, void *ws,
                size_t ws_bytes, void *stream);
int prt_join_td(const prt_td_tree *tree, const float *queries, int64_t nq,
                uint32_t *pairs, int64_t cap, int64_t *count_d, void *ws,
                size_t ws_bytes, void *stream);

/* prt_join_bu with no allocation and no workspace argument: for callers that
 * cannot allocate (code under graph capture, serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  The workspace is a
 * fixed device array of the library (15 MiB per device, held until the
 * process ends), so nq is at most 262,144 (PRT_E_UNSUPPORTED beyond).  Every
 * call on a device shares that array: calls that overlap on one device
 * (two streams, two host threads) race. */
int prt_join_bu_noalloc(const prt_bu_tree *tree, const float *queries, int64_t nq,
                        uint32_t *pairs, int64_t cap, int64_t *count_d, void *stream);

/* ---- the reference's search statistics --------------------------------
 * Counters of the reference's CPU search over every query, as its main()
 * prints them (BU:990-1027, TD:760-798): stats[0..5] = sum and max of
 * `num` (entries tested), sum and max of the search's return value, sum and
 * max of `num1` (leaf hits).  Replaces RtreeSearch BU:439-483 / TD:56-76.
 * prt_sibling_overlap_*: *count = pairs of overlapping entries that share a
 * node (RtreeSearch1 BU:574-590 / TD:78-93).  Outputs are device int64. */
int prt_search_stats_bu(const prt_bu_tree *tree, const float *queries, int64_t nq,
                        int64_t *stats, void *stream);
int prt_search_stats_td(const prt_td_tree *tree, const float *queries, int64_t nq,
                        int64_t *stats, void *stream);
int prt_sibling_overlap_bu(const prt_bu_tree *tree, int64_t *count, void *stream);
int prt_sibling_overlap_td(const prt_td_tree *tree, int64_t *count, void *stream);

/* ---- overlay: area of A[pairs[2t]] intersect B[pairs[2t+1]], FP64 ----
 * Polygons are CSR: xy = interleaved (x, y) doubles, off[i]..off[i+1]. */
size_t prt_overlay_plain_workspace_size(int64_t npairs, int32_t max_verts);
int prt_overlay_plain(const uint32_t *pairs, int64_t npairs, const double *xyA,
                      const int64_t *offA, const double *xyB, const int64_t *offB,
                      int32_t max_verts, double *area, void *ws, size_t ws_bytes,
                      void *stream);

/* Area of A[pairs[2t]] op B[pairs[2t+1]] for the other GPC operations
 * (op = PRT_OP_DIFF / INT / XOR / UNION; a polygon of fewer than 3
 * vertices is empty): the same sweep with the operation's y range and gap
 * predicate.  No reference counterpart beyond gpc_op (the reference only
 * calls GPC_INT, OV:111). */
int prt_overlay_plain_op(const uint32_t *pairs, int64_t npairs, const double *xyA,
                         const int64_t *offA, const double *xyB, const int64_t *offB,
                         int32_t max_verts, int32_t op, double *area, void *ws, size_t ws_bytes,
                         void *stream);

/* Batched overlay: same areas as prt_overlay_plain (bit-identical), pairs
 * bucketed by n1+n2 and run by persistent work-stealing launches, one per
 * bucket, with the team shape of `tune` (NULL = prt_tune_default).  Also
 * replaces Execute OV:81-113 / gpc_polygon_clip OV:115-750. */
int prt_tune_default(prt_tune *tune);
size_t prt_overlay_batched_workspace_size(int64_t npairs, int32_t max_verts);
int prt_overlay_batched(const uint32_t *pairs, int64_t npairs, const double *xyA,
                        const int64_t *offA, const double *xyB, const int64_t *offB,
                        int32_t max_verts, const prt_tune *tune, double *area, void *ws,
                        size_t ws_bytes, void *stream);
int prt_overlay_batched_op(const uint32_t *pairs, int64_t npairs, const double *xyA,
                           const int64_t *offA, const double *xyB, const int64_t *offB,
                           int32_t max_verts, int32_t op, const prt_tune *tune, double *area,
                           void *ws, size_t ws_bytes, void *stream);
/* Pairs of the last prt_overlay_batched call on workspace `ws` that went to
 * the plain path because they did not fit the team's LDS (synchronises). */
int prt_overlay_batched_overflow(const void *ws, size_t ws_bytes, int64_t npairs,
                                 int32_t max_verts, int64_t *count_h, void *stream);
/* ---- differentiable overlay (no reference counterpart: the reference
 * computes no area) ----------------------------------------------------
 * Accumulates into gA / gB (caller zeroes them) weights[t] * dArea_t /
 * d(x, y) of every vertex of A[pairs[2t]] and B[pairs[2t+1]], with the
 * clipping topology fixed (weights NULL = all ones); green_area[t] (if not
 * NULL) gets the area by Green's theorem from the same parts.
 * max_poly_verts = the largest vertex count of any single polygon. */
int prt_overlay_area_grad(const uint32_t *pairs, int64_t npairs, const double *xyA,
                          const int64_t *offA, const double *xyB, const int64_t *offB,
                          int32_t max_poly_verts, const double *weights, double *gA, double *gB,
                          double *green_area, void *stream);
int prt_overlay_area_grad_f32(const uint32_t *pairs, int64_t npairs, const float *xyA,
                              const int64_t *offA, const float *xyB, const int64_t *offB,
                              int32_t max_poly_verts, const float *weights, float *gA, float *gB,
                              float *green_area, void *stream);

/* As prt_overlay_batched for min(*npairs_d, cap) pairs, the count read on
 * the device (written there by a join): no host round trip, so the call can
 * follow a join on one stream and be captured with it into a graph.
 * Workspace: prt_overlay_batched_workspace_size(cap, max_verts). */
int prt_overlay_batched_dev(const uint32_t *pairs, const int64_t *npairs_d, int64_t cap,
                            const double *xyA, const int64_t *offA, const double *xyB,
                            const int64_t *offB, int32_t max_verts, const prt_tune *tune,
                            double *area, void *ws, size_t ws_bytes, void *stream);

/* Sum of area[0 .. n) and the number of entries > 0 (the reference's count1
 * of non-empty results, OV:45-58), n = min(*npairs_d, cap) or cap when
 * npairs_d is NULL.  Deterministic (fixed reduction order for a given n).
 * total / nonempty may each be NULL. */
size_t prt_overlay_total_workspace_size(void);
int prt_overlay_total(const double *area, const int64_t *npairs_d, int64_t cap, double *total,
                      int64_t *nonempty, void *ws, size_t ws_bytes, void *stream);

/* fp32 MBRs (xmin, ymin, xmax, ymax) of FP64 CSR polygons, each bound
 * rounded outward so the box contains the polygon (the reference builds its
 * tree on (int)(v*1e5) boxes, OV:805-808). */
int prt_poly_mbr_dev(const double *xy, const int64_t *off, int64_t n, float *mbr, void *stream);

/* ---- result contours of A[pairs[2t]] op B[pairs[2t+1]] ---------------
 * GPC's output polygon (gpc_polygon_clip's result, OV:527-750, copied out
 * OV:720-728), in two passes.  prt_overlay_contours_count writes per pair
 * the number of contours and of vertices (-1 in both for a pair of more
 * than max_verts vertices).  The caller scans them into cont_off / vert_off
 * (np + 1 exclusive offsets, device) and calls prt_overlay_contours_write,
 * which writes per contour its first vertex (cont_start, absolute index
 * into xy) and hole flag, and the vertices (xy, interleaved doubles).
 * Outer contours run counter-clockwise, holes clockwise; contours of a pair
 * come out in the order the sweep closes them.  *status (device int32) is
 * set to 1 when a pair's output did not match its counts.  The workspace of
 * the count pass is prt_overlay_contours_workspace_size(np, max_verts, 0).
 * The vertex order is this library's own (see the kernel's header), not
 * GPC's.
 *
 * stage (device, optional, NULL for none): the count pass keeps there the
 * vertices in a pool of P slots, taken 8 at a time by the pairs, P the
 * largest value (< 2^31) with prt_overlay_contours_stage_size(np, P) <=
 * stage_bytes (a stage too small for P = 8 is an error).  The write pass,
 * given the same stage unchanged, copies the kept pairs out and sweeps again
 * only the pairs that did not fit, so the two passes must see the same
 * pairs, polygons and op (the count pass stamps the stage with its
 * arguments; a write pass with other arguments, or given a stage the count
 * pass did not fill, sets *status).  The output is the same with or without
 * a stage;
 * a pass takes as long as its slowest pair, so size P for every pair to
 * fit (about the total vertex count plus 4 per pair). */
size_t prt_overlay_contours_workspace_size(int64_t npairs, int32_t max_verts,
                                           int64_t total_vertices);
size_t prt_overlay_contours_stage_size(int64_t npairs, int64_t pool_verts);
int prt_overlay_contours_count(const uint32_t *pairs, int64_t npairs, const double *xyA,
                               const int64_t *offA, const double *xyB, const int64_t *offB,
                               int32_t max_verts, int32_t op, int64_t *ncont, int64_t *nvert,
                               void *stage, size_t stage_bytes, void *ws, size_t ws_bytes,
                               void *stream);
int prt_overlay_contours_write(const uint32_t *pairs, int64_t npairs, const double *xyA,
                               const int64_t *offA, const double *xyB, const int64_t *offB,
                               int32_t max_verts, int32_t op, const int64_t *cont_off,
                               const int64_t *vert_off, int64_t total_vertices,
                               int64_t *cont_start, int32_t *cont_hole, double *xy,
                               int32_t *status, const void *stage, size_t stage_bytes, void *ws,
                               size_t ws_bytes, void *stream);

/* ---- overlay pipeline: MBRs of A and B -> R-tree over B's MBRs -> join
 * with A's MBRs -> batched FP64 clip of every pair -> total area ----------
 * Pairs come out as (A id, B id) sorted by (A id, B id); count[0] = pairs
 * found (only the first cap are kept and clipped: count[0] > cap means the
 * capacity was too small), count[1] = 0 (reserved).  area[t] is the FP64 area of
 * pair t; total / nonempty (optional) get the sum and the count of area > 0.
 * prt_pipeline_run enqueues the sequence; prt_pipeline_capture records it
 * once into a hipGraph on `stream` (a created stream, not the null stream)
 * and prt_pipeline_launch replays it on any stream.  The graph holds the
 * pointers of io and ws: replays read whatever those buffers then hold. */
typedef struct prt_pipeline_io {
    const double *xyA;
    const int64_t *offA;
    int64_t nA;
    const double *xyB;
    const int64_t *offB;
    int64_t nB;
    int32_t max_verts; /* >= largest (n_A + n_B) vertex count of any pair */
    int32_t fanout;    /* 2..64 */
    int32_t tree;      /* 0 bottom-up, 1 top-down */
    int32_t reserved;
    int64_t cap;       /* pair capacity */
    uint32_t *pairs;   /* [2*cap] */
    double *area;      /* [cap] */
    int64_t *count;    /* [2] */
    double *total;     /* [1] or NULL */
    int64_t *nonempty; /* [1] or NULL */
} prt_pipeline_io;
typedef struct prt_pipeline prt_pipeline;
size_t prt_pipeline_workspace_size(const prt_pipeline_io *io);
int prt_pipeline_run(const prt_pipeline_io *io, const prt_tune *tune, void *ws, size_t ws_bytes,
                     void *stream);
int prt_pipeline_capture(const prt_pipeline_io *io, const prt_tune *tune, void *ws,
                         size_t ws_bytes, void *stream, prt_pipeline **out);
int prt_pipeline_launch(prt_pipeline *p, void *stream);
int prt_pipeline_destroy(prt_pipeline *p);

/* ---- the reference's text formats (host memory, no device work) -------
 * Rect files (BU:668-691, TD:486-497): n, then n records "x1 x2 y1 y2 key";
 * xyxy_h gets (x1, y1, x2, y2) per rect.  Polygon files (OV:766-818): n,
 * then per polygon "num hole", a box of four doubles, num vertices "x y";
 * off_h[n + 1] CSR offsets, xy_h[2 * nverts], hole_h[n], box_h[4 * n] =
 * (int)(v * 100000) of the box values, min_xy_h[2] = smallest box[0] and
 * box[1] (OV:805-814).  Readers: call with the arrays NULL for the sizes,
 * then with arrays of those sizes. */
int prt_read_rects_text(const char *path, int64_t *n, int64_t *xyxy_h, int32_t *key_h);
int prt_write_rects_text(const char *path, int64_t n, const int64_t *xyxy_h,
                         const int32_t *key_h);
/* BU:682-721: shift rects by the smallest x1 / y1 (in place) and find the
 * key scale 10^k (range / 10^k <= 1e4, integer division). */
int prt_rects_normalize(int64_t *xyxy_h, int64_t n, int64_t *min_x, int64_t *min_y,
                        int64_t *scale);
int prt_read_polys_text(const char *path, int64_t *n, int64_t *nverts, int64_t *off_h,
                        double *xy_h, int32_t *hole_h, int32_t *box_h, int32_t *min_xy_h);
int prt_write_polys_text(const char *path, int64_t n, const int64_t *off_h, const double *xy_h,
                         const int32_t *hole_h);

/* ---- reference-shaped reader: readData, OV:752-888 ---------------------
 * The same argument list as the reference's readData (gpc_vertex arrays as
 * interleaved {x, y} doubles, long as int64): reads the base and overlay
 * polygon files; rect1..4 / rect*_query = (int)(v * 100000) of each
 * polygon's box, num_* vertex counts, prefix_* first-vertex offsets, hole_*
 * flags (host arrays from malloc: release with free()); vertex_* = device
 * arrays from hipMalloc (release with hipFree).  *min_x / *min_y are lowered
 * to the smallest base box values (the caller initialises them, OV:805-810).
 * Unlike the reference it returns a status instead of exiting. */
int prt_read_data(const char *base_filename, const char *overlay_filename, int32_t **rect1,
                  int32_t **rect2, int32_t **rect3, int32_t **rect4, int32_t *num_contours,
                  int32_t **rect1_query, int32_t **rect2_query, int32_t **rect3_query,
                  int32_t **rect4_query, int32_t *num_querys, int32_t **num_base,
                  int64_t **prefix_base, double **vertex_base, int32_t **num_overlay,
                  int64_t **prefix_overlay, double **vertex_overlay, int32_t *min_x,
                  int32_t *min_y, int32_t **hole_base, int32_t **hole_overlay);

/* ---- reference-shaped clip entry: run(), OV:61-79 ----------------------
 * Record t pairs base polygon zorder_index[pairs[t].nodeid - offset] (the
 * Z-order position -> id map, OV:89; a bottom-up tree's perm array) with
 * overlay polygon pairs[t].queryid.  Polygons as the reference's readData
 * holds them: num[] counts, prefix[] (running sums of num) and {x, y}
 * double vertices; hole flags are accepted and do not change the area.
 * area (optional, device) gets the FP64 area per record; *count1_h the
 * number of non-empty intersections (printfOutput OV:45-58).  Allocates
 * device memory and synchronises, like the reference.  The layout of the
 * reference's rtree_search record is not in the reference (its common.h is
 * absent): this one is {nodeid, queryid}. */
typedef struct prt_rtree_search {
    int32_t nodeid;
    int32_t queryid;
} prt_rtree_search;
int prt_run(const prt_rtree_search *pairs, int32_t offset, const int32_t *num_base,
            const int64_t *prefix_base, const double *vertex_base, const int32_t *num_overlay,
            const int64_t *prefix_overlay, const double *vertex_overlay, int32_t num_output,
            const int32_t *hole_base, const int32_t *hole_overlay, const int32_t *zorder_index,
            double *area, int64_t *count1_h);

/* Largest LDS allocation of one block on the current device (bytes). */
int prt_device_lds_limit(void);
/* LDS bytes of one batched-overlay team for pairs of up to nv vertices
 * with the beam-list capacity list_factor * nv. */
int64_t prt_overlay_team_lds(int32_t nv, int32_t list_factor);

#ifdef __cplusplus
}
#endif
#endif
