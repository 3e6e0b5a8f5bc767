/* C ABI of the MI355X-native metric-dependent CCH path (libnwcch.so).
 *
 * Plain pointers and sizes only.  Every function returns 0 (NWCCH_OK) or a
 * negative error code; nwcch_last_error() gives the message of the calling
 * thread's last failure.  No C++ exception crosses this boundary.
 *
 * Conventions
 *   - Vertices of the hierarchy are ranks (0 = lowest), as in python/cch_nb.py.
 *   - Weights are int32.  NWCCH_INF = 2^31-1 means "no arc / unreachable" and
 *     absorbs additions; INT32_MIN is not a valid weight.  Finite sums outside
 *     (-INF, INF) set NWCCH_STATUS_OVERFLOW instead of saturating silently.
 *   - Customized shortcut weights of one metric are one int32 [mp][2] array
 *     `ud`: ud[2a] = up (lower -> higher rank), ud[2a+1] = down.  The reference
 *     keeps them as two arrays (upw, dnw); a drop-in exposes them as two strided
 *     views of `ud`.
 *   - Device pointers (`dev` functions) are HIP device memory; `stream` is a
 *     hipStream_t (NULL = default stream).  Calls are stream-ordered: they
 *     enqueue work and return; results are valid after the stream reaches them.
 *     Calls on one handle from several host threads must use distinct streams
 *     and distinct output buffers.
 */
#ifndef NWCCH_H_
#define NWCCH_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NWCCH_OK 0
#define NWCCH_ERR_ARG (-1)      /* bad argument: null, size, shape */
#define NWCCH_ERR_ORDER (-2)    /* order is not a permutation (python/cch.py:125) */
#define NWCCH_ERR_CHORDAL (-3)  /* an arc (y, z) missing: not a chordal completion (python/cch.py:218) */
#define NWCCH_ERR_RUNTIME (-4)  /* HIP or internal failure; see nwcch_last_error() */
#define NWCCH_ERR_CAPACITY (-5) /* a size beyond what the library supports */

#define NWCCH_INF 2147483647

#define NWCCH_STATUS_OVERFLOW 1u
#define NWCCH_STATUS_BAD_WEIGHT 2u
#define NWCCH_STATUS_BAD_QUERY 4u
#define NWCCH_STATUS_QUEUE_FULL 8u
#define NWCCH_STATUS_PATH_LIMIT 16u /* a path unpacking passed 4n + 16 vertices: the metric has a negative cycle */

typedef struct nwcch_topo nwcch_topo; /* host topology (metric independent) */
typedef struct nwcch_dev nwcch_dev;   /* device copy of one topology */

/* Message of the calling thread's last failure ("" if none). */
const char* nwcch_last_error(void);

/* ------------------------------------------------------------ host side ----- */

/* Nested-dissection order with inertial-flow bisection on the undirected graph
 * underlying the directed CSR (off[n+1], tgt[m]); rank_out[v] = rank of v.
 * Replaces cch_nb.nested_dissection (python/cch_nb.py:336-426) and
 * cch.nested_dissection_order (python/cch.py:42-106); cpp/cch.cpp:323-363. */
int nwcch_order(int64_t n, const int64_t* off, const int32_t* tgt, const double* x,
                const double* y, int32_t leaf, int32_t* rank_out);

/* Symbolic contraction + arc map + triangle index + level schedule from a graph
 * and an order.  Replaces CCH.__init__ (python/cch.py:112-208) and
 * cch_nb.symbolic + edge_arcs (python/cch_nb.py:431-515); cpp/cch.cpp:377-435. */
int nwcch_topo_create(int64_t n, const int64_t* off, const int32_t* tgt, const int32_t* rank,
                      nwcch_topo** out);

/* nwcch_topo_create with options.  flags & NWCCH_TOPO_DEVICE_INDEX: the
 * lower-triangle index (python/cch.py:168-186, by_c cch.py:308) is built on GPU
 * `device` instead of the host.  The host then counts nothing and keeps no
 * triangle list: GPU `device` counts every arc's lower triangles, the host
 * orders the schedule from the counts, nwcch_dev_create fills the triangle
 * lists in device memory, and nwcch_update_prepare fills the side lists
 * (python/cch.py:309-324) there.  Every array equals the host-built one.
 * nwcch_topo_export_triangles then serves sched_arc and tri_off only;
 * nwcch_dev_export_triangles reads the list back.  A GPU is needed.
 *
 * flags & NWCCH_TOPO_DEVICE_SYMBOLIC: the symbolic contraction
 * (python/cch.py:136-166, python/cch_nb.py:431-481, cpp/cch.cpp:377-420) and
 * the original-arc map (python/cch.py:188-204, python/cch_nb.py:496-515) are
 * built on GPU `device`.  The host computes the elimination tree from the
 * original graph first (Liu's union-find walk, near-linear); the GPU then
 * builds the rows by ET height, leaves first.  first/head/e_arc/e_up are
 * identical to the host build.  An elimination tree deeper than 15,360 is
 * contracted on the host instead (the GPU's per-vertex bitmap lives in LDS).
 * Either flag may be given without the other.  A GPU is needed. */
#define NWCCH_TOPO_DEVICE_INDEX 1
#define NWCCH_TOPO_DEVICE_SYMBOLIC 2
int nwcch_topo_create_ex(int64_t n, const int64_t* off, const int32_t* tgt, const int32_t* rank,
                         int32_t flags, int device, nwcch_topo** out);

/* The same from an existing upward CSR (first[n+1], head[mp]) and arc map
 * (e_arc[m] with -1 for loops, e_up[m]), i.e. the arrays cch_nb.customize takes
 * (python/cch_nb.py:530).  e_arc may be NULL (m = 0): no original arcs. */
int nwcch_topo_from_csr(int64_t n, const int64_t* first, const int32_t* head, int64_t m,
                        const int64_t* e_arc, const int8_t* e_up, nwcch_topo** out);

/* nwcch_topo_from_csr with the options of nwcch_topo_create_ex
 * (NWCCH_TOPO_DEVICE_INDEX: the triangle index on GPU `device`).  A CSR that
 * is not a chordal completion is still rejected (NWCCH_ERR_CHORDAL): the
 * device counts are checked against sum d(d-1)/2 before anything else. */
int nwcch_topo_from_csr_ex(int64_t n, const int64_t* first, const int32_t* head, int64_t m,
                           const int64_t* e_arc, const int8_t* e_up, int32_t flags, int device,
                           nwcch_topo** out);

void nwcch_topo_destroy(nwcch_topo* t);

/* out8 = {n, m, mp, triangles, depth, levels, loop_edges, mapped_edges}
 * (the structure line of cpp/cch.cpp:779-781 and python/cch.py:206-208). */
int nwcch_topo_info(const nwcch_topo* t, int64_t* out8);

/* Copies of the host arrays; any pointer may be NULL to skip it.
 * first[n+1] (python/cch.py:157), head[mp], parent[n] (-1 = root), e_arc[m],
 * e_up[m], level[n] (ET height, cpp/cch.cpp:468-480), depth[n] (ET depth). */
int nwcch_topo_export(const nwcch_topo* t, int64_t* first, int32_t* head, int32_t* parent,
                      int64_t* e_arc, int8_t* e_up, int32_t* level, int32_t* depth);

/* Lower triangles in pull form: for schedule position p, arc sched_arc[p] has
 * triangles tri_off[p]..tri_off[p+1] of (a, b) pairs in tri_ab.  The same set as
 * tri_a/tri_b/tri_c grouped by c (python/cch.py:168-186, 308). */
int nwcch_topo_export_triangles(const nwcch_topo* t, int32_t* sched_arc, int64_t* tri_off,
                                int32_t* tri_ab);

/* ---------------------------------------------------------- device side ----- */

/* Upload a topology to GPU `device`.  The handle owns its device memory. */
int nwcch_dev_create(const nwcch_topo* t, int device, nwcch_dev** out);
void nwcch_dev_destroy(nwcch_dev* h);

/* out6 = {n, m, mp, triangles, depth, levels} */
int nwcch_dev_info(const nwcch_dev* h, int64_t* out6);

/* Host copy of the handle's triangle list in schedule order, tri_ab[2T] as
 * nwcch_topo_export_triangles gives it (the device-built list of a topology
 * made with NWCCH_TOPO_DEVICE_INDEX included).  Synchronous. */
int nwcch_dev_export_triangles(const nwcch_dev* h, int32_t* tri_ab);

/* Host copy of the partial-update side lists after nwcch_update_prepare:
 * side_off[mp+1], side_top[side_off[mp]] (= 2T) -- for arc a the top arcs of
 * the lower triangles a is a side of (python/cch.py:309-324).  Either pointer
 * may be NULL.  Synchronous. */
int nwcch_dev_export_sides(const nwcch_dev* h, int64_t* side_off, int32_t* side_top);

/* Full customization of one metric: w[m] original arc weights (device) ->
 * ud[mp][2] (device).  Replaces cch_nb.customize (python/cch_nb.py:529-563),
 * CCH.customize (python/cch.py:249-287, without the witness list) and
 * CCH::customize / customize_parallel (cpp/cch.cpp:437-463, 496-529). */
int nwcch_customize(nwcch_dev* h, const int32_t* w, int32_t* ud, void* stream);
/* The base pass and the level launches (one per ET level) are captured once per
 * stream into a HIP graph and replayed; the buffers of each call reach the
 * kernels through a per-stream argument slot written by a one-thread kernel in
 * stream order, so any w and ud may be passed.  Inside a caller's own stream
 * capture the launches are recorded one by one into that capture.  Tuning
 * changes drop the captured graphs (the device is synchronised then).  Up to
 * 256 streams per handle keep their own slot; a further stream takes the
 * oldest slot over after a device synchronisation. */

/* The values python/cch.py's customize(g, stop_on_negative_cycle=True) leaves
 * when it stops (python/cch.py:259-278): from the complete customization
 * ud_full (device) of weights w (device), ud_out (device, not ud_full) gets the
 * final values of every arc whose lower end has rank <= x_stop, and for every
 * other arc its base value folded with the lower triangles whose lower vertex
 * is <= x_stop.  The reference stops at the row of its first witness arc, so
 * x_stop = that arc's lower end; x_stop = -1 gives the base values it leaves
 * for a negative self-loop (python/cch.py:259-263). */
int nwcch_customize_stop_state(nwcch_dev* h, const int32_t* w, const int32_t* ud_full, int32_t x_stop,
                               int32_t* ud_out, void* stream);

/* Negative-cycle scan of ud[mp][2] (any metric, no handle needed): out2
 * (device int64[2]) = {number of arcs with up + down < 0 (INF never counts),
 * smallest such arc or -1}.  Replaces cch_nb.negative_cycle_arcs
 * (python/cch_nb.py:566-572), CCH::negative_cycle_arcs (cpp/cch.cpp:531-536)
 * and the scan inside CCH.has_negative_cycle (python/cch.py:662-664). */
int nwcch_negcycle_scan(const int32_t* ud, int64_t mp, int64_t* out2, void* stream);

/* Every witness arc of the scan, in no particular order: list (device int64,
 * capacity cap), *count (device int64) = total number (may exceed cap).  With
 * a sort it is the list CCH.customize returns (python/cch.py:287). */
int nwcch_negcycle_list(const int32_t* ud, int64_t mp, int64_t* list, int64_t cap, int64_t* count,
                        void* stream);

/* Elimination-tree queries between ranks s_rank[i] -> t_rank[i], i < nq, one
 * per wavefront: dist[i] (NWCCH_INF if unreachable) and, if scanned != NULL,
 * the arcs relaxed.  Replaces cch_nb.query (python/cch_nb.py:577-615),
 * CCH.query (python/cch.py:675-732, which returns (dist, scanned)) and
 * Query::run (cpp/cch.cpp:549-575). */
int nwcch_query(nwcch_dev* h, const int32_t* ud, const int32_t* s_rank, const int32_t* t_rank,
                int64_t nq, int32_t* dist, int32_t* scanned, void* stream);

/* The same on a metric stored with a stride: arc a's (up, down) pair is at
 * ud + 2 * a * ud_stride (for metric k of a batched result with slab width SW:
 * ud = UD + 2 * ((k / SW) * mp * SW + k % SW), ud_stride = SW).  `order`
 * (device int32[nq] permutation, or NULL) is the query packing: wavefront i
 * serves query order[i]; results stay at the query's own index. */
int nwcch_query_strided(nwcch_dev* h, const int32_t* ud, int64_t ud_stride, const int32_t* s_rank,
                        const int32_t* t_rank, int64_t nq, const int32_t* order, int32_t* dist,
                        int32_t* scanned, void* stream);

/* One metric's interleaved (up, down) pairs ud[mp][2] (device) split into the
 * separate arrays up[mp] and dn[mp] (device), the layout in which the
 * reference keeps a customization (python/cch_nb.py:529-563 returns upw and
 * dnw; cpp/cch.cpp:437-463 keeps vectors up and dn).  Asynchronous on
 * `stream`. */
int nwcch_metric_split(const int32_t* ud, int64_t mp, int32_t* up, int32_t* dn, void* stream);

/* ET queries on a metric given as separate up[mp] and dn[mp] arrays (device):
 * the arguments of cch_nb.query(s_rank, t_rank, first, head, parent, upw, dnw,
 * ...) python/cch_nb.py:577-615 and Query::run cpp/cch.cpp:549-575, for nq
 * pairs at once; results as nwcch_query.  Each search then reads 6 bytes per
 * scanned arc (its own weight and a 16-bit head depth) instead of a pair and a
 * 32-bit depth.  Runs the blocked query kernel, so the elimination tree may be
 * at most about 7,000 levels deep (NWCCH_ERR_CAPACITY beyond). */
int nwcch_query_split(nwcch_dev* h, const int32_t* up, const int32_t* dn, const int32_t* s_rank,
                      const int32_t* t_rank, int64_t nq, const int32_t* order, int32_t* dist,
                      int32_t* scanned, void* stream);

/* Slab width used for K metrics when `slab` (a power of two <= 64, 0 = the
 * default 64) is requested: min(slab, K rounded up to a power of two). */
int nwcch_batch_slab_width(int32_t K, int32_t slab);

/* Queries on several resident one-metric customizations in one launch: query i
 * uses metric qmetric[i] (device int32), whose ud[mp][2] device address is
 * ud_table[qmetric[i]] (device int64[n_metrics]).  A metric index outside
 * [0, n_metrics) gives INF and NWCCH_STATUS_BAD_QUERY.  The serving loop packs
 * all metrics' admitted queries into one such launch. */
int nwcch_query_multi(nwcch_dev* h, const int64_t* ud_table, int32_t n_metrics, const int32_t* qmetric,
                      const int32_t* s_rank, const int32_t* t_rank, int64_t nq, int32_t* dist, int32_t* scanned,
                      void* stream);

/* Batched customization of K metrics on one topology in one level sweep:
 * W[K][m] weights (device, metric-major) -> UD (device) in metric slabs:
 * int32 [ceil(K/SW)][mp][SW][2] with SW = nwcch_batch_slab_width(K, slab); with
 * the default slab and K <= 64 this is [1][mp][K'][2], the K' >= K pairs of one
 * arc contiguous.  Equals K calls of nwcch_customize, i.e. K runs of
 * cch_nb.customize (python/cch_nb.py:529-563) or of CCH::customize
 * (cpp/cch.cpp:437-463), which the reference runs one metric at a time
 * (python/run_nb.py:121-124; cpp/cch.cpp:805-816). */
int nwcch_customize_batch(nwcch_dev* h, const int32_t* W, int32_t K, int32_t slab, int32_t* UD, void* stream);

/* Per-metric negative-cycle scan of a batched UD (same K and slab): out
 * (device int64[2K]) = {count of metric 0..K-1, smallest witness arc of metric
 * 0..K-1 or -1} (python/cch_nb.py:566-572 per metric). */
int nwcch_negcycle_scan_batch(const int32_t* UD, int64_t mp, int32_t K, int32_t slab, int64_t* out,
                              void* stream);

/* Shortest paths with their vertices: for query i, dist[i] as nwcch_query and
 * the ranks of the unpacked path from s_rank[i] to t_rank[i] in
 * path[i * max_len ...] (device int32 [nq][max_len]); path_len[i] = the number
 * of vertices (-1 if unreachable).  A longer path keeps its first max_len
 * vertices and raises NWCCH_STATUS_QUEUE_FULL.  On a metric with a negative
 * cycle the unpacking can grow exponentially: past 4n + 16 vertices it stops,
 * path_len[i] = -2 and NWCCH_STATUS_PATH_LIMIT is raised.  Shortcuts are split at the
 * first lower triangle, x ascending, whose sides sum to the shortcut's value,
 * as CCH.query_path / _unpack do (python/cch.py:406-507). */
int nwcch_query_path(nwcch_dev* h, const int32_t* ud, const int32_t* s_rank, const int32_t* t_rank, int64_t nq,
                     int32_t max_len, int32_t* dist, int32_t* path_len, int32_t* path, void* stream);

/* CCH potential of a customized metric: p (device int32[n], indexed by rank)
 * = min(0, min over u of dist(u, v)) for the vertex of each rank, the
 * virtual-source distances Johnson's algorithm computes, in two level-by-level
 * sweeps over the hierarchy.  Replaces cch_potential (cpp/cch.cpp:635-655) and
 * CCH.potential (python/cch.py:734-755, which returns it by original id).
 * Valid for metrics without negative cycles; a value below -(2^31-1) raises
 * NWCCH_STATUS_OVERFLOW.  Its launches are replayed from a per-stream captured
 * graph, as nwcch_customize's are. */
int nwcch_potential(nwcch_dev* h, const int32_t* ud, int32_t* p, void* stream);

/* Build (once, on the host, or on the device for a topology made with
 * NWCCH_TOPO_DEVICE_INDEX) and upload the partial-update index of the topology
 * the handle was created from: for every arc the top arcs of the lower
 * triangles it is a side of, arc heights and frontier buckets.  The lazy
 * build_update_index of python/cch.py:289-347 (by_c, as_side, tail,
 * arc_edges). */
int nwcch_update_prepare(nwcch_dev* h, nwcch_topo* t);

/* Partial re-customization of one metric in place.  w[m] and ud[mp][2] are
 * the metric's device weights and its current customization; the k changes
 * (edge_idx[i], new_w[i]) are HOST arrays of distinct original-arc indices (the
 * order of python/cch.py:257) and their new weights.  Writes w, re-evaluates
 * only the arcs whose value can change, lowest first, and updates ud.
 * result (device int64[4]) = {arcs evaluated, arcs changed, changed arcs with
 * up + down < 0, smallest such arc or -1}; the changed arcs (unordered) go to
 * `changed` (device int32, capacity changed_cap) or, if NULL, to an internal
 * buffer read by nwcch_update_changed.  mode 0 = auto, 1 = one persistent
 * workgroup, 2 = one launch per ET height, 3 = a full recustomization of the
 * new weights diffed against the old values (the same values, evaluated count,
 * changed set and status bits as the frontier; auto picks it from
 * upd_full_min changes on, the persistent workgroup below).  Updates on one handle must be issued on one
 * stream.  Replaces CCH.update_edges (python/cch.py:349-402) and
 * CCH.update_created_negative_cycle (python/cch.py:666-671). */
int nwcch_update(nwcch_dev* h, int32_t* w, int32_t* ud, const int64_t* edge_idx, const int32_t* new_w,
                 int64_t k, int32_t mode, int32_t* changed, int64_t changed_cap, int64_t* result,
                 void* stream);

/* nwcch_update without any allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate device or pinned memory mid-run (tight
 * serving loops, code under HIP graph capture, harnesses that forbid it).
 * Same arguments, semantics and results as nwcch_update, except that the k
 * changes are staged in a static device array of the library (one per GPU,
 * held until the process ends) instead of the library's growable pinned and
 * device slots, and its per-height launches go out one by one (mode 3 runs
 * as mode 2): k is at most 65536 (NWCCH_ERR_CAPACITY beyond), and the host
 * arrays edge_idx and new_w must stay unchanged until the stream has reached
 * the update.  All calls on one GPU share that array: calls that overlap on
 * one GPU (several streams or threads) race.  nwcch_update_prepare must have
 * been called (it also locates the array on the handle's GPU). */
int nwcch_update_noalloc(nwcch_dev* h, int32_t* w, int32_t* ud, const int64_t* edge_idx, const int32_t* new_w,
                         int64_t k, int32_t mode, int32_t* changed, int64_t changed_cap, int64_t* result,
                         void* stream);

/* Copy the first `cap` changed arcs of the last update into out (device). */
int nwcch_update_changed(nwcch_dev* h, int32_t* out, int64_t cap, void* stream);

/* Read (and optionally clear) the sticky status bits of a handle.  Synchronises
 * `stream`.  The reference raises Python exceptions / asserts instead
 * (python/cch.py:125, 218; python/preprocess.py:9-10). */
int nwcch_status(nwcch_dev* h, uint32_t* out, int clear, void* stream);

/* Tunables of the level kernels (no counterpart in the reference; set by the
 * autotuner nwcch.tune):
 *   "flat_min_tri"  single-metric sweep: a level whose heaviest arc has at least
 *                   this many lower triangles runs triangle-parallel (one lane
 *                   per triangle); 0 = every level (default), 2^30 = none.
 *   "flat_store"    triangle-parallel levels: 0 = every arc folded with atomic
 *                   minima (default); 1 = arcs lying in one wavefront stored
 *                   with a plain store, 2 = with a non-temporal store.
 *   "flat_xcd"      triangle-parallel levels: 1 = each XCD takes one contiguous
 *                   slice of the level, 0 = workgroups interleaved (default).
 *   "tri_per_lane"  the other levels: an arc gets the smallest power-of-two
 *                   team of lanes, up to 64, with at most this many lower
 *                   triangles per lane; default 4, values 1..256.
 *   "batch_tri_per_lane"  batched sweep: wavefronts per arc, up to 16, likewise;
 *                   default 16, values 1..256.
 *   "batch_order"   batched sweep: 1 = the arcs of each wave class taken in id
 *                   order (a position table built per shape; the first call
 *                   after a shape change synchronises the device), 0 = in
 *                   schedule order (default).
 *   "batch_xcd"     batched sweep: 1 = each XCD takes one contiguous slice of a
 *                   wave class, 0 = workgroups interleaved (default).
 *   "upd_max_blocks"  workgroups of one per-height launch of a large update,
 *                   at most (default 1024).
 *   "customize_graph"  1 = nwcch_customize and nwcch_potential replay a
 *                   captured graph per stream (default), 0 = they launch
 *                   kernel by kernel.
 *   "query_wpb"     queries (wavefronts) per workgroup of the query kernels:
 *                   1 (default), 2 or 4, fewer where their labels would pass
 *                   64 KB of LDS.
 *   "upd_persistent_max"  nwcch_update in mode 0 runs updates of at most this
 *                   many changed arcs in one persistent workgroup, larger ones
 *                   one launch per height; default 2.
 *   "upd_full_min"  nwcch_update in mode 0 recustomizes fully (mode 3) from
 *                   this many changed arcs on (above upd_persistent_max);
 *                   default 3, 0 = never.
 *   "update_graph"  1 = the per-height launches of nwcch_update replay
 *                   captured graphs of 32 heights each, per stream (default);
 *                   0 = they go out one by one.  nwcch_update_noalloc always
 *                   launches them one by one (capturing allocates).
 *   "rows_min_len", "rows_min_rows", "rows_max_rows", "rows_min_level_tri"
 *                   nwcch_customize runs a level in row-merge form (one
 *                   workgroup per row, the row's values in LDS, the suffixes
 *                   of the rows below streamed) when its suffix pieces average
 *                   at least rows_min_len arcs (default 16; 65 = never), it
 *                   has rows_min_rows..rows_max_rows rows with triangles
 *                   (default 3..200) and at least rows_min_level_tri
 *                   triangles (default 262,144).
 *   "rows_batch", "rows_batch_min_rows", "rows_batch_max_rows",
 *   "rows_batch_min_level_tri", "rows_batch_min_len"
 *                   the same form in nwcch_customize_batch for slabs of 8
 *                   metrics and K a multiple of 8 (rows_batch 1 = on,
 *                   default; 8..40 rows, 262,144 triangles, pieces of 16 arcs
 *                   by default).  Every form gives identical values.
 *   "query_blk", "query_blk_small", "query_blk_waves", "query_blk_unroll"
 *                   the blocked query kernel (blocks of 32 or 64 ancestors, a
 *                   block-sized (min, +) matrix in LDS): 1 = used for hierarchies deeper
 *                   than 1,024 and for launches of at most query_blk_small
 *                   queries (default 4,096), 0 = never; wavefronts per search
 *                   side 1, 2, 4 or 8 (0 = by launch size, default); 64-arc
 *                   chunks in flight per wavefront 4, 8 (default) or 16.
 *   "query_blk64"   1 = with 8 wavefronts per side the blocked kernel walks
 *                   the chains in blocks of 64 ancestors (default), 0 = 32.
 *   "query_deep_max"  query launches of at most this many queries use the
 *                   latency-oriented kernels (deeper prefetch), larger launches
 *                   the throughput-oriented one (more wavefronts per CU);
 *                   default 32768.
 *   "query_chunk"   1: the latency-oriented launches read ancestor chains 32
 *                   vertices at a time and prefetch rows four chain vertices
 *                   ahead; 0: they walk parent pointers, prefetching 128 arcs
 *                   of the next row.  Default 1 for ET depths above 1024. */
int nwcch_set_tuning(nwcch_dev* h, const char* key, int32_t value);
int nwcch_get_tuning(const nwcch_dev* h, const char* key, int32_t* value);

/* Launch profiling (the reference's Timing/now() role, cpp/cch.cpp:42-45,
 * 714-729, at kernel granularity): after nwcch_profile_begin every kernel launch
 * of the handle is bracketed by two HIP events (up to max_launches launches);
 * nwcch_profile_end waits for them and writes, per launch, the kernel id
 * (0 customize level, 2 query, 3 update, 4 batched customize) and its duration
 * in ms; *count = launches recorded. */
int nwcch_profile_begin(nwcch_dev* h, int64_t max_launches);
int nwcch_profile_end(nwcch_dev* h, int32_t* kernel, float* ms, int64_t cap, int64_t* count);

#ifdef __cplusplus
}
#endif

#endif /* NWCCH_H_ */
