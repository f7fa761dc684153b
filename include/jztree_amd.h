/*
 * jztree_amd.h -- C ABI of the MI355X (gfx950) tree hot path.
 *
 * One entry point per XLA-FFI target that jztree binds for this path.  Each declaration cites
 * the reference interface it replaces (paths relative to the jztree repository):
 *   handler signature  src/jztree_cuda/generated/ffi_<name>.cu
 *   host function      src/jztree_cuda/<name>.cuh
 *   Python call site   src/jztree/<name>.py
 *
 * Conventions (mirroring the reference's FFI contract, see ffi_knn.cu:159-168):
 *   - buffer sizes the FFI handlers infer from buffer dimensions are explicit arguments here;
 *     interaction-list readers clamp every segment to the list size they are given, so an
 *     overflowing list (reported by the caller afterwards, as the reference does) never makes a
 *     kernel read past its buffer;
 *   - every buffer is a device pointer owned by the caller; outputs are preallocated with the
 *     shapes the Python layer computes; nothing is allocated inside an entry point;
 *   - scratch space arrives as an explicit `tmp` buffer whose size is queried with the matching
 *     *_tmp_bytes function (the reference passes tmp buffers as extra FFI outputs);
 *   - all work is enqueued on `stream`; no entry point synchronises or reads device data on the
 *     host, so a sequence of calls can be captured into a hipGraph;
 *   - the return value is 0 on success; otherwise jzt_last_error() holds the message (the
 *     reference returns std::string errors mapped to ffi::Error::Internal);
 *   - positions are AoS (n, dim) arrays of `dtype` (JZT_F32 or JZT_F64); index outputs of the
 *     sort are int32 for f32 and int64 for f64 (same_width_int, src/jztree/data.py:13-19).
 */
#ifndef JZTREE_AMD_H
#define JZTREE_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define JZT_ABI_VERSION 1

typedef struct ihipStream_t* jzt_stream_t; /* a hipStream_t */

enum { JZT_F32 = 0, JZT_F64 = 1, JZT_I32 = 2, JZT_I64 = 3 }; /* integer positions: sort and SearchSortedZ */

/* message of the last failed call on this host thread */
const char* jzt_last_error(void);
int jzt_abi_version(void);

/* Tuning knobs (no counterpart in the reference, which has no autotuner): process-wide settings
 * read at each launch; results never depend on them.  Set them between calls, not while another
 * host thread launches.
 *   JZT_TUNE_LEAF_WAVES  waves per EU the leaf kNN kernel is compiled for when k <= 32 (1, 2, 4;
 *                        default 4): the register budget per lane against occupancy
 *   JZT_TUNE_SORT_KPT    keys per thread of the onesweep radix passes (8, 12, 16; default 12):
 *                        tile = 256 * kpt keys
 *   JZT_TUNE_XCD         order of the query leaves in the kNN and FoF leaf kernels (0, 1; default
 *                        1): 1 gives each of the 8 XCDs a contiguous eighth of the z-ordered
 *                        leaves, so that neighbouring leaves share an L2; 0 deals consecutive
 *                        leaves round-robin to the XCDs */
#define JZT_TUNE_LEAF_WAVES 0
#define JZT_TUNE_SORT_KPT 1
#define JZT_TUNE_XCD 2
int jzt_set_tuning(int knob, int value);
int jzt_get_tuning(int knob);

/* ------------------------------------------------------------------------------------------ */
/* sort                                                                                         */
/* ------------------------------------------------------------------------------------------ */

/* PosZorderSort -- replaces ffi target "PosZorderSort" (generated/ffi_sort.cu:62-164, host
 * sort.cuh:40-76, caller tree.py:75-93).  pos_out[i] = pos_in[id_out[i]] in z-order; pos_out may
 * alias pos_in (the reference aliases {0:0}).  Order: exact Morton order of the float bits, ties by
 * input index.  NaN rows last.  dim in [2, 3]. */
size_t jzt_zsort_tmp_bytes(int64_t n, int dim, int dtype);
int jzt_pos_zorder_sort(jzt_stream_t stream, const void* pos_in, void* pos_out, void* id_out, void* tmp,
                        size_t tmp_bytes, int64_t n, int dim, int dtype);

/* PosZorderSort without a workspace -- no counterpart in the reference.  For callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): same inputs, outputs and order as jzt_pos_zorder_sort, but the scratch is a
 * fixed 64 MiB device array inside the library (one per device, held until the process ends), so
 * nothing is allocated and no tmp buffer is passed.  n is limited to what fits that array
 * (jzt_zsort_tmp_bytes(n, dim, dtype) <= 64 MiB: up to 1,427,456 points for f32 3D,
 * 1,137,141 for f64 3D); larger n fails
 * with a message.  Calls that overlap on one device share the array and race: serialise them. */
int jzt_pos_zorder_sort_noalloc(jzt_stream_t stream, const void* pos_in, void* pos_out, void* id_out, int64_t n,
                                int dim, int dtype);

/* Batched PosZorderSort -- no counterpart in the reference, whose batching is vmap with
 * vmap_method="sequential" (tree.py:88): one set at a time.  The n points hold nsets independent
 * sets, set s in rows [set_offsets[s], set_offsets[s+1]) (device int64, nsets+1 entries,
 * set_offsets[0] = 0, set_offsets[nsets] = n).  Every set stays in its rows and is ordered exactly
 * as jzt_pos_zorder_sort orders it alone; id_out holds input row indices (subtract the set's
 * offset for set-local ids).  1 <= nsets <= 65536, n < 2^30. */
size_t jzt_zsort_segmented_tmp_bytes(int64_t n, int64_t nsets, int dim, int dtype);
int jzt_pos_zorder_sort_segmented(jzt_stream_t stream, const void* pos_in, void* pos_out, void* id_out,
                                  const int64_t* set_offsets, int64_t nsets, void* tmp, size_t tmp_bytes,
                                  int64_t n, int dim, int dtype);

/* SearchSortedZ -- replaces ffi target "SearchSortedZ" (kernel sort.cuh:82-132, caller tree.py:143-158).
 * indices[i] = left insertion point of posz_query[i] into the z-sorted posz_have (have[idx-1] <
 * x <= have[idx] in the z-order of jzt_pos_zorder_sort).  leaf_search != 0: posz_have holds one
 * point per leaf; the index of the neighbouring leaf with the smaller level difference to x. */
int jzt_search_sorted_z(jzt_stream_t stream, const void* posz_have, const void* posz_query, int64_t n_have,
                        int64_t n_query, int dim, int dtype, int leaf_search, int32_t* indices);

/* SegmentSort -- replaces ffi target "SegmentSort" (generated/ffi_knn.cu:412-423, knn.cuh:700-722,
 * caller knn.py:112-127).  Sorts (key, val) within segments spl[i]..spl[i+1] by (key, val). */
int jzt_segment_sort(jzt_stream_t stream, const int32_t* spl, const float* key, const int32_t* val,
                     float* key_out, int32_t* val_out, int64_t size_segs, int64_t size_keys);

/* ------------------------------------------------------------------------------------------ */
/* tree                                                                                         */
/* ------------------------------------------------------------------------------------------ */

/* FlagLeafBoundaries -- replaces ffi target "FlagLeafBoundaries" (generated/ffi_tree.cu:31-141,
 * kernel tree.cuh:13-98, caller tree.py:189-262).  split_flags/lvl have size_part+1 entries;
 * lvl_bound (2 int32) and npart (1 int32) are device scalars. */
int jzt_flag_leaf_boundaries(jzt_stream_t stream, const void* posz, const uint8_t* ptype,
                             const int32_t* lvl_bound, const int32_t* npart, int8_t* split_flags,
                             int32_t* lvl, int64_t size_part, int dim, int dtype, int max_size,
                             int scan_size, int num_types);

/* FindNodeBoundaries -- replaces ffi target "FindNodeBoundaries" (generated/ffi_tree.cu:148-257,
 * kernel tree.cuh:104-199, caller tree.py:264-284).  pos_in has size_nodes-1 rows; pos_boundary
 * 2 rows; nleaves a device scalar; outputs have size_nodes entries. */
int jzt_find_node_boundaries(jzt_stream_t stream, const void* pos_in, const void* pos_boundary,
                             const int32_t* nleaves, int32_t* nodes_levels, int32_t* nodes_lbound,
                             int32_t* nodes_rbound, int64_t size_nodes, int dim, int dtype, int lvl_max,
                             int lvl_invalid);

/* Batched tree construction -- no counterpart in the reference (see jzt_pos_zorder_sort_segmented).
 * As jzt_flag_leaf_boundaries (single particle type) and jzt_find_node_boundaries, with set_id
 * (int32, one per row of posz / pos_in, non-decreasing) separating the sets of a batch: the
 * boundary between rows of different sets gets level 2^16, above every Morton level, so it is
 * always a leaf split and no node extends across it.  Inside a set the flags are those of a
 * single-set call on that set. */
int jzt_flag_leaf_boundaries_sets(jzt_stream_t stream, const void* posz, const int32_t* set_id,
                                  const int32_t* lvl_bound, const int32_t* npart, int8_t* split_flags,
                                  int32_t* lvl, int64_t size_part, int dim, int dtype, int max_size,
                                  int scan_size);
int jzt_find_node_boundaries_sets(jzt_stream_t stream, const void* pos_in, const int32_t* set_id,
                                  const void* pos_boundary, const int32_t* nleaves, int32_t* nodes_levels,
                                  int32_t* nodes_lbound, int32_t* nodes_rbound, int64_t size_nodes, int dim,
                                  int dtype, int lvl_invalid);

/* GetNodeGeometry -- replaces ffi target "GetNodeGeometry" (generated/ffi_tree.cu:409-529, kernel
 * tree.cuh:312-372, caller tree.py:286-329).  mode_flags: 1 level, 2 center, 4 extent (outputs not
 * selected may be NULL).  nnodes a device scalar. */
int jzt_get_node_geometry(jzt_stream_t stream, const void* pos, const int32_t* lbound, const int32_t* rbound,
                          const int32_t* nnodes, int32_t* level, void* center, void* extent,
                          int64_t size_nodes, int64_t size_part, int dim, int dtype, int lvl_invalid,
                          uint32_t mode_flags, int upper_extent);

/* CenterOfMass -- replaces ffi target "CenterOfMass" (kernel tree.cuh:378-429, caller
 * tree.py:479-499 and :736-754).  Per node i < nnodes: the mass-weighted centre of the points
 * pos[isplit[i] .. isplit[i+1]) (row-major (n, dim)) with weights mass[] and their total mass,
 * written as com_out[i*(dim+1) .. ] = (centre[dim], mass); Kahan-compensated sums in point order
 * when kahan != 0.  Empty ranges and non-positive totals give NaN rows.  (The reference handles
 * float32 3-D; float64 and 2-D are instantiated too.) */
int jzt_center_of_mass(jzt_stream_t stream, const int32_t* isplit, const void* pos, const void* mass, void* com_out,
                       int64_t nnodes, int kahan, int dim, int dtype);

/* jzt_geometry_to_f32 -- no counterpart in the reference (extension for float64 positions with
 * float32 distance kernels).  Per node of a float64 tree (centres cent (n, dim), levels lvl): the
 * smallest enclosing Morton cell whose boundaries and centre are float32 numbers (per-dim side at
 * least 2^(E-23) for coordinates below 2^E in magnitude), as float32 centres and levels.  It
 * contains the float32-rounded points of the node (rounding is monotone and fixes those
 * boundaries), so the float32 node kernels prune exactly for the rounded points.  Invalid slots
 * (lvl_invalid or non-finite centre) are copied. */
int jzt_geometry_to_f32(jzt_stream_t stream, const double* cent, const int32_t* lvl, int64_t n, int dim,
                        int lvl_invalid, float* cent_out, int32_t* lvl_out);

/* The JAX glue of the tree build, as device ops (no reference FFI counterpart: the reference runs
 * these steps as XLA ops inside tree.py).
 *
 * jzt_where_flags: tree.py:257 `jnp.where(flags, size=alloc, fill_value=npart)`.  out[j] is the
 * index of the j-th nonzero flag, padded with `fill` (or *fill_dev if non-NULL); *count (device,
 * may be NULL) = number of nonzero flags. */
size_t jzt_where_tmp_bytes(int64_t n);
int jzt_where_flags(jzt_stream_t stream, const int8_t* flags, int64_t n, int32_t* out, int64_t size,
                    int32_t fill, const int32_t* fill_dev, int32_t* count, void* tmp, size_t tmp_bytes);

/* jzt_count_not_nan: data.py:82-90 get_num's default `jnp.sum(~jnp.isnan(pos[..., 0]))`, the
 * filled particle count, as one device op: *out (device int32) = number of rows whose first
 * coordinate is not NaN (a kernel zero-fills it first; no memset, so it can be graph-captured). */
int jzt_count_not_nan(jzt_stream_t stream, const void* pos, int64_t n, int dim, int dtype, int32_t* out);

/* jzt_regularization_level: tree.py:55-68 + :160-187 find_regularization_level.  Histogram of
 * `lvl` weighted by `weights` (NULL = 1) over [min_lvl, max_lvl], then the first level whose
 * volume exceeds max_volume_fac times the mean volume below the min_percentile level, in FP32. */
size_t jzt_regularization_tmp_bytes(int min_lvl, int max_lvl);
int jzt_regularization_level(jzt_stream_t stream, const int32_t* lvl, const int32_t* weights, int64_t n,
                             int32_t* out_level, int min_lvl, int max_lvl, float max_volume_fac,
                             float min_percentile, void* tmp, size_t tmp_bytes);

/* jzt_dense_ilist: tree.py:868-930 `grouped_dense_interaction_list` as one device op.  With
 * nn = *nnodes top-plane nodes grouped into nsuper = ceil(nn/ngroup) super-nodes:
 * spl_super[i] = min(i*ngroup, nn) and ispl[i] = min(i*nsuper, nsuper^2) for i <= size_super,
 * ilist[j] = j % nsuper and r2[j] = 0 (r2 may be NULL) for j < min(nsuper^2, size_ilist);
 * *nsuper (device) receives nsuper.  The caller checks nsuper^2 <= size_ilist. */
int jzt_dense_ilist(jzt_stream_t stream, const int32_t* nnodes, int ngroup, int32_t* spl_super, int32_t* ispl,
                    int64_t size_super, int32_t* ilist, float* r2, int64_t size_ilist, int32_t* nsuper);

/* jzt_dense_ilist_sets: jzt_dense_ilist for a batch (no counterpart in the reference).  node_set
 * (int32, non-decreasing over the first *nnodes entries, values in [0, nsets)) names the set of
 * each top-plane node.  The c_s nodes of set s form ceil(c_s/ngroup) super-nodes (numbered set
 * by set), and the list of a super-node holds every super-node of its own set, in order.  tmp
 * holds jzt_dense_ilist_sets_tmp_bytes(nsets) bytes.  The caller checks sum_s nsuper_s^2 <=
 * size_ilist via ispl[size_super]. */
size_t jzt_dense_ilist_sets_tmp_bytes(int64_t nsets);
int jzt_dense_ilist_sets(jzt_stream_t stream, const int32_t* nnodes, const int32_t* node_set, int64_t nsets,
                         int ngroup, int32_t* spl_super, int32_t* ispl, int64_t size_super, int32_t* ilist,
                         float* r2, int64_t size_ilist, int32_t* nsuper, void* tmp, size_t tmp_bytes);

/* ------------------------------------------------------------------------------------------ */
/* kNN                                                                                          */
/* ------------------------------------------------------------------------------------------ */

/* KnnNode2Node -- replaces ffi target "KnnNode2Node" (generated/ffi_knn.cu:191-370, host
 * knn.cuh:602-694, caller knn.py:62-110).  nodes: (size_nodes, dim+1) of dtype, the last column
 * holding the node level bit-cast (PosLvl.pos_lvl, data.py:231-239).  Outputs: node_rmax2
 * (size_nodes), node_ilist_spl (size_nodes+1), node_ilist_isrc/_r2 (node_ilist_size).  Each node's
 * list is sorted ascending by (r2, isrc).  *overflow semantics as the reference: spl[-1] may exceed
 * node_ilist_size, the caller checks. */
int jzt_knn_node2node(jzt_stream_t stream, const int32_t* parent_ilist_spl, const int32_t* parent_ilist_isrc,
                      const float* parent_ilist_r2, const int32_t* parent_spl, const void* nodes,
                      const int32_t* nodes_npart, float* node_rmax2, int32_t* node_ilist_spl,
                      int32_t* node_ilist_isrc, float* node_ilist_r2, int64_t size_parents, int64_t size_nodes,
                      int64_t parent_ilist_size, int64_t node_ilist_size, int k, float boxsize, int dim, int dtype);

/* KnnNode2Node with a caller-given search radius (extension, no reference counterpart;
 * jzt_knn_node2node is this call with node_r2_in = NULL).  When node_r2_in (size_nodes float) is
 * non-NULL, the guaranteed-radius search is skipped: node i's list holds every node whose
 * distance bound is <= node_r2_in[i] * (1 + 2^-16), and node_rmax2 receives that scaled radius.
 * Used with a proposed radius r2_0 per leaf, so that jzt_knn_leaf2leaf_ex(init_r2 = r2_0) can
 * rely on the lists holding every leaf within r2_0. */
int jzt_knn_node2node_ex(jzt_stream_t stream, const int32_t* parent_ilist_spl, const int32_t* parent_ilist_isrc,
                         const float* parent_ilist_r2, const int32_t* parent_spl, const void* nodes,
                         const int32_t* nodes_npart, const float* node_r2_in, float* node_rmax2,
                         int32_t* node_ilist_spl, int32_t* node_ilist_isrc, float* node_ilist_r2,
                         int64_t size_parents, int64_t size_nodes, int64_t parent_ilist_size,
                         int64_t node_ilist_size, int k, float boxsize, int dim, int dtype);


/* Regularization's extra leaf splits (tree.py:236-243 of the reference, one kernel here):
 * flags[i] = flags[i] | (lvl[i] >= *lvl_max && i <= *npart) as 0/1 bytes, i < n; lvl_max and npart
 * are device scalars. */
int jzt_flags_or_level(jzt_stream_t stream, int8_t* flags, const int32_t* lvl, int64_t n, const int32_t* lvl_max,
                       const int32_t* npart);

/* One plane of a per-plane packed array (extension; the data.py:251-345 PackedArray.get of the
 * reference, which jax evaluates as an index + select): out row i = data row ispl[level] + i
 * while that is below ispl[level + 1], else every element of the row = *fill.  Rows are
 * row_elems elements of elem_bytes (4 or 8) bytes; ispl is int64 when ispl64, else int32 (device). */
int jzt_packed_get(jzt_stream_t stream, const void* data, int64_t nrows, const void* ispl, int ispl64, int level,
                   const void* fill, int elem_bytes, int row_elems, void* out, int64_t size);

/* Guaranteed k-neighbour radius^2 of the nodes of one tree plane (extension, the start bound of
 * the bounded kNN search), planes taken from the top down: cap[c] = squared diagonal (float64) of
 * node c's cell if it holds >= k particles (npart = spl[n2l[c+1]] - spl[n2l[c]], indices clamped
 * to nspl - 1), else cap_up of its parent in the plane above (the last P with n2c_up[P] <= c,
 * clamped into [0, size)); +inf at the top plane (n2c_up = cap_up = NULL).  n2l has size+1
 * entries, n2c_up size+1 (non-decreasing). */
int jzt_plane_cap(jzt_stream_t stream, const int32_t* n2l, const int32_t* spl, int64_t nspl, const int32_t* lvl,
                  const int32_t* n2c_up, const double* cap_up, int64_t size, int k, int dim, double* cap);

/* Segment maximum (extension, used to pass search radii up the tree planes): out[i] = max of
 * vals over [spl[i], spl[i+1]) clamped to [0, nvals), or `empty` for an empty segment.  spl has
 * nseg+1 entries. */
int jzt_segment_max_f32(jzt_stream_t stream, const int32_t* spl, const float* vals, int64_t nvals, float* out,
                        int64_t nseg, float empty);

/* KnnLeaf2Leaf -- replaces ffi target "KnnLeaf2Leaf" (generated/ffi_knn.cu:31-188, host
 * knn.cuh:309-356, caller knn.py:28-60).  knn_rad (nq, k) of dtype, knn_id (nq, k) int32.
 * Neighbours ascending by (canonical r2, id); id = z-index in xT, or src_ids[z-index] when src_ids
 * is non-NULL (an extension: lets the caller report input indices directly).  dtype JZT_F32 or
 * JZT_F64, dim 2-10 (the reference's instances, _generate_ffi.py:9-21).  tmp: at least
 * jzt_knn_leaf2leaf_tmp_bytes bytes (the multi-pass state, then 256 bytes of work-queue counters
 * that balance the query leaves over the waves); a smaller tmp (down to NULL for one pass) runs
 * the leaves in a fixed order, with the same results.  Calls that share a tmp must not overlap. */
size_t jzt_knn_leaf2leaf_tmp_bytes(int64_t size_part_query, int k, int dtype);
int jzt_knn_leaf2leaf(jzt_stream_t stream, const int32_t* ilist_spl, const int32_t* ilist_isrc,
                      const float* ilist_r2, int64_t ilist_size, const int32_t* splT, const void* xT,
                      const int32_t* splQ, const void* xQ, const int32_t* src_ids, void* knn_rad, int32_t* knn_id,
                      void* tmp, size_t tmp_bytes, int64_t size_leaves_query, int64_t size_part_query, int k,
                      float boxsize, int dim, int dtype);

/* KnnLeaf2Leaf with three optional extensions (no reference counterpart; jzt_knn_leaf2leaf is this
 * call with all three NULL and out_rows = size_part_query):
 *   dst_rows (size_part_query int32): output row of each query (rows outside [0, out_rows) are not
 *     written), so results can land in input order directly;
 *   init_r2 (size_leaves_query float, k <= 64): starting bound per query leaf.  Only candidates with
 *     r2 <= init_r2 enter the top-k.  The caller guarantees that each query leaf's list holds every
 *     leaf within init_r2 of it; then every query that finds k neighbours is exact.  A leaf with
 *     init_r2 < 0 (or NaN) is skipped: its rows and retry flags are left as they are;
 *   retry (size_part_query int32, with init_r2): 1 for queries left with fewer than k neighbours
 *     inside the bound (their rows hold sentinels and must be recomputed), else 0. */
int jzt_knn_leaf2leaf_ex(jzt_stream_t stream, const int32_t* ilist_spl, const int32_t* ilist_isrc,
                         const float* ilist_r2, int64_t ilist_size, const int32_t* splT, const void* xT,
                         const int32_t* splQ, const void* xQ, const int32_t* src_ids, const int32_t* dst_rows,
                         const float* init_r2, int32_t* retry, void* knn_rad, int32_t* knn_id, void* tmp,
                         size_t tmp_bytes, int64_t size_leaves_query, int64_t size_part_query, int64_t out_rows,
                         int k, float boxsize, int dim, int dtype);

/* KnnFused -- no single reference counterpart: replaces the leaf-plane KnnNode2Node together with
 * KnnLeaf2Leaf (knn.cuh:309-356, :602-694; callers knn.py:186 and :325) by one kernel that walks
 * the plane-1 lists (parent_ilist_*, r2-sorted, as produced by jzt_knn_node2node for plane 1) and
 * evaluates child leaves on the fly.  parent_child_spl: plane-1 node -> leaf splits; leaf_parent:
 * leaf -> plane-1 node; leaf_nodes: (size_leaves, dim+1) leaf centre + level bits; leaf_spl: leaf ->
 * particle splits; xz: z-ordered positions (sources = queries).  src_ids (optional): reported id
 * per particle; dst_rows (optional): output row per query.  init_r2 (optional, per leaf): starting
 * bound; queries left with < k neighbours inside it are flagged in `retry` (n int32, required with
 * init_r2) and recomputed without bound before return.  k <= 64.  Results equal the unfused ops. */
int jzt_knn_fused(jzt_stream_t stream, const int32_t* parent_ilist_spl, const int32_t* parent_ilist_isrc,
                  const float* parent_ilist_r2, int64_t parent_ilist_size, const int32_t* parent_child_spl,
                  const int32_t* leaf_parent, const void* leaf_nodes, const int32_t* leaf_spl, const void* xz,
                  const int32_t* src_ids, const int32_t* dst_rows, const float* init_r2, int32_t* retry,
                  void* knn_rad, int32_t* knn_id, int64_t size_leaves, int64_t size_part, int k, float boxsize,
                  int dim, int dtype);

/* KnnFused retry pass (extension): recomputes exactly the queries flagged in `retry` (size_part
 * int32, input) and writes their rows; other rows are left untouched.  Same arguments as
 * jzt_knn_fused, plus bound_r2 (per leaf, may be NULL): a radius^2 within which every flagged
 * query of the leaf is known to have k neighbours (e.g. the guaranteed cell radius); the search
 * starts there, and a chunk whose flagged queries come up short after all is redone unbounded.
 * NULL = unbounded.  This is the exact fallback after any bounded pass (the fused one or
 * jzt_knn_leaf2leaf_ex with init_r2). */
int jzt_knn_fused_retry(jzt_stream_t stream, const int32_t* parent_ilist_spl, const int32_t* parent_ilist_isrc,
                        const float* parent_ilist_r2, int64_t parent_ilist_size, const int32_t* parent_child_spl,
                        const int32_t* leaf_parent, const void* leaf_nodes, const int32_t* leaf_spl, const void* xz,
                        const int32_t* src_ids, const int32_t* dst_rows, const float* bound_r2, const int32_t* retry,
                        void* knn_rad, int32_t* knn_id, int64_t size_leaves, int64_t size_part, int k, float boxsize,
                        int dim, int dtype);

/* ------------------------------------------------------------------------------------------ */
/* learned radius proposer (no reference counterpart)                                          */
/* ------------------------------------------------------------------------------------------ */

/* A 12-64-64-64-1 GELU MLP with bf16 weights (fp32 arithmetic) predicts, per leaf, log2 of the
 * largest k-th neighbour distance among the leaf's particles.  exp2(prediction + margin)^2 is the
 * starting bound r2_0 of jzt_knn_node2node_ex / jzt_knn_leaf2leaf_ex, whose retry flags feed
 * jzt_knn_fused_retry: kNN results never depend on the network.  Flat parameter layout
 * (input-major weights): W1[12][64] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[64] b4. */
#define JZT_PROPOSER_NFEAT 12
#define JZT_PROPOSER_NPARAM 9217
int jzt_proposer_num_params(void);

/* 12 features per leaf slot (feat: size x 12 float; slots >= *nleaves get zeros).  Plane-0 arrays
 * npart0/lvl0 (size), leaf_parent (size, plane-1 node of each leaf), plane-1 arrays npart1/lvl1,
 * parent_child_spl (plane-1 -> leaf splits), and optionally parent_parent (plane-1 -> plane-2 node)
 * with npart2/lvl2 (NULL when the tree has no plane 2).  nleaves: device scalar. */
int jzt_leaf_features(jzt_stream_t stream, const int32_t* npart0, const int32_t* lvl0, const int32_t* leaf_parent,
                      const int32_t* npart1, const int32_t* lvl1, const int32_t* parent_child_spl,
                      const int32_t* parent_parent, const int32_t* npart2, const int32_t* lvl2,
                      const int32_t* nleaves, int64_t size, float* feat);

/* Inference.  norm (26 float): feature mean[12], feature std[12], target mean, target std (the
 * network works on standardised features and targets).  Writes, per sample i < n,
 * out_log2r[i] = prediction (optional) and out_r2[i] = exp2(prediction + margin)^2 (optional; 0
 * where npart != NULL and npart[i] <= 0). */
int jzt_mlp_infer(jzt_stream_t stream, const uint16_t* weights, const float* norm, const float* feat,
                  const int32_t* npart, int64_t n, float margin, float* out_r2, float* out_log2r);

/* Training forward + backward: grad (9217 float) = d/dW of loss = (1/batch) sum (y - t)^2 over
 * standardised features x (batch x 12) and targets t, taken at the bf16 weights; loss (1 float).
 * Both are overwritten. */
int jzt_mlp_grad(jzt_stream_t stream, const uint16_t* weights, const float* x, const float* t, int64_t batch,
                 float* grad, float* loss);

/* One Adam step on the fp32 master weights w (bias-corrected, step counts from 1), then
 * wq = bf16(w) rounded to nearest-even.  jzt_to_bf16: wq = bf16(w) only. */
int jzt_adam_step(jzt_stream_t stream, float* w, uint16_t* wq, float* m, float* v, const float* grad, int64_t n,
                  float lr, float beta1, float beta2, float eps, int step);
int jzt_to_bf16(jzt_stream_t stream, const float* w, uint16_t* wq, int64_t n);

/* ------------------------------------------------------------------------------------------ */
/* friends-of-friends                                                                           */
/* ------------------------------------------------------------------------------------------ */

/* NodeToChildLabel -- replaces ffi target "NodeToChildLabel" (generated/ffi_fof.cu:31-124, kernel
 * fof.cuh:42-71, caller fof.py:85-109).  node_igroup (size_child) is zero-filled first. */
int jzt_node_to_child_label(jzt_stream_t stream, const int32_t* parent_igroup, const uint8_t* parent_is_local,
                            const int32_t* parent_lvl, const int32_t* parent_spl, int32_t* node_igroup,
                            int64_t size_parent, int64_t size_child, float r2link, int dim, int dtype);

/* FofNode2Node -- replaces ffi target "FofNode2Node" (generated/ffi_fof.cu:127-274, host
 * fof.cuh:210-284, caller fof.py:31-70).  node_igroup may alias node_igroup_in ({4:0}). */
int jzt_fof_node2node(jzt_stream_t stream, const int32_t* parent_ilist_spl, const int32_t* parent_ilist,
                      const int32_t* parent_spl, const void* nodes, const int32_t* node_igroup_in,
                      int32_t* node_igroup, int32_t* node_ilist_spl, int32_t* node_ilist, int64_t size_parent,
                      int64_t size_node, int64_t parent_ilist_size, int64_t size_node_ilist, float r2link,
                      float boxsize, int dim, int dtype);

/* FofLeaf2Leaf -- replaces ffi target "FofLeaf2Leaf" (generated/ffi_fof.cu:277-406, host
 * fof.cuh:349-379, caller fof.py:153-169).  part_igroup = min z-index of each particle's group. */
int jzt_fof_leaf2leaf(jzt_stream_t stream, const int32_t* ilist_spl, const int32_t* ilist, int64_t ilist_size,
                      const int32_t* spl, const void* pos, const int32_t* part_igroup_in, int32_t* part_igroup,
                      int64_t size_leaves, int64_t size_part, float r2link, float boxsize, int dim, int dtype);
/* FofLeaf2Leaf with the leaves' cells (extension, no counterpart in the reference): leaf_cells
 * holds (centre[dim], level bits) per leaf as jzt_fof_node2node's nodes do.  A list entry whose
 * cell lies beyond rlink of a query chunk's bounding box (and needs no periodic wrap) is dropped
 * without loading its particles.  Same labels as jzt_fof_leaf2leaf. */
int jzt_fof_leaf2leaf_cells(jzt_stream_t stream, const int32_t* ilist_spl, const int32_t* ilist, int64_t ilist_size,
                            const int32_t* spl, const void* pos, const void* leaf_cells,
                            const int32_t* part_igroup_in, int32_t* part_igroup, int64_t size_leaves,
                            int64_t size_part, float r2link, float boxsize, int dim, int dtype);

/* FoF over a batch of sets, each with its own linking length (no counterpart in the reference;
 * see jzt_pos_zorder_sort_segmented).  As the three calls above with the scalar r2link replaced
 * by one float r2 per parent node (NodeToChildLabel, FofNode2Node) or per query leaf
 * (FofLeaf2Leaf): the r2link of the set the node or leaf belongs to.  Every node of a parent's
 * or leaf's interaction list lies in the same set, so each set is linked exactly as a
 * single-set call with its own r2link links it. */
int jzt_node_to_child_label_sets(jzt_stream_t stream, const int32_t* parent_igroup, const int32_t* parent_lvl,
                                 const int32_t* parent_spl, const float* parent_r2link, int32_t* node_igroup,
                                 int64_t size_parent, int64_t size_child, int dim, int dtype);
int jzt_fof_node2node_sets(jzt_stream_t stream, const int32_t* parent_ilist_spl, const int32_t* parent_ilist,
                           const int32_t* parent_spl, const float* parent_r2link, const void* nodes,
                           const int32_t* node_igroup_in, int32_t* node_igroup, int32_t* node_ilist_spl,
                           int32_t* node_ilist, int64_t size_parent, int64_t size_node, int64_t parent_ilist_size,
                           int64_t size_node_ilist, float boxsize, int dim, int dtype);
int jzt_fof_leaf2leaf_sets(jzt_stream_t stream, const int32_t* ilist_spl, const int32_t* ilist, int64_t ilist_size,
                           const int32_t* spl, const float* leaf_r2link, const void* pos,
                           const int32_t* part_igroup_in, int32_t* part_igroup, int64_t size_leaves,
                           int64_t size_part, float boxsize, int dim, int dtype);

/* ------------------------------------------------------------------------------------------ */
/* FoF catalogue                                                                                */
/* ------------------------------------------------------------------------------------------ */

/* jzt_fof_order -- the group order of the reference's _fof_order (fof.py:510-519), which runs as
 * XLA ops (argsort + scatter-add) with no FFI target.  igroup: labels of the z-ordered particles
 * (each the index of the group's first member); npart (device int32, NULL = n): the filled count.
 * isort = stable argsort of (igroup[i] for i < npart, n otherwise); counts[i] = number of filled
 * particles labelled isort[i] (the group size at each group's first member, 0 elsewhere). */
size_t jzt_fof_order_tmp_bytes(int64_t n);
int jzt_fof_order(jzt_stream_t stream, const int32_t* igroup, const int32_t* npart, int64_t n, int32_t* isort,
                  int32_t* counts, void* tmp, size_t tmp_bytes);

/* jzt_fof_catalogue -- the reductions of _fof_catalogue_from_groups (fof.py:650-766, one GPU),
 * XLA ops in the reference.  Particles in group order (AoS pos/vel of `dtype`, optional per-particle
 * mass, v_rad, zplus1; mass == NULL means every particle has mass mass_scalar) with counts[i] the
 * group size at each group's first member (jzt_fof_order).  Groups with >= npart_min members are
 * catalogued in particle order: *ngroups (device), count, offset (first particle; size_part past
 * ngroups), mass (mass array: sum; else count * mass_scalar), com_pos (centre of mass, periodic
 * differences to the first member when boxsize > 0, wrapped into the box), com_inertia_radius,
 * com_vel, v_rad, scale_factor (mass-weighted mean of 1/zplus1).  Every output but ngroups, count
 * and offset may be NULL (not computed; the inertia radius needs com_pos internally only).  Slots
 * past ngroups: count 0, offset size_part, mass 0, NaN elsewhere.  Sums run in float64. */
size_t jzt_fof_catalogue_tmp_bytes(int64_t size_part, int64_t size_cata, int dim, int dtype);
int jzt_fof_catalogue(jzt_stream_t stream, const void* pos, const void* mass, double mass_scalar, const void* vel,
                      const void* v_rad, const void* zplus1, const int32_t* counts, int64_t size_part,
                      int64_t size_cata, int npart_min, double boxsize, int dim, int dtype, int32_t* ngroups,
                      int32_t* count, int32_t* offset, void* mass_out, void* com_pos, void* com_inertia_radius,
                      void* com_vel, void* v_rad_out, void* scale_factor, void* tmp, size_t tmp_bytes);

/* jzt_graph_count_nodes: diagnostic, no counterpart in the reference.  counts[0..3] = kernel,
 * memcpy, memset and other nodes of a HIP graph (hipGraph_t passed as void*).  Lets a caller that
 * captures the pipeline check that it recorded no memset nodes. */
int jzt_graph_count_nodes(void* graph, int64_t* counts);

#ifdef __cplusplus
}
#endif

#endif /* JZTREE_AMD_H */
