/* cubql_capi.h -- C ABI of the MI355X (gfx950) BVH build-and-query library.
 *
 * Plain C: pointers, sizes and status codes; no C++ or torch types.  Every
 * entry point names the reference interface it replaces (file:line in the
 * cuBQL sources).  All device pointers are HIP device (or managed) memory.
 * Work is enqueued on the given stream; no entry point blocks the host
 * unless stated.  Return value: 0 on success, a negative cubql_status on
 * failure (cubql_last_error_string() gives the message for the calling
 * thread).  No exception or signal crosses this boundary.
 */
#ifndef CUBQL_AMD_CAPI_H
#define CUBQL_AMD_CAPI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t *cubql_stream_t; /* == hipStream_t */

/* ---------------------------------------------------------------- types */

/* box_t<float,3> (math/box.h:60-196): lower xyz then upper xyz, 24 bytes.
   A box with lower[0] > upper[0] marks an invalid (ignored) primitive. */
typedef struct { float lower[3]; float upper[3]; } cubql_box3f;

/* BinaryBVH<float,3>::Node (bvh.h:78-102): bounds plus one 64-bit word with
   offset in bits [0,48) and count in bits [48,64).  count==0: inner node,
   children at offset and offset+1; count>0: leaf with primIDs
   [offset, offset+count).  32 bytes, 16-byte aligned. */
typedef struct {
  cubql_box3f bounds;
  uint64_t offsetAndCountBits;
} cubql_node3f;

/* BinaryBVH<float,3> (bvh.h:71-111): byte-identical, 32 bytes.  Node 0 is
   the root; node 1 is unused by the GPU builders. */
typedef struct {
  cubql_node3f *nodes;
  uint32_t numNodes;
  uint32_t *primIDs;
  uint32_t numPrims;
} cubql_bvh3f;

/* WideBVH<float,3,W>::Node::Child (bvh.h:117-130): bounds plus one 64-bit
   word with valid in bit 0, offset in bits [1,46) and count in bits
   [46,62).  A valid child with count 0 is inner (its wide node is
   nodes[offset]); otherwise a leaf over primIDs [offset, offset+count).
   32 bytes, 16-byte aligned. */
typedef struct {
  cubql_box3f bounds;
  uint64_t validOffsetCount;
} cubql_wide_child3f;

/* WideBVH<float,3,W> (bvh.h:113-141), byte-identical for W = 4 and 8:
   node i is children [i*W, i*W+W) of `nodes`; node 0 is the root. */
typedef struct {
  cubql_wide_child3f *nodes;
  uint32_t numNodes;
  uint32_t *primIDs;
  uint32_t numPrims;
} cubql_wbvh3f;

/* BuildConfig (bvh.h:31-64), byte-identical: maxAllowedLeafSize,
   makeLeafThreshold, buildMethod (0 spatial median, 1 SAH, 2 ELH). */
typedef struct {
  int maxAllowedLeafSize;
  int makeLeafThreshold;
  int buildMethod;
} cubql_build_config;

enum { CUBQL_SPATIAL_MEDIAN = 0, CUBQL_SAH = 1, CUBQL_ELH = 2 };

/* GpuMemoryResource (builder/cuda.h:31-34) as a callback table.  A NULL
   resource means hipMallocAsync/hipFreeAsync on the device's default pool
   with an unbounded release threshold (builder/cuda.h:52-67).  Callbacks
   return 0 on success. */
typedef struct {
  int (*malloc_fn)(void **ptr, size_t size, cubql_stream_t s, void *user);
  int (*free_fn)(void *ptr, cubql_stream_t s, void *user);
  void *user;
} cubql_mem_resource;

typedef enum {
  CUBQL_OK = 0,
  CUBQL_ERR_INVALID_ARG = -1,
  CUBQL_ERR_HIP = -2,
  CUBQL_ERR_ALLOC = -3,
  CUBQL_ERR_UNSUPPORTED = -4,
  CUBQL_ERR_INTERNAL = -5
} cubql_status;

/* Message for the last failing call on this thread ("" if none). */
const char *cubql_last_error_string(void);
/* Library version string. */
const char *cubql_version(void);

/* ------------------------------------------------------------- builders */

/* cuBQL::gpuBuilder(BinaryBVH<float,3>&, const box3f*, uint32_t, BuildConfig,
   cudaStream_t, GpuMemoryResource&)  -- builder/cuda.h:91-99,
   dispatch gpu_builder.h:31-67.  buildMethod SPATIAL_MEDIAN runs the
   adaptive spatial-median builder (sm_builder.h:469-614), SAH the binned
   SAH builder (sah_builder.h:396-537) and ELH the binned edge-length-
   heuristic builder (elh_builder.h:437-557: 8 bins, cost = sum of box
   extents x count); makeLeafThreshold 0 means 1.  SAH and ELH honour
   maxAllowedLeafSize as cubql_sah_builder_f3 describes.
   numBoxes == 0 leaves *bvh untouched.  Spatial median: bvh->numPrims =
   numBoxes, invalid prims listed after all valid ones (sm_builder.h:590).
   Blocks the host once at the end to read back the node count. */
int cubql_gpu_builder_f3(cubql_bvh3f *bvh, const cubql_box3f *d_boxes,
                         uint32_t numBoxes, const cubql_build_config *cfg,
                         cubql_stream_t s, const cubql_mem_resource *mem);

/* cuBQL::cuda::sahBuilder -- builder/cuda.h:122-130, sah_builder.h:396-552,
   with the bin count as a parameter (the reference fixes SAHBins::numBins
   = 16, sah_builder.h:53-61; gpuBuilder with buildMethod SAH uses 16).
   numBins in [2, 64].  Leaves never exceed cfg->maxAllowedLeafSize
   (bvh.h:50-53; <= 0 means 65535): where the SAH prefers a larger leaf the
   best split is taken, and a node no plane can split is cut in the middle
   of its prim range.  bvh->numPrims = numPrims, invalid prims (lower.x >
   upper.x, sah_builder.h:140) listed after all valid ones. */
int cubql_sah_builder_f3(cubql_bvh3f *bvh, const cubql_box3f *d_boxes,
                         uint32_t numPrims, const cubql_build_config *cfg,
                         int numBins, cubql_stream_t s,
                         const cubql_mem_resource *mem);

/* cuBQL::cuda::radixBuilder -- builder/cuda.h:134-140, radix.h:591-771.
   63-bit Morton keys (21 bits per axis, radix.h:30-33).  bvh->numPrims =
   number of valid prims; primIDs in Morton order, ties by primID. */
int cubql_radix_builder_f3(cubql_bvh3f *bvh, const cubql_box3f *d_boxes,
                           uint32_t numPrims, const cubql_build_config *cfg,
                           cubql_stream_t s, const cubql_mem_resource *mem);

/* Same as cubql_radix_builder_f3 with 30-bit keys (10 bits per axis): the
   reference's radix.h:34-38 variant and the key width of
   rebinMortonBuilder.h:33-36. */
int cubql_radix30_builder_f3(cubql_bvh3f *bvh, const cubql_box3f *d_boxes,
                             uint32_t numPrims, const cubql_build_config *cfg,
                             cubql_stream_t s, const cubql_mem_resource *mem);

/* cuBQL::cuda::rebinRadixBuilder -- builder/cuda.h:148-154,
   rebinMortonBuilder.h:1166-1487 (NO_BLOCK_PHASE, its default).  30-bit
   Morton keys; a range of more than 32 prims whose keys are all equal is
   re-keyed in its own centroid domain and re-sorted (broad phase); ranges
   of <= 32 prims re-key in place with the reference's exchange sort
   (thread phase).  Leaf threshold = min(makeLeafThreshold,
   maxAllowedLeafSize), 0 -> 1.  bvh->numPrims = number of valid prims.  A
   range whose prims all fall in one cell of their own domain (extent far
   below 1e-20) is keyed by position, where the reference would not
   terminate. */
int cubql_rebin_radix_builder_f3(cubql_bvh3f *bvh, const cubql_box3f *d_boxes,
                                 uint32_t numPrims, const cubql_build_config *cfg,
                                 cubql_stream_t s, const cubql_mem_resource *mem);

/* cuBQL::cuda::free(BinaryBVH&, stream, GpuMemoryResource&) --
   builder/cuda.h:155-161, gpu_builder.h:251-261.  Frees nodes and primIDs
   with the same resource used to build, then synchronises the stream. */
int cubql_free_f3(cubql_bvh3f *bvh, cubql_stream_t s,
                  const cubql_mem_resource *mem);

/* Bottom-up refit of all node boxes from prim boxes (sm_builder.h:616-693). */
int cubql_refit_f3(cubql_bvh3f *bvh, const cubql_box3f *d_boxes,
                   cubql_stream_t s, const cubql_mem_resource *mem);

/* cuBQL::gpuBuilder(WideBVH<float,3,W>&, const box3f*, uint32_t,
   BuildConfig, cudaStream_t, GpuMemoryResource&) -- builder/cuda.h:
   107-115, wide_gpu_builder.h:162-215: builds the binary BVH as
   cubql_gpu_builder_f3 does (cfg->buildMethod selects spatial median or
   SAH), then collapses it (cubql_collapse_wide_f3) and keeps its primIDs.
   width is 4 or 8.  Blocks the host to read the node counts. */
int cubql_gpu_builder_wide_f3(cubql_wbvh3f *bvh, int width, const cubql_box3f *d_boxes,
                              uint32_t numBoxes, const cubql_build_config *cfg,
                              cubql_stream_t s, const cubql_mem_resource *mem);

/* Collapse of any BinaryBVH into a WideBVH of width 4 or 8, the
   collapseInit/collapseSummarize/collapseExecute steps of
   wide_gpu_builder.h:39-160: the binary root and every inner node whose
   depth is a multiple of log2(width) each become one wide node, whose
   children are the leaves and such roots met by a depth-first walk of its
   treelet (second child first); unused child slots are invalid (empty box,
   valid bit 0).  Wide node ids increase with the binary root's index
   (the reference assigns them with an atomic counter).  The wide tree gets
   its own copy of the primIDs; free it with cubql_free_wide_f3. */
int cubql_collapse_wide_f3(const cubql_bvh3f *binary, int width, cubql_wbvh3f *wide,
                           cubql_stream_t s, const cubql_mem_resource *mem);

/* cuBQL::cuda::free(WideBVH&, ...) -- builder/cuda.h:164-170: frees nodes
   and primIDs with the resource used to build, then synchronises. */
int cubql_free_wide_f3(cubql_wbvh3f *bvh, cubql_stream_t s, const cubql_mem_resource *mem);

/* Stable ascending sort of (key, value) pairs on key bits [0, keyBits) --
   the operation the reference builders obtain from
   cub::DeviceRadixSort::SortPairs (builder/cuda/radix.h:666-687,
   rebinMortonBuilder.h:1242-1263).  Result in (d_keys, d_vals); the Tmp
   arrays (n entries each) are scratch. */
int cubql_radix_sort_pairs_u32(uint32_t *d_keys, uint32_t *d_vals,
                               uint32_t *d_keysTmp, uint32_t *d_valsTmp,
                               uint32_t n, int keyBits, cubql_stream_t s);
int cubql_radix_sort_pairs_u64(uint64_t *d_keys, uint32_t *d_vals,
                               uint64_t *d_keysTmp, uint32_t *d_valsTmp,
                               uint32_t n, int keyBits, cubql_stream_t s);

/* -------------------------------------------------------------- queries */

/* Batched cuBQL::points::findClosest (queries/points/findClosest.h:49-62,
   181-226 over shrinkingRadiusQuery::forEachPrim, traversal/
   shrinkingRadiusQuery.h:409-508, 766-801).  For each query i: outIDs[i] =
   closest data point within sqrt(sqrMaxSearchDist) or -1; outSqrDist[i]
   = its squared distance (x*x+y*y)+z*z in float32 (or +inf).  Among equal
   distances the lowest primID is returned.  outSqrDist may be NULL.
   points/queries are float3 arrays (12-byte stride).  A query whose
   traversal needs more than 1024 stack entries (a tree deeper than 1024
   levels) is incomplete and gets id CUBQL_INCOMPLETE (-2), distance +inf;
   the reference truncates such a traversal silently
   (shrinkingRadiusQuery.h:475-478). */
#define CUBQL_INCOMPLETE (-2)
int cubql_fcp_points_f3(const cubql_bvh3f *bvh, const float *d_points,
                        const float *d_queries, uint32_t numQueries,
                        float sqrMaxSearchDist, int32_t *d_outIDs,
                        float *d_outSqrDist, cubql_stream_t s);

/* No counterpart in the reference: cubql_fcp_points_f3 for callers that
   cannot allocate (code under HIP graph capture, serving loops, harnesses
   that forbid hipMalloc).  Same arguments and results; the call allocates
   nothing and keeps its temporaries (overflow list, query ordering) in a
   static 24 MiB device array that each GPU holds until the process ends.
   At most CUBQL_FCP_NOALLOC_MAX_QUERIES queries per call (more:
   CUBQL_ERR_INVALID_ARG).  Every call on a GPU shares that array, so calls
   that overlap on one GPU (two streams or threads) race. */
#define CUBQL_FCP_NOALLOC_MAX_QUERIES 1048576u
int cubql_fcp_points_noalloc_f3(const cubql_bvh3f *bvh, const float *d_points,
                                const float *d_queries, uint32_t numQueries,
                                float sqrMaxSearchDist, int32_t *d_outIDs,
                                float *d_outSqrDist, cubql_stream_t s);

/* Batched cuBQL::points::findKNN (queries/points/knn.h:49-116 with
   knn::insert_linear, queries/common/knn.h:57-78).  Output rows of k
   entries sorted by (sqrDist, primID); unfilled entries are -1/+inf.
   d_numFound may be NULL.  1 <= k <= 64.  An incomplete query (see
   cubql_fcp_points_f3) gets numFound -1 and ids CUBQL_INCOMPLETE. */
int cubql_knn_points_f3(const cubql_bvh3f *bvh, const float *d_points,
                        const float *d_queries, uint32_t numQueries, int k,
                        float sqrMaxSearchDist, int32_t *d_outIDs,
                        float *d_outSqrDist, int32_t *d_numFound,
                        cubql_stream_t s);

/* Batched points::findClosest / findKNN over a WideBVH<float,3,W> built
   over points (queries/points/findClosest.h:82-90, 228-258 with the wide
   forEachLeaf of traversal/shrinkingRadiusQuery.h:553-636; width 4 or 8).
   Same results as cubql_fcp_points_f3 / cubql_knn_points_f3: the lowest
   primID wins equal distances, so ids and distances do not depend on the
   tree's width. */
int cubql_fcp_points_wide_f3(const cubql_wbvh3f *bvh, int width, const float *d_points,
                             const float *d_queries, uint32_t numQueries,
                             float sqrMaxSearchDist, int32_t *d_outIDs,
                             float *d_outSqrDist, cubql_stream_t s);
int cubql_knn_points_wide_f3(const cubql_wbvh3f *bvh, int width, const float *d_points,
                             const float *d_queries, uint32_t numQueries, int k,
                             float sqrMaxSearchDist, int32_t *d_outIDs,
                             float *d_outSqrDist, int32_t *d_numFound,
                             cubql_stream_t s);

/* Batched cuBQL::fixedBoxQuery::forEachPrim (traversal/fixedBoxQuery.h:
   42-160, 215-230) with a per-prim box test: for each query box i,
   d_counts[i] = number of prims whose box (d_primBoxes, the boxes the tree
   was built over) overlaps it, closed boxes as box_t::overlaps
   (math/box.h:208-212); the first min(count, maxPerQuery) of their primIDs
   go to d_outIDs[i*maxPerQuery ...] in traversal order (depth first, first
   child first).  d_outIDs may be NULL (counts only): call once for the
   counts, then again with maxPerQuery = their maximum.  A query that would
   need more than 1024 stack entries reports count 0xffffffff. */
int cubql_box_query_f3(const cubql_bvh3f *bvh, const cubql_box3f *d_primBoxes,
                       const cubql_box3f *d_queryBoxes, uint32_t numQueries,
                       uint32_t maxPerQuery, uint32_t *d_outIDs, uint32_t *d_counts,
                       cubql_stream_t s);
/* The same over a WideBVH of width 4 or 8 (fixedBoxQuery.h:163-212:
   children in slot order). */
int cubql_box_query_wide_f3(const cubql_wbvh3f *bvh, int width, const cubql_box3f *d_primBoxes,
                            const cubql_box3f *d_queryBoxes, uint32_t numQueries,
                            uint32_t maxPerQuery, uint32_t *d_outIDs, uint32_t *d_counts,
                            cubql_stream_t s);

/* Triangle i = (verts[tris[3i]], verts[tris[3i+1]], verts[tris[3i+2]]); with
   d_tris == NULL it is (verts[3*(3i)], ...): the reference's non-indexed
   Triangle array (queries/triangles/Triangle.h:26-32). */

/* Builder input boxes for triangles: Triangle::bounds() per triangle
   (Triangle.h:35-36), the generateBoxes kernel of
   samples/s02_distanceToTriangleMesh/distanceToTriangleMesh.cu:55-64. */
int cubql_triangle_boxes_f3(const int32_t *d_tris, const float *d_verts,
                            uint32_t numTris, cubql_box3f *d_boxes,
                            cubql_stream_t s);

/* Batched cuBQL::triangles::findClosest (queries/triangles/findClosest.h:
   132-152; per-triangle test closestPoint :93-122).  For each query:
   outTriIDs[i] = closest triangle within sqrt(sqrMax) or -1 (lowest id on
   equal distances), outSqrDist[i] its squared distance (float32, no FMA
   contraction), outClosest[3i..3i+2] the closest point FCPResult::P.
   outSqrDist and outClosest may be NULL.  An incomplete query (see
   cubql_fcp_points_f3) gets id CUBQL_INCOMPLETE. */
int cubql_fcp_triangles_f3(const cubql_bvh3f *bvh, const int32_t *d_tris,
                           const float *d_verts, const float *d_queries,
                           uint32_t numQueries, float sqrMaxSearchDist,
                           int32_t *d_outTriIDs, float *d_outSqrDist,
                           float *d_outClosest, cubql_stream_t s);

/* Adjoint of the unsigned distance d_i = |q_i - P_i| of a forward
   cubql_fcp_triangles_f3 call (no reference counterpart: the reference has
   no differentiation).  Recomputes the closest feature from the saved
   triangle ids; with upstream gradients g_i = dL/dd_i:
     gradQueries[i] = g_i (q_i - P_i)/d_i                 (overwritten)
     gradVerts[v]  += -g_i w_ij (q_i - P_i)/d_i           (accumulated)
   where w are the barycentric weights of P on the triangle; 0 at d = 0 and
   for ids < 0.  Either output may be NULL. */
int cubql_fcp_triangles_backward_f3(const int32_t *d_tris, const float *d_verts,
                                    const float *d_queries, uint32_t numQueries,
                                    const int32_t *d_triIDs,
                                    const float *d_gradDist,
                                    float *d_gradQueries, float *d_gradVerts,
                                    cubql_stream_t s);

/* Adjoint of point distances d_ij = |q_i - p_{ids[i*k+j]}| from a forward
   fcp (k = 1) or kNN call (no reference counterpart):
     gradQueries[i]  = sum_j g_ij (q_i - p)/d_ij          (overwritten)
     gradPoints[p]  += -g_ij (q_i - p)/d_ij               (accumulated)
   ids < 0 and d = 0 contribute nothing.  Either output may be NULL.  The
   point gradient is gathered per point (entries sorted by point id, one
   sum per point) rather than scattered with a float atomic per entry; its
   temporaries (28 bytes per entry plus the sort's) come from mem (NULL:
   hipMallocAsync).  numQueries * k < 2^32. */
int cubql_points_distance_backward_f3(const float *d_points, uint32_t numPoints,
                                      const float *d_queries,
                                      uint32_t numQueries, int k,
                                      const int32_t *d_ids,
                                      const float *d_gradDist,
                                      float *d_gradQueries,
                                      float *d_gradPoints, cubql_stream_t s,
                                      const cubql_mem_resource *mem);

/* ------------------------------------------------------------- batched */

/* Tunables of the batched path (chosen by cubql_amd.autotune):
   leafSize      leaf threshold T of every per-cloud tree (>= 1);
   blockSize     query workgroup size: 64, 128 or 256;
   stackDepth    traversal stack entries before the deep fallback: 8, 16, 32
                 or 64 (queries that need more rerun with a deeper stack);
   maxCloudSize  optional upper bound on any cloud's size (0 = unknown); it
                 only bounds the number of hierarchy launches. */
typedef struct {
  int leafSize;
  int blockSize;
  int stackDepth;
  uint32_t maxCloudSize;
} cubql_batch_params;

/* Workspace bytes for cubql_batch_chamfer_f3 (no reference counterpart). */
size_t cubql_batch_chamfer_workspace_f3(uint32_t numClouds, uint32_t numSrc,
                                        uint32_t numTgt);

/* Batched bidirectional closest point + Chamfer adjoint over numClouds
   independent cloud pairs (no reference counterpart).  Cloud c holds source
   points [srcOffsets[c], srcOffsets[c+1]) and target points likewise
   (device arrays of numClouds+1 entries that start at 0, end at numSrc or
   numTgt and never decrease: the clouds split each side).  Builds one 30-bit Morton tree per
   cloud and side (each identical to cubql_radix30_builder_f3 on that cloud
   with makeLeafThreshold = leafSize), then for every source point the
   closest target point of the same cloud (global index, -1 if the cloud has
   no target points) and vice versa, with the lowest index winning ties.
   If any of loss/gradSrc/gradTgt is given:
     loss[c] = mean_{i in S_c} |s_i - nn(s_i)| + mean_{j in T_c} |t_j - nn(t_j)|
   and gradSrc/gradTgt = d(sum_c loss[c]) / d(points) (overwritten).
   No host synchronisation and no allocation: the call can be captured in a
   HIP graph (see cubql_batch_chamfer_graph_f3). */
int cubql_batch_chamfer_f3(const float *d_src, const uint32_t *d_srcOffsets,
                           uint32_t numSrc, const float *d_tgt,
                           const uint32_t *d_tgtOffsets, uint32_t numTgt,
                           uint32_t numClouds, const cubql_batch_params *params,
                           void *d_workspace, size_t workspaceBytes,
                           int32_t *d_srcNN, float *d_srcSqrDist,
                           int32_t *d_tgtNN, float *d_tgtSqrDist, float *d_loss,
                           float *d_gradSrc, float *d_gradTgt, cubql_stream_t s);

/* The same launch sequence captured once into an executable HIP graph bound
   to these buffers; replay it with cubql_graph_launch. */
typedef struct cubql_graph_s *cubql_graph;
int cubql_batch_chamfer_graph_f3(const float *d_src, const uint32_t *d_srcOffsets,
                                 uint32_t numSrc, const float *d_tgt,
                                 const uint32_t *d_tgtOffsets, uint32_t numTgt,
                                 uint32_t numClouds,
                                 const cubql_batch_params *params,
                                 void *d_workspace, size_t workspaceBytes,
                                 int32_t *d_srcNN, float *d_srcSqrDist,
                                 int32_t *d_tgtNN, float *d_tgtSqrDist,
                                 float *d_loss, float *d_gradSrc,
                                 float *d_gradTgt, cubql_graph *out);
int cubql_graph_launch(cubql_graph g, cubql_stream_t s);
int cubql_graph_destroy(cubql_graph g);

/* ------------------------------------------------------------ profiling */

/* Event-based per-kernel timing, the counterpart of the reference's
   compile-time CUBQL_PROFILE timers (builder/cuda/sm_builder.h:24-66).  Off
   by default.  When on, builders and queries bracket their dominant kernels
   with HIP events on the caller's stream and accumulate, per kernel name,
   the number of (non-empty) launches, total milliseconds, algorithmic bytes
   and work units (prims or queries).  Kernel names: "sm_partition",
   "sm_small", "fcp_points", "knn_points", "radix_onesweep", ... */
void cubql_profile_enable(int on);
void cubql_profile_reset(void);
/* 0 on success, -1 if the kernel has no record. */
int cubql_profile_query(const char *kernel, uint64_t *launches, double *totalMs,
                        double *bytes, uint64_t *units);

/* Traversal statistics, the counterpart of the nodeDist / primCode counters
   of the reference's traversal-quality test (testing/bvh/buildQual.h:65-98,
   NUM_TRAVERSAL_STEPS): runs the same queries as cubql_fcp_points_f3
   (k == 0), cubql_knn_points_f3 (1 <= k <= 64) or cubql_fcp_triangles_f3
   with per-query counters and writes three device uint64 sums to d_out:
   bytes of node data loaded (64 per child pair expanded, 32 per node popped
   from the stack), prims (points or triangles) tested, and queries.  A
   measurement aid: slower than the plain queries, results discarded. */
int cubql_query_stats_points_f3(const cubql_bvh3f *bvh, const float *d_points,
                                const float *d_queries, uint32_t numQueries, int k,
                                float sqrMaxSearchDist, uint64_t *d_out,
                                cubql_stream_t s);
int cubql_query_stats_triangles_f3(const cubql_bvh3f *bvh, const int32_t *d_tris,
                                   const float *d_verts, const float *d_queries,
                                   uint32_t numQueries, float sqrMaxSearchDist,
                                   uint64_t *d_out, cubql_stream_t s);

#ifdef __cplusplus
}
#endif
#endif /* CUBQL_AMD_CAPI_H */
