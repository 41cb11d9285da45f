// oracle/oracle.h -- TEST INFRASTRUCTURE ONLY.
//
// CPU restatement of the reference's BVH build-and-query semantics. It is the
// checker for the HIP library: tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg are the only callers. The product library never links,
// loads or calls anything in this directory.
//
// Every routine cites the reference file:line it restates (paths relative to
// the reference repository root). Arithmetic is float32 with FP contraction
// OFF (the Makefile passes -ffp-contract=off), matching the reference host
// path (g++ without FMA).
#pragma once
#include <cstdint>
#include <cstddef>

extern "C" {

/* Layout shared with the product: 24-byte box, 32-byte node whose last 8
   bytes hold offset (low 48 bits) and count (high 16 bits); bvh.h:78-102. */
typedef struct { float lo[3]; float hi[3]; } orc_box3f;
typedef struct { orc_box3f b; uint64_t oc; } orc_node3f;

/* LCG<8>(seed,index) uniform generator; random.h:24-68, Generator.cpp:202-218.
   out[i*dims+d] = float(lo[d] + rng()*(hi[d]-lo[d])) computed in double. */
void orc_lcg_uniform(uint32_t seed, int64_t n, int dims, const double *lo,
                     const double *hi, float *out);

/* Serial recursive spatial-median builder; cpu/spatialMedian.h:73-168.
   nodes capacity must be >= 2*n+2.  Output numbering: root-first DFS, no
   unused node 1, primIDs holds valid prims only (numPrims = #valid). */
int orc_sm_build(const orc_box3f *boxes, int n, int leafThreshold,
                 orc_node3f *nodes, uint32_t *primIDs, int *numNodes,
                 int *numPrims);

/* Binned SAH builder; restates cuda/sah_builder.h:55-106 (evaluateSAH),
   152-205 (binPrims), 261-330 (selectSplits/updatePrims) as a recursion.
   Bins: bin = clamp(int(rel*numBins)) with rel = (c-lo)/(hi-lo+1e-20f), bin 0
   when hi <= lo; a bin accumulates the prim BOXES and a count.  A node with
   count <= min(leafThreshold,leafCap) is a leaf; otherwise the first minimum
   of [leaf cost A*N, then for dim 0..2, bin 0..numBins-1 the split cost
   A_L*N_L + A_R*N_R over splits with both sides non-empty] decides, with
   strict '<' as in the reference.  Leaf cap (no counterpart in the
   reference's SAH builder, which ignores BuildConfig::maxAllowedLeafSize): a
   leaf choice over more than leafCap prims takes the best split instead;
   with no valid split the node's current order is cut in the middle.
   Partition order: left side stable, right side reversed (the GPU
   builder's order), so primIDs match position for position.  Node
   numbering: root-first DFS without an unused node 1. */
int orc_sah_build(const orc_box3f *boxes, int n, int leafThreshold, int numBins,
                  int leafCap, orc_node3f *nodes, uint32_t *primIDs, int *numNodes,
                  int *numPrims);

/* One SAH split decision over an explicit set of prim boxes (the root step
   of orc_sah_build) with centroid bounds taken over all n boxes.  Returns
   the split dim and *bin, -1 for a leaf, or -2 when allowLeaf is 0 and no
   split has both sides non-empty (the middle-split fallback). */
int orc_sah_decide(const orc_box3f *boxes, int n, int numBins, int allowLeaf, int *bin);

/* Binned edge-length-heuristic builder; restates cuda/elh_builder.h:
   56-135 (ELHBins, edgeLengths, evaluateELH), 184-230 (binPrims), 258-350
   (selectSplits/updatePrims) as orc_sah_build does for SAH: 8 bins,
   rel = (c-lo)/(hi-lo) (no epsilon), the box measure is the sum of the
   extents, the same leaf-first strict-'<' evaluation and the same leaf-cap
   rule (the reference's ELH builder has no cap either). */
int orc_elh_build(const orc_box3f *boxes, int n, int leafThreshold, int leafCap, orc_node3f *nodes,
                  uint32_t *primIDs, int *numNodes, int *numPrims);
int orc_elh_decide(const orc_box3f *boxes, int n, int allowLeaf, int *bin);

/* Robust rebinning Morton builder; cuda/rebinMortonBuilder.h:595-1476 with
   NO_BLOCK_PHASE (its default, :22).  30-bit keys over the valid prims
   (ascending primID before a stable sort), then per node: leaf if size <=
   leafThreshold; a range of > 32 prims splits at findSplit (no equal-keys
   fallback) or, when all its keys are equal, is re-keyed in its own
   centroid domain (the local index if that domain is one point) and stably
   re-sorted, then examined again; a range of <= 32 prims (thread phase)
   re-keys in place and applies the reference's exchange sort, and becomes a
   leaf if it still cannot split.  Numbering: root-first DFS, no unused
   node 1. */
int orc_rebin_build(const orc_box3f *boxes, int n, int leafThreshold, orc_node3f *nodes,
                    uint32_t *primIDs, int *numNodes, int *numPrims);

/* Bottom-up exact refit of node bounds from prim boxes (sm_builder.h:633-675,
   spatialMedian.h:117-134). Only nodes reachable from node 0 are written. */
void orc_refit(orc_node3f *nodes, int numNodes, const uint32_t *primIDs,
               const orc_box3f *boxes);

/* TreeChecker (testing/bvh/checkBuilders.cu:64-108) plus structural limits:
   returns 0 if every reachable node's box equals exactly the union of its
   children / leaf prim boxes, every valid prim appears exactly once, invalid
   prims appear nowhere, leaf counts are in [1, maxLeaf] (maxLeaf<=0: no cap),
   and no index is out of range.  Negative codes name the first failure. */
int orc_check_tree(const orc_node3f *nodes, int numNodes,
                   const uint32_t *primIDs, int numPrims,
                   const orc_box3f *boxes, int n, int maxLeaf,
                   int *maxDepth, int *numLeaves);

/* Canonical serialisation of a tree (node indices and leaf-internal order
   ignored): pre-order, a leaf emits [count, sorted ids...], an inner node
   emits [-1] then its children ordered by their minimum primID.  Returns the
   number of int64 written (cap = capacity), or -1 if cap is too small. */
int64_t orc_canonical(const orc_node3f *nodes, const uint32_t *primIDs,
                      int64_t *out, int64_t cap);

/* Closest point over a point BVH.  mode 0 = reference semantics (first found
   wins ties; shrinkingRadiusQuery.h:409-508 with the 64-entry stack and
   findClosest.h:181-226); mode 1 = lowest primID among equal distances (no
   stack limit).  stats[0] += node pairs visited, stats[1] += prims visited. */
void orc_fcp(const orc_node3f *nodes, const uint32_t *primIDs,
             const float *points, const float *queries, int64_t nq,
             float sqrMax, int mode, int32_t *ids, float *sqrDist,
             int64_t *stats);

/* kNN over a point BVH; mode 0 = reference insert_linear (common/knn.h:57-78,
   points/knn.h:75-116); mode 1 = (sqrDist, primID) lexicographic order.
   ids/sqrDist are [nq*k]; unfilled slots get -1 / +inf. */
void orc_knn(const orc_node3f *nodes, const uint32_t *primIDs,
             const float *points, const float *queries, int64_t nq, int k,
             float sqrMax, int mode, int32_t *ids, float *sqrDist,
             int32_t *numFound);

/* Brute force closest point, float32 reference formula (dx*dx+dy*dy)+dz*dz
   (conservativeDistances.h:61-90) with lowest-id ties. */
void orc_fcp_brute(const float *points, int64_t n, const float *queries,
                   int64_t nq, int32_t *ids, float *sqrDist);
/* Brute force in float64: exact closest distance (lowest id on ties). */
void orc_fcp_brute_f64(const float *points, int64_t n, const float *queries,
                       int64_t nq, int32_t *ids, double *sqrDist);
/* Brute force kNN, float32 formula, (d,id) order. */
void orc_knn_brute(const float *points, int64_t n, const float *queries,
                   int64_t nq, int k, int32_t *ids, float *sqrDist);

/* Morton keys; radix.h:41-70 (Quantizer), 199-241 (centroid bounds),
   283-421 (interleave, key kernel).  bitsPerDim = 21 (63-bit) or 10
   (30-bit).  Valid prims are compacted in ascending primID order.
   Returns the number of valid prims. */
int orc_morton_keys(const orc_box3f *boxes, int n, int bitsPerDim,
                    uint64_t *keys, uint32_t *ids);
/* Stable sort of (key,id) pairs by key (cub DeviceRadixSort semantics). */
void orc_sort_pairs(uint64_t *keys, uint32_t *ids, int n);
/* Top-down hierarchy over sorted keys; radix.h:446-570.  Breadth-first
   numbering, node 1 unused.  Returns numNodes. */
int orc_radix_hierarchy(const uint64_t *sortedKeys, int n, int leafThreshold,
                        orc_node3f *nodes);

/* ---------------------------------------------------------- triangles
   Triangle i is (verts[tris[3i]], verts[tris[3i+1]], verts[tris[3i+2]]);
   with tris == NULL it is verts[3i..3i+2] (the reference's non-indexed
   Triangle array).  */

/* Closest point on one triangle, float32, contraction off
   (queries/triangles/findClosest.h:74-122): returns the squared distance,
   writes the closest point P and the barycentric weights of P on (a,b,c)
   (face: projection barycentrics; edge: (1-t, t); vertex: one-hot). */
float orc_tri_closest(const float *a, const float *b, const float *c,
                      const float *q, float *P, float *w);

/* Triangle boxes = Triangle::bounds() (queries/triangles/Triangle.h:35-36). */
void orc_tri_boxes(const int32_t *tris, const float *verts, int64_t numTris,
                   orc_box3f *boxes);

/* Closest triangle over a BVH (queries/triangles/findClosest.h:132-152)
   with the lowest-id tie rule; P may be NULL. */
void orc_fcp_tris(const orc_node3f *nodes, const uint32_t *primIDs,
                  const int32_t *tris, const float *verts, const float *queries,
                  int64_t nq, float sqrMax, int32_t *ids, float *sqrDist,
                  float *P);

/* FP64 brute force: exact closest distance (textbook region method) over
   all triangles; id = lowest triangle with the minimal FP64 distance. */
void orc_fcp_tris_brute_f64(const int32_t *tris, const float *verts,
                            int64_t numTris, const float *queries, int64_t nq,
                            int32_t *ids, double *sqrDist);

/* FP64 gradient of sum_i g_i * dist(q_i, tri_{id_i}) w.r.t. queries and
   vertices (envelope theorem: d/dq = (q-P)/d, d/dv_j = -w_j (q-P)/d, with
   P and w from the FP64 closest point); zero at d == 0.  gradQ [nq*3] is
   overwritten, gradV [nv*3] accumulated. */
void orc_tri_grad_f64(const int32_t *tris, const float *verts,
                      const float *queries, int64_t nq, const int32_t *ids,
                      const double *g, double *gradQ, double *gradV);

/* FP64 gradient for points: d/dq = (q-p)/d, d/dp = -(q-p)/d (k entries per
   query, ids < 0 skipped); gradQ overwritten, gradP accumulated. */
void orc_point_grad_f64(const float *points, const float *queries, int64_t nq,
                        int k, const int32_t *ids, const double *g,
                        double *gradQ, double *gradP);

/* AtomicBox order-preserving encode/decode; builder_common.h:102-117. */
int32_t orc_atomicbox_encode(float f);
float orc_atomicbox_decode(int32_t i);

/* Collapse of a binary BVH into a W-wide one (W = 4 or 8);
   wide_gpu_builder.h:39-160.  Children are 32-byte records (box + one
   64-bit word: valid bit 0, offset bits 1-45, count bits 46-61), W per wide
   node, wide ids in increasing order of their binary root.  Returns the
   wide node count, or -1 if it exceeds cap. */
int orc_wide_collapse(const orc_node3f *nodes, int numNodes, int W, orc_node3f *wide, int cap);

/* Fixed-box queries; traversal/fixedBoxQuery.h:74-139 (binary, wide ==
   NULL) and :163-212 (wide, W children per node): for each query box,
   counts[q] = number of prims whose box overlaps it (closed boxes,
   math/box.h:208-212) and out[q*maxOut ..] the first min(count, maxOut)
   of them in traversal order (depth first, first child / lowest slot
   first). */
void orc_box_query(const orc_node3f *nodes, int numNodes, const orc_node3f *wide, int numWide, int W,
                   const uint32_t *primIDs, const orc_box3f *primBoxes, const orc_box3f *queries, int64_t nq,
                   int maxOut, uint32_t *out, uint32_t *counts);

}  // extern "C"
