// oracle/oracle.cpp -- TEST INFRASTRUCTURE ONLY (see oracle.h).
//
// CPU restatement of the reference semantics, used as the checker for the
// HIP library.  Compiled with -ffp-contract=off so float expressions round
// exactly as written.
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <numeric>
#include <functional>
#include <vector>

namespace {

constexpr uint64_t kOffMask = (1ull << 48) - 1;
inline uint64_t pack_oc(uint64_t off, uint64_t cnt) { return (off & kOffMask) | (cnt << 48); }
inline uint64_t oc_off(uint64_t oc) { return oc & kOffMask; }
inline uint32_t oc_cnt(uint64_t oc) { return (uint32_t)(oc >> 48); }

inline bool box_valid(const orc_box3f &b) { return b.lo[0] <= b.hi[0]; }

// box_t::center(): (lower+upper)/2 per dim (box.h:146-147).
inline void box_center(const orc_box3f &b, float c[3]) {
  for (int d = 0; d < 3; d++) c[d] = (b.lo[d] + b.hi[d]) / 2.f;
}

inline orc_box3f empty_box() {
  orc_box3f b;
  for (int d = 0; d < 3; d++) {
    b.lo[d] = std::numeric_limits<float>::infinity();
    b.hi[d] = -std::numeric_limits<float>::infinity();
  }
  return b;
}
inline void grow(orc_box3f &b, const orc_box3f &o) {
  for (int d = 0; d < 3; d++) {
    b.lo[d] = std::min(b.lo[d], o.lo[d]);
    b.hi[d] = std::max(b.hi[d], o.hi[d]);
  }
}
inline bool box_eq(const orc_box3f &a, const orc_box3f &b) {
  for (int d = 0; d < 3; d++)
    if (a.lo[d] != b.lo[d] || a.hi[d] != b.hi[d]) return false;
  return true;
}

// fSqrDistance_rd(vec3f,vec3f) = fSquare(x)+fSquare(y)+fSquare(z), evaluated
// left to right without contraction (conservativeDistances.h:61-90).
inline float sqr_dist_pt(const float *p, const float *q) {
  float dx = p[0] - q[0], dy = p[1] - q[1], dz = p[2] - q[2];
  return dx * dx + dy * dy + dz * dz;
}
// fSqrDistance_rd(vec3f, box3f): clamp then dot (shrinkingRadiusQuery.h:253-259).
inline float sqr_dist_box(const float *q, const orc_box3f &b) {
  float v[3];
  for (int d = 0; d < 3; d++) {
    float p = std::min(std::max(q[d], b.lo[d]), b.hi[d]);
    v[d] = p - q[d];
  }
  return v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
}

// ---------------------------------------------------------------------------
// spatial median (cpu/spatialMedian.h:73-115)
struct Topo { int offset, count; };

void sm_rec(int nodeID, int begin, int end, std::vector<Topo> &topo,
            std::vector<int> &ids, std::vector<int> &alt,
            const orc_box3f *boxes, int T) {
  if (end - begin <= T) { topo[nodeID] = {begin, end - begin}; return; }
  float clo[3], chi[3];
  for (int d = 0; d < 3; d++) { clo[d] = INFINITY; chi[d] = -INFINITY; }
  for (int i = begin; i < end; i++) {
    float c[3];
    box_center(boxes[ids[i]], c);
    for (int d = 0; d < 3; d++) { clo[d] = std::min(clo[d], c[d]); chi[d] = std::max(chi[d], c[d]); }
  }
  float sz[3] = {chi[0] - clo[0], chi[1] - clo[1], chi[2] - clo[2]};
  // arg_max for 3 dims (vec.h:639-643): ties go to the lower dimension
  int dim = sz[2] > std::max(sz[0], sz[1]) ? 2 : (sz[1] > sz[0] ? 1 : 0);
  float pos = (clo[dim] + chi[dim]) / 2.f;
  int nl = 0, nr = 0;
  for (int i = begin; i < end; i++) {
    int pid = ids[i];
    float c[3];
    box_center(boxes[pid], c);
    if (c[dim] < pos) alt[begin + nl++] = pid;
    else alt[end - ++nr] = pid;
  }
  int mid;
  if (nl && nr) mid = begin + nl;
  else if (end - begin <= std::max(1, T)) { topo[nodeID] = {begin, end - begin}; return; }
  else mid = (begin + end) / 2;
  for (int i = begin; i < end; i++) ids[i] = alt[i];
  int child = (int)topo.size();
  topo.push_back({0, 0});
  topo.push_back({0, 0});
  topo[nodeID] = {child, 0};
  sm_rec(child + 0, begin, mid, topo, ids, alt, boxes, T);
  sm_rec(child + 1, mid, end, topo, ids, alt, boxes, T);
}

// ---------------------------------------------------------------------------
// binned SAH (cuda/sah_builder.h)

// surfaceArea(box3f) (box.h:199-205)
inline float surface_area(const orc_box3f &b) {
  const float sx = b.hi[0] - b.lo[0];
  const float sy = b.hi[1] - b.lo[1];
  const float sz = b.hi[2] - b.lo[2];
  return sx * sy + sx * sz + sy * sz;
}

// edgeLengths(box) (elh_builder.h:66-72, its active #if branch): the sum of
// the box's extents, accumulated from 0 in dimension order.
inline float edge_lengths(const orc_box3f &b) {
  float sum = 0.f;
  for (int d = 0; d < 3; d++) sum += b.hi[d] - b.lo[d];
  return sum;
}

// The two binned heuristics share everything but the bin denominator's
// epsilon and the box measure: SAH (sah_builder.h) uses (hi-lo+1e-20f) and
// surfaceArea, ELH (elh_builder.h) (hi-lo) and edgeLengths.
enum { kHeurSAH = 0, kHeurELH = 1 };
inline float heur_eps(int h) { return h == kHeurELH ? 0.f : 1e-20f; }
inline float heur_measure(int h, const orc_box3f &b) {
  return h == kHeurELH ? edge_lengths(b) : surface_area(b);
}

// bin of a centre coordinate (sah_builder.h:181-187, 301-308; elh_builder.h:
// 202-213, 329-336).  The float clamp equals max(0,min(NB-1,int(rel*NB)))
// for every finite in-range value and gives the saturating GPU conversion's
// answer for NaN and +-inf.  (hi-lo)+0 is hi-lo exactly.
inline int sah_bin(float c, float lo, float hi, int nb, float eps = 1e-20f) {
  if (!(hi > lo)) return 0;
  const float rel = (c - lo) / (hi - lo + eps);
  const float f = rel * (float)nb;
  if (f >= (float)(nb - 1)) return nb - 1;
  if (f > 0.f) return (int)f;
  return 0;
}

struct SahBin { orc_box3f b; int count; };

// evaluateSAH (sah_builder.h:62-106) or evaluateELH (elh_builder.h:89-135),
// the same loop over a different box measure; allowLeaf=false drops the
// leaf candidate.  Returns the split dim or -1.
int sah_evaluate(const std::vector<SahBin> &bins, int nb, bool allowLeaf, int &splitBin, int heur = kHeurSAH) {
  float bestCost = INFINITY;
  int splitDim = -1;
  std::vector<float> rAreas(nb);
  for (int d = 0; d < 3; d++) {
    const SahBin *db = bins.data() + d * nb;
    orc_box3f box = empty_box();
    int rCount = 0;
    for (int b = nb - 1; b >= 0; --b) {
      grow(box, db[b].b);
      rCount += db[b].count;
      rAreas[b] = heur_measure(heur, box);
    }
    const float leafCost = rAreas[0] * (float)rCount;
    if (allowLeaf && leafCost < bestCost) {
      bestCost = leafCost;
      splitDim = -1;
    }
    box = empty_box();
    int lCount = 0;
    for (int b = 0; b < nb; b++) {
      const float rArea = rAreas[b];
      const float lArea = heur_measure(heur, box);
      if (lCount > 0 && rCount > 0) {
        const float cost = lArea * (float)lCount + rArea * (float)rCount;
        if (cost < bestCost) {
          bestCost = cost;
          splitDim = d;
          splitBin = b;
        }
      }
      grow(box, db[b].b);
      lCount += db[b].count;
      rCount -= db[b].count;
    }
  }
  return splitDim;
}

struct SahCtx {
  const orc_box3f *boxes;
  int T, nb, cap;
  int heur = kHeurSAH;
  std::vector<int> ids, alt;
  std::vector<Topo> topo;
};

// centroid bounds, bins, decision for the prims ids[begin,end)
// returns dim (-1 leaf, -2 middle split)
int sah_decide_range(const SahCtx &c, const int *ids, int begin, int end, float clo[3],
                     float chi[3], int &splitBin) {
  const int n = end - begin;
  for (int d = 0; d < 3; d++) { clo[d] = INFINITY; chi[d] = -INFINITY; }
  for (int i = begin; i < end; i++) {
    float ce[3];
    box_center(c.boxes[ids[i]], ce);
    for (int d = 0; d < 3; d++) { clo[d] = std::min(clo[d], ce[d]); chi[d] = std::max(chi[d], ce[d]); }
  }
  std::vector<SahBin> bins(3 * c.nb, SahBin{empty_box(), 0});
  for (int i = begin; i < end; i++) {
    const orc_box3f &pb = c.boxes[ids[i]];
    for (int d = 0; d < 3; d++) {
      // binPrims uses 0.5f*(lo+hi), equal to center() for every input
      const float pd = 0.5f * (pb.lo[d] + pb.hi[d]);
      SahBin &bn = bins[d * c.nb + sah_bin(pd, clo[d], chi[d], c.nb, heur_eps(c.heur))];
      grow(bn.b, pb);
      bn.count++;
    }
  }
  int dim = sah_evaluate(bins, c.nb, true, splitBin, c.heur);
  if (dim < 0 && n > c.cap) {
    dim = sah_evaluate(bins, c.nb, false, splitBin, c.heur);
    if (dim < 0) dim = -2;
  }
  return dim;
}

void sah_rec(SahCtx &c, int nodeID, int begin, int end) {
  const int n = end - begin;
  if (n <= std::min(c.T, c.cap)) { c.topo[nodeID] = {begin, n}; return; }
  float clo[3], chi[3];
  int bin = 0;
  const int dim = sah_decide_range(c, c.ids.data(), begin, end, clo, chi, bin);
  if (dim == -1) { c.topo[nodeID] = {begin, n}; return; }
  const int mid0 = begin + n / 2;
  int nl = 0, nr = 0;
  for (int i = begin; i < end; i++) {
    const int pid = c.ids[i];
    bool left;
    if (dim == -2) {
      left = i < mid0;
    } else {
      const orc_box3f &pb = c.boxes[pid];
      const float pd = 0.5f * (pb.lo[dim] + pb.hi[dim]);
      left = sah_bin(pd, clo[dim], chi[dim], c.nb, heur_eps(c.heur)) < bin;  // side = (bin >= splitBin)
    }
    if (left) c.alt[begin + nl++] = pid;
    else c.alt[end - ++nr] = pid;
  }
  for (int i = begin; i < end; i++) c.ids[i] = c.alt[i];
  const int mid = begin + nl;
  const int child = (int)c.topo.size();
  c.topo.push_back({0, 0});
  c.topo.push_back({0, 0});
  c.topo[nodeID] = {child, 0};
  sah_rec(c, child + 0, begin, mid);
  sah_rec(c, child + 1, mid, end);
}

orc_box3f refit_rec(orc_node3f *nodes, uint64_t id, const uint32_t *primIDs,
                    const orc_box3f *boxes) {
  orc_node3f &n = nodes[id];
  orc_box3f b = empty_box();
  uint32_t cnt = oc_cnt(n.oc);
  uint64_t off = oc_off(n.oc);
  if (cnt == 0) {
    orc_box3f l = refit_rec(nodes, off + 0, primIDs, boxes);
    orc_box3f r = refit_rec(nodes, off + 1, primIDs, boxes);
    grow(b, l);
    grow(b, r);
  } else {
    for (uint32_t i = 0; i < cnt; i++) grow(b, boxes[primIDs[off + i]]);
  }
  n.b = b;
  return b;
}

// ---------------------------------------------------------------------------
// shrinking-radius traversal (shrinkingRadiusQuery.h:409-508).  `lowestId`
// switches the cull tests from the reference's `>=`/`<` to `>`/`<=` so that
// nodes at exactly the current radius are still visited (needed for the
// lowest-primID tie rule), and lifts the 64-entry stack limit.
template <typename LeafFn>
void for_each_leaf(const orc_node3f *nodes, const uint32_t *primIDs,
                   const float *q, float sqrMax, bool lowestId, LeafFn &&leaf,
                   int64_t *stats) {
  struct Entry { uint32_t idx; float dist; };
  std::vector<Entry> stack;
  stack.reserve(64);
  const size_t stackCap = lowestId ? (size_t)-1 : 64;
  float cull = sqrMax;
  uint64_t node = nodes[0].oc;
  while (true) {
    while (true) {
      if (oc_cnt(node) != 0) break;
      uint32_t i0 = (uint32_t)oc_off(node) + 0, i1 = (uint32_t)oc_off(node) + 1;
      const orc_node3f &n0 = nodes[i0];
      const orc_node3f &n1 = nodes[i1];
      if (stats) stats[0]++;
      float d0 = sqr_dist_box(q, n0.b);
      float d1 = sqr_dist_box(q, n1.b);
      float dmin = std::min(d0, d1), dmax = std::max(d0, d1);
      bool dead = lowestId ? (dmin > cull) : (dmin >= cull);
      if (dead) { node = 0; break; }  // count 0 + dead flag via offset 0 below
      uint32_t farID;
      if (d0 < d1) { node = n0.oc; farID = i1; }
      else { node = n1.oc; farID = i0; }
      bool both = lowestId ? (dmax <= cull) : (dmax < cull);
      if (both) {
        if (stack.size() >= stackCap) return;  // reference: silent truncation
        stack.push_back({farID, dmax});
      }
    }
    if (oc_cnt(node) != 0) {
      float r = leaf(primIDs + oc_off(node), oc_cnt(node));
      if (r < 0.f) return;
      cull = std::min(cull, r);
    }
    while (true) {
      if (stack.empty()) return;
      Entry e = stack.back();
      stack.pop_back();
      if (e.dist <= cull) { node = nodes[e.idx].oc; break; }
    }
  }
}

}  // namespace

extern "C" {

void orc_lcg_uniform(uint32_t seed, int64_t n, int dims, const double *lo,
                     const double *hi, float *out) {
  for (int64_t tid = 0; tid < n; tid++) {
    // LCG<8>::init (random.h:31-44)
    uint32_t v0 = seed, v1 = (uint32_t)tid, s0 = 0;
    for (int r = 0; r < 8; r++) {
      s0 += 0x9e3779b9u;
      v0 += ((v1 << 4) + 0xa341316cu) ^ (v1 + s0) ^ ((v1 >> 5) + 0xc8013ea4u);
      v1 += ((v0 << 4) + 0xad90777du) ^ (v0 + s0) ^ ((v0 >> 5) + 0x7e95761eu);
    }
    uint32_t state = v0;
    for (int d = 0; d < dims; d++) {
      state = 1664525u * state + 1013904223u;       // random.h:59-65
      float r = (float)(state & 0x00FFFFFFu) / (float)0x01000000;
      double x = lo[d] + (double)r * (hi[d] - lo[d]);  // Generator.cpp:216
      out[tid * dims + d] = (float)x;                   // IO.h:61-66 convert
    }
  }
}

int orc_sm_build(const orc_box3f *boxes, int n, int leafThreshold,
                 orc_node3f *nodes, uint32_t *primIDs, int *numNodes,
                 int *numPrims) {
  std::vector<int> ids;
  ids.reserve(n);
  for (int i = 0; i < n; i++)
    if (box_valid(boxes[i])) ids.push_back(i);
  std::vector<int> alt(ids.size());
  std::vector<Topo> topo(1, Topo{0, 0});
  sm_rec(0, 0, (int)ids.size(), topo, ids, alt, boxes, leafThreshold);
  for (size_t i = 0; i < ids.size(); i++) primIDs[i] = (uint32_t)ids[i];
  for (size_t i = 0; i < topo.size(); i++) {
    nodes[i].oc = pack_oc((uint64_t)topo[i].offset, (uint64_t)topo[i].count);
    nodes[i].b = empty_box();
  }
  *numNodes = (int)topo.size();
  *numPrims = (int)ids.size();
  if (!ids.empty()) refit_rec(nodes, 0, primIDs, boxes);
  return 0;
}

static int binned_build(int heur, const orc_box3f *boxes, int n, int leafThreshold, int numBins, int leafCap,
                        orc_node3f *nodes, uint32_t *primIDs, int *numNodes, int *numPrims) {
  SahCtx c;
  c.heur = heur;
  c.boxes = boxes;
  c.T = std::max(1, leafThreshold);
  c.nb = numBins;
  c.cap = leafCap;
  for (int i = 0; i < n; i++)
    if (box_valid(boxes[i])) c.ids.push_back(i);
  c.alt.resize(c.ids.size());
  c.topo.assign(1, Topo{0, 0});
  if (!c.ids.empty()) sah_rec(c, 0, 0, (int)c.ids.size());
  for (size_t i = 0; i < c.ids.size(); i++) primIDs[i] = (uint32_t)c.ids[i];
  for (size_t i = 0; i < c.topo.size(); i++) {
    nodes[i].oc = pack_oc((uint64_t)c.topo[i].offset, (uint64_t)c.topo[i].count);
    nodes[i].b = empty_box();
  }
  *numNodes = c.ids.empty() ? 0 : (int)c.topo.size();
  *numPrims = (int)c.ids.size();
  if (!c.ids.empty()) refit_rec(nodes, 0, primIDs, boxes);
  return 0;
}

int orc_sah_build(const orc_box3f *boxes, int n, int leafThreshold, int numBins,
                  int leafCap, orc_node3f *nodes, uint32_t *primIDs, int *numNodes,
                  int *numPrims) {
  return binned_build(kHeurSAH, boxes, n, leafThreshold, numBins, leafCap, nodes, primIDs, numNodes, numPrims);
}

int orc_elh_build(const orc_box3f *boxes, int n, int leafThreshold, int leafCap, orc_node3f *nodes,
                  uint32_t *primIDs, int *numNodes, int *numPrims) {
  return binned_build(kHeurELH, boxes, n, leafThreshold, 8, leafCap, nodes, primIDs, numNodes, numPrims);
}

static int binned_decide(int heur, const orc_box3f *boxes, int n, int numBins, int allowLeaf, int *bin) {
  SahCtx c;
  c.heur = heur;
  c.boxes = boxes;
  c.nb = numBins;
  c.cap = allowLeaf ? n : 0;
  c.T = 0;
  std::vector<int> ids(n);
  std::iota(ids.begin(), ids.end(), 0);
  float clo[3], chi[3];
  return sah_decide_range(c, ids.data(), 0, n, clo, chi, *bin);
}

int orc_sah_decide(const orc_box3f *boxes, int n, int numBins, int allowLeaf, int *bin) {
  return binned_decide(kHeurSAH, boxes, n, numBins, allowLeaf, bin);
}

int orc_elh_decide(const orc_box3f *boxes, int n, int allowLeaf, int *bin) {
  return binned_decide(kHeurELH, boxes, n, 8, allowLeaf, bin);
}

void orc_refit(orc_node3f *nodes, int numNodes, const uint32_t *primIDs,
               const orc_box3f *boxes) {
  (void)numNodes;
  refit_rec(nodes, 0, primIDs, boxes);
}

int orc_check_tree(const orc_node3f *nodes, int numNodes,
                   const uint32_t *primIDs, int numPrims,
                   const orc_box3f *boxes, int n, int maxLeaf, int *maxDepth,
                   int *numLeaves) {
  std::vector<char> seen(n, 0);
  int depthMax = 0, leaves = 0;
  struct Item { uint64_t id; int depth; };
  // iterative post-order to compute exact bounds and compare
  struct Frame { uint64_t id; int depth; int state; orc_box3f acc; };
  std::vector<Frame> st;
  st.push_back({0, 1, 0, empty_box()});
  std::vector<orc_box3f> ret;  // returned child boxes
  int64_t visited = 0;
  while (!st.empty()) {
    Frame &f = st.back();
    if (f.id >= (uint64_t)numNodes) return -1;  // node index out of range
    const orc_node3f &nd = nodes[f.id];
    uint32_t cnt = oc_cnt(nd.oc);
    uint64_t off = oc_off(nd.oc);
    if (cnt) {
      if (++visited > (int64_t)numNodes) return -2;  // cycle
      if (maxLeaf > 0 && (int)cnt > maxLeaf) return -3;  // leaf too large
      if (off + cnt > (uint64_t)numPrims) return -4;      // prim range out of range
      orc_box3f b = empty_box();
      for (uint32_t i = 0; i < cnt; i++) {
        uint32_t p = primIDs[off + i];
        if (p >= (uint32_t)n) return -5;
        if (!box_valid(boxes[p])) return -6;  // invalid prim referenced
        if (seen[p]) return -7;               // prim referenced twice
        seen[p] = 1;
        grow(b, boxes[p]);
      }
      if (!box_eq(b, nd.b)) return -8;  // leaf box mismatch
      leaves++;
      depthMax = std::max(depthMax, f.depth);
      ret.push_back(b);
      st.pop_back();
      continue;
    }
    if (f.state == 0) {
      if (++visited > (int64_t)numNodes) return -2;
      if (off + 1 >= (uint64_t)numNodes || off == 0) return -1;
      f.state = 1;
      int d = f.depth;
      st.push_back({off + 0, d + 1, 0, empty_box()});
      continue;
    }
    if (f.state == 1) {
      f.state = 2;
      int d = f.depth;
      st.push_back({off + 1, d + 1, 0, empty_box()});
      continue;
    }
    // both children done: last two entries of ret
    orc_box3f r = ret.back(); ret.pop_back();
    orc_box3f l = ret.back(); ret.pop_back();
    orc_box3f b = empty_box();
    grow(b, l);
    grow(b, r);
    if (!box_eq(b, nd.b)) return -9;  // inner box mismatch
    ret.push_back(b);
    st.pop_back();
  }
  for (int i = 0; i < n; i++)
    if (box_valid(boxes[i]) && !seen[i]) return -10;  // valid prim missing
  if (maxDepth) *maxDepth = depthMax;
  if (numLeaves) *numLeaves = leaves;
  return 0;
}

int64_t orc_canonical(const orc_node3f *nodes, const uint32_t *primIDs,
                      int64_t *out, int64_t cap) {
  // compute min primID per node (post-order), then emit pre-order
  struct Rec {
    static uint32_t minId(const orc_node3f *nodes, const uint32_t *primIDs,
                          uint64_t id, std::vector<std::pair<uint64_t, uint32_t>> &memo) {
      const orc_node3f &nd = nodes[id];
      uint32_t cnt = oc_cnt(nd.oc);
      uint64_t off = oc_off(nd.oc);
      uint32_t m = 0xffffffffu;
      if (cnt) {
        for (uint32_t i = 0; i < cnt; i++) m = std::min(m, primIDs[off + i]);
      } else {
        m = std::min(minId(nodes, primIDs, off, memo), minId(nodes, primIDs, off + 1, memo));
      }
      memo.push_back({id, m});
      return m;
    }
  };
  // iterative emit with on-demand min computation (memoised in a hash-free
  // way: recompute is O(subtree), so precompute all mins via explicit map)
  std::vector<std::pair<uint64_t, uint32_t>> memo;
  Rec::minId(nodes, primIDs, 0, memo);
  std::sort(memo.begin(), memo.end());
  auto getMin = [&](uint64_t id) {
    auto it = std::lower_bound(memo.begin(), memo.end(), std::make_pair(id, 0u));
    return it->second;
  };
  int64_t w = 0;
  std::vector<uint64_t> st{0};
  std::vector<uint32_t> tmp;
  while (!st.empty()) {
    uint64_t id = st.back();
    st.pop_back();
    const orc_node3f &nd = nodes[id];
    uint32_t cnt = oc_cnt(nd.oc);
    uint64_t off = oc_off(nd.oc);
    if (cnt) {
      if (w + 1 + cnt > cap) return -1;
      out[w++] = cnt;
      tmp.assign(primIDs + off, primIDs + off + cnt);
      std::sort(tmp.begin(), tmp.end());
      for (uint32_t v : tmp) out[w++] = v;
    } else {
      if (w + 1 > cap) return -1;
      out[w++] = -1;
      uint64_t a = off, b = off + 1;
      if (getMin(b) < getMin(a)) std::swap(a, b);
      st.push_back(b);  // emitted second
      st.push_back(a);  // emitted first
    }
  }
  return w;
}

void orc_fcp(const orc_node3f *nodes, const uint32_t *primIDs,
             const float *points, const float *queries, int64_t nq,
             float sqrMax, int mode, int32_t *ids, float *sqrDist,
             int64_t *stats) {
  const bool lowest = mode != 0;
  for (int64_t qi = 0; qi < nq; qi++) {
    const float *q = queries + 3 * qi;
    int32_t best = -1;
    float bestD = sqrMax;
    auto leaf = [&](const uint32_t *prims, uint32_t cnt) -> float {
      float r = INFINITY;
      for (uint32_t i = 0; i < cnt; i++) {
        int32_t pid = (int32_t)prims[i];
        if (stats) stats[1]++;
        float d = sqr_dist_pt(points + 3 * (int64_t)pid, q);
        bool take = lowest ? (d < bestD || (d == bestD && best >= 0 && pid < best))
                           : !(d >= bestD);  // findClosest.h:211-218
        if (take) { bestD = d; best = pid; }
        r = std::min(r, bestD);
      }
      return r;
    };
    for_each_leaf(nodes, primIDs, q, sqrMax, lowest, leaf, stats);
    ids[qi] = best;
    sqrDist[qi] = best >= 0 ? bestD : INFINITY;
  }
}

void orc_knn(const orc_node3f *nodes, const uint32_t *primIDs,
             const float *points, const float *queries, int64_t nq, int k,
             float sqrMax, int mode, int32_t *ids, float *sqrDist,
             int32_t *numFound) {
  const bool lowest = mode != 0;
  std::vector<int32_t> cid(k);
  std::vector<float> cd(k);
  for (int64_t qi = 0; qi < nq; qi++) {
    const float *q = queries + 3 * qi;
    int found = 0;
    float maxD = sqrMax;
    auto insert = [&](int32_t pid, float d) {
      int pos;
      if (lowest) {
        if (!(d < sqrMax)) return;
        if (found < k) pos = found++;
        else {
          if (d > cd[k - 1] || (d == cd[k - 1] && pid > cid[k - 1])) return;
          pos = k - 1;
        }
        while (pos > 0 && (cd[pos - 1] > d || (cd[pos - 1] == d && cid[pos - 1] > pid))) {
          cd[pos] = cd[pos - 1];
          cid[pos] = cid[pos - 1];
          --pos;
        }
      } else {
        // knn::insert_linear (common/knn.h:57-78)
        if (found < k) pos = found++;
        else {
          if (d >= maxD) return;
          pos = k - 1;
        }
        while (pos > 0 && cd[pos - 1] > d) {
          cd[pos] = cd[pos - 1];
          cid[pos] = cid[pos - 1];
          --pos;
        }
      }
      cd[pos] = d;
      cid[pos] = pid;
      if (found == k) maxD = cd[k - 1];
    };
    auto leaf = [&](const uint32_t *prims, uint32_t cnt) -> float {
      float r = INFINITY;
      for (uint32_t i = 0; i < cnt; i++) {
        int32_t pid = (int32_t)prims[i];
        float d = sqr_dist_pt(points + 3 * (int64_t)pid, q);
        insert(pid, d);
        r = std::min(r, maxD);
      }
      return r;
    };
    for_each_leaf(nodes, primIDs, q, sqrMax, lowest, leaf, nullptr);
    for (int j = 0; j < k; j++) {
      ids[qi * k + j] = j < found ? cid[j] : -1;
      sqrDist[qi * k + j] = j < found ? cd[j] : INFINITY;
    }
    if (numFound) numFound[qi] = found;
  }
}

void orc_fcp_brute(const float *points, int64_t n, const float *queries,
                   int64_t nq, int32_t *ids, float *sqrDist) {
  for (int64_t qi = 0; qi < nq; qi++) {
    const float *q = queries + 3 * qi;
    int32_t best = -1;
    float bd = INFINITY;
    for (int64_t i = 0; i < n; i++) {
      float d = sqr_dist_pt(points + 3 * i, q);
      if (d < bd) { bd = d; best = (int32_t)i; }
    }
    ids[qi] = best;
    sqrDist[qi] = bd;
  }
}

void orc_fcp_brute_f64(const float *points, int64_t n, const float *queries,
                       int64_t nq, int32_t *ids, double *sqrDist) {
  for (int64_t qi = 0; qi < nq; qi++) {
    const float *q = queries + 3 * qi;
    int32_t best = -1;
    double bd = INFINITY;
    for (int64_t i = 0; i < n; i++) {
      double dx = (double)points[3 * i + 0] - (double)q[0];
      double dy = (double)points[3 * i + 1] - (double)q[1];
      double dz = (double)points[3 * i + 2] - (double)q[2];
      double d = dx * dx + dy * dy + dz * dz;
      if (d < bd) { bd = d; best = (int32_t)i; }
    }
    ids[qi] = best;
    sqrDist[qi] = bd;
  }
}

void orc_knn_brute(const float *points, int64_t n, const float *queries,
                   int64_t nq, int k, int32_t *ids, float *sqrDist) {
  std::vector<std::pair<float, int32_t>> all(n);
  for (int64_t qi = 0; qi < nq; qi++) {
    const float *q = queries + 3 * qi;
    for (int64_t i = 0; i < n; i++) all[i] = {sqr_dist_pt(points + 3 * i, q), (int32_t)i};
    int64_t kk = std::min<int64_t>(k, n);
    std::partial_sort(all.begin(), all.begin() + kk, all.end());
    for (int j = 0; j < k; j++) {
      ids[qi * k + j] = j < kk ? all[j].second : -1;
      sqrDist[qi * k + j] = j < kk ? all[j].first : INFINITY;
    }
  }
}

// ---------------------------------------------------------------------------
// Morton (radix.h)
static inline uint64_t spread3(uint32_t v, int bits) {
  uint64_t r = 0;
  for (int i = 0; i < bits; i++) r |= (uint64_t)((v >> i) & 1u) << (3 * i);
  return r;
}

int orc_morton_keys(const orc_box3f *boxes, int n, int bitsPerDim,
                    uint64_t *keys, uint32_t *ids) {
  // centroid bounds of valid prims (radix.h:199-227); AtomicBox is exact
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int i = 0; i < n; i++) {
    if (!box_valid(boxes[i])) continue;
    float c[3];
    for (int d = 0; d < 3; d++) c[d] = 0.5f * (boxes[i].lo[d] + boxes[i].hi[d]);
    for (int d = 0; d < 3; d++) { lo[d] = std::min(lo[d], c[d]); hi[d] = std::max(hi[d], c[d]); }
  }
  // Quantizer::init (radix.h:46-53): uniform scale from the largest extent
  float ext = std::max(std::max(hi[0] - lo[0], hi[1] - lo[1]), hi[2] - lo[2]);
  float inv = 1.f / std::max(ext, 1e-20f);
  float scale = (float)(1u << bitsPerDim) * inv;
  const uint32_t cmax = (1u << bitsPerDim) - 1;
  int w = 0;
  for (int i = 0; i < n; i++) {
    if (!box_valid(boxes[i])) continue;
    uint32_t cell[3];
    for (int d = 0; d < 3; d++) {
      float c = 0.5f * (boxes[i].lo[d] + boxes[i].hi[d]);
      float f = (c - lo[d]) * scale;            // radix.h:59
      uint32_t u = (uint32_t)f;
      cell[d] = std::min(u, cmax);              // radix.h:60
    }
    keys[w] = spread3(cell[0], bitsPerDim) | (spread3(cell[1], bitsPerDim) << 1) |
              (spread3(cell[2], bitsPerDim) << 2);
    ids[w] = (uint32_t)i;
    w++;
  }
  return w;
}

void orc_sort_pairs(uint64_t *keys, uint32_t *ids, int n) {
  std::vector<int> perm(n);
  std::iota(perm.begin(), perm.end(), 0);
  std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return keys[a] < keys[b]; });
  std::vector<uint64_t> k2(n);
  std::vector<uint32_t> i2(n);
  for (int i = 0; i < n; i++) { k2[i] = keys[perm[i]]; i2[i] = ids[perm[i]]; }
  std::copy(k2.begin(), k2.end(), keys);
  std::copy(i2.begin(), i2.end(), ids);
}

int orc_radix_hierarchy(const uint64_t *keys, int n, int T, orc_node3f *nodes) {
  // level-by-level like the reference's createNodes loop (radix.h:501-570),
  // with deterministic breadth-first numbering
  struct Range { uint32_t b, e; };
  std::vector<Range> open{{0u, (uint32_t)n}};
  std::vector<uint64_t> openIds{0};
  nodes[1].oc = 0;
  nodes[1].b = empty_box();
  int numNodes = 2;
  if (T <= 0) T = 1;
  while (!open.empty()) {
    std::vector<Range> next;
    std::vector<uint64_t> nextIds;
    for (size_t i = 0; i < open.size(); i++) {
      uint32_t b = open[i].b, e = open[i].e;
      uint64_t id = openIds[i];
      nodes[id].b = empty_box();
      int size = (int)(e - b);
      int split = -1;
      if (size > T) {
        uint64_t fk = keys[b], lk = keys[e - 1];
        if (fk == lk) split = (int)((b + e) / 2);  // radix.h:455-465
        else {
          int nmatch = __builtin_clzll(fk ^ lk);
          uint64_t search = lk & (~0ull << (63 - nmatch));
          uint32_t lo = b, hi = e;
          while (hi > lo) {
            uint32_t mid = (lo + hi) / 2;
            if (keys[mid] < search) lo = mid + 1;
            else hi = mid;
          }
          split = (int)lo;
        }
      }
      if (split < 0) {
        nodes[id].oc = pack_oc(b, (uint64_t)size);
      } else {
        uint64_t c = (uint64_t)numNodes;
        numNodes += 2;
        nodes[id].oc = pack_oc(c, 0);
        next.push_back({b, (uint32_t)split});
        nextIds.push_back(c);
        next.push_back({(uint32_t)split, e});
        nextIds.push_back(c + 1);
      }
    }
    open.swap(next);
    openIds.swap(nextIds);
  }
  return numNodes;
}

// ---------------------------------------------------------------------------
// robust rebinning Morton builder (cuda/rebinMortonBuilder.h)
namespace {

struct RebinCtx {
  const orc_box3f *boxes;
  int T;
  std::vector<uint32_t> keys, ids;
  std::vector<Topo> topo;
};

inline void center_of(const orc_box3f &b, float c[3]) {
  for (int d = 0; d < 3; d++) c[d] = (b.lo[d] + b.hi[d]) / 2.f;  // box.h:146-147
}

// Quantizer<float,3> (rebinMortonBuilder.h:38-60) + interleaveBits
// (:416-437): 10 bits per axis, bit i of x,y,z -> bits 3i,3i+1,3i+2.
struct Quant30 {
  float lo[3], scale;
  void init(const float dlo[3], const float dhi[3]) {
    for (int d = 0; d < 3; d++) lo[d] = dlo[d];
    const float ext = std::max(std::max(dhi[0] - dlo[0], dhi[1] - dlo[1]), dhi[2] - dlo[2]);
    scale = (float)(1u << 10) * (1.f / std::max(ext, 1e-20f));
  }
  uint32_t key(const float c[3]) const {
    uint32_t cell[3];
    for (int d = 0; d < 3; d++) {
      const float f = (c[d] - lo[d]) * scale;
      cell[d] = std::min((uint32_t)f, 1023u);
    }
    return (uint32_t)(spread3(cell[0], 10) | (spread3(cell[1], 10) << 1) | (spread3(cell[2], 10) << 2));
  }
};

// centroid domain of ids[b,e); returns true when lower == upper (all axes)
bool rebin_domain(const RebinCtx &c, uint32_t b, uint32_t e, Quant30 &q) {
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (uint32_t i = b; i < e; i++) {
    float ce[3];
    center_of(c.boxes[c.ids[i]], ce);
    for (int d = 0; d < 3; d++) { lo[d] = std::min(lo[d], ce[d]); hi[d] = std::max(hi[d], ce[d]); }
  }
  q.init(lo, hi);
  return lo[0] == hi[0] && lo[1] == hi[1] && lo[2] == hi[2];
}

// findSplit (rebinMortonBuilder.h:595-622): no equal-keys fallback
bool rebin_find_split(const std::vector<uint32_t> &keys, uint32_t b, uint32_t e, uint32_t &split) {
  const uint32_t fk = keys[b], lk = keys[e - 1];
  if (fk == lk) return false;
  const int nm = __builtin_clz(fk ^ lk);
  const uint32_t search = lk & (uint32_t)(0xffffffffull << (31 - nm));
  uint32_t lo = b, hi = e;
  while (hi > lo) {
    const uint32_t mid = (lo + hi) / 2;
    if (keys[mid] < search) lo = mid + 1;
    else hi = mid;
  }
  split = lo;
  return true;
}

// Broad-phase rebin of one range (rebinMortonBuilder.h:1027-1063, 1305-1405):
// new key = Morton code in the range's own centroid domain (local index if
// that domain is a single point), then a stable sort by the new key.  If a
// non-degenerate domain still gives one key for every prim (extent far
// below 1e-20), the local index is used instead -- the reference would
// rebin that range forever.
void broad_rebin(RebinCtx &c, uint32_t b, uint32_t e) {
  Quant30 q;
  const bool degenerate = rebin_domain(c, b, e, q);
  std::vector<std::pair<uint32_t, uint32_t>> kv(e - b);
  for (uint32_t i = b; i < e; i++) {
    float ce[3];
    center_of(c.boxes[c.ids[i]], ce);
    kv[i - b] = {degenerate ? i - b : q.key(ce), c.ids[i]};
  }
  bool allEqual = true;
  for (auto &p : kv) allEqual = allEqual && p.first == kv[0].first;
  if (allEqual)
    for (uint32_t i = 0; i < e - b; i++) kv[i].first = i;
  std::stable_sort(kv.begin(), kv.end(),
                   [](const std::pair<uint32_t, uint32_t> &x, const std::pair<uint32_t, uint32_t> &y) {
                     return x.first < y.first;
                   });
  for (uint32_t i = b; i < e; i++) { c.keys[i] = kv[i - b].first; c.ids[i] = kv[i - b].second; }
}

// Thread-phase rebin (rebinMortonBuilder.h:820-844): re-key in the range's
// own domain (or key = i-begin), then bubbleSort (:758-771), an unstable
// exchange sort whose exact swap sequence fixes the order of equal keys.
void thread_rebin(RebinCtx &c, uint32_t b, uint32_t e) {
  Quant30 q;
  const bool degenerate = rebin_domain(c, b, e, q);
  for (uint32_t i = b; i < e; i++) {
    if (degenerate) {
      c.keys[i] = i - b;
    } else {
      float ce[3];
      center_of(c.boxes[c.ids[i]], ce);
      c.keys[i] = q.key(ce);
    }
  }
  for (uint32_t i = b; i < e; i++)
    for (uint32_t j = i + 1; j < e; j++)
      if (c.keys[j] < c.keys[i]) {
        std::swap(c.keys[i], c.keys[j]);
        std::swap(c.ids[i], c.ids[j]);
      }
}

// A node's fate (broadPhaseCreateNodes :664-749, threadPhaseCreateNodes
// :773-880): leaf if size <= T; ranges of <= 32 prims are finished in the
// thread phase; larger ranges split by findSplit or are rebinned and
// examined again.
void rebin_rec(RebinCtx &c, int nodeID, uint32_t b, uint32_t e) {
  const uint32_t n = e - b;
  if ((int)n <= c.T) { c.topo[nodeID] = {(int)b, (int)n}; return; }
  uint32_t split = 0;
  if (n > 32) {
    while (!rebin_find_split(c.keys, b, e, split)) broad_rebin(c, b, e);
  } else if (!rebin_find_split(c.keys, b, e, split)) {
    thread_rebin(c, b, e);
    if (!rebin_find_split(c.keys, b, e, split)) { c.topo[nodeID] = {(int)b, (int)n}; return; }
  }
  const int child = (int)c.topo.size();
  c.topo.push_back({0, 0});
  c.topo.push_back({0, 0});
  c.topo[nodeID] = {child, 0};
  rebin_rec(c, child + 0, b, split);
  rebin_rec(c, child + 1, split, e);
}

}  // namespace

int orc_rebin_build(const orc_box3f *boxes, int n, int leafThreshold, orc_node3f *nodes,
                    uint32_t *primIDs, int *numNodes, int *numPrims) {
  RebinCtx c;
  c.boxes = boxes;
  c.T = std::max(1, leafThreshold);
  std::vector<uint64_t> k64(std::max(n, 1));
  std::vector<uint32_t> ids(std::max(n, 1));
  const int nv = orc_morton_keys(boxes, n, 10, k64.data(), ids.data());
  orc_sort_pairs(k64.data(), ids.data(), nv);
  c.keys.resize(nv);
  c.ids.assign(ids.begin(), ids.begin() + nv);
  for (int i = 0; i < nv; i++) c.keys[i] = (uint32_t)k64[i];
  c.topo.assign(1, Topo{0, 0});
  if (nv > 0) rebin_rec(c, 0, 0, (uint32_t)nv);
  for (int i = 0; i < nv; i++) primIDs[i] = c.ids[i];
  for (size_t i = 0; i < c.topo.size(); i++) {
    nodes[i].oc = pack_oc((uint64_t)c.topo[i].offset, (uint64_t)c.topo[i].count);
    nodes[i].b = empty_box();
  }
  *numNodes = nv ? (int)c.topo.size() : 0;
  *numPrims = nv;
  if (nv) refit_rec(nodes, 0, primIDs, boxes);
  return 0;
}

// ---------------------------------------------------------------------------
// triangles (queries/triangles/findClosest.h:56-152)
namespace {
struct V3 { float x, y, z; };
inline V3 v3(const float *p) { return {p[0], p[1], p[2]}; }
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 mul(float s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
// dot: ((0 + x*x') + y*y') + z*z' (vec.h:458-466); the leading 0 is exact
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// cross (vec.h:232-237)
inline V3 cross(V3 a, V3 b) {
  return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
// clamp(t) = min(max(t,0),1) with the device's NaN-ignoring fminf/fmaxf
inline float clamp01(float t) { return std::fmin(std::fmax(t, 0.f), 1.f); }

// Edge::closestPoint (findClosest.h:74-88): distance and parameter t
inline float edge_closest(V3 a, V3 b, V3 p, float &t, V3 &cp) {
  V3 ab = sub(b, a);
  t = dot(sub(p, a), ab) / dot(ab, ab);
  t = clamp01(t);
  cp = add(a, mul(t, ab));
  V3 d = sub(cp, p);
  return dot(d, d);
}

void tri_verts(const int32_t *tris, const float *verts, int64_t i, const float *&a,
               const float *&b, const float *&c) {
  if (tris) {
    a = verts + 3 * (int64_t)tris[3 * i + 0];
    b = verts + 3 * (int64_t)tris[3 * i + 1];
    c = verts + 3 * (int64_t)tris[3 * i + 2];
  } else {
    a = verts + 9 * i;
    b = verts + 9 * i + 3;
    c = verts + 9 * i + 6;
  }
}

// exact (FP64) closest point on a triangle, textbook region method
// (Ericson, Real-Time Collision Detection, 5.1.5); returns squared distance
double tri_closest_f64(const float *af, const float *bf, const float *cf, const float *qf,
                       double P[3], double w[3]) {
  double a[3], b[3], c[3], p[3];
  for (int i = 0; i < 3; i++) { a[i] = af[i]; b[i] = bf[i]; c[i] = cf[i]; p[i] = qf[i]; }
  auto d3 = [](const double *x, const double *y) { return x[0] * y[0] + x[1] * y[1] + x[2] * y[2]; };
  double ab[3], ac[3], ap[3], bp[3], cp[3];
  for (int i = 0; i < 3; i++) {
    ab[i] = b[i] - a[i]; ac[i] = c[i] - a[i]; ap[i] = p[i] - a[i];
    bp[i] = p[i] - b[i]; cp[i] = p[i] - c[i];
  }
  double d1 = d3(ab, ap), d2 = d3(ac, ap), d3v = d3(ab, bp), d4 = d3(ac, bp);
  double d5 = d3(ab, cp), d6 = d3(ac, cp);
  double u = 1, v = 0, ww = 0;  // weights of a, b, c
  if (d1 <= 0 && d2 <= 0) { u = 1; v = 0; ww = 0; }
  else if (d3v >= 0 && d4 <= d3v) { u = 0; v = 1; ww = 0; }
  else {
    double vc = d1 * d4 - d3v * d2;
    if (vc <= 0 && d1 >= 0 && d3v <= 0) { double t = d1 / (d1 - d3v); u = 1 - t; v = t; ww = 0; }
    else if (d6 >= 0 && d5 <= d6) { u = 0; v = 0; ww = 1; }
    else {
      double vb = d5 * d2 - d1 * d6;
      if (vb <= 0 && d2 >= 0 && d6 <= 0) { double t = d2 / (d2 - d6); u = 1 - t; v = 0; ww = t; }
      else {
        double va = d3v * d6 - d5 * d4;
        if (va <= 0 && (d4 - d3v) >= 0 && (d5 - d6) >= 0) {
          double t = (d4 - d3v) / ((d4 - d3v) + (d5 - d6)); u = 0; v = 1 - t; ww = t;
        } else {
          double den = 1.0 / (va + vb + vc);
          v = vb * den; ww = vc * den; u = 1 - v - ww;
        }
      }
    }
  }
  double s = 0;
  for (int i = 0; i < 3; i++) {
    P[i] = u * a[i] + v * b[i] + ww * c[i];
    double d = p[i] - P[i];
    s += d * d;
  }
  w[0] = u; w[1] = v; w[2] = ww;
  return s;
}
}  // namespace

float orc_tri_closest(const float *af, const float *bf, const float *cf, const float *qf, float *P,
                      float *w) {
  V3 a = v3(af), b = v3(bf), c = v3(cf), q = v3(qf);
  V3 N = cross(sub(b, a), sub(c, a));
  bool outside = (N.x == 0.f && N.y == 0.f && N.z == 0.f) ||
                 (dot(sub(q, a), cross(sub(b, a), N)) >= 0.f) ||
                 (dot(sub(q, b), cross(sub(c, b), N)) >= 0.f) ||
                 (dot(sub(q, c), cross(sub(a, c), N)) >= 0.f);
  if (outside) {
    // three edge tests in order, strictly-closer acceptance (findClosest.h:109-112)
    float best = INFINITY, t;
    V3 cp, bestP{0, 0, 0};
    float bw[3] = {0, 0, 0};
    float d = edge_closest(a, b, q, t, cp);
    if (d < best) { best = d; bestP = cp; bw[0] = 1.f - t; bw[1] = t; bw[2] = 0.f; }
    d = edge_closest(b, c, q, t, cp);
    if (d < best) { best = d; bestP = cp; bw[0] = 0.f; bw[1] = 1.f - t; bw[2] = t; }
    d = edge_closest(c, a, q, t, cp);
    if (d < best) { best = d; bestP = cp; bw[0] = t; bw[1] = 0.f; bw[2] = 1.f - t; }
    if (P) { P[0] = bestP.x; P[1] = bestP.y; P[2] = bestP.z; }
    if (w) { w[0] = bw[0]; w[1] = bw[1]; w[2] = bw[2]; }
    return best;
  }
  // normalize(N) = N * (1/sqrt(dot(N,N))) (vec.h:580-582)
  float nn = dot(N, N);
  V3 Nn = mul(1.f / std::sqrt(nn), N);
  float sd = dot(sub(q, a), Nn);
  V3 Pp = sub(q, mul(sd, Nn));
  if (P) { P[0] = Pp.x; P[1] = Pp.y; P[2] = Pp.z; }
  if (w) {
    float wa = dot(cross(sub(c, b), sub(Pp, b)), N) / nn;
    float wb = dot(cross(sub(a, c), sub(Pp, c)), N) / nn;
    w[0] = wa; w[1] = wb; w[2] = 1.f - wa - wb;
  }
  return sd * sd;
}

void orc_tri_boxes(const int32_t *tris, const float *verts, int64_t numTris, orc_box3f *boxes) {
  for (int64_t i = 0; i < numTris; i++) {
    const float *a, *b, *c;
    tri_verts(tris, verts, i, a, b, c);
    for (int d = 0; d < 3; d++) {
      boxes[i].lo[d] = std::min(std::min(a[d], b[d]), c[d]);
      boxes[i].hi[d] = std::max(std::max(a[d], b[d]), c[d]);
    }
  }
}

void orc_fcp_tris(const orc_node3f *nodes, const uint32_t *primIDs, const int32_t *tris,
                  const float *verts, const float *queries, int64_t nq, float sqrMax,
                  int32_t *ids, float *sqrDist, float *Pout) {
  for (int64_t qi = 0; qi < nq; qi++) {
    const float *q = queries + 3 * qi;
    int32_t best = -1;
    float bestD = sqrMax;
    float bestP[3] = {0, 0, 0};
    auto leaf = [&](const uint32_t *prims, uint32_t cnt) -> float {
      for (uint32_t i = 0; i < cnt; i++) {
        int32_t pid = (int32_t)prims[i];
        const float *a, *b, *c;
        tri_verts(tris, verts, pid, a, b, c);
        float P[3];
        float d = orc_tri_closest(a, b, c, q, P, nullptr);
        if (d < bestD || (d == bestD && best >= 0 && pid < best)) {
          bestD = d;
          best = pid;
          bestP[0] = P[0]; bestP[1] = P[1]; bestP[2] = P[2];
        }
      }
      return bestD;
    };
    for_each_leaf(nodes, primIDs, q, sqrMax, true, leaf, nullptr);
    ids[qi] = best;
    sqrDist[qi] = best >= 0 ? bestD : INFINITY;
    if (Pout) for (int d = 0; d < 3; d++) Pout[3 * qi + d] = bestP[d];
  }
}

void orc_fcp_tris_brute_f64(const int32_t *tris, const float *verts, int64_t numTris,
                            const float *queries, int64_t nq, int32_t *ids, double *sqrDist) {
  for (int64_t qi = 0; qi < nq; qi++) {
    double best = INFINITY;
    int32_t bi = -1;
    for (int64_t i = 0; i < numTris; i++) {
      const float *a, *b, *c;
      tri_verts(tris, verts, i, a, b, c);
      double P[3], w[3];
      double d = tri_closest_f64(a, b, c, queries + 3 * qi, P, w);
      if (d < best) { best = d; bi = (int32_t)i; }
    }
    ids[qi] = bi;
    sqrDist[qi] = best;
  }
}

void orc_tri_grad_f64(const int32_t *tris, const float *verts, const float *queries, int64_t nq,
                      const int32_t *ids, const double *g, double *gradQ, double *gradV) {
  for (int64_t qi = 0; qi < nq; qi++) {
    for (int d = 0; d < 3; d++) gradQ[3 * qi + d] = 0.0;
    const int32_t t = ids[qi];
    if (t < 0) continue;
    const float *a, *b, *c;
    tri_verts(tris, verts, t, a, b, c);
    double P[3], w[3];
    double s = tri_closest_f64(a, b, c, queries + 3 * qi, P, w);
    double dist = std::sqrt(s);
    if (!(dist > 0)) continue;
    double dir[3];
    for (int d = 0; d < 3; d++) dir[d] = ((double)queries[3 * qi + d] - P[d]) / dist;
    for (int d = 0; d < 3; d++) gradQ[3 * qi + d] = g[qi] * dir[d];
    int64_t vi[3];
    for (int j = 0; j < 3; j++) vi[j] = tris ? (int64_t)tris[3 * t + j] : 3 * (int64_t)t + j;
    for (int j = 0; j < 3; j++)
      for (int d = 0; d < 3; d++) gradV[3 * vi[j] + d] -= g[qi] * w[j] * dir[d];
  }
}

void orc_point_grad_f64(const float *points, const float *queries, int64_t nq, int k,
                        const int32_t *ids, const double *g, double *gradQ, double *gradP) {
  for (int64_t qi = 0; qi < nq; qi++) {
    for (int d = 0; d < 3; d++) gradQ[3 * qi + d] = 0.0;
    for (int j = 0; j < k; j++) {
      const int32_t p = ids[qi * k + j];
      if (p < 0) continue;
      double diff[3], s = 0;
      for (int d = 0; d < 3; d++) {
        diff[d] = (double)queries[3 * qi + d] - (double)points[3 * (int64_t)p + d];
        s += diff[d] * diff[d];
      }
      double dist = std::sqrt(s);
      if (!(dist > 0)) continue;
      for (int d = 0; d < 3; d++) {
        double gg = g[qi * k + j] * diff[d] / dist;
        gradQ[3 * qi + d] += gg;
        gradP[3 * (int64_t)p + d] -= gg;
      }
    }
  }
}

int32_t orc_atomicbox_encode(float f) {
  int32_t bits;
  std::memcpy(&bits, &f, 4);
  if (bits & (int32_t)0x80000000) bits ^= 0x7fffffff;
  return bits;
}
float orc_atomicbox_decode(int32_t bits) {
  if (bits & (int32_t)0x80000000) bits ^= 0x7fffffff;
  float f;
  std::memcpy(&f, &bits, 4);
  return f;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Wide BVH collapse; builder/cuda/wide_gpu_builder.h:39-160 restated as a
// serial pass.  A binary node is a wide root if it is node 0 or an inner
// node whose depth (root = 0) is a multiple of log2(W) (:77-98).  A wide
// node's children come from a depth-first walk of its treelet that pushes
// offset+0 then offset+1 and stops at leaves and at other wide roots
// (:117-146); unused slots are invalid with empty bounds (:147-158).  Wide
// ids: increasing binary root index (the reference takes them from an
// atomic counter, which leaves their order open).  Writes numWide*W
// children (box + valid:1|offset:45|count:16 word) and returns numWide,
// or -1 if `cap` wide nodes are too few.
int orc_wide_collapse(const orc_node3f *nodes, int numNodes, int W, orc_node3f *wide, int cap) {
  if (numNodes <= 0) return 0;
  int logW = W == 4 ? 2 : 3;
  bool node1Used = (nodes[0].oc >> 48) == 0 && (nodes[0].oc & ((1ull << 48) - 1)) == 1;
  std::vector<int> depth(numNodes, -1), wideId(numNodes, -1);
  // depth by a walk from the root (unreachable nodes keep -1)
  std::vector<int> stack = {0};
  depth[0] = 0;
  while (!stack.empty()) {
    int i = stack.back();
    stack.pop_back();
    uint64_t oc = nodes[i].oc;
    if ((oc >> 48) != 0) continue;
    int off = (int)(oc & ((1ull << 48) - 1));
    for (int c = 0; c < 2; c++) {
      depth[off + c] = depth[i] + 1;
      stack.push_back(off + c);
    }
  }
  (void)node1Used;
  int numWide = 0;
  for (int i = 0; i < numNodes; i++) {
    bool inner = (nodes[i].oc >> 48) == 0;
    if (i == 0 || (depth[i] > 0 && inner && depth[i] % logW == 0)) wideId[i] = numWide++;
  }
  if (numWide > cap) return -1;
  for (int i = 0; i < numNodes; i++) {
    if (wideId[i] < 0) continue;
    orc_node3f *target = wide + (size_t)wideId[i] * W;
    int written = 0;
    std::vector<int> st = {i};
    while (!st.empty()) {
      int x = st.back();
      st.pop_back();
      uint64_t oc = nodes[x].oc;
      uint64_t cnt = oc >> 48, off = oc & ((1ull << 48) - 1);
      if (cnt != 0 || (x != i && wideId[x] >= 0)) {
        if (written >= W) return -1;
        orc_node3f c;
        c.b = nodes[x].b;
        uint64_t o = cnt ? off : (uint64_t)wideId[x];
        c.oc = 1ull | ((o & ((1ull << 45) - 1)) << 1) | (cnt << 46);
        target[written++] = c;
      } else {
        st.push_back((int)off + 0);
        st.push_back((int)off + 1);
      }
    }
    for (; written < W; written++) {
      orc_node3f c;
      for (int d = 0; d < 3; d++) {
        c.b.lo[d] = INFINITY;
        c.b.hi[d] = -INFINITY;
      }
      c.oc = ((uint64_t)0xffffffffu << 1) | (0xffffull << 46);
      target[written] = c;
    }
  }
  return numWide;
}

// ---------------------------------------------------------------------------
// Fixed-box queries; traversal/fixedBoxQuery.h:74-139 (binary: depth first,
// the first child before the second when both overlap) and :163-212 (wide:
// children in slot order), with the per-prim test box_t::overlaps
// (math/box.h:208-212: closed boxes).  For each query, counts[q] = number of
// overlapping prims and out[q*maxOut ..] the first min(count, maxOut) of
// them in traversal order.  wide == nullptr selects the binary tree.
static bool orc_overlaps(const float *alo, const float *ahi, const float *blo, const float *bhi) {
  for (int d = 0; d < 3; d++)
    if (ahi[d] < blo[d] || bhi[d] < alo[d]) return false;
  return true;
}

static void orc_box_query_wide_rec(const orc_node3f *wide, int W, uint32_t n, const orc_box3f &Q,
                                   const std::function<void(uint64_t, uint64_t)> &leaf) {
  for (int s = 0; s < W; s++) {
    const orc_node3f &c = wide[(size_t)n * W + s];
    uint64_t w = c.oc;
    if (!(w & 1ull) || !orc_overlaps(Q.lo, Q.hi, c.b.lo, c.b.hi)) continue;
    uint64_t off = (w >> 1) & ((1ull << 45) - 1), cnt = (w >> 46) & 0xffff;
    if (cnt)
      leaf(off, cnt);
    else
      orc_box_query_wide_rec(wide, W, (uint32_t)off, Q, leaf);
  }
}

void orc_box_query(const orc_node3f *nodes, int numNodes, const orc_node3f *wide, int numWide, int W,
                   const uint32_t *primIDs, const orc_box3f *primBoxes, const orc_box3f *queries, int64_t nq,
                   int maxOut, uint32_t *out, uint32_t *counts) {
  for (int64_t q = 0; q < nq; q++) {
    const orc_box3f &Q = queries[q];
    uint32_t count = 0;
    std::function<void(uint64_t, uint64_t)> leaf = [&](uint64_t off, uint64_t cnt) {
      for (uint64_t j = 0; j < cnt; j++) {
        uint32_t p = primIDs[off + j];
        if (orc_overlaps(Q.lo, Q.hi, primBoxes[p].lo, primBoxes[p].hi)) {
          if ((int)count < maxOut) out[q * maxOut + count] = p;
          count++;
        }
      }
    };
    if (!wide && numNodes > 0) {
      // depth first; when both children overlap the first is walked first
      std::vector<uint32_t> stack = {0};
      while (!stack.empty()) {
        uint32_t i = stack.back();
        stack.pop_back();
        uint64_t oc = nodes[i].oc, cnt = oc >> 48, off = oc & ((1ull << 48) - 1);
        if (cnt) {
          leaf(off, cnt);
          continue;
        }
        bool o0 = orc_overlaps(Q.lo, Q.hi, nodes[off].b.lo, nodes[off].b.hi);
        bool o1 = orc_overlaps(Q.lo, Q.hi, nodes[off + 1].b.lo, nodes[off + 1].b.hi);
        if (o1) stack.push_back((uint32_t)off + 1);
        if (o0) stack.push_back((uint32_t)off);
      }
    } else if (wide && numWide > 0) {
      // children in slot order, each inner child's subtree before the next slot
      orc_box_query_wide_rec(wide, W, 0, Q, leaf);
    }
    counts[q] = count;
  }
}
