// cuBQL/queries/common/knn.h -- kNN candidate list (queries/common/knn.h:
// 40-78): Result{numFound, sqrDistMax}, Candidate{primID, sqrDist} and
// insert_linear.  The list is kept sorted by (sqrDist, primID): an entry
// equal in distance to existing ones goes after those with a lower primID
// and before those with a higher one (the reference sorts by distance only,
// so equal distances keep arrival order), and when the list is full a
// candidate is rejected unless it sorts before the last entry.  That makes
// the list independent of the traversal order.
#pragma once
#include "cuBQL/bvh.h"

namespace cuBQL {
namespace knn {

struct Result {
  int numFound;
  float sqrDistMax;
};

struct Candidate {
  int primID;
  float sqrDist;
};

CUBQL_BOTH bool before(const Candidate &a, const Candidate &b) {
  return a.sqrDist < b.sqrDist || (a.sqrDist == b.sqrDist && a.primID < b.primID);
}

CUBQL_BOTH void insert_linear(Candidate *candidates, int maxCandidates, Candidate newCandidate, Result &result) {
  int insertPos;
  if (result.numFound < maxCandidates) {
    insertPos = result.numFound++;
  } else {
    if (!before(newCandidate, candidates[maxCandidates - 1])) return;
    insertPos = maxCandidates - 1;
  }
  while (insertPos > 0 && before(newCandidate, candidates[insertPos - 1])) {
    candidates[insertPos] = candidates[insertPos - 1];
    --insertPos;
  }
  candidates[insertPos] = newCandidate;
  if (result.numFound == maxCandidates) result.sqrDistMax = candidates[maxCandidates - 1].sqrDist;
}

}  // namespace knn
}  // namespace cuBQL
