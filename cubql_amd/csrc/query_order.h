// query_order.h -- Morton ordering of a batch of queries.
//
// Queries are processed one per lane; lanes of a wave that start from nearby
// points walk nearly the same nodes, so sorting a large batch by the Morton
// code of its queries (cells of the tree's root box) makes node loads
// coherent and cuts divergence.  Results are written back to each query's
// own slot, so the output is independent of the order.
#pragma once
#include "common.h"

namespace cubql {

// Batches below this size run in input order.
constexpr uint32_t kQueryOrderMin = 32768;

// On success *order is a device permutation (sorted position -> query
// index) valid until `scr` is destroyed, or nullptr for small batches.
int query_order(const cubql_bvh3f &bvh, const float *queries, uint32_t nq, hipStream_t s, Scratch &scr,
                const uint32_t **order);

// The same on a caller-provided workspace of query_order_ws_bytes(nq)
// bytes (256-byte aligned); *order points into it.  Allocates nothing.
size_t query_order_ws_bytes(uint32_t nq);
int query_order_ws(const cubql_bvh3f &bvh, const float *queries, uint32_t nq, hipStream_t s, void *ws,
                   const uint32_t **order);
// The same with the cells taken over a device box (rootBox == nullptr:
// no ordering).
int query_order_box_ws(const cubql_box3f *rootBox, const float *queries, uint32_t nq, hipStream_t s, void *ws,
                       const uint32_t **order);

}  // namespace cubql
