"""Writes tests/golden/*.npz from the reference's own host path.

Run in the build container (needs oracle/_ref/libcubql_ref.so, compiled from
the reference sources by oracle/Makefile):

    python tests/golden/make_golden.py

Each fixture is data only: seeded inputs (or the spec to regenerate them) and
the reference's outputs:

* ``lcg_uniform.npz``  -- first values of UniformPointGenerator (LCG<8>) for
  the seeds/domains the workloads use (Generator.cpp:202-218).
* ``sm_tree_*.npz``    -- cpu::spatialMedian trees (nodes + primIDs) for
  seeded inputs, plus host findClosest / findKNN results on seeded queries.
* ``c1_fcp.npz``       -- the 100K/100K closest-point workload: reference
  ids and squared distances for every query, and a digest of the reference
  tree.
"""
from __future__ import annotations

import hashlib
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle as O  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def canon_digest(tree) -> str:
    return hashlib.sha256(O.canonical(tree).tobytes()).hexdigest()


def save(name, **arrays):
    path = os.path.join(OUT, name)
    np.savez_compressed(path, **arrays)
    print(f"wrote {path} ({os.path.getsize(path)} bytes)")


def main():
    O.build()
    if not O.ref_available():
        raise SystemExit("oracle/_ref/libcubql_ref.so missing: needs the reference sources")

    # --- generator
    gens = {}
    for seed, lo, hi in [(0, 0.0, 1.0), (1, -0.1, 1.1), (2, 0.0, 1.0), (3, 0.0, 1.0), (290374, 0.0, 1.0)]:
        gens[f"s{seed}"] = O.ref_uniform(seed, 1024, lo, hi)
        gens[f"s{seed}_dom"] = np.array([lo, hi])
    save("lcg_uniform.npz", **gens)

    # --- small spatial-median trees + queries
    cases = {}
    P = O.ref_uniform(0, 8000, 0.0, 1.0)
    Q = O.ref_uniform(1, 3000, -0.1, 1.1)
    cases["uniform8k"] = (O.points_to_boxes(P), P, Q, 8)
    # duplicates: 1000 copies of 3 distinct points + 10 scattered ones
    rng = np.random.default_rng(7)
    base = rng.random((3, 3), dtype=np.float32)
    dup = np.concatenate([np.repeat(base, 1000, axis=0), rng.random((10, 3), dtype=np.float32)])
    cases["dupes"] = (O.points_to_boxes(dup), dup, rng.random((500, 3), dtype=np.float32), 8)
    # all identical points
    same = np.full((777, 3), 0.25, dtype=np.float32)
    cases["identical"] = (O.points_to_boxes(same), same, rng.random((100, 3), dtype=np.float32), 4)
    # adjacent floats on one axis: midpoint rounds onto an endpoint
    a = np.float32(0.5)
    b = np.nextafter(a, np.float32(1))
    adj = np.zeros((600, 3), dtype=np.float32)
    adj[:, 0] = np.where(np.arange(600) % 3 == 0, a, b)
    adj[:, 1] = rng.random(600, dtype=np.float32) * np.float32(1e-7)
    cases["adjacent"] = (O.points_to_boxes(adj), adj, rng.random((100, 3), dtype=np.float32), 2)
    # boxes (not points) with some invalid (inverted) entries
    lo = rng.random((5000, 3), dtype=np.float32)
    ext = rng.random((5000, 3), dtype=np.float32) * np.float32(0.01)
    bx = np.concatenate([lo, lo + ext], axis=1).astype(np.float32)
    inv = rng.random(5000) < 0.05
    bx[inv, 0], bx[inv, 3] = np.float32(1.0), np.float32(0.0)
    cases["boxes_invalid"] = (bx, None, None, 3)
    # clustered skew: wide-range magnitudes
    cl = np.concatenate([rng.normal(0, 1e-3, (3000, 3)), rng.normal(100, 1e-1, (3000, 3)),
                         rng.normal(-1e4, 10, (50, 3))]).astype(np.float32)
    cases["clustered"] = (O.points_to_boxes(cl), cl, cl[::7] + np.float32(1e-4), 8)

    for name, (boxes, pts, qs, T) in cases.items():
        tree = O.ref_sm_build(boxes, T)
        arrs = dict(boxes=boxes, leaf_threshold=np.array(T), nodes=tree.nodes.view(np.uint8),
                    prim_ids=tree.prim_ids, digest=np.array(canon_digest(tree)))
        if pts is not None:
            ids, d2 = O.ref_fcp(tree, pts, qs)
            kid, kd, nf = O.ref_knn(tree, pts, qs, 8)
            arrs.update(points=pts, queries=qs, fcp_ids=ids, fcp_d2=d2, knn_ids=kid, knn_d2=kd,
                        knn_found=nf)
        save(f"sm_tree_{name}.npz", **arrs)

    # --- C1 workload: 100K points seed 0, 100K queries seed 1 in [-0.1,1.1]
    P = O.ref_uniform(0, 100000, 0.0, 1.0)
    Q = O.ref_uniform(1, 100000, -0.1, 1.1)
    tree = O.ref_sm_build(O.points_to_boxes(P), 8)
    ids, d2 = O.ref_fcp(tree, P, Q)
    kid, kd, nf = O.ref_knn(tree, P, Q[:4000], 8)
    save("c1_fcp.npz", fcp_ids=ids, fcp_d2=d2, knn_ids=kid, knn_d2=kd,
         digest=np.array(canon_digest(tree)), num_nodes=np.array(len(tree.nodes)),
         points_sha=np.array(hashlib.sha256(P.tobytes()).hexdigest()),
         queries_sha=np.array(hashlib.sha256(Q.tobytes()).hexdigest()))


if __name__ == "__main__":
    main()
