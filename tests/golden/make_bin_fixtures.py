"""Writes tests/golden/bin/: reference-format point and gold files.

Run in the build container (needs oracle/_ref, compiled from the reference
sources by oracle/Makefile):

    python tests/golden/make_bin_fixtures.py

* ``uniform_data_2000.bin`` / ``uniform_queries_1000.bin`` -- written by the
  reference's own point generator (testing/common/generatePoints.cpp, the
  'uniform' generator, seeds 290374 and 7): size_t count + vec3d items.
* ``closestPoint_gold.bin`` -- for each query, the distance to the closest
  data point: the reference's host points::findClosest over a
  cpu::spatialMedian tree, then sqrtf of fSqrDistance_rd (the quantity
  testing/points/closestPoint.h:36-45 writes with --rebuild-gold).
* ``knn5_gold.bin`` -- sqrtf of the 5th nearest squared distance from the
  reference's host points::findKNN (testing/points/knn5.h:34-46).

The reference's own closestPoint / knn5 programs do not compile from this
snapshot (they include a cuBQL/builder/host.h the sources do not hold), so
the gold values come from the same host functions through oracle/_ref.
"""
from __future__ import annotations

import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle as O  # noqa: E402
from cubql_amd import io  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "bin")


def main():
    O.build()
    gen = os.path.join(ROOT, "oracle", "_ref", "cubql_generate_points")
    if not (O.ref_available() and os.path.exists(gen)):
        raise SystemExit("oracle/_ref missing: needs the reference sources")
    os.makedirs(OUT, exist_ok=True)
    data_f = os.path.join(OUT, "uniform_data_2000.bin")
    query_f = os.path.join(OUT, "uniform_queries_1000.bin")
    for path, n, seed in [(data_f, 2000, 290374), (query_f, 1000, 7)]:
        subprocess.run([gen, "-n", str(n), "-s", str(seed), "-g", "uniform", "-o", path], check=True,
                       capture_output=True)
    P, Q = io.load_points(data_f), io.load_points(query_f)
    tree = O.ref_sm_build(O.points_to_boxes(P), 8)
    ids, d2 = O.ref_fcp(tree, P, Q)
    io.save_results(os.path.join(OUT, "closestPoint_gold.bin"), np.sqrt(d2.astype(np.float32)))
    kid, kd2, nf = O.ref_knn(tree, P, Q, 5)
    assert np.all(nf == 5)
    io.save_results(os.path.join(OUT, "knn5_gold.bin"), np.sqrt(kd2[:, 4].astype(np.float32)))
    for f in sorted(os.listdir(OUT)):
        print(f"wrote {f} ({os.path.getsize(os.path.join(OUT, f))} bytes)")


if __name__ == "__main__":
    main()
