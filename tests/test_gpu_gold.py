"""The reference's gold-file point tests on the GPU path
(cubql_amd/cubql_testing_points, the counterpart of testing/points/
closestPoint and knn5): reference-generated point files, gold files from
the reference's host queries.

* every builder passes closestPoint and knn5 against the gold files (the
  reference's 1e-3 rule), and the results are in fact identical: a
  --rebuild-gold run writes the committed gold files byte for byte;
* a gold file with one value moved by 0.01 fails with exit status 1 and
  names the index.
"""
import os
import subprocess

import numpy as np
import pytest

from cubql_amd import io

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "golden", "bin")
EXE = os.path.join(os.path.dirname(HERE), "cubql_amd", "cubql_testing_points")
DATA = os.path.join(BIN, "uniform_data_2000.bin")
QUERIES = os.path.join(BIN, "uniform_queries_1000.bin")


def _run(test, gold, *extra):
    print(f"[gpu] cubql_testing_points {test} {' '.join(extra)}", flush=True)
    return subprocess.run([EXE, test, "-d", DATA, "-q", QUERIES, "-g", gold, *extra], capture_output=True,
                          text=True, timeout=120)


@pytest.mark.parametrize("builder", ["sm", "radix", "radix30", "rebin", "sah"])
@pytest.mark.parametrize("test", ["closestPoint", "knn5"])
def test_gold_files_pass(test, builder, tmp_path):
    gold = os.path.join(BIN, f"{test}_gold.bin")
    r = _run(test, gold, "--builder", builder)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "PASSED" in r.stdout
    out = tmp_path / "rebuilt.bin"
    r = _run(test, str(out), "--builder", builder, "--rebuild-gold")
    assert r.returncode == 0, r.stdout + r.stderr
    assert out.read_bytes() == open(gold, "rb").read()


def test_gold_file_mismatch_fails(tmp_path):
    g = io.load_results(os.path.join(BIN, "closestPoint_gold.bin")).copy()
    g[17] += np.float32(0.01)
    bad = tmp_path / "bad.bin"
    io.save_results(str(bad), g)
    r = _run("closestPoint", str(bad))
    assert r.returncode == 1
    assert "different result for index 17" in r.stdout
