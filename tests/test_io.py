"""The reference's binary dump format (cubql_amd/io.py, include/cubql_amd/
io.hpp) and the gold fixtures under tests/golden/bin.

* the point files were written by the reference's own generatePoints
  program: the loader must return exactly the doubles our LCG<8> uniform
  generator gives for the same seeds on the generator's default domain
  [-1e5, 1e5] (Generator.cpp:202-218, 330-340) -- pins loader and generator;
* the gold files must equal what brute force gives for the same float
  arithmetic: sqrt of the smallest (closestPoint) or 5th smallest (knn5)
  squared distance (x*x+y*y)+z*z in float32;
* the C++ header reads the same files and writes byte-identical ones;
* format: an 8-byte item count, then the raw items; truncated files fail.
"""
import os
import subprocess

import numpy as np
import pytest

from cubql_amd import io, synth

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "golden", "bin")
ROOT = os.path.dirname(HERE)


def test_reference_point_files_match_generator():
    for name, n, seed in [("uniform_data_2000.bin", 2000, 290374), ("uniform_queries_1000.bin", 1000, 7)]:
        raw = io.load_binary(os.path.join(BIN, name), "<f8", (3,))
        assert raw.shape == (n, 3)
        assert np.array_equal(raw, synth.lcg_uniform(seed, n, -1e5, 1e5, dtype=np.float64))
        assert np.array_equal(io.load_points(os.path.join(BIN, name)),
                              synth.lcg_uniform(seed, n, -1e5, 1e5))


def _sq(P, Q):
    d = Q[:, None, :] - P[None, :, :]
    return (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]


def test_gold_files_equal_brute_force():
    P = io.load_points(os.path.join(BIN, "uniform_data_2000.bin"))
    Q = io.load_points(os.path.join(BIN, "uniform_queries_1000.bin"))
    d2 = _sq(P, Q).astype(np.float32)
    assert d2.dtype == np.float32
    gold = io.load_results(os.path.join(BIN, "closestPoint_gold.bin"))
    assert np.array_equal(gold, np.sqrt(d2.min(axis=1)))
    gold5 = io.load_results(os.path.join(BIN, "knn5_gold.bin"))
    assert np.array_equal(gold5, np.sqrt(np.sort(d2, axis=1)[:, 4]))


def test_format_round_trip_and_truncation(tmp_path):
    p = tmp_path / "r.bin"
    r = np.array([1.5, np.inf, 0.0], np.float32)
    io.save_results(str(p), r)
    b = p.read_bytes()
    assert b[:8] == (3).to_bytes(8, "little") and len(b) == 8 + 12
    assert np.array_equal(io.load_results(str(p)), r)
    pts = np.arange(12, dtype=np.float64).reshape(4, 3) / 7
    io.save_points(str(tmp_path / "p.bin"), pts)
    assert np.array_equal(io.load_binary(str(tmp_path / "p.bin"), "<f8", (3,)), pts)
    (tmp_path / "t.bin").write_bytes(b[:-4])
    with pytest.raises(ValueError):
        io.load_results(str(tmp_path / "t.bin"))


def test_cpp_header_reads_and_writes_the_format(tmp_path):
    src = tmp_path / "io_check.cpp"
    src.write_text(
        '#include "cubql_amd/io.hpp"\n'
        "int main(int, char **av) {\n"
        "  auto p = cuBQL::samples::loadPointsAsFloat3(av[1]);\n"
        "  cuBQL::samples::saveBinary(av[2], p);\n"
        "  auto g = cuBQL::samples::loadBinary<float>(av[3]);\n"
        "  cuBQL::samples::saveBinary(av[4], g);\n"
        "  return 0;\n}\n")
    exe = tmp_path / "io_check"
    subprocess.run(["g++", "-std=c++17", "-O1", "-I" + os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    out_p, out_g = tmp_path / "p.bin", tmp_path / "g.bin"
    subprocess.run([str(exe), os.path.join(BIN, "uniform_data_2000.bin"), str(out_p),
                    os.path.join(BIN, "knn5_gold.bin"), str(out_g)], check=True)
    flat = io.load_binary(str(out_p), "<f4")
    assert np.array_equal(flat.reshape(-1, 3), io.load_points(os.path.join(BIN, "uniform_data_2000.bin")))
    assert out_g.read_bytes() == open(os.path.join(BIN, "knn5_gold.bin"), "rb").read()


def test_device_generator_equals_numpy_generator():
    """lcg_uniform_torch (the on-device generator the benchmark uses for its
    10M points and queries) gives exactly lcg_uniform's values; run here
    with torch on the CPU, the same integer ops as on the GPU."""
    import torch
    for seed, n, lo, hi in [(2, 300_000, 0.0, 1.0), (3, 100_000, 0.0, 1.0), (1, 50_000, -0.1, 1.1),
                            (290374, 20_000, -1e5, 1e5)]:
        t = synth.lcg_uniform_torch(seed, n, lo, hi, device="cpu").numpy()
        assert np.array_equal(t, synth.lcg_uniform(seed, n, lo, hi)), seed
