"""The reference's binary dump format for points and results.

Counterpart of cuBQL::samples::loadBinary / saveBinary / convert
(samples/common/IO.h:31-66): a file holds a host size_t item count N (8
bytes, little endian) followed by N raw items.  Point files hold three
doubles per point (testing/common/generatePoints.cpp:50-55); gold files one
float per query (testing/points/closestPoint.h:150-160).  The C++ spelling
is include/cubql_amd/io.hpp.
"""
from __future__ import annotations

import numpy as np

_COUNT = np.dtype("<u8")


def load_binary(path: str, dtype, item_shape=()) -> np.ndarray:
    """Items of `dtype` (each of shape `item_shape`) from a dump file."""
    dt = np.dtype(dtype)
    with open(path, "rb") as f:
        head = f.read(8)
        if len(head) != 8:
            raise ValueError(f"{path}: no item count")
        n = int(np.frombuffer(head, _COUNT)[0])
        per = int(np.prod(item_shape)) if item_shape else 1
        data = np.fromfile(f, dtype=dt, count=n * per)
    if data.size != n * per:
        raise ValueError(f"{path}: {data.size // per} of {n} items present")
    return data.reshape((n,) + tuple(item_shape))


def save_binary(path: str, items: np.ndarray) -> None:
    """Writes items (first axis = item index) as a dump file."""
    a = np.ascontiguousarray(items)
    with open(path, "wb") as f:
        f.write(np.array([a.shape[0]], _COUNT).tobytes())
        f.write(a.tobytes())


def load_points(path: str) -> np.ndarray:
    """[n, 3] float32 points from a vec3d point file (each component
    converted with round-to-nearest, as convert<float,3> does)."""
    return load_binary(path, "<f8", (3,)).astype(np.float32)


def save_points(path: str, points: np.ndarray) -> None:
    """[n, 3] points as a vec3d point file."""
    save_binary(path, np.asarray(points, dtype="<f8").reshape(-1, 3))


def load_results(path: str) -> np.ndarray:
    """float32 results of a gold file."""
    return load_binary(path, "<f4")


def save_results(path: str, results: np.ndarray) -> None:
    save_binary(path, np.asarray(results, dtype="<f4").reshape(-1))
