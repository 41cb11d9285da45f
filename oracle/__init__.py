"""oracle -- TEST INFRASTRUCTURE ONLY.

numpy/ctypes front end for the CPU checker libraries built by oracle/Makefile:

* ``liboracle.so``  -- our C++ restatement of the reference semantics
  (oracle.cpp; each routine cites the reference file:line it follows);
* ``_ref/libcubql_ref.so`` -- the reference's own host path (cpu::spatialMedian,
  host points::findClosest / findKNN, UniformPointGenerator), compiled from the
  reference sources in place.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
import this package, and only as the checker.  The product (``cubql_amd``)
never imports it.

Parity status per function (what pins it) is listed in the module docstrings
of tests/test_oracle.py.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
NODE_DTYPE = np.dtype([("lo", "<f4", 3), ("hi", "<f4", 3), ("oc", "<u8")])
assert NODE_DTYPE.itemsize == 32

_lib = None
_ref = None


def build(quiet: bool = True) -> None:
    """Compile the oracle (and the reference shim when the reference is present)."""
    out = subprocess.run(["make", "-C", HERE, "-j4"], capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + out.stdout + out.stderr)
    if not quiet:
        print(out.stdout)


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        _lib = C.CDLL(path)
        V, I, I64, F, U32 = C.c_void_p, C.c_int, C.c_int64, C.c_float, C.c_uint32
        sig = {
            "orc_lcg_uniform": (None, [U32, I64, I, V, V, V]),
            "orc_sm_build": (I, [V, I, I, V, V, V, V]),
            "orc_sah_build": (I, [V, I, I, I, I, V, V, V, V]),
            "orc_sah_decide": (I, [V, I, I, I, V]),
            "orc_elh_build": (I, [V, I, I, I, V, V, V, V]),
            "orc_elh_decide": (I, [V, I, I, V]),
            "orc_rebin_build": (I, [V, I, I, V, V, V, V]),
            "orc_refit": (None, [V, I, V, V]),
            "orc_check_tree": (I, [V, I, V, I, V, I, I, V, V]),
            "orc_canonical": (I64, [V, V, V, I64]),
            "orc_fcp": (None, [V, V, V, V, I64, F, I, V, V, V]),
            "orc_knn": (None, [V, V, V, V, I64, I, F, I, V, V, V]),
            "orc_fcp_brute": (None, [V, I64, V, I64, V, V]),
            "orc_fcp_brute_f64": (None, [V, I64, V, I64, V, V]),
            "orc_knn_brute": (None, [V, I64, V, I64, I, V, V]),
            "orc_morton_keys": (I, [V, I, I, V, V]),
            "orc_sort_pairs": (None, [V, V, I]),
            "orc_radix_hierarchy": (I, [V, I, I, V]),
            "orc_wide_collapse": (I, [V, I, I, V, I]),
            "orc_box_query": (None, [V, I, V, I, I, V, V, V, I64, I, V, V]),
            "orc_atomicbox_encode": (C.c_int32, [F]),
            "orc_atomicbox_decode": (F, [C.c_int32]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(_lib, name)
            fn.restype = res
            fn.argtypes = args
    return _lib


def ref_available() -> bool:
    return os.path.exists(os.path.join(HERE, "_ref", "libcubql_ref.so"))


def ref():
    """The compiled reference host path (raises if it was never built)."""
    global _ref
    if _ref is None:
        path = os.path.join(HERE, "_ref", "libcubql_ref.so")
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        _ref = C.CDLL(path)
        V, I, F, D = C.c_void_p, C.c_int, C.c_float, C.c_double
        sig = {
            "ref_uniform_points": (I, [I, I, D, D, V]),
            "ref_cpu_build": (I, [V, I, I, V, V, V, V]),
            "ref_fcp": (None, [V, I, V, I, V, V, I, F, V, V]),
            "ref_knn": (None, [V, I, V, I, V, V, I, I, F, V, V, V]),
            "ref_time_cpu_build": (D, [V, I, I]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(_ref, name)
            fn.restype = res
            fn.argtypes = args
    return _ref


# ----------------------------------------------------------------------------
# helpers
def points_to_boxes(points: np.ndarray) -> np.ndarray:
    """box3f().including(p) for each point (closestPoint.cu:32-39): [n,6] f32."""
    p = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
    return np.ascontiguousarray(np.concatenate([p, p], axis=1))


def lcg_uniform(seed: int, n: int, lo=0.0, hi=1.0, dims: int = 3) -> np.ndarray:
    lo_a = np.full(dims, lo, dtype=np.float64)
    hi_a = np.full(dims, hi, dtype=np.float64)
    out = np.empty((n, dims), dtype=np.float32)
    lib().orc_lcg_uniform(seed & 0xFFFFFFFF, n, dims, _p(lo_a), _p(hi_a), _p(out))
    return out


class Tree:
    """A BVH in host memory: nodes (NODE_DTYPE) + primIDs (uint32)."""

    def __init__(self, nodes: np.ndarray, prim_ids: np.ndarray):
        self.nodes = np.ascontiguousarray(nodes.view(NODE_DTYPE).reshape(-1))
        self.prim_ids = np.ascontiguousarray(prim_ids, dtype=np.uint32)

    @property
    def counts(self):
        return (self.nodes["oc"] >> np.uint64(48)).astype(np.int64)

    @property
    def offsets(self):
        return (self.nodes["oc"] & np.uint64((1 << 48) - 1)).astype(np.int64)


def sm_build(boxes: np.ndarray, leaf_threshold: int) -> Tree:
    b = np.ascontiguousarray(boxes, dtype=np.float32).reshape(-1, 6)
    n = b.shape[0]
    nodes = np.zeros(2 * n + 2, dtype=NODE_DTYPE)
    ids = np.zeros(max(n, 1), dtype=np.uint32)
    nn = C.c_int()
    npr = C.c_int()
    lib().orc_sm_build(_p(b), n, leaf_threshold, _p(nodes), _p(ids), C.byref(nn), C.byref(npr))
    return Tree(nodes[: nn.value], ids[: npr.value])


def sah_build(boxes: np.ndarray, leaf_threshold: int, num_bins: int = 16,
              leaf_cap: int = 1 << 15) -> Tree:
    """Binned SAH tree (restates cuda/sah_builder.h; see oracle.h)."""
    b = np.ascontiguousarray(boxes, dtype=np.float32).reshape(-1, 6)
    n = b.shape[0]
    nodes = np.zeros(2 * n + 2, dtype=NODE_DTYPE)
    ids = np.zeros(max(n, 1), dtype=np.uint32)
    nn = C.c_int()
    npr = C.c_int()
    lib().orc_sah_build(_p(b), n, leaf_threshold, num_bins, leaf_cap, _p(nodes), _p(ids),
                        C.byref(nn), C.byref(npr))
    return Tree(nodes[: nn.value], ids[: npr.value])


def elh_build(boxes: np.ndarray, leaf_threshold: int, leaf_cap: int = 1 << 15) -> Tree:
    """Binned edge-length-heuristic tree (restates cuda/elh_builder.h; see oracle.h)."""
    b = np.ascontiguousarray(boxes, dtype=np.float32).reshape(-1, 6)
    n = b.shape[0]
    nodes = np.zeros(2 * n + 2, dtype=NODE_DTYPE)
    ids = np.zeros(max(n, 1), dtype=np.uint32)
    nn = C.c_int()
    npr = C.c_int()
    lib().orc_elh_build(_p(b), n, leaf_threshold, leaf_cap, _p(nodes), _p(ids), C.byref(nn), C.byref(npr))
    return Tree(nodes[: nn.value], ids[: npr.value])


def elh_decide(boxes: np.ndarray, allow_leaf: bool = True):
    """(dim, bin) of one ELH decision over all boxes; dim -1 = leaf, -2 = no split."""
    b = np.ascontiguousarray(boxes, dtype=np.float32).reshape(-1, 6)
    bn = C.c_int(0)
    d = lib().orc_elh_decide(_p(b), b.shape[0], 1 if allow_leaf else 0, C.byref(bn))
    return d, bn.value


def rebin_build(boxes: np.ndarray, leaf_threshold: int) -> Tree:
    """Robust rebinning Morton tree (restates cuda/rebinMortonBuilder.h)."""
    b = np.ascontiguousarray(boxes, dtype=np.float32).reshape(-1, 6)
    n = b.shape[0]
    nodes = np.zeros(2 * n + 2, dtype=NODE_DTYPE)
    ids = np.zeros(max(n, 1), dtype=np.uint32)
    nn = C.c_int()
    npr = C.c_int()
    lib().orc_rebin_build(_p(b), n, leaf_threshold, _p(nodes), _p(ids), C.byref(nn), C.byref(npr))
    return Tree(nodes[: nn.value], ids[: npr.value])


def sah_decide(boxes: np.ndarray, num_bins: int, allow_leaf: bool = True):
    """(dim, bin) of one SAH decision over all boxes; dim -1 = leaf, -2 = no split."""
    b = np.ascontiguousarray(boxes, dtype=np.float32).reshape(-1, 6)
    bn = C.c_int(0)
    d = lib().orc_sah_decide(_p(b), b.shape[0], num_bins, 1 if allow_leaf else 0, C.byref(bn))
    return d, bn.value


def ref_sm_build(boxes: np.ndarray, leaf_threshold: int) -> Tree:
    b = np.ascontiguousarray(boxes, dtype=np.float32).reshape(-1, 6)
    n = b.shape[0]
    nodes = np.zeros(2 * n + 2, dtype=NODE_DTYPE)
    ids = np.zeros(max(n, 1), dtype=np.uint32)
    nn = C.c_int()
    npr = C.c_int()
    ref().ref_cpu_build(_p(b), n, leaf_threshold, _p(nodes), _p(ids), C.byref(nn), C.byref(npr))
    return Tree(nodes[: nn.value], ids[: npr.value])


def check_tree(tree: Tree, boxes: np.ndarray, max_leaf: int = 0):
    """Returns (code, max_depth, num_leaves); code 0 means valid."""
    b = np.ascontiguousarray(boxes, dtype=np.float32).reshape(-1, 6)
    md = C.c_int()
    nl = C.c_int()
    code = lib().orc_check_tree(_p(tree.nodes), len(tree.nodes), _p(tree.prim_ids),
                                len(tree.prim_ids), _p(b), b.shape[0], max_leaf,
                                C.byref(md), C.byref(nl))
    return code, md.value, nl.value


def canonical(tree: Tree) -> np.ndarray:
    cap = 2 * len(tree.prim_ids) + 2 * len(tree.nodes) + 16
    out = np.empty(cap, dtype=np.int64)
    w = lib().orc_canonical(_p(tree.nodes), _p(tree.prim_ids), _p(out), cap)
    if w < 0:
        raise RuntimeError("canonical buffer too small")
    return out[:w].copy()


def fcp(tree: Tree, points, queries, sqr_max=np.inf, lowest_id=True, stats=False):
    P = np.ascontiguousarray(points, dtype=np.float32)
    Q = np.ascontiguousarray(queries, dtype=np.float32)
    nq = Q.shape[0]
    ids = np.empty(nq, dtype=np.int32)
    d2 = np.empty(nq, dtype=np.float32)
    st = np.zeros(2, dtype=np.int64)
    lib().orc_fcp(_p(tree.nodes), _p(tree.prim_ids), _p(P), _p(Q), nq, float(sqr_max),
                  1 if lowest_id else 0, _p(ids), _p(d2), _p(st))
    return (ids, d2, st) if stats else (ids, d2)


def knn(tree: Tree, points, queries, k, sqr_max=np.inf, lowest_id=True):
    P = np.ascontiguousarray(points, dtype=np.float32)
    Q = np.ascontiguousarray(queries, dtype=np.float32)
    nq = Q.shape[0]
    ids = np.empty((nq, k), dtype=np.int32)
    d2 = np.empty((nq, k), dtype=np.float32)
    nf = np.empty(nq, dtype=np.int32)
    lib().orc_knn(_p(tree.nodes), _p(tree.prim_ids), _p(P), _p(Q), nq, k, float(sqr_max),
                  1 if lowest_id else 0, _p(ids), _p(d2), _p(nf))
    return ids, d2, nf


def ref_fcp(tree: Tree, points, queries, sqr_max=np.inf):
    P = np.ascontiguousarray(points, dtype=np.float32)
    Q = np.ascontiguousarray(queries, dtype=np.float32)
    nq = Q.shape[0]
    ids = np.empty(nq, dtype=np.int32)
    d2 = np.empty(nq, dtype=np.float32)
    ref().ref_fcp(_p(tree.nodes), len(tree.nodes), _p(tree.prim_ids), len(tree.prim_ids),
                  _p(P), _p(Q), nq, float(sqr_max), _p(ids), _p(d2))
    return ids, d2


def ref_knn(tree: Tree, points, queries, k, sqr_max=np.inf):
    P = np.ascontiguousarray(points, dtype=np.float32)
    Q = np.ascontiguousarray(queries, dtype=np.float32)
    nq = Q.shape[0]
    ids = np.empty((nq, k), dtype=np.int32)
    d2 = np.empty((nq, k), dtype=np.float32)
    nf = np.empty(nq, dtype=np.int32)
    ref().ref_knn(_p(tree.nodes), len(tree.nodes), _p(tree.prim_ids), len(tree.prim_ids),
                  _p(P), _p(Q), nq, k, float(sqr_max), _p(ids), _p(d2), _p(nf))
    return ids, d2, nf


def ref_uniform(seed: int, n: int, lo: float, hi: float) -> np.ndarray:
    out = np.empty((n, 3), dtype=np.float32)
    ref().ref_uniform_points(seed, n, lo, hi, _p(out))
    return out


def fcp_brute(points, queries):
    P = np.ascontiguousarray(points, dtype=np.float32)
    Q = np.ascontiguousarray(queries, dtype=np.float32)
    ids = np.empty(Q.shape[0], dtype=np.int32)
    d2 = np.empty(Q.shape[0], dtype=np.float32)
    lib().orc_fcp_brute(_p(P), P.shape[0], _p(Q), Q.shape[0], _p(ids), _p(d2))
    return ids, d2


def fcp_brute_f64(points, queries):
    P = np.ascontiguousarray(points, dtype=np.float32)
    Q = np.ascontiguousarray(queries, dtype=np.float32)
    ids = np.empty(Q.shape[0], dtype=np.int32)
    d2 = np.empty(Q.shape[0], dtype=np.float64)
    lib().orc_fcp_brute_f64(_p(P), P.shape[0], _p(Q), Q.shape[0], _p(ids), _p(d2))
    return ids, d2


def knn_brute(points, queries, k):
    P = np.ascontiguousarray(points, dtype=np.float32)
    Q = np.ascontiguousarray(queries, dtype=np.float32)
    ids = np.empty((Q.shape[0], k), dtype=np.int32)
    d2 = np.empty((Q.shape[0], k), dtype=np.float32)
    lib().orc_knn_brute(_p(P), P.shape[0], _p(Q), Q.shape[0], k, _p(ids), _p(d2))
    return ids, d2


def morton_keys(boxes, bits_per_dim: int):
    b = np.ascontiguousarray(boxes, dtype=np.float32).reshape(-1, 6)
    n = b.shape[0]
    keys = np.empty(n, dtype=np.uint64)
    ids = np.empty(n, dtype=np.uint32)
    nv = lib().orc_morton_keys(_p(b), n, bits_per_dim, _p(keys), _p(ids))
    return keys[:nv].copy(), ids[:nv].copy()


def sort_pairs(keys, ids):
    k = np.ascontiguousarray(keys, dtype=np.uint64).copy()
    v = np.ascontiguousarray(ids, dtype=np.uint32).copy()
    lib().orc_sort_pairs(_p(k), _p(v), len(k))
    return k, v


def radix_build(boxes, bits_per_dim: int, leaf_threshold: int) -> Tree:
    """Morton keys -> stable sort -> top-down hierarchy -> exact refit."""
    keys, ids = morton_keys(boxes, bits_per_dim)
    keys, ids = sort_pairs(keys, ids)
    n = len(keys)
    nodes = np.zeros(2 * n + 2, dtype=NODE_DTYPE)
    nn = lib().orc_radix_hierarchy(_p(keys), n, leaf_threshold, _p(nodes))
    tree = Tree(nodes[:nn], ids)
    b = np.ascontiguousarray(boxes, dtype=np.float32).reshape(-1, 6)
    lib().orc_refit(_p(tree.nodes), nn, _p(tree.prim_ids), _p(b))
    return tree


def wide_collapse(tree: Tree, width: int):
    """W-wide collapse of a binary tree (restates wide_gpu_builder.h:39-160):
    returns [numWide, width] children as NODE_DTYPE records whose 'oc' word
    is valid:1 | offset:45 | count:16."""
    cap = max(1, len(tree.nodes))
    out = np.zeros(cap * width, dtype=NODE_DTYPE)
    n = lib().orc_wide_collapse(_p(tree.nodes), len(tree.nodes), width, _p(out), cap)
    if n < 0:
        raise RuntimeError("wide collapse failed")
    return out[: n * width].reshape(n, width)


def box_query(tree: Tree, prim_boxes, query_boxes, max_out: int, wide=None, width: int = 4):
    """(ids [nq, max_out] uint32, counts [nq] uint32) of the fixed-box query
    in traversal order (restates traversal/fixedBoxQuery.h)."""
    pb = np.ascontiguousarray(prim_boxes, dtype=np.float32).reshape(-1, 6)
    qb = np.ascontiguousarray(query_boxes, dtype=np.float32).reshape(-1, 6)
    nq = qb.shape[0]
    out = np.zeros((nq, max(max_out, 1)), dtype=np.uint32)
    cnt = np.zeros(nq, dtype=np.uint32)
    w = None if wide is None else np.ascontiguousarray(wide.reshape(-1)).view(NODE_DTYPE)
    lib().orc_box_query(_p(tree.nodes), len(tree.nodes), _p(w), 0 if wide is None else wide.shape[0], width,
                        _p(tree.prim_ids), _p(pb), _p(qb), nq, max_out, _p(out), _p(cnt))
    return out[:, :max_out], cnt


def wide_fields(children):
    """(valid, offset, count) arrays of wide child records."""
    w = children["oc"].astype(np.uint64)
    return ((w & np.uint64(1)).astype(bool), ((w >> np.uint64(1)) & np.uint64((1 << 45) - 1)).astype(np.int64),
            ((w >> np.uint64(46)) & np.uint64(0xffff)).astype(np.int64))


# ----------------------------------------------------------------------------
# triangles and gradients
def _tri_sigs():
    L = lib()
    V, I64, F = C.c_void_p, C.c_int64, C.c_float
    for name, (res, args) in {
        "orc_tri_closest": (F, [V, V, V, V, V, V]),
        "orc_tri_boxes": (None, [V, V, I64, V]),
        "orc_fcp_tris": (None, [V, V, V, V, V, I64, F, V, V, V]),
        "orc_fcp_tris_brute_f64": (None, [V, V, I64, V, I64, V, V]),
        "orc_tri_grad_f64": (None, [V, V, V, I64, V, V, V, V]),
        "orc_point_grad_f64": (None, [V, V, I64, C.c_int, V, V, V, V]),
    }.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    return L


def _tris_arg(tris):
    return None if tris is None else np.ascontiguousarray(tris, dtype=np.int32)


def tri_boxes(tris, verts):
    t, v = _tris_arg(tris), np.ascontiguousarray(verts, dtype=np.float32)
    n = t.shape[0] if t is not None else v.shape[0] // 3
    out = np.empty((n, 6), dtype=np.float32)
    _tri_sigs().orc_tri_boxes(_p(t), _p(v), n, _p(out))
    return out


def fcp_tris(tree: Tree, tris, verts, queries, sqr_max=np.inf):
    t, v = _tris_arg(tris), np.ascontiguousarray(verts, dtype=np.float32)
    Q = np.ascontiguousarray(queries, dtype=np.float32)
    nq = Q.shape[0]
    ids = np.empty(nq, dtype=np.int32)
    d2 = np.empty(nq, dtype=np.float32)
    P = np.empty((nq, 3), dtype=np.float32)
    _tri_sigs().orc_fcp_tris(_p(tree.nodes), _p(tree.prim_ids), _p(t), _p(v), _p(Q), nq,
                             float(sqr_max), _p(ids), _p(d2), _p(P))
    return ids, d2, P


def fcp_tris_brute_f64(tris, verts, queries):
    t, v = _tris_arg(tris), np.ascontiguousarray(verts, dtype=np.float32)
    Q = np.ascontiguousarray(queries, dtype=np.float32)
    n = t.shape[0] if t is not None else v.shape[0] // 3
    ids = np.empty(Q.shape[0], dtype=np.int32)
    d2 = np.empty(Q.shape[0], dtype=np.float64)
    _tri_sigs().orc_fcp_tris_brute_f64(_p(t), _p(v), n, _p(Q), Q.shape[0], _p(ids), _p(d2))
    return ids, d2


def tri_grad_f64(tris, verts, queries, ids, g):
    t, v = _tris_arg(tris), np.ascontiguousarray(verts, dtype=np.float32)
    Q = np.ascontiguousarray(queries, dtype=np.float32)
    gq = np.zeros((Q.shape[0], 3), dtype=np.float64)
    gv = np.zeros((v.shape[0], 3), dtype=np.float64)
    _tri_sigs().orc_tri_grad_f64(_p(t), _p(v), _p(Q), Q.shape[0],
                                 _p(np.ascontiguousarray(ids, dtype=np.int32)),
                                 _p(np.ascontiguousarray(g, dtype=np.float64)), _p(gq), _p(gv))
    return gq, gv


def point_grad_f64(points, queries, ids, g):
    P = np.ascontiguousarray(points, dtype=np.float32)
    Q = np.ascontiguousarray(queries, dtype=np.float32)
    ids2 = np.ascontiguousarray(ids, dtype=np.int32)
    k = 1 if ids2.ndim == 1 else ids2.shape[1]
    gq = np.zeros((Q.shape[0], 3), dtype=np.float64)
    gp = np.zeros((P.shape[0], 3), dtype=np.float64)
    _tri_sigs().orc_point_grad_f64(_p(P), _p(Q), Q.shape[0], k, _p(ids2),
                                   _p(np.ascontiguousarray(g, dtype=np.float64)), _p(gq), _p(gp))
    return gq, gp


def tri_closest(a, b, c, q):
    """(sqrDist, P[3], w[3]) of the float32 restatement for one triangle."""
    f = lambda x: np.ascontiguousarray(x, dtype=np.float32)  # noqa: E731
    a, b, c, q = f(a), f(b), f(c), f(q)
    P = np.empty(3, dtype=np.float32)
    w = np.empty(3, dtype=np.float32)
    d = _tri_sigs().orc_tri_closest(_p(a), _p(b), _p(c), _p(q), _p(P), _p(w))
    return d, P, w
