"""cubql_amd -- MI355X (gfx950) BVH build-and-query, Python host binding.

Python mirror of the reference's host interface for the build-and-query
path, over the C ABI in ``include/cubql_amd/cubql_capi.h``:

* ``gpuBuilder(boxes, BuildConfig(...))``  -- cuBQL::gpuBuilder
  (builder/cuda.h:91-99)
* ``cuda.radixBuilder / cuda.radix30Builder`` -- cuBQL::cuda::radixBuilder
  (builder/cuda.h:134-140)
* ``cuda.free(bvh)`` -- cuBQL::cuda::free (builder/cuda.h:155-161)
* ``points.findClosest / points.findKNN`` -- batched versions of the device
  functions in queries/points/findClosest.h and queries/points/knn.h

Device memory and streams come from PyTorch (plumbing only): BVH arrays are
allocated through a memory resource backed by torch's caching allocator,
and work is enqueued on torch's current stream unless a stream is given.

The compute path is the HIP library ``libcubql_amd.so``; importing this
package on a machine where it is missing raises immediately -- there is no
CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import math
import os
from typing import Optional

import torch

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcubql_amd.so")

SPATIAL_MEDIAN, SAH, ELH = 0, 1, 2


class CubqlError(RuntimeError):
    pass


class Box3f(C.Structure):
    _fields_ = [("lower", C.c_float * 3), ("upper", C.c_float * 3)]


class BVH3f(C.Structure):
    """Byte-identical to BinaryBVH<float,3> (bvh.h:71-111)."""
    _fields_ = [("nodes", C.c_void_p), ("numNodes", C.c_uint32),
                ("primIDs", C.c_void_p), ("numPrims", C.c_uint32)]


class WBVH3f(C.Structure):
    """Byte-identical to WideBVH<float,3,W> (bvh.h:113-141)."""
    _fields_ = [("nodes", C.c_void_p), ("numNodes", C.c_uint32),
                ("primIDs", C.c_void_p), ("numPrims", C.c_uint32)]


class BuildConfig(C.Structure):
    """Byte-identical to cuBQL::BuildConfig (bvh.h:31-64)."""
    _fields_ = [("maxAllowedLeafSize", C.c_int), ("makeLeafThreshold", C.c_int),
                ("buildMethod", C.c_int)]

    def __init__(self, makeLeafThreshold: int = 0, maxAllowedLeafSize: int = 1 << 15,
                 buildMethod: int = SPATIAL_MEDIAN):
        super().__init__(maxAllowedLeafSize, makeLeafThreshold, buildMethod)

    def enableSAH(self):
        self.buildMethod = SAH
        return self

    def enableELH(self):
        """BuildConfig::enableELH (bvh.h:33): the edge-length-heuristic builder."""
        self.buildMethod = ELH
        return self


_MALLOC_CB = C.CFUNCTYPE(C.c_int, C.POINTER(C.c_void_p), C.c_size_t, C.c_void_p, C.c_void_p)
_FREE_CB = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_void_p, C.c_void_p)


class BatchParams(C.Structure):
    """cubql_batch_params: leafSize, blockSize (64/128/256), stackDepth (8/16/32/64), maxCloudSize."""
    _fields_ = [("leafSize", C.c_int), ("blockSize", C.c_int), ("stackDepth", C.c_int),
                ("maxCloudSize", C.c_uint32)]

    def __init__(self, leafSize: int = 8, blockSize: int = 256, stackDepth: int = 64,
                 maxCloudSize: int = 0):
        super().__init__(leafSize, blockSize, stackDepth, maxCloudSize)

    def as_tuple(self):
        return (self.leafSize, self.blockSize, self.stackDepth, self.maxCloudSize)


class MemResource(C.Structure):
    _fields_ = [("malloc_fn", _MALLOC_CB), ("free_fn", _FREE_CB), ("user", C.c_void_p)]


_lib = None


def lib():
    """Load libcubql_amd.so (raises if it was never built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise CubqlError(f"{LIB_PATH} not found: build it with __graft_entry__.build() "
                         "(or make -C cubql_amd/csrc); there is no CPU fallback")
    L = C.CDLL(LIB_PATH)
    V, U32, I, F = C.c_void_p, C.c_uint32, C.c_int, C.c_float
    P = C.POINTER
    sigs = {
        "cubql_last_error_string": (C.c_char_p, []),
        "cubql_version": (C.c_char_p, []),
        "cubql_gpu_builder_f3": (I, [P(BVH3f), V, U32, P(BuildConfig), V, P(MemResource)]),
        "cubql_radix_builder_f3": (I, [P(BVH3f), V, U32, P(BuildConfig), V, P(MemResource)]),
        "cubql_radix30_builder_f3": (I, [P(BVH3f), V, U32, P(BuildConfig), V, P(MemResource)]),
        "cubql_rebin_radix_builder_f3": (I, [P(BVH3f), V, U32, P(BuildConfig), V, P(MemResource)]),
        "cubql_sah_builder_f3": (I, [P(BVH3f), V, U32, P(BuildConfig), C.c_int, V, P(MemResource)]),
        "cubql_free_f3": (I, [P(BVH3f), V, P(MemResource)]),
        "cubql_refit_f3": (I, [P(BVH3f), V, V, P(MemResource)]),
        "cubql_fcp_points_f3": (I, [P(BVH3f), V, V, U32, F, V, V, V]),
        "cubql_fcp_points_noalloc_f3": (I, [P(BVH3f), V, V, U32, F, V, V, V]),
        "cubql_knn_points_f3": (I, [P(BVH3f), V, V, U32, I, F, V, V, V, V]),
        "cubql_gpu_builder_wide_f3": (I, [P(WBVH3f), I, V, U32, P(BuildConfig), V, P(MemResource)]),
        "cubql_collapse_wide_f3": (I, [P(BVH3f), I, P(WBVH3f), V, P(MemResource)]),
        "cubql_free_wide_f3": (I, [P(WBVH3f), V, P(MemResource)]),
        "cubql_fcp_points_wide_f3": (I, [P(WBVH3f), I, V, V, U32, F, V, V, V]),
        "cubql_box_query_f3": (I, [P(BVH3f), V, V, U32, U32, V, V, V]),
        "cubql_box_query_wide_f3": (I, [P(WBVH3f), I, V, V, U32, U32, V, V, V]),
        "cubql_knn_points_wide_f3": (I, [P(WBVH3f), I, V, V, U32, I, F, V, V, V, V]),
        "cubql_radix_sort_pairs_u32": (I, [V, V, V, V, U32, I, V]),
        "cubql_radix_sort_pairs_u64": (I, [V, V, V, V, U32, I, V]),
        "cubql_triangle_boxes_f3": (I, [V, V, U32, V, V]),
        "cubql_fcp_triangles_f3": (I, [P(BVH3f), V, V, V, U32, F, V, V, V, V]),
        "cubql_fcp_triangles_backward_f3": (I, [V, V, V, U32, V, V, V, V, V]),
        "cubql_points_distance_backward_f3": (I, [V, U32, V, U32, I, V, V, V, V, V, P(MemResource)]),
        "cubql_batch_chamfer_workspace_f3": (C.c_size_t, [U32, U32, U32]),
        "cubql_batch_chamfer_f3": (I, [V, V, U32, V, V, U32, U32, P(BatchParams), V, C.c_size_t,
                                       V, V, V, V, V, V, V, V]),
        "cubql_batch_chamfer_graph_f3": (I, [V, V, U32, V, V, U32, U32, P(BatchParams), V,
                                             C.c_size_t, V, V, V, V, V, V, V, P(V)]),
        "cubql_graph_launch": (I, [V, V]),
        "cubql_graph_destroy": (I, [V]),
        "cubql_query_stats_points_f3": (I, [P(BVH3f), V, V, U32, I, F, V, V]),
        "cubql_query_stats_triangles_f3": (I, [P(BVH3f), V, V, V, U32, F, V, V]),
        "cubql_profile_enable": (None, [I]),
        "cubql_profile_reset": (None, []),
        "cubql_profile_query": (I, [C.c_char_p, P(C.c_uint64), P(C.c_double), P(C.c_double),
                                    P(C.c_uint64)]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(L, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def exported_symbols():
    """Names declared in the C ABI header that the loaded library exports."""
    L = lib()
    return {n for n in _declared_symbols() if hasattr(L, n)}


def _declared_symbols():
    hdr = os.path.join(os.path.dirname(HERE), "include", "cubql_amd", "cubql_capi.h")
    import re
    names = set()
    with open(hdr) as f:
        for m in re.finditer(r"\b(cubql_[a-z0-9_]+)\s*\(", f.read()):
            names.add(m.group(1))
    return names


def _check(status: int, what: str):
    if status != 0:
        msg = lib().cubql_last_error_string()
        raise CubqlError(f"{what} failed ({status}): {msg.decode() if msg else ''}")


def _stream_handle(stream) -> int:
    if stream is None:
        stream = torch.cuda.current_stream()
    return stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream)


def _ptr(t: Optional[torch.Tensor]):
    return None if t is None else C.c_void_p(t.data_ptr())


class TorchMemory:
    """cubql_mem_resource backed by torch's caching allocator.

    Each block is allocated with the library's stream as torch's current
    stream, so the caching allocator ties it to that stream: once the
    library frees it (at the end of a call, while its kernels may still be
    queued), torch hands it out again only to allocations on the same
    stream, which run after those kernels (stream-ordered reuse, like
    hipMallocAsync/hipFreeAsync)."""

    def __init__(self, device=None):
        self.device = torch.device(device) if device is not None else torch.device(
            "cuda", torch.cuda.current_device())
        self.live = {}

        def _malloc(pp, size, stream, user):
            try:
                st = (torch.cuda.ExternalStream(stream, device=self.device) if stream
                      else torch.cuda.default_stream(self.device))
                with torch.cuda.stream(st):
                    t = torch.empty(int(size), dtype=torch.uint8, device=self.device)
                self.live[t.data_ptr()] = t
                pp[0] = t.data_ptr()
                return 0
            except Exception:  # noqa: BLE001 -- reported as a status code
                return 1

        def _free(p, stream, user):
            self.live.pop(p, None)
            return 0

        self._m = _MALLOC_CB(_malloc)
        self._f = _FREE_CB(_free)
        self.resource = MemResource(self._m, self._f, None)

    def tensor(self, ptr: int) -> torch.Tensor:
        return self.live[ptr]


class BinaryBVH:
    """A BinaryBVH<float,3> whose arrays live in torch-managed device memory."""

    def __init__(self, mem: TorchMemory):
        self.mem = mem
        self.c = BVH3f()

    @property
    def numNodes(self) -> int:
        return int(self.c.numNodes)

    @property
    def numPrims(self) -> int:
        return int(self.c.numPrims)

    def nodes_tensor(self) -> torch.Tensor:
        """[numNodes, 8] float32 view (lo xyz, hi xyz, admin word as 2 x u32)."""
        if not self.c.nodes or self.numNodes == 0:
            return torch.empty((0, 8), dtype=torch.float32, device=self.mem.device)
        t = self.mem.tensor(self.c.nodes)
        return t[: self.numNodes * 32].view(torch.float32).view(self.numNodes, 8)

    def prim_ids_tensor(self) -> torch.Tensor:
        if not self.c.primIDs:
            return torch.empty((0,), dtype=torch.int32, device=self.mem.device)
        t = self.mem.tensor(self.c.primIDs)
        return t[: self.numPrims * 4].view(torch.int32)

    def free(self, stream=None):
        if self.c.nodes or self.c.primIDs:
            _check(lib().cubql_free_f3(C.byref(self.c), C.c_void_p(_stream_handle(stream)),
                                       C.byref(self.mem.resource)), "cuda::free")

    def __del__(self):
        try:
            if torch is not None and (self.c.nodes or self.c.primIDs):
                self.mem.live.pop(self.c.nodes, None)
                self.mem.live.pop(self.c.primIDs, None)
        except Exception:  # noqa: BLE001
            pass


class WideBVH:
    """A WideBVH<float,3,W> (W = 4 or 8) in torch-managed device memory."""

    def __init__(self, mem: TorchMemory, width: int):
        if width not in (4, 8):
            raise CubqlError("WideBVH width must be 4 or 8")
        self.mem = mem
        self.width = width
        self.c = WBVH3f()

    @property
    def numNodes(self) -> int:
        return int(self.c.numNodes)

    @property
    def numPrims(self) -> int:
        return int(self.c.numPrims)

    def nodes_tensor(self) -> torch.Tensor:
        """[numNodes, W, 8] float32 view of the children (lo xyz, hi xyz,
        valid:1 | offset:45 | count:16 word as 2 x u32)."""
        if not self.c.nodes or self.numNodes == 0:
            return torch.empty((0, self.width, 8), dtype=torch.float32, device=self.mem.device)
        t = self.mem.tensor(self.c.nodes)
        n = self.numNodes * self.width
        return t[: n * 32].view(torch.float32).view(self.numNodes, self.width, 8)

    def prim_ids_tensor(self) -> torch.Tensor:
        if not self.c.primIDs:
            return torch.empty((0,), dtype=torch.int32, device=self.mem.device)
        return self.mem.tensor(self.c.primIDs)[: self.numPrims * 4].view(torch.int32)

    def free(self, stream=None):
        if self.c.nodes or self.c.primIDs:
            _check(lib().cubql_free_wide_f3(C.byref(self.c), C.c_void_p(_stream_handle(stream)),
                                            C.byref(self.mem.resource)), "cuda::free(WideBVH)")

    def __del__(self):
        try:
            if torch is not None and (self.c.nodes or self.c.primIDs):
                self.mem.live.pop(self.c.nodes, None)
                self.mem.live.pop(self.c.primIDs, None)
        except Exception:  # noqa: BLE001
            pass


def wideGpuBuilder(boxes: torch.Tensor, width: int = 4, config: Optional[BuildConfig] = None,
                   stream=None, mem: Optional[TorchMemory] = None) -> WideBVH:
    """cuBQL::gpuBuilder(WideBVH<float,3,W>&, ...) (builder/cuda.h:107-115):
    binary build as gpuBuilder, collapsed to width 4 or 8."""
    boxes = _as_boxes(boxes)
    config = config or BuildConfig()
    mem = mem or TorchMemory(boxes.device)
    bvh = WideBVH(mem, width)
    _check(lib().cubql_gpu_builder_wide_f3(C.byref(bvh.c), int(width), _ptr(boxes), boxes.shape[0],
                                           C.byref(config), C.c_void_p(_stream_handle(stream)),
                                           C.byref(mem.resource)), "gpuBuilder(WideBVH)")
    return bvh


def collapse_wide(bvh: "BinaryBVH", width: int = 4, stream=None) -> WideBVH:
    """WideBVH of width 4 or 8 collapsed from any BinaryBVH (its own primIDs copy)."""
    w = WideBVH(bvh.mem, width)
    _check(lib().cubql_collapse_wide_f3(C.byref(bvh.c), int(width), C.byref(w.c),
                                        C.c_void_p(_stream_handle(stream)), C.byref(bvh.mem.resource)),
           "collapse_wide")
    return w


def _as_boxes(boxes: torch.Tensor) -> torch.Tensor:
    if boxes.dtype != torch.float32 or not boxes.is_cuda:
        raise CubqlError("boxes must be a float32 CUDA tensor of shape [n,6] (lower xyz, upper xyz)")
    return boxes.contiguous().view(-1, 6)


def _build(fn_name: str, boxes: torch.Tensor, config: Optional[BuildConfig], stream, mem,
           *extra):
    boxes = _as_boxes(boxes)
    config = config or BuildConfig()
    mem = mem or TorchMemory(boxes.device)
    bvh = BinaryBVH(mem)
    fn = getattr(lib(), fn_name)
    _check(fn(C.byref(bvh.c), _ptr(boxes), boxes.shape[0], C.byref(config), *extra,
              C.c_void_p(_stream_handle(stream)), C.byref(mem.resource)), fn_name)
    return bvh


def gpuBuilder(boxes: torch.Tensor, config: Optional[BuildConfig] = None, stream=None,
               mem: Optional[TorchMemory] = None) -> BinaryBVH:
    """cuBQL::gpuBuilder (builder/cuda.h:91-99) over [n,6] float32 boxes."""
    return _build("cubql_gpu_builder_f3", boxes, config, stream, mem)


def refit(bvh: BinaryBVH, boxes: torch.Tensor, stream=None):
    boxes = _as_boxes(boxes)
    _check(lib().cubql_refit_f3(C.byref(bvh.c), _ptr(boxes), C.c_void_p(_stream_handle(stream)),
                                C.byref(bvh.mem.resource)), "refit")


class cuda:  # noqa: N801 -- mirrors the reference namespace cuBQL::cuda
    @staticmethod
    def radixBuilder(boxes, config=None, stream=None, mem=None) -> BinaryBVH:
        """cuBQL::cuda::radixBuilder (builder/cuda.h:134-140), 63-bit keys."""
        return _build("cubql_radix_builder_f3", boxes, config, stream, mem)

    @staticmethod
    def rebinRadixBuilder(boxes, config=None, stream=None, mem=None) -> BinaryBVH:
        """cuBQL::cuda::rebinRadixBuilder (builder/cuda.h:148-154): robust
        30-bit Morton builder that re-keys ranges of equal keys."""
        return _build("cubql_rebin_radix_builder_f3", boxes, config, stream, mem)

    @staticmethod
    def sahBuilder(boxes, config=None, stream=None, mem=None, numBins: int = 16) -> BinaryBVH:
        """cuBQL::cuda::sahBuilder (builder/cuda.h:122-130) with a bin count
        parameter (the reference fixes 16); leaves <= config.maxAllowedLeafSize."""
        return _build("cubql_sah_builder_f3", boxes, config, stream, mem, C.c_int(numBins))

    @staticmethod
    def radix30Builder(boxes, config=None, stream=None, mem=None) -> BinaryBVH:
        """30-bit Morton variant of radixBuilder (radix.h:34-38)."""
        return _build("cubql_radix30_builder_f3", boxes, config, stream, mem)

    @staticmethod
    def free(bvh: BinaryBVH, stream=None):
        bvh.free(stream)


def _as_points(p: torch.Tensor, what: str) -> torch.Tensor:
    if p.dtype != torch.float32 or not p.is_cuda:
        raise CubqlError(f"{what} must be a float32 CUDA tensor of shape [n,3]")
    return p.contiguous().view(-1, 3)


class points:  # noqa: N801 -- mirrors cuBQL::points
    @staticmethod
    def findClosest(bvh: BinaryBVH, data: torch.Tensor, queries: torch.Tensor,
                    sqrMaxSearchDist: float = math.inf, stream=None):
        """Batched points::findClosest: returns (int32 ids, float32 sqrDist)."""
        data = _as_points(data, "data")
        queries = _as_points(queries, "queries")
        nq = queries.shape[0]
        ids = torch.empty(nq, dtype=torch.int32, device=queries.device)
        d2 = torch.empty(nq, dtype=torch.float32, device=queries.device)
        st = C.c_void_p(_stream_handle(stream))
        if isinstance(bvh, WideBVH):
            _check(lib().cubql_fcp_points_wide_f3(C.byref(bvh.c), bvh.width, _ptr(data), _ptr(queries), nq,
                                                  float(sqrMaxSearchDist), _ptr(ids), _ptr(d2), st),
                   "findClosest(WideBVH)")
        else:
            _check(lib().cubql_fcp_points_f3(C.byref(bvh.c), _ptr(data), _ptr(queries), nq,
                                             float(sqrMaxSearchDist), _ptr(ids), _ptr(d2), st), "findClosest")
        return ids, d2

    @staticmethod
    def findKNN(bvh: BinaryBVH, data: torch.Tensor, queries: torch.Tensor, k: int,
                sqrMaxSearchDist: float = math.inf, stream=None):
        """Batched points::findKNN: (ids [nq,k], sqrDist [nq,k], numFound [nq])."""
        data = _as_points(data, "data")
        queries = _as_points(queries, "queries")
        nq = queries.shape[0]
        ids = torch.empty((nq, k), dtype=torch.int32, device=queries.device)
        d2 = torch.empty((nq, k), dtype=torch.float32, device=queries.device)
        nf = torch.empty(nq, dtype=torch.int32, device=queries.device)
        st = C.c_void_p(_stream_handle(stream))
        if isinstance(bvh, WideBVH):
            _check(lib().cubql_knn_points_wide_f3(C.byref(bvh.c), bvh.width, _ptr(data), _ptr(queries), nq,
                                                  int(k), float(sqrMaxSearchDist), _ptr(ids), _ptr(d2),
                                                  _ptr(nf), st), "findKNN(WideBVH)")
        else:
            _check(lib().cubql_knn_points_f3(C.byref(bvh.c), _ptr(data), _ptr(queries), nq, int(k),
                                             float(sqrMaxSearchDist), _ptr(ids), _ptr(d2), _ptr(nf), st),
                   "findKNN")
        return ids, d2, nf


class fixedBoxQuery:  # noqa: N801 -- mirrors cuBQL::fixedBoxQuery
    @staticmethod
    def forEachPrim(bvh, prim_boxes: torch.Tensor, query_boxes: torch.Tensor, max_per_query: int = 0,
                    stream=None):
        """Batched fixed-box query: (ids [nq, max_per_query] int32 in
        traversal order, counts [nq] int32) of the prims whose boxes overlap
        each query box; max_per_query = 0 returns the counts only."""
        pb, qb = _as_boxes(prim_boxes), _as_boxes(query_boxes)
        nq = qb.shape[0]
        counts = torch.empty(nq, dtype=torch.int32, device=qb.device)
        ids = (torch.full((nq, max_per_query), -1, dtype=torch.int32, device=qb.device)
               if max_per_query > 0 else None)
        st = C.c_void_p(_stream_handle(stream))
        if isinstance(bvh, WideBVH):
            _check(lib().cubql_box_query_wide_f3(C.byref(bvh.c), bvh.width, _ptr(pb), _ptr(qb), nq,
                                                 int(max_per_query), _ptr(ids), _ptr(counts), st),
                   "fixedBoxQuery(WideBVH)")
        else:
            _check(lib().cubql_box_query_f3(C.byref(bvh.c), _ptr(pb), _ptr(qb), nq, int(max_per_query),
                                            _ptr(ids), _ptr(counts), st), "fixedBoxQuery")
        return ids, counts


class triangles:  # noqa: N801 -- mirrors cuBQL::triangles
    @staticmethod
    def boxes(tris: Optional[torch.Tensor], verts: torch.Tensor, stream=None) -> torch.Tensor:
        """Triangle::bounds() per triangle: [numTris, 6] float32."""
        verts = _as_points(verts, "verts")
        n = (tris.shape[0] if tris is not None else verts.shape[0] // 3)
        out = torch.empty((n, 6), dtype=torch.float32, device=verts.device)
        _check(lib().cubql_triangle_boxes_f3(_ptr(_as_tris(tris)), _ptr(verts), n, _ptr(out),
                                             C.c_void_p(_stream_handle(stream))), "triangle_boxes")
        return out

    @staticmethod
    def findClosest(bvh: BinaryBVH, tris: Optional[torch.Tensor], verts: torch.Tensor,
                    queries: torch.Tensor, sqrMaxSearchDist: float = math.inf,
                    want_points: bool = True, stream=None):
        """Batched triangles::findClosest: (int32 triIDs, float32 sqrDist, closest points)."""
        verts = _as_points(verts, "verts")
        queries = _as_points(queries, "queries")
        nq = queries.shape[0]
        ids = torch.empty(nq, dtype=torch.int32, device=queries.device)
        d2 = torch.empty(nq, dtype=torch.float32, device=queries.device)
        cp = torch.empty((nq, 3), dtype=torch.float32, device=queries.device) if want_points else None
        _check(lib().cubql_fcp_triangles_f3(C.byref(bvh.c), _ptr(_as_tris(tris)), _ptr(verts),
                                            _ptr(queries), nq, float(sqrMaxSearchDist), _ptr(ids),
                                            _ptr(d2), _ptr(cp), C.c_void_p(_stream_handle(stream))),
               "triangles.findClosest")
        return ids, d2, cp

    @staticmethod
    def backward(tris: Optional[torch.Tensor], verts: torch.Tensor, queries: torch.Tensor,
                 tri_ids: torch.Tensor, grad_dist: torch.Tensor, grad_queries=None,
                 grad_verts=None, stream=None):
        """Adjoint of the unsigned distances (cubql_fcp_triangles_backward_f3)."""
        verts = _as_points(verts, "verts")
        queries = _as_points(queries, "queries")
        _check(lib().cubql_fcp_triangles_backward_f3(
            _ptr(_as_tris(tris)), _ptr(verts), _ptr(queries), queries.shape[0],
            _ptr(tri_ids.contiguous()), _ptr(grad_dist.contiguous()), _ptr(grad_queries),
            _ptr(grad_verts), C.c_void_p(_stream_handle(stream))), "triangles.backward")
        return grad_queries, grad_verts


def points_backward(data: torch.Tensor, queries: torch.Tensor, ids: torch.Tensor,
                    grad_dist: torch.Tensor, grad_queries=None, grad_points=None, stream=None,
                    mem: Optional["TorchMemory"] = None):
    """Adjoint of point distances for fcp (ids [nq]) or kNN (ids [nq,k]);
    grad_points accumulates, grad_queries is overwritten."""
    data = _as_points(data, "data")
    queries = _as_points(queries, "queries")
    nq = queries.shape[0]
    k = 1 if ids.dim() == 1 else ids.shape[1]
    _check(lib().cubql_points_distance_backward_f3(
        _ptr(data), data.shape[0], _ptr(queries), nq, k, _ptr(ids.contiguous()), _ptr(grad_dist.contiguous()),
        _ptr(grad_queries), _ptr(grad_points), C.c_void_p(_stream_handle(stream)),
        C.byref(mem.resource) if mem is not None else None),
        "points_backward")
    return grad_queries, grad_points


def _as_tris(t: Optional[torch.Tensor]):
    if t is None:
        return None
    if t.dtype != torch.int32 or not t.is_cuda:
        raise CubqlError("tris must be an int32 CUDA tensor of shape [n,3]")
    return t.contiguous()


class ChamferBatch:
    """Batched bidirectional closest point + Chamfer adjoint over many cloud
    pairs (cubql_batch_chamfer_f3).  Buffers are allocated once; run() is
    eager, capture() records the launch sequence into a HIP graph and
    replay() launches it."""

    def __init__(self, src: torch.Tensor, src_off: torch.Tensor, tgt: torch.Tensor,
                 tgt_off: torch.Tensor, params: Optional[BatchParams] = None, want_grad=True):
        self.src = _as_points(src, "src")
        self.tgt = _as_points(tgt, "tgt")
        self.src_off = src_off.to(torch.int32).contiguous()
        self.tgt_off = tgt_off.to(torch.int32).contiguous()
        self.num_clouds = self.src_off.numel() - 1
        self.params = params or BatchParams()
        dev = self.src.device
        ns, nt = self.src.shape[0], self.tgt.shape[0]
        for name, off, n in (("src_off", self.src_off, ns), ("tgt_off", self.tgt_off, nt)):
            o = off.cpu()
            if (self.num_clouds < 0 or o.numel() != self.num_clouds + 1 or int(o[0]) != 0
                    or int(o[-1]) != n or bool((o[1:] < o[:-1]).any())):
                raise CubqlError(f"{name} must hold {max(self.num_clouds, 0) + 1} non-decreasing "
                                 f"offsets from 0 to {n}")
        wsb = lib().cubql_batch_chamfer_workspace_f3(self.num_clouds, ns, nt)
        self.ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
        self.src_nn = torch.empty(ns, dtype=torch.int32, device=dev)
        self.tgt_nn = torch.empty(nt, dtype=torch.int32, device=dev)
        self.src_d2 = torch.empty(ns, dtype=torch.float32, device=dev)
        self.tgt_d2 = torch.empty(nt, dtype=torch.float32, device=dev)
        self.loss = torch.empty(self.num_clouds, dtype=torch.float32, device=dev)
        self.grad_src = torch.empty_like(self.src) if want_grad else None
        self.grad_tgt = torch.empty_like(self.tgt) if want_grad else None
        self.graph = None

    def _args(self):
        return (_ptr(self.src), _ptr(self.src_off), self.src.shape[0], _ptr(self.tgt),
                _ptr(self.tgt_off), self.tgt.shape[0], self.num_clouds, C.byref(self.params),
                _ptr(self.ws), self.ws.numel(), _ptr(self.src_nn), _ptr(self.src_d2),
                _ptr(self.tgt_nn), _ptr(self.tgt_d2), _ptr(self.loss), _ptr(self.grad_src),
                _ptr(self.grad_tgt))

    def run(self, stream=None):
        _check(lib().cubql_batch_chamfer_f3(*self._args(), C.c_void_p(_stream_handle(stream))),
               "batch_chamfer")
        return self

    def capture(self):
        g = C.c_void_p()
        _check(lib().cubql_batch_chamfer_graph_f3(*self._args(), C.byref(g)), "batch_chamfer_graph")
        self.graph = g
        return self

    def replay(self, stream=None):
        _check(lib().cubql_graph_launch(self.graph, C.c_void_p(_stream_handle(stream))), "graph_launch")
        return self

    def __del__(self):
        try:
            if self.graph:
                lib().cubql_graph_destroy(self.graph)
        except Exception:  # noqa: BLE001
            pass


class profile:  # noqa: N801
    """Event-based per-kernel timing inside the library (cubql_profile_*)."""

    @staticmethod
    def enable(on: bool = True):
        lib().cubql_profile_enable(1 if on else 0)

    @staticmethod
    def reset():
        lib().cubql_profile_reset()

    @staticmethod
    def query(kernel: str):
        """{'launches', 'ms', 'bytes', 'units'} for a kernel name, or None."""
        n, ms, by, un = C.c_uint64(), C.c_double(), C.c_double(), C.c_uint64()
        if lib().cubql_profile_query(kernel.encode(), C.byref(n), C.byref(ms), C.byref(by),
                                     C.byref(un)) != 0:
            return None
        return {"launches": n.value, "ms": ms.value, "bytes": by.value, "units": un.value}


def query_stats(bvh: BinaryBVH, queries: torch.Tensor, data: Optional[torch.Tensor] = None, k: int = 0,
                tris: Optional[torch.Tensor] = None, verts: Optional[torch.Tensor] = None,
                sqrMaxSearchDist: float = math.inf, stream=None):
    """Traversal statistics of a query batch (cubql_query_stats_*): points
    closest-point (k=0) or kNN (k>0) over `data`, or closest-triangle over
    (tris, verts).  Returns {'node_bytes', 'prims', 'queries'} summed over
    the batch (the reference's buildQual NUM_TRAVERSAL_STEPS counters)."""
    queries = _as_points(queries, "queries")
    nq = queries.shape[0]
    out = torch.zeros(3, dtype=torch.int64, device=queries.device)
    st = C.c_void_p(_stream_handle(stream))
    if verts is not None:
        v = _as_points(verts, "verts")
        _check(lib().cubql_query_stats_triangles_f3(C.byref(bvh.c), _ptr(_as_tris(tris)), _ptr(v), _ptr(queries),
                                                    nq, float(sqrMaxSearchDist), _ptr(out), st),
               "query_stats(triangles)")
    else:
        data = _as_points(data, "data")
        _check(lib().cubql_query_stats_points_f3(C.byref(bvh.c), _ptr(data), _ptr(queries), nq, int(k),
                                                 float(sqrMaxSearchDist), _ptr(out), st), "query_stats(points)")
    o = out.cpu().tolist()
    return {"node_bytes": int(o[0]), "prims": int(o[1]), "queries": int(o[2])}


def radix_sort_pairs(keys: torch.Tensor, vals: torch.Tensor, key_bits: int, stream=None):
    """Stable sort of (keys, vals) in place on key bits [0, key_bits).

    keys: int32 (read as uint32) or int64 (read as uint64); vals: int32."""
    if keys.dtype not in (torch.int32, torch.int64) or vals.dtype != torch.int32:
        raise CubqlError("keys must be int32/int64 and vals int32")
    n = keys.numel()
    kt, vt = torch.empty_like(keys), torch.empty_like(vals)
    fn = lib().cubql_radix_sort_pairs_u32 if keys.dtype == torch.int32 else lib().cubql_radix_sort_pairs_u64
    _check(fn(_ptr(keys), _ptr(vals), _ptr(kt), _ptr(vt), n, int(key_bits),
              C.c_void_p(_stream_handle(stream))), "radix_sort_pairs")
    return keys, vals


def points_to_boxes(p: torch.Tensor) -> torch.Tensor:
    """box3f().including(p) per point (samples/s01_closestPoint_points_gpu/closestPoint.cu:16-22)."""
    p = p.contiguous().view(-1, 3)
    return torch.cat([p, p], dim=1).contiguous()


def wide_nodes_to_numpy(bvh: WideBVH):
    """Host copy of wide children as [numNodes, W] structured records (lo[3],
    hi[3], valid:1 | offset:45 | count:16 word)."""
    import numpy as np
    dt = np.dtype([("lo", "<f4", 3), ("hi", "<f4", 3), ("oc", "<u8")])
    raw = bvh.nodes_tensor().contiguous().cpu().numpy()
    return raw.view(dt).reshape(bvh.numNodes, bvh.width)


def nodes_to_numpy(bvh: BinaryBVH):
    """Host copy of nodes as a structured array (lo[3], hi[3], oc u64)."""
    import numpy as np
    dt = np.dtype([("lo", "<f4", 3), ("hi", "<f4", 3), ("oc", "<u8")])
    raw = bvh.nodes_tensor().contiguous().cpu().numpy()
    return raw.view(dt).reshape(-1)
