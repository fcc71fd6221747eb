"""ctypes binding of the C ABI in ``include/pptk_amd.h`` (``lib/libpptk_amd.so``).

Used for device-resident work (device pointers from torch tensors, explicit
streams) by the benchmark, the batch/serving layers and the GPU tests; the
drop-in Python surface is the ``kdtree`` / ``estimate_normals`` extension
modules. This is the same binding a maintainer of pptk would add to call the
library from Python without the CPython modules.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libpptk_amd.so")

# pptk_status
OK, E_VALUE, E_TYPE, E_RUNTIME, E_MEMORY, E_DEVICE, E_NOTIMPL = range(7)
# pptk_dtype
F32, F64, I8, I16, I32, I64, U8, U16, U32, U64 = range(10)
# pptk_mem
MEM_HOST, MEM_DEVICE = 0, 1
# pptk_query_kind
Q_ALL, Q_INDICES, Q_POINTS = 0, 1, 2
# pptk_precision
PREC_AUTO, PREC_MIXED, PREC_FP64 = 0, 1, 2

_NP2DT = {
    np.dtype(np.float32): F32, np.dtype(np.float64): F64,
    np.dtype(np.int8): I8, np.dtype(np.int16): I16, np.dtype(np.int32): I32,
    np.dtype(np.int64): I64, np.dtype(np.uint8): U8, np.dtype(np.uint16): U16,
    np.dtype(np.uint32): U32, np.dtype(np.uint64): U64,
}

EXPORTED = [
    "pptk_last_error", "pptk_device_count", "pptk_trim", "pptk_tree_build", "pptk_tree_free",
    "pptk_tree_check", "pptk_tree_get_info", "pptk_tree_export", "pptk_knn", "pptk_knn_noalloc",
    "pptk_rnear_count",
    "pptk_rnear_fill", "pptk_normals", "pptk_knn_normals", "pptk_sort_pairs_u64",
    "pptk_exclusive_scan_i32",
    "pptk_batch_create", "pptk_batch_run", "pptk_batch_get_info", "pptk_batch_free",
    "pptk_tree_serialize", "pptk_tree_deserialize",
]


class BuildParams(ctypes.Structure):
    _fields_ = [("max_leaf_size", ctypes.c_int), ("num_proc", ctypes.c_int),
                ("empty_split", ctypes.c_double)]


class Queries(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("m", ctypes.c_int64), ("indices", ctypes.c_void_p),
                ("points", ctypes.c_void_p), ("dtype", ctypes.c_int), ("mem", ctypes.c_int)]


class TreeInfo(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int), ("dtype", ctypes.c_int), ("dist_dtype", ctypes.c_int),
                ("num_input", ctypes.c_int64), ("num_valid", ctypes.c_int64),
                ("leaf_size", ctypes.c_int), ("num_leaves", ctypes.c_int64),
                ("num_nodes", ctypes.c_int64), ("depth", ctypes.c_int),
                ("bbox_lo", ctypes.c_double * 4), ("bbox_hi", ctypes.c_double * 4),
                ("device_bytes", ctypes.c_int64)]


class BatchParams(ctypes.Structure):
    _fields_ = [("k", ctypes.c_int), ("normals_k", ctypes.c_int), ("dmax", ctypes.c_double),
                ("max_leaf_size", ctypes.c_int), ("precision", ctypes.c_int),
                ("use_graph", ctypes.c_int), ("query_block", ctypes.c_int),
                ("timing", ctypes.c_int)]


class BatchInfo(ctypes.Structure):
    _fields_ = [("num_clouds", ctypes.c_int64), ("num_points", ctypes.c_int64),
                ("num_valid", ctypes.c_int64), ("num_nodes", ctypes.c_int64),
                ("depth", ctypes.c_int), ("leaf_size", ctypes.c_int),
                ("morton_bits", ctypes.c_int), ("graph_nodes", ctypes.c_int),
                ("device_bytes", ctypes.c_int64), ("num_fp64", ctypes.c_int64),
                ("build_ms", ctypes.c_double), ("query_ms", ctypes.c_double)]


_lib = None


def lib():
    """Loads libpptk_amd.so (raises OSError if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    from ._runtime import preload_hip_runtime
    preload_hip_runtime()
    L = ctypes.CDLL(LIB_PATH)
    vp, i64, ci, cd = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_double
    L.pptk_last_error.restype = ctypes.c_char_p
    L.pptk_device_count.restype = ci
    L.pptk_trim.restype = ci
    L.pptk_tree_build.restype = ci
    L.pptk_tree_build.argtypes = [vp, i64, ci, ci, ctypes.POINTER(BuildParams), ci, vp,
                                  ctypes.POINTER(vp)]
    L.pptk_tree_free.argtypes = [vp]
    L.pptk_tree_check.argtypes = [vp]
    L.pptk_tree_check.restype = ci
    L.pptk_tree_get_info.restype = ci
    L.pptk_tree_get_info.argtypes = [vp, ctypes.POINTER(TreeInfo)]
    L.pptk_tree_export.restype = ci
    L.pptk_tree_export.argtypes = [vp, vp, vp, vp, vp, vp, ci]
    L.pptk_knn.restype = ci
    L.pptk_knn.argtypes = [vp, ctypes.POINTER(Queries), ci, cd, vp, vp, vp, ci, vp]
    L.pptk_knn_noalloc.restype = ci
    L.pptk_knn_noalloc.argtypes = [vp, ctypes.POINTER(Queries), ci, cd, vp, vp, vp, vp]
    L.pptk_rnear_count.restype = ci
    L.pptk_rnear_count.argtypes = [vp, ctypes.POINTER(Queries), cd, vp, ctypes.POINTER(i64), ci,
                                   vp]
    L.pptk_rnear_fill.restype = ci
    L.pptk_rnear_fill.argtypes = [vp, ctypes.POINTER(Queries), cd, vp, vp, ci, vp]
    L.pptk_normals.restype = ci
    L.pptk_normals.argtypes = [vp, ci, cd, vp, i64, ci, ci, vp, vp, vp, ctypes.POINTER(i64), ci,
                               vp]
    L.pptk_knn_normals.restype = ci
    L.pptk_knn_normals.argtypes = [vp, ci, cd, ci, cd, ci, ci, vp, vp, vp, vp, vp, vp,
                                   ctypes.POINTER(i64), ci, vp]
    L.pptk_sort_pairs_u64.restype = ci
    L.pptk_sort_pairs_u64.argtypes = [vp, vp, i64, ci, ci]
    L.pptk_exclusive_scan_i32.restype = ci
    L.pptk_exclusive_scan_i32.argtypes = [vp, vp, i64, ci]
    L.pptk_batch_create.restype = ci
    L.pptk_batch_create.argtypes = [vp, i64, ci, ci, ctypes.POINTER(BatchParams),
                                    ctypes.POINTER(vp)]
    L.pptk_batch_run.restype = ci
    L.pptk_batch_run.argtypes = [vp, vp, vp, vp, vp, vp, vp]
    L.pptk_batch_get_info.restype = ci
    L.pptk_batch_get_info.argtypes = [vp, ctypes.POINTER(BatchInfo)]
    L.pptk_batch_free.argtypes = [vp]
    L.pptk_tree_serialize.restype = ci
    L.pptk_tree_serialize.argtypes = [vp, vp, i64, ctypes.POINTER(i64)]
    L.pptk_tree_deserialize.restype = ci
    L.pptk_tree_deserialize.argtypes = [vp, i64, vp, ctypes.POINTER(vp)]
    _lib = L
    return L


class PptkError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("pptk status %d: %s" % (status, msg))
        self.status = status


_EXC = {E_VALUE: ValueError, E_TYPE: TypeError, E_MEMORY: MemoryError,
        E_NOTIMPL: NotImplementedError}


def check(status):
    if status != OK:
        msg = lib().pptk_last_error().decode("utf-8", "replace")
        exc = _EXC.get(status)
        if exc is not None:
            raise exc(msg)
        raise PptkError(status, msg)


def dtype_code(dt):
    try:
        return _NP2DT[np.dtype(dt)]
    except KeyError:
        raise TypeError("unsupported dtype %s" % (dt,))


def _p(a):
    """Pointer of a numpy array or an int device address (None -> NULL)."""
    if a is None:
        return None
    if isinstance(a, int):
        return ctypes.c_void_p(a)
    return a.ctypes.data_as(ctypes.c_void_p)


class Tree:
    """Owning handle of a device-resident tree."""

    def __init__(self, handle):
        self.handle = handle
        self._info = None

    @classmethod
    def build(cls, points=None, *, dev_ptr=None, n=None, dim=3, dtype=np.float32, leaf=0,
              stream=None):
        """Build from a host (n, dim) array, or from a device pointer."""
        h = ctypes.c_void_p()
        bp = BuildParams(int(leaf), -1, 0.2)
        if dev_ptr is not None:
            st = lib().pptk_tree_build(ctypes.c_void_p(dev_ptr), int(n), int(dim),
                                       dtype_code(dtype), ctypes.byref(bp), MEM_DEVICE, stream,
                                       ctypes.byref(h))
        else:
            pts = np.ascontiguousarray(points)
            st = lib().pptk_tree_build(_p(pts), pts.shape[0], pts.shape[1],
                                       dtype_code(pts.dtype), ctypes.byref(bp), MEM_HOST, stream,
                                       ctypes.byref(h))
        check(st)
        return cls(h.value)

    def free(self):
        if self.handle:
            lib().pptk_tree_free(ctypes.c_void_p(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    @property
    def info(self):
        if self._info is None:
            inf = TreeInfo()
            check(lib().pptk_tree_get_info(ctypes.c_void_p(self.handle), ctypes.byref(inf)))
            self._info = inf
        return self._info

    @property
    def dist_dtype(self):
        return np.float64 if self.info.dist_dtype == F64 else np.float32

    def serialize(self):
        """The tree in the persistent flat format (bytes)."""
        n = ctypes.c_int64()
        check(lib().pptk_tree_serialize(ctypes.c_void_p(self.handle), None, 0, ctypes.byref(n)))
        buf = np.empty(n.value, np.uint8)
        check(lib().pptk_tree_serialize(ctypes.c_void_p(self.handle), _p(buf), n.value,
                                        ctypes.byref(n)))
        return buf.tobytes()

    @classmethod
    def deserialize(cls, data, stream=None):
        """A device tree from serialize()'s bytes (validated before upload)."""
        buf = np.frombuffer(data, np.uint8)
        h = ctypes.c_void_p()
        check(lib().pptk_tree_deserialize(_p(buf), buf.size, stream, ctypes.byref(h)))
        return cls(h.value)

    def save(self, path):
        with open(path, "wb") as f:
            f.write(self.serialize())

    @classmethod
    def load(cls, path, stream=None):
        with open(path, "rb") as f:
            return cls.deserialize(f.read(), stream=stream)

    def export(self):
        """Host copies: points (tree order), reverse, forward, boxes (nn, 2, dim),
        nodes (nn, 4) = [a, b, child, parent]."""
        inf = self.info
        dt = self.dist_dtype
        pts = np.empty((inf.num_valid, inf.dim), dt)
        rev = np.empty(inf.num_valid, np.int32)
        fwd = np.empty(inf.num_input, np.int32)
        box = np.empty((inf.num_nodes, 2, inf.dim), dt)
        nodes = np.empty((inf.num_nodes, 4), np.int32)
        check(lib().pptk_tree_export(ctypes.c_void_p(self.handle), _p(pts), _p(rev), _p(fwd),
                                     _p(box), _p(nodes), MEM_HOST))
        return pts, rev, fwd, box, nodes


def make_queries(queries=None, *, dev_ptr=None, m=None, kind=None, dtype=None, mem=MEM_HOST):
    """Query descriptor. queries: None (all), 1-D int array (indices), 2-D array (points).

    The returned tuple keeps referenced arrays alive."""
    q = Queries()
    keep = None
    if dev_ptr is not None:
        q.kind = kind
        q.m = int(m)
        q.mem = MEM_DEVICE
        if kind == Q_INDICES:
            q.indices = dev_ptr
        else:
            q.points = dev_ptr
            q.dtype = dtype_code(dtype)
        return q, keep
    if queries is None:
        q.kind = Q_ALL
    else:
        a = np.asarray(queries)
        if a.ndim == 2:
            keep = np.ascontiguousarray(a)
            q.kind = Q_POINTS
            q.m = keep.shape[0]
            q.points = keep.ctypes.data
            q.dtype = dtype_code(keep.dtype)
        else:
            keep = np.ascontiguousarray(a.astype(np.int32).ravel())
            q.kind = Q_INDICES
            q.m = keep.shape[0]
            q.indices = keep.ctypes.data
    q.mem = mem
    return q, keep


def knn(tree, queries, k, dmax=np.inf, with_dist=False, stream=None):
    """Host-memory kNN: (idx (m,k) int32 -1 padded, counts (m,) [, dist])."""
    q, keep = make_queries(queries)
    m = tree.info.num_valid if q.kind == Q_ALL else q.m
    idx = np.empty((m, k), np.int32)
    cnt = np.empty(m, np.int32)
    dist = np.empty((m, k), tree.dist_dtype) if with_dist else None
    check(lib().pptk_knn(ctypes.c_void_p(tree.handle), ctypes.byref(q), int(k), float(dmax),
                         _p(idx), _p(cnt), _p(dist), MEM_HOST, stream))
    return (idx, cnt, dist) if with_dist else (idx, cnt)


def rnear(tree, queries, dmax, stream=None, max_total=None):
    """Host-memory r-near: (offsets (m+1,) int64, idx int32). With max_total,
    raises MemoryError before the fill when the lists hold more entries."""
    q, keep = make_queries(queries)
    m = tree.info.num_valid if q.kind == Q_ALL else q.m
    off = np.empty(m + 1, np.int64)
    total = ctypes.c_int64()
    check(lib().pptk_rnear_count(ctypes.c_void_p(tree.handle), ctypes.byref(q), float(dmax),
                                 _p(off), ctypes.byref(total), MEM_HOST, stream))
    if max_total is not None and total.value > max_total:
        raise MemoryError("r-near lists hold %d entries, more than the %d allowed"
                          % (total.value, max_total))
    idx = np.empty(int(total.value), np.int32)
    check(lib().pptk_rnear_fill(ctypes.c_void_p(tree.handle), ctypes.byref(q), float(dmax),
                                _p(off), _p(idx), MEM_HOST, stream))
    return off, idx


def normals(tree, k, r, rows=None, num_eigen=1, precision=PREC_AUTO, want_evals=True,
            want_sizes=True, stream=None):
    """Host-memory normals: (evecs, evals, sizes, num_fp64)."""
    inf = tree.info
    dt = np.float64 if inf.dtype == F64 else np.float32
    rr = None if rows is None else np.ascontiguousarray(np.asarray(rows, np.int32))
    m = inf.num_input if rr is None else rr.shape[0]
    ev = np.empty((m, 3, 3) if num_eigen == 3 else (m, 3), dt)
    ew = (np.empty((m, 3) if num_eigen == 3 else (m,), dt)) if want_evals else None
    sz = np.empty(m, np.int32) if want_sizes else None
    nfb = ctypes.c_int64()
    check(lib().pptk_normals(ctypes.c_void_p(tree.handle), int(k), float(r), _p(rr),
                             0 if rr is None else rr.shape[0], int(num_eigen), int(precision),
                             _p(ev), _p(ew), _p(sz), ctypes.byref(nfb), MEM_HOST, stream))
    return ev, ew, sz, int(nfb.value)


def knn_normals(tree, k, nk, dmax=np.inf, r=np.inf, num_eigen=1, precision=PREC_AUTO,
                with_dist=False, want_evals=True, want_sizes=True, stream=None):
    """Host-memory self-query kNN + normals from one search (pptk_knn_normals):
    (idx, cnt, dist or None, evecs, evals, sizes, num_fp64), each as knn()
    with queries=None and normals() with rows=None return it."""
    inf = tree.info
    dt = np.float64 if inf.dtype == F64 else np.float32
    m = inf.num_valid
    idx = np.empty((m, k), np.int32)
    cnt = np.empty(m, np.int32)
    dist = np.empty((m, k), tree.dist_dtype) if with_dist else None
    n = inf.num_input
    ev = np.empty((n, 3, 3) if num_eigen == 3 else (n, 3), dt)
    ew = (np.empty((n, 3) if num_eigen == 3 else (n,), dt)) if want_evals else None
    sz = np.empty(n, np.int32) if want_sizes else None
    nfb = ctypes.c_int64()
    check(lib().pptk_knn_normals(ctypes.c_void_p(tree.handle), int(k), float(dmax), int(nk),
                                 float(r), int(num_eigen), int(precision), _p(idx), _p(cnt),
                                 _p(dist), _p(ev), _p(ew), _p(sz), ctypes.byref(nfb), MEM_HOST,
                                 stream))
    return idx, cnt, dist, ev, ew, sz, int(nfb.value)


class Batch:
    """Owning handle of a batch plan (pptk_batch_*): fixed cloud sizes, every
    device buffer allocated up front, runs on device pointers."""

    def __init__(self, sizes, *, dim=3, dtype=np.float32, k=16, normals_k=16, dmax=np.inf,
                 leaf=0, precision=PREC_AUTO, use_graph=True, query_block=0, timing=False):
        sz = np.ascontiguousarray(np.asarray(sizes, np.int64))
        self.sizes = sz
        self.offsets = np.concatenate([[0], np.cumsum(sz)]).astype(np.int64)
        self.dim = int(dim)
        self.dtype = np.dtype(dtype)
        self.k = int(k)
        self.normals_k = int(normals_k)
        prm = BatchParams(self.k, self.normals_k, float(dmax), int(leaf), int(precision),
                          1 if use_graph else 0, int(query_block), 1 if timing else 0)
        h = ctypes.c_void_p()
        self.handle = None
        check(lib().pptk_batch_create(_p(sz), sz.shape[0], self.dim, dtype_code(self.dtype),
                                      ctypes.byref(prm), ctypes.byref(h)))
        self.handle = h.value

    @property
    def num_points(self):
        return int(self.offsets[-1])

    def run(self, pts_ptr, idx_ptr=None, cnt_ptr=None, nrm_ptr=None, eval_ptr=None,
            stream=None):
        """One run on device pointers (ints); see pptk_batch_run."""
        check(lib().pptk_batch_run(ctypes.c_void_p(self.handle), _p(pts_ptr), _p(idx_ptr),
                                   _p(cnt_ptr), _p(nrm_ptr), _p(eval_ptr), stream))

    def info(self):
        inf = BatchInfo()
        check(lib().pptk_batch_get_info(ctypes.c_void_p(self.handle), ctypes.byref(inf)))
        return inf

    def free(self):
        if self.handle:
            lib().pptk_batch_free(ctypes.c_void_p(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
