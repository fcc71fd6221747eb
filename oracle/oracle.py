"""TEST INFRASTRUCTURE ONLY -- the oracle.

Python access to the CPU checkers. Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg import this module; the product path never does.

* ``OracleTree``: C++ restatement of pptk's k-d tree (oracle/pptk_oracle.cpp,
  following pptk/kdtree/src/kdtree-impl.h), kNN with the (distance, index)
  order, r-near, FP64 PCA normals.
* ``RefTree``: pptk's own k-d tree code compiled from the reference sources
  (oracle/ref_driver.cpp -> oracle/_ref/libpptk_ref.so), when built.
* ``brute_knn`` / ``brute_rnear``: exhaustive NumPy search with the
  reference's distance formula (kdtree-impl.h:331-339) evaluated in the
  distance dtype, left to right, without FMA.
* ``normals_fp64``: FP64 covariance of given neighbourhoods + numpy eigh.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

c_int_p = ctypes.POINTER(ctypes.c_int)


def _load():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise ImportError("oracle not built: run `make -C oracle`")
        lib = ctypes.CDLL(path)
        lib.oracle_build.restype = ctypes.c_void_p
        lib.oracle_build.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_double]
        lib.oracle_free.argtypes = [ctypes.c_void_p]
        lib.oracle_num_valid.argtypes = [ctypes.c_void_p]
        lib.oracle_num_valid.restype = ctypes.c_int
        lib.oracle_knn.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                   ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_void_p,
                                   ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
        lib.oracle_rnear.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                     ctypes.c_void_p, ctypes.c_int, ctypes.c_double,
                                     ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
        lib.oracle_normals.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double,
                                       ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                       ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
        lib.oracle_csr_first_mismatch.restype = ctypes.c_longlong
        lib.oracle_csr_first_mismatch.argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_longlong,
                                                                         ctypes.c_int]
        _LIB = lib
    return _LIB


def csr_first_mismatch(off_a, idx_a, off_b, idx_b, nthreads=0):
    """-1 if the two CSR lists hold the same set in every row, else the
    first differing row (each row compared after sorting copies)."""
    off_a = np.ascontiguousarray(off_a, np.int64)
    off_b = np.ascontiguousarray(off_b, np.int64)
    idx_a = np.ascontiguousarray(idx_a, np.int32)
    idx_b = np.ascontiguousarray(idx_b, np.int32)
    m = len(off_a) - 1
    if len(off_b) - 1 != m:
        raise ValueError("row counts differ")
    if off_a[-1] > len(idx_a) or off_b[-1] > len(idx_b):
        raise ValueError("offsets beyond the lists")
    return int(_load().oracle_csr_first_mismatch(_ptr(off_a), _ptr(idx_a), _ptr(off_b),
                                                 _ptr(idx_b), m, nthreads))


def dist_dtype(dtype):
    """Accumulator<T>::Type (pptk/kdtree/src/accumulator.h:12-27)."""
    return np.float64 if np.dtype(dtype) == np.float64 else np.float32


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class OracleTree:
    """CPU restatement of pointkd::KdTree<T, dim> (kdtree.h:44-249)."""

    def __init__(self, points, leaf=10, empty_split=0.2):
        pts = np.asarray(points)
        if pts.ndim != 2 or pts.shape[1] not in (2, 3, 4):
            raise ValueError("points must be n x {2,3,4}")
        self.dim = pts.shape[1]
        self.dtype = dist_dtype(pts.dtype)
        self.points = np.ascontiguousarray(pts.astype(self.dtype))
        self.n_input = pts.shape[0]
        lib = _load()
        self._h = lib.oracle_build(_ptr(self.points), self.n_input, self.dim,
                                   int(self.dtype == np.float64), int(leaf), float(empty_split))
        self.num_valid = lib.oracle_num_valid(self._h)

    def __del__(self):
        if getattr(self, "_h", None):
            _load().oracle_free(self._h)
            self._h = None

    def _queries(self, queries):
        if queries is None:
            return 0, None, None, self.num_valid
        q = np.asarray(queries)
        if q.ndim == 2:
            qp = np.ascontiguousarray(q.astype(self.dtype))
            return 2, None, qp, q.shape[0]
        idx = np.ascontiguousarray(q.astype(np.int32).ravel())
        return 1, idx, None, idx.shape[0]

    def knn(self, queries, k, dmax=np.inf, nthreads=0, with_dist=False):
        """(idx (m,k) int32 -1 padded, counts (m,), [dist (m,k)])."""
        mode, lst, qp, m = self._queries(queries)
        idx = np.empty((m, k), np.int32)
        cnt = np.empty(m, np.int32)
        dist = np.empty((m, k), self.dtype) if with_dist else None
        _load().oracle_knn(self._h, mode, _ptr(lst), _ptr(qp), m, k, float(dmax), _ptr(idx),
                           _ptr(cnt), _ptr(dist), nthreads)
        return (idx, cnt, dist) if with_dist else (idx, cnt)

    def rnear(self, queries, dmax, nthreads=0):
        """(offsets (m+1,) int64, idx int32) with each list sorted."""
        mode, lst, qp, m = self._queries(queries)
        off = np.zeros(m + 1, np.int64)
        lib = _load()
        lib.oracle_rnear(self._h, mode, _ptr(lst), _ptr(qp), m, float(dmax), _ptr(off), None,
                         nthreads)
        idx = np.empty(int(off[-1]), np.int32)
        lib.oracle_rnear(self._h, mode, _ptr(lst), _ptr(qp), m, float(dmax), _ptr(off), _ptr(idx),
                         nthreads)
        return off, idx

    def normals(self, k, r, rows=None, nthreads=0):
        """FP64 PCA: (evecs (m,3,3) [i,j,:] = j-th eigvec, evals (m,3), sizes (m,))."""
        if self.dim != 3:
            raise ValueError("normals need 3-D points")
        rr = None if rows is None else np.ascontiguousarray(np.asarray(rows, np.int32))
        m = self.n_input if rr is None else rr.shape[0]
        ev = np.empty((m, 3, 3), np.float64)
        ew = np.empty((m, 3), np.float64)
        sz = np.empty(m, np.int32)
        _load().oracle_normals(self._h, int(k), float(r), _ptr(rr), m, _ptr(ev), _ptr(ew),
                               _ptr(sz), nthreads)
        return ev, ew, sz


def ref_available():
    return os.path.exists(os.path.join(HERE, "_ref", "libpptk_ref.so"))


def _load_ref():
    global _REF
    if _REF is None:
        path = os.path.join(HERE, "_ref", "libpptk_ref.so")
        if not os.path.exists(path):
            raise ImportError("reference driver not built (oracle/_ref)")
        lib = ctypes.CDLL(path)
        for t in ("f32", "f64"):
            b = getattr(lib, "ref_build_" + t)
            b.restype = ctypes.c_void_p
            b.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
            getattr(lib, "ref_free_" + t).argtypes = [ctypes.c_void_p]
            np_ = getattr(lib, "ref_num_points_" + t)
            np_.argtypes = [ctypes.c_void_p]
            np_.restype = ctypes.c_int
            getattr(lib, "ref_knn_" + t).argtypes = [
                ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
            getattr(lib, "ref_rnear_" + t).argtypes = [
                ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_double,
                ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
        _REF = lib
    return _REF


class RefTree:
    """pptk's own KdTree<T, 3> (float32 or float64), built serially."""

    def __init__(self, points, leaf=10):
        pts = np.asarray(points)
        if pts.dtype not in (np.float32, np.float64) or pts.ndim != 2 or pts.shape[1] != 3:
            raise ValueError("RefTree takes n x 3 float32/float64")
        self.t = "f64" if pts.dtype == np.float64 else "f32"
        self.points = np.ascontiguousarray(pts)
        lib = _load_ref()
        self._h = getattr(lib, "ref_build_" + self.t)(_ptr(self.points), pts.shape[0], int(leaf))
        self.num_valid = getattr(lib, "ref_num_points_" + self.t)(self._h)

    def __del__(self):
        if getattr(self, "_h", None):
            getattr(_load_ref(), "ref_free_" + self.t)(self._h)
            self._h = None

    def knn(self, queries, k, dmax=np.inf, nthreads=0, m_self=None):
        """queries: None (self queries over original indices 0..m_self-1,
        default all) or an m x 3 array of query points."""
        lib = _load_ref()
        if queries is None:
            mode, qp = 0, None
            m = self.num_valid if m_self is None else int(m_self)
        else:
            qp = np.ascontiguousarray(np.asarray(queries, self.points.dtype))
            mode, m = 1, qp.shape[0]
        idx = np.empty((m, k), np.int32)
        cnt = np.empty(m, np.int32)
        getattr(lib, "ref_knn_" + self.t)(self._h, mode, _ptr(qp), m, k, float(dmax), _ptr(idx),
                                          _ptr(cnt), nthreads)
        return idx, cnt

    def rnear(self, queries, dmax, nthreads=0):
        lib = _load_ref()
        if queries is None:
            mode, qp, m = 0, None, self.num_valid
        else:
            qp = np.ascontiguousarray(np.asarray(queries, self.points.dtype))
            mode, m = 1, qp.shape[0]
        off = np.zeros(m + 1, np.int64)
        fn = getattr(lib, "ref_rnear_" + self.t)
        fn(self._h, mode, _ptr(qp), m, float(dmax), _ptr(off), None, nthreads)
        idx = np.empty(int(off[-1]), np.int32)
        fn(self._h, mode, _ptr(qp), m, float(dmax), _ptr(off), _ptr(idx), nthreads)
        return off, idx


# ------------------------------------------------------------ brute force ----

def pair_dist2(q, p, dtype):
    """((0 + t0*t0) + t1*t1) + ... in `dtype`, t = p - q (kdtree-impl.h:331-339).

    NumPy evaluates each elementwise operation in IEEE arithmetic of the array
    dtype without contraction, so this is the reference formula exactly."""
    q = np.asarray(q, dtype)
    p = np.asarray(p, dtype)
    s = np.zeros(np.broadcast_shapes(q.shape[:-1], p.shape[:-1]), dtype)
    # non-finite rows are inputs here: their inf/NaN distances are expected
    with np.errstate(over="ignore", invalid="ignore"):
        for i in range(q.shape[-1]):
            t = p[..., i] - q[..., i]
            s = s + t * t
    return s


def brute_knn(points, queries, k, dmax=np.inf, self_rows=None, chunk=256):
    """Exhaustive kNN ordered by (d, index); d < dmax^2 in the distance dtype.

    queries: m x dim points; self_rows: optional original index per query to
    exclude (self queries). Returns (idx (m,k) -1 padded, counts)."""
    pts = np.asarray(points)
    dt = dist_dtype(pts.dtype)
    P = pts.astype(dt)
    valid = np.all(np.isfinite(P), axis=1)
    Q = np.asarray(queries).astype(dt)
    r = dt(dmax)
    r2 = r * r
    m = Q.shape[0]
    out = np.full((m, k), -1, np.int32)
    cnt = np.zeros(m, np.int32)
    ids = np.arange(P.shape[0])
    for s in range(0, m, chunk):
        qs = Q[s:s + chunk]
        d = pair_dist2(qs[:, None, :], P[None, :, :], dt)
        ok = (d < r2) & valid[None, :]
        ok &= np.all(np.isfinite(qs), axis=1)[:, None]
        if self_rows is not None:
            sr = np.asarray(self_rows)[s:s + chunk]
            ok &= ids[None, :] != sr[:, None]
        for j in range(qs.shape[0]):
            cand = np.nonzero(ok[j])[0]
            order = np.lexsort((cand, d[j, cand]))[:k]
            sel = cand[order]
            out[s + j, :len(sel)] = sel
            cnt[s + j] = len(sel)
    return out, cnt


def brute_rnear(points, queries, dmax, self_rows=None):
    """Exhaustive r-near; lists sorted by index. Returns list of arrays."""
    pts = np.asarray(points)
    dt = dist_dtype(pts.dtype)
    P = pts.astype(dt)
    valid = np.all(np.isfinite(P), axis=1)
    Q = np.asarray(queries).astype(dt)
    r = dt(dmax)
    r2 = r * r
    res = []
    for j in range(Q.shape[0]):
        d = pair_dist2(Q[j][None, :], P, dt)
        ok = (d < r2) & valid & np.all(np.isfinite(Q[j]))
        if self_rows is not None:
            ok[self_rows[j]] = False
        res.append(np.nonzero(ok)[0].astype(np.int32))
    return res


def normals_fp64(points, nbr_idx, nbr_cnt):
    """FP64 PCA of given neighbourhoods: (evecs (m,3,3) rows = eigvecs, evals (m,3)).

    C = X^T X / n of the centred neighbourhood
    (estimate_normals.cpp:84-89), eigen-decomposition by LAPACK (numpy.eigh)."""
    P = np.asarray(points, np.float64)
    m = nbr_idx.shape[0]
    ev = np.full((m, 3, 3), np.nan)
    ew = np.full((m, 3), np.nan)
    for c in np.unique(nbr_cnt):
        if c == 0:
            continue
        rows = np.nonzero(nbr_cnt == c)[0]
        X = P[nbr_idx[rows, :c]]                      # (r, c, 3)
        X = X - X.mean(axis=1, keepdims=True)
        C = np.einsum("rci,rcj->rij", X, X) / c
        w, v = np.linalg.eigh(C)
        ew[rows] = w
        ev[rows] = np.transpose(v, (0, 2, 1))
    return ev, ew


def angle_deg(a, b):
    """Sign-invariant angle between unit vectors, degrees."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    c = np.abs(np.sum(a * b, axis=-1))
    c /= np.linalg.norm(a, axis=-1) * np.linalg.norm(b, axis=-1)
    return np.degrees(np.arccos(np.clip(c, 0.0, 1.0)))
