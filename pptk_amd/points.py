"""Points: a write-locked point array with a lazily rebuilt device tree.

Host-side mirror of pptk's direct caller of the k-d tree (pptk/points/
points.py:86-166 and the nbhds_op expression, pptk/points/expr.py:397-424):

* a Points object is a numpy array that can only be modified through its own
  methods (item assignment, in-place operators); every modification stamps
  the underlying memory block as changed;
* the first neighbourhood query after a change rebuilds the tree (here a
  device-resident tree through pptk_amd.kdtree._build), so repeated queries
  on unchanged points reuse the resident tree;
* nbhds(queries, k, r) answers at once (kdtree._query); NBHDS(...) returns a
  lazy expression evaluated in chunks when iterated or indexed, or in one
  device call by evaluate().

Differences from pptk, on purpose: modification stamps come from a counter
rather than the wall clock (two changes within one clock tick are still
ordered); ones_like/empty_like return ones/empties (pptk returns zeros);
Points cannot be subclassed (pptk's Final metaclass has no effect in
Python 3).
"""
import itertools
import math
import threading

import numpy as np

from . import kdtree

__all__ = ["Points", "points", "zeros", "ones", "empty", "zeros_like", "ones_like",
           "empty_like", "rand", "load", "nbhds_op"]

_stamp = itertools.count(1)
_stamp_lock = threading.Lock()


def _next_stamp():
    with _stamp_lock:
        return next(_stamp)


def _memaddr(a):
    return a.__array_interface__["data"][0]


def points(obj, dtype=None, copy=True):
    """A Points array from `obj` (a view when copy=False and obj is an
    ndarray of that dtype)."""
    if not copy and isinstance(obj, np.ndarray) and (dtype is None or dtype == obj.dtype):
        # as in pptk, only a Points array can be viewed as Points
        return obj.view(Points)
    tmp = np.asarray(obj, dtype=dtype)
    ret = empty(tmp.shape, dtype=tmp.dtype)
    ret[:] = tmp
    return ret


def empty(shape, dtype=float):
    return Points(shape=shape, dtype=dtype)


def zeros(shape, dtype=float):
    ret = empty(shape, dtype)
    ret[...] = 0
    return ret


def ones(shape, dtype=float):
    ret = empty(shape, dtype)
    ret[...] = 1
    return ret


def zeros_like(a, dtype=None):
    return zeros(a.shape, a.dtype if dtype is None else dtype)


def ones_like(a, dtype=None):
    return ones(a.shape, a.dtype if dtype is None else dtype)


def empty_like(a, dtype=None):
    return empty(a.shape, a.dtype if dtype is None else dtype)


def rand(*dims):
    ret = empty(dims, float)
    ret[...] = np.random.rand(*dims)
    return ret


def load(file, **kwargs):
    return points(np.load(file, **kwargs))


class Points(np.ndarray):
    """Write-locked array of points with a cached device tree."""

    # memory block address -> stamp of its last modification
    _last_modified = {}

    def __init_subclass__(cls, **kwargs):
        raise TypeError("type Points is not an acceptable base type")

    def __new__(cls, *args, **kwargs):
        obj = super().__new__(cls, *args, **kwargs)
        obj.flags.writeable = False
        return obj

    def __array_finalize__(self, obj):
        self._tree = None
        self._built_stamp = None
        if obj is not None and not isinstance(obj, Points):
            raise TypeError("Detected attempt at creating Points-type view on "
                            "non-Points object.")
        if obj is None:
            # fresh allocation (__new__)
            if not self.flags.owndata:
                raise TypeError("Detected attempt at creating Points-type view on "
                                "buffer object via __new__(buffer=...)")
            self._memloc = _memaddr(self)
            self._memsize = self.size * self.dtype.itemsize
        elif not (obj._memloc <= _memaddr(self) < obj._memloc + max(obj._memsize, 1)):
            # a copy: its own memory block
            self._memloc = _memaddr(self)
            self._memsize = self.size * self.dtype.itemsize
        else:
            # a slice or view of a Points object: same memory block
            self._memloc = obj._memloc
            self._memsize = obj._memsize

    def copy(self, order="C"):
        x = super().copy(order)
        x.flags.writeable = False
        return x

    # ---- modification tracking ----
    def _record_modify(self):
        Points._last_modified[self._memloc] = _next_stamp()

    def _guard(self, f):
        def guarded(*args, **kwargs):
            bases = []
            b = self.base
            while isinstance(b, np.ndarray):
                bases.append(b)
                b = b.base
            for x in reversed(bases):  # the owner first
                x.flags.writeable = True
            self.flags.writeable = True
            try:
                return f(*args, **kwargs)
            finally:
                self.flags.writeable = False
                for x in bases:
                    if isinstance(x, Points):
                        x.flags.writeable = False
                self._record_modify()
        return guarded

    def __setitem__(self, key, value):
        return self._guard(super().__setitem__)(key, value)

    def __delitem__(self, key):
        return self._guard(super().__delitem__)(key)

    def __iadd__(self, o):
        return self._guard(super().__iadd__)(o)

    def __isub__(self, o):
        return self._guard(super().__isub__)(o)

    def __imul__(self, o):
        return self._guard(super().__imul__)(o)

    def __itruediv__(self, o):
        return self._guard(super().__itruediv__)(o)

    def __ifloordiv__(self, o):
        return self._guard(super().__ifloordiv__)(o)

    def __imod__(self, o):
        return self._guard(super().__imod__)(o)

    def __ipow__(self, o):
        return self._guard(super().__ipow__)(o)

    def __ilshift__(self, o):
        return self._guard(super().__ilshift__)(o)

    def __irshift__(self, o):
        return self._guard(super().__irshift__)(o)

    def __iand__(self, o):
        return self._guard(super().__iand__)(o)

    def __ixor__(self, o):
        return self._guard(super().__ixor__)(o)

    def __ior__(self, o):
        return self._guard(super().__ior__)(o)

    # ---- tree cache (points.py:140-159) ----
    def _update_kd_tree(self):
        """Rebuilds the device tree if the memory block changed since the
        last build; returns True when it rebuilt."""
        last = Points._last_modified.get(self._memloc)
        if last is None:
            last = Points._last_modified[self._memloc] = _next_stamp()
        if self._built_stamp is None or self._built_stamp < last:
            self._tree = kdtree._build(np.asarray(self))
            self._built_stamp = _next_stamp()
            return True
        return False

    def nbhds(self, queries=None, k=1, r=None, verbose=False):
        """kdtree._query on the cached tree (points.py:161-163)."""
        self._update_kd_tree()
        return kdtree._query(self._tree, queries=queries, k=k, dmax=r)

    def NBHDS(self, queries=None, k=1, r=None, verbose=False):
        """Lazy neighbourhood expression (points.py:165-166)."""
        return nbhds_op(self, queries, k, r)


# chunking of lazy evaluation (expr.py:22-24)
_NUM_CHUNKS = 128
_MIN_CHUNK = 1000
_MAX_CHUNK = 5000


class nbhds_op:
    """Lazy neighbourhoods of `data` (expr.py:397-424): one result per query
    (or per point when queries is None), computed chunk by chunk on access."""

    def __init__(self, data, queries, k, r):
        self.data = data
        self.queries = queries
        self.k = k
        self.r = r

    def __len__(self):
        return self.data.shape[0] if self.queries is None else len(self.queries)

    def _evaluate_chunk(self, index, size, use_cache=False):
        size = max(0, min(size, len(self) - index))
        self.data._update_kd_tree()
        if self.queries is None:
            q = np.arange(index, index + size)
        else:
            q = self.queries[index:index + size]
        return kdtree._query(self.data._tree, q, self.k, self.r)

    def _chunk_size(self, num_chunks):
        n = len(self)
        return max(_MIN_CHUNK, min(_MAX_CHUNK, int(math.ceil(float(n) / num_chunks))))

    def evaluate(self, num_chunks=_NUM_CHUNKS, use_cache=True):
        """All results as a list. The device answers every query in one call
        (chunking only bounds host memory in pptk)."""
        n = len(self)
        if n == 0:
            return []
        return list(self._evaluate_chunk(0, n))

    def __iter__(self):
        n = len(self)
        step = self._chunk_size(_NUM_CHUNKS)
        for i in range(0, n, step):
            for x in self._evaluate_chunk(i, step):
                yield x

    def __getitem__(self, key):
        if isinstance(key, (int, np.integer)):
            if key >= len(self) or key < -len(self):
                raise IndexError("list index out of range")
            key = int(key) % len(self)
            return self._evaluate_chunk(key, 1)[0]
        raise TypeError("unrecognized index key type %s" % type(key))

    def __repr__(self):
        n = len(self)
        if n > 10:
            beg = self._evaluate_chunk(0, 3)
            end = self._evaluate_chunk(n - 3, 3)
            body = ",\n".join(repr(x) for x in beg) + ",\n...\n" + \
                ",\n".join(repr(x) for x in end)
        else:
            body = ",\n".join(repr(x) for x in self._evaluate_chunk(0, n))
        return "Expression %s (%d elements)[\n%s\n]" % (type(self), n, body)
