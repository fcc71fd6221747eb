"""Serving layer: a request loop over device-resident k-d trees.

pptk has no server; its callers build a tree per cloud and query it
(kdtree._build / kdtree._query, pptk/kdtree/kdtree_wrapper.cpp:161-304) or
call estimate_normals (pptk/processing/estimate_normals/estimate_normals.cpp:
259-345), one call at a time. This module runs many such requests at once on
one GPU:

* Requests name a cloud. A cloud's tree is built by the first request that
  brings its points and stays resident in HBM for follow-up requests; idle
  trees are evicted least-recently-used when memory is needed.
* Admission is strict FIFO by HBM footprint: the request at the head of the
  queue is admitted when a slot is free and its estimated device footprint
  (build temporaries, tree, query outputs) fits the budget, after evicting
  idle trees if needed. Later requests never overtake it, so a large cloud is
  not starved by a stream of small ones.
* Each slot is a worker thread with its own HIP stream (default priority);
  admitted requests run concurrently and their kernels overlap on the device
  (continuous batching: a request that arrives while others are on the GPU
  starts as soon as a slot and memory are free, not at a batch boundary).
* Each request computes on its own tree and buffers, so its result does not
  depend on what else is in flight (the serving tests check this bit for
  bit).

The device work goes through a backend object (GpuBackend: the C ABI of
libpptk_amd through pptk_amd.capi); the scheduler itself is plain Python and
is tested on the CPU with a fake backend.
"""
import collections
import ctypes
import threading
import time
from dataclasses import dataclass

import numpy as np

KINDS = ("knn", "rnear", "normals")


@dataclass
class Request:
    """One request.

    kind    : "knn" (self query of every point, k neighbours), "rnear" (self
              query of `queries`, all neighbours with d < dmax) or "normals"
              (estimate_normals(k, r=dmax) of every point).
    cloud   : cloud id; a resident tree for it is reused.
    points  : (n, 3) float32 points; needed when the cloud is not resident.
    k, dmax : query parameters.
    queries : rnear only: int32 point indices of the cloud (self queries);
              None = every point.
    """
    kind: str
    cloud: object
    points: np.ndarray = None
    k: int = 16
    dmax: float = np.inf
    queries: np.ndarray = None
    tag: object = None
    size: int = 0  # points the request queries (throughput accounting)


class Cancelled(Exception):
    pass


class Ticket:
    """Handle of a submitted request: result(), cancel(), timings."""

    def __init__(self, req, now):
        self.req = req
        self.t_submit = now
        self.t_start = None
        self.t_done = None
        self.built = False      # this request built the cloud's tree
        self._done = threading.Event()
        self._result = None
        self._error = None
        self._state = "queued"  # queued | running | done | cancelled

    def _finish(self, result=None, error=None, now=None):
        self._result, self._error = result, error
        self.t_done = now if now is not None else time.perf_counter()
        self._done.set()

    def done(self):
        return self._done.is_set()

    def result(self, timeout=None):
        if not self._done.wait(timeout):
            raise TimeoutError("request not finished")
        if self._error is not None:
            raise self._error
        return self._result

    @property
    def error(self):
        return self._error

    @property
    def latency(self):
        return None if self.t_done is None else self.t_done - self.t_submit


# ---------------------------------------------------------------- footprint ----

def tree_bytes_estimate(n):
    """Resident tree: point records (16 B), reverse/forward/leaf maps (12 B),
    nodes + boxes (48 B per node, about n/32 nodes at 64 points per leaf)."""
    return 28 * n + 48 * (n // 16 + 1) + 4096


def build_temp_bytes(n):
    """Build temporaries on top of the tree: staged input (12 B), Morton keys
    and values (12 B), sort buffers (12 B), node array at capacity 2n (32 B)."""
    return 68 * n + (1 << 20)


def query_bytes(req, n):
    """Device bytes of the query outputs and temporaries."""
    if req.kind == "knn":
        return n * (4 * req.k + 4) + (1 << 20)
    if req.kind == "normals":
        return n * (12 + 4 + 4) + (1 << 20)
    m = n if req.queries is None else len(req.queries)
    # offsets, counts, order keys; the neighbour lists themselves are bounded
    # per request by the rnear_cap the server was given
    return m * 24 + (1 << 20)


# ------------------------------------------------------------------- budget ----

class _Entry:
    __slots__ = ("tree", "bytes", "users", "last_use", "ready", "failed", "n")

    def __init__(self):
        self.tree = None
        self.bytes = 0
        self.users = 0
        self.last_use = 0.0
        self.ready = threading.Event()
        self.failed = None
        self.n = 0


class Server:
    """Request loop over `slots` worker threads sharing one HBM budget."""

    def __init__(self, backend, budget_bytes=200 * 10 ** 9, slots=8, rnear_cap=1 << 30,
                 clock=time.perf_counter):
        """rnear_cap: device bytes reserved for one r-near request's neighbour
        lists; a request whose lists need more fails alone (MemoryError)."""
        if slots < 1:
            raise ValueError("slots must be >= 1")
        self.backend = backend
        self.budget = int(budget_bytes)
        self.slots = slots
        self.rnear_cap = int(rnear_cap)
        self.clock = clock
        self._cv = threading.Condition()
        self._queue = collections.deque()
        self._trees = {}          # cloud id -> _Entry
        self._reserved = 0        # bytes held by running requests
        self._resident = 0        # bytes held by resident trees
        self._peak = 0
        self._admitting = False
        self._closed = False
        self._busy = 0
        self.completed = []
        self.evictions = 0
        self._threads = [threading.Thread(target=self._worker, args=(s,), daemon=True)
                         for s in range(slots)]
        for t in self._threads:
            t.start()

    # ---- client side ----
    def submit(self, req):
        """Queues `req`; returns its Ticket. Invalid requests complete at once
        with their error and never reach the device."""
        t = Ticket(req, self.clock())
        err = self._validate(req)
        if err is not None:
            t._state = "done"
            t._finish(error=err, now=self.clock())
            return t
        with self._cv:
            if self._closed:
                raise RuntimeError("server is closed")
            self._queue.append(t)
            self._cv.notify_all()
        return t

    def cancel(self, ticket):
        """Cancels a queued request; a running one completes normally."""
        with self._cv:
            if ticket._state != "queued":
                return False
            try:
                self._queue.remove(ticket)
            except ValueError:
                return False
            ticket._state = "cancelled"
        ticket._finish(error=Cancelled("request cancelled"), now=self.clock())
        return True

    def close(self, wait=True):
        with self._cv:
            self._closed = True
            self._cv.notify_all()
        if wait:
            for t in self._threads:
                t.join()
        with self._cv:
            for e in self._trees.values():
                if e.tree is not None:
                    self.backend.free(e.tree)
            self._trees.clear()
            self._resident = 0

    def drain(self):
        """Blocks until the queue is empty and no request is running."""
        with self._cv:
            while self._queue or self._busy:
                self._cv.wait()

    def checkpoint(self, directory):
        """Writes every resident tree (persistent tree format) and an index
        to `directory`; returns the number of trees written. Cloud ids must
        be strings or integers."""
        import json
        import os
        os.makedirs(directory, exist_ok=True)
        with self._cv:
            items = [(cid, e) for cid, e in self._trees.items() if e.tree is not None]
            for _, e in items:
                e.users += 1  # pinned while written
        index = []
        try:
            for i, (cid, e) in enumerate(items):
                if not isinstance(cid, (str, int)):
                    raise TypeError("cloud id %r cannot be checkpointed" % (cid,))
                name = "tree%06d.pptktree" % i
                with open(os.path.join(directory, name), "wb") as f:
                    f.write(self.backend.serialize(e.tree))
                index.append({"cloud": cid, "file": name, "points": e.n})
        finally:
            with self._cv:
                for _, e in items:
                    e.users -= 1
                self._cv.notify_all()
        with open(os.path.join(directory, "index.json"), "w") as f:
            json.dump(index, f)
        return len(index)

    def restore(self, directory):
        """Loads the trees of checkpoint(directory) as resident trees (those
        that fit the budget, evicting idle ones); returns how many."""
        import json
        import os
        with open(os.path.join(directory, "index.json")) as f:
            index = json.load(f)
        loaded = 0
        for item in index:
            with open(os.path.join(directory, item["file"]), "rb") as f:
                blob = f.read()
            tree = self.backend.deserialize(blob, 0)
            nbytes = int(self.backend.tree_bytes(tree))
            with self._cv:
                if item["cloud"] in self._trees or not self._evict_for(nbytes):
                    self.backend.free(tree)
                    continue
                e = _Entry()
                e.tree, e.bytes, e.n = tree, nbytes, int(item["points"])
                e.last_use = self.clock()
                e.ready.set()
                self._trees[item["cloud"]] = e
                self._resident += nbytes
                loaded += 1
        return loaded

    def memory(self):
        with self._cv:
            return {"reserved": self._reserved, "resident": self._resident,
                    "peak": self._peak, "budget": self.budget,
                    "resident_trees": sum(1 for e in self._trees.values() if e.tree)}

    # ---- validation and footprint ----
    @staticmethod
    def _validate(req):
        if req.kind not in KINDS:
            return ValueError("unknown request kind %r" % (req.kind,))
        if req.kind in ("knn", "normals") and not (1 <= int(req.k) <= 1024):
            return ValueError("k must be in [1, 1024]")
        if not (req.dmax >= 0):
            return ValueError("dmax must be non-negative")
        if req.points is not None:
            p = req.points
            if not isinstance(p, np.ndarray) or p.ndim != 2 or p.shape[1] != 3:
                return ValueError("points must be an (n, 3) array")
            if p.dtype != np.float32:
                return TypeError("points must be float32")
        if req.kind == "rnear" and req.queries is not None:
            q = np.asarray(req.queries)
            if q.ndim != 1 or not np.issubdtype(q.dtype, np.integer):
                return ValueError("rnear queries must be a 1-D integer index array")
        return None

    def _need(self, req, entry, n):
        need = query_bytes(req, n)
        if entry is None:
            need += tree_bytes_estimate(n) + build_temp_bytes(n)
        if req.kind == "rnear":
            need += self.rnear_cap
        return need

    def _evict_for(self, need):
        """Evicts idle resident trees (LRU) until `need` fits; under the lock."""
        while self._reserved + self._resident + need > self.budget:
            idle = [(e.last_use, cid) for cid, e in self._trees.items()
                    if e.users == 0 and e.tree is not None]
            if not idle:
                return False
            _, cid = min(idle)
            e = self._trees.pop(cid)
            self._resident -= e.bytes
            self.backend.free(e.tree)
            self.evictions += 1
        return True

    # ---- worker ----
    def _admit(self):
        """Takes the head request once it fits (strict FIFO). Returns
        (ticket, entry, need, must_build) or None when closed."""
        with self._cv:
            while True:
                while not self._closed and (self._admitting or not self._queue):
                    self._cv.wait()
                if self._closed and not self._queue:
                    return None
                if self._admitting:
                    continue
                self._admitting = True
                break
            try:
                t = self._queue[0]
                req = t.req
                # an entry that is not ready yet is being built by a running
                # request; this one waits for it on the worker side
                entry = self._trees.get(req.cloud)
                n = (entry.n if entry is not None and entry.n else
                     (len(req.points) if req.points is not None else 0))
                if entry is None and req.points is None:
                    self._queue.popleft()
                    t._state = "done"
                    t._finish(error=KeyError("cloud %r is not resident and the request "
                                             "carries no points" % (req.cloud,)),
                              now=self.clock())
                    return "skip"
                need = self._need(req, entry, n)
                if need > self.budget:
                    self._queue.popleft()
                    t._state = "done"
                    t._finish(error=MemoryError("request needs %d bytes, budget is %d"
                                                % (need, self.budget)), now=self.clock())
                    return "skip"
                # need <= budget: once the running requests finish and the
                # idle trees are evicted it fits, so this wait ends
                while not self._evict_for(need):
                    self._cv.wait()
                    if not self._queue or self._queue[0] is not t:
                        return "skip"  # cancelled meanwhile: admit the new head
                    if self._trees.get(req.cloud) is not entry:
                        return "skip"  # built or evicted meanwhile: recompute
                self._queue.popleft()
                must_build = entry is None
                if must_build:
                    entry = _Entry()
                    entry.n = n
                    self._trees[req.cloud] = entry
                entry.users += 1
                self._reserved += need
                self._peak = max(self._peak, self._reserved + self._resident)
                self._busy += 1
                t._state = "running"
                t.t_start = self.clock()
                return t, entry, need, must_build
            finally:
                self._admitting = False
                self._cv.notify_all()

    def _worker(self, slot):
        while True:
            got = self._admit()
            if got is None:
                return
            if got == "skip":
                continue
            t, entry, need, must_build = got
            req = t.req
            result, error = None, None
            try:
                if must_build:
                    # Requests for the same cloud wait on entry.ready; it is
                    # set only once entry.tree is assigned (or the build has
                    # failed), never in between.
                    try:
                        tree = self.backend.build(req.points, slot)
                        nbytes = int(self.backend.tree_bytes(tree))
                        with self._cv:
                            entry.tree = tree
                            entry.bytes = nbytes
                            self._resident += nbytes
                            # the tree's estimate and the build temporaries
                            # are now the resident tree
                            est = tree_bytes_estimate(entry.n) + build_temp_bytes(entry.n)
                            self._reserved -= est
                            need -= est
                            self._cv.notify_all()
                    except BaseException as e:
                        entry.failed = e
                        with self._cv:
                            self._trees.pop(req.cloud, None)
                            self._cv.notify_all()
                        raise
                    finally:
                        entry.ready.set()
                    t.built = True
                else:
                    entry.ready.wait()
                    if entry.failed is not None:
                        raise RuntimeError("building cloud %r failed: %s"
                                           % (req.cloud, entry.failed))
                result = self.backend.run(req, entry.tree, slot, self.rnear_cap)
            except Exception as e:  # the request's own error; others go on
                error = e
            finally:
                with self._cv:
                    self._reserved -= need
                    entry.users -= 1
                    entry.last_use = self.clock()
                    self._busy -= 1
                    self._cv.notify_all()
            t._state = "done"
            t._finish(result=result, error=error, now=self.clock())
            with self._cv:
                self.completed.append(t)


# ------------------------------------------------------------------ backend ----

class GpuBackend:
    """Device work through the C ABI: one HIP stream per slot.

    outputs="host": results are numpy arrays (copied over PCIe, as pptk's
    Python calls return them); outputs="device": results stay in HBM as torch
    tensors (idx/cnt, offsets/lists, normals) for a consumer on the GPU."""

    def __init__(self, slots, outputs="host"):
        import torch  # torch's HIP runtime must start before the library's
        if not torch.cuda.is_available():
            raise RuntimeError("GpuBackend needs a GPU")
        if outputs not in ("host", "device"):
            raise ValueError("outputs must be 'host' or 'device'")
        from . import capi
        self.torch = torch
        self.capi = capi
        self.lib = capi.lib()
        self.outputs = outputs
        self.dev = torch.device("cuda:0")
        self.streams = [torch.cuda.Stream() for _ in range(slots)]

    def _s(self, slot):
        return ctypes.c_void_p(self.streams[slot].cuda_stream)

    def build(self, points, slot):
        return self.capi.Tree.build(np.ascontiguousarray(points), stream=self._s(slot))

    def tree_bytes(self, tree):
        return tree.info.device_bytes

    def free(self, tree):
        tree.free()

    def serialize(self, tree):
        return tree.serialize()

    def deserialize(self, blob, slot):
        return self.capi.Tree.deserialize(blob, stream=self._s(slot))

    def run(self, req, tree, slot, rnear_cap):
        c = self.capi
        s = self._s(slot)
        if self.outputs == "device":
            # Outputs are allocated on the slot stream, the stream the library
            # writes them on, so torch's caching allocator orders their reuse
            # against that work. A consumer on another stream calls
            # tensor.record_stream() before freeing them.
            with self.torch.cuda.stream(self.streams[slot]):
                return self._run_device(req, tree, s, rnear_cap)
        if req.kind == "knn":
            return c.knn(tree, None, int(req.k), float(req.dmax), stream=s)
        if req.kind == "normals":
            ev, _, _, _ = c.normals(tree, int(req.k), float(req.dmax), want_evals=False,
                                    want_sizes=False, stream=s)
            return ev
        q = None if req.queries is None else np.asarray(req.queries, np.int32)
        return c.rnear(tree, q, float(req.dmax), stream=s, max_total=rnear_cap // 4)

    def _run_device(self, req, tree, s, rnear_cap):
        torch, c, L = self.torch, self.capi, self.lib
        h = ctypes.c_void_p(tree.handle)
        n = tree.info.num_valid
        if req.kind == "knn":
            k = int(req.k)
            idx = torch.empty((n, k), dtype=torch.int32, device=self.dev)
            cnt = torch.empty(n, dtype=torch.int32, device=self.dev)
            q = c.Queries(c.Q_ALL, 0, None, None, 0, c.MEM_DEVICE)
            c.check(L.pptk_knn(h, ctypes.byref(q), k, float(req.dmax),
                               ctypes.c_void_p(idx.data_ptr()), ctypes.c_void_p(cnt.data_ptr()),
                               None, c.MEM_DEVICE, s))
            return idx, cnt
        if req.kind == "normals":
            ni = tree.info.num_input
            ev = torch.empty((ni, 3), dtype=torch.float32, device=self.dev)
            c.check(L.pptk_normals(h, int(req.k), float(req.dmax), None, 0, 1, c.PREC_AUTO,
                                   ctypes.c_void_p(ev.data_ptr()), None, None, None,
                                   c.MEM_DEVICE, s))
            return ev
        if req.queries is None:
            q = c.Queries(c.Q_ALL, 0, None, None, 0, c.MEM_DEVICE)
            m = n
            keep = None
        else:
            keep = torch.from_numpy(np.ascontiguousarray(req.queries, np.int32)).to(self.dev)
            m = keep.numel()
            q = c.Queries(c.Q_INDICES, m, keep.data_ptr(), None, 0, c.MEM_DEVICE)
        off = torch.empty(m + 1, dtype=torch.int64, device=self.dev)
        total = ctypes.c_int64()
        c.check(L.pptk_rnear_count(h, ctypes.byref(q), float(req.dmax),
                                   ctypes.c_void_p(off.data_ptr()), ctypes.byref(total),
                                   c.MEM_DEVICE, s))
        if total.value > rnear_cap // 4:
            raise MemoryError("r-near lists hold %d entries, more than the %d allowed"
                              % (total.value, rnear_cap // 4))
        lst = torch.empty(max(int(total.value), 1), dtype=torch.int32, device=self.dev)
        c.check(L.pptk_rnear_fill(h, ctypes.byref(q), float(req.dmax),
                                  ctypes.c_void_p(off.data_ptr()), ctypes.c_void_p(lst.data_ptr()),
                                  c.MEM_DEVICE, s))
        return off, lst[:int(total.value)]


# ---------------------------------------------------------------- workload ----

def request_mix(rng, clouds, n_requests, follow_up=0.5):
    """C4-style request sequence over a pool of clouds: kind uniform over
    kNN k in {8, 16, 32}, r-near dmax in {0.05, 0.2} m (20,000 sampled points
    of the cloud as self queries) and normals k=16; with probability
    `follow_up` a request reuses one of the 16 most recent clouds (its tree is
    resident), else it brings the next cloud of the pool."""
    kinds = [("knn", 8, np.inf), ("knn", 16, np.inf), ("knn", 32, np.inf),
             ("rnear", 0, 0.05), ("rnear", 0, 0.2), ("normals", 16, np.inf)]
    reqs = []
    seen = []
    nxt = 0
    for i in range(n_requests):
        if seen and rng.random() < follow_up:
            cid = seen[-1 - int(rng.integers(0, min(16, len(seen))))]
            pts = None
        else:
            cid = "cloud%d" % nxt
            pts = clouds[nxt % len(clouds)]
            nxt += 1
            seen.append(cid)
        kind, k, dmax = kinds[int(rng.integers(0, len(kinds)))]
        n = len(clouds[int(cid[5:]) % len(clouds)])
        q = None
        if kind == "rnear":
            q = np.sort(rng.choice(n, size=min(n, 20_000), replace=False)).astype(np.int32)
        reqs.append(Request(kind, cid, pts, k=k if k else 16,
                            dmax=float(np.float32(dmax)) if np.isfinite(dmax) else dmax,
                            queries=q, tag=i, size=n if q is None else len(q)))
    return reqs


def percentile_stats(tickets, t0, t1):
    """Requests/s, queried points/s (each request counts the points it
    queries: its cloud for kNN and normals, its query set for r-near) and
    latency percentiles (submit to completion, queueing included)."""
    lat = np.array([t.latency for t in tickets if t.error is None and t.latency is not None])
    pts = sum(t.req.size or (len(t.req.points) if t.req.points is not None else 0)
              for t in tickets if t.error is None)
    span = max(t1 - t0, 1e-9)
    return {
        "requests": len(tickets),
        "errors": sum(1 for t in tickets if t.error is not None),
        "requests_per_s": len(lat) / span,
        "points_per_s": pts / span,
        "p50_ms": float(np.percentile(lat, 50) * 1e3) if len(lat) else None,
        "p99_ms": float(np.percentile(lat, 99) * 1e3) if len(lat) else None,
        "mean_ms": float(lat.mean() * 1e3) if len(lat) else None,
    }


def run_open_loop(server, reqs, rate, rng, clock=time.perf_counter):
    """Submits `reqs` with Poisson arrivals at `rate` requests/s (open loop:
    arrivals do not wait for completions); returns (tickets, t0, t_end)."""
    gaps = rng.exponential(1.0 / rate, len(reqs))
    t0 = clock()
    due = t0
    tickets = []
    for r, g in zip(reqs, gaps):
        due += g
        d = due - clock()
        if d > 0:
            time.sleep(d)
        tickets.append(server.submit(r))
    for t in tickets:
        t._done.wait()
    t_end = max(t.t_done for t in tickets)
    return tickets, t0, t_end
