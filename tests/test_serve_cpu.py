"""CPU: the serving scheduler (pptk_amd.serve.Server) with a fake backend.

The fake backend computes brute-force results with numpy, sleeps to make
requests overlap, checks the server's memory accounting on every call, and
injects failures on demand (a failing build, an allocation failure inside a
run), so admission, residency, eviction, cancellation and error isolation
are tested without a GPU.
"""
import threading
import time

import numpy as np
import pytest

from pptk_amd import serve


class FakeTree:
    def __init__(self, pts):
        self.pts = pts
        self.n = len(pts)


class FakeBackend:
    def __init__(self, delay=0.002, fail_build=(), fail_run=(), tree_bytes=None):
        self.delay = delay
        self.fail_build = list(fail_build)
        self.fail_run = set(fail_run)
        self.tree_bytes_fn = tree_bytes
        self.server = None
        self.builds = 0
        self.runs = []
        self.active = 0
        self.max_active = 0
        self.over_budget = []
        self.lock = threading.Lock()

    def _enter(self):
        with self.lock:
            self.active += 1
            self.max_active = max(self.max_active, self.active)
        m = self.server.memory()
        if m["reserved"] + m["resident"] > m["budget"]:
            self.over_budget.append(m)

    def _leave(self):
        with self.lock:
            self.active -= 1

    def build(self, points, slot):
        self._enter()
        try:
            time.sleep(self.delay)
            if any(np.array_equal(points, p) for p in self.fail_build):
                raise MemoryError("injected allocation failure in build")
            with self.lock:
                self.builds += 1
            return FakeTree(np.array(points))
        finally:
            self._leave()

    def tree_bytes(self, tree):
        return (self.tree_bytes_fn(tree.n) if self.tree_bytes_fn
                else serve.tree_bytes_estimate(tree.n))

    def free(self, tree):
        tree.pts = None

    def serialize(self, tree):
        return np.ascontiguousarray(tree.pts, np.float32).tobytes()

    def deserialize(self, blob, slot):
        return FakeTree(np.frombuffer(blob, np.float32).reshape(-1, 3).copy())

    def run(self, req, tree, slot, rnear_cap):
        self._enter()
        try:
            time.sleep(self.delay)
            with self.lock:
                self.runs.append(req.tag)
            if req.tag in self.fail_run:
                raise MemoryError("injected device allocation failure")
            return brute(req, tree.pts)
        finally:
            self._leave()


def brute(req, pts):
    d = ((pts[:, None, :].astype(np.float64) - pts[None, :, :]) ** 2).sum(-1)
    if req.kind == "knn":
        np.fill_diagonal(d, np.inf)
        idx = np.argsort(d, axis=1, kind="stable")[:, :req.k].astype(np.int32)
        return idx, np.full(len(pts), min(req.k, len(pts) - 1), np.int32)
    if req.kind == "normals":
        out = np.empty((len(pts), 3))
        for i in range(len(pts)):
            nb = np.argsort(d[i], kind="stable")[:req.k]
            x = pts[nb] - pts[nb].mean(0)
            w, v = np.linalg.eigh(x.T @ x)
            out[i] = v[:, 0] * np.sign(v[np.argmax(np.abs(v[:, 0])), 0])
        return out
    q = np.arange(len(pts)) if req.queries is None else req.queries
    lists = [np.flatnonzero((d[i] < req.dmax ** 2) & (np.arange(len(pts)) != i)) for i in q]
    off = np.concatenate([[0], np.cumsum([len(x) for x in lists])])
    return off, (np.concatenate(lists) if lists else np.zeros(0, np.int64)).astype(np.int32)


def make_server(backend, **kw):
    s = serve.Server(backend, **kw)
    backend.server = s
    return s


def clouds(rng, count, lo=20, hi=120):
    return [rng.uniform(0, 1, (int(rng.integers(lo, hi)), 3)).astype(np.float32)
            for _ in range(count)]


def mixed_requests(rng, cl, n):
    reqs = []
    for i in range(n):
        c = int(rng.integers(0, len(cl)))
        kind = ["knn", "rnear", "normals"][i % 3]
        q = np.arange(0, len(cl[c]), 3, dtype=np.int32) if kind == "rnear" else None
        reqs.append(serve.Request(kind, "c%d" % c, cl[c], k=4 + (i % 3), dmax=0.3, queries=q,
                                  tag=i))
    return reqs


def same(a, b):
    if isinstance(a, tuple):
        return all(np.array_equal(x, y) for x, y in zip(a, b))
    return np.array_equal(a, b)


def test_results_alone_and_mixed_are_identical():
    rng = np.random.default_rng(1)
    cl = clouds(rng, 6)
    reqs = mixed_requests(rng, cl, 30)
    be1 = FakeBackend(delay=0.0)
    s1 = make_server(be1, slots=1)
    alone = [s1.submit(r).result() for r in reqs]
    s1.close()
    be = FakeBackend(delay=0.003)
    s = make_server(be, slots=6)
    tickets = [s.submit(r) for r in reqs]
    mixed = [t.result(timeout=30) for t in tickets]
    s.close()
    assert be.max_active > 1
    for a, b in zip(alone, mixed):
        assert same(a, b)
    for r, a in zip(reqs, alone):
        assert same(a, brute(r, r.points))


def test_budget_is_never_exceeded_and_trees_are_evicted():
    rng = np.random.default_rng(2)
    cl = clouds(rng, 10, 200, 400)
    big = max(serve.tree_bytes_estimate(len(c)) + serve.build_temp_bytes(len(c)) +
              serve.query_bytes(serve.Request("knn", 0, None, k=8), len(c)) for c in cl)
    budget = int(2.5 * big)
    be = FakeBackend(delay=0.002)
    s = make_server(be, slots=4, budget_bytes=budget)
    ts = [s.submit(serve.Request("knn", "c%d" % i, c, k=8, tag=i)) for i, c in enumerate(cl)]
    for t in ts:
        t.result(timeout=30)
    s.drain()
    m = s.memory()
    assert not be.over_budget
    assert m["peak"] <= budget
    assert s.evictions >= 1
    assert m["reserved"] == 0
    s.close()


def test_strict_fifo_admission():
    """A request that must wait for memory is admitted before any request
    submitted after it, even small ones that would fit."""
    rng = np.random.default_rng(3)
    small = [rng.uniform(0, 1, (20, 3)).astype(np.float32) for _ in range(6)]
    large = rng.uniform(0, 1, (300, 3)).astype(np.float32)
    need_large = (serve.tree_bytes_estimate(300) + serve.build_temp_bytes(300) +
                  serve.query_bytes(serve.Request("knn", 0, None, k=4), 300))
    need_small = (serve.tree_bytes_estimate(20) + serve.build_temp_bytes(20) +
                  serve.query_bytes(serve.Request("knn", 0, None, k=4), 20))
    budget = need_large + need_small  # the large one fits only beside one small
    be = FakeBackend(delay=0.01)
    s = make_server(be, slots=4, budget_bytes=budget)
    first = [s.submit(serve.Request("knn", "s%d" % i, small[i], k=4, tag=i)) for i in range(2)]
    tl = s.submit(serve.Request("knn", "L", large, k=4, tag="L"))
    later = [s.submit(serve.Request("knn", "s%d" % i, small[i], k=4, tag=i))
             for i in range(2, 6)]
    for t in first + [tl] + later:
        t.result(timeout=30)
    assert all(t.t_start >= tl.t_start for t in later)
    s.close()


def test_validation_failure_does_not_poison_others():
    rng = np.random.default_rng(4)
    cl = clouds(rng, 3)
    be = FakeBackend()
    s = make_server(be, slots=3)
    good = [s.submit(serve.Request("knn", "c%d" % i, c, k=3, tag=i)) for i, c in enumerate(cl)]
    bad = [s.submit(serve.Request("nope", "c0", cl[0])),
           s.submit(serve.Request("knn", "c1", cl[1], k=0)),
           s.submit(serve.Request("knn", "c2", cl[2].astype(np.float64), k=3)),
           s.submit(serve.Request("rnear", "c0", cl[0], dmax=-1.0))]
    for t in good:
        t.result(timeout=30)
    for t in bad:
        with pytest.raises((ValueError, TypeError)):
            t.result(timeout=5)
    s.close()


def test_injected_failures_are_isolated_and_release_memory():
    rng = np.random.default_rng(5)
    cl = clouds(rng, 4)
    be = FakeBackend(fail_build=[cl[1]], fail_run=[2])
    s = make_server(be, slots=4)
    ts = [s.submit(serve.Request("knn", "c%d" % i, c, k=3, tag=i)) for i, c in enumerate(cl)]
    with pytest.raises(MemoryError):
        ts[1].result(timeout=30)
    with pytest.raises(MemoryError):
        ts[2].result(timeout=30)
    ts[0].result(timeout=30)
    ts[3].result(timeout=30)
    s.drain()
    assert s.memory()["reserved"] == 0
    # the failed cloud is not resident: a follow-up without points fails alone
    t = s.submit(serve.Request("knn", "c1", None, k=3))
    with pytest.raises(KeyError):
        t.result(timeout=30)
    # cloud 2 built fine (its run failed): follow-ups reuse its tree
    builds = be.builds
    t = s.submit(serve.Request("normals", "c2", None, k=5, tag=99))
    assert same(t.result(timeout=30), brute(t.req, cl[2]))
    assert be.builds == builds
    s.close()


def test_cancel_queued_request():
    rng = np.random.default_rng(6)
    cl = clouds(rng, 8)
    be = FakeBackend(delay=0.05)
    s = make_server(be, slots=1)
    ts = [s.submit(serve.Request("knn", "c%d" % i, c, k=3, tag=i)) for i, c in enumerate(cl)]
    assert s.cancel(ts[-1])
    with pytest.raises(serve.Cancelled):
        ts[-1].result(timeout=5)
    for t in ts[:-1]:
        t.result(timeout=30)
    assert len(cl) - 1 not in be.runs
    s.close()


def test_request_larger_than_budget_fails_at_once():
    rng = np.random.default_rng(7)
    c = rng.uniform(0, 1, (1000, 3)).astype(np.float32)
    be = FakeBackend()
    s = make_server(be, slots=2, budget_bytes=10_000)
    t = s.submit(serve.Request("knn", "c", c, k=3))
    with pytest.raises(MemoryError):
        t.result(timeout=5)
    s.close()


def test_follow_ups_reuse_resident_tree_and_concurrent_first_use_builds_once():
    rng = np.random.default_rng(8)
    c = rng.uniform(0, 1, (150, 3)).astype(np.float32)
    be = FakeBackend(delay=0.02)
    s = make_server(be, slots=4)
    ts = [s.submit(serve.Request("knn", "c", c, k=4, tag=0))]
    ts += [s.submit(serve.Request(k_, "c", None, k=4, dmax=0.2, tag=i + 1))
           for i, k_ in enumerate(["knn", "normals", "rnear", "knn"])]
    for t in ts:
        t.result(timeout=30)
    assert be.builds == 1
    assert ts[0].built and not any(t.built for t in ts[1:])
    s.close()


def test_open_loop_stats():
    rng = np.random.default_rng(9)
    cl = clouds(rng, 5)
    be = FakeBackend(delay=0.001)
    s = make_server(be, slots=4)
    reqs = serve.request_mix(rng, cl, 40)
    for r in reqs:  # the mix samples 20,000 r-near queries; keep them in range
        if r.queries is not None:
            n = len(cl[int(r.cloud[5:]) % len(cl)])
            r.queries = r.queries[r.queries < n]
    tickets, t0, t1 = serve.run_open_loop(s, reqs, rate=2000.0, rng=rng)
    st = serve.percentile_stats(tickets, t0, t1)
    assert st["errors"] == 0 and st["requests"] == 40
    assert st["p50_ms"] <= st["p99_ms"]
    s.close()


def test_checkpoint_and_restore_resident_trees(tmp_path):
    rng = np.random.default_rng(10)
    cl = clouds(rng, 4)
    be = FakeBackend(delay=0.0)
    s = make_server(be, slots=2)
    for i, c in enumerate(cl):
        s.submit(serve.Request("knn", "c%d" % i, c, k=3)).result(timeout=30)
    assert s.checkpoint(str(tmp_path)) == 4
    s.close()
    be2 = FakeBackend(delay=0.0)
    s2 = make_server(be2, slots=2)
    assert s2.restore(str(tmp_path)) == 4
    # follow-ups without points run on the restored trees, nothing rebuilt
    for i, c in enumerate(cl):
        r = s2.submit(serve.Request("normals", "c%d" % i, None, k=4))
        assert same(r.result(timeout=30), brute(r.req, c))
    assert be2.builds == 0
    s2.close()


class LockHoldingBackend(FakeBackend):
    """Holds the server's lock from another thread for `hold` seconds, taken
    just before a build returns. That widens the window between a finished
    build and the assignment of its tree to the cloud entry, where requests
    waiting for the same cloud must not yet run."""

    def __init__(self, hold=0.1, **kw):
        super().__init__(**kw)
        self.hold = hold
        self.none_trees = 0

    def build(self, points, slot):
        tree = super().build(points, slot)
        got = threading.Event()

        def hold():
            with self.server._cv:
                got.set()
                time.sleep(self.hold)

        threading.Thread(target=hold, daemon=True).start()
        got.wait()
        return tree

    def run(self, req, tree, slot, rnear_cap):
        if tree is None:
            with self.lock:
                self.none_trees += 1
        return super().run(req, tree, slot, rnear_cap)


def test_waiters_never_see_a_built_tree_before_it_is_published():
    rng = np.random.default_rng(11)
    c = rng.uniform(0, 1, (80, 3)).astype(np.float32)
    be = LockHoldingBackend(hold=0.1, delay=0.05)
    s = make_server(be, slots=4)
    ts = [s.submit(serve.Request("knn", "c", c, k=4, tag=0))]
    ts += [s.submit(serve.Request("knn", "c", None, k=4, tag=i)) for i in (1, 2, 3)]
    for t in ts:
        assert same(t.result(timeout=30), brute(t.req, c))
    assert be.none_trees == 0 and be.builds == 1
    s.close()
