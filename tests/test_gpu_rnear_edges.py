"""GPU: r-near count and fill passes around their edges, against the oracle
(RNearNeighbors, pptk/kdtree/src/kdtree-impl.h:978-1069; lists compared as
sets):
* the query array rewritten in place between count and fill (same pointer,
  same count): the fill reuses the count's processing order (a permutation
  kept for locality) and must still answer the rewritten rows;
* balls that straddle every leaf of a thin shell (more than 10M neighbours,
  most from partly covered leaves) next to balls that cover a whole blob;
* a count on other queries, or at another dmax, between a count and its
  fill; self queries of every point through the drop-in.
"""
import ctypes

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _count(tree, q, dmax):
    from pptk_amd import capi
    m = tree.info.num_valid if q.kind == capi.Q_ALL else q.m
    off = np.empty(m + 1, np.int64)
    tot = ctypes.c_int64()
    capi.check(capi.lib().pptk_rnear_count(ctypes.c_void_p(tree.handle), ctypes.byref(q),
                                           float(dmax), capi._p(off), ctypes.byref(tot),
                                           capi.MEM_HOST, None))
    return off


def _fill(tree, q, dmax, off):
    from pptk_amd import capi
    idx = np.empty(max(int(off[-1]), 1), np.int32)
    capi.check(capi.lib().pptk_rnear_fill(ctypes.c_void_p(tree.handle), ctypes.byref(q),
                                          float(dmax), capi._p(off), capi._p(idx),
                                          capi.MEM_HOST, None))
    return idx[: int(off[-1])]


def _same_sets(off, idx, woff, widx):
    assert np.array_equal(off, woff)
    assert O.csr_first_mismatch(off, idx, woff, widx) == -1


def test_fill_after_rewriting_the_queries_in_place(gpu):
    """Rows i and j of equal count swapped in place after the count: the old
    offsets still fit, and the fill must answer the rewritten rows."""
    from pptk_amd import capi
    rng = np.random.default_rng(41)
    pts = rng.random((200_000, 3)).astype(np.float32)
    tree = capi.Tree.build(pts)
    qp = np.ascontiguousarray(pts[rng.choice(len(pts), 20_000, replace=False)])
    q, keep = capi.make_queries(qp)
    dmax = float(np.float32(0.02))
    off = _count(tree, q, dmax)
    cnt = np.diff(off)
    # swap pairs of rows with equal counts but different neighbours
    swapped = 0
    for c in np.unique(cnt[cnt > 3])[:50]:
        rows = np.flatnonzero(cnt == c)
        if len(rows) >= 2:
            i, j = rows[0], rows[1]
            keep[[i, j]] = keep[[j, i]]
            swapped += 1
    assert swapped >= 10
    idx = _fill(tree, q, dmax, off)
    ot = O.OracleTree(pts)
    woff, widx = ot.rnear(keep, dmax)
    _same_sets(off, idx, woff, widx)
    tree.free()


@pytest.mark.parametrize("m_centre", [16, 48])
def test_straddling_shell_and_covered_blob(gpu, m_centre):
    """Queries at the centre of a thin spherical shell with the shell's
    radius: every leaf of the shell straddles every such ball. Queries inside
    a separate blob take it whole."""
    from pptk_amd import capi
    rng = np.random.default_rng(42)
    d = rng.normal(size=(100_000, 3))
    shell = d / np.linalg.norm(d, axis=1, keepdims=True) * (1.0 + rng.normal(0, 1e-3, (100_000, 1)))
    blob = 5.0 + 0.5 * rng.random((50_000, 3))
    pts = np.concatenate([shell, blob]).astype(np.float32)
    tree = capi.Tree.build(pts)
    qp = np.concatenate([rng.normal(0, 1e-3, (m_centre, 3)),
                         blob[rng.choice(len(blob), 200, replace=False)]]).astype(np.float32)
    q, keep = capi.make_queries(qp)
    dmax = float(np.float32(1.0))
    off = _count(tree, q, dmax)
    idx = _fill(tree, q, dmax, off)
    woff, widx = O.OracleTree(pts).rnear(keep, dmax)
    assert woff[-1] > 10_000_000
    _same_sets(off, idx, woff, widx)
    tree.free()


def test_interleaved_counts_and_another_dmax(gpu):
    from pptk_amd import capi
    rng = np.random.default_rng(43)
    pts = rng.random((100_000, 3)).astype(np.float32)
    tree = capi.Tree.build(pts)
    ot = O.OracleTree(pts)
    qa = np.ascontiguousarray(pts[:8000])
    qb = np.ascontiguousarray(pts[50_000:58_000])
    a, ka = capi.make_queries(qa)
    b, kb = capi.make_queries(qb)
    d1, d2 = float(np.float32(0.03)), float(np.float32(0.05))
    offa = _count(tree, a, d1)
    offb = _count(tree, b, d1)
    _same_sets(offa, _fill(tree, a, d1, offa), *ot.rnear(qa, d1))
    _same_sets(offb, _fill(tree, b, d1, offb), *ot.rnear(qb, d1))
    # a count at one dmax, a fill at another (with that dmax's offsets)
    _count(tree, a, d1)
    w2 = ot.rnear(qa, d2)
    _same_sets(w2[0], _fill(tree, a, d2, w2[0]), *w2)
    # self queries of every point: count then fill through the drop-in
    import pptk_amd
    kd = pptk_amd.kdtree
    t2 = kd._build(pts)
    lists = kd._query(t2, None, k=None, dmax=d1)
    off = np.concatenate([[0], np.cumsum([len(x) for x in lists])]).astype(np.int64)
    woff, widx = ot.rnear(None, d1)
    _same_sets(off, np.concatenate(lists), woff, widx)
    tree.free()


def _lidar_case(n, m, seed):
    from pptk_amd import scenes
    rng = np.random.default_rng(seed)
    pts = scenes.lidar_scene(n, seed=seed)
    qp = np.ascontiguousarray(pts[rng.choice(len(pts), m, replace=False)])
    return pts, qp


@pytest.mark.parametrize("kind", ["points", "indices", "all", "points64"])
def test_replayed_fill_equals_walked_fill(gpu, kind):
    """The first fill after a count replays the count's event lists (float32
    and float64 trees); they are dropped once used, so a second fill with the
    same offsets walks the tree. Both write every list in the same order
    (tree order), element for element, and hold the oracle's sets."""
    from pptk_amd import capi
    pts, qp = _lidar_case(300_000, 30_000, 44)
    if kind == "points64":
        pts, qp = pts.astype(np.float64), qp.astype(np.float64)
    tree = capi.Tree.build(pts)
    if kind in ("points", "points64"):
        src = qp
    elif kind == "indices":
        src = np.random.default_rng(45).choice(len(pts), 30_000, replace=False).astype(np.int32)
    else:
        src = None
    q, keep = capi.make_queries(src)
    dmax = float(np.float32(0.1))
    off = _count(tree, q, dmax)
    replayed = _fill(tree, q, dmax, off)
    walked = _fill(tree, q, dmax, off)
    assert off[-1] > 1_000_000
    assert np.array_equal(replayed, walked)
    woff, widx = O.OracleTree(pts).rnear(src, dmax)
    _same_sets(off, replayed, woff, widx)
    tree.free()


@pytest.mark.parametrize("chunks", ["0", "1", "37"])
def test_fill_when_the_event_pool_runs_out(gpu, monkeypatch, chunks):
    """PPTK_RNEAR_EVENT_CHUNKS caps the event pool of a count: with no chunks
    no list is kept, with too few the waves that could not take a chunk keep
    none and their fill waves walk the tree; the rest replay. Results equal
    those of an uncapped count and fill, element for element."""
    from pptk_amd import capi
    pts, qp = _lidar_case(200_000, 20_000, 46)
    tree = capi.Tree.build(pts)
    q, keep = capi.make_queries(qp)
    dmax = float(np.float32(0.1))
    off0 = _count(tree, q, dmax)
    want = _fill(tree, q, dmax, off0)
    monkeypatch.setenv("PPTK_RNEAR_EVENT_CHUNKS", chunks)
    off = _count(tree, q, dmax)
    got = _fill(tree, q, dmax, off)
    monkeypatch.delenv("PPTK_RNEAR_EVENT_CHUNKS")
    assert np.array_equal(off, off0)
    assert np.array_equal(got, want)
    tree.free()


def test_fill_of_other_queries_at_the_same_address(gpu):
    """A fill whose queries are not the count's (same array, new contents,
    different per-row sizes) must not replay the count's lists: the fill's
    offsets come from a count of the new contents made on another tree, so
    the rows only line up if the fill answers the new contents."""
    from pptk_amd import capi
    pts, qp = _lidar_case(200_000, 20_000, 47)
    tree = capi.Tree.build(pts)
    other = capi.Tree.build(pts)
    q, keep = capi.make_queries(qp)
    dmax = float(np.float32(0.1))
    _count(tree, q, dmax)
    keep[:] = keep[::-1].copy()  # rows reversed in place
    off = _count(other, q, dmax)
    idx = _fill(tree, q, dmax, off)
    woff, widx = O.OracleTree(pts).rnear(keep, dmax)
    _same_sets(off, idx, woff, widx)
    tree.free()
    other.free()


def test_concurrent_counts_and_fills_on_one_tree(gpu):
    """Two host threads, each on its own stream with its own query set, run
    count then fill on the same tree over and over: each thread's count
    replaces the tree's event lists, so a fill replays its own count's lists
    or finds the other thread's and walks the tree. Every result must hold the
    oracle's sets."""
    import threading

    import torch

    from pptk_amd import capi
    pts, _ = _lidar_case(200_000, 1, 48)
    tree = capi.Tree.build(pts)
    ot = O.OracleTree(pts)
    rng = np.random.default_rng(49)
    sets = [np.ascontiguousarray(pts[rng.choice(len(pts), 12_000, replace=False)]) for _ in range(2)]
    dmax = float(np.float32(0.1))
    want = [ot.rnear(s, dmax) for s in sets]
    errors = []

    def worker(i):
        try:
            stream = torch.cuda.Stream()
            sh = ctypes.c_void_p(stream.cuda_stream)
            q, keep = capi.make_queries(sets[i])
            for _ in range(6):
                off = np.empty(q.m + 1, np.int64)
                tot = ctypes.c_int64()
                capi.check(capi.lib().pptk_rnear_count(ctypes.c_void_p(tree.handle), ctypes.byref(q),
                                                       dmax, capi._p(off), ctypes.byref(tot),
                                                       capi.MEM_HOST, sh))
                idx = np.empty(max(int(off[-1]), 1), np.int32)
                capi.check(capi.lib().pptk_rnear_fill(ctypes.c_void_p(tree.handle), ctypes.byref(q),
                                                      dmax, capi._p(off), capi._p(idx),
                                                      capi.MEM_HOST, sh))
                _same_sets(off, idx[: int(off[-1])], *want[i])
        except Exception as e:  # reported by the main thread
            errors.append(e)

    threads = [threading.Thread(target=worker, args=(i,)) for i in range(2)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    tree.free()
    assert not errors, errors[0]
