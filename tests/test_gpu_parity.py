"""GPU parity: the HIP path (through the drop-in modules and the C ABI)
against the oracle on the same seeded inputs. Index work is compared
bit-exactly; normals by angle against FP64."""
import ctypes

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _kd():
    import pptk_amd
    return pptk_amd.kdtree


def _rows(lst, width=None):
    """List of int32 arrays -> (m, width) matrix padded with -1, + counts."""
    cnt = np.array([len(x) for x in lst], np.int32)
    k = width if width is not None else (int(cnt.max()) if len(cnt) else 1)
    out = np.full((len(lst), max(k, 1)), -1, np.int32)
    for i, x in enumerate(lst):
        assert x.dtype == np.int32 and x.ndim == 1
        out[i, :len(x)] = x
    return out, cnt


def assert_knn_equal(lst, want, wcnt):
    got, cnt = _rows(lst, want.shape[1])
    assert np.array_equal(cnt, wcnt)
    assert np.array_equal(got, want)


# ------------------------------------------------------------ primitives ----

def test_sort_pairs_matches_stable_argsort(gpu):
    from pptk_amd import capi
    rng = np.random.default_rng(5)
    # (n, key bits, sorted bits): even and odd pass counts, one tile and
    # hundreds, keys sharing their top digits; back-to-back calls reuse pool
    # memory holding the last call's histograms and offsets
    cases = [(1, 63, 64), (1000, 63, 64), (4097, 20, 20), (300_000, 63, 64),
             (1_000_000, 32, 32), (1_000_000, 45, 45), (3_000_000, 41, 41),
             (2_500_000, 24, 24), (2_500_000, 24, 24), (65_536, 7, 8)]
    for n, bits, end_bit in cases:
        keys = rng.integers(0, 2 ** min(bits, 63), n, dtype=np.uint64)
        keys[: n // 3] = keys[n // 2]  # duplicates exercise stability
        vals = np.arange(n, dtype=np.uint32)
        k2, v2 = keys.copy(), vals.copy()
        capi.check(capi.lib().pptk_sort_pairs_u64(capi._p(k2), capi._p(v2), n, end_bit,
                                                  capi.MEM_HOST))
        order = np.argsort(keys, kind="stable")
        assert np.array_equal(k2, keys[order]), (n, end_bit)
        assert np.array_equal(v2, vals[order]), (n, end_bit)


def test_exclusive_scan(gpu):
    from pptk_amd import capi
    rng = np.random.default_rng(6)
    # up to 733 chunks of 4096: more chunk sums than a scan block has threads
    for n in [0, 1, 4095, 4096, 4097, 1_000_003, 1_100_000, 3_000_017]:
        x = rng.integers(0, 1000, n).astype(np.int32)
        out = np.empty(n + 1, np.int64)
        capi.check(capi.lib().pptk_exclusive_scan_i32(capi._p(x), capi._p(out), n,
                                                      capi.MEM_HOST))
        want = np.concatenate([[0], np.cumsum(x, dtype=np.int64)])
        assert np.array_equal(out, want), n


# ----------------------------------------------------------------- build ----

@pytest.mark.parametrize("dim,dtype", [(3, np.float32), (3, np.float64), (2, np.float32),
                                       (4, np.float64), (3, np.int16)])
def test_tree_invariants(gpu, dim, dtype):
    from pptk_amd import capi
    rng = np.random.default_rng(dim)
    n = 50_000
    pts = (rng.random((n, dim)) * 1000).astype(dtype)
    if np.issubdtype(dtype, np.floating):
        pts[::97, dim - 1] = np.nan
        pts[::89, 0] = np.inf
    t = capi.Tree.build(pts, leaf=13)
    inf = t.info
    P, rev, fwd, box, nodes = t.export()
    valid = np.all(np.isfinite(pts.astype(np.float64)), axis=1)
    assert inf.num_valid == valid.sum()
    # reverse is a permutation of the valid rows; forward inverts it
    assert np.array_equal(np.sort(rev), np.nonzero(valid)[0])
    assert np.array_equal(fwd[rev], np.arange(inf.num_valid))
    assert np.all(fwd[~valid] == -1)
    dd = O.dist_dtype(dtype)
    assert np.array_equal(P, pts[rev].astype(dd))
    check_tree_structure(P, box, nodes, inf.leaf_size)


def check_tree_structure(P, box, nodes, leaf):
    """Every node's points are its range; children partition the parent's
    range in order; leaves hold <= leaf points and cover every point once;
    boxes are the tight boxes of their points."""
    a, b, child, parent = nodes.T
    nn = len(nodes)
    assert a[0] == 0 and b[0] == len(P) and parent[0] == -1
    inner = np.nonzero(child >= 0)[0]
    c0, c1 = child[inner], child[inner] + 1
    assert np.all(c0 % 2 == 1) and np.all(c1 < nn)
    assert np.all(parent[c0] == inner) and np.all(parent[c1] == inner)
    assert np.array_equal(a[c0], a[inner]) and np.array_equal(b[c1], b[inner])
    assert np.array_equal(b[c0], a[c1]) and np.all(a[c0] < b[c0]) and np.all(a[c1] < b[c1])
    leaves = np.nonzero(child < 0)[0]
    assert np.all(b[leaves] - a[leaves] <= leaf)
    assert (b[leaves] - a[leaves]).sum() == len(P)
    assert np.all(b[inner] - a[inner] > leaf)
    # each node is reachable from the root exactly once
    seen = np.zeros(nn, np.int32)
    seen[0] += 1
    seen[c0] += 1
    seen[c1] += 1
    assert np.all(seen == 1)
    # tight leaf boxes, parent boxes are unions of child boxes
    for i in leaves[:: max(1, len(leaves) // 500)]:
        seg = P[a[i]:b[i]]
        assert np.array_equal(box[i, 0], seg.min(0)) and np.array_equal(box[i, 1], seg.max(0))
    assert np.array_equal(box[inner, 0], np.minimum(box[c0, 0], box[c1, 0]))
    assert np.array_equal(box[inner, 1], np.maximum(box[c0, 1], box[c1, 1]))


def test_tree_structure_duplicates(gpu):
    """Many identical points: equal Morton codes split at the middle."""
    from pptk_amd import capi
    pts = np.concatenate([np.zeros((5000, 3)), np.ones((3000, 3)),
                          np.random.default_rng(0).random((100, 3))]).astype(np.float32)
    for leaf in (1, 7, 64, 200):
        t = capi.Tree.build(pts, leaf=leaf)
        P, rev, fwd, box, nodes = t.export()
        check_tree_structure(P, box, nodes, leaf)


# ------------------------------------------------------------------- kNN ----

def test_knn_self_c0_float64_bit_exact(gpu):
    """Uniform cube workload, 100k float64 points (seed 0), k=8 self query."""
    pts = np.random.RandomState(0).rand(100_000, 3)
    kd = _kd()
    tree = kd._build(pts)
    got, cnt = _rows(kd._query(tree, None, k=8))
    want, wcnt = O.OracleTree(pts).knn(None, 8)
    assert np.array_equal(cnt, wcnt)
    assert np.array_equal(got, want)


def test_normals_c0_float64_every_row_vs_fp64(gpu):
    """The uniform-cube workload's normals leg on its own input: 100k float64
    points (seed 0), estimate_normals(pts, 16, inf) (FP64 for float64 input,
    as the reference computes in T = double, estimate_normals.cpp:84-89),
    against LAPACK on the oracle's neighbourhoods, every row. The angle
    distribution goes to $PPTK_REPORT_DIR/c0_normals.json when that
    directory exists."""
    import json
    import os

    import pptk_amd
    pts = np.random.RandomState(0).rand(100_000, 3)
    ev, ew, sz = pptk_amd.estimate_normals(pts, 16, np.inf, output_eigenvalues=True,
                                           output_neighborhood_sizes=True, verbose=False)
    assert ev.dtype == np.float64 and ev.shape == (100_000, 3)
    nb, cnt = O.OracleTree(pts).knn(pts, 16)
    assert np.array_equal(sz, cnt)
    fev, few = O.normals_fp64(pts, nb, cnt)
    ang = O.angle_deg(ev, fev[:, 0])
    q = {p: float(np.percentile(ang, v)) for p, v in (("p50", 50), ("p90", 90), ("p99", 99))}
    q["max"] = float(ang.max())
    gap = (few[:, 1] - few[:, 0]) / few[:, 2]
    report = {"workload": "uniform cube 100000 float64 points (seed 0), normals k=16 r=inf",
              "rows": 100_000, "angle_deg": q,
              "fraction_eigengap_lt_1e-3": float(np.mean(gap < 1e-3))}
    d = os.environ.get("PPTK_REPORT_DIR", "reports")
    if os.path.isdir(d):
        with open(os.path.join(d, "c0_normals.json"), "w") as f:
            json.dump(report, f, indent=1)
    assert q["max"] < 0.5, q
    # smallest eigenvalue to FP64 accuracy relative to the largest
    assert np.all(np.abs(ew - few[:, 0]) <= 1e-9 * few[:, 0] + 1e-12 * few[:, 2])


@pytest.mark.parametrize("k", [1, 2, 5, 8, 16, 17, 32, 33, 64])
def test_knn_self_float32_k(gpu, k):
    rng = np.random.default_rng(k)
    pts = rng.random((30_000, 3)).astype(np.float32)
    kd = _kd()
    got, cnt = _rows(kd._query(kd._build(pts), None, k=k))
    want, wcnt = O.OracleTree(pts).knn(None, k)
    assert np.array_equal(cnt, wcnt)
    assert np.array_equal(got[:, :k], want)


@pytest.mark.parametrize("leaf", [1, 4, 10, 16, 32, 64, 100, 256])
def test_knn_leaf_size_does_not_change_results(gpu, leaf):
    rng = np.random.default_rng(11)
    pts = rng.normal(size=(20_000, 3))
    kd = _kd()
    got, _ = _rows(kd._query(kd._build(pts, maxleafsize=leaf), None, k=12))
    want, _ = O.OracleTree(pts).knn(None, 12)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("leaf", [96, 256])
def test_large_leaves_rnear_and_normals(gpu, leaf):
    """Leaves above one wavefront (scanned in chunks of 64): r-near sets and
    normals unchanged."""
    from pptk_amd import capi
    rng = np.random.default_rng(12)
    pts = rng.normal(size=(30_000, 3)).astype(np.float32)
    t = capi.Tree.build(pts, leaf=leaf)
    t16 = capi.Tree.build(pts, leaf=16)
    qi = rng.choice(len(pts), 3000, replace=False)
    dm = float(np.float32(0.15))
    off, idx = capi.rnear(t, qi, dm)
    woff, widx = O.OracleTree(pts).rnear(qi, dm)
    assert np.array_equal(off, woff)
    for i in range(len(qi)):
        assert np.array_equal(np.sort(idx[off[i]:off[i + 1]]), widx[woff[i]:woff[i + 1]])
    a = capi.normals(t, 16, np.inf)
    b = capi.normals(t16, 16, np.inf)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_knn_dmax_and_points_queries(gpu):
    rng = np.random.default_rng(3)
    pts = rng.random((40_000, 3))
    q = rng.random((5000, 3))
    kd = _kd()
    tree = kd._build(pts)
    ot = O.OracleTree(pts)
    for dmax in [0.01, 0.02, 0.05]:
        assert_knn_equal(kd._query(tree, q, k=10, dmax=dmax),
                         *ot.knn(q, 10, dmax=float(np.float32(dmax))))
    # float32 queries against a float64 tree
    assert_knn_equal(kd._query(tree, q.astype(np.float32), k=6),
                     *ot.knn(q.astype(np.float32).astype(np.float64), 6))


def test_knn_index_queries(gpu):
    rng = np.random.default_rng(4)
    n = 20_000
    pts = rng.random((n, 3)).astype(np.float32)
    kd = _kd()
    tree = kd._build(pts)
    ot = O.OracleTree(pts)
    idx = rng.integers(-n, n, 6000)
    got, _ = _rows(kd._query(tree, list(idx), k=7))
    want, _ = ot.knn(np.where(idx < 0, idx + n, idx), 7)
    assert np.array_equal(got, want)
    mask = rng.random(n) < 0.1
    got, _ = _rows(kd._query(tree, mask, k=7))
    want, _ = ot.knn(np.nonzero(mask)[0], 7)
    assert np.array_equal(got, want)
    got, _ = _rows(kd._query(tree, np.int64(5), k=3))
    assert np.array_equal(got, ot.knn(np.array([5]), 3)[0])


def test_knn_ties_duplicates_grid(gpu):
    """Integer grid with duplicated points: many exact distance ties."""
    g = np.stack(np.meshgrid(np.arange(20), np.arange(20), np.arange(10)), -1).reshape(-1, 3)
    pts = np.concatenate([g, g[::3], g[::7]]).astype(np.float64)
    kd = _kd()
    tree = kd._build(pts)
    ot = O.OracleTree(pts)
    for k in [1, 6, 7, 26, 27, 40]:
        got, cnt = _rows(kd._query(tree, None, k=k))
        want, wcnt = ot.knn(None, k)
        assert np.array_equal(cnt, wcnt), k
        assert np.array_equal(got[:, :k], want), k
    bw, _ = O.brute_knn(pts, pts[:500], 9, self_rows=np.arange(500))
    got, _ = _rows(kd._query(tree, np.arange(500), k=9))
    assert np.array_equal(got, bw)


def test_knn_float32_ties_and_denormal_distances(gpu):
    """float32 trees order neighbours by a packed (distance bits, index) key:
    exact ties on a duplicated integer grid, and a grid scaled by 3e-21 whose
    squared distances (multiples of 9e-42) are float denormals."""
    g = np.stack(np.meshgrid(np.arange(20), np.arange(20), np.arange(10)), -1).reshape(-1, 3)
    grid = np.concatenate([g, g[::3], g[::7]]).astype(np.float32)
    tiny = (g.astype(np.float32) * np.float32(3e-21)).astype(np.float32)
    assert 0 < np.float32(3e-21) ** 2 < np.finfo(np.float32).tiny
    kd = _kd()
    for pts in (grid, tiny):
        tree = kd._build(pts)
        ot = O.OracleTree(pts)
        for k in [1, 7, 16, 27, 40]:
            got, cnt = _rows(kd._query(tree, None, k=k))
            want, wcnt = ot.knn(None, k)
            assert np.array_equal(cnt, wcnt), k
            assert np.array_equal(got[:, :k], want), k
        bw, _ = O.brute_knn(pts, pts[:400], 9, self_rows=np.arange(400))
        got, _ = _rows(kd._query(tree, np.arange(400), k=9))
        assert np.array_equal(got, bw)


def test_knn_float32_self_dmax_short_rows(gpu):
    """float32 self queries over every point with a finite dmax: a dense
    cluster in a sparse background, so that many rows end before k
    neighbours (padded with -1, distance +inf) and the search bound starts
    from below dmax."""
    rng = np.random.default_rng(11)
    pts = np.concatenate([rng.random((30_000, 3)) * 0.1,
                          rng.random((10_000, 3))]).astype(np.float32)
    kd = _kd()
    tree = kd._build(pts)
    ot = O.OracleTree(pts)
    for k in [4, 10, 16, 32]:
        full = False
        for dmax in [0.004, 0.02, 0.08]:
            r = float(np.float32(dmax))
            got, cnt = _rows(kd._query(tree, None, k=k, dmax=dmax), width=k)
            want, wcnt = ot.knn(None, k, dmax=r)
            assert np.array_equal(cnt, wcnt), (k, dmax)
            assert (cnt < k).any(), (k, dmax)
            assert np.array_equal(got, want), (k, dmax)
            full = full or bool((cnt == k).any())
        assert full, k


@pytest.mark.parametrize("dtype", [np.int8, np.int32, np.uint16, np.int64, np.uint64])
def test_knn_integer_dtypes(gpu, dtype):
    rng = np.random.default_rng(8)
    hi = 100 if dtype == np.int8 else 5000
    pts = rng.integers(0, hi, (8000, 3)).astype(dtype)
    kd = _kd()
    got, cnt = _rows(kd._query(kd._build(pts), None, k=5))
    want, wcnt = O.OracleTree(pts).knn(None, 5)
    assert np.array_equal(cnt, wcnt)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("dim", [2, 4])
def test_knn_dims(gpu, dim):
    rng = np.random.default_rng(dim)
    pts = rng.random((25_000, dim))
    kd = _kd()
    tree = kd._build(pts)
    ot = O.OracleTree(pts)
    got, _ = _rows(kd._query(tree, None, k=9))
    assert np.array_equal(got, ot.knn(None, 9)[0])
    q = rng.random((2000, dim)).astype(np.float32)
    assert_knn_equal(kd._query(tree, q, k=4, dmax=0.1),
                     *ot.knn(q.astype(np.float64), 4, dmax=float(np.float32(0.1))))


def test_knn_nan_rows(gpu):
    rng = np.random.default_rng(9)
    pts = rng.random((10_000, 3))
    pts[rng.integers(0, 10_000, 300), rng.integers(0, 3, 300)] = np.nan
    pts[5, 1] = -np.inf
    kd = _kd()
    tree = kd._build(pts)
    res = kd._query(tree, None, k=4)
    ot = O.OracleTree(pts)
    assert len(res) == ot.num_valid
    got, cnt = _rows(res)
    want, wcnt = ot.knn(None, 4)
    assert np.array_equal(cnt, wcnt)
    assert np.array_equal(got, want)


def test_small_and_empty(gpu):
    kd = _kd()
    pts = np.random.default_rng(1).random((5, 3))
    got = kd._query(kd._build(pts), None, k=10)
    assert [len(x) for x in got] == [4] * 5
    want, _ = O.OracleTree(pts).knn(None, 10)
    assert np.array_equal(_rows(got)[0], want[:, :4])
    empty = np.zeros((0, 3))
    assert kd._query(kd._build(empty), None, k=3) == []
    allnan = np.full((4, 3), np.nan)
    assert kd._query(kd._build(allnan), np.zeros((2, 3)), k=3)[0].size == 0


def test_deterministic(gpu):
    rng = np.random.default_rng(12)
    pts = rng.random((200_000, 3)).astype(np.float32)
    kd = _kd()
    a = np.vstack(kd._query(kd._build(pts), None, k=8))
    b = np.vstack(kd._query(kd._build(pts), None, k=8))
    assert np.array_equal(a, b)


# ---------------------------------------------------------------- r-near ----

@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_rnear_self_and_points(gpu, dtype):
    rng = np.random.default_rng(21)
    pts = rng.random((30_000, 3)).astype(dtype)
    kd = _kd()
    tree = kd._build(pts)
    ot = O.OracleTree(pts)
    for dmax in [0.02, 0.05]:
        r = float(np.float32(dmax))
        res = kd._query(tree, None, k=None, dmax=dmax)
        off, idx = ot.rnear(None, r)
        assert len(res) == len(off) - 1
        for i in range(0, len(res), 7):
            assert np.array_equal(np.sort(res[i]), idx[off[i]:off[i + 1]]), i
        q = rng.random((6000, 3))
        res = kd._query(tree, q, k=None, dmax=dmax)
        off, idx = ot.rnear(q.astype(ot.dtype), r)
        for i in range(len(res)):
            assert np.array_equal(np.sort(res[i]), idx[off[i]:off[i + 1]]), i


def test_rnear_ragged_waves_and_heavy_queries(gpu):
    """Query counts that leave the last wave of the count and fill passes
    partly filled (the passes take 32 and 16 queries per wave), below and
    above the size where queries are Morton-ordered, on a cloud whose dense
    cluster gives thousands of hits per query next to sparse ones (the fill
    launches its heaviest waves first)."""
    rng = np.random.default_rng(23)
    pts = np.concatenate([rng.random((12_000, 3)),
                          0.5 + 0.03 * rng.random((6_000, 3))]).astype(np.float32)
    kd = _kd()
    tree = kd._build(pts)
    ot = O.OracleTree(pts)
    r = float(np.float32(0.04))
    for m in [1, 15, 17, 31, 33, 4096 + 13]:
        q = np.concatenate([rng.random((m // 2, 3)), 0.5 + 0.03 * rng.random((m - m // 2, 3))])
        res = kd._query(tree, q, k=None, dmax=r)
        off, idx = ot.rnear(q.astype(ot.dtype), r)
        assert len(res) == m
        for i in range(m):
            assert np.array_equal(np.sort(res[i]), idx[off[i]:off[i + 1]]), (m, i)
    res = kd._query(tree, None, k=None, dmax=r)
    off, idx = ot.rnear(None, r)
    assert int(off[-1]) > 50 * len(pts)
    for i in range(0, len(res), 3):
        assert np.array_equal(np.sort(res[i]), idx[off[i]:off[i + 1]]), i


def test_query_order_reuse_same_array_new_contents(gpu):
    """A tree keeps the processing order of its last query array for the next
    call on the same array and count. Rewriting the array in place between
    calls leaves a stale order behind: results must not change (only the
    order of work does)."""
    import torch
    from pptk_amd import capi
    rng = np.random.default_rng(24)
    pts = rng.random((40_000, 3)).astype(np.float32)
    t = capi.Tree.build(pts)
    ot = O.OracleTree(pts)
    r = float(np.float32(0.03))
    m = 6000
    dev = torch.device("cuda:0")
    q_d = torch.empty((m, 3), dtype=torch.float32, device=dev)
    off_d = torch.empty(m + 1, dtype=torch.int64, device=dev)
    qd, _ = capi.make_queries(dev_ptr=q_d.data_ptr(), m=m, kind=capi.Q_POINTS, dtype=np.float32)
    L = capi.lib()
    for rep in range(3):
        q = rng.random((m, 3)).astype(np.float32) * (1.0 if rep != 1 else 0.25)
        q_d.copy_(torch.from_numpy(q))
        torch.cuda.synchronize()
        total = ctypes.c_int64()
        capi.check(L.pptk_rnear_count(ctypes.c_void_p(t.handle), ctypes.byref(qd), r,
                                      ctypes.c_void_p(off_d.data_ptr()), ctypes.byref(total),
                                      capi.MEM_DEVICE, None))
        idx_d = torch.empty(max(1, total.value), dtype=torch.int32, device=dev)
        capi.check(L.pptk_rnear_fill(ctypes.c_void_p(t.handle), ctypes.byref(qd), r,
                                     ctypes.c_void_p(off_d.data_ptr()),
                                     ctypes.c_void_p(idx_d.data_ptr()), capi.MEM_DEVICE, None))
        torch.cuda.synchronize()
        off, idx = off_d.cpu().numpy(), idx_d.cpu().numpy()
        woff, widx = ot.rnear(q, r)
        assert np.array_equal(np.diff(off), np.diff(woff)), rep
        for i in range(0, m, 5):
            assert np.array_equal(np.sort(idx[off[i]:off[i + 1]]), widx[woff[i]:woff[i + 1]]), (rep, i)
    # index queries: one host array rewritten between kNN calls
    qi = rng.integers(0, len(pts), 5000).astype(np.int32)
    qs = capi.Queries()
    qs.kind, qs.m, qs.indices, qs.mem = capi.Q_INDICES, len(qi), qi.ctypes.data, capi.MEM_HOST
    gi = np.empty((len(qi), 8), np.int32)
    gc = np.empty(len(qi), np.int32)
    for rep in range(3):
        capi.check(L.pptk_knn(ctypes.c_void_p(t.handle), ctypes.byref(qs), 8, float("inf"),
                              ctypes.c_void_p(gi.ctypes.data), ctypes.c_void_p(gc.ctypes.data),
                              None, capi.MEM_HOST, None))
        wi, wc = ot.knn(qi, 8)
        assert np.array_equal(gi, wi) and np.array_equal(gc, wc), rep
        qi[:] = rng.integers(0, len(pts), len(qi))


@pytest.mark.parametrize("kind", ["points", "self"])
def test_rnear_fill_stays_inside_output_with_bad_offsets(gpu, kind):
    """Offsets that do not come from the count pass (too large, decreasing)
    must not make the fill write outside [0, offsets[m])."""
    import torch
    from pptk_amd import capi
    rng = np.random.default_rng(25)
    pts = rng.random((20_000, 3)).astype(np.float32)
    t = capi.Tree.build(pts)
    r = float(np.float32(0.05))
    dev = torch.device("cuda:0")
    if kind == "points":
        q = torch.from_numpy(rng.random((5000, 3)).astype(np.float32)).to(dev)
        qd, _ = capi.make_queries(dev_ptr=q.data_ptr(), m=5000, kind=capi.Q_POINTS,
                                  dtype=np.float32)
        m = 5000
    else:
        qd, _ = capi.make_queries(None, mem=capi.MEM_DEVICE)
        qd.kind = capi.Q_ALL
        m = len(pts)
    L = capi.lib()
    off_d = torch.empty(m + 1, dtype=torch.int64, device=dev)
    total = ctypes.c_int64()
    capi.check(L.pptk_rnear_count(ctypes.c_void_p(t.handle), ctypes.byref(qd), r,
                                  ctypes.c_void_p(off_d.data_ptr()), ctypes.byref(total),
                                  capi.MEM_DEVICE, None))
    good = off_d.cpu().numpy()
    bad = good.copy()
    bad[1:-1] = bad[1:-1][::-1]              # decreasing inside
    bad[m // 2] = good[-1] + 100_000         # past the end
    bad[-1] = good[-1] // 2                  # a smaller total
    off_d.copy_(torch.from_numpy(bad))
    guard = 1 << 16
    out = torch.full((int(bad[-1]) + guard,), -7, dtype=torch.int32, device=dev)
    capi.check(L.pptk_rnear_fill(ctypes.c_void_p(t.handle), ctypes.byref(qd), r,
                                 ctypes.c_void_p(off_d.data_ptr()),
                                 ctypes.c_void_p(out.data_ptr()), capi.MEM_DEVICE, None))
    torch.cuda.synchronize()
    assert bool((out[int(bad[-1]):] == -7).all())


def test_rnear_large_radius_whole_subtrees(gpu):
    rng = np.random.default_rng(22)
    pts = rng.random((5000, 3)).astype(np.float32)
    kd = _kd()
    res = kd._query(kd._build(pts), np.arange(0, 5000, 50), k=None, dmax=0.6)
    off, idx = O.OracleTree(pts).rnear(np.arange(0, 5000, 50), float(np.float32(0.6)))
    for i in range(len(res)):
        assert np.array_equal(np.sort(res[i]), idx[off[i]:off[i + 1]])


# --------------------------------------------------------------- normals ----

def _collinear():
    y = np.zeros((10, 3), dtype=np.float64)
    y[:, 0] = np.arange(10)
    return y


def test_normals_reference_output_types(gpu):
    """pptk tests/test_estimate_normals.py:38-138 (closed form, 8 combos)."""
    import pptk_amd
    y = _collinear()
    evals = np.array([2. / 3] + [5. / 4] * 8 + [2. / 3])
    evecs = np.array([[1., 0., 0.]] * 10)
    sizes = np.array([3] + [4] * 8 + [3])
    for ev in (0, 1):
        for ae in (0, 1):
            for ns in (0, 1):
                z = pptk_amd.estimate_normals(y, 4, 3, output_eigenvalues=ev,
                                              output_all_eigenvectors=ae,
                                              output_neighborhood_sizes=ns, verbose=False)
                if not ev and not ns:
                    assert isinstance(z, np.ndarray)
                    assert z.shape == ((10, 3, 3) if ae else (10, 3))
                    continue
                assert isinstance(z, tuple) and len(z) == 3
                assert z[0].shape == ((10, 3, 3) if ae else (10, 3))
                assert (z[1] is None) == (not ev)
                assert (z[2] is None) == (not ns)
                if ns:
                    assert np.all(z[2] == sizes)
                if ae:
                    assert np.allclose(z[0][:, -1, :], evecs)
                    if ev:
                        assert np.allclose(z[1][:, -1], evals)


def test_normals_reference_subsample(gpu):
    """pptk tests/test_estimate_normals.py:143-183."""
    import pptk_amd
    y = _collinear()
    evals = np.array([2. / 3] + [5. / 4] * 8 + [2. / 3])
    evecs = np.array([[1., 0., 0.]] * 10)
    sizes = np.array([3] + [4] * 8 + [3])
    s = [0, 3, 4, -1]
    mask = np.zeros(10, dtype=bool)
    mask[s] = True
    kw = dict(output_eigenvalues=True, output_all_eigenvectors=True,
              output_neighborhood_sizes=True, verbose=False)
    z = pptk_amd.estimate_normals(y, 4, 3, subsample=mask, **kw)
    sm = np.nonzero(mask)[0]
    assert np.allclose(z[0][:, -1, :], evecs[sm])
    assert np.allclose(z[1][:, -1], evals[sm])
    assert np.all(z[2] == sizes[sm])
    with pytest.raises(ValueError):
        pptk_amd.estimate_normals(y, 4, 3, subsample=np.zeros(2, dtype=bool), **kw)
    z = pptk_amd.estimate_normals(y, 4, 3, subsample=s, **kw)
    assert np.allclose(z[0][:, -1, :], evecs[s])
    assert np.allclose(z[1][:, -1], evals[s])
    assert np.all(z[2] == sizes[s])


@pytest.mark.parametrize("dtype,prec", [(np.float32, None), (np.float32, "fp64"),
                                        (np.float64, None), (np.float64, "mixed")])
def test_normals_vs_fp64_oracle(gpu, dtype, prec):
    import pptk_amd
    rng = np.random.default_rng(31)
    # noisy planes: well-defined normals
    n = 40_000
    uv = rng.random((n, 2)) * 10
    pts = np.stack([uv[:, 0], uv[:, 1], 0.3 * uv[:, 0] + rng.normal(0, 0.01, n)], 1)
    pts = pts.astype(dtype)
    kw = {} if prec is None else {"precision": prec}
    ev, ew, sz = pptk_amd.estimate_normals(pts, 16, np.inf, output_eigenvalues=True,
                                           output_all_eigenvectors=True,
                                           output_neighborhood_sizes=True, verbose=False, **kw)
    ot = O.OracleTree(pts)
    idx, cnt = ot.knn(pts, 16)   # neighbourhood includes the point itself
    assert np.all(sz == cnt)
    fev, few = O.normals_fp64(pts, idx, cnt)
    ang = O.angle_deg(ev[:, 0], fev[:, 0])
    assert ang.max() < 0.5, ang.max()
    assert np.allclose(ew, few, rtol=1e-3, atol=1e-6 * np.abs(few).max())


def test_normals_float32_k_with_radius_short_rows(gpu):
    """float32 normals, k = 16 within a finite radius: many neighbourhoods end
    before k points. Sizes match the oracle's exactly; the normals match an
    FP64 solve of the oracle's neighbourhoods."""
    import pptk_amd
    rng = np.random.default_rng(12)
    pts = np.concatenate([rng.random((15_000, 3)) * 0.1,
                          rng.random((5_000, 3))]).astype(np.float32)
    r = float(np.float32(0.02))
    ev, ew, sz = pptk_amd.estimate_normals(pts, 16, 0.02, output_eigenvalues=True,
                                           output_all_eigenvectors=True,
                                           output_neighborhood_sizes=True, verbose=False)
    ot = O.OracleTree(pts)
    idx, cnt = ot.knn(pts, 16, dmax=r)   # the point itself included
    assert np.array_equal(sz, cnt)
    assert (cnt < 16).any() and (cnt == 16).any()
    ok = cnt >= 4
    fev, few = O.normals_fp64(pts, idx[ok], cnt[ok])
    ang = O.angle_deg(ev[ok][:, 0], fev[:, 0])
    assert np.median(ang) < 0.01 and np.percentile(ang, 90) < 0.5, np.percentile(ang, [50, 90])


def test_normals_radius_and_k_minus_one(gpu):
    import pptk_amd
    rng = np.random.default_rng(32)
    pts = rng.random((20_000, 3))
    ot = O.OracleTree(pts)
    # k with a finite radius
    ev, ew, sz = pptk_amd.estimate_normals(pts, 10, 0.03, output_eigenvalues=True,
                                           output_all_eigenvectors=True,
                                           output_neighborhood_sizes=True, verbose=False)
    oev, oew, osz = ot.normals(10, float(np.float32(0.03)))
    assert np.array_equal(sz, osz)
    ok = sz >= 3
    assert np.allclose(ew[ok], oew[ok], rtol=1e-6, atol=1e-12)
    # pure r-near neighbourhoods
    ev, ew, sz = pptk_amd.estimate_normals(pts, -1, 0.05, output_eigenvalues=True,
                                           output_all_eigenvectors=True,
                                           output_neighborhood_sizes=True, verbose=False)
    oev, oew, osz = ot.normals(-1, float(np.float32(0.05)))
    assert np.array_equal(sz, osz)
    ok = sz >= 3
    assert np.allclose(ew[ok], oew[ok], rtol=1e-6, atol=1e-12)
    # k < -1: empty neighbourhoods, NaN output (estimate_normals.cpp:79-82)
    z = pptk_amd.estimate_normals(pts[:100], -2, 0.05, output_neighborhood_sizes=True,
                                  verbose=False)
    assert np.all(z[2] == 0) and np.all(np.isnan(z[0]))


def test_normals_large_k(gpu):
    import pptk_amd
    rng = np.random.default_rng(33)
    pts = rng.random((5000, 3))
    ev, ew, sz = pptk_amd.estimate_normals(pts, 50, np.inf, output_eigenvalues=True,
                                           output_all_eigenvectors=True,
                                           output_neighborhood_sizes=True, verbose=False)
    oev, oew, osz = O.OracleTree(pts).normals(50, np.inf)
    assert np.array_equal(sz, osz)
    assert np.allclose(ew, oew, rtol=1e-9, atol=1e-14)
