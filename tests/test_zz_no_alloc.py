"""GPU: the allocation-free kNN entry point (pptk_knn_noalloc) against
pptk_knn, bit for bit, through the C ABI.

pptk_knn_noalloc keeps its temporaries (the query order of point and index
queries and its radix sort; the heaps of k > 32 searches) in one static
device array per GPU instead of allocating them. Every case calls both entry
points on the same tree and device buffers and compares indices, counts and
distances exactly; the calls come in a row with shrinking inputs, so the
second call finds the first one's leftovers in the shared array.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX_ORDERED = 2_739_072  # point/index queries the 64 MiB array can order


def _torch():
    import torch
    return torch


def _queries(kind, m, ptr, dtype=np.float32):
    from pptk_amd import capi
    q, _ = capi.make_queries(dev_ptr=ptr, m=m, kind=kind, dtype=dtype)
    return q


def _run(entry, tree, q, nrows, k, dmax, dist_dtype):
    """Calls pptk_knn (entry='knn') or pptk_knn_noalloc on device outputs."""
    torch = _torch()
    from pptk_amd import capi
    L = capi.lib()
    tdt = torch.float32 if dist_dtype == np.float32 else torch.float64
    idx = torch.full((nrows, k), -7, dtype=torch.int32, device="cuda")
    cnt = torch.full((nrows,), -7, dtype=torch.int32, device="cuda")
    dist = torch.full((nrows, k), 123.0, dtype=tdt, device="cuda")
    torch.cuda.synchronize()
    h = ctypes.c_void_p(tree.handle)
    args = (ctypes.c_void_p(idx.data_ptr()), ctypes.c_void_p(cnt.data_ptr()),
            ctypes.c_void_p(dist.data_ptr()))
    if entry == "knn":
        st = L.pptk_knn(h, ctypes.byref(q), int(k), float(dmax), *args, capi.MEM_DEVICE, None)
    else:
        st = L.pptk_knn_noalloc(h, ctypes.byref(q), int(k), float(dmax), *args, None)
    capi.check(st)
    return idx.cpu().numpy(), cnt.cpu().numpy(), dist.cpu().numpy()


def _same(a, b):
    for x, y in zip(a, b):
        assert x.shape == y.shape
        assert np.array_equal(x, y, equal_nan=True)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_noalloc_matches_knn_points_then_smaller(gpu, dtype):
    torch = _torch()
    from pptk_amd import capi
    rng = np.random.default_rng(21)
    pts = rng.uniform(-5, 5, (300_000, 3)).astype(dtype)
    tree = capi.Tree.build(pts)
    for m, k, dmax in [(200_000, 16, np.inf), (50_000, 8, float(np.float32(0.4))),
                       (6_000, 40, np.inf), (4_000, 3, np.inf)]:
        qp = torch.from_numpy(rng.uniform(-6, 6, (m, 3)).astype(dtype)).cuda()
        q = _queries(capi.Q_POINTS, m, qp.data_ptr(), dtype)
        want = _run("knn", tree, q, m, k, dmax, dtype)
        got = _run("noalloc", tree, q, m, k, dmax, dtype)
        _same(got, want)
        assert (want[1] > 0).any()
    tree.free()


def test_noalloc_matches_knn_indices_and_all(gpu):
    torch = _torch()
    from pptk_amd import capi
    rng = np.random.default_rng(22)
    pts = rng.normal(size=(120_000, 3)).astype(np.float32)
    pts[::997] = np.nan  # invalid rows: empty self-query rows
    tree = capi.Tree.build(pts)
    for m, k in [(100_000, 12), (30_000, 33)]:
        qi = torch.from_numpy(rng.choice(len(pts), m, replace=False).astype(np.int32)).cuda()
        q = _queries(capi.Q_INDICES, m, qi.data_ptr())
        _same(_run("noalloc", tree, q, m, k, np.inf, np.float32),
              _run("knn", tree, q, m, k, np.inf, np.float32))
    nv = tree.info.num_valid
    for k in (16, 5):
        q, _ = capi.make_queries(None)
        q.mem = capi.MEM_DEVICE
        _same(_run("noalloc", tree, q, nv, k, np.inf, np.float32),
              _run("knn", tree, q, nv, k, np.inf, np.float32))
    tree.free()


def test_noalloc_rejects_what_it_cannot_hold(gpu):
    torch = _torch()
    from pptk_amd import capi
    rng = np.random.default_rng(23)
    pts = rng.uniform(0, 1, (20_000, 3)).astype(np.float32)
    tree = capi.Tree.build(pts)
    m = MAX_ORDERED + 1
    qp = torch.zeros((m, 3), dtype=torch.float32, device="cuda")
    q = _queries(capi.Q_POINTS, m, qp.data_ptr())
    with pytest.raises(ValueError, match="at most %d" % MAX_ORDERED):
        _run("noalloc", tree, q, m, 1, np.inf, np.float32)
    # host-memory queries are for pptk_knn
    hq, keep = capi.make_queries(rng.uniform(0, 1, (5000, 3)).astype(np.float32))
    with pytest.raises(ValueError, match="device memory"):
        _run("noalloc", tree, hq, 5000, 4, np.inf, np.float32)
    # the library still answers after the rejections
    qp = torch.from_numpy(pts[:5000]).cuda()
    q = _queries(capi.Q_POINTS, 5000, qp.data_ptr())
    _same(_run("noalloc", tree, q, 5000, 8, np.inf, np.float32),
          _run("knn", tree, q, 5000, 8, np.inf, np.float32))
    tree.free()
